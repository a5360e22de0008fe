"""ctypes bindings of the CPU oracle and of the reference exporter.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker. The product (ls-mps_amd/) never
imports this module.

``Oracle`` wraps oracle/lib/liblsmps_oracle.so (the from-scratch restatement
of the reference's src/lsmps.cpp). ``RefHarness`` wraps
oracle/_ref/libref_harness.so, the reference's own src/lsmps.cpp compiled by
oracle/Makefile (present only where it was built).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "lib", "liblsmps_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_harness.so")
REF_BIN = os.path.join(HERE, "_ref", "lsmps-bench")

_p = C.c_void_p
_u64 = C.c_uint64
_i32 = C.c_int32
_d = C.c_double


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class WlsStats(C.Structure):
    _fields_ = [("sampled", _u64), ("ok", _u64), ("regularized", _u64), ("failed", _u64),
                ("virtual_points", _u64), ("max_gradient_error", _d), ("rms_gradient_error", _d),
                ("max_laplacian_error", _d), ("rms_laplacian_error", _d)]


class PpeStats(C.Structure):
    _fields_ = [("unknowns", _u64), ("nnz", _u64), ("dirichlet", _u64), ("neumann", _u64),
                ("failed_rows", _u64), ("iterations", _i32), ("converged", _i32),
                ("initial_residual", _d), ("final_residual", _d), ("relative_l2_error", _d),
                ("max_error", _d), ("assembly_ms", _d), ("solve_ms", _d), ("graph_rows", _u64)]


class RunStats(C.Structure):
    _fields_ = [("particles", _u64), ("pairs", _u64), ("truncated", _u64), ("type_count", _u64 * 5),
                ("neighbor_mean", _d), ("p50", _i32), ("p90", _i32), ("p99", _i32),
                ("neighbor_ms", _d), ("classify_ms", _d), ("wls_ms", _d), ("wls_all_ms", _d)]


def _struct_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        out[name] = list(v) if not isinstance(v, (int, float)) else v
    return out


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"oracle library not built: {ORACLE_SO} (run make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        L.orc_generate_case.restype = _u64
        L.orc_generate_case.argtypes = [C.c_char_p, _u64, C.c_uint32, _d, C.c_int, _p, _p]
        L.orc_new.restype = _p
        L.orc_new.argtypes = [_p, _p, _u64]
        L.orc_free.argtypes = [_p]
        L.orc_neighbors.argtypes = [_p, _d, C.c_int, C.POINTER(_u64), C.POINTER(_u64)]
        L.orc_neighbors_brute.argtypes = [_p, _d, C.c_int, C.POINTER(_u64), C.POINTER(_u64)]
        L.orc_get_neighbors.argtypes = [_p, _p, _p]
        L.orc_set_neighbors.argtypes = [_p, _p, _p]
        L.orc_classify.argtypes = [_p, _d]
        L.orc_get_types.argtypes = [_p, _p]
        L.orc_set_types.argtypes = [_p, _p]
        L.orc_wls_all.argtypes = [_p, _p, _d, _d, _d, _p, _p, _p]
        L.orc_wls_moments.restype = _d
        L.orc_wls_moments.argtypes = [_p, _p, _d, _d, _u64, _p, _p]
        L.orc_wls_test.argtypes = [_p, _d, _d, _d, C.c_int, C.POINTER(WlsStats)]
        L.orc_assemble_ppe.argtypes = [_p, _d, C.c_int, C.c_int, C.c_uint32, _d, C.POINTER(PpeStats)]
        L.orc_assemble_ppe_wls.argtypes = [_p, _d, _d, _d, C.c_int, C.c_int, C.c_uint32, _d, C.POINTER(PpeStats)]
        L.orc_get_ppe.argtypes = [_p, _p, _p, _p, _p, _p, _p]
        L.orc_solve_ppe.argtypes = [_p, C.c_int, _d, C.c_int, _p, C.POINTER(PpeStats), _p]
        L.orc_run.argtypes = [_p, _d, _d, C.c_int, C.c_int, C.c_int, C.c_int, _d, C.c_int,
                              C.POINTER(RunStats), C.POINTER(WlsStats), C.POINTER(PpeStats)]
        L.orc_features.argtypes = [_p, _d, _p, C.POINTER(_d)]
        L.orc_hashed_uniform.restype = _d
        L.orc_hashed_uniform.argtypes = [C.c_uint32, _u64]
        L.orc_mlp_forward.argtypes = [_p, _p, _u64, _p]
        L.orc_mlp_loss_grad.restype = _d
        L.orc_mlp_loss_grad.argtypes = [_p, _p, _p, _p, _u64, _p]
        _LIB = L
    return _LIB


def generate_case(case_name, particles, seed, l0=0.01, film_layers=4):
    """Reference generator restated (src/lsmps.cpp:68,72) -> (xyz[n,3], type[n])."""
    L = lib()
    n = L.orc_generate_case(case_name.encode(), particles, seed, l0, film_layers, None, None)
    xyz = np.empty((n, 3), np.float64)
    t = np.empty(n, np.uint8)
    L.orc_generate_case(case_name.encode(), particles, seed, l0, film_layers, _ptr(xyz), _ptr(t))
    return xyz, t


class Oracle:
    """One particle set in the CPU oracle (handle around ``orc_state``)."""

    def __init__(self, xyz, types):
        self.xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        self.n = self.xyz.shape[0]
        t = np.ascontiguousarray(types, np.uint8)
        self.L = lib()
        self.h = self.L.orc_new(_ptr(self.xyz), _ptr(t), self.n)
        self.pairs = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_free(self.h)
            self.h = None

    def neighbors(self, re, maxn, brute=False):
        pairs, trunc = _u64(), _u64()
        fn = self.L.orc_neighbors_brute if brute else self.L.orc_neighbors
        fn(self.h, re, maxn, C.byref(pairs), C.byref(trunc))
        self.pairs = pairs.value
        return pairs.value, trunc.value

    def get_neighbors(self):
        row = np.empty(self.n + 1, np.int32)
        col = np.empty(max(self.pairs, 1), np.int32)
        self.L.orc_get_neighbors(self.h, _ptr(row), _ptr(col))
        return row, col[: self.pairs]

    def set_neighbors(self, row, col):
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        self.L.orc_set_neighbors(self.h, _ptr(row), _ptr(col))
        self.pairs = int(row[-1])

    def classify(self, l0):
        self.L.orc_classify(self.h, l0)

    def types(self):
        t = np.empty(self.n, np.uint8)
        self.L.orc_get_types(self.h, _ptr(t))
        return t

    def set_types(self, t):
        t = np.ascontiguousarray(t, np.uint8)
        self.L.orc_set_types(self.h, _ptr(t))

    def wls_all(self, f, re, l0, reg=1e-8):
        f = np.ascontiguousarray(f, np.float64)
        g = np.empty((self.n, 3), np.float64)
        lap = np.empty(self.n, np.float64)
        st = np.empty(self.n, np.uint8)
        self.L.orc_wls_all(self.h, _ptr(f), re, l0, reg, _ptr(g), _ptr(lap), _ptr(st))
        return g, lap, st

    def wls_moments(self, f, re, l0, i):
        f = np.ascontiguousarray(f, np.float64)
        m = np.empty((9, 9), np.float64)
        b = np.empty(9, np.float64)
        ws = self.L.orc_wls_moments(self.h, _ptr(f), re, l0, i, _ptr(m), _ptr(b))
        return ws, m, b

    def wls_test(self, re, l0, samples, reg=1e-8):
        st = WlsStats()
        self.L.orc_wls_test(self.h, re, l0, reg, samples, C.byref(st))
        return _struct_dict(st)

    def assemble(self, re, max_unknowns, xmode=0, xseed=0, shift=0.0):
        st = PpeStats()
        self.L.orc_assemble_ppe(self.h, re, max_unknowns, xmode, xseed, shift, C.byref(st))
        self._ppe = _struct_dict(st)
        return self._ppe

    def assemble_wls(self, re, l0, max_unknowns, xmode=0, xseed=0, shift=0.0, reg=1e-8):
        st = PpeStats()
        self.L.orc_assemble_ppe_wls(self.h, re, l0, reg, max_unknowns, xmode, xseed, shift, C.byref(st))
        self._ppe = _struct_dict(st)
        return self._ppe

    def get_ppe(self):
        m, nnz = self._ppe["unknowns"], self._ppe["nnz"]
        row = np.empty(m + 1, np.int32)
        col = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1), np.float64)
        b = np.empty(max(m, 1), np.float64)
        xe = np.empty(max(m, 1), np.float64)
        rev = np.empty(max(m, 1), np.int32)
        self.L.orc_get_ppe(self.h, _ptr(row), _ptr(col), _ptr(val), _ptr(b), _ptr(xe), _ptr(rev))
        return dict(row=row, col=col[:nnz], val=val[:nnz], b=b[:m], x_exact=xe[:m], rev=rev[:m])

    def features(self, re):
        """(m x 16 learned-preconditioner features in row order, s_b)."""
        m = len(self.get_ppe()["b"])
        f = np.empty((max(m, 1), 16), np.float32)
        sb = _d()
        self.L.orc_features(self.h, re, _ptr(f), C.byref(sb))
        return f[:m], sb.value

    def solve(self, max_iter, tol, precond=0, x0=None):
        st = PpeStats()
        m = self._ppe["unknowns"]
        x = np.empty(max(m, 1), np.float64)
        x0a = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        self.L.orc_solve_ppe(self.h, max_iter, tol, precond, _ptr(x0a), C.byref(st), _ptr(x))
        return _struct_dict(st), x[:m]

    def run(self, l0, re_ratio, maxn, wls_samples, max_unknowns, max_iter, tol, do_solve=True):
        rs, ws, ps = RunStats(), WlsStats(), PpeStats()
        self.L.orc_run(self.h, l0, re_ratio, maxn, wls_samples, max_unknowns, max_iter, tol,
                       1 if do_solve else 0, C.byref(rs), C.byref(ws), C.byref(ps))
        return _struct_dict(rs), _struct_dict(ws), _struct_dict(ps)


def mlp_forward(params, feat):
    L = lib()
    params = np.ascontiguousarray(params, np.float32)
    feat = np.ascontiguousarray(feat, np.float32)
    n = feat.shape[0]
    out = np.empty((n, 2), np.float32)
    L.orc_mlp_forward(_ptr(params), _ptr(feat), n, _ptr(out))
    return out


def mlp_loss_grad(params, feat, target, wk):
    L = lib()
    params = np.ascontiguousarray(params, np.float32)
    feat = np.ascontiguousarray(feat, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    wk = np.ascontiguousarray(wk, np.float32)
    g = np.empty_like(params)
    loss = L.orc_mlp_loss_grad(_ptr(params), _ptr(feat), _ptr(target), _ptr(wk), feat.shape[0], _ptr(g))
    return loss, g


class RefHarness:
    """The reference's own src/lsmps.cpp behind a C ABI (oracle/_ref)."""

    def __init__(self):
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO)
        L = C.CDLL(REF_SO)
        L.ref_setup.restype = _u64
        L.ref_setup.argtypes = [C.c_char_p, _u64, C.c_uint32, _d, _d, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_int, _d]
        L.ref_setup_particles.argtypes = [_p, _p, _u64, _d, _d, C.c_int, C.c_int, C.c_int, C.c_int, _d]
        L.ref_get_particles.argtypes = [_p, _p]
        L.ref_neighbors.restype = _u64
        L.ref_neighbors.argtypes = [C.POINTER(_u64)]
        L.ref_get_neighbors.argtypes = [_p, _p]
        L.ref_wls_test.argtypes = [_p, _p]
        L.ref_wls_particles.argtypes = [_p, _u64, _p, _p, _p]
        L.ref_assemble.restype = _u64
        L.ref_assemble.argtypes = [C.POINTER(_u64)] * 4
        L.ref_get_ppe.argtypes = [_p] * 5
        L.ref_solve.argtypes = [C.POINTER(_i32), C.POINTER(_i32), _p]
        L.ref_run_config.restype = _d
        L.ref_run_config.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
        L.ref_time_stages.argtypes = [_p]
        self.L = L
        self.n = 0

    def setup(self, case, particles, seed, l0, re_ratio=2.1, film_layers=4, maxn=96, wls_samples=4096,
              max_unknowns=6000, max_iter=1000, tol=1e-7):
        self.n = self.L.ref_setup(case.encode(), particles, seed, l0, re_ratio, film_layers, maxn, wls_samples,
                                  max_unknowns, max_iter, tol)
        return self.n

    def setup_particles(self, xyz, types, l0, re_ratio=2.1, maxn=96, wls_samples=4096, max_unknowns=6000,
                        max_iter=1000, tol=1e-7):
        xyz = np.ascontiguousarray(xyz, np.float64)
        t = np.ascontiguousarray(types, np.uint8)
        self.n = xyz.shape[0]
        self.L.ref_setup_particles(_ptr(xyz), _ptr(t), self.n, l0, re_ratio, maxn, wls_samples, max_unknowns,
                                   max_iter, tol)

    def particles(self):
        xyz = np.empty((self.n, 3), np.float64)
        t = np.empty(self.n, np.uint8)
        self.L.ref_get_particles(_ptr(xyz), _ptr(t))
        return xyz, t

    def neighbors(self):
        tr = _u64()
        pairs = self.L.ref_neighbors(C.byref(tr))
        row = np.empty(self.n + 1, np.int32)
        col = np.empty(max(pairs, 1), np.int32)
        self.L.ref_get_neighbors(_ptr(row), _ptr(col))
        return row, col[:pairs], tr.value

    def classify(self):
        self.L.ref_classify()

    def wls_test(self):
        c = np.empty(5, np.uint64)
        e = np.empty(4, np.float64)
        self.L.ref_wls_test(_ptr(c), _ptr(e))
        keys = ["sampled", "ok", "regularized", "failed", "virtual_points"]
        out = {k: int(v) for k, v in zip(keys, c)}
        out.update(dict(zip(["max_gradient_error", "rms_gradient_error", "max_laplacian_error",
                             "rms_laplacian_error"], map(float, e))))
        return out

    def wls_particles(self, idx):
        idx = np.ascontiguousarray(idx, np.uint64)
        g = np.empty((len(idx), 3))
        lap = np.empty(len(idx))
        st = np.empty(len(idx), np.uint8)
        self.L.ref_wls_particles(_ptr(idx), len(idx), _ptr(g), _ptr(lap), _ptr(st))
        return g, lap, st

    def assemble(self):
        u, f, ne, di = _u64(), _u64(), _u64(), _u64()
        nnz = self.L.ref_assemble(C.byref(u), C.byref(f), C.byref(ne), C.byref(di))
        m = u.value
        row = np.empty(m + 1, np.int32)
        col = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1))
        b = np.empty(max(m, 1))
        xe = np.empty(max(m, 1))
        self.L.ref_get_ppe(_ptr(row), _ptr(col), _ptr(val), _ptr(b), _ptr(xe))
        return dict(row=row, col=col[:nnz], val=val[:nnz], b=b[:m], x_exact=xe[:m], unknowns=m,
                    failed_rows=f.value, neumann=ne.value, dirichlet=di.value, nnz=nnz)

    def solve(self):
        it, cv = _i32(), _i32()
        o = np.empty(5)
        self.L.ref_solve(C.byref(it), C.byref(cv), _ptr(o))
        return dict(iterations=it.value, converged=cv.value, initial_residual=o[0], final_residual=o[1],
                    relative_l2_error=o[2], max_error=o[3], solve_ms=o[4])

    def time_stages(self):
        o = np.empty(4)
        self.L.ref_time_stages(_ptr(o))
        return dict(neighbor_ms=o[0], classify_ms=o[1], wls_ms=o[2], assembly_ms=o[3])
