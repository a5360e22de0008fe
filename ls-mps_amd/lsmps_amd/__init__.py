"""Python binding of the MI355X (gfx950) LS-MPS per-step particle pipeline.

A thin ctypes layer over the C ABI in include/lsmps_hip.h
(ls-mps_amd/lib/liblsmps_hip.so). It mirrors the stages of the reference's
run_benchmark (src/lsmps.cpp:73): neighbors -> classify -> WLS -> PPE
assembly -> BiCGStab, with the reference's names and statistics. There is
no CPU fallback: if the HIP library is missing or no GPU is present, the
constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .cases import REFERENCE_FIELD, C1_FIELD, generate_case  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "liblsmps_hip.so")

INTERIOR, FREE_SURFACE, SPLASH, NEAR_WALL, WALL = range(5)
WLS_OK, WLS_REGULARIZED, WLS_FAILED, WLS_SKIPPED, WLS_NO_WEIGHT = range(5)
WLS_FP64, WLS_MIXED = 0, 1
WLS_FP32_PIVOT_DEFAULT = 0.2  # LSMPS_WLS_FP32_PIVOT_DEFAULT
SOLVE_PLAIN, SOLVE_JACOBI, SOLVE_MIXED, SOLVE_LEARNED = range(4)
XSTAR_QUADRATIC, XSTAR_RANDOM = 0, 1
TIMING_STAGES, TIMING_KERNELS = 1, 2  # LSMPS_TIMING_* flags of Pipeline.set_timing

_u64, _i32, _d, _f, _p = C.c_uint64, C.c_int32, C.c_double, C.c_float, C.c_void_p


class LsmpsError(RuntimeError):
    pass


class CnlStats(C.Structure):
    _fields_ = [("pairs", _u64), ("truncated", _u64), ("mean", _d), ("p50", _i32), ("p90", _i32), ("p99", _i32),
                ("max_count", _i32), ("grid", _i32 * 3), ("ms", _f), ("kernel_ms", _f)]


class TypeStats(C.Structure):
    _fields_ = [("count", _u64 * 5), ("ms", _f)]


class QuadField(C.Structure):
    _fields_ = [("c", _d * 10), ("lap_exact", _d)]


class WlsParams(C.Structure):
    _fields_ = [("re", _d), ("l0", _d), ("reg", _d), ("precision", _i32), ("fp32_pivot_min", _d)]


class StepParams(C.Structure):
    _fields_ = [("re", _d), ("max_neighbors", _i32), ("l0", _d), ("wls", WlsParams)]


class WlsResult(C.Structure):
    _fields_ = [("ok", _u64), ("regularized", _u64), ("failed", _u64), ("skipped", _u64), ("fp64_redo", _u64),
                ("ms", _f), ("kernel_ms", _f)]


class WlsStats(C.Structure):
    _fields_ = [("sampled", _u64), ("ok", _u64), ("regularized", _u64), ("failed", _u64), ("virtual_points", _u64),
                ("max_gradient_error", _d), ("rms_gradient_error", _d), ("max_laplacian_error", _d),
                ("rms_laplacian_error", _d)]


class PpeParams(C.Structure):
    _fields_ = [("re", _d), ("max_unknowns", _i32), ("xmode", _i32), ("xseed", C.c_uint32), ("shift", _d)]


class PpeResult(C.Structure):
    _fields_ = [("unknowns", _u64), ("nnz", _u64), ("dirichlet", _u64), ("neumann", _u64), ("failed_rows", _u64),
                ("ms", _f), ("kernel_ms", _f), ("padded_nnz", _u64), ("graph_rows", _u64)]


class ReplayInfo(C.Structure):
    _fields_ = [("rows", _u64), ("edges", _u64), ("max_row", _i32), ("pad", _i32), ("self_edges", _u64),
                ("types", _u64 * 5), ("flagged", _u64), ("ms", _f)]


class SolveParams(C.Structure):
    _fields_ = [("method", _i32), ("max_iter", _i32), ("tol", _d), ("use_x0", _i32), ("chunk", _i32)]


class SolveResult(C.Structure):
    _fields_ = [("iterations", _i32), ("converged", _i32), ("state", _i32), ("replacements", _i32),
                ("fallback", _i32), ("initial_residual", _d), ("final_residual", _d), ("relative_l2_error", _d),
                ("max_error", _d), ("ms", _f), ("breakdown_restarts", _i32)]


class TrainParams(C.Structure):
    _fields_ = [("steps", _i32), ("batch", _i32), ("lr", _f), ("wk0", _f), ("wk1", _f), ("seed", C.c_uint32)]


class TrainResult(C.Structure):
    _fields_ = [("loss_first", _d), ("loss_last", _d), ("samples", _u64), ("ms", _f)]


MLP_FEATURES, MLP_OUTPUTS, MLP_PARAMS = 16, 2, 5378


def as_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


# Every symbol the C ABI exports (include/lsmps_hip.h), with ctypes types.
ABI = {
    "lsmps_abi_version": (C.c_int, []),
    "lsmps_create": (C.c_int, [C.c_int, C.POINTER(_p)]),
    "lsmps_destroy": (None, [_p]),
    "lsmps_last_error": (C.c_char_p, [_p]),
    "lsmps_synchronize": (C.c_int, [_p]),
    "lsmps_set_timing": (C.c_int, [_p, _i32]),
    "lsmps_device_name": (C.c_int, [_p, C.c_char_p, _i32]),
    "lsmps_upload_particles": (C.c_int, [_p, _p, _p, _u64]),
    "lsmps_build_neighbors": (C.c_int, [_p, _d, _i32, C.POINTER(CnlStats)]),
    "lsmps_build_neighbors_classify": (C.c_int, [_p, _d, _i32, _d, C.POINTER(CnlStats), C.POINTER(TypeStats)]),
    "lsmps_classify": (C.c_int, [_p, _d, C.POINTER(TypeStats)]),
    "lsmps_set_field_quadratic": (C.c_int, [_p, C.POINTER(QuadField)]),
    "lsmps_set_field_values": (C.c_int, [_p, _p, _u64]),
    "lsmps_wls": (C.c_int, [_p, C.POINTER(WlsParams), C.POINTER(WlsResult)]),
    "lsmps_step": (C.c_int, [_p, C.POINTER(StepParams), C.POINTER(CnlStats), C.POINTER(TypeStats),
                             C.POINTER(WlsResult)]),
    "lsmps_wls_noalloc": (C.c_int, [_p, C.POINTER(WlsParams), _p, _p, _p, C.POINTER(WlsResult)]),
    "lsmps_wls_sampled_stats": (C.c_int, [_p, _i32, C.POINTER(WlsStats)]),
    "lsmps_assemble_ppe": (C.c_int, [_p, C.POINTER(PpeParams), C.POINTER(PpeResult)]),
    "lsmps_assemble_ppe_wls": (C.c_int, [_p, C.POINTER(PpeParams), C.c_double, C.c_double, C.POINTER(PpeResult)]),
    "lsmps_replay_load": (C.c_int, [_p, _p, _p, _p, _p, _u64, C.c_uint32, C.POINTER(ReplayInfo)]),
    "lsmps_replay_wls": (C.c_int, [_p, C.c_double, C.c_double, C.c_double, C.POINTER(WlsResult)]),
    "lsmps_replay_assemble": (C.c_int, [_p, C.c_int32, C.c_uint32, C.c_double, C.POINTER(PpeResult)]),
    "lsmps_solve_ppe": (C.c_int, [_p, C.POINTER(SolveParams), C.POINTER(SolveResult)]),
    "lsmps_upload_x0": (C.c_int, [_p, _p, _u64]),
    "lsmps_download_types": (C.c_int, [_p, _p]),
    "lsmps_download_positions": (C.c_int, [_p, _p]),
    "lsmps_download_neighbors": (C.c_int, [_p, _p, _p]),
    "lsmps_download_wls": (C.c_int, [_p, _p, _p, _p]),
    "lsmps_download_ppe": (C.c_int, [_p, _p, _p, _p, _p, _p, _p]),
    "lsmps_download_solution": (C.c_int, [_p, _p]),
    "lsmps_build_neighbors_skin": (C.c_int, [_p, _d, _d, _i32, _d, _i32, C.POINTER(CnlStats), C.POINTER(TypeStats)]),
    "lsmps_set_velocities": (C.c_int, [_p, _p, _u64]),
    "lsmps_append_particles": (C.c_int, [_p, _p, _p, _p, _u64]),
    "lsmps_advance": (C.c_int, [_p, _d]),
    "lsmps_remove_outside": (C.c_int, [_p, _p, _p, C.POINTER(_u64)]),
    "lsmps_features": (C.c_int, [_p, _p, C.POINTER(_d)]),
    "lsmps_mlp_set_params": (C.c_int, [_p, _p, _u64]),
    "lsmps_mlp_get_params": (C.c_int, [_p, _p, _u64]),
    "lsmps_mlp_predict": (C.c_int, [_p, _p, _u64, _p]),
    "lsmps_mlp_loss_grad": (C.c_int, [_p, _p, _p, _u64, _p, _p, C.POINTER(_d)]),
    "lsmps_train_collect": (C.c_int, [_p, _i32, C.POINTER(_u64)]),
    "lsmps_train_clear": (C.c_int, [_p]),
    "lsmps_mlp_train": (C.c_int, [_p, C.POINTER(TrainParams), C.POINTER(TrainResult)]),
}

_LIB = None


def load_library(path: str = LIB_PATH):
    """Loads the HIP library and binds every ABI symbol (raises if absent)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise LsmpsError(f"HIP library not built: {path} (run make -C ls-mps_amd)")
    lib = C.CDLL(path)
    for name, (res, args) in ABI.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


class Pipeline:
    """One device context holding one particle set and its stage results."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = _p()
        rc = self.L.lsmps_create(device, C.byref(h))
        if rc != 0:
            raise LsmpsError(f"lsmps_create({device}) failed with {rc}: no usable HIP device")
        self.h = h
        self.n = 0
        self.pairs = 0
        self.m = 0
        self.nnz = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.lsmps_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.lsmps_last_error(self.h).decode(errors="replace")
            raise LsmpsError(f"{what} failed ({rc}): {msg}")

    # -- stages --------------------------------------------------------------
    def upload(self, xyz, types):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(types, np.uint8)
        assert t.shape[0] == xyz.shape[0]
        self._check(self.L.lsmps_upload_particles(self.h, _ptr(xyz), _ptr(t), xyz.shape[0]), "upload")
        self.n = xyz.shape[0]

    def neighbors(self, re, max_neighbors):
        s = CnlStats()
        self._check(self.L.lsmps_build_neighbors(self.h, re, max_neighbors, C.byref(s)), "build_neighbors")
        self.pairs = s.pairs
        return as_dict(s)

    def neighbors_classify(self, re, max_neighbors, l0):
        """Neighbors and classification fused in one pass."""
        s = CnlStats()
        t = TypeStats()
        self._check(self.L.lsmps_build_neighbors_classify(self.h, re, max_neighbors, l0, C.byref(s), C.byref(t)),
                    "build_neighbors_classify")
        self.pairs = s.pairs
        return as_dict(s), as_dict(t)

    def classify(self, l0):
        s = TypeStats()
        self._check(self.L.lsmps_classify(self.h, l0, C.byref(s)), "classify")
        return as_dict(s)

    def set_field(self, field):
        """field: (c[10], lap_exact) quadratic, or an array of n values."""
        if isinstance(field, tuple):
            q = QuadField()
            for k, v in enumerate(field[0]):
                q.c[k] = v
            q.lap_exact = field[1]
            self._check(self.L.lsmps_set_field_quadratic(self.h, C.byref(q)), "set_field_quadratic")
        else:
            f = np.ascontiguousarray(field, np.float64)
            self._check(self.L.lsmps_set_field_values(self.h, _ptr(f), f.shape[0]), "set_field_values")

    def wls(self, re, l0, precision=WLS_FP64, reg=1e-8, fp32_pivot_min=0.0):
        """fp32_pivot_min (WLS_MIXED): rows whose FP32 Cholesky pivot ratio is at
        or below it are computed in FP64; 0 selects the library default."""
        p = WlsParams(re, l0, reg, precision, fp32_pivot_min)
        s = WlsResult()
        self._check(self.L.lsmps_wls(self.h, C.byref(p), C.byref(s)), "wls")
        return as_dict(s)

    def set_timing(self, flags):
        """Which HIP events neighbor builds and WLS stages record
        (TIMING_STAGES: their ms, TIMING_KERNELS: their kernel_ms)."""
        self._check(self.L.lsmps_set_timing(self.h, flags), "set_timing")

    def step(self, re, max_neighbors, l0, precision=WLS_FP64, reg=1e-8, fp32_pivot_min=0.0, wls_re=None):
        """neighbors_classify(re, max_neighbors, l0) then wls(wls_re or re, l0,
        precision, ...) as one call (one wait for both stages' readbacks)."""
        p = StepParams(re, max_neighbors, l0, WlsParams(re if wls_re is None else wls_re, l0, reg, precision,
                                                        fp32_pivot_min))
        s, t, w = CnlStats(), TypeStats(), WlsResult()
        self._check(self.L.lsmps_step(self.h, C.byref(p), C.byref(s), C.byref(t), C.byref(w)), "step")
        self.pairs = s.pairs
        return as_dict(s), as_dict(t), as_dict(w)

    def wls_sampled_stats(self, wls_samples):
        s = WlsStats()
        self._check(self.L.lsmps_wls_sampled_stats(self.h, wls_samples, C.byref(s)), "wls_sampled_stats")
        return as_dict(s)

    def assemble(self, re, max_unknowns, xmode=XSTAR_QUADRATIC, xseed=0, shift=0.0):
        p = PpeParams(re, max_unknowns, xmode, xseed, shift)
        s = PpeResult()
        self._check(self.L.lsmps_assemble_ppe(self.h, C.byref(p), C.byref(s)), "assemble_ppe")
        self.m, self.nnz = s.unknowns, s.nnz
        return as_dict(s)

    # -- synthetic CSR replay ------------------------------------------------------
    def replay_load(self, row_ptr, col, particle_type, wls_ill_conditioned, seed=0):
        """Loads a replay graph (arrays in the calibration tool's dtypes)."""
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        n = row_ptr.shape[0] - 1
        col = np.ascontiguousarray(col, np.int32)
        t = np.ascontiguousarray(particle_type, np.uint8)
        f = np.ascontiguousarray(wls_ill_conditioned, np.uint8)
        if t.shape != (n,) or f.shape != (n,):
            raise ValueError("particle_type and wls_ill_conditioned need one entry per row")
        s = ReplayInfo()
        self._check(self.L.lsmps_replay_load(self.h, _ptr(row_ptr), _ptr(col), _ptr(t), _ptr(f), n, seed,
                                             C.byref(s)), "replay_load")
        self.n = n
        self.pairs = s.edges
        return as_dict(s)

    def replay_wls(self, re, l0, reg=1e-8):
        s = WlsResult()
        self._check(self.L.lsmps_replay_wls(self.h, re, l0, reg, C.byref(s)), "replay_wls")
        return as_dict(s)

    def replay_assemble(self, max_unknowns, xseed, l0):
        s = PpeResult()
        self._check(self.L.lsmps_replay_assemble(self.h, max_unknowns, xseed, l0, C.byref(s)), "replay_assemble")
        self.m, self.nnz = s.unknowns, s.nnz
        return as_dict(s)

    def assemble_wls(self, re, l0, max_unknowns, xmode=XSTAR_QUADRATIC, xseed=0, shift=0.0, reg=1e-8):
        """WLS-Laplacian PPE rows (lsmps_assemble_ppe_wls)."""
        p = PpeParams(re, max_unknowns, xmode, xseed, shift)
        s = PpeResult()
        self._check(self.L.lsmps_assemble_ppe_wls(self.h, C.byref(p), l0, reg, C.byref(s)), "assemble_ppe_wls")
        self.m, self.nnz = s.unknowns, s.nnz
        return as_dict(s)

    def upload_x0(self, x0):
        x0 = np.ascontiguousarray(x0, np.float64)
        self._check(self.L.lsmps_upload_x0(self.h, _ptr(x0), x0.shape[0]), "upload_x0")

    def solve(self, method=SOLVE_PLAIN, max_iter=1000, tol=1e-7, use_x0=False, chunk=0):
        p = SolveParams(method, max_iter, tol, 1 if use_x0 else 0, chunk)
        s = SolveResult()
        self._check(self.L.lsmps_solve_ppe(self.h, C.byref(p), C.byref(s)), "solve_ppe")
        return as_dict(s)

    # -- downloads -------------------------------------------------------------
    def types(self):
        t = np.empty(self.n, np.uint8)
        self._check(self.L.lsmps_download_types(self.h, _ptr(t)), "download_types")
        return t

    def positions(self):
        xyz = np.empty((max(self.n, 1), 3), np.float64)
        self._check(self.L.lsmps_download_positions(self.h, _ptr(xyz)), "download_positions")
        return xyz[: self.n]

    def neighbor_lists(self):
        row = np.empty(self.n + 1, np.int32)
        col = np.empty(max(self.pairs, 1), np.int32)
        self._check(self.L.lsmps_download_neighbors(self.h, _ptr(row), _ptr(col)), "download_neighbors")
        return row, col[: self.pairs]

    def wls_results(self):
        g = np.empty((self.n, 3), np.float64)
        lap = np.empty(self.n, np.float64)
        st = np.empty(self.n, np.uint8)
        self._check(self.L.lsmps_download_wls(self.h, _ptr(g), _ptr(lap), _ptr(st)), "download_wls")
        return g, lap, st

    def ppe(self):
        m, nnz = self.m, self.nnz
        row = np.empty(m + 1, np.int32)
        col = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1), np.float64)
        b = np.empty(max(m, 1), np.float64)
        xe = np.empty(max(m, 1), np.float64)
        rev = np.empty(max(m, 1), np.int32)
        self._check(self.L.lsmps_download_ppe(self.h, _ptr(row), _ptr(col), _ptr(val), _ptr(b), _ptr(xe), _ptr(rev)),
                    "download_ppe")
        return dict(row=row, col=col[:nnz], val=val[:nnz], b=b[:m], x_exact=xe[:m], rev=rev[:m])

    def solution(self):
        x = np.empty(max(self.m, 1), np.float64)
        self._check(self.L.lsmps_download_solution(self.h, _ptr(x)), "download_solution")
        return x[: self.m]

    # -- time stepping -------------------------------------------------------------
    def neighbors_skin(self, re, skin, max_neighbors, l0, classify=True):
        s = CnlStats()
        t = TypeStats()
        self._check(self.L.lsmps_build_neighbors_skin(self.h, re, skin, max_neighbors, l0, 1 if classify else 0,
                                                      C.byref(s), C.byref(t)), "build_neighbors_skin")
        self.pairs = s.pairs
        return as_dict(s), as_dict(t)

    def set_velocities(self, uvw):
        u = np.ascontiguousarray(uvw, np.float64).reshape(-1, 3)
        self._check(self.L.lsmps_set_velocities(self.h, _ptr(u), u.shape[0]), "set_velocities")

    def append(self, xyz, types, uvw=None):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(types, np.uint8)
        u = None if uvw is None else np.ascontiguousarray(uvw, np.float64).reshape(-1, 3)
        self._check(self.L.lsmps_append_particles(self.h, _ptr(xyz), _ptr(t), _ptr(u), xyz.shape[0]), "append")
        self.n += xyz.shape[0]

    def advance(self, dt):
        self._check(self.L.lsmps_advance(self.h, dt), "advance")

    def remove_outside(self, lo, hi):
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        r = _u64()
        self._check(self.L.lsmps_remove_outside(self.h, _ptr(lo), _ptr(hi), C.byref(r)), "remove_outside")
        self.n -= r.value
        return r.value

    # -- learned preconditioner --------------------------------------------------
    def features(self):
        """(m x 16 features in reference row order, s_b)."""
        f = np.empty((max(self.m, 1), MLP_FEATURES), np.float32)
        sb = _d()
        self._check(self.L.lsmps_features(self.h, _ptr(f), C.byref(sb)), "features")
        return f[: self.m], sb.value

    def set_params(self, params):
        w = np.ascontiguousarray(params, np.float32)
        self._check(self.L.lsmps_mlp_set_params(self.h, _ptr(w), w.shape[0]), "mlp_set_params")

    def get_params(self):
        w = np.empty(MLP_PARAMS, np.float32)
        self._check(self.L.lsmps_mlp_get_params(self.h, _ptr(w), MLP_PARAMS), "mlp_get_params")
        return w

    def predict(self, feat):
        f = np.ascontiguousarray(feat, np.float32).reshape(-1, MLP_FEATURES)
        out = np.empty((max(f.shape[0], 1), MLP_OUTPUTS), np.float32)
        self._check(self.L.lsmps_mlp_predict(self.h, _ptr(f), f.shape[0], _ptr(out)), "mlp_predict")
        return out[: f.shape[0]]

    def loss_grad(self, feat, target, wk=(1.0, 1.0)):
        f = np.ascontiguousarray(feat, np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(target, np.float32).reshape(-1, MLP_OUTPUTS)
        w = np.ascontiguousarray(wk, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = _d()
        self._check(self.L.lsmps_mlp_loss_grad(self.h, _ptr(f), _ptr(t), f.shape[0], _ptr(w), _ptr(g), C.byref(loss)),
                    "mlp_loss_grad")
        return loss.value, g

    def train_collect(self, stride=1):
        tot = _u64()
        self._check(self.L.lsmps_train_collect(self.h, stride, C.byref(tot)), "train_collect")
        return tot.value

    def train_clear(self):
        self._check(self.L.lsmps_train_clear(self.h), "train_clear")

    def train(self, steps, batch=4096, lr=1e-3, wk=(1.0, 1.0), seed=1):
        p = TrainParams(steps, batch, lr, wk[0], wk[1], seed)
        r = TrainResult()
        self._check(self.L.lsmps_mlp_train(self.h, C.byref(p), C.byref(r)), "mlp_train")
        return as_dict(r)
