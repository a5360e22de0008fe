"""lsmps_wls_noalloc: the WLS stage without device allocation (scratch in a
fixed device array). Its outputs, written into caller-owned device memory in
particle order, equal lsmps_wls's downloaded results bit for bit, for both
precisions, on two calls in a row with the second on fewer particles; a
context beyond LSMPS_NOALLOC_MAX_PARTICLES is refused."""
import ctypes as C
import sys

import numpy as np
import pytest

import lsmps_amd as L

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 2097152


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


class DeviceArray:
    """Caller-owned device memory (hipMalloc through the HIP runtime), filled
    with a byte pattern so unwritten entries show."""

    def __init__(self, hip, nbytes):
        self.hip, self.nbytes = hip, max(int(nbytes), 8)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        assert hip.hipMemset(self.ptr, 0xFF, C.c_size_t(self.nbytes)) == 0

    def get(self, dtype, count):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0
        return out[:count]

    def free(self):
        self.hip.hipFree(self.ptr)


def noalloc(p, re, l0, precision):
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    n = p.n
    g, lap, st = DeviceArray(hip, 24 * n), DeviceArray(hip, 8 * n), DeviceArray(hip, n)
    try:
        assert hip.hipDeviceSynchronize() == 0
        prm = L.WlsParams(re, l0, 1e-8, precision)
        res = L.WlsResult()
        rc = p.L.lsmps_wls_noalloc(p.h, C.byref(prm), g.ptr, lap.ptr, st.ptr, C.byref(res))
        p._check(rc, "wls_noalloc")
        return (g.get(np.float64, 3 * n).reshape(-1, 3), lap.get(np.float64, n), st.get(np.uint8, n),
                L.as_dict(res))
    finally:
        for a in (g, lap, st):
            a.free()


@pytest.mark.parametrize("precision", [L.WLS_FP64, L.WLS_MIXED])
def test_wls_noalloc_matches_wls(precision):
    l0 = 0.004
    re = 2.1 * l0
    p = L.Pipeline(0)
    try:
        for n in (60000, 25000):  # the second call on fewer particles
            say("noalloc", precision, n)
            xyz, t = L.generate_case("industrial_mix", n, 31, l0)
            p.upload(xyz, t)
            p.neighbors_classify(re, 128, l0)
            p.set_field(L.C1_FIELD)
            r0 = p.wls(re, l0, precision)
            g0, lap0, st0 = p.wls_results()
            g1, lap1, st1, r1 = noalloc(p, re, l0, precision)
            assert np.array_equal(st0, st1)
            assert np.array_equal(g0, g1) and np.array_equal(lap0, lap1)
            for k in ("ok", "regularized", "failed", "skipped", "fp64_redo"):
                assert r0[k] == r1[k], (k, r0, r1)
            assert (st1 == L.WLS_OK).sum() > 0.5 * n
    finally:
        p.close()


def test_wls_noalloc_refuses_beyond_limit():
    l0 = 0.004
    re = 2.1 * l0
    xyz, t = L.generate_case("industrial_mix", NOALLOC_MAX + 1000, 5, l0)
    assert len(xyz) > NOALLOC_MAX
    say("noalloc limit", len(xyz))
    p = L.Pipeline(0)
    try:
        p.upload(xyz, t)
        p.neighbors_classify(re, 128, l0)
        p.set_field(L.C1_FIELD)
        prm = L.WlsParams(re, l0, 1e-8, L.WLS_MIXED)
        res = L.WlsResult()
        rc = p.L.lsmps_wls_noalloc(p.h, C.byref(prm), None, None, None, C.byref(res))
        assert rc == -4
        assert "LSMPS_NOALLOC_MAX_PARTICLES" in p.L.lsmps_last_error(p.h).decode()
    finally:
        p.close()
