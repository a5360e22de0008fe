"""Synthetic particle sets and manufactured fields.

``generate_case`` calls the host generator of liblsmps_hip.so
(lsmps_generate_case_c, include/lsmps_host.h), which is the library's
generate_case: the reference's cases (src/lsmps.cpp:68,72) bit for bit,
plus the measurement scenes bulk_lattice, rain_plate and industrial_mix.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "liblsmps_hip.so")

# Quadratic fields as (c0, cx, cy, cz, cxx, cyy, czz, cxy, cxz, cyz), exact Laplacian.
# The reference's test field (src/lsmps.cpp:28): Laplacian 1.8.
REFERENCE_FIELD = ((1.0, 2.0, -3.0, 0.5, 0.7, -0.2, 0.4, 0.3, -0.1, 0.25), 1.8)
# p = 1 + x + 2y - z + x^2 + yz + z^2: Laplacian 4.
C1_FIELD = ((1.0, 1.0, 2.0, -1.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.0), 4.0)

_GEN = None


def _gen():
    global _GEN
    if _GEN is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"library not built: {LIB_PATH}")
        lib = C.CDLL(LIB_PATH)
        fn = lib.lsmps_generate_case_c
        fn.restype = C.c_uint64
        fn.argtypes = [C.c_char_p, C.c_uint64, C.c_uint32, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
        _GEN = fn
    return _GEN


def generate_case(case_name: str, particles: int, seed: int, l0: float = 0.01, film_layers: int = 4):
    """-> (xyz float64[n,3], type uint8[n])"""
    fn = _gen()
    n = fn(case_name.encode(), particles, seed, l0, film_layers, None, None)
    xyz = np.empty((n, 3), np.float64)
    t = np.empty(n, np.uint8)
    fn(case_name.encode(), particles, seed, l0, film_layers, xyz.ctypes.data_as(C.c_void_p),
       t.ctypes.data_as(C.c_void_p))
    return xyz, t


def field_values(xyz, field):
    """Evaluates a quadratic field at positions, in the library's summation order."""
    c, _ = field
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * x + c[5] * y * y + c[6] * z * z + c[7] * x * y
            + c[8] * x * z + c[9] * y * z)


def field_gradient(xyz, field):
    c, _ = field
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    return np.stack([c[1] + 2 * c[4] * x + c[7] * y + c[8] * z, c[2] + 2 * c[5] * y + c[7] * x + c[9] * z,
                     c[3] + 2 * c[6] * z + c[8] * x + c[9] * y], 1)
