"""Parity at the benchmark's full size: the 3,000,000-particle industrial_mix
scene (l0 = 0.004, re = 2.1 l0, max_neighbors 128), GPU against the CPU
oracle, bit for bit: neighbor lists in the reference's order
(src/lsmps.cpp:26), types (src/lsmps.cpp:27), the FP64 WLS gradient,
Laplacian and status of every particle (src/lsmps.cpp:28) and the PPE CSR,
b and x* (src/lsmps.cpp:29-64; x* from the seeded hash, as the benchmark
solves it). Also at that size: the mixed WLS against the
manufactured quadratic, and the mixed BiCGStab's solution against 1e-8 on
the true residual recomputed on the host. The oracle takes about 10 s of
CPU here.
"""
import os
import sys
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "ls-mps_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import lsmps_amd as L  # noqa: E402
import oracle_py as O  # noqa: E402

pytestmark = pytest.mark.gpu

N, L0, MAXN = 3_000_000, 0.004, 128
RE = 2.1 * L0


def say(*a):
    print("[gpu]", *a, flush=True)


def test_industrial_mix_3m_bitexact():
    t0 = time.time()
    xyz, types = L.generate_case("industrial_mix", N, 23, L0)
    say("industrial_mix", N, "particles generated")
    p = L.Pipeline(0)
    try:
        p.upload(xyz, types)
        cs, _ = p.neighbors_classify(RE, MAXN, L0)
        row, col = p.neighbor_lists()
        gtypes = p.types()
        # The same particles again: this build reads them in the last
        # build's sorted order; lists and types are the same, and the stages
        # below (checked against the oracle) run on it.
        p.neighbors_classify(RE, MAXN, L0)
        row2, col2 = p.neighbor_lists()
        assert np.array_equal(row, row2) and np.array_equal(col, col2)
        assert np.array_equal(gtypes, p.types())
        p.set_field(L.C1_FIELD)
        p.wls(RE, L0, L.WLS_MIXED)
        gm, lapm, stm = p.wls_results()
        p.wls(RE, L0, L.WLS_FP64)
        g, lap, st = p.wls_results()
        # x* ~ U(-1, 1) with seed 11, as the benchmark solves it.
        p.assemble(RE, 1 << 30, L.XSTAR_RANDOM, 11)
        A = p.ppe()
        ss = p.solve(L.SOLVE_MIXED, 20000, 1e-8)
        x = p.solution()
    finally:
        p.close()
    say("GPU done: pairs", int(row[-1]), "unknowns", len(A["b"]), "mixed solve", ss, f"{time.time() - t0:.1f} s")
    # Mixed WLS: the same particles skipped, and FP32 moments keep the
    # manufactured quadratic's gradient and Laplacian (4). The gradient bound
    # is the small cases' 1e-3. The Laplacian's RMS error measured here is
    # 0.061 (l0 = 0.004 on a 3.1 m domain; the small cases at l0 = 0.01 stay
    # below 0.05), so its bound here is 0.075.
    assert np.array_equal(stm == L.WLS_SKIPPED, st == L.WLS_SKIPPED)
    ok = (st <= 1) & (stm <= 1)
    ge = L.cases.field_gradient(xyz, L.C1_FIELD)
    eg = np.sqrt(np.mean((gm[ok] - ge[ok]) ** 2))
    el = np.sqrt(np.mean((lapm[ok] - 4.0) ** 2))
    el64 = np.sqrt(np.mean((lap[ok] - 4.0) ** 2))
    say("mixed WLS RMS error: gradient", eg, "Laplacian", el, "(FP64 Laplacian", el64, ")")
    assert eg < 1e-3
    assert el < 0.075
    # The mixed solve meets 1e-8 on the true FP64 residual, recomputed here
    # from the downloaded CSR.
    assert ss["converged"] == 1
    m = len(A["b"])
    ax = np.add.reduceat(A["val"] * x[A["col"]], A["row"][:-1]) if m else np.zeros(0)
    res = np.linalg.norm(A["b"] - ax)
    say("mixed solve: true residual", res, "bound", 1e-8 * max(1.0, np.linalg.norm(A["b"])))
    assert res <= 1e-8 * max(1.0, np.linalg.norm(A["b"])) * 1.01

    o = O.Oracle(xyz, types)
    pairs, trunc = o.neighbors(RE, MAXN)
    orow, ocol = o.get_neighbors()
    assert cs["truncated"] == trunc
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    del orow, ocol, col
    say("neighbor lists equal", f"{time.time() - t0:.1f} s")
    o.classify(L0)
    assert np.array_equal(gtypes, o.types())
    og, ol, ost = o.wls_all(L.cases.field_values(xyz, L.C1_FIELD), RE, L0)
    assert np.array_equal(st, ost)
    assert np.array_equal(g, og) and np.array_equal(lap, ol)
    say("types and FP64 WLS equal", f"{time.time() - t0:.1f} s")
    o.assemble(RE, 1 << 30, xmode=1, xseed=11)
    B = o.get_ppe()
    for k in ("row", "col", "val", "b", "x_exact"):
        assert np.array_equal(A[k], B[k]), k
    say("PPE equal, nnz", len(B["val"]), f"{time.time() - t0:.1f} s")
