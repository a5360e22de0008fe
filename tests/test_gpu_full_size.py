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
        wm = p.wls(RE, L0, L.WLS_MIXED)
        gm, lapm, stm = p.wls_results()
        p.wls(RE, L0, L.WLS_MIXED, fp32_pivot_min=1e-3)
        gm3, lapm3, stm3 = p.wls_results()
        p.wls(RE, L0, L.WLS_FP64)
        g, lap, st = p.wls_results()
        # x* ~ U(-1, 1) with seed 11, as the benchmark solves it.
        p.assemble(RE, 1 << 30, L.XSTAR_RANDOM, 11)
        A = p.ppe()
        ss = p.solve(L.SOLVE_MIXED, 20000, 1e-8)
        x = p.solution()
        # The reference's own smooth x* = x^2 + y^2: |b| = 4.9e5, so the
        # target residual is far below what the FP32 matrix resolves on
        # |A||x|; the mixed refinement stalls and FP64 Jacobi finishes from
        # its iterate (fallback), which meets 1e-8 on the true residual.
        p.assemble(RE, 1 << 30, L.XSTAR_QUADRATIC, 0)
        Aq = p.ppe()
        sq = p.solve(L.SOLVE_MIXED, 20000, 1e-8)
        xq = p.solution()
    finally:
        p.close()
    say("x* = x^2 + y^2, mixed solve:", sq)
    assert sq["converged"] == 1 and sq["fallback"] == 1
    axq = np.add.reduceat(Aq["val"] * xq[Aq["col"]], Aq["row"][:-1])
    resq = np.linalg.norm(Aq["b"] - axq)
    assert resq <= 1e-8 * max(1.0, np.linalg.norm(Aq["b"])) * 1.01, resq
    del Aq, xq, axq
    say("GPU done: pairs", int(row[-1]), "unknowns", len(A["b"]), "mixed solve", ss, f"{time.time() - t0:.1f} s")
    # Mixed WLS. An FP32 row's Laplacian error grows as its smallest
    # Cholesky pivot ratio falls: the curvature is a small part of the fit
    # next to the gradient (|grad| ~ 5 against a Laplacian of 4 over
    # re = 8.4e-3), so the FP32 rounding of the linear part lands on it,
    # amplified by the conditioning. Rows with a pivot ratio at or below the
    # default 0.2 (about 10% here, the one-sided surface stencils) are
    # computed in FP64 and are the FP64 kernel's bits. Measured on MI355X:
    # RMS Laplacian error 7.9e-5 and largest 8e-3 (with the ratio at 1e-3,
    # the three-digit rule: 6.1e-2 and 39.5). Bounds: 1e-3 and 5e-2.
    assert np.array_equal(stm == L.WLS_SKIPPED, st == L.WLS_SKIPPED)
    assert np.array_equal(stm, st)
    same = np.all(gm == g, axis=1) & (lapm == lap)
    say("mixed WLS:", wm, "rows equal to FP64:", int(same.sum()))
    assert int(same.sum()) >= wm["fp64_redo"] + wm["skipped"]
    ok = (st <= 1) & (stm <= 1)
    ge = L.cases.field_gradient(xyz, L.C1_FIELD)
    eg = np.sqrt(np.mean((gm[ok] - ge[ok]) ** 2))
    dl = np.abs(lapm[ok] - 4.0)
    el = np.sqrt(np.mean(dl ** 2))
    el64 = np.sqrt(np.mean((lap[ok] - 4.0) ** 2))
    el3 = np.sqrt(np.mean((lapm3[ok] - 4.0) ** 2))
    say("mixed WLS RMS error: gradient", eg, "Laplacian", el, "max", dl.max(), "(FP64 Laplacian", el64,
        "; pivot ratio 1e-3:", el3, ")")
    assert eg < 1e-5
    assert el < 1e-3 and dl.max() < 5e-2
    # With the ratio at 1e-3 (the three-digit rule) the FP32 error of the
    # ill-conditioned rows shows: the RMS is far larger.
    assert el3 > 10 * el
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
