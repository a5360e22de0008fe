"""The five measurement configurations at their stated sizes, on the GPU.

C1  bulk_lattice 50,000 (50 x 50 x 20, jitter 0.1 l0, seed 7): the FP64 WLS
    of the C1 field p = 1 + x + 2y - z + x^2 + yz + z^2 on every particle,
    bit-exact against the oracle, and the PPE with the diagonal shift on the
    first 6,000 unknowns, bit-exact against the oracle's assemble_ppe with
    the same shift, solved to 1e-10.
C2  rain on a plate (tests/test_gpu_c2.py).
C3  industrial_mix 3M (tests/test_gpu_full_size.py and bench.py).
C4  the shipped learned weights on the 20 held-out seeds at 1M particles:
    every solve converges to the true relative residual 1e-8.
C5  the narrow-gap timeline: 1.5M particles, 200 steps, 2,000 droplets
    injected per step, lists rebuilt every 10 steps: particle conservation,
    every step's solve converged, per-step p50/p99.
"""
import json
import os
import sys

import numpy as np
import pytest

import lsmps_amd as L
import oracle_py as O
from lsmps_amd import learn, timeline

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pipe():
    p = L.Pipeline(0)
    yield p
    p.close()


@pytest.fixture(scope="module")
def c1():
    l0 = 0.01
    xyz, t = L.generate_case("bulk_lattice", 50000, 7, l0)
    o = O.Oracle(xyz, t)
    o.neighbors(2.1 * l0, 96)
    o.classify(l0)
    return xyz, t, o, l0


def test_c1_wls_field_full_size(pipe, c1):
    """C1 at 50k: FP64 WLS of the C1 field on every particle equals the
    oracle's bit for bit; its error against the exact gradient and
    Laplacian (4) is within 10x of the reference's stated RMS (1e-11 and
    1e-9)."""
    xyz, t, o, l0 = c1
    re = 2.1 * l0
    say("C1 WLS", len(xyz))
    pipe.upload(xyz, t)
    pipe.neighbors_classify(re, 96, l0)
    assert np.array_equal(pipe.types(), o.types())
    pipe.set_field(L.C1_FIELD)
    pipe.wls(re, l0, L.WLS_FP64)
    g, lap, st = pipe.wls_results()
    og, ol, ost = o.wls_all(L.cases.field_values(xyz, L.C1_FIELD), re, l0)
    assert np.array_equal(st, ost)
    assert np.array_equal(g, og) and np.array_equal(lap, ol)
    ok = st <= L.WLS_REGULARIZED
    # Bulk block: no Wall; only corner particles with fewer than 10
    # neighbors are Splash and skipped.
    assert ok.sum() >= len(xyz) - 8
    ge = L.cases.field_gradient(xyz, L.C1_FIELD)
    rms_g = np.sqrt(np.mean(np.sum((g[ok] - ge[ok]) ** 2, axis=1)))
    rms_l = np.sqrt(np.mean((lap[ok] - 4.0) ** 2))
    say("C1 rms gradient", rms_g, "laplacian", rms_l)
    assert rms_g < 1e-10 and rms_l < 1e-8


@pytest.mark.parametrize("shift", [0.05, 0.5])
def test_c1_diagonal_shift_bitexact(pipe, c1, shift):
    """C1's PPE: graph Laplacian with the diagonal shift on the first 6,000
    unknowns, seeded x*: CSR, b and x* equal the oracle's assemble_ppe with
    the same shift bit for bit; the shifted system solves to 1e-10 (true
    residual recomputed on the host)."""
    xyz, t, o, l0 = c1
    re = 2.1 * l0
    say("C1 shift", shift)
    pipe.upload(xyz, t)
    pipe.neighbors_classify(re, 96, l0)
    pr = pipe.assemble(re, 6000, L.XSTAR_RANDOM, 5, shift)
    A = pipe.ppe()
    o.assemble(re, 6000, 1, 5, shift)
    B = o.get_ppe()
    assert pr["unknowns"] == 6000
    for k in ("row", "col", "val", "b", "x_exact", "rev"):
        assert np.array_equal(A[k], B[k]), k
    # The shift scales every diagonal but the pin by exactly (1 + shift).
    d = A["val"][A["row"][1:] - 1]
    assert d[0] == 1.0
    for method in (L.SOLVE_JACOBI, L.SOLVE_MIXED):
        st = pipe.solve(method, 20000, 1e-10)
        x = pipe.solution()
        n = len(A["b"])
        ax = np.zeros(n)
        np.add.at(ax, np.repeat(np.arange(n), np.diff(A["row"])), A["val"] * x[A["col"]])
        res = np.linalg.norm(ax - A["b"])
        say("C1 shift", shift, "method", method, st["iterations"], "it, residual", res)
        assert st["converged"] == 1
        assert res <= 1e-10 * max(1.0, np.linalg.norm(A["b"])) * 1.01
        assert np.linalg.norm(x - A["x_exact"]) <= 1e-6 * np.linalg.norm(A["x_exact"])


def test_c4_learned_on_held_out_seeds(pipe):
    """C4: the shipped weights on the 20 held-out seeds (1001-1020) at 1M
    particles. Every FP64 Jacobi, mixed and learned + mixed solve reaches the
    true relative residual 1e-8; iterations and times are printed."""
    params = learn.load_params(learn.DEFAULT_WEIGHTS)
    pipe.set_params(params)
    seeds = list(range(1001, 1021))
    say("C4 held-out seeds", seeds[0], "..", seeds[-1], "at 1M")
    rows = learn.evaluate(pipe, seeds, 1_000_000, repeats=0)
    for r in rows:
        say("C4 seed", r["seed"], {k: (r[k]["iterations"], round(r[k]["ms"], 1), r[k]["converged"])
                                   for k in ("fp64_jacobi", "mixed", "learned")})
        for k in ("fp64_jacobi", "mixed", "learned"):
            assert r[k]["converged"] == 1, (r["seed"], k, r[k])
    s = learn.summarize(rows)
    say("C4 summary", json.dumps(s))
    assert s["fp64_jacobi"]["all_converged"] and s["mixed"]["all_converged"] and s["learned"]["all_converged"]


@pytest.mark.timeout(900)
def test_c5_timeline_full_size():
    """C5: 1.5M particles between walls 4 l0 apart, 2,000 droplets per step,
    outflow removal and neighbor rebuilds every 10 steps, 200 steps, the
    learned start with the mixed fallback: particles are conserved
    (start + injected - removed), every step's solve converges, 20 rebuilds;
    per-step p50 / p99 are printed."""
    params = learn.load_params(learn.DEFAULT_WEIGHTS)
    say("C5 timeline 1.5M x 200 steps")
    s, rows = timeline.run(particles=1_500_000, steps=200, inject=2000, rebuild=10, weights=params,
                           verbose=False)
    say("C5 summary", json.dumps(s))
    assert s["particles_end"] == s["particles_start"] + s["injected"] - s["removed"]
    assert s["injected"] == 200 * 2000
    assert sum(r["rebuild"] for r in rows) == 20
    assert s["all_converged"]
    assert s["removed"] > 0
