"""The learned preconditioner on the GPU, against the CPU oracle.

- Features of every PPE row: equal to the oracle's (same definitions, FP64
  sums in another order, stored as FP32).
- The network's forward pass and the training kernels' loss and gradient:
  equal to the oracle's FP32/FP64 restatement evaluated at the same
  BF16-rounded weight matrices.
- Training on self-generated solves lowers the loss; on held-out scenes
  the learned solve meets the same true-residual tolerance as the plain
  method. The test prints what the network gains (initial residual,
  iterations and time against the plain mixed solver) and asserts that
  the learned start lowers the initial residual.
- A network whose start raises the residual is detected and the plain
  mixed method runs instead (fallback = 1).
"""
import sys

import numpy as np
import pytest

import lsmps_amd as L
import oracle_py as O
from lsmps_amd import learn

pytestmark = pytest.mark.gpu

W_SLICES = ((0, 1024), (1088, 5184), (5248, 5376))  # weight matrices; biases stay FP32


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def bf16_weights(p):
    """The parameters as the forward pass sees them: weight matrices rounded
    to BF16 (nearest even), biases unchanged."""
    q = np.array(p, np.float32, copy=True)
    for a, b in W_SLICES:
        u = q[a:b].view(np.uint32).astype(np.uint64)
        u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
        q[a:b] = u.astype(np.uint32).view(np.float32)
    return q


@pytest.fixture(scope="module")
def pipe():
    p = L.Pipeline(0)
    yield p
    p.close()


@pytest.mark.parametrize("case,n", [("wall_film", 4000), ("industrial_mix", 6000), ("rain_injection", 3000)])
def test_features_match_oracle(pipe, case, n):
    say("features", case, n)
    l0, re = 0.01, 0.021
    xyz, t = L.generate_case(case, n, 7, l0)
    pipe.upload(xyz, t)
    pipe.neighbors_classify(re, 96, l0)
    pipe.assemble(re, 1 << 30, L.XSTAR_RANDOM, 5)
    fg, sg = pipe.features()
    o = O.Oracle(xyz, t)
    o.neighbors(re, 96)
    o.classify(l0)
    o.assemble(re, 1 << 30, 1, 5)
    fo, so = o.features(re)
    assert fg.shape == fo.shape and fg.shape[0] > 0
    assert abs(sg - so) <= 1e-12 * so
    np.testing.assert_allclose(fg, fo, rtol=2e-6, atol=2e-6)


def test_mlp_forward_and_gradient_match_oracle(pipe):
    rng = np.random.default_rng(3)
    p = learn.init_params(3)
    p[5248:5376] *= 10.0  # a last layer of ordinary size
    p[1024:1088] = rng.normal(0, 0.1, 64)
    p[5184:5248] = rng.normal(0, 0.1, 64)
    p[5376:] = (0.3, -0.2)
    pipe.set_params(p)
    feat = rng.normal(0, 1, (3000, 16)).astype(np.float32)
    target = rng.normal(0, 1, (3000, 2)).astype(np.float32)
    q = bf16_weights(p)
    yg = pipe.predict(feat)
    yo = O.mlp_forward(q, feat)
    np.testing.assert_allclose(yg, yo, rtol=1e-4, atol=1e-5)
    wk = (1.0, 0.5)
    lg, gg = pipe.loss_grad(feat, target, wk)
    lo, go = O.mlp_loss_grad(q, feat, target, wk)
    assert abs(lg - lo) <= 1e-4 * abs(lo)
    for a, b in ((0, 1024), (1024, 1088), (1088, 5184), (5184, 5248), (5248, 5376), (5376, 5378)):
        scale = np.abs(go[a:b]).max()
        assert np.abs(gg[a:b] - go[a:b]).max() <= 1e-3 * scale + 1e-7, (a, b)
    # A finite difference of the GPU loss along the GPU gradient, in FP32
    # master weights (the BF16 rounding is steady for a small step here
    # only on the biases, so the step moves the biases).
    d = np.zeros_like(p)
    d[1024:1088] = gg[1024:1088]
    d[5184:5248] = gg[5184:5248]
    d[5376:] = gg[5376:]
    eps = 1e-3 / max(1e-12, np.linalg.norm(d))
    pipe.set_params(p + eps * d)
    lp, _ = pipe.loss_grad(feat, target, wk)
    pipe.set_params(p - eps * d)
    lm, _ = pipe.loss_grad(feat, target, wk)
    fd = (lp - lm) / (2 * eps)
    assert abs(fd - float(d @ d)) <= 2e-2 * float(d @ d)


def test_training_and_learned_solve(pipe):
    pipe.set_params(learn.init_params(0))
    pipe.train_clear()
    n_train = learn.collect(pipe, range(1, 9), stride=2, sizes=[20000] * 8)
    say("training rows", n_train)
    curve = learn.train(pipe, 600, batch=2048, lr=3e-3, chunks=3)
    first, last = curve[0][1], curve[-1][2]
    say("loss", first, "->", last)
    assert last < 0.8 * first
    rows = learn.evaluate(pipe, [101, 102, 103], 30000, tol=1e-8)
    for r in rows:
        say("held-out", r["seed"], {k: (r[k]["iterations"], round(r[k]["ms"], 2), r[k]["fallback"])
                                    for k in ("fp64_jacobi", "mixed", "learned")})
        for k in ("fp64_jacobi", "mixed", "learned"):
            assert r[k]["converged"] == 1
    s = learn.summarize(rows)
    say("summary", s)
    # What the network gains: the learned start lowers the true initial
    # residual below the zero start's (which is ||b||).
    ratios = [r["learned"]["initial_residual"] / r["mixed"]["initial_residual"] for r in rows]
    say("initial residual learned / zero start", ratios)
    assert np.mean(ratios) < 0.9
    # The final answers meet the same tolerance: check one on the host.
    learn.prepare(pipe, "bulk_lattice", 30000, 101)
    st = pipe.solve(L.SOLVE_LEARNED, 20000, 1e-8)
    A = pipe.ppe()
    x = pipe.solution()
    m = len(A["b"])
    ax = np.zeros(m)
    np.add.at(ax, np.repeat(np.arange(m), np.diff(A["row"])), A["val"] * x[A["col"]])
    assert np.linalg.norm(ax - A["b"]) <= 1e-8 * max(1.0, np.linalg.norm(A["b"])) * 1.0001
    assert st["converged"] == 1


def test_learned_fallback_when_guess_raises_residual(pipe):
    rng = np.random.default_rng(5)
    p = learn.init_params(1)
    p[5248:5312] = rng.normal(0, 50.0, 64)  # a wild initial-guess output
    pipe.set_params(p)
    learn.prepare(pipe, "bulk_lattice", 20000, 77)
    st = pipe.solve(L.SOLVE_LEARNED, 20000, 1e-8)
    say("fallback", st)
    assert st["fallback"] == 1 and st["converged"] == 1
