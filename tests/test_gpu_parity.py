"""GPU parity of the HIP path against the reference's golden vectors and
the CPU oracle. Every call goes through the C ABI (lsmps_amd -> liblsmps_hip.so).

Bar: bit-exact for neighbor lists, types, PPE structure/values/b and the FP64
WLS operator; BiCGStab by tolerance. BiCGStab's iteration count and final
error on these pinned graph Laplacians are chaotic in the last bits of the
dot products, and the GPU sums them in parallel (with fused multiply-adds).
tests/golden/solver_bands.json holds the spread of the reference's own
algorithm when every dot product is perturbed by one rounding (40 seeded
trials, make_solver_bands.py): e.g. 306..526 iterations for rain_2000 whose
unperturbed count is 356. The GPU must converge, land in that iteration
range (15% margin) and have a relative L2 error against x* no larger than
1.5x the band's largest.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest
from conftest import GOLDEN, SMALL_FIXTURES, load_fixture

import lsmps_amd as L
import oracle_py as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def pipe():
    p = L.Pipeline(0)
    yield p
    p.close()


BANDS = json.load(open(os.path.join(GOLDEN, "solver_bands.json")))


def check_solve(st, ref, what):
    assert st["converged"] == ref["converged"], (what, st, ref)
    if ref["converged"]:
        b = BANDS[what]
        it = st["iterations"]
        assert 0.85 * b["iter_min"] - 3 <= it <= 1.15 * b["iter_max"] + 3, (what, it, b)
        assert st["relative_l2_error"] <= 1.5 * b["rel_l2_max"], (what, st["relative_l2_error"], b)


@pytest.mark.parametrize("name", SMALL_FIXTURES)
def test_small_fixture_bitexact(pipe, name):
    d = load_fixture(name)
    m = d["meta"]
    re = m["re_ratio"] * m["l0"]
    say("fixture", name, "n", len(d["xyz"]))
    pipe.upload(d["xyz"], d["type_in"])
    cs = pipe.neighbors(re, m["max_neighbors"])
    row, col = pipe.neighbor_lists()
    assert np.array_equal(row, d["nbr_row"]) and np.array_equal(col, d["nbr_col"])
    assert cs["truncated"] == m["truncated"]
    pipe.classify(m["l0"])
    assert np.array_equal(pipe.types(), d["type_out"])
    pipe.set_field(L.REFERENCE_FIELD)
    pipe.wls(re, m["l0"], L.WLS_FP64)
    g, lap, st = pipe.wls_results()
    assert np.array_equal(st, d["wls_status"])
    assert np.array_equal(g, d["wls_grad"]) and np.array_equal(lap, d["wls_lap"])
    ws = pipe.wls_sampled_stats(m["wls_samples"])
    for k, v in m["wls"].items():
        assert ws[k] == v, (k, ws[k], v)
    pr = pipe.assemble(re, m["max_unknowns"])
    for k, v in m["ppe"].items():
        assert pr[k] == v, k
    A = pipe.ppe()
    for k in ("row", "col", "val", "b", "x_exact"):
        assert np.array_equal(A[k], d["ppe_" + k]), k
    if pr["unknowns"]:
        st = pipe.solve(L.SOLVE_PLAIN, m["max_iter"], m["tol"])
        check_solve(st, m["solve"], name)


@pytest.mark.parametrize("name", ["film_trunc_1000", "rain_2000", "edge"])
def test_fused_neighbors_classify(pipe, name):
    d = load_fixture(name)
    m = d["meta"]
    re = m["re_ratio"] * m["l0"]
    pipe.upload(d["xyz"], d["type_in"])
    cs, ts = pipe.neighbors_classify(re, m["max_neighbors"], m["l0"])
    row, col = pipe.neighbor_lists()
    assert np.array_equal(row, d["nbr_row"]) and np.array_equal(col, d["nbr_col"])
    assert cs["truncated"] == m["truncated"]
    assert np.array_equal(pipe.types(), d["type_out"])
    assert ts["count"] == np.bincount(d["type_out"], minlength=5).tolist()


def test_reference_configs(pipe, ref_configs):
    for name, rec in ref_configs.items():
        c = rec["config"]
        l0 = float(c["l0"])
        re = 2.1 * l0
        say("config", name)
        xyz, t = L.generate_case(c["case"], int(c["particles"]), int(c["seed"]), l0, int(c.get("film_layers", 4)))
        pipe.upload(xyz, t)
        cs, ts = pipe.neighbors_classify(re, int(c["max_neighbors"]), l0)
        assert cs["pairs"] == rec["pairs"] and cs["truncated"] == rec["truncated"]
        row, col = pipe.neighbor_lists()
        assert sha(row) == rec["sha"]["nbr_row"] and sha(col) == rec["sha"]["nbr_col"], name
        assert np.bincount(np.diff(row)).tolist() == rec["neighbor_count_hist"]
        assert ts["count"] == rec["type_counts"]
        assert sha(pipe.types()) == rec["sha"]["type_out"]
        pipe.set_field(L.REFERENCE_FIELD)
        pipe.wls(re, l0, L.WLS_FP64)
        ws = pipe.wls_sampled_stats(int(c["wls_samples"]))
        for k, v in rec["wls"].items():
            assert ws[k] == v, (name, k, ws[k], v)
        pr = pipe.assemble(re, int(c["ppe_max_unknowns"]))
        for k, v in rec["ppe"].items():
            assert pr[k] == v, (name, k)
        A = pipe.ppe()
        for k in ("row", "col", "val", "b", "x_exact"):
            assert sha(A[k]) == rec["sha"]["ppe_" + k], (name, k)
        st = pipe.solve(L.SOLVE_PLAIN, int(c["ppe_max_iter"]), float(c["ppe_tol"]))
        check_solve(st, rec["solve"], "config:" + name)


def test_neighbors_against_oracle_random_sets(pipe):
    rng = np.random.default_rng(11)
    for n, maxn, spread in [(1, 8, 0.1), (2, 8, 0.001), (700, 96, 0.1), (3000, 10, 0.12), (5000, 40, 0.3)]:
        xyz = rng.uniform(-spread, spread, (n, 3))
        if n > 100:
            xyz[:20] = xyz[20:40]  # coincident pairs
        t = rng.integers(0, 5, n).astype(np.uint8)
        say("random set", n, maxn)
        pipe.upload(xyz, t)
        cs = pipe.neighbors(0.021, maxn)
        row, col = pipe.neighbor_lists()
        o = O.Oracle(xyz, t)
        pairs, trunc = o.neighbors(0.021, maxn)
        orow, ocol = o.get_neighbors()
        assert np.array_equal(row, orow) and np.array_equal(col, ocol)
        assert cs["truncated"] == trunc
        pipe.classify(0.01)
        o.classify(0.01)
        assert np.array_equal(pipe.types(), o.types())


def test_empty_input(pipe):
    pipe.upload(np.zeros((0, 3)), np.zeros(0, np.uint8))
    cs = pipe.neighbors(0.021, 96)
    assert cs["pairs"] == 0
    pipe.classify(0.01)
    pr = pipe.assemble(0.021, 100)
    assert pr["unknowns"] == 0 and pr["nnz"] == 0


def test_wls_other_field_and_mixed(pipe):
    """FP64 operator reproduces the C1 quadratic; FP32 mixed mode close."""
    xyz, t = L.generate_case("bulk_lattice", 20000, 3, 0.01)
    say("wls fields", len(xyz))
    pipe.upload(xyz, t)
    pipe.neighbors(0.021, 96)
    pipe.classify(0.01)
    pipe.set_field(L.C1_FIELD)
    r64 = pipe.wls(0.021, 0.01, L.WLS_FP64)
    g64, l64, s64 = pipe.wls_results()
    ok = s64 <= 1
    ge = L.cases.field_gradient(xyz, L.C1_FIELD)
    assert np.sqrt(np.mean((g64[ok] - ge[ok]) ** 2)) < 1e-10
    assert np.sqrt(np.mean((l64[ok] - 4.0) ** 2)) < 1e-8
    # Oracle agreement (bit-exact FP64).
    o = O.Oracle(xyz, t)
    o.neighbors(0.021, 96)
    o.classify(0.01)
    og, ol, ost = o.wls_all(L.cases.field_values(xyz, L.C1_FIELD), 0.021, 0.01)
    f_dev = L.cases.field_values(xyz, L.C1_FIELD)
    assert np.array_equal(ost, s64)
    assert np.array_equal(og, g64) and np.array_equal(ol, l64), "FP64 WLS must be bit-exact"
    assert f_dev.shape[0] == len(xyz)
    r32 = pipe.wls(0.021, 0.01, L.WLS_MIXED)
    g32, l32, s32 = pipe.wls_results()
    say("mixed wls", r32)
    assert np.array_equal(s32 == L.WLS_SKIPPED, s64 == L.WLS_SKIPPED)
    okm = ok & (s32 <= 1)
    # FP32 moments: relative accuracy ~1e-5 of the field scale.
    assert np.sqrt(np.mean((g32[okm] - ge[okm]) ** 2)) < 1e-3
    assert np.sqrt(np.mean((l32[okm] - 4.0) ** 2)) < 5e-2


def test_fp64_wls_tiny_quotients_take_plain_divisions(pipe):
    """The FP64 WLS forms its quotients from one reciprocal per divisor and
    sends a row whose quotient is too small for that (|q| < 2^-900) to the
    FP64 redo kernel, which divides. A field of 1e-300 times the C1 values
    on half the block (x below the median) puts every right-hand-side
    quotient of those rows there: those rows are redone, the others are
    not, and every row equals the oracle's bit for bit."""
    xyz, t = L.generate_case("bulk_lattice", 20000, 3, 0.01)
    vals = L.cases.field_values(xyz, L.C1_FIELD)
    tiny = xyz[:, 0] < np.median(xyz[:, 0])
    vals = np.where(tiny, vals * 1e-300, vals)
    say("tiny quotients", len(xyz), int(tiny.sum()))
    pipe.upload(xyz, t)
    pipe.neighbors(0.021, 96)
    pipe.classify(0.01)
    pipe.set_field(vals)
    r64 = pipe.wls(0.021, 0.01, L.WLS_FP64)
    g64, l64, s64 = pipe.wls_results()
    o = O.Oracle(xyz, t)
    o.neighbors(0.021, 96)
    o.classify(0.01)
    og, ol, ost = o.wls_all(vals, 0.021, 0.01)
    say("redone", r64["fp64_redo"])
    assert np.array_equal(ost, s64)
    assert np.array_equal(og, g64) and np.array_equal(ol, l64)
    solved = s64 <= L.WLS_REGULARIZED
    assert int((solved & tiny).sum()) > 1000
    assert int((solved & tiny).sum()) <= r64["fp64_redo"] < int(solved.sum())
    deep = xyz[:, 0] < np.median(xyz[:, 0]) - 0.05  # neighbors all in the scaled half
    assert np.all(np.abs(l64[solved & deep]) < 1e-290)


def test_mixed_wls_redo_rows_equal_fp64(pipe):
    """Rows the mixed WLS redoes in FP64 (eight lanes per row, the five
    regularisation steps side by side) carry exactly the FP64 kernel's
    results: every row that needed the ladder or failed in FP64, and at
    least as many rows as were redone, match bit for bit."""
    l0 = 0.004
    re = 2.1 * l0
    xyz, t = L.generate_case("rain_plate", 30000, 5, l0)
    say("mixed redo rows", len(xyz))
    pipe.upload(xyz, t)
    pipe.neighbors_classify(re, 128, l0)
    pipe.set_field(L.C1_FIELD)
    pipe.wls(re, l0, L.WLS_FP64)
    g64, l64, s64 = pipe.wls_results()
    r32 = pipe.wls(re, l0, L.WLS_MIXED)
    g32, l32, s32 = pipe.wls_results()
    say("mixed wls", r32)
    same = (s32 == s64) & np.all(g32 == g64, axis=1) & (l32 == l64)
    assert r32["fp64_redo"] > 0
    assert int(same.sum()) >= r32["fp64_redo"]
    ladder = (s64 == L.WLS_REGULARIZED) | (s64 == L.WLS_FAILED) | (s64 == L.WLS_NO_WEIGHT)
    say("ladder rows", int(ladder.sum()), "redone", r32["fp64_redo"])
    assert np.all(same[ladder])
    # A pivot-ratio threshold of 1 sends every row to FP64: the FP64 results.
    r1 = pipe.wls(re, l0, L.WLS_MIXED, fp32_pivot_min=1.0)
    g1, l1, s1 = pipe.wls_results()
    assert r1["fp64_redo"] == r1["ok"] + r1["regularized"] + r1["failed"]
    assert np.array_equal(s1, s64) and np.array_equal(g1, g64) and np.array_equal(l1, l64)
    for bad in (-0.1, 1.5, float("nan")):
        with pytest.raises(L.LsmpsError):
            pipe.wls(re, l0, L.WLS_MIXED, fp32_pivot_min=bad)


def test_timing_flags_leave_results_unchanged():
    """lsmps_set_timing only decides which HIP events the neighbor builds and
    WLS stages record: the same lists, types and WLS results with events off;
    a field whose events are off reads 0, the others are timed."""
    l0 = 0.01
    re = 2.1 * l0
    xyz, t = L.generate_case("wall_film", 8000, 4, l0)
    say("timing flags", len(xyz))
    out = {}
    for flags in (L.TIMING_STAGES | L.TIMING_KERNELS, 0, L.TIMING_KERNELS, L.TIMING_STAGES):
        p = L.Pipeline(0)
        try:
            p.upload(xyz, t)
            p.set_field(L.C1_FIELD)
            p.set_timing(flags)
            cs, _, wr = p.step(re, 96, l0, L.WLS_MIXED)
            cs2, _ = p.neighbors_classify(re, 96, l0)  # kept grid, the separate calls
            wr2 = p.wls(re, l0, L.WLS_FP64)
            for c in (cs, cs2):
                assert (c["ms"] > 0) == bool(flags & L.TIMING_STAGES)
                assert (c["kernel_ms"] > 0) == bool(flags & L.TIMING_KERNELS)
            for w in (wr, wr2):
                assert (w["ms"] > 0) == bool(flags & L.TIMING_STAGES)
                assert (w["kernel_ms"] > 0) == bool(flags & L.TIMING_KERNELS)
            out[flags] = (p.neighbor_lists(), p.types(), p.wls_results())
            with pytest.raises(L.LsmpsError):
                p.set_timing(4)
        finally:
            p.close()
    ref = out[L.TIMING_STAGES | L.TIMING_KERNELS]
    for flags, (lists, types, wls) in out.items():
        assert all(np.array_equal(a, b) for a, b in zip(lists, ref[0]))
        assert np.array_equal(types, ref[1])
        assert all(a.tobytes() == b.tobytes() for a, b in zip(wls, ref[2]))


def test_mixed_wls_scale_and_range():
    """The mixed WLS works in the caller's length unit: the same scene scaled
    by 2^8 reproduces the quadratic field x^2 + 2y - z as accurately as the
    unscaled one (gradient (2x, 2, -1), Laplacian 2, in each unit). With a
    radius beyond 1e12 (scene scaled by 2^48) every row is computed in FP64
    and the results are the FP64 kernel's bit for bit."""
    l0 = 0.004
    re = 2.1 * l0
    xyz, t = L.generate_case("rain_plate", 20000, 5, l0)
    say("mixed WLS scale", len(xyz))

    def run(k):
        p = L.Pipeline(0)
        try:
            s = 2.0 ** k
            pts = xyz * s
            p.upload(pts, t)
            p.neighbors_classify(re * s, 128, l0 * s)
            p.set_field(pts[:, 0] ** 2 + 2.0 * pts[:, 1] - pts[:, 2])
            r = p.wls(re * s, l0 * s, L.WLS_MIXED)
            out = p.wls_results()
            p.wls(re * s, l0 * s, L.WLS_FP64)
            return pts, r, out, p.wls_results()
        finally:
            p.close()

    errs = []
    for k in (0, 8):
        pts, r, (g, lap, st), _ = run(k)
        ok = st == L.WLS_OK
        assert ok.sum() > len(xyz) // 2
        ge = np.stack([2.0 * pts[:, 0], np.full(len(pts), 2.0), np.full(len(pts), -1.0)], 1)
        eg = np.sqrt(np.mean(np.sum((g[ok] - ge[ok]) ** 2, 1) / np.sum(ge[ok] ** 2, 1)))
        el = np.sqrt(np.mean((lap[ok] - 2.0) ** 2))
        say("scale 2^%d" % k, "FP64 rows", r["fp64_redo"], "rms gradient (relative)", eg, "rms Laplacian", el)
        assert eg < 1e-3 and el < 5e-2
        errs.append((eg, el, r["fp64_redo"]))
    assert abs(errs[0][2] - errs[1][2]) <= max(5, errs[0][2] // 100)
    _, rbig, (gb, lb, sb), (g64, l64, s64) = run(48)
    assert rbig["fp64_redo"] == rbig["ok"] + rbig["regularized"] + rbig["failed"]
    assert np.array_equal(sb, s64) and np.array_equal(gb, g64) and np.array_equal(lb, l64)


def test_kept_grid_reused_and_outgrown(pipe):
    """A build reuses the last build's grid (box plus a margin) without a
    round trip for the box; lists and types match a fresh context's whether
    the particles stay inside that grid or some leave it (the build is then
    redone on the exact box)."""
    l0 = 0.01
    re = 2.1 * l0
    xyz, t = L.generate_case("wall_film", 6000, 3, l0)
    say("kept grid", len(xyz))

    def fresh(points, types):
        q = L.Pipeline(0)
        q.upload(points, types)
        q.neighbors_classify(re, 96, l0)
        out = q.neighbor_lists(), q.types()
        q.close()
        return out

    pipe.upload(xyz, t)
    pipe.neighbors_classify(re, 96, l0)
    (r0, c0), t0 = fresh(xyz, t)
    # Same particles again: the kept grid holds them, and the build reads
    # them in the last build's sorted order.
    pipe.neighbors_classify(re, 96, l0)
    r1, c1 = pipe.neighbor_lists()
    assert np.array_equal(r0, r1) and np.array_equal(c0, c1)
    assert np.array_equal(t0, pipe.types())
    # Moved by lsmps_advance, then built again in that order: the same
    # lists, types and FP64 WLS (bit for bit) as a fresh context given the
    # positions moved on the host by the same x += u dt.
    u = np.random.default_rng(8).uniform(-1.0, 1.0, xyz.shape)
    dt = 0.3 * l0
    pipe.set_velocities(u)
    pipe.advance(dt)
    moved = xyz + u * dt
    pipe.neighbors_classify(re, 96, l0)
    pipe.set_field(L.C1_FIELD)
    pipe.wls(re, l0, L.WLS_FP64)
    wm = pipe.wls_results()
    q = L.Pipeline(0)
    q.upload(moved, t)
    q.neighbors_classify(re, 96, l0)
    rq, cq = q.neighbor_lists()
    tq = q.types()
    q.set_field(L.C1_FIELD)
    q.wls(re, l0, L.WLS_FP64)
    wq = q.wls_results()
    q.close()
    r4, c4 = pipe.neighbor_lists()
    assert np.array_equal(rq, r4) and np.array_equal(cq, c4)
    assert np.array_equal(tq, pipe.types())
    for a, b in zip(wm, wq):
        assert np.array_equal(a, b)
    # Particles well outside the last box: the kept grid is outgrown.
    lo, hi = xyz.min(0), xyz.max(0)
    extra = lo + (hi - lo) * np.array([[1.0, 0.5, 0.5]]) + np.array([[12 * re, 0.0, 0.0]])
    extra = extra + l0 * np.arange(3)[:, None] * np.array([[1.0, 0.0, 0.0]])
    te = np.zeros(len(extra), np.uint8)
    # The context has velocities: new particles need theirs too.
    with pytest.raises(L.LsmpsError):
        pipe.append(extra, te)
    pipe.append(extra, te, np.zeros_like(extra))
    types_now = pipe.types()
    st = pipe.neighbors_classify(re, 96, l0)
    say("after append", st[0]["pairs"])
    both = np.concatenate([moved, extra])
    (r2, c2), t2 = fresh(both, types_now)
    r3, c3 = pipe.neighbor_lists()
    assert np.array_equal(r2, r3) and np.array_equal(c2, c3)
    assert np.array_equal(t2, pipe.types())


def test_step_matches_separate_calls():
    """lsmps_step queues the neighbor build and the WLS stage before one wait;
    it gives what lsmps_build_neighbors_classify followed by lsmps_wls gives,
    bit for bit (lists, types, statistics, WLS results): on a first build, on
    repeated builds (kept grid, last build's order), in both precisions, and
    when particles have left the kept grid (build and WLS redone on the exact
    box)."""
    l0 = 0.01
    re = 2.1 * l0
    xyz, t = L.generate_case("wall_film", 20000, 5, l0)
    say("step against separate calls", len(xyz))
    a, b = L.Pipeline(0), L.Pipeline(0)
    try:
        for q in (a, b):
            q.upload(xyz, t)
            q.set_field(L.C1_FIELD)

        def same(sa, sb):
            (ca, ta, wa), (cb, tb, wb) = sa, sb
            for k in ("pairs", "truncated", "mean", "p50", "p90", "p99", "max_count", "grid"):
                assert ca[k] == cb[k], k
            assert ta["count"] == tb["count"]
            for k in ("ok", "regularized", "failed", "skipped", "fp64_redo"):
                assert wa[k] == wb[k], k
            assert ca["ms"] > 0 and wa["ms"] > 0 and wa["kernel_ms"] > 0
            ra, ka = a.neighbor_lists()
            rb, kb = b.neighbor_lists()
            assert np.array_equal(ra, rb) and np.array_equal(ka, kb)
            assert np.array_equal(a.types(), b.types())
            for x, y in zip(a.wls_results(), b.wls_results()):
                assert x.tobytes() == y.tobytes()

        def both(prec):
            sa = a.step(re, 96, l0, prec)
            cb, tb = b.neighbors_classify(re, 96, l0)
            same(sa, (cb, tb, b.wls(re, l0, prec)))

        for prec in (L.WLS_FP64, L.WLS_MIXED, L.WLS_MIXED, L.WLS_FP64):
            both(prec)
        # A few particles pushed ten list radii along x: outside the kept grid.
        u = np.zeros_like(xyz)
        u[:5, 0] = 1.0
        for q in (a, b):
            q.set_velocities(u)
            q.advance(10 * re)
        both(L.WLS_MIXED)
        both(L.WLS_FP64)
        # A smaller WLS radius than the lists'.
        sa = a.step(re, 96, l0, L.WLS_FP64, wls_re=1.9 * l0)
        cb, tb = b.neighbors_classify(re, 96, l0)
        same(sa, (cb, tb, b.wls(1.9 * l0, l0, L.WLS_FP64)))
        # Checked before anything is queued.
        with pytest.raises(L.LsmpsError):
            a.step(re, 96, l0, L.WLS_FP64, wls_re=1.1 * re)
        with pytest.raises(L.LsmpsError):
            a.step(re, 96, l0, 7)
        c = L.Pipeline(0)
        c.upload(xyz, t)
        with pytest.raises(L.LsmpsError):
            c.step(re, 96, l0, L.WLS_FP64)  # no field
        c.close()
    finally:
        a.close()
        b.close()


def test_field_values_and_quadratic_paths_agree(pipe):
    """The field reaches the sorted records four ways: evaluated (quadratic)
    or gathered (values), by the neighbor build when the field is set before
    it, else by the WLS stage. The host values of the quadratic are the
    device's bit for bit, so all four FP64 WLS results are identical."""
    l0 = 0.01
    re = 2.1 * l0
    xyz, t = L.generate_case("wall_film", 8000, 3, l0)
    vals = L.cases.field_values(xyz, L.C1_FIELD)
    say("field paths", len(xyz))
    pipe.upload(xyz, t)
    pipe.set_field(vals)  # values, gathered by the build
    pipe.neighbors_classify(re, 96, l0)
    pipe.wls(re, l0, L.WLS_FP64)
    ref = pipe.wls_results()
    pipe.set_field(L.C1_FIELD)  # quadratic, evaluated by the WLS stage
    pipe.wls(re, l0, L.WLS_FP64)
    out = [pipe.wls_results()]
    pipe.set_field(vals)  # values, gathered by the WLS stage
    pipe.wls(re, l0, L.WLS_FP64)
    out.append(pipe.wls_results())
    pipe.set_field(L.C1_FIELD)  # quadratic, evaluated by the build
    pipe.neighbors_classify(re, 96, l0)
    pipe.wls(re, l0, L.WLS_FP64)
    out.append(pipe.wls_results())
    for g, lp, st in out:
        assert np.array_equal(g, ref[0]) and np.array_equal(lp, ref[1]) and np.array_equal(st, ref[2])
    assert int((ref[2] <= 1).sum()) > 0


def test_solvers_reach_tolerance_and_agree(pipe):
    xyz, t = L.generate_case("bulk_lattice", 50000, 7, 0.01)
    say("solvers C1", len(xyz))
    pipe.upload(xyz, t)
    pipe.neighbors(0.021, 96)
    pipe.classify(0.01)
    pr = pipe.assemble(0.021, 6000, L.XSTAR_RANDOM, 5)
    A = pipe.ppe()
    o = O.Oracle(xyz, t)
    o.neighbors(0.021, 96)
    o.classify(0.01)
    o.assemble(0.021, 6000, 1, 5)
    B = o.get_ppe()
    for k in ("row", "col", "val", "b", "x_exact"):
        assert np.array_equal(A[k], B[k]), k
    ost, ox = o.solve(20000, 1e-10, precond=1)
    bnorm = np.linalg.norm(A["b"])
    for method in (L.SOLVE_JACOBI, L.SOLVE_MIXED, L.SOLVE_PLAIN):
        st = pipe.solve(method, 20000, 1e-10)
        x = pipe.solution()
        say("method", method, st)
        assert st["converged"] == 1
        # True residual recomputed on the host in FP64.
        n = len(A["b"])
        ax = np.zeros(n)
        rows = np.repeat(np.arange(n), np.diff(A["row"]))
        np.add.at(ax, rows, A["val"] * x[A["col"]])
        res = np.linalg.norm(ax - A["b"])
        assert res <= 1e-10 * max(1.0, bnorm) * (1.0 if method != L.SOLVE_PLAIN else 10.0)
        # Two solutions that both meet the residual tolerance differ by at most
        # cond(A) * tol relative to |x|: cond ~1e5 for this pinned Laplacian and
        # tol 1e-10 give 1e-5. The Jacobi paths land near 1e-7; the plain path
        # stops on its recurrence residual (the reference's rule) near 1e-6.
        bound = 1e-5
        assert np.linalg.norm(x - ox) <= bound * np.linalg.norm(ox)
        if method == L.SOLVE_JACOBI:
            assert abs(st["iterations"] - ost["iterations"]) <= max(5, 0.15 * ost["iterations"])


def test_solver_max_iter_and_breakdown_paths(pipe):
    d = load_fixture("film_1200")
    m = d["meta"]
    re = m["re_ratio"] * m["l0"]
    pipe.upload(d["xyz"], d["type_in"])
    pipe.neighbors(re, m["max_neighbors"])
    pipe.classify(m["l0"])
    pipe.assemble(re, m["max_unknowns"])
    st = pipe.solve(L.SOLVE_PLAIN, 1, 1e-7)
    assert st["converged"] == 0 and st["iterations"] == 1
    st = pipe.solve(L.SOLVE_PLAIN, 0, 1e-7)
    assert st["converged"] == 0 and st["iterations"] == 0


def test_plain_solver_stops_on_half_step_residual(pipe):
    # One unknown: the identity pin row, A = [1]. The reference's BiCGStab
    # (src/lsmps.cpp:67) then has alpha = 1 and s = r - alpha A p = 0 in its
    # first iteration and stops on ||s||, with x = alpha p = b = x*, exactly.
    d = load_fixture("film_1200")
    m = d["meta"]
    re = m["re_ratio"] * m["l0"]
    pipe.upload(d["xyz"], d["type_in"])
    pipe.neighbors(re, m["max_neighbors"])
    pipe.classify(m["l0"])
    pipe.assemble(re, 1)
    A = pipe.ppe()
    assert len(A["b"]) == 1 and A["val"].tolist() == [1.0]
    st = pipe.solve(L.SOLVE_PLAIN, 100, 1e-7)
    say("one-row plain solve", st)
    assert st["converged"] == 1 and st["iterations"] == 1
    assert pipe.solution().tolist() == A["x_exact"].tolist()


def test_deterministic(pipe):
    xyz, t = L.generate_case("rain_injection", 30000, 2, 0.01)
    outs = []
    for _ in range(2):
        pipe.upload(xyz, t)
        pipe.neighbors(0.021, 96)
        pipe.classify(0.01)
        pipe.set_field(L.REFERENCE_FIELD)
        pipe.wls(0.021, 0.01, L.WLS_MIXED)
        g, lap, s = pipe.wls_results()
        pipe.assemble(0.021, 10 ** 9)
        st = pipe.solve(L.SOLVE_MIXED, 5000, 1e-8)
        outs.append((g.copy(), lap.copy(), st["iterations"], pipe.solution().copy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2] and np.array_equal(outs[0][3], outs[1][3])


def test_solver_without_graphs_gives_the_same_solves():
    """LSMPS_SOLVER_GRAPHS=0 (read at context creation) launches the solver's
    iterations directly instead of replaying captured graphs: the same
    kernels and arguments, so every method gives the same iterations and the
    same solution bit for bit."""
    xyz, t = L.generate_case("rain_injection", 30000, 2, 0.01)
    runs = []
    for graphs in ("1", "0"):
        os.environ["LSMPS_SOLVER_GRAPHS"] = graphs
        try:
            p = L.Pipeline(0)
        finally:
            del os.environ["LSMPS_SOLVER_GRAPHS"]
        try:
            p.upload(xyz, t)
            p.neighbors_classify(0.021, 96, 0.01)
            p.assemble(0.021, 10 ** 9, L.XSTAR_RANDOM, 5)
            out = []
            for method in (L.SOLVE_PLAIN, L.SOLVE_JACOBI, L.SOLVE_MIXED):
                st = p.solve(method, 5000, 1e-8, chunk=16)
                out.append((st["iterations"], st["converged"], st["final_residual"], p.solution().copy()))
            runs.append(out)
        finally:
            p.close()
    say("graphs vs direct:", [(a[0], b[0]) for a, b in zip(*runs)])
    for a, b in zip(*runs):
        assert a[:3] == b[:3] and np.array_equal(a[3], b[3])


def test_build_self_checks():
    """LSMPS_CHECK_BUILD=1: every neighbor build downloads its cell list and
    lists and checks their invariants on the host (cell starts an exclusive
    scan to n, each slot inside its cell, perm and iperm inverse, every list
    ascending in [0, n) without the row itself and within the list radius,
    lengths adding up to the pair count). Exercised on the build sequences
    the tests and the timeline run: a fused build, the same particles again
    (kept grid, sorted reads), motion, a skin build, appended particles,
    outflow removal, a one-call step, and a scene with far outliers
    (compressed grid)."""
    os.environ["LSMPS_CHECK_BUILD"] = "1"
    try:
        p = L.Pipeline(0)
    finally:
        del os.environ["LSMPS_CHECK_BUILD"]
    try:
        l0 = 0.01
        re = 2.1 * l0
        xyz, t = L.generate_case("wall_film", 20000, 3, l0)
        say("build checks", len(xyz))
        p.upload(xyz, t)
        p.neighbors_classify(re, 96, l0)
        p.neighbors_classify(re, 96, l0)  # kept grid, sorted reads
        rng = np.random.default_rng(2)
        u = rng.uniform(-1.0, 1.0, xyz.shape)
        p.set_velocities(u)
        p.advance(0.1 * l0)
        p.neighbors_classify(re, 96, l0)  # moved particles on the kept grid
        p.advance(0.1 * l0)
        p.neighbors_skin(re, 0.2 * re, 128, l0, True)
        p.append(xyz[:500] + np.array([0.0, 0.0, 0.5 * l0]), t[:500], u[:500])
        p.neighbors_skin(re, 0.2 * re, 128, l0, True)
        lo, hi = xyz.min(axis=0), xyz.max(axis=0)
        p.remove_outside(lo, hi - np.array([0.1, 0.0, 0.0]))
        p.neighbors_classify(re, 96, l0)
        p.set_field(L.C1_FIELD)
        p.step(re, 96, l0, L.WLS_MIXED)  # the one-call step (checks after its WLS stage)
        far = np.concatenate([xyz, xyz[:1] + np.array([[1e6 * re, 0.0, 0.0]])])
        p.upload(far, np.concatenate([t, t[:1]]))
        p.neighbors_classify(re, 96, l0)  # compressed grid
    finally:
        p.close()


def test_far_outliers_take_the_compressed_grid(pipe):
    """Particles 10^6 re and 3-4 x 10^5 re away from the rest (a box far too
    large for the dense grid): the build compresses the occupied cell
    coordinates per axis and still gives the oracle's lists and types (the
    reference hashes cells, so any extent works), and the FP64 WLS and the
    PPE stay bit-exact; a second build on the same particles agrees. A cloud
    whose cells are scattered on every axis is refused with LSMPS_E_LIMIT."""
    l0 = 0.01
    re = 2.1 * l0
    xyz, t = L.generate_case("wall_film", 6000, 3, l0)
    far = np.array([[1e6 * re, 0.0, 0.0], [-3e5 * re, 2e5 * re, 4e5 * re], [-3e5 * re, 2e5 * re, 4e5 * re + 0.5 * re]])
    pts = np.concatenate([xyz, xyz[:1] + far])
    ty = np.concatenate([t, np.zeros(3, np.uint8)])
    say("far outliers", len(pts))
    o = O.Oracle(pts, ty)
    o.neighbors(re, 96)
    orow, ocol = o.get_neighbors()
    o.classify(l0)
    og, ol, ost = o.wls_all(L.cases.field_values(pts, L.C1_FIELD), re, l0)
    o.assemble(re, 1 << 30, 1, 3)
    B = o.get_ppe()
    pipe.upload(pts, ty)
    for _ in range(2):
        cs, _ = pipe.neighbors_classify(re, 96, l0)
        row, col = pipe.neighbor_lists()
        assert np.array_equal(row, orow) and np.array_equal(col, ocol)
        assert np.array_equal(pipe.types(), o.types())
    # The last two outliers are neighbors of each other only.
    assert sorted(col[row[-2]:row[-1]].tolist()) == [len(pts) - 2]
    pipe.set_field(L.C1_FIELD)
    pipe.wls(re, l0, L.WLS_FP64)
    g, lap, st = pipe.wls_results()
    assert np.array_equal(st, ost) and np.array_equal(g, og) and np.array_equal(lap, ol)
    pipe.assemble(re, 1 << 30, L.XSTAR_RANDOM, 3)
    A = pipe.ppe()
    for k in ("row", "col", "val", "b", "x_exact", "rev"):
        assert np.array_equal(A[k], B[k]), k
    rng = np.random.default_rng(2)
    cloud = rng.uniform(0.0, 5e4 * re, (3000, 3))
    pipe.upload(cloud, np.zeros(len(cloud), np.uint8))
    with pytest.raises(L.LsmpsError) as e:
        pipe.neighbors(re, 96)
    assert "(-4)" in str(e.value)
