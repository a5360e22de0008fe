"""CPU emulation of the index arithmetic of the GPU cell sort and neighbor
pass (cnl.hip: k_cell_coords, k_cell_key, the scan, k_scatter, k_cell_rank,
k_gather_sorted, the run bounds and the per-wavefront staging of
k_cnl_pass), on the time-step sequence that builds skin lists, moves the
particles and builds again. It checks every index the kernels form against
the buffer it addresses and the assumptions the staging relies on:

  * sum of cell counts = n, cell starts non-decreasing, start[ncells] = n;
  * every run bound read (start[c + row offset]) lies in [0, ncells];
  * every run lies in [0, n];
  * within a wavefront of 8 consecutive sorted particles, the non-empty runs
    of each row start and end in lane order (the kernel stages their union
    [first start, last end) in one load);
  * staged loads (u0 + lane) stay below n + CNL_PAD, LDS stage reads below
    2 x 64, direct loads below n + CNL_PAD;
  * the unsort of the previous build's types reads a valid permutation;
  * every particle's candidate slots cover exactly its 27 cells.

It runs the kernels' integer math in numpy, so it says nothing about timing,
only that no index leaves its buffer and no loop bound is wrong for this
sequence."""
import numpy as np
import pytest

from lsmps_amd import timeline

CNL_PAD, WAVE, GROUP = 64, 64, 8
XLO, XHI, YLO, YHI, ZLO, ZHI = 1, 2, 4, 8, 16, 32


def build(x, rl):
    """Cell sort of k_cell_coords / k_cell_key / scan / k_scatter / k_cell_rank
    on the exact box; returns the sorted order and the grid."""
    n = len(x)
    c3 = np.floor(x / rl).astype(np.int64)  # (int)floor(x / re), per axis
    lo, hi = c3.min(0), c3.max(0)
    g = hi - lo + 1
    u, v, w = (c3 - lo).T
    cell = (w * g[1] + v) * g[0] + u
    ncells = int(np.prod(g))
    count = np.bincount(cell, minlength=ncells)
    start = np.zeros(ncells + 1, np.int64)
    start[1:] = np.cumsum(count)
    assert count.sum() == n and start[-1] == n and np.all(np.diff(start) >= 0)
    perm = np.argsort(cell, kind="stable")  # k_cell_rank: ascending index inside a cell
    iperm = np.empty(n, np.int64)
    iperm[perm] = np.arange(n)
    skey = cell[perm]
    assert np.all(np.diff(skey) >= 0)
    su, sv, sw = u[perm], v[perm], w[perm]
    flags = ((su > 0) * XLO | (su < g[0] - 1) * XHI | (sv > 0) * YLO | (sv < g[1] - 1) * YHI
             | (sw > 0) * ZLO | (sw < g[2] - 1) * ZHI)
    return dict(n=n, g=g, start=start, ncells=ncells, perm=perm, iperm=iperm, skey=skey, flags=flags, cell=cell,
                c3=c3)


def check_pass(B):
    """Run bounds and staging of k_cnl_pass for every wavefront."""
    n, g, start, ncells = B["n"], B["g"], B["start"], B["ncells"]
    gx, gy = int(g[0]), int(g[1])
    nw = (n + GROUP - 1) // GROUP
    cover = np.zeros(n, np.int64)  # candidates per particle
    for r in range(9):
        dz, dy = r // 3 - 1, r % 3 - 1
        need = (ZLO if dz < 0 else ZHI if dz > 0 else 0) | (YLO if dy < 0 else YHI if dy > 0 else 0)
        s = np.arange(nw * GROUP)
        live = s < n
        sl = np.minimum(s, n - 1)
        c = B["skey"][sl]
        fl = B["flags"][sl]
        ok = live & ((fl & need) == need)
        clo = c - ((fl & XLO) != 0)
        chi = c + np.where((fl & XHI) != 0, 2, 1)
        off = (dz * gy + dy) * gx
        ilo, ihi = clo + off, chi + off
        assert np.all(ilo[ok] >= 0) and np.all(ihi[ok] <= ncells), "run bound index outside cstart"
        b = np.where(ok, start[np.clip(ilo, 0, ncells)], 0)
        e = np.where(ok, start[np.clip(ihi, 0, ncells)], 0)
        assert np.all((b >= 0) & (b <= e) & (e <= n))
        cover[sl[live]] += (e - b)[live]
        # Wavefront staging: 8 particles (groups) per wavefront.
        bw, ew = b.reshape(nw, GROUP), e.reshape(nw, GROUP)
        ne = bw < ew
        for k in range(nw):
            idx = np.flatnonzero(ne[k])
            if idx.size == 0:
                continue
            bb, ee = bw[k, idx], ew[k, idx]
            assert np.all(np.diff(bb) >= 0) and np.all(np.diff(ee) >= 0), "runs out of lane order"
            u0, u1 = bb[0], ee[-1]
            assert np.all(bb >= u0) and np.all(ee <= u1)
            if u1 - u0 <= WAVE:
                assert u0 + WAVE - 1 < n + CNL_PAD, "staged load past the padding"
                # lane g reads stage[(bq - u0) + g + 8 j] while any lane's run continues
                rem = ee - bb
                jmax = int(np.ceil(rem.max() / GROUP))
                assert (bb - u0).max() + GROUP - 1 + GROUP * (jmax - 1) < 2 * WAVE, "LDS stage read out of range"
            else:
                assert (ee + GROUP - 1).max() < n + CNL_PAD + GROUP
    return cover


def brute_cover(B):
    """Particles in the 27 cells around each particle's cell."""
    c3 = B["c3"]
    key = {}
    for i, t in enumerate(map(tuple, c3)):
        key[t] = key.get(t, 0) + 1
    out = np.zeros(len(c3), np.int64)
    for i, t in enumerate(map(tuple, c3)):
        tot = 0
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    tot += key.get((t[0] + dx, t[1] + dy, t[2] + dz), 0)
        out[i] = tot
    return out


@pytest.mark.parametrize("particles", [4000, 40000])
def test_skin_then_plain_build_indices(particles):
    l0 = 0.01
    re = 2.1 * l0
    skin = 0.2 * re
    xyz, t, layer, lx, ly = timeline.gap_scene(particles, l0, 5)
    u = np.zeros_like(xyz)
    u[:, 0] = np.asarray(timeline.LAYER_SPEED)[layer]
    dt = 0.5 * skin / (3 * 1.2)
    A = build(xyz, re + skin)
    cover = check_pass(A)
    assert np.array_equal(cover[A["iperm"]], brute_cover(A)) if particles <= 4000 else True
    x = xyz.copy()
    for _ in range(3):
        x += u * dt
    B = build(x, re)
    # The second build brings the first build's sorted types back to caller
    # order through the first build's inverse permutation.
    assert np.array_equal(np.sort(A["iperm"]), np.arange(len(x)))
    cover = check_pass(B)
    if particles <= 4000:
        assert np.array_equal(cover[B["iperm"]], brute_cover(B))


@pytest.mark.parametrize("case", ["wall_film", "industrial_mix"])
def test_skin_then_fresh_build_indices_random_motion(case):
    """The sequences of the two earlier skin tests: 20k particles, skin
    lists, three steps of random motion below skin / 2, then a plain build."""
    import lsmps_amd as L

    l0 = 0.01
    re = 2.1 * l0
    skin = 0.2 * re
    xyz, t = L.generate_case(case, 20000, 3, l0)
    rng = np.random.default_rng(4)
    u = rng.uniform(-1.0, 1.0, xyz.shape)
    dt = skin / (2.0 * 3 * 2.0 * np.sqrt(3.0)) * 0.99
    A = build(xyz, re + skin)
    check_pass(A)
    x = xyz.copy()
    for _ in range(3):
        x += u * dt
    check_pass(build(x, re))
