"""Host side of the step timeline: the gap scene and the droplet sites.

Droplets take each site of the open layer at most once, and with the
prescribed per-layer velocities no two particles (film, wall or droplet)
come closer than 0.6 l0 over a run; the smallest distance is found by
brute force over a k-d tree.
"""
import numpy as np
from scipy.spatial import cKDTree

from lsmps_amd import timeline as T


def test_gap_scene_layers():
    l0 = 0.01
    xyz, types, layer, lx, ly = T.gap_scene(4000, l0, 3)
    side = int(np.ceil(np.sqrt(4000 / 4.0)))
    assert xyz.shape == (4 * side * side, 3) and lx == ly == side * l0
    assert set(np.unique(layer)) == {0, 1, 2, 4}
    assert np.all((types == T.WALL) == np.isin(layer, (0, 4)))
    assert np.abs(xyz[:, 2] - layer * l0).max() <= 0.02 * l0
    assert xyz[:, 0].min() >= 0.0 and xyz[:, 0].max() < lx


def test_droplet_sites_unique_and_spaced():
    l0, speed, rebuild = 0.01, 1.0, 10
    xyz, _, layer, lx, ly = T.gap_scene(6000, l0, 5)
    sites = T.DropletSites(lx, ly, l0, speed, 5)
    dt = 0.5 * 0.2 * 2.1 * l0 / (speed * rebuild * 1.2)
    pos = xyz.copy()
    vel = np.zeros_like(pos)
    vel[:, 0] = speed * np.asarray(T.LAYER_SPEED)[layer]
    handed = 0
    for step in range(40):
        d = sites.take(50, step * dt)
        handed += d.shape[0]
        assert np.all(d[:, 0] < lx)
        pos = np.vstack([pos, d])
        vel = np.vstack([vel, np.tile([T.LAYER_SPEED[T.DROPLET_LAYER] * speed, 0.0, 0.0], (d.shape[0], 1))])
        pos += vel * dt
        if step % rebuild == rebuild - 1:
            keep = pos[:, 0] < lx
            pos, vel = pos[keep], vel[keep]
            dist, _ = cKDTree(pos).query(pos, k=2)
            assert dist[:, 1].min() >= 0.6 * l0
    # Each site once: the handed-out count never exceeds the sites passed.
    assert handed <= sites.next <= sites.sx * sites.sy
    assert handed == 40 * 50
