"""Step timeline on the GPU: the narrow-gap scene with rain injection.

The reference models a time-stepped run in its Python tooling
(python/rain_plate_lsmps_cpu.py:311-381: inject, neighbors, classify, WLS
operators, PPE, move, remove outflow, every step; the injection of
python/rain_plate_cpu.py:67-87 and the outflow test of :146-154; the
neighbor-rebuild cadence of python/workload_calibration.py:266-280). This
driver runs that loop through the GPU pipeline, with the pressure solve of
every step and the neighbor lists rebuilt on a cadence:

  scene     a gap between two Wall layers 4 l0 apart (z = 0 and z = 4 l0),
            jittered lattice at spacing l0: a fluid film on the layers
            z = l0 and z = 2 l0, the layer z = 3 l0 open for the droplets;
            about 1.5M particles by default;
  motion    prescribed plane-Poiseuille velocities along +x, one speed per
            layer (walls 0, film 0.75 U and U, droplet layer 0.75 U), so
            particles of one layer never approach each other; x += u dt;
  inject    `inject` droplets per step, each on a free site of a seeded
            permutation of the droplet layer's sites (spacing 0.75 l0,
            moving with that layer, jitter 0.05 l0). Uniformly random
            positions would put droplets almost on top of film particles:
            the PPE weight re / r - 1 then diverges and the solve needs
            tens of thousands of iterations; on sites, no two particles come
            closer than about 0.6 l0;
  outflow   particles with x >= the initial length leave at rebuild steps;
  cadence   neighbor lists (and classification, and outflow removal, and
            the entry of injected particles) every `rebuild` steps, with a
            skin of 0.2 re: the lists hold every pair within re + skin, and
            the stages use only pairs within re. dt keeps the relative
            motion over one cadence below skin / 2;
  per step  WLS (mixed precision) of a manufactured quadratic, PPE assembly
            over all fluid unknowns with a seeded pressure, and the solve to
            a true relative residual of 1e-8: by default the learned start
            and diagonal in the mixed solver, with the plain mixed method as
            the fallback (--method mixed / jacobi for the plain mixed or the
            FP64 Jacobi solver).

    python -m lsmps_amd.timeline --particles 1500000 --steps 200 [--method jacobi]

prints one JSON summary: p50 / p99 / max of the step wall time and of each
stage, particle counts and solver statistics.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import (C1_FIELD, INTERIOR, SOLVE_JACOBI, SOLVE_LEARNED, SOLVE_MIXED, WALL, WLS_MIXED, XSTAR_RANDOM,
               Pipeline)


def log(*a):
    print("[timeline]", *a, file=sys.stderr, flush=True)


def percentiles(v):
    v = np.asarray(v, np.float64)
    if v.size == 0:
        return {}
    return {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99)), "max": float(v.max()),
            "mean": float(v.mean())}


LAYER_SPEED = (0.0, 0.75, 1.0, 0.75, 0.0)  # x-velocity / U of the layers z = 0 .. 4 l0
DROPLET_LAYER = 3
SITE_SPACING = 0.75  # droplet sites, in l0


def gap_scene(particles, l0, seed):
    """Walls at z = 0 and 4 l0, film on z = l0 and 2 l0 (jitter 0.02 l0),
    the layer z = 3 l0 empty. Returns positions, types, per-layer index and
    the lengths (lx, ly) of the gap."""
    rng = np.random.default_rng(seed)
    side = max(4, int(np.ceil(np.sqrt(particles / 4.0))))
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    i, j = i.ravel(), j.ravel()
    xyz, types, layer = [], [], []
    for k in (0, 1, 2, 4):
        p = np.column_stack([(i + 0.5) * l0, (j + 0.5) * l0, np.full(i.size, k * l0)])
        p += rng.uniform(-0.02 * l0, 0.02 * l0, p.shape)
        xyz.append(p)
        types.append(np.full(i.size, WALL if k in (0, 4) else INTERIOR, np.uint8))
        layer.append(np.full(i.size, k, np.int8))
    return np.concatenate(xyz), np.concatenate(types), np.concatenate(layer), side * l0, side * l0


class DropletSites:
    """Sites of the droplet layer on a square lattice of spacing
    SITE_SPACING l0, fixed in the frame of that layer, handed out once each
    in a seeded order; a site that has flowed to within one spacing of the
    outflow is skipped for good."""

    def __init__(self, lx, ly, l0, speed, seed):
        h = SITE_SPACING * l0
        self.sx, self.sy = int(lx // h), int(ly // h)
        self.h, self.l0, self.lx = h, l0, lx
        self.u = LAYER_SPEED[DROPLET_LAYER] * speed
        self.rng = np.random.default_rng(seed + 7)
        self.order = self.rng.permutation(self.sx * self.sy)
        self.next = 0

    def take(self, count, t):
        """Up to `count` droplet positions at time t."""
        got = []
        while count > 0 and self.next < self.order.size:
            idx = self.order[self.next: self.next + max(count, 1024)]
            x = (idx // self.sy + 0.5) * self.h + self.u * t
            ok = np.flatnonzero(x < self.lx - self.h)
            if ok.size > count:
                ok = ok[:count]
                self.next += int(ok[-1]) + 1
            else:
                self.next += idx.size
            got.append(idx[ok])
            count -= ok.size
        idx = np.concatenate(got) if got else np.zeros(0, np.int64)
        p = np.column_stack([(idx // self.sy + 0.5) * self.h + self.u * t, (idx % self.sy + 0.5) * self.h,
                             np.full(idx.size, DROPLET_LAYER * self.l0)])
        return p + self.rng.uniform(-0.05 * self.l0, 0.05 * self.l0, p.shape)


def run(particles=1_500_000, steps=200, inject=2000, rebuild=10, l0=0.004, speed=1.0, seed=5, maxn=128,
        weights=None, tol=1e-8, max_iter=20000, pipe=None, verbose=True, method=None):
    re = 2.1 * l0
    skin = 0.2 * re
    # The largest relative speed of two particles within re + skin is U (the
    # middle film layer against a wall); over one cadence it moves them less
    # than skin / 2.
    dt = 0.5 * skin / (speed * rebuild * 1.2)
    own = pipe is None
    p = pipe or Pipeline(0)
    xyz, types, layer, lx, ly = gap_scene(particles, l0, seed)
    u = np.zeros_like(xyz)
    u[:, 0] = speed * np.asarray(LAYER_SPEED)[layer]
    n0 = xyz.shape[0]
    p.upload(xyz, types)
    p.set_velocities(u)
    del xyz, types, layer, u
    if method is None:
        method = SOLVE_LEARNED if weights is not None else SOLVE_MIXED
    if method == SOLVE_LEARNED:
        if weights is None:
            raise ValueError("the learned method needs weights")
        p.set_params(weights)
    sites = DropletSites(lx, ly, l0, speed, seed)
    du = np.array([LAYER_SPEED[DROPLET_LAYER] * speed, 0.0, 0.0])
    box_lo = np.array([-1e30, -1e30, -1e30])
    box_hi = np.array([lx, 1e30, 1e30])
    rows = []
    injected = removed = 0
    p.L.lsmps_synchronize(p.h)
    for step in range(steps):
        t0 = time.perf_counter()
        rec = {"step": step}
        if step % rebuild == 0:
            r = p.remove_outside(box_lo, box_hi)
            removed += r
            cs, ts = p.neighbors_skin(re, skin, maxn, l0, True)
            rec.update(rebuild=1, removed=r, cnl_ms=cs["ms"], pairs=cs["pairs"], truncated=cs["truncated"])
        else:
            rec.update(rebuild=0, removed=0, cnl_ms=0.0)
        p.set_field(C1_FIELD)
        wr = p.wls(re, l0, WLS_MIXED)
        pr = p.assemble(re, 1 << 30, XSTAR_RANDOM, seed * 1000 + step)
        st = p.solve(method, max_iter, tol, chunk=32)
        # Droplets enter, then everything moves.
        d = sites.take(inject, step * dt)
        if d.shape[0]:
            p.append(d, np.full(d.shape[0], INTERIOR, np.uint8), np.tile(du, (d.shape[0], 1)))
        injected += d.shape[0]
        p.advance(dt)
        p.L.lsmps_synchronize(p.h)
        wall = (time.perf_counter() - t0) * 1e3
        rec.update(particles=p.n, unknowns=pr["unknowns"], nnz=pr["nnz"], wls_ms=wr["ms"], assembly_ms=pr["ms"],
                   solve_ms=st["ms"], iterations=st["iterations"], converged=st["converged"],
                   fallback=st["fallback"], step_ms=wall)
        rows.append(rec)
        if verbose:
            log(f"step {step}: n={p.n} unknowns={pr['unknowns']} it={st['iterations']} conv={st['converged']} "
                f"fb={st['fallback']} step {wall:.1f} ms (cnl {rec['cnl_ms']:.2f}, wls {wr['ms']:.2f}, "
                f"asm {pr['ms']:.2f}, solve {st['ms']:.1f})")
    summary = {
        "scene": f"gap between walls 4 l0 apart, film on 2 layers, droplet layer open; l0={l0}, re=2.1 l0, "
                 f"skin=0.2 re (lists to re + skin, stages filter r < re)",
        "steps": steps, "inject_per_step": inject, "rebuild_every": rebuild, "dt": dt, "speed": speed,
        "method": {SOLVE_LEARNED: "learned+mixed", SOLVE_MIXED: "mixed", SOLVE_JACOBI: "fp64 jacobi"}[method],
        "particles_start": n0, "particles_end": int(p.n), "injected": injected, "removed": removed,
        "all_converged": bool(all(r["converged"] for r in rows)),
        "fallbacks": int(sum(r["fallback"] for r in rows)),
        "iterations": percentiles([r["iterations"] for r in rows]),
        "step_ms": percentiles([r["step_ms"] for r in rows]),
        "solve_ms": percentiles([r["solve_ms"] for r in rows]),
        "assembly_ms": percentiles([r["assembly_ms"] for r in rows]),
        "wls_ms": percentiles([r["wls_ms"] for r in rows]),
        "cnl_ms_rebuild_steps": percentiles([r["cnl_ms"] for r in rows if r["rebuild"]]),
    }
    if own:
        p.close()
    return summary, rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--particles", type=int, default=1_500_000)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--inject", type=int, default=2000)
    ap.add_argument("--rebuild", type=int, default=10)
    ap.add_argument("--weights", default=None,
                    help="learned-preconditioner weights (.npy); default: the weights shipped with the package")
    ap.add_argument("--method", choices=("learned", "mixed", "jacobi"), default="learned",
                    help="learned start + mixed solver (default), plain mixed, or FP64 Jacobi BiCGStab")
    ap.add_argument("--out", default=None, help="write the per-step rows as JSON here")
    args = ap.parse_args(argv)
    w = None
    if args.method == "learned":
        from .learn import DEFAULT_WEIGHTS, load_params

        w = load_params(args.weights or DEFAULT_WEIGHTS)
    method = {"learned": SOLVE_LEARNED, "mixed": SOLVE_MIXED, "jacobi": SOLVE_JACOBI}[args.method]
    summary, rows = run(args.particles, args.steps, args.inject, args.rebuild, weights=w, method=method)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"summary": summary, "rows": rows}, f, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
