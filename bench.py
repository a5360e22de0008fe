"""Benchmark of the MI355X LS-MPS per-step particle pipeline (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--particles N]

Workload: the industrial_mix scene of 3,000,000 particles (70% jittered bulk
block, 20% wall film 2-4 l0 thick on a plane tilted 12 deg, 10% Poisson-disk
splash at density 0.2 / l0^3), l0 = 0.004, re = 2.1 l0, max_neighbors 128.
Particles are uploaded to HBM before timing. One step = cell-list neighbor
search + classification + WLS gradient/Laplacian of a quadratic field for
every particle (FP32 moments, FP64 redo of weak rows); the timed loop runs
K steps, `value` = particles / step time (Mparticles/s).

Also measured, outside the timed steps: PPE assembly over all fluid
unknowns, and the PPE solve to a true relative residual of 1e-8 with plain
FP64 right-Jacobi BiCGStab and with mixed-precision BiCGStab (FP32 Krylov
vectors and matrix, FP64 reliable updates), both on the same matrix and
right-hand side (seeded x* ~ U(-1, 1)).

The CPU baseline is the repository's single-threaded CPU oracle (a port of
the reference's algorithm) running neighbors + classification + WLS on a
bounded sample of the same scene, on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ls-mps_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=3_000_000)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solve-max-iter", type=int, default=40000)
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def wls_bytes_per_particle(k_mean):
    # positions 24, row_ptr 4, neighbor indices 4k, field 8, type 1,
    # gradient 24, Laplacian 8, status 1 (each read or written once).
    return 24 + 4 + 4 * k_mean + 8 + 1 + 24 + 8 + 1


def cnl_bytes_per_particle(k_mean):
    # positions 24, row pointer 4, neighbor indices 4k, type read+write 2.
    return 24 + 4 + 4 * k_mean + 2


def cpu_baseline(n_sample, l0, maxn):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    import lsmps_amd as L

    xyz, t = L.generate_case("industrial_mix", n_sample, 23, l0)
    o = O.Oracle(xyz, t)
    rs, ws, ps = o.run(l0, 2.1, maxn, 8192, 0, 0, 0.0, do_solve=False)
    cpu_ms = rs["neighbor_ms"] + rs["classify_ms"] + rs["wls_all_ms"]
    return {
        "value": round(n_sample / cpu_ms / 1e3, 4),
        "unit": "Mparticles/s",
        "cores": 1,
        "kind": "port",
        "sample": f"industrial_mix scene of {n_sample} particles (same generator, 1/{3_000_000 // n_sample} size): "
                  f"neighbors {rs['neighbor_ms']:.0f} ms + classify {rs['classify_ms']:.0f} ms + "
                  f"WLS over all particles {rs['wls_all_ms']:.0f} ms, single thread",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (this build has no "
              "multi-GPU path: the pipeline would shard by spatial domain, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import lsmps_amd as L

    l0, re_ratio, maxn = 0.004, 2.1, 128
    re = re_ratio * l0
    n = args.particles
    log(f"generating industrial_mix n={n}")
    xyz, types = L.generate_case("industrial_mix", n, 23, l0)
    p = L.Pipeline(0)
    p.upload(xyz, types)
    p.set_field(L.C1_FIELD)

    def step():
        cs, ts = p.neighbors_classify(re, maxn, l0)
        wr = p.wls(re, l0, L.WLS_MIXED)
        return cs, ts, wr

    for _ in range(args.warmup):
        step()
    log(f"warmup done ({args.warmup}); timing {args.steps} steps")
    p.L.lsmps_synchronize(p.h)
    t0 = time.perf_counter()
    stage = {"cnl": [], "classify": [], "wls": []}
    for _ in range(args.steps):
        cs, ts, wr = step()
        stage["cnl"].append(cs["ms"])
        stage["classify"].append(ts["ms"])
        stage["wls"].append(wr["ms"])
    p.L.lsmps_synchronize(p.h)
    wall_ms = (time.perf_counter() - t0) * 1e3 / max(1, args.steps)
    med = {k: float(np.median(v)) if v else 0.0 for k, v in stage.items()}
    gpu_step_ms = med["cnl"] + med["classify"] + med["wls"]
    kmean = cs["mean"]
    log(f"step wall {wall_ms:.3f} ms, gpu stages {med}, k_mean {kmean:.2f}, wls {wr}")

    # Roofline of the dominant kernel (the WLS pass, one launch per step).
    wls_bytes = n * wls_bytes_per_particle(kmean)
    wls_gbs = wls_bytes / (med["wls"] * 1e-3) / 1e9
    cnl_bytes = n * cnl_bytes_per_particle(kmean)
    cnl_gbs = cnl_bytes / (med["cnl"] * 1e-3) / 1e9

    extra = {"stage_ms": med, "neighbor_mean": kmean, "pairs": cs["pairs"], "types": ts["count"],
             "wls_fp64_redo": wr["fp64_redo"], "cnl_algorithmic_GBs": round(cnl_gbs, 1),
             "wls_algorithmic_GBs": round(wls_gbs, 1)}

    if not args.no_solve:
        pr = p.assemble(re, 1 << 30, L.XSTAR_RANDOM, 11)
        extra["ppe"] = {"unknowns": pr["unknowns"], "nnz": pr["nnz"], "assembly_ms": pr["ms"]}
        log(f"assembly {pr}")
        solves = {}
        for name, method in (("fp64_jacobi", L.SOLVE_JACOBI), ("mixed", L.SOLVE_MIXED)):
            st = p.solve(method, args.solve_max_iter, 1e-8, chunk=32)
            log(f"solve {name}: {st}")
            solves[name] = {"ms": round(st["ms"], 3), "iterations": st["iterations"], "converged": st["converged"],
                            "final_residual": st["final_residual"], "replacements": st["replacements"]}
        if solves["mixed"]["ms"] > 0:
            solves["speedup_mixed_vs_fp64"] = round(solves["fp64_jacobi"]["ms"] / solves["mixed"]["ms"], 3)
        extra["ppe_time_to_1e-8"] = solves
    p.close()

    out = {
        "metric": "CNL+WLS Mparticles/s (PPE time-to-1e-8 residual, mixed vs plain FP64, in ppe_time_to_1e-8)",
        "value": round(n / gpu_step_ms / 1e3, 3),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(gpu_step_ms, 4),
        "wall_ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 moments / fp64 positions, CSR and redo",
        "data": "synthetic",
        "config": {"workload": f"industrial_mix {n} particles: 70% jittered bulk, 20% tilted wall film 2-4 l0, "
                               f"10% Poisson-disk splash at 0.2/l0^3; l0=0.004, re=2.1 l0, max_neighbors=128",
                   "particles": n, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(wls_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(wls_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_wls_mixed (+ k_wls_exact_list redo)",
                     "bytes_per_launch": int(wls_bytes)},
    }
    out.update(extra)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(max(10_000, n // 10), l0, maxn)
        except Exception as e:  # the oracle library may be absent on a bare checkout
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
