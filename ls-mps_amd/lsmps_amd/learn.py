"""Training and evaluation of the learned preconditioner.

The network (16 -> 64 -> 64 -> 2, tanh) is trained on the GPU by the
library's own kernels (lsmps_mlp_train) on PPE solves that the pipeline
generates itself: for each training scene, neighbors + classification +
assembly with a seeded manufactured pressure, an FP64 right-Jacobi solve to
1e-10, then lsmps_train_collect takes every stride-th row's features and
targets (x / s_b and the log of x diag / b, clamped).

    python -m lsmps_amd.learn --out ls-mps_amd/weights/mlp_bulk_lattice.npy

trains on bulk-lattice scenes of 100k-300k particles (seeds 1-200, the
50x50x20-style jittered block at l0 = 0.01) and evaluates on held-out seeds
at 1M particles: iterations and time to a true relative residual of 1e-8
for the learned mixed solver against plain mixed and FP64 Jacobi.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import (MLP_PARAMS, SOLVE_JACOBI, SOLVE_LEARNED, SOLVE_MIXED, XSTAR_RANDOM, Pipeline, generate_case)

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_DIR = os.path.join(os.path.dirname(HERE), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "mlp_bulk_lattice.npy")

L0 = 0.01
RE = 2.1 * L0
MAXN = 96
# Scene families the network is trained and evaluated on: particle spacing
# l0 and max_neighbors of each (re = 2.1 l0).
SCENES = {"bulk_lattice": (0.01, 96), "industrial_mix": (0.004, 128)}


def log(*a):
    print("[learn]", *a, file=sys.stderr, flush=True)


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weight matrices, zero biases; the last layer is scaled
    down so an untrained network starts near x0 = 0 and the Jacobi diagonal."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out, gain in ((16, 64, 1.0), (64, 64, 1.0), (64, 2, 0.1)):
        lim = gain * np.sqrt(6.0 / (fan_in + fan_out))
        parts.append(rng.uniform(-lim, lim, fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    p = np.concatenate(parts).astype(np.float32)
    assert p.shape[0] == MLP_PARAMS
    return p


def save_params(path: str, params: np.ndarray, meta: dict) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.ascontiguousarray(params, np.float32))
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


def load_params(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    p = np.load(path, allow_pickle=False)
    if p.dtype != np.float32 or p.shape != (MLP_PARAMS,):
        raise ValueError(f"{path}: expected {MLP_PARAMS} float32 parameters")
    return p


def scene_size(seed: int, lo: int = 100_000, hi: int = 300_000) -> int:
    """Particle count of training scene `seed`, spread over [lo, hi]."""
    return lo + (seed * 7919) % (hi - lo + 1)


def prepare(pipe: Pipeline, case: str, n: int, seed: int, l0: float = None, maxn: int = None):
    if l0 is None or maxn is None:
        l0, maxn = SCENES.get(case, (L0, MAXN))
    xyz, t = generate_case(case, n, seed, l0)
    pipe.upload(xyz, t)
    pipe.neighbors_classify(2.1 * l0, maxn, l0)
    return pipe.assemble(2.1 * l0, 1 << 30, XSTAR_RANDOM, seed)


def collect(pipe: Pipeline, seeds, case: str = "bulk_lattice", stride: int = 4, sizes=None, tol: float = 1e-10):
    """Adds the rows of one FP64 solve per seed to the context's training set."""
    total = 0
    for k, seed in enumerate(seeds):
        n = sizes[k] if sizes is not None else scene_size(seed)
        pr = prepare(pipe, case, n, seed)
        st = pipe.solve(SOLVE_JACOBI, 20000, tol, chunk=32)
        if not st["converged"]:
            log(f"seed {seed}: FP64 solve did not converge ({st['iterations']} iterations); skipped")
            continue
        total = pipe.train_collect(stride)
        if k % 10 == 0 or k == len(seeds) - 1:
            log(f"scene {k + 1}/{len(seeds)} seed {seed} n={n} unknowns={pr['unknowns']} "
                f"iterations={st['iterations']} samples={total}")
    return total


def train(pipe: Pipeline, steps: int, batch: int = 4096, lr: float = 2e-3, seed: int = 1, wk=(1.0, 1.0),
          chunks: int = 10):
    """Adam in `chunks` calls (progress lines between them); returns the loss curve."""
    curve = []
    per = max(1, steps // chunks)
    done = 0
    while done < steps:
        k = min(per, steps - done)
        r = pipe.train(k, batch, lr, wk, seed + done)
        curve.append((done + k, r["loss_first"], r["loss_last"], r["ms"]))
        done += k
        log(f"step {done}/{steps}: loss {r['loss_last']:.5f} ({r['ms']:.0f} ms)")
    return curve


def evaluate(pipe: Pipeline, seeds, n: int, case: str = "bulk_lattice", tol: float = 1e-8, repeats: int = 1):
    """Per seed: iterations and time to tol for FP64 Jacobi, mixed, learned."""
    rows = []
    for seed in seeds:
        pr = prepare(pipe, case, n, seed)
        rec = {"seed": seed, "unknowns": pr["unknowns"]}
        for name, method in (("fp64_jacobi", SOLVE_JACOBI), ("mixed", SOLVE_MIXED), ("learned", SOLVE_LEARNED)):
            st = None
            for _ in range(1 + repeats):  # the first call allocates and captures
                st = pipe.solve(method, 20000, tol, chunk=32)
            rec[name] = {"iterations": st["iterations"], "ms": st["ms"], "converged": st["converged"],
                         "fallback": st["fallback"], "initial_residual": st["initial_residual"],
                         "final_residual": st["final_residual"]}
        rows.append(rec)
        log(f"eval seed {seed}: " + ", ".join(f"{k} {rec[k]['iterations']} it {rec[k]['ms']:.1f} ms"
                                              for k in ("fp64_jacobi", "mixed", "learned")))
    return rows


def summarize(rows):
    out = {}
    for k in ("fp64_jacobi", "mixed", "learned"):
        out[k] = {"mean_iterations": float(np.mean([r[k]["iterations"] for r in rows])),
                  "mean_ms": float(np.mean([r[k]["ms"] for r in rows])),
                  "all_converged": bool(all(r[k]["converged"] for r in rows)),
                  "fallbacks": int(sum(r[k]["fallback"] for r in rows))}
    out["speedup_learned_vs_fp64"] = out["fp64_jacobi"]["mean_ms"] / out["learned"]["mean_ms"]
    out["speedup_mixed_vs_fp64"] = out["fp64_jacobi"]["mean_ms"] / out["mixed"]["mean_ms"]
    out["learned_gain_vs_mixed"] = out["mixed"]["mean_ms"] / out["learned"]["mean_ms"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--case", default="bulk_lattice", choices=sorted(SCENES),
                    help="scene family of the training solves (and of the evaluation unless --eval-case)")
    ap.add_argument("--eval-case", default=None, choices=sorted(SCENES))
    ap.add_argument("--init", default=None, help="start from these weights instead of a fresh initialisation")
    ap.add_argument("--train-seeds", type=int, default=200)
    ap.add_argument("--train-particles", default="100000,300000",
                    help="LO,HI: training scene sizes, spread over [LO, HI] by seed")
    ap.add_argument("--eval-seeds", type=int, default=20)
    ap.add_argument("--eval-particles", type=int, default=1_000_000)
    ap.add_argument("--stride", type=int, default=4)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--report", default=None, help="also write the evaluation as JSON here")
    ap.add_argument("--eval-only", action="store_true",
                    help="evaluate the weights at --out on the held-out seeds (no training); needs --report")
    args = ap.parse_args(argv)
    t0 = time.time()
    pipe = Pipeline(0)
    if args.eval_only:
        if not args.report:
            ap.error("--eval-only needs --report")
        params = load_params(args.out)
        pipe.set_params(params)
        eval_case = args.eval_case or args.case
        eval_seeds = list(range(1001, 1001 + args.eval_seeds))
        rows = evaluate(pipe, eval_seeds, args.eval_particles, case=eval_case)
        summary = summarize(rows)
        meta_path = os.path.splitext(args.out)[0] + ".json"
        meta = {}
        if os.path.exists(meta_path):
            with open(meta_path) as f:
                meta = json.load(f)
        meta["eval"] = {"case": eval_case, "seeds": eval_seeds, "particles": args.eval_particles, "tol": 1e-8,
                        "summary": summary}
        # The weights' metadata carries the latest evaluation, so the two
        # files never disagree.
        with open(meta_path, "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        with open(args.report, "w") as f:
            json.dump({"meta": meta, "rows": rows}, f, indent=1)
        log(f"evaluated {args.out}; summary {json.dumps(summary)}; {time.time() - t0:.0f} s")
        pipe.close()
        return
    pipe.set_params(load_params(args.init) if args.init else init_params(0))
    train_seeds = list(range(1, args.train_seeds + 1))
    lo, hi = (int(v) for v in args.train_particles.split(","))
    sizes = [scene_size(sd, lo, hi) for sd in train_seeds]
    n_samples = collect(pipe, train_seeds, case=args.case, stride=args.stride, sizes=sizes)
    log(f"training set: {n_samples} rows from {len(train_seeds)} scenes ({time.time() - t0:.0f} s)")
    curve = train(pipe, args.steps, args.batch, args.lr)
    params = pipe.get_params()
    eval_case = args.eval_case or args.case
    eval_seeds = list(range(1001, 1001 + args.eval_seeds))
    rows = evaluate(pipe, eval_seeds, args.eval_particles, case=eval_case)
    summary = summarize(rows)
    l0, maxn = SCENES[args.case]
    meta = {"train": {"case": args.case, "l0": l0, "re": 2.1 * l0, "max_neighbors": maxn, "seeds": [1, args.train_seeds],
                      "particles": [lo, hi], "stride": args.stride, "samples": n_samples,
                      "steps": args.steps, "batch": args.batch, "lr": args.lr, "solve": "FP64 Jacobi to 1e-10",
                      "loss_curve": curve},
            "init": os.path.basename(args.init) if args.init else None,
            "eval": {"case": eval_case, "seeds": eval_seeds, "particles": args.eval_particles, "tol": 1e-8,
                     "summary": summary}}
    save_params(args.out, params, meta)
    if args.report:
        with open(args.report, "w") as f:
            json.dump({"meta": meta, "rows": rows}, f, indent=1)
    log(f"saved {args.out}; summary {json.dumps(summary)}; {time.time() - t0:.0f} s")
    pipe.close()


if __name__ == "__main__":
    main()
