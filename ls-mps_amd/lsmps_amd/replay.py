"""Synthetic CSR replay of a customer-profile proxy on the GPU.

The reference's calibration tool writes a statistical proxy of a customer
workload as a directory of `.npy` files plus a step timeline
(python/workload_calibration.py:236-301):

  particle_type.npy            uint8 [n]    0 interior, 1 near wall, 2 free
                                            surface, 3 splash, 4 wall/other
  neighbor_row_length.npy      int32 [n]
  csr_row_ptr.npy              int64 [n+1]
  csr_col_idx.npy (optional)   int32 [nnz]  written with --materialize-csr
  virtual_neighbor_count.npy   int16 [n]
  wls_ill_conditioned.npy      bool  [n]
  timeline.csv                 step, injected, deleted, cnl_rebuild,
                               ppe_iterations
  metadata.json                the realized statistics

`load_contract` reads and checks them (dtypes and shapes as the tool writes
them; no pickles). Without csr_col_idx the columns are drawn here with the
tool's rule (python/workload_calibration.py:255-262: offsets rint(N(0,
max(8, sqrt n))) around the row, wrapped mod n) from a seeded numpy
generator: the same distribution, not the tool's draws, whose generator
state the files do not record.

`run` replays the timeline on the GPU through the C ABI: the graph becomes
the neighbor table (lsmps_replay_load), then per step the WLS stage on
synthetic edge offsets (lsmps_replay_wls; flagged rows take the FP64
fallback), the PPE assembly on rebuild steps (lsmps_replay_assemble) and
exactly `ppe_iterations` BiCGStab iterations (lsmps_solve_ppe with tol = 0).
Injection and deletion counts are reported, not applied: the proxy graph
carries no geometry to place new particles in.

    python -m lsmps_amd.replay <synthetic_workload dir> [--steps K]
    python -m lsmps_amd.replay <dir> --generate <profile.json>   (write, then replay)
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import sys
import time

import numpy as np

from . import SOLVE_JACOBI, SOLVE_MIXED, Pipeline

CONTRACT = {
    "particle_type": (np.uint8, "n"),
    "neighbor_row_length": (np.int32, "n"),
    "csr_row_ptr": (np.int64, "n+1"),
    "virtual_neighbor_count": (np.int16, "n"),
    "wls_ill_conditioned": (np.bool_, "n"),
}


def log(*a):
    print("[replay]", *a, file=sys.stderr, flush=True)


def synthesize_columns(row_len: np.ndarray, seed: int) -> np.ndarray:
    """Locality-biased columns with the calibration tool's distribution."""
    n = row_len.shape[0]
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n, dtype=np.int64), row_len.astype(np.int64))
    off = np.rint(rng.normal(0.0, max(8.0, math.sqrt(n)), rows.shape[0])).astype(np.int64)
    return ((rows + off) % n).astype(np.int32)


def load_contract(path: str, column_seed: int = 1) -> dict:
    """Reads and checks a synthetic-workload directory."""
    out = {}
    for name, (dtype, _) in CONTRACT.items():
        a = np.load(os.path.join(path, name + ".npy"), allow_pickle=False)
        if a.dtype != dtype or a.ndim != 1:
            raise ValueError(f"{name}.npy: expected 1-D {np.dtype(dtype).name}, got {a.ndim}-D {a.dtype}")
        out[name] = a
    n = out["particle_type"].shape[0]
    for name, (_, shape) in CONTRACT.items():
        want = n + 1 if shape == "n+1" else n
        if out[name].shape[0] != want:
            raise ValueError(f"{name}.npy: {out[name].shape[0]} entries, expected {want}")
    rp = out["csr_row_ptr"]
    if rp[0] != 0 or np.any(np.diff(rp) < 0):
        raise ValueError("csr_row_ptr.npy: must start at 0 and not decrease")
    if not np.array_equal(np.diff(rp), out["neighbor_row_length"].astype(np.int64)):
        raise ValueError("neighbor_row_length.npy disagrees with csr_row_ptr.npy")
    if out["particle_type"].max(initial=0) > 4:
        raise ValueError("particle_type.npy: values outside 0..4")
    cpath = os.path.join(path, "csr_col_idx.npy")
    if os.path.exists(cpath):
        col = np.load(cpath, allow_pickle=False)
        if col.dtype != np.int32 or col.shape != (int(rp[-1]),):
            raise ValueError(f"csr_col_idx.npy: expected int32 [{int(rp[-1])}], got {col.dtype} {col.shape}")
        if col.size and (col.min() < 0 or col.max() >= n):
            raise ValueError("csr_col_idx.npy: columns outside [0, n)")
        out["csr_col_idx"] = col
        out["columns"] = "file"
    else:
        out["csr_col_idx"] = synthesize_columns(out["neighbor_row_length"], column_seed)
        out["columns"] = f"synthesized (seed {column_seed})"
    out["timeline"] = []
    tpath = os.path.join(path, "timeline.csv")
    if os.path.exists(tpath):
        with open(tpath, newline="") as f:
            for r in csv.DictReader(f):
                out["timeline"].append({k: int(r[k]) for k in ("step", "injected", "deleted", "cnl_rebuild",
                                                                "ppe_iterations")})
    mpath = os.path.join(path, "metadata.json")
    out["metadata"] = json.load(open(mpath)) if os.path.exists(mpath) else {}
    return out


def write_contract(path: str, row_len, particle_type, wls_bad, virtual=None, col=None, timeline=None):
    """Writes arrays in the tool's file layout and dtypes (for tests and
    for callers that build a proxy themselves)."""
    os.makedirs(path, exist_ok=True)
    row_len = np.asarray(row_len, np.int32)
    rp = np.zeros(row_len.shape[0] + 1, np.int64)
    np.cumsum(row_len, out=rp[1:])
    np.save(os.path.join(path, "particle_type.npy"), np.asarray(particle_type, np.uint8))
    np.save(os.path.join(path, "neighbor_row_length.npy"), row_len)
    np.save(os.path.join(path, "csr_row_ptr.npy"), rp)
    np.save(os.path.join(path, "virtual_neighbor_count.npy"),
            np.asarray(virtual if virtual is not None else np.zeros(row_len.shape[0]), np.int16))
    np.save(os.path.join(path, "wls_ill_conditioned.npy"), np.asarray(wls_bad, np.bool_))
    if col is not None:
        np.save(os.path.join(path, "csr_col_idx.npy"), np.asarray(col, np.int32))
    if timeline is not None:
        with open(os.path.join(path, "timeline.csv"), "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=["step", "injected", "deleted", "cnl_rebuild", "ppe_iterations"])
            w.writeheader()
            for r in timeline:
                w.writerow(r)


def profile_row_lengths(n: int, mean: float, p90: float, p99: float, rng) -> np.ndarray:
    """Row lengths with a profile's mean, P90 and P99, built the way the
    calibration tool builds them (python/workload_calibration.py:175-209):
    a four-band mixture (50% low, 40% middle, 9% high, 1% tail band) and a
    final +-1 adjustment of random rows to the requested mean."""
    u = rng.random(n)
    low = max(1.0, 0.55 * mean)
    mid = max(low, min(p90, 1.06 * mean))
    high = max(mid, p90)
    tail = max(high, p99)
    bands = ((u < 0.50, 0.75 * low, 1.05 * low), ((u >= 0.50) & (u < 0.90), 0.90 * mid, 1.02 * high),
             ((u >= 0.90) & (u < 0.99), high, max(high + 1.0, 0.98 * tail)), (u >= 0.99, tail, 1.08 * tail))
    x = np.empty(n)
    for sel, a, b in bands:
        x[sel] = rng.uniform(a, b, int(sel.sum()))
    x = np.maximum(1, np.rint(x)).astype(np.int64)
    delta = int(round(mean * n)) - int(x.sum())
    if delta:
        order = rng.permutation(n)
        step = 1 if delta > 0 else -1
        k = 0
        while delta and k < 100 * n:
            i = order[k % n]
            if step > 0 or x[i] > 1:
                x[i] += step
                delta -= step
            k += 1
    return x.astype(np.int32)


def make_workload(metrics: dict, path: str, seed: int = 43, materialize: bool = False) -> dict:
    """A synthetic workload directory in the tool's layout from a profile's
    metrics (config/customer_car_rain_profile.example.json:5-18): row
    lengths as profile_row_lengths, types in the profile's ratios (near
    wall, free surface, splash; the rest interior) shuffled, flags with the
    ill-conditioned ratio, and a timeline of Poisson injection/deletion
    counts, the rebuild cadence and the PPE iteration count."""
    rng = np.random.default_rng(seed)
    n = int(round(metrics["particles"]))
    row_len = profile_row_lengths(n, metrics["neighbor_mean"], metrics["neighbor_p90"], metrics["neighbor_p99"], rng)
    ratios = [min(max(metrics.get(k, 0.0), 0.0), 1.0) for k in ("near_wall_ratio", "free_surface_ratio",
                                                               "splash_ratio")]
    if sum(ratios) > 0.95:
        ratios = [r * 0.95 / sum(ratios) for r in ratios]
    counts = [int(round(n * r)) for r in ratios]
    t = np.concatenate([np.zeros(n - sum(counts), np.uint8)] +
                       [np.full(c, k + 1, np.uint8) for k, c in enumerate(counts)])
    rng.shuffle(t)
    bad = rng.random(n) < metrics.get("wls_ill_conditioned_ratio", 0.0)
    virtual = rng.binomial(np.maximum(row_len, 1), min(metrics.get("virtual_particle_ratio", 0.0), 0.8))
    steps = int(metrics.get("timeline_steps", 200))
    rebuild = max(1, int(round(metrics.get("cnl_rebuild_interval", 10))))
    inj, dele = metrics.get("injected_per_step", 0), metrics.get("deleted_per_step", 0)
    it = metrics.get("ppe_iterations", 1)
    tl = [{"step": s, "injected": int(rng.poisson(inj)) if inj else 0, "deleted": int(rng.poisson(dele)) if dele else 0,
           "cnl_rebuild": int(s % rebuild == 0),
           "ppe_iterations": max(1, int(round(rng.normal(it, max(1.0, 0.05 * it)))))} for s in range(steps)]
    col = synthesize_columns(row_len, seed + 1) if materialize else None
    write_contract(path, row_len, t, bad, virtual=virtual.astype(np.int16), col=col, timeline=tl)
    meta = {"particles": n, "neighbor_mean": float(row_len.mean()),
            "neighbor_p90": float(np.quantile(row_len, 0.9, method="higher")),
            "neighbor_p99": float(np.quantile(row_len, 0.99, method="higher")),
            "wls_ill_conditioned_ratio": float(bad.mean()), "csr_nnz": int(row_len.sum())}
    with open(os.path.join(path, "metadata.json"), "w") as f:
        json.dump(meta, f, indent=1)
    return meta


def percentiles(v):
    v = np.asarray(v, np.float64)
    if v.size == 0:
        return {}
    return {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99)), "max": float(v.max()),
            "mean": float(v.mean())}


def run(path: str, steps: int | None = None, l0: float = 0.004, re_ratio: float = 2.1, method: str = "mixed",
        seed: int = 7, pipe: Pipeline | None = None, verbose: bool = True):
    """Replays the workload's timeline; returns (summary, per-step rows)."""
    d = load_contract(path, column_seed=seed)
    tl = d["timeline"] or [{"step": 0, "injected": 0, "deleted": 0, "cnl_rebuild": 1, "ppe_iterations": 1}]
    if steps is not None:
        tl = tl[:steps]
    own = pipe is None
    p = pipe or Pipeline(0)
    re = re_ratio * l0
    solve = {"mixed": SOLVE_MIXED, "jacobi": SOLVE_JACOBI}[method]
    t0 = time.perf_counter()
    info = p.replay_load(d["csr_row_ptr"], d["csr_col_idx"], d["particle_type"], d["wls_ill_conditioned"], seed)
    load_s = time.perf_counter() - t0
    if verbose:
        log(f"loaded {info['rows']} rows, {info['edges']} edges (columns {d['columns']}), longest row "
            f"{info['max_row']}, {info['flagged']} flagged, in {load_s:.2f} s")
    rows = []
    pr = None
    for k, ev in enumerate(tl):
        t1 = time.perf_counter()
        rec = dict(ev)
        if pr is None or ev["cnl_rebuild"]:
            pr = p.replay_assemble(1 << 30, seed * 1000 + k, l0)
            rec["assembly_ms"] = pr["ms"]
        else:
            rec["assembly_ms"] = 0.0
        wr = p.replay_wls(re, l0)
        st = p.solve(solve, max(1, ev["ppe_iterations"]), 0.0, chunk=8)
        p.L.lsmps_synchronize(p.h)
        rec.update(wls_ms=wr["ms"], fp64_redo=wr["fp64_redo"], regularized=wr["regularized"], solve_ms=st["ms"],
                   iterations=st["iterations"], unknowns=pr["unknowns"], nnz=pr["nnz"],
                   step_ms=(time.perf_counter() - t1) * 1e3)
        rows.append(rec)
        if verbose:
            log(f"step {ev['step']}: wls {wr['ms']:.2f} ms ({wr['fp64_redo']} redo), asm {rec['assembly_ms']:.2f} ms, "
                f"{st['iterations']} it {st['ms']:.2f} ms, step {rec['step_ms']:.1f} ms")
    summary = {
        "rows": info["rows"], "edges": info["edges"], "columns": d["columns"], "max_row": info["max_row"],
        "types": info["types"], "flagged": info["flagged"], "self_edges": info["self_edges"],
        "unknowns": pr["unknowns"], "ppe_nnz": pr["nnz"], "method": method, "steps": len(rows),
        "injected_total": int(sum(r["injected"] for r in rows)), "deleted_total": int(sum(r["deleted"] for r in rows)),
        "step_ms": percentiles([r["step_ms"] for r in rows]), "solve_ms": percentiles([r["solve_ms"] for r in rows]),
        "wls_ms": percentiles([r["wls_ms"] for r in rows]),
        "assembly_ms_rebuild_steps": percentiles([r["assembly_ms"] for r in rows if r["cnl_rebuild"] or
                                                  r is rows[0]]),
        "load_s": load_s,
    }
    if own:
        p.close()
    return summary, rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("path", help="synthetic_workload directory written by the calibration tool")
    ap.add_argument("--generate", default=None, metavar="PROFILE_JSON",
                    help="first write a workload into PATH from a profile's metrics")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--l0", type=float, default=0.004)
    ap.add_argument("--method", choices=("mixed", "jacobi"), default="mixed")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=None, help="write the per-step rows as JSON here")
    args = ap.parse_args(argv)
    if args.generate:
        doc = json.load(open(args.generate))
        meta = make_workload(doc.get("metrics", doc), args.path)
        log(f"wrote {args.path}: {meta}")
    summary, rows = run(args.path, args.steps, args.l0, method=args.method, seed=args.seed)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"summary": summary, "rows": rows}, f, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
