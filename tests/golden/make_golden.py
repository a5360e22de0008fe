"""Writes the golden vectors under tests/golden/ from the reference itself.

Runs the reference's own src/lsmps.cpp, compiled by oracle/Makefile (target
``ref``) into oracle/_ref/, in the container that holds the reference tree:

    make -C oracle ref && python tests/golden/make_golden.py

Outputs (data only: inputs and the reference's outputs on them):
  * ``<name>.npz``  small cases with full arrays: positions, types before and
    after classification, neighbor CSR in the reference's order, per-particle
    WLS results of the reference's test field, PPE CSR/b/x*, solve stats.
  * ``ref_configs.json``  the reference's five shipped configs at full size:
    SHA-256 of every array plus all scalar stats, and the text of the
    reference CLI's JSON/CSV/stdout with the wall-clock fields removed.
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle_py as O  # noqa: E402

REF_CFG_DIR = "/root/reference/config"

# name, case, particles, seed, l0, film_layers, max_neighbors, wls_samples, max_unknowns
SMALL = [
    ("bulk_trunc_1500", "bulk_uniform", 1500, 3, 0.01, 4, 12, 4096, 1000),
    ("film_1200", "wall_film", 1200, 5, 0.01, 4, 96, 4096, 6000),
    ("rain_2000", "rain_injection", 2000, 9, 0.01, 4, 96, 4096, 6000),
    ("gap_1500", "narrow_gap", 1500, 17, 0.008, 4, 128, 4096, 6000),
    ("film_trunc_1000", "wall_film", 1000, 21, 0.01, 2, 16, 4096, 6000),
]
# The reference's shipped configs (file name, case, particles, seed, l0,
# film_layers, max_neighbors, wls_samples, max_unknowns, max_iter, tol).
CONFIGS = ["smoke", "bulk_uniform", "wall_film", "rain_injection", "narrow_gap"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def parse_cfg(path):
    out = {}
    for line in open(path):
        line = line.split("#", 1)[0]
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        out[k.strip()] = v.strip()
    return out


def run_case(R, case, n, seed, l0, fl, maxn, ws, cap, max_iter=1000, tol=1e-7):
    R.setup(case, n, seed, l0, 2.1, fl, maxn, ws, cap, max_iter, tol)
    xyz, t0 = R.particles()
    row, col, trunc = R.neighbors()
    R.classify()
    _, t1 = R.particles()
    wst = R.wls_test()
    g, lap, st = R.wls_particles(np.arange(len(t1), dtype=np.uint64))
    A = R.assemble()
    S = R.solve()
    return dict(xyz=xyz, type_in=t0, type_out=t1, nbr_row=row, nbr_col=col, truncated=trunc,
                wls=wst, wls_grad=g, wls_lap=lap, wls_status=st, ppe=A, solve=S)


def edge_particles():
    """Hand-built set that drives the failure paths of the reference:
    a single-layer plane (coplanar neighbors -> regularization ladder), a
    3-D jittered block, and two clusters of 12 coincident particles (zero
    weight sum -> failed WLS; one sits below 1.5*l0 so it becomes a NearWall
    PPE unknown with no usable neighbor -> failed PPE row)."""
    l0 = 0.01
    rng = np.random.default_rng(1234)
    pts = []
    for j in range(18):
        for i in range(18):
            pts.append(((i + 0.5) * l0 + rng.uniform(-.05, .05) * l0, (j + 0.5) * l0 + rng.uniform(-.05, .05) * l0,
                        0.08))
    for k in range(6):
        for j in range(10):
            for i in range(10):
                pts.append(((i + 0.5 + rng.uniform(-.1, .1)) * l0 + 0.3, (j + 0.5 + rng.uniform(-.1, .1)) * l0,
                            (k + 0.5 + rng.uniform(-.1, .1)) * l0))
    for c in [(0.8, 0.8, 0.5), (0.8, 0.5, 0.005)]:
        pts.extend([c] * 12)
    xyz = np.array(pts, np.float64)
    t = np.zeros(len(xyz), np.uint8)
    return xyz, t, l0


def strip_times(text):
    text = re.sub(r'"(assembly_ms|solve_ms)": [-0-9.e+]+', r'"\1": null', text)
    return text


def main():
    R = O.RefHarness()
    for name, case, n, seed, l0, fl, maxn, ws, cap in SMALL:
        d = run_case(R, case, n, seed, l0, fl, maxn, ws, cap)
        A, S, W = d["ppe"], d["solve"], d["wls"]
        meta = dict(case=case, particles=n, seed=seed, l0=l0, re_ratio=2.1, film_layers=fl, max_neighbors=maxn,
                    wls_samples=ws, max_unknowns=cap, max_iter=1000, tol=1e-7, truncated=d["truncated"],
                    wls=W, ppe={k: A[k] for k in ("unknowns", "nnz", "failed_rows", "neumann", "dirichlet")},
                    solve={k: S[k] for k in S if k != "solve_ms"})
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), xyz=d["xyz"], type_in=d["type_in"], type_out=d["type_out"],
            nbr_row=d["nbr_row"], nbr_col=d["nbr_col"], wls_grad=d["wls_grad"], wls_lap=d["wls_lap"],
            wls_status=d["wls_status"], ppe_row=A["row"], ppe_col=A["col"], ppe_val=A["val"], ppe_b=A["b"],
            ppe_x_exact=A["x_exact"], meta=np.array(json.dumps(meta)))
        print(name, "pairs", len(d["nbr_col"]), "trunc", d["truncated"], "nnz", A["nnz"], "iters", S["iterations"])

    xyz, t, l0 = edge_particles()
    R.setup_particles(xyz, t, l0, 2.1, 96, 4096, 6000, 1000, 1e-7)
    row, col, trunc = R.neighbors()
    R.classify()
    _, t1 = R.particles()
    wst = R.wls_test()
    g, lap, st = R.wls_particles(np.arange(len(t1), dtype=np.uint64))
    A = R.assemble()
    S = R.solve()
    meta = dict(case="edge", particles=len(xyz), seed=0, l0=l0, re_ratio=2.1, film_layers=4, max_neighbors=96,
                wls_samples=4096, max_unknowns=6000, max_iter=1000, tol=1e-7, truncated=trunc, wls=wst,
                ppe={k: A[k] for k in ("unknowns", "nnz", "failed_rows", "neumann", "dirichlet")},
                solve={k: S[k] for k in S if k != "solve_ms"})
    np.savez_compressed(os.path.join(HERE, "edge.npz"), xyz=xyz, type_in=t, type_out=t1, nbr_row=row, nbr_col=col,
                        wls_grad=g, wls_lap=lap, wls_status=st, ppe_row=A["row"], ppe_col=A["col"],
                        ppe_val=A["val"], ppe_b=A["b"], ppe_x_exact=A["x_exact"], meta=np.array(json.dumps(meta)))
    print("edge", "status counts", np.bincount(st, minlength=4), "failed_rows", A["failed_rows"], "wls", wst)

    out = {}
    for cname in CONFIGS:
        cfg = parse_cfg(os.path.join(REF_CFG_DIR, cname + ".cfg"))
        case = cfg["case"]
        n = int(cfg["particles"])
        seed = int(cfg["seed"])
        l0 = float(cfg["l0"])
        fl = int(cfg.get("film_layers", 4))
        maxn = int(cfg["max_neighbors"])
        ws = int(cfg["wls_samples"])
        cap = int(cfg["ppe_max_unknowns"])
        it = int(cfg["ppe_max_iter"])
        tol = float(cfg["ppe_tol"])
        d = run_case(R, case, n, seed, l0, fl, maxn, ws, cap, it, tol)
        A, S = d["ppe"], d["solve"]
        cnt = np.diff(d["nbr_row"])
        rec = dict(config=cfg, particles=n, pairs=int(len(d["nbr_col"])), truncated=int(d["truncated"]),
                   type_counts=[int(x) for x in np.bincount(d["type_out"], minlength=5)],
                   neighbor_count_hist=[int(x) for x in np.bincount(cnt)],
                   sha={"xyz": sha(d["xyz"]), "type_in": sha(d["type_in"]), "type_out": sha(d["type_out"]),
                        "nbr_row": sha(d["nbr_row"]), "nbr_col": sha(d["nbr_col"]), "ppe_row": sha(A["row"]),
                        "ppe_col": sha(A["col"]), "ppe_val": sha(A["val"]), "ppe_b": sha(A["b"]),
                        "ppe_x_exact": sha(A["x_exact"])},
                   wls=d["wls"], ppe={k: A[k] for k in ("unknowns", "nnz", "failed_rows", "neumann", "dirichlet")},
                   solve={k: S[k] for k in S if k != "solve_ms"})
        with tempfile.TemporaryDirectory() as td:
            proc = subprocess.run([O.REF_BIN, os.path.join(REF_CFG_DIR, cname + ".cfg"), td], capture_output=True,
                                  text=True, check=False)
            rec["cli_exit"] = proc.returncode
            rec["cli_stdout"] = [ln for ln in proc.stdout.splitlines() if not ln.startswith(("PPE timing", "output="))]
            rec["cli_json"] = strip_times(open(os.path.join(td, case + ".json")).read())
            csv = open(os.path.join(td, case + ".csv")).read().splitlines()
            rec["cli_csv_header"] = csv[0]
            rec["cli_csv_row_untimed"] = ",".join(csv[1].split(",")[:-2])
        out[cname] = rec
        print(cname, "pairs", rec["pairs"], "nnz", A["nnz"], "iters", S["iterations"], "exit", rec["cli_exit"])
    with open(os.path.join(HERE, "ref_configs.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
