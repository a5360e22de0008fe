"""Writes tests/golden/solver_bands.json: the spread of the reference's
BiCGStab (src/lsmps.cpp:67) under rounding-level perturbations.

BiCGStab's iteration count and final error on these pinned graph Laplacians
are chaotic in the last bits of the dot products: any change of summation
order (pairwise, blocked, parallel on a GPU) moves them. To say what
"matching the reference" means for a solver that reduces in a different
order, this script reruns the reference's algorithm (same updates, guards
and stopping rule) on each golden PPE system 40 times, each time
multiplying every dot product by (1 + 2e-16 * N(0,1)) from a seeded
generator -- a perturbation of the size of one rounding -- and records the
range of iteration counts and of the relative L2 error against x*.

    python tests/golden/make_solver_bands.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))

TRIALS = 40
EPS = 2e-16


def bicgstab_perturbed(A, b, tol, maxit, rng):
    n = len(b)
    x = np.zeros(n)
    r = b.copy()
    r0 = r.copy()
    p = np.zeros(n)
    v = np.zeros(n)

    def dot(a, c):
        return float(np.dot(a, c)) * (1.0 + EPS * rng.standard_normal())

    thr = tol * max(1.0, np.sqrt(dot(b, b)))
    rho0 = alpha = omega = 1.0
    for it in range(maxit):
        rho = dot(r0, r)
        if abs(rho) < 1e-30:
            return it, False, x
        beta = (rho / rho0) * (alpha / omega)
        p = r + beta * (p - omega * v)
        v = A @ p
        den = dot(r0, v)
        if abs(den) < 1e-30:
            return it, False, x
        alpha = rho / den
        s = r - alpha * v
        if np.sqrt(dot(s, s)) <= thr:
            return it + 1, True, x + alpha * p
        t = A @ s
        tt = dot(t, t)
        if tt < 1e-30:
            return it, False, x
        omega = dot(t, s) / tt
        x = x + (alpha * p + omega * s)
        r = s - omega * t
        if np.sqrt(dot(r, r)) <= thr:
            return it + 1, True, x
        rho0 = rho
    return maxit, False, x


def band(row, col, val, b, xe, tol, maxit):
    n = len(b)
    A = sp.csr_matrix((val, col, row), shape=(n, n))
    its, rels, conv = [], [], []
    for trial in range(TRIALS):
        it, ok, x = bicgstab_perturbed(A, b, tol, maxit, np.random.default_rng(trial))
        its.append(it)
        conv.append(ok)
        rels.append(float(np.linalg.norm(x - xe) / max(np.linalg.norm(xe), 1e-300)))
    return {"iter_min": int(min(its)), "iter_max": int(max(its)), "rel_l2_min": min(rels), "rel_l2_max": max(rels),
            "converged_all": bool(all(conv)), "trials": TRIALS, "eps": EPS}


def main():
    from conftest import SMALL_FIXTURES, load_fixture, load_ref_configs

    import oracle_py as O

    out = {}
    for name in SMALL_FIXTURES:
        d = load_fixture(name)
        m = d["meta"]
        out[name] = band(d["ppe_row"], d["ppe_col"], d["ppe_val"], d["ppe_b"], d["ppe_x_exact"], m["tol"],
                         m["max_iter"])
        print(name, out[name], flush=True)
    for name, rec in load_ref_configs().items():
        c = rec["config"]
        l0 = float(c["l0"])
        xyz, t = O.generate_case(c["case"], int(c["particles"]), int(c["seed"]), l0, int(c.get("film_layers", 4)))
        o = O.Oracle(xyz, t)
        o.neighbors(2.1 * l0, int(c["max_neighbors"]))
        o.classify(l0)
        o.assemble(2.1 * l0, int(c["ppe_max_unknowns"]))
        A = o.get_ppe()
        out["config:" + name] = band(A["row"], A["col"], A["val"], A["b"], A["x_exact"], float(c["ppe_tol"]),
                                     int(c["ppe_max_iter"]))
        print(name, out["config:" + name], flush=True)
    with open(os.path.join(HERE, "solver_bands.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
