"""Benchmark of the MI355X LS-MPS per-step particle pipeline (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--particles N]

Workload: the industrial_mix scene of 3,000,000 particles (70% jittered bulk
block, 20% wall film 2-4 l0 thick on a plane tilted 12 deg, 10% Poisson-disk
splash at density 0.2 / l0^3), l0 = 0.004, re = 2.1 l0, max_neighbors 128.
Particles are uploaded to HBM before timing. One step = cell-list neighbor
search + classification + WLS gradient/Laplacian of a quadratic field for
every particle (FP32 moments, FP64 for ill-conditioned rows); the timed loop runs
K steps, `value` = particles / step time (Mparticles/s).

Also measured, outside the timed steps: PPE assembly over all fluid
unknowns, and the PPE solve to a true relative residual of 1e-8 with plain
FP64 right-Jacobi BiCGStab, with mixed-precision BiCGStab (FP32 iteration
matrix, FP64 Krylov vectors and dot products, FP64 true-residual
replacements and restarts), and with the learned start and
diagonal inside the mixed solver (network weights from
ls-mps_amd/weights/mlp_industrial_mix.npy, trained by python -m
lsmps_amd.learn --case industrial_mix), all on the same matrix, over five
seeded right-hand sides x* ~ U(-1, 1) (medians reported), then once on the
reference's x* = x^2 + y^2.

The CPU baseline is the reference's own src/lsmps.cpp (built into
oracle/_ref where the reference tree was present) running neighbors +
classification + WLS over the same particles on one host core; without that
build, the repository's CPU oracle on a tenth-size scene.

`roofline` is for the step's dominant kernel (the neighbor pass or the WLS
kernel, whichever takes longer): algorithmic bytes of one launch over its
average duration from HIP events around it. `traffic` comes from the
committed counter summary under profiles/ when its particle count matches.
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
# Generator seed of the timed industrial_mix scene.
SCENE_SEED = 23
# Untimed steps with every HIP event, for the stage times.
STAGE_STEPS = 10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--particles", type=int, default=3_000_000)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--solvers", default="fp64_jacobi,mixed,learned_mixed,wls_rows,xstar_quadratic",
                    help="comma list of the PPE solves to time (fp64_jacobi, mixed, learned_mixed; wls_rows: "
                         "the WLS-Laplacian rows, assembled and solved with FP64 Jacobi and mixed; "
                         "xstar_quadratic: the reference's x* = x^2 + y^2)")
    ap.add_argument("--xstar-seeds", type=int, default=5,
                    help="seeded right-hand sides over which time-to-1e-8 is reported (median)")
    ap.add_argument("--scene-seeds", default="24,25,26,27",
                    help="further industrial_mix scenes (generator seeds) for the time-to-1e-8 sample "
                         "(ppe_time_to_1e-8_scenes, with the timed scene's first --scene-xstar-seeds solves); "
                         "empty to skip")
    ap.add_argument("--scene-xstar-seeds", type=int, default=3,
                    help="seeded right-hand sides per scene in that sample")
    ap.add_argument("--no-components", action="store_true",
                    help="skip the per-component error of the solution (needs the CSR on the host)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solve-max-iter", type=int, default=40000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run 3 device copies of MIB MiB (for counter calibration)")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def wls_bytes_per_particle(k_mean):
    # k_wls_mixed: position 24 + field 8, list length 4, neighbor indices 4k,
    # type 1 read; gradient 24, Laplacian 8, status 1 written.
    return 24 + 8 + 4 + 4 * k_mean + 1 + 24 + 8 + 1


def cnl_bytes_per_particle(k_mean):
    # k_cnl_pass (neighbor search + classification): position 24 and type 1
    # read; neighbor indices 4k, list length 4, new type 1 written.
    return 24 + 1 + 4 * k_mean + 4 + 1


def wls_exact_bytes_per_particle(k_mean):
    # k_wls_exact: the same inputs and outputs as the mixed kernel.
    return wls_bytes_per_particle(k_mean)


def assembly_bytes(n, m, nnz, k_mean):
    # Types 1 per particle; per unknown row: neighbor list 4k + length 4 +
    # position 24 + row maps 4; the matrix nnz * (4 + 8); diagonal, x* and b
    # 24 per row.
    return n * 1 + m * (4 * k_mean + 4 + 24 + 4) + nnz * 12 + m * 24


COUNTER_FILE = os.path.join(ROOT, "profiles", "counters_industrial_mix_3M.txt")
# VALU issue rate of MI355X: 256 CUs, one wavefront instruction per CU and
# cycle, 2.4 GHz.
VALU_PEAK_PER_S = 256 * 2.4e9


def _counter_lines(tag):
    try:
        with open(COUNTER_FILE) as f:
            return [ln.split() for ln in f if ln.startswith(tag + " ")]
    except OSError:
        return []


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the counter
    reads for a device copy of known size), or None when the summary is
    absent or was taken on another particle count."""
    for fields in _counter_lines("traffic"):
        # traffic <kernel> <particles> <fetch_bytes> <write_bytes>
        if len(fields) == 5 and fields[1] == kernel and int(fields[2]) == n:
            return int(float(fields[3]) + float(fields[4]))
    return None


def measured_valu(kernel, n):
    """VALU wavefront instructions per launch of `kernel` (SQ_INSTS_VALU)
    from the committed counter summary, or None."""
    for fields in _counter_lines("valu"):
        if len(fields) == 4 and fields[1] == kernel and int(fields[2]) == n:
            return int(fields[3])
    return None


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB through the HIP runtime (its
    blit kernel), so counter runs can scale FETCH_SIZE / WRITE_SIZE by what
    they read for a transfer of known size."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    nbytes = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, C.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy calibration: 3 device copies of {nbytes} bytes")


def cpu_baseline(xyz, types, l0, maxn):
    """Neighbor search + classification + WLS over every particle on the
    host, single thread. Prefers the reference's own src/lsmps.cpp built into
    oracle/_ref (kind "reference"); falls back to the repository's oracle
    (kind "port") on a bounded sample when that build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C

    import numpy as np
    import oracle_py as O

    n = len(types)
    try:
        ref = O.RefHarness()
    except (FileNotFoundError, OSError):
        ref = None
    # One host core, as taskset -c would pin it: the process's affinity is
    # narrowed to the first core it may run on for the timed CPU stages.
    allowed = sorted(os.sched_getaffinity(0))
    core = allowed[0]
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    host = {"cores": 1, "threads": 1, "pinned_core": core, "cpu_model": model, "host_cpus": len(allowed)}
    os.sched_setaffinity(0, {core})
    try:
        return _cpu_baseline_timed(ref, xyz, types, l0, maxn, n, host)
    finally:
        os.sched_setaffinity(0, set(allowed))


def _cpu_baseline_timed(ref, xyz, types, l0, maxn, n, host):
    import ctypes as C

    import numpy as np
    import oracle_py as O

    if ref is not None:
        # wls_samples = n makes the reference's wls_test visit every particle
        # (stride 1), so its WLS stage covers the whole workload.
        # max_iter 10 and tol 0: the reference's solve_ppe runs exactly ten
        # BiCGStab iterations on the full PPE (a whole solve would take an
        # hour on one core); its time per iteration is reported beside the
        # GPU's.
        ref.setup_particles(xyz, types, l0, 2.1, maxn=maxn, wls_samples=n, max_unknowns=n, max_iter=10, tol=0.0)
        out = np.zeros(4)
        ref.L.ref_time_stages(out.ctypes.data_as(C.c_void_p))
        cpu_ms = out[0] + out[1] + out[2]
        sv = ref.solve()
        per_it = sv["solve_ms"] / max(1, sv["iterations"])
        return dict({
            "value": round(n / cpu_ms / 1e3, 4),
            "unit": "Mparticles/s",
            "kind": "reference",
            "precision": "FP64 WLS (compare value_fp64)",
            "stage_ms": {"neighbors": round(out[0], 1), "classify": round(out[1], 1), "wls_all": round(out[2], 1),
                         "assemble_ppe": round(out[3], 1)},
            "ppe_bicgstab": {"iterations": sv["iterations"], "ms_per_iteration": round(per_it, 1),
                             "note": "the reference's solve_ppe (unpreconditioned FP64 BiCGStab on its CSR) over "
                                     "all unknowns, ten iterations, the final residual product included"},
            "sample": f"the full {n}-particle workload through the reference's src/lsmps.cpp (g++ -O3, its own "
                      f"flags), one thread pinned to core {host['pinned_core']}: neighbors {out[0]:.0f} ms + "
                      f"classify {out[1]:.0f} ms + wls_test over all particles {out[2]:.0f} ms; its assemble_ppe "
                      f"over all unknowns took {out[3]:.0f} ms (not in value)",
        }, **host)
    import lsmps_amd as L

    n_sample = max(10_000, n // 10)
    xyz_s, t_s = L.generate_case("industrial_mix", n_sample, 23, l0)
    o = O.Oracle(xyz_s, t_s)
    rs, _, _ = o.run(l0, 2.1, maxn, 8192, 0, 0, 0.0, do_solve=False)
    cpu_ms = rs["neighbor_ms"] + rs["classify_ms"] + rs["wls_all_ms"]
    return {
        **host,
        "value": round(n_sample / cpu_ms / 1e3, 4),
        "unit": "Mparticles/s",
        "kind": "port",
        "sample": f"industrial_mix scene of {n_sample} particles (same generator): neighbors "
                  f"{rs['neighbor_ms']:.0f} ms + classify {rs['classify_ms']:.0f} ms + WLS over all particles "
                  f"{rs['wls_all_ms']:.0f} ms, single thread",
    }


def scene_sample(n, l0, re, maxn, scene_seeds, xseeds, methods, weights, max_iter, first_rows):
    """Time-to-1e-8 over several scenes and right-hand sides: iteration
    counts (and so times) on these pinned graph Laplacians are chaotic, and
    the FP64-to-learned ratio of one solve ranges over ~1.3-2.5, so one
    scene's medians are a small sample. Each further scene is generated,
    built and assembled like the timed one, with the same solvers and
    weights; first_rows holds the timed scene's solves for the same seeds."""
    import numpy as np

    import lsmps_amd as L

    rows = list(first_rows)
    params = None
    if weights:
        from lsmps_amd import learn

        params = learn.load_params(weights)
    for sc in scene_seeds:
        xyz, t = L.generate_case("industrial_mix", n, sc, l0)
        q = L.Pipeline(0)
        try:
            q.upload(xyz, t)
            q.neighbors_classify(re, maxn, l0)
            if params is not None:
                q.set_params(params)
            q.assemble(re, 1 << 30, L.XSTAR_RANDOM, xseeds[0])
            for _, m in methods:  # allocation and graph capture
                q.solve(m, max_iter, 1e-8, chunk=32)
            for x in xseeds:
                q.assemble(re, 1 << 30, L.XSTAR_RANDOM, x)
                r = {"scene": sc, "xstar_seed": x}
                for name, m in methods:
                    st = q.solve(m, max_iter, 1e-8, chunk=32)
                    r[name] = {"ms": round(st["ms"], 3), "iterations": st["iterations"],
                               "converged": st["converged"], "fallback": st["fallback"]}
                log(f"scene {sc} x* seed {x}: {r}")
                rows.append(r)
        finally:
            q.close()
    names = [nm for nm, _ in methods]
    out = {"scenes": sorted({r["scene"] for r in rows}), "xstar_seeds": list(xseeds), "solves": rows,
           "all_converged": all(r[nm]["converged"] for r in rows for nm in names),
           "median_ms": {nm: round(float(np.median([r[nm]["ms"] for r in rows])), 3) for nm in names}}
    if "fp64_jacobi" in names:
        for nm in names:
            if nm == "fp64_jacobi":
                continue
            ratios = [r["fp64_jacobi"]["ms"] / r[nm]["ms"] for r in rows]
            out[f"speedup_{nm}_vs_fp64"] = {
                "of_medians": round(out["median_ms"]["fp64_jacobi"] / out["median_ms"][nm], 3),
                "median": round(float(np.median(ratios)), 3), "min": round(min(ratios), 3),
                "max": round(max(ratios), 3)}
    return out


def component_error(A, x):
    """Relative L2 and max error of x against x* after removing, on every
    connected component of the PPE graph, the mean of x - x*: the graph
    Laplacian pinned only at row 0 determines the solution on the other
    components up to a constant, which any converged solver may carry."""
    import numpy as np
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    m = len(A["b"])
    g = csr_matrix((np.ones(len(A["col"]), np.int8), A["col"], A["row"]), shape=(m, m))
    ncomp, lab = connected_components(g, directed=False)
    e = x - A["x_exact"]
    mean = np.bincount(lab, weights=e, minlength=ncomp) / np.maximum(np.bincount(lab, minlength=ncomp), 1)
    mean[lab[0]] = 0.0  # the pinned component is determined exactly
    e = e - mean[lab]
    return {"components": int(ncomp), "relative_l2_error": float(np.linalg.norm(e) / np.linalg.norm(A["x_exact"])),
            "max_error": float(np.abs(e).max())}


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
    xyz, types = L.generate_case("industrial_mix", n, SCENE_SEED, l0)
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
    p = L.Pipeline(0)
    p.upload(xyz, types)
    p.set_field(L.C1_FIELD)

    full = L.TIMING_STAGES | L.TIMING_KERNELS

    def run_steps(precision, k, timing=full):
        stage = {"cnl": [], "wls": [], "k_cnl_pass": [], "k_wls": []}
        p.set_timing(timing)
        p.L.lsmps_synchronize(p.h)
        t0 = time.perf_counter()
        for _ in range(k):
            # lsmps_step: the neighbor build and the WLS stage queued
            # together, one wait for both readbacks.
            cs, ts, wr = p.step(re, maxn, l0, precision)
            stage["cnl"].append(cs["ms"])
            stage["wls"].append(wr["ms"])
            stage["k_cnl_pass"].append(cs["kernel_ms"])
            stage["k_wls"].append(wr["kernel_ms"])
        p.L.lsmps_synchronize(p.h)
        wall = (time.perf_counter() - t0) * 1e3 / max(1, k)
        return wall, {key: float(np.mean(v)) if v else 0.0 for key, v in stage.items()}, cs, ts, wr

    # The step at the reference's precision (FP64 WLS, bit-exact with
    # wls_test) and with the mixed WLS (FP32 moments; rows whose FP32 pivots
    # fall to 0.2 of their diagonal, which include those the reference would
    # regularise, in FP64); the same neighbor pass in both.
    # The timed steps record only the HIP events around the stages' main
    # kernels (each recorded event costs the stream a few microseconds); the
    # stage times come from ten more steps with every event, after them.
    run_steps(L.WLS_FP64, args.warmup)
    log(f"warmup done ({args.warmup}); timing {args.steps} FP64 steps")
    wall64, avg64, cs, ts, wr64 = run_steps(L.WLS_FP64, args.steps, L.TIMING_KERNELS)
    _, st64, _, _, _ = run_steps(L.WLS_FP64, STAGE_STEPS)
    run_steps(L.WLS_MIXED, args.warmup)
    log(f"timing {args.steps} mixed steps")
    wall32, avg32, cs, ts, wr = run_steps(L.WLS_MIXED, args.steps, L.TIMING_KERNELS)
    _, st32, _, _, _ = run_steps(L.WLS_MIXED, STAGE_STEPS)
    for a, b in ((avg64, st64), (avg32, st32)):
        a["cnl"], a["wls"] = b["cnl"], b["wls"]
    kmean = cs["mean"]
    log(f"step wall fp64 {wall64:.3f} ms {avg64}; mixed {wall32:.3f} ms {avg32}; k_mean {kmean:.2f}; wls {wr}")
    avg = {"cnl": avg32["cnl"], "wls_mixed": avg32["wls"], "wls_fp64": avg64["wls"],
           "k_cnl_pass": 0.5 * (avg32["k_cnl_pass"] + avg64["k_cnl_pass"]), "k_wls_mixed": avg32["k_wls"],
           "k_wls_exact": avg64["k_wls"]}

    # Roofline of the dominant kernel of the headline step: algorithmic bytes
    # of one launch over its average duration (HIP events around that launch
    # on the library's stream).
    per_particle = {"k_cnl_pass": cnl_bytes_per_particle(kmean), "k_wls_mixed": wls_bytes_per_particle(kmean),
                    "k_wls_exact": wls_exact_bytes_per_particle(kmean)}
    kernel_gbs = {k: round(n * per_particle[k] / (avg[k] * 1e-3) / 1e9, 1) for k in per_particle if avg[k] > 0}
    dom = max(("k_cnl_pass", "k_wls_mixed"), key=lambda k: avg[k])
    dom_bytes = n * per_particle[dom]
    dom_gbs = dom_bytes / (avg[dom] * 1e-3) / 1e9

    extra = {"value_fp64": round(n / wall64 / 1e3, 3), "ms_per_step_fp64": round(wall64, 4),
             "stage_ms": {k: round(v, 4) for k, v in avg.items()}, "neighbor_mean": kmean, "pairs": cs["pairs"],
             "types": ts["count"], "wls_fp64_redo": wr["fp64_redo"], "kernel_algorithmic_GBs": kernel_gbs}

    if not args.no_solve:
        # Steady state, as inside a time-step loop: the first call of each
        # stage also allocates its buffers (and captures the solver's HIP
        # graph); that call is reported as first_call_ms, the second is timed.
        pr0 = p.assemble(re, 1 << 30, L.XSTAR_RANDOM, 11)
        pr = p.assemble(re, 1 << 30, L.XSTAR_RANDOM, 11)
        asm_bytes = assembly_bytes(n, pr["unknowns"], pr["nnz"], kmean)
        extra["ppe"] = {"unknowns": pr["unknowns"], "nnz": pr["nnz"], "padded_nnz": pr["padded_nnz"],
                        "assembly_ms": round(pr["ms"], 3), "fill_kernel_ms": round(pr["kernel_ms"], 4),
                        "first_call_ms": round(pr0["ms"], 3), "algorithmic_bytes": int(asm_bytes),
                        "algorithmic_GBs": round(asm_bytes / (pr["ms"] * 1e-3) / 1e9, 1)}
        log(f"assembly {pr}")
        # The target's aggregate: neighbor search + classification, WLS and
        # PPE assembly, algorithmic bytes over the three stage times.
        for tag, wkey, wk in (("mixed", "wls_mixed", "k_wls_mixed"), ("fp64", "wls_fp64", "k_wls_exact")):
            b = n * (per_particle["k_cnl_pass"] + per_particle[wk]) + asm_bytes
            t_ms = avg["cnl"] + avg[wkey] + pr["ms"]
            extra.setdefault("aggregate", {})[tag] = {
                "bytes": int(b), "ms": round(t_ms, 4), "GBs": round(b / (t_ms * 1e-3) / 1e9, 1),
                "frac": round(b / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        solves = {}
        wanted = set(args.solvers.split(","))
        methods = [(k, m) for k, m in (("fp64_jacobi", L.SOLVE_JACOBI), ("mixed", L.SOLVE_MIXED)) if k in wanted]
        # Weights trained on solves of this scene family (industrial_mix,
        # 100k-300k particles); the bulk-lattice weights otherwise.
        weights = os.path.join(ROOT, "ls-mps_amd", "weights", "mlp_industrial_mix.npy")
        if not os.path.exists(weights):
            weights = os.path.join(ROOT, "ls-mps_amd", "weights", "mlp_bulk_lattice.npy")
        extra["learned_weights"] = os.path.relpath(weights, ROOT)
        if os.path.exists(weights) and "learned_mixed" in wanted:
            from lsmps_amd import learn

            p.set_params(learn.load_params(weights))
            methods.append(("learned_mixed", L.SOLVE_LEARNED))
        for name, method in methods:  # allocation and graph capture
            st0 = p.solve(method, args.solve_max_iter, 1e-8, chunk=32)
            solves[name] = {"first_call_ms": round(st0["ms"], 3), "per_seed": []}
        seeds = [11 + k for k in range(max(1, args.xstar_seeds))]
        for seed in seeds:
            if seed != 11:
                p.assemble(re, 1 << 30, L.XSTAR_RANDOM, seed)
            for name, method in methods:
                st = p.solve(method, args.solve_max_iter, 1e-8, chunk=32)
                log(f"x* seed {seed} solve {name}: {st}")
                solves[name]["per_seed"].append(
                    {"seed": seed, "ms": round(st["ms"], 3), "iterations": st["iterations"],
                     "converged": st["converged"], "final_residual": st["final_residual"],
                     "replacements": st["replacements"], "breakdown_restarts": st["breakdown_restarts"],
                     "fallback": st["fallback"], "initial_residual": st["initial_residual"]})
                if seed == seeds[-1] and name == methods[-1][0] and not args.no_components:
                    A = p.ppe()
                    for nm, mt in methods:
                        p.solve(mt, args.solve_max_iter, 1e-8, chunk=32)
                        solves[nm]["error_modulo_components"] = component_error(A, p.solution())
                    del A
        for name in solves:
            ps = solves[name]["per_seed"]
            solves[name]["median_ms"] = round(float(np.median([r["ms"] for r in ps])), 3)
            solves[name]["median_iterations"] = float(np.median([r["iterations"] for r in ps]))
            solves[name]["all_converged"] = all(r["converged"] for r in ps)
        for name in ("mixed", "learned_mixed"):
            if name in solves and "fp64_jacobi" in solves:
                solves[f"speedup_{name}_vs_fp64"] = round(solves["fp64_jacobi"]["median_ms"] /
                                                          solves[name]["median_ms"], 3)
        extra["ppe_time_to_1e-8"] = solves
        if "xstar_quadratic" in wanted:
            # The reference's own manufactured pressure x* = x^2 + y^2
            # (src/lsmps.cpp:29): a smooth right-hand side with a small norm.
            p.assemble(re, 1 << 30, L.XSTAR_QUADRATIC, 0)
            quad = {}
            for name, method in methods:
                st = p.solve(method, args.solve_max_iter, 1e-8, chunk=32)
                log(f"x* = x^2 + y^2 solve {name}: {st}")
                quad[name] = {"ms": round(st["ms"], 3), "iterations": st["iterations"], "converged": st["converged"],
                              "final_residual": st["final_residual"], "breakdown_restarts": st["breakdown_restarts"],
                              "fallback": st["fallback"]}
            extra["ppe_xstar_quadratic"] = quad
        if "wls_rows" in wanted:
            # The WLS-Laplacian rows on the same particles (steady state).
            p.assemble_wls(re, l0, 1 << 30, L.XSTAR_RANDOM, 11)
            pw = p.assemble_wls(re, l0, 1 << 30, L.XSTAR_RANDOM, 11)
            log(f"wls rows {pw}")
            wls_rows = {"unknowns": pw["unknowns"], "nnz": pw["nnz"], "graph_rows": pw["graph_rows"],
                        "failed_rows": pw["failed_rows"], "assembly_ms": round(pw["ms"], 3),
                        "fill_kernel_ms": round(pw["kernel_ms"], 4)}
            for name, method in (("fp64_jacobi", L.SOLVE_JACOBI), ("mixed", L.SOLVE_MIXED)):
                p.solve(method, args.solve_max_iter, 1e-8, chunk=32)
                st = p.solve(method, args.solve_max_iter, 1e-8, chunk=32)
                log(f"wls rows solve {name}: {st}")
                wls_rows[name] = {"ms": round(st["ms"], 3), "iterations": st["iterations"],
                                  "converged": st["converged"]}
            extra["ppe_wls_rows"] = wls_rows
    p.close()
    scene_seeds = [int(v) for v in args.scene_seeds.split(",") if v.strip()]
    if not args.no_solve and scene_seeds and args.scene_xstar_seeds > 0 and "ppe_time_to_1e-8" in extra:
        solves = extra["ppe_time_to_1e-8"]
        xseeds = [11 + k for k in range(args.scene_xstar_seeds)]
        first = []
        for x in xseeds:
            r = {"scene": SCENE_SEED, "xstar_seed": x}
            for name, _ in methods:
                hit = [e for e in solves[name]["per_seed"] if e["seed"] == x]
                if not hit:
                    break
                r[name] = {k: hit[0][k] for k in ("ms", "iterations", "converged", "fallback")}
            else:
                first.append(r)
        extra["ppe_time_to_1e-8_scenes"] = scene_sample(
            n, l0, re, maxn, scene_seeds, xseeds, methods,
            weights if any(nm == "learned_mixed" for nm, _ in methods) else None, args.solve_max_iter, first)

    out = {
        "metric": "CNL+WLS Mparticles/s (PPE time-to-1e-8 residual, mixed vs plain FP64, in ppe_time_to_1e-8)",
        "value": round(n / wall32 / 1e3, 3),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 moments (f64 positions; f64 rows where the f32 pivot ratio is <= 0.2); value_fp64: f64",
        "data": "synthetic",
        "config": {"workload": f"industrial_mix {n} particles: 70% jittered bulk, 20% tilted wall film 2-4 l0, "
                               f"10% Poisson-disk splash at 0.2/l0^3; l0=0.004, re=2.1 l0, max_neighbors=128; "
                               f"step = neighbor search + classification + WLS gradient/Laplacian",
                   "particles": n, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(dom_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom_gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, n),
                     "kernel": dom, "kernel_ms": round(avg[dom], 4), "algorithmic_bytes": int(dom_bytes)},
    }
    # The same kernels against the VALU issue rate (the neighbor pass and the
    # WLS kernels are VALU-bound: their HBM fraction stays low by design).
    valu = {}
    for k in per_particle:
        inst = measured_valu(k, n)
        if inst and avg[k] > 0:
            valu[k] = {"instructions": inst, "per_particle": round(inst / n, 1),
                       "issue_frac": round(inst / (VALU_PEAK_PER_S * avg[k] * 1e-3), 4)}
    if valu:
        out["roofline"]["valu"] = valu
    if "aggregate" in extra:
        out["roofline"]["aggregate"] = extra.pop("aggregate")
    out.update(extra)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(xyz, types, l0, maxn)
        except Exception as e:  # the oracle library may be absent on a bare checkout
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
