"""The reference's command-line program on the GPU library, end to end.

For each reference configuration: write its .cfg, run the CLI, and compare
with what the reference's own binary printed and wrote (tests/golden/
ref_configs.json). Two executables: this repository's CLI
(ls-mps_amd/lib/lsmps-bench) and, where the reference tree was present at
build time, the reference's unmodified src/main.cpp compiled against
include/lsmps.hpp and linked to liblsmps_hip.so (oracle/_ref/lsmps-bench-on-hip,
built by oracle/Makefile). Structure, types and the WLS statistics (FP64 path)
must match to the printed digit; the BiCGStab iteration count and errors must
fall in the reference's rounding band (see test_gpu_parity.py).
"""
import json
import os
import subprocess

import pytest
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

BANDS = json.load(open(os.path.join(GOLDEN, "solver_bands.json")))
EXES = {
    "own": os.path.join(ROOT, "ls-mps_amd", "lib", "lsmps-bench"),
    "reference_main": os.path.join(ROOT, "oracle", "_ref", "lsmps-bench-on-hip"),
}


def ppe_fields(line):
    return dict(kv.split("=") for kv in line.split()[1:])


@pytest.mark.parametrize("which", sorted(EXES))
def test_cli_matches_reference_output(ref_configs, which, tmp_path):
    exe = EXES[which]
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    for name, rec in ref_configs.items():
        print("[gpu] cli", which, name, flush=True)
        cfg = tmp_path / f"{name}.cfg"
        cfg.write_text("".join(f"{k}={v}\n" for k, v in rec["config"].items()))
        out = tmp_path / name
        p = subprocess.run([exe, str(cfg), str(out)], capture_output=True, text=True, timeout=300)
        assert p.returncode == rec["cli_exit"], (name, p.returncode, p.stderr)
        lines = p.stdout.splitlines()
        ref = rec["cli_stdout"]
        assert lines[:5] == ref[:5], (name, lines[:5])
        got, want = ppe_fields(lines[5]), ppe_fields(ref[5])
        for k in ("unknowns", "nnz", "converged"):
            assert got[k] == want[k], (name, k)
        band = BANDS["config:" + name]
        it = int(got["iterations"])
        assert 0.85 * band["iter_min"] - 3 <= it <= 1.15 * band["iter_max"] + 3, (name, it, band)
        assert float(got["rel_l2"]) <= 1.5 * band["rel_l2_max"], (name, got["rel_l2"])
        assert lines[6].startswith("PPE timing(ms) assembly=") and lines[7] == f"output={out}/{rec['config']['case']}.{{json,csv}}"
        j = json.loads((out / f"{rec['config']['case']}.json").read_text())
        r = json.loads(rec["cli_json"])
        for k in ("case", "particles", "types", "neighbors", "wls"):
            assert j[k] == r[k], (name, k)
        for k in ("unknowns", "nnz", "dirichlet", "neumann", "failed_rows", "converged"):
            assert j["ppe"][k] == r["ppe"][k], (name, k)
        csv = (out / f"{rec['config']['case']}.csv").read_text().splitlines()
        assert csv[0] == rec["cli_csv_header"]
        # The untimed prefix up to the iteration count is deterministic.
        row, want_row = csv[1].split(","), rec["cli_csv_row_untimed"].split(",")
        assert row[:14] == want_row[:14], (name, row[:14])
        if which == "own":
            g = json.loads((out / f"{rec['config']['case']}.gpu.json").read_text())
            assert g["particles"] == r["particles"] and g["nnz"] == r["ppe"]["nnz"]
            for st in ("neighbors_classify", "wls", "ppe_assembly", "ppe_solve_per_iteration"):
                assert g["stages"][st]["ms"] > 0 and g["stages"][st]["algorithmic_bytes"] > 0, st
