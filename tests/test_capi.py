"""CPU-side checks of the product library (no GPU compute calls).

The C-ABI library loads and exports every function include/*.h declares;
the host-side generate_case reproduces the reference's particle sets bit
for bit; creating a context without a GPU fails with an error code instead
of crashing; the config parser follows the reference's rules.
"""
import hashlib
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest
from conftest import ROOT, SMALL_FIXTURES, load_fixture

import lsmps_amd as L

HEADERS = ["lsmps_hip.h", "lsmps_host.h"]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lsmps_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib_path = L.LIB_PATH
    assert os.path.exists(lib_path), "run __graft_entry__.build() first"
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for h in HEADERS:
        for fn in declared_functions(os.path.join(ROOT, "include", h)):
            assert fn in exported, f"{fn} declared in {h} but not exported"
    # The C++ API of lsmps.hpp is exported too (mangled names).
    for sym in ("load_config", "generate_case", "run_benchmark", "write_json", "write_csv", "type_name"):
        assert any(sym in e and "lsmps" in e for e in exported), sym


def test_ctypes_binding_covers_the_abi():
    declared = set(declared_functions(os.path.join(ROOT, "include", "lsmps_hip.h")))
    assert declared == set(L.ABI), declared ^ set(L.ABI)
    lib = L.load_library()
    assert lib.lsmps_abi_version() == 8


def test_generate_case_matches_reference(ref_configs):
    for name, rec in ref_configs.items():
        c = rec["config"]
        xyz, t = L.generate_case(c["case"], int(c["particles"]), int(c["seed"]), float(c["l0"]),
                                 int(c.get("film_layers", 4)))
        assert hashlib.sha256(xyz.tobytes()).hexdigest() == rec["sha"]["xyz"], name
        assert hashlib.sha256(t.tobytes()).hexdigest() == rec["sha"]["type_in"], name
    for name in SMALL_FIXTURES:
        d = load_fixture(name)
        m = d["meta"]
        if m["case"] == "edge":
            continue
        xyz, t = L.generate_case(m["case"], m["particles"], m["seed"], m["l0"], m["film_layers"])
        assert np.array_equal(xyz, d["xyz"]) and np.array_equal(t, d["type_in"])


def test_measurement_scenes_shapes():
    xyz, t = L.generate_case("bulk_lattice", 50000, 7, 0.01)
    assert xyz.shape == (50000, 3)
    # 50 x 50 x 20 block with +-0.1 l0 jitter.
    assert xyz[:, 0].max() < 0.5 and xyz[:, 2].max() < 0.2
    xyz, t = L.generate_case("rain_plate", 100000, 7, 0.01)
    assert np.bincount(t, minlength=5).tolist() == [55000, 0, 30000, 0, 15000]
    xyz2, t2 = L.generate_case("rain_plate", 100000, 7, 0.01)
    assert np.array_equal(xyz, xyz2)


def test_context_without_gpu_fails_cleanly():
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(L.LsmpsError):
        L.Pipeline(0)


def test_cli_error_exit_code():
    cli = os.path.join(ROOT, "ls-mps_amd", "lib", "lsmps-bench")
    with tempfile.TemporaryDirectory() as td:
        p = subprocess.run([cli, os.path.join(td, "missing.cfg"), td], capture_output=True, text=True)
        assert p.returncode == 1
        assert "cannot open config" in p.stderr
        bad = os.path.join(td, "bad.cfg")
        open(bad, "w").write("particles = lots\n")
        p = subprocess.run([cli, bad, td], capture_output=True, text=True)
        assert p.returncode == 1
