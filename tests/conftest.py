import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "ls-mps_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


SMALL_FIXTURES = ["bulk_trunc_1500", "film_1200", "rain_2000", "gap_1500", "film_trunc_1000", "edge"]


def load_fixture(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    out = {k: d[k] for k in d.files if k != "meta"}
    out["meta"] = json.loads(str(d["meta"]))
    return out


def load_ref_configs():
    with open(os.path.join(GOLDEN, "ref_configs.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ref_configs():
    return load_ref_configs()
