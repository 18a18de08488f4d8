import gzip
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

ge.package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


def golden_spec(name):
    with open(os.path.join(GOLDEN, name + "_spec.json")) as f:
        return json.load(f)


def golden_meta(name):
    with open(os.path.join(GOLDEN, name + "_golden.json")) as f:
        return json.load(f)


def golden_f64(fname):
    with open(os.path.join(GOLDEN, fname), "rb") as f:
        return np.frombuffer(gzip.decompress(f.read()), dtype="<f8").copy()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as orc
    if not os.path.exists(orc.LIB_PATH):
        orc.build()
    return orc
