import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURES = os.path.join(GOLDEN, "Test_input")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(GOLDEN, "goldens.json")) as f:
        return json.load(f)["cases"]


def fixture_path(name):
    return os.path.join(FIXTURES, name)


def boost_close(actual, expected, tol):
    """Boost.Test `tolerance(t)` on floating values: the strong check
    |a-e| <= t*|a| and |a-e| <= t*|e|, so an expected 0 must be exactly 0."""
    a = np.asarray(actual, np.float64)
    e = np.asarray(expected, np.float64)
    d = np.abs(a - e)
    return (d <= tol * np.abs(a)) & (d <= tol * np.abs(e)) | (d == 0)


def load_case_matrix(case):
    """Build the golden case's matrix with the oracle's loaders (CRS)."""
    import oracle
    m = case["matrix"]
    if m["kind"] == "poisson":
        return oracle.poisson_crs(m["m"], m["n"]), None
    if m["kind"] == "mtx":
        t = oracle.load_mm(fixture_path(m["file"]), m["has_data"], m["symmetric"], m["random_fill"])
    else:
        t = oracle.load_bin(fixture_path(m["file"]), m["size"], m["symmetric"], m["random"])
    return oracle.triplet_to_crs(t), t
