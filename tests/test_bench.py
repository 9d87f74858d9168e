"""CPU checks of bench.py's measurement contract: the algorithmic bytes it
divides by (the per-SpMV formula b_v*Z + 4Z + 4(N+1) + 2*b_x*N, pinned to
the configurations' closed-form sizes), its Poisson generator against the
oracle's, the served stream's request law, and its refusal of more than one
GPU."""
import os
import subprocess
import sys

import numpy as np

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_algorithmic_bytes_of_the_configurations():
    # 2D Poisson 1024x1024: Z = 5mn - 2m - 2n
    z0 = 5 * 1024 * 1024 - 2 * 1024 - 2 * 1024
    assert z0 == 5238784
    assert bench.algorithmic_bytes(1024 * 1024, z0) == 83836932
    # 9-point stencil on 2896^2 plus 2 random entries per row: (3*2896-2)^2 + 2N
    n2 = 2896 * 2896
    z2 = (3 * 2896 - 2) ** 2 + 2 * n2
    assert z2 == 92220228
    assert bench.algorithmic_bytes(n2, z2) == 1274379060
    assert bench.algorithmic_bytes(n2, z2, 4, 4) == 838403620  # FP32 values and vectors


def test_bench_poisson_equals_the_oracle():
    for m, n in ((1, 1), (3, 2), (17, 9), (64, 48)):
        N, rs, ci, dv = bench.poisson_crs(m, n)
        A = oracle.poisson_crs(m, n)
        assert N == A.nor
        assert np.array_equal(rs, A.row_start) and np.array_equal(ci, A.col_ind) and np.array_equal(dv, A.data)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr and r.stdout == ""


def test_served_requests_are_the_same_at_every_rate():
    """The served-stream sweep compares rates on one job mix: only the
    arrival times change, in proportion to 1/rate."""
    a = bench._serve_requests(60, 10.0, 12)
    b = bench._serve_requests(60, 2000.0, 12)
    strip = [{k: v for k, v in r.items() if k != "t"} for r in a]
    assert strip == [{k: v for k, v in r.items() if k != "t"} for r in b]
    ta, tb = np.array([r["t"] for r in a]), np.array([r["t"] for r in b])
    assert np.all(np.diff(ta) > 0) and np.allclose(ta * 10.0, tb * 2000.0, rtol=1e-12)
    assert {r["mat"] for r in a} <= set(range(12))
    solves = [r for r in a if r["tol"] > 0]
    assert all(r["steps"] == 1000 for r in solves) and all(1 <= r["steps"] <= 64 for r in a if r["tol"] < 0)
