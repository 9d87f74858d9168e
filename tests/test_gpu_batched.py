"""GPU parity of the batched power method (pwm_batch_*) against the oracle.

Every instance of a batch is checked against the FP64 oracle
(CRS::powerIteration, Matrix/CRS.hpp:118-130) from the same start vector:
the eigenvalue estimate (Rayleigh quotient of the last normalised iterate)
within 1e-10 (FP64 batch) or 1e-6 (FP32/BF16 values, FP32 vectors), and the
final normalised vector within 1e-12 / 1e-4."""
import numpy as np
import pytest

import oracle

pw = pytest.importorskip("thesis_spmv_amd")
pytestmark = pytest.mark.gpu


def _rq(A, v):
    return float(v @ oracle.crs_mv(A, v) / (v @ v))


def _instances(rng, with_powerlaw=True):
    mats = [oracle.poisson_crs(5, 5), oracle.poisson_crs(10, 5), oracle.poisson_crs(37, 23), oracle.poisson_crs(64, 64)]
    if with_powerlaw:
        for n in (300, 2000, 9000):
            rs, ci, dv = pw.gen_spd_powerlaw(n, seed=int(rng.integers(1 << 30)))
            mats.append(oracle.CRSMatrix(n, n, rs, ci, dv))
        # nonsymmetric, nonnegative, with rows longer than one 4096-item tile
        n = 3000
        lens = rng.integers(0, 12, n)
        lens[7], lens[1500] = 9000, 5000
        rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        ci = rng.integers(0, n, rs[-1]).astype(np.int32)
        mats.append(oracle.CRSMatrix(n, n, rs, ci, rng.random(rs[-1]) + 0.1))
    return mats


@pytest.mark.parametrize("values,tol_lam,tol_vec", [("f64", 1e-10, 1e-12), ("f32", 1e-6, 1e-4)])
def test_fixed_iterations_against_oracle(values, tol_lam, tol_vec):
    rng = np.random.default_rng(0)
    mats = _instances(rng)
    b = pw.Batch(values)
    ids = [b.add_crs(A.row_start, A.col_ind, A.data) for A in mats]
    K = 60
    lam, its, ms = b.solve(K, -1.0, seed=11)
    print(f"batch {values}: {len(mats)} instances, {ms:.2f} ms", flush=True)
    assert np.all(its == K)
    for i, A in zip(ids, mats):
        x0 = b.start_vector(i, A.nor, seed=11)
        xr, yr = oracle.crs_power_iteration(A, x0, K - 1)
        v = xr if (K - 1) % 2 == 0 else yr
        ref = _rq(A, v)
        # the last instance has 9000-entry rows: FP32 row sums of that length
        # carry ~1e-6 relative error of their own
        tl = tol_lam if (values == "f64" or i < len(mats) - 1) else 1e-5
        err = abs(lam[i] - ref) / abs(ref)
        print(f"  instance {i} n={A.nor} nnz={A.nnz}: lambda rel.err {err:.2e}", flush=True)
        assert err <= tl, (i, A.nor, lam[i], ref)
        vec, _ = b.vector(i, A.nor)
        xf, yf = oracle.crs_power_iteration(A, x0, K)
        vf = xf if K % 2 == 0 else yf
        assert np.max(np.abs(vec - vf)) <= tol_vec, (i, A.nor)


def test_bf16_poisson_batch():
    b = pw.Batch("bf16")
    sizes = [(3, 3), (8, 5), (40, 40), (100, 7)]
    ids = [b.add_poisson(m, n) for m, n in sizes]
    lam, its, _ = b.solve(80, -1.0, seed=3)
    b64 = pw.Batch("f64")
    for m, n in sizes:
        b64.add_poisson(m, n)
    lam64, _, _ = b64.solve(80, -1.0, seed=3)
    for i, (m, n) in zip(ids, sizes):
        A = oracle.poisson_crs(m, n)
        x0 = b64.start_vector(i, A.nor, seed=3)
        xr, yr = oracle.crs_power_iteration(A, x0, 79)
        ref = _rq(A, yr)
        assert abs(lam64[i] - ref) <= 1e-10 * ref
        print(f"poisson {m}x{n}: bf16 lambda rel.err vs FP64 oracle {abs(lam[i] - ref) / ref:.2e}", flush=True)
        assert abs(lam[i] - ref) <= 1e-6 * ref


def test_converged_eigenvalues_closed_form():
    b = pw.Batch("f64")
    sizes = [(4, 3), (9, 7), (16, 11)]
    for m, n in sizes:
        b.add_poisson(m, n)
    lam, its, _ = b.solve(200000, 1e-13, seed=5)
    for i, (m, n) in enumerate(sizes):
        exact = 4 + 2 * np.cos(np.pi / (m + 1)) + 2 * np.cos(np.pi / (n + 1))
        assert its[i] < 200000 and abs(lam[i] - exact) <= 1e-9 * exact, (m, n, its[i], lam[i], exact)


def test_queue_with_more_instances_than_workgroups():
    """More instances than resident workgroups: each workgroup takes several."""
    b = pw.Batch("f32")
    rng = np.random.default_rng(2)
    sizes = [(int(rng.integers(2, 9)), int(rng.integers(2, 9))) for _ in range(1300)]
    for m, n in sizes:
        b.add_poisson(m, n)
    lam, its, ms = b.solve(30, -1.0, seed=9)
    print(f"1300 small instances: {ms:.2f} ms", flush=True)
    for i in range(0, 1300, 37):
        m, n = sizes[i]
        A = oracle.poisson_crs(m, n)
        x0 = b.start_vector(i, A.nor, seed=9)
        xr, yr = oracle.crs_power_iteration(A, x0, 29)
        ref = _rq(A, yr)
        assert abs(lam[i] - ref) <= 1e-5 * abs(ref)


def test_errors():
    b = pw.Batch("bf16")
    with pytest.raises(pw.PwmError):
        b.add_crs([0, 1, 2], [0, 1], [1.0, 0.1])  # 0.1 is not exact in BF16
    with pytest.raises(pw.PwmError):
        pw.Batch("f32").solve(10)  # no instances
    with pytest.raises(pw.PwmError):
        pw.Batch("f32").add_crs([0, 1], [3], [1.0])  # column out of range
