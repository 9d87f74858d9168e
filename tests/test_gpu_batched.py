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


def test_many_small_instances():
    """More instances than the chip has compute units, most of one tile."""
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


def _mixed_batch(values, mats, **opt):
    b = pw.Batch(values, **opt)
    for A in mats:
        b.add_crs(A.row_start, A.col_ind, A.data)
    return b


@pytest.mark.parametrize("values", ["f64", "f32"])
def test_alone_equals_in_the_mix(values):
    """An instance's tiles are aligned to its own start and its reductions run
    in a fixed order, so its result is bit-identical alone and in a batch."""
    rng = np.random.default_rng(7)
    mats = _instances(rng)
    b = _mixed_batch(values, mats)
    lam, its, _ = b.solve(300, 1e-9, seed=21)
    assert len(set(its.tolist())) > 1  # instances stop at different steps
    for i in (0, 3, 5, len(mats) - 1):
        # the start vector hashes (seed, instance id, row): keep the id by
        # placing 2x2 Poisson instances before it (instance 0 runs truly alone)
        solo = pw.Batch(values)
        for _ in range(i):
            solo.add_poisson(2, 2)
        solo.add_crs(mats[i].row_start, mats[i].col_ind, mats[i].data)
        lam1, its1, _ = solo.solve(300, 1e-9, seed=21)
        assert its1[i] == its[i] and lam1[i] == lam[i], (i, lam1[i], lam[i])
        v_mix, yn_mix = b.vector(i, mats[i].nor)
        v_solo, yn_solo = solo.vector(i, mats[i].nor)
        assert np.array_equal(v_mix, v_solo) and yn_mix == yn_solo


def test_graph_and_chunking_do_not_change_results():
    rng = np.random.default_rng(8)
    mats = _instances(rng)
    ref = None
    for opt in ({"graph": 1, "chunk": 32}, {"graph": 0, "chunk": 32}, {"graph": 1, "chunk": 7},
                {"graph": 0, "chunk": 1}):
        b = _mixed_batch("f32", mats, **opt)
        out = []
        for rep in range(2):  # graph replays of a second solve too
            lam, its, _ = b.solve(150, 1e-8, seed=4)
            vecs = [b.vector(i, A.nor)[0] for i, A in enumerate(mats)]
            out.append((lam, its, vecs))
        st = b.stats()
        print(f"options {opt}: stats {st}", flush=True)
        if opt["graph"]:
            assert st["graphs"] >= 1
        for lam, its, vecs in out:
            if ref is None:
                ref = (lam, its, vecs)
            assert np.array_equal(lam, ref[0]) and np.array_equal(its, ref[1])
            for v, w in zip(vecs, ref[2]):
                assert np.array_equal(v, w)


@pytest.mark.parametrize("ipt", [4, 16])
def test_tile_size_variants(ipt):
    rng = np.random.default_rng(9)
    mats = _instances(rng)
    b = _mixed_batch("f64", mats, items_per_thread=ipt)
    K = 40
    lam, its, _ = b.solve(K, -1.0, seed=2)
    for i, A in enumerate(mats):
        x0 = b.start_vector(i, A.nor, seed=2)
        xr, yr = oracle.crs_power_iteration(A, x0, K - 1)
        ref = _rq(A, xr if (K - 1) % 2 == 0 else yr)
        assert abs(lam[i] - ref) <= 1e-10 * abs(ref), (ipt, i)


@pytest.mark.parametrize("values,opt", [("f64", {}), ("f32", {}), ("f32", {"bf16_when_exact": 1})])
def test_autotune_keeps_a_timed_schedule(values, opt):
    """The first solve after the upload times the candidate tile schedules and
    keeps the fastest; the results are those of a batch fixed to that
    schedule, bit for bit, and a second solve does not tune again."""
    rng = np.random.default_rng(12)
    mats = _instances(rng)
    b = _mixed_batch(values, mats, autotune=1, **opt)
    assert not b.schedule()["tuned"]
    lam, its, _ = b.solve(200, 1e-9, seed=5)
    s = b.schedule()
    print(f"{values} {opt}: schedule {s}", flush=True)
    assert s["tuned"]
    ms = s["candidate_ms"]
    expect = {(8, False), (16, False)} if opt else {(8, False), (8, True), (16, False), (16, True)}
    assert set(ms) == expect and all(v > 0 for v in ms.values())
    assert ms[(s["items_per_thread"], s["persistent"])] == min(ms.values())
    fixed = _mixed_batch(values, mats, items_per_thread=s["items_per_thread"], persistent=int(s["persistent"]), **opt)
    lam2, its2, _ = fixed.solve(200, 1e-9, seed=5)
    assert np.array_equal(lam, lam2) and np.array_equal(its, its2)
    for i, A in enumerate(mats):
        assert np.array_equal(b.vector(i, A.nor)[0], fixed.vector(i, A.nor)[0])
    lam3, its3, _ = b.solve(200, 1e-9, seed=5)
    assert b.schedule() == s
    assert np.array_equal(lam, lam3) and np.array_equal(its, its3)
    # against the oracle: K plain steps
    K = 40
    lam4, _, _ = b.solve(K, -1.0, seed=2)
    tol = 1e-10 if values == "f64" else 1e-5  # FP32 row sums of 9000-entry rows
    for i, A in enumerate(mats):
        x0 = b.start_vector(i, A.nor, seed=2)
        xr, yr = oracle.crs_power_iteration(A, x0, K - 1)
        ref = _rq(A, xr if (K - 1) % 2 == 0 else yr)
        assert abs(lam4[i] - ref) <= tol * abs(ref), (i, lam4[i], ref)


def test_autotune_edge_cases():
    """Autotuning a batch whose caller fixed 256 x 4 tiles (more tile slots
    than any candidate needs), a one-instance batch, and a batch that gets a
    new instance after it was tuned (the next solve uploads and tunes again)."""
    rng = np.random.default_rng(13)
    mats = _instances(rng)
    b = _mixed_batch("f32", mats, items_per_thread=4, autotune=1)
    lam, its, _ = b.solve(30, -1.0, seed=3)
    s = b.schedule()
    assert s["tuned"] and s["items_per_thread"] in (8, 16)
    fixed = _mixed_batch("f32", mats, items_per_thread=s["items_per_thread"], persistent=int(s["persistent"]))
    lam2, _, _ = fixed.solve(30, -1.0, seed=3)
    assert np.array_equal(lam, lam2)
    one = pw.Batch("f64", autotune=1)
    one.add_poisson(3, 2)
    lam1, its1, _ = one.solve(25, -1.0, seed=1)
    A = oracle.poisson_crs(3, 2)
    xr, yr = oracle.crs_power_iteration(A, one.start_vector(0, seed=1), 24)
    assert abs(lam1[0] - _rq(A, xr)) <= 1e-10 * abs(lam1[0]) and one.schedule()["tuned"]  # 24 steps: in x
    b.add_poisson(20, 20)
    assert not b.schedule()["tuned"]
    lam3, its3, _ = b.solve(30, -1.0, seed=3)
    assert b.schedule()["tuned"] and len(lam3) == len(mats) + 1


def test_errors():
    b = pw.Batch("bf16")
    with pytest.raises(pw.PwmError):
        b.add_crs([0, 1, 2], [0, 1], [1.0, 0.1])  # 0.1 is not exact in BF16
    with pytest.raises(pw.PwmError):
        pw.Batch("f32").solve(10)  # no instances
    with pytest.raises(pw.PwmError):
        pw.Batch("f32").add_crs([0, 1], [3], [1.0])  # column out of range


def test_bf16_where_exact_is_bitwise_the_pure_batches():
    """An FP32 batch with bf16_when_exact stores the Poisson instances (4 and
    -1) as BF16 and the power-law ones as FP32, in one launch per step: each
    instance's result is bit-identical to the same instance in a pure BF16
    (Poisson) or pure FP32 (power-law) batch, and every eigenvalue is within
    1e-6 of the FP64 oracle."""
    rng = np.random.default_rng(9)
    mats = _instances(rng)
    mixed = pw.Batch("f32", bf16_when_exact=1)
    ids = [mixed.add_crs(A.row_start, A.col_ind, A.data) for A in mats]
    K = 50
    lam_m, _, _ = mixed.solve(K, -1.0, seed=5)
    f32 = pw.Batch("f32")
    for A in mats:
        f32.add_crs(A.row_start, A.col_ind, A.data)
    lam_f, _, _ = f32.solve(K, -1.0, seed=5)
    # _instances lists the four Poisson matrices first: in a BF16 batch of
    # just those they keep their ids (and so their start vectors)
    bf2 = pw.Batch("bf16")
    for A in mats[:4]:
        bf2.add_crs(A.row_start, A.col_ind, A.data)
    lam_b, _, _ = bf2.solve(K, -1.0, seed=5)
    for i, A in zip(ids, mats):
        vm, _ = mixed.vector(i, A.nor)
        if i < 4:  # the Poisson instances: BF16 values
            vb, _ = bf2.vector(i, A.nor)
            assert lam_m[i] == lam_b[i] and np.array_equal(vm, vb), i
        else:
            vf, _ = f32.vector(i, A.nor)
            assert lam_m[i] == lam_f[i] and np.array_equal(vm, vf), i
        x0 = mixed.start_vector(i, A.nor, seed=5)
        xr, yr = oracle.crs_power_iteration(A, x0, K - 1)
        ref = _rq(A, xr if (K - 1) % 2 == 0 else yr)
        tl = 1e-6 if i < len(mats) - 1 else 1e-5
        assert abs(lam_m[i] - ref) <= tl * abs(ref), (i, lam_m[i], ref)
    with pytest.raises(pw.PwmError):
        pw.Batch("f64", bf16_when_exact=1)
