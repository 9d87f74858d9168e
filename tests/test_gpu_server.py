"""GPU tests of the request server (pwm_server_*): results against the FP64
oracle (CRS::powerIteration, Matrix/CRS.hpp:118-130), bit-identical results
alone and in a mix of concurrently running jobs, admission and eviction
under a device-memory budget, and error handling."""
import numpy as np
import pytest

import oracle

pw = pytest.importorskip("thesis_spmv_amd")
pytestmark = pytest.mark.gpu


def _rq(A, v):
    return float(v @ oracle.crs_mv(A, v) / (v @ v))


def _drain(srv, jobs, limit=100000):
    out = {}
    for _ in range(limit):
        srv.step()
        for j in list(jobs):
            try:
                out[j] = srv.result(j, jobs[j])
                del jobs[j]
            except pw.PwmError as e:
                assert e.code == -6  # still running
        if not jobs:
            return out
    raise AssertionError("jobs did not finish")


def _powerlaw(rng, n, long_rows=()):
    lens = rng.integers(0, 14, n)
    for r, L in long_rows:
        lens[r] = L
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, int(L), replace=False)) for L in lens]).astype(np.int32)
    return oracle.CRSMatrix(n, n, rs, ci, rng.random(int(rs[-1])) + 0.1)  # Perron root: a real gap


@pytest.mark.parametrize("values,tol_lam,tol_vec", [("f64", 1e-10, 1e-12), ("f32", 2e-6, 1e-4)])
def test_chains_and_solves_against_oracle(values, tol_lam, tol_vec):
    rng = np.random.default_rng(1)
    srv = pw.Server(values, budget_bytes=256 << 20, max_jobs=64)
    mats = {srv.add_poisson(30, 20): oracle.poisson_crs(30, 20)}
    A = _powerlaw(rng, 5000, long_rows=[(10, 4000), (4000, 2500)])
    mats[srv.add_matrix(A.row_start, A.col_ind, A.data)] = A
    jobs, spec = {}, {}
    for k, (mat, L) in enumerate([(0, 1), (0, 7), (1, 1), (1, 33), (0, 64), (1, 2)]):
        j = srv.submit(mat, L, -1.0, seed=100 + k)
        jobs[j] = mats[mat].nor
        spec[j] = (mat, L, 100 + k)
    res = _drain(srv, jobs)
    for j, (lam, yn, its, vec) in res.items():
        mat, L, seed = spec[j]
        A = mats[mat]
        assert its == L
        x0 = pw.server_start_vector(A.nor, seed)
        if values != "f64":
            x0 = x0.astype(np.float32).astype(np.float64)
        # L steps: y_L = A^L x0 / prod|y_k|, normalised; lambda_L = RQ of
        # the iterate before the last step
        xr, yr = oracle.crs_power_iteration(A, x0, L - 1) if L > 1 else (x0, None)
        prev = xr if (L - 1) % 2 == 0 else yr
        ref_lam = _rq(A, prev)
        assert abs(lam - ref_lam) <= tol_lam * abs(ref_lam), (j, lam, ref_lam)
        y = oracle.crs_mv(A, prev / (np.linalg.norm(prev) if L > 1 else 1.0))
        assert abs(yn - np.linalg.norm(y)) <= tol_lam * 10 * np.linalg.norm(y)
        assert np.max(np.abs(vec - y / np.linalg.norm(y))) <= tol_vec


def test_power_solves_converge_to_closed_form():
    srv = pw.Server("f64", budget_bytes=64 << 20, max_jobs=8)
    sizes = [(6, 5), (12, 9)]
    ids = [srv.add_poisson(m, n) for m, n in sizes]
    jobs = {srv.submit(i, 200000, 1e-13, seed=3): m * n for i, (m, n) in zip(ids, sizes)}
    res = _drain(srv, dict(jobs))
    for (j, _), (m, n) in zip(jobs.items(), sizes):
        lam, _, its, _ = res[j]
        exact = 4 + 2 * np.cos(np.pi / (m + 1)) + 2 * np.cos(np.pi / (n + 1))
        assert its < 200000 and abs(lam - exact) <= 1e-9 * exact


@pytest.mark.parametrize("overlap", [1, 0])
@pytest.mark.parametrize("values", ["f32", "f64"])
def test_alone_equals_in_the_mix(values, overlap):
    """A job's result is bit-identical whether it runs alone or joins while
    other jobs are running and others arrive after it, with the joining jobs
    staged while a chunk runs (overlap 1) or before it (overlap 0)."""
    rng = np.random.default_rng(5)
    A = _powerlaw(rng, 3000, long_rows=[(5, 3000)])
    B = _powerlaw(rng, 800)

    def fresh():
        srv = pw.Server(values, budget_bytes=128 << 20, max_jobs=32, chunk=3, overlap=overlap)
        a = srv.add_matrix(A.row_start, A.col_ind, A.data)
        b = srv.add_matrix(B.row_start, B.col_ind, B.data)
        p = srv.add_poisson(40, 25)
        return srv, a, b, p

    srv, a, b, p = fresh()
    alone = _drain(srv, {srv.submit(a, 500, 1e-9, seed=77): A.nor})
    (lam0, yn0, it0, v0), = alone.values()

    srv, a, b, p = fresh()
    others = {srv.submit(b, 40, -1.0, seed=1): B.nor, srv.submit(p, 300, 1e-10, seed=2): 1000}
    running_seen = []
    for _ in range(2):
        running_seen.append(srv.step()[0])
    target = srv.submit(a, 500, 1e-9, seed=77)  # joins while others run
    more = {srv.submit(b, 9, -1.0, seed=3): B.nor, srv.submit(a, 12, -1.0, seed=4): A.nor}
    res = _drain(srv, {**others, target: A.nor, **more})
    lam1, yn1, it1, v1 = res[target]
    assert max(running_seen) >= 2
    assert it1 == it0 and lam1 == lam0 and yn1 == yn0 and np.array_equal(v1, v0)


def test_budget_admission_and_eviction():
    # three 200x200 Poisson matrices (1.75 MB each in FP32, job vectors
    # 0.32 MB) under a budget that holds only two of them
    srv = pw.Server("f32", budget_bytes=4 << 20, max_jobs=16)
    ids = [srv.add_poisson(200, 200) for _ in range(3)]
    refs = {}
    for k, i in enumerate(ids):
        j = srv.submit(i, 20, -1.0, seed=k)
        refs[j] = _drain(srv, {j: 40000})[j]
        assert srv.resident(i)
    st = srv.stats()
    assert st["resident"] == 2 and st["used_bytes"] <= 4 << 20
    assert not srv.resident(ids[0])  # least recently used, evicted
    # the evicted matrix comes back (regenerated) with identical results
    j = srv.submit(ids[0], 20, -1.0, seed=0)
    again = _drain(srv, {j: 40000})[j]
    first = refs[min(refs)]
    assert again[0] == first[0] and np.array_equal(again[3], first[3])
    # many jobs at once: they queue and run as memory frees up (staged while
    # other jobs run, evicting only matrices no staged or running job uses)
    jobs = {srv.submit(ids[k % 3], 15, -1.0, seed=k): 40000 for k in range(9)}
    res = _drain(srv, dict(jobs))
    assert len(res) == 9
    assert all(its == 15 and np.isfinite(lam) for lam, yn, its, vec in res.values())
    # the first three jobs again, submitted together: identical results
    one = {srv.submit(ids[k % 3], 20, -1.0, seed=k): 40000 for k in range(3)}
    for j, r in _drain(srv, dict(one)).items():
        ref = refs[min(refs) + (j - min(one))]
        assert r[0] == ref[0] and np.array_equal(r[3], ref[3])


def test_errors():
    srv = pw.Server("bf16", budget_bytes=1 << 20, max_jobs=4)
    with pytest.raises(pw.PwmError) as e:
        srv.add_matrix([0, 1, 2], [0, 1], [1.0, 0.1])  # 0.1 is not exact in BF16
    assert e.value.code == -1
    with pytest.raises(pw.PwmError) as e:
        srv.add_poisson(1000, 1000)  # larger than the whole budget
    assert e.value.code == -7
    p = srv.add_poisson(10, 10)
    with pytest.raises(pw.PwmError):
        srv.submit(p, 0)
    with pytest.raises(pw.PwmError):
        srv.submit(99, 10)
    j = srv.submit(p, 50)
    with pytest.raises(pw.PwmError) as e:
        srv.result(j)
    assert e.value.code == -6
    with pytest.raises(pw.PwmError) as e:
        srv.remove_matrix(p)  # a job is queued on it
    assert e.value.code == -6
    res = _drain(srv, {j: 100})
    assert res[j][2] == 50
    srv.remove_matrix(p)
    with pytest.raises(pw.PwmError):
        srv.submit(p, 10)
    with pytest.raises(pw.PwmError):
        srv.result(j)  # already collected


def test_bf16_where_exact_is_bitwise_the_pure_servers():
    """With bf16_when_exact an FP32 server stores the Poisson matrices and the
    matrices whose values are exact in BF16 as BF16, the others as FP32; each
    job's result equals the same job on a pure BF16 or pure FP32 server, bit
    for bit."""
    rng = np.random.default_rng(9)
    A = _powerlaw(rng, 2500, long_rows=[(7, 2000)])  # values in (0.1, 1.1): FP32 storage
    B = _powerlaw(rng, 1200)
    B = oracle.CRSMatrix(B.nor, B.noc, B.row_start, B.col_ind, np.round(B.data * 8) + 1.0)  # small integers: BF16

    def run(values, **opt):
        srv = pw.Server(values, budget_bytes=128 << 20, max_jobs=16, chunk=4, **opt)
        ids = {}
        if values != "bf16":
            ids["A"] = srv.add_matrix(A.row_start, A.col_ind, A.data)
        ids["B"] = srv.add_matrix(B.row_start, B.col_ind, B.data)
        ids["P"] = srv.add_poisson(45, 30)
        n = {"A": A.nor, "B": B.nor, "P": 45 * 30}
        jobs = {}
        for k, (name, L, tol) in enumerate([("A", 300, 1e-9), ("B", 40, -1.0), ("P", 400, 1e-10), ("B", 250, 1e-9),
                                            ("A", 17, -1.0), ("P", 5, -1.0)]):
            if name in ids:
                jobs[srv.submit(ids[name], L, tol, seed=50 + k)] = (k, n[name])
        res = _drain(srv, {j: nn for j, (k, nn) in jobs.items()})
        return {jobs[j][0]: r for j, r in res.items()}

    mixed = run("f32", bf16_when_exact=1)
    f32 = run("f32")
    b16 = run("bf16")
    for k, (lam, yn, its, vec) in mixed.items():
        ref = b16[k] if k in b16 else f32[k]
        if k in (0, 4):
            ref = f32[k]
        assert its == ref[2] and lam == ref[0] and yn == ref[1] and np.array_equal(vec, ref[3]), k
    with pytest.raises(pw.PwmError) as e:
        pw.Server("f64", budget_bytes=1 << 20, max_jobs=2, bf16_when_exact=1)
    assert e.value.code == -1
    srv = pw.Server("f32", budget_bytes=1 << 20, max_jobs=2)
    srv.add_poisson(4, 4)
    with pytest.raises(pw.PwmError) as e:
        srv.set_option("bf16_when_exact", 1)
    assert e.value.code == -6
    # a borrowed matrix stored as BF16 whose caller changes a value to one
    # BF16 cannot hold: the upload refuses it instead of narrowing it
    srv = pw.Server("f32", budget_bytes=16 << 20, max_jobs=2, bf16_when_exact=1)
    rs = np.array([0, 2, 3, 5], np.int32)
    ci = np.array([0, 2, 1, 0, 2], np.int32)
    dv = np.array([2.0, 1.0, 3.0, 1.0, 2.0])
    mat = srv.add_matrix(rs, ci, dv, borrow=True)
    srv._held[mat][2][1] = 0.1
    srv.submit(mat, 5, -1.0, seed=1)
    with pytest.raises(pw.PwmError) as e:
        srv.step()
    assert e.value.code == -1


@pytest.mark.parametrize("values", ["f32", "f64"])
def test_borrowed_columns_checked_at_upload(values):
    """A borrowed matrix whose caller moves a column index out of range after
    registration is refused when it is uploaded (the columns are checked as
    they are copied), before any kernel reads it; once repaired it serves."""
    A = oracle.poisson_crs(30, 20)
    for bad in (A.nor, -1):
        srv = pw.Server(values, budget_bytes=16 << 20, max_jobs=2)
        rs, ci, dv = A.row_start.astype(np.int32), A.col_ind.astype(np.int32), A.data.copy()
        mat = srv.add_matrix(rs, ci, dv, borrow=True)
        held = srv._held[mat][1]
        held[len(held) // 2] = bad
        j = srv.submit(mat, 5, -1.0, seed=1)
        with pytest.raises(pw.PwmError, match="column index out of range") as e:
            srv.step()
        assert e.value.code == -1 and not srv.resident(mat)
        held[len(held) // 2] = A.col_ind[len(held) // 2]
        res = _drain(srv, {j: A.nor})  # the queued job runs on the repaired arrays
        assert res[j][2] == 5


def test_many_queued_jobs_match_alone():
    """600 jobs submitted at once on a server with 64 slots: they queue, are
    staged while earlier ones run, and every copy of a job gives the result
    of that job run alone, bit for bit."""
    rng = np.random.default_rng(21)
    A = _powerlaw(rng, 2000)
    combos = [("P", 7, -1.0, 1), ("A", 30, 1e-9, 2), ("Q", 3, -1.0, 3), ("P", 200, 1e-10, 4), ("A", 1, -1.0, 5),
              ("Q", 64, -1.0, 6)]

    def fresh():
        srv = pw.Server("f32", budget_bytes=64 << 20, max_jobs=64, chunk=4)
        ids = {"P": srv.add_poisson(40, 30), "A": srv.add_matrix(A.row_start, A.col_ind, A.data),
               "Q": srv.add_poisson(20, 20)}
        return srv, ids, {"P": 1200, "A": A.nor, "Q": 400}

    alone = []
    for name, L, tol, seed in combos:
        srv, ids, n = fresh()
        j = srv.submit(ids[name], L, tol, seed=seed)
        alone.append(_drain(srv, {j: n[name]})[j])
    srv, ids, n = fresh()
    jobs, which = {}, {}
    for copy in range(100):
        for c, (name, L, tol, seed) in enumerate(combos):
            j = srv.submit(ids[name], L, tol, seed=seed)
            jobs[j] = n[name]
            which[j] = c
    assert srv.stats()["queued"] == 600
    res = {}
    for _ in range(100000):
        srv.step()
        for j in srv.finished():
            res[j] = srv.result(j, jobs[j])
        if len(res) == 600:
            break
    assert len(res) == 600 and srv.stats()["queued"] == 0
    for j, (lam, yn, its, vec) in res.items():
        ref = alone[which[j]]
        assert its == ref[2] and lam == ref[0] and yn == ref[1] and np.array_equal(vec, ref[3])


def test_refused_upload_while_others_run():
    """A borrowed matrix whose columns went out of range is refused while
    another job runs (its upload is staged during that job's chunks): the
    step reports the error, the running job goes on unharmed, and once the
    arrays are repaired the refused job is served."""
    A = oracle.poisson_crs(40, 30)
    srv = pw.Server("f32", budget_bytes=64 << 20, max_jobs=4, chunk=2)
    p = srv.add_poisson(50, 50)
    first = srv.submit(p, 200, -1.0, seed=3)
    srv.step()  # the first job runs
    rs, ci, dv = A.row_start.astype(np.int32), A.col_ind.astype(np.int32), A.data.copy()
    mat = srv.add_matrix(rs, ci, dv, borrow=True)
    held = srv._held[mat][1]
    good = int(held[17])
    held[17] = A.nor + 5
    second = srv.submit(mat, 6, -1.0, seed=4)
    with pytest.raises(pw.PwmError, match="column index out of range") as e:
        srv.step()
    assert e.value.code == -1
    held[17] = good
    res = _drain(srv, {first: 2500, second: A.nor})
    assert res[first][2] == 200 and res[second][2] == 6
    ref = pw.Server("f32", budget_bytes=64 << 20, max_jobs=4, chunk=2)
    p2 = ref.add_poisson(50, 50)
    alone = _drain(ref, {ref.submit(p2, 200, -1.0, seed=3): 2500})
    (lam0, yn0, it0, v0), = alone.values()
    assert res[first][0] == lam0 and np.array_equal(res[first][3], v0)
