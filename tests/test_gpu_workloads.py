"""GPU parity at the benchmark workloads' own sizes, through the C ABI.

- Power-law CRS, N = 4 194 304 (~67M nnz), every format: FP64 y within
  1e-13 relative max-norm of the oracle's sequential CRS::mv
  (Matrix/CRS.hpp:99-109); merge-path split points bit-exact
  (SOA_Implementations/CRS_Merge.hpp:96-112).
- 9-point stencil 2896^2 + 2 random columns per row (~92M nnz), every
  format: FP64 y within 1e-13; FP32 storage within the rounding-error bound
  of an FP32 dot product of the row's length (checked per row) and its
  relative max-norm error against FP64 reported.
- The batched power method on the 512-instance uneven SPD batch (256 Poisson
  grids with sides U{32..1024}, 256 symmetric power-law matrices with N
  log-uniform in [1e3, 2e5]): 30 fixed steps, sampled instances (the largest
  grid and the largest power-law matrix among them) against the oracle's
  powerIteration (Matrix/CRS.hpp:118-130) and its Rayleigh quotient.
- The served stream: requests drawn by the serving workload's law (SpMV
  chains of 1-64 steps and power solves on matrices of 1e4..1e7 rows,
  repeats kept resident, 64 GB budget) against the oracle, and each sampled
  result bit-identical to the same job served alone."""
import numpy as np
import pytest

import oracle

pw = pytest.importorskip("thesis_spmv_amd")
pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24  # unit roundoff of FP32


def _relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def _rq(A, v):
    return float(v @ oracle.crs_mv(A, v) / (v @ v))


FORMATS = [
    ("crs_auto", pw.CRSHip, {}),
    ("crs_vector", pw.CRSHip, {"crs_kernel": 1}),
    ("crs_tiles", pw.CRSHip, {"crs_kernel": 2}),
    ("merge_path", pw.CRSMergeHip, {}),
    ("csb_auto", pw.CSBHip, {}),
    ("csb_beta14_8groups", pw.CSBHip, {"block_bits": 14, "csb_groups": 8, "csb_nt": 1}),
    ("csb_swept", pw.CSBHip, {"csb_layout": 4}),
    ("blockcoh", pw.BlockCOHHip, {}),
]


@pytest.fixture(scope="module")
def powerlaw():
    """The power-law workload: N = 4 194 304, Pareto(1.8) row lengths of mean
    16 capped at 50 000, uniform distinct columns, N(0,1) values; x ~ U[-1,1]
    from seed 0."""
    n = 4194304
    rs, ci, dv = pw.gen_powerlaw(n, alpha=1.8, xm=7.3333, cap=50000, struct_seed=1, value_seed=2)
    A = oracle.CRSMatrix(n, n, rs, ci, dv)
    x = pw.gen_uniform(n, 0)
    return A, x, oracle.crs_mv(A, x)


@pytest.fixture(scope="module")
def stencil9():
    """The blocked workload: 9-point stencil on 2896 x 2896 plus 2 uniform
    off-band columns per row (seed 3); x ~ U[-1,1] from seed 0."""
    side = 2896
    rs, ci, dv = pw.gen_stencil9(side, 2, 3)
    A = oracle.CRSMatrix(side * side, side * side, rs, ci, dv)
    x = pw.gen_uniform(A.nor, 0)
    return A, x, oracle.crs_mv(A, x)


def _load(cls, A, values="f64", **opt):
    m = cls(-1, values, **opt)
    m.loadCRS(A.nor, A.noc, A.row_start, A.col_ind, A.data)
    return m


@pytest.mark.parametrize("name,cls,opt", FORMATS, ids=[f[0] for f in FORMATS])
def test_powerlaw_full_size(powerlaw, name, cls, opt):
    A, x, ref = powerlaw
    print(f"power-law N={A.nor} nnz={A.nnz} format {name}", flush=True)
    m = _load(cls, A, **opt)
    y = np.zeros(A.nor)
    m.mv(x, y)
    err = _relerr(y, ref)
    info = m.info()
    print(f"  {name}: rel.err {err:.3e}; group_size {info['group_size']} block_bits {info['block_bits']} "
          f"non_temporal {info['non_temporal']}", flush=True)
    assert err <= 1e-13
    if cls is pw.CRSMergeHip:
        a, b = m.merge_splits()
        ra, rb = oracle.merge_splits(A, len(a) - 1)
        assert np.array_equal(a, ra) and np.array_equal(b, rb)
    if name == "csb_beta14_8groups":
        assert info["group_size"] == 8 and info["block_bits"] == 14


@pytest.mark.parametrize("name,cls,opt", FORMATS, ids=[f[0] for f in FORMATS])
def test_stencil9_full_size(stencil9, name, cls, opt):
    A, x, ref = stencil9
    print(f"stencil9 N={A.nor} nnz={A.nnz} format {name}", flush=True)
    if name == "csb_swept":
        # 8.4M FP64 rows (67 MB) exceed the LDS of all the chip's CUs together
        # (256 x 160 KB): the swept layout refuses them; FP32 rows fit
        with pytest.raises(pw.PwmError) as e:
            _load(cls, A, **opt)
        assert e.value.code == -8
    else:
        m = _load(cls, A, **opt)
        y = np.zeros(A.nor)
        m.mv(x, y)
        err = _relerr(y, ref)
        print(f"  {name} FP64: rel.err {err:.3e}", flush=True)
        assert err <= 1e-13
        del m
    # FP32 values, vectors and row sums: each row within the first-order
    # error bound of an FP32 dot product of its length over FP32-rounded
    # operands, (len + 4) u sum_j |a_ij x_j|
    m32 = _load(cls, A, "f32", **opt)
    y32 = np.zeros(A.nor)
    m32.mv(x, y32)
    absA = oracle.CRSMatrix(A.nor, A.noc, A.row_start, A.col_ind, np.abs(A.data))
    bound = (np.diff(A.row_start) + 4) * U32 * oracle.crs_mv(absA, np.abs(x))
    over = np.abs(y32 - ref) > bound
    err32 = _relerr(y32, ref)
    print(f"  {name} FP32: rel.err vs FP64 {err32:.3e}; rows over the rounding bound: {int(over.sum())}", flush=True)
    assert not over.any()


# ---------------------------------------------------------------- batch

def _batch_workload():
    """The batch workload's instances, drawn as the benchmark draws them."""
    rng = np.random.default_rng(4)
    sides = rng.integers(32, 1025, 256)
    ns = np.exp(rng.uniform(np.log(1e3), np.log(2e5), 256)).astype(np.int64)
    spd = [pw.gen_spd_powerlaw(int(n), seed=1000 + i) for i, n in enumerate(ns)]
    return sides, spd


@pytest.fixture(scope="module")
def batch_workload():
    return _batch_workload()


def _instance(sides, spd, i):
    if i < len(sides):
        return oracle.poisson_crs(int(sides[i]), int(sides[i]))
    rs, ci, dv = spd[i - len(sides)]
    return oracle.CRSMatrix(len(rs) - 1, len(rs) - 1, rs, ci, dv)


@pytest.mark.parametrize("values,tol,opt", [("f32", 1e-6, {}), ("f64", 1e-10, {}),
                                            ("f32", 1e-6, {"bf16_when_exact": 1, "autotune": 1})],
                         ids=["f32", "f64", "f32-bf16_where_exact-autotuned"])
def test_batch_512_instances(batch_workload, values, tol, opt):
    """The benchmark's 512 instances, K plain steps, lambda of sampled
    instances (the largest grid and power-law instance included) against the
    FP64 oracle; the last case is the benchmark's own configuration."""
    sides, spd = batch_workload
    b = pw.Batch(values, **opt)
    for s in sides:
        b.add_poisson(int(s), int(s))
    for rs, ci, dv in spd:
        b.add_crs(rs, ci, dv)
    n_inst, rows, nnz = b.size()
    print(f"batch {values}: {n_inst} instances, {rows} rows, {nnz} nnz", flush=True)
    assert n_inst == 512
    K = 30
    lam, its, ms = b.solve(K, -1.0, seed=0)
    print(f"  {K} steps in {ms:.1f} ms; schedule {b.schedule()}", flush=True)
    assert np.all(its == K)
    big_side = int(np.argmax(sides))
    big_n = len(sides) + int(np.argmax([len(m[0]) for m in spd]))
    sample = sorted({big_side, big_n, *range(0, 512, 37)})
    for i in sample:
        A = _instance(sides, spd, i)
        x0 = b.start_vector(i, seed=0)
        xr, yr = oracle.crs_power_iteration(A, x0, K - 1)
        ref = _rq(A, xr if (K - 1) % 2 == 0 else yr)
        err = abs(lam[i] - ref) / abs(ref)
        print(f"  instance {i} n={A.nor} nnz={A.nnz}: lambda {lam[i]:.10g} rel.err {err:.2e}", flush=True)
        assert err <= tol, (i, lam[i], ref)


def test_batch_workload_alone_equals_in_the_mix(batch_workload):
    """The largest grid and the largest power-law instance of the 512 give
    bit-identical results alone and in the full batch (same id: tiny
    instances stand in front of it)."""
    sides, spd = batch_workload
    b = pw.Batch("f32")
    for s in sides:
        b.add_poisson(int(s), int(s))
    for rs, ci, dv in spd:
        b.add_crs(rs, ci, dv)
    lam, its, _ = b.solve(400, 1e-7, seed=1)
    for i in (int(np.argmax(sides)), len(sides) + int(np.argmax([len(m[0]) for m in spd]))):
        solo = pw.Batch("f32")
        for _ in range(i):
            solo.add_poisson(2, 2)
        A = _instance(sides, spd, i)
        if i < len(sides):
            solo.add_poisson(int(sides[i]), int(sides[i]))
        else:
            solo.add_crs(A.row_start, A.col_ind, A.data)
        lam1, its1, _ = solo.solve(400, 1e-7, seed=1)
        print(f"instance {i} n={A.nor}: steps {its[i]} / {its1[i]}, lambda {lam[i]!r} / {lam1[i]!r}", flush=True)
        assert its1[i] == its[i] and lam1[i] == lam[i]
        assert np.array_equal(b.vector(i)[0], solo.vector(i)[0])


# ---------------------------------------------------------------- server

def _serve_pool(n_pool=12, max_n=1e7):
    """Matrices of the serving workload: N log-uniform in [1e4, max_n],
    Poisson grids, power-law and stencil9 matrices in turn (seed 5)."""
    rng = np.random.default_rng(5)
    ns = np.exp(rng.uniform(np.log(1e4), np.log(max_n), n_pool)).astype(np.int64)
    pool = []
    for i, n in enumerate(ns):
        kind = ("poisson", "powerlaw", "stencil9")[i % 3]
        if kind == "poisson":
            side = max(2, int(np.sqrt(n)))
            pool.append({"kind": kind, "side": side, "n": side * side})
        elif kind == "powerlaw":
            rs, ci, dv = pw.gen_powerlaw(int(n), alpha=1.8, xm=7.3333, cap=50000, struct_seed=100 + i,
                                         value_seed=200 + i)
            pool.append({"kind": kind, "n": int(n), "crs": (rs, ci, dv)})
        else:
            side = max(3, int(np.sqrt(n)))
            rs, ci, dv = pw.gen_stencil9(side, 2, 300 + i)
            pool.append({"kind": kind, "n": side * side, "crs": (rs, ci, dv)})
    return pool


def _serve_requests(rng, n_req, n_pool):
    """The serving workload's request law: 75% SpMV chains of U{1..64} steps,
    25% power solves (tol 1e-7, at most 1000 steps); 20% reuse a resident
    matrix, the others register a fresh copy."""
    reqs = []
    for i in range(n_req):
        chain = rng.random() < 0.75
        reqs.append({"mat": int(rng.integers(n_pool)), "repeat": bool(rng.random() < 0.2),
                     "steps": int(rng.integers(1, 65)) if chain else 1000, "tol": -1.0 if chain else 1e-7,
                     "seed": 1000 + i})
    return reqs


def _oracle_matrix(p):
    if p["kind"] == "poisson":
        return oracle.poisson_crs(p["side"], p["side"])
    rs, ci, dv = p["crs"]
    return oracle.CRSMatrix(len(rs) - 1, len(rs) - 1, rs, ci, dv)


def _register(srv, p):
    if p["kind"] == "poisson":
        return srv.add_poisson(p["side"], p["side"])
    return srv.add_matrix(*p["crs"], borrow=True)


def test_served_stream_at_workload_law():
    """40 requests of the serving law, arriving four at a time between server
    steps (so that later ones join running ones), under a 64 GB budget."""
    pool = _serve_pool()
    print("pool N: " + ", ".join(f"{p['kind']} {p['n']}" for p in pool), flush=True)
    reqs = _serve_requests(np.random.default_rng(50), 40, len(pool))
    srv = pw.Server("f32", budget_bytes=64 << 30, max_jobs=1024)
    resident, jobs, results = {}, {}, {}
    k = 0
    while k < len(reqs) or jobs:
        for r in reqs[k:k + 4]:
            p = pool[r["mat"]]
            ids = [m for m in resident.get(r["mat"], []) if srv.resident(m)]
            mid = ids[0] if r["repeat"] and ids else _register(srv, p)
            resident.setdefault(r["mat"], []).append(mid)
            jobs[srv.submit(mid, r["steps"], r["tol"], seed=r["seed"])] = r
        k += 4
        srv.step()
        for j in srv.finished():
            r = jobs.pop(j)
            results[j] = (r, srv.result(j, pool[r["mat"]]["n"]))
    assert len(results) == len(reqs)
    # oracle checks on the jobs whose sequential FP64 replay is cheap
    # (steps x nnz below 4e8), chains and solves alike
    checked = 0
    for j, (r, (lam, yn, its, vec)) in sorted(results.items()):
        p = pool[r["mat"]]
        A = _oracle_matrix(p)
        if r["tol"] < 0:
            assert its == r["steps"]
        if its * A.nnz > 4e8:
            continue
        x0 = pw.server_start_vector(A.nor, r["seed"]).astype(np.float32).astype(np.float64)
        xr, yr = oracle.crs_power_iteration(A, x0, its - 1) if its > 1 else (x0, None)
        prev = xr if (its - 1) % 2 == 0 else yr
        ref = _rq(A, prev)
        # the power-law matrices here are not symmetric (N(0,1) values) and
        # a short chain's Rayleigh quotient can sit near 0, where a relative
        # error says nothing about FP32 rounding; the error is therefore
        # measured against max(|lambda|, |v|^T |A| |v| / v^T v), the scale of
        # the rounding errors of the quotient's sums (= lambda for the SPD
        # Poisson grids)
        absA = oracle.CRSMatrix(A.nor, A.noc, A.row_start, A.col_ind, np.abs(A.data))
        av = np.abs(prev)
        scale = max(abs(ref), float(av @ oracle.crs_mv(absA, av) / (prev @ prev)))
        err = abs(lam - ref) / scale
        print(f"  job {j} {p['kind']} n={A.nor} steps {its} (tol {r['tol']}): lambda {lam:.9g}, error / scale "
              f"{err:.2e} (relative {abs(lam - ref) / abs(ref):.2e})", flush=True)
        assert err <= 2e-6, (j, lam, ref, scale)
        checked += 1
    assert checked >= 10
    # alone == in the mix: the three longest-running results, replayed on a
    # fresh server with nothing else in flight
    longest = sorted(results, key=lambda j: -results[j][1][2] * pool[results[j][0]["mat"]]["n"])[:3]
    for j in longest:
        r, (lam, yn, its, vec) = results[j]
        solo = pw.Server("f32", budget_bytes=64 << 30, max_jobs=8)
        mid = _register(solo, pool[r["mat"]])
        js = solo.submit(mid, r["steps"], r["tol"], seed=r["seed"])
        while True:
            solo.step()
            if js in solo.finished():
                break
        lam1, yn1, its1, vec1 = solo.result(js, pool[r["mat"]]["n"])
        print(f"  job {j} n={pool[r['mat']]['n']} steps {its}: alone {its1} steps, lambda {lam1!r} vs {lam!r}",
              flush=True)
        assert its1 == its and lam1 == lam and yn1 == yn and np.array_equal(vec1, vec)
        del solo
