"""GPU parity of the CRS vector kernel (CRSHip) and the merge-path kernel
(CRSMergeHip) against the oracle, through the C ABI.

Bars: integer artefacts (device-built CRS, merge split points) bit-exact;
FP64 y within 1e-13 relative max-norm; golden vectors at the reference's own
Boost tolerances (1e-14 mv, 1e-12 power); eigenvalues within 1e-10."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from conftest import ROOT, boost_close, load_case_matrix

pw = pytest.importorskip("thesis_spmv_amd")
pytestmark = pytest.mark.gpu

FORMATS = [pw.CRSHip, pw.CRSMergeHip]


def _relerr(a, b):
    m = np.max(np.abs(b)) if len(b) else 0.0
    return float(np.max(np.abs(a - b)) / (m if m > 0 else 1.0)) if len(b) else 0.0


def _load(cls, A, threads=-1, values="f64", **opt):
    m = cls(threads, values, **opt)
    m.loadCRS(A.nor, A.noc, A.row_start, A.col_ind, A.data)
    return m


# ---------------------------------------------------------------- goldens

@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
@pytest.mark.parametrize("idx", range(24))
def test_golden(goldens, cls, idx):
    """Every reference test case (Tests/poisson.cpp, Tests/input.cpp) with the
    same threads sweep the reference uses for its parallel classes."""
    case = goldens[idx]
    A, t = load_case_matrix(case)
    print(f"case {case['name']} {case['index_type']} {cls.__name__}", flush=True)
    threads_list = (1, 2, 3, 4, 8, -1)
    for threads in threads_list:
        m = cls(threads)
        if case["matrix"]["kind"] == "poisson":
            m.generatePoissonMatrix(case["matrix"]["m"], case["matrix"]["n"], 1)
        else:
            rows = t.rows.astype(np.uint32) if case["index_type"] == "uint32" else t.rows
            cols = t.cols.astype(np.uint32) if case["index_type"] == "uint32" else t.cols

            class T:  # the reference Triplet's members (Matrix/Triplet.hpp:26-45)
                row_coord, col_coord, data, row_size, col_size = rows, cols, t.vals, t.nor, t.noc
            m.loadFromTriplets(T, 1)
        x = np.ones(A.nor)
        y = np.zeros(A.nor)
        if case["op"] == "mv":
            m.mv(x, y)
            got = y
        else:
            m.powerIteration(x, y, case["iterations"])
            got = x if case["checks"] == "x" else y
        ok = boost_close(got, case["expected"], case["tolerance"])
        assert ok.all(), (case["name"], threads, np.flatnonzero(~ok)[:5])


# ---------------------------------------------------------------- integer artefacts

@pytest.mark.parametrize("m,n", [(1, 1), (3, 3), (9, 3), (1, 17), (17, 1), (64, 48), (1024, 1024)])
def test_device_poisson_bit_exact(m, n):
    A = oracle.poisson_crs(m, n)
    h = pw.CRSHip()
    h.generatePoissonMatrix(m, n, 1)
    rs, ci, dv = h.crs_export()
    assert np.array_equal(rs, A.row_start) and np.array_equal(ci, A.col_ind) and np.array_equal(dv, A.data)


def test_triplets_to_crs_bit_exact():
    rng = np.random.default_rng(1)
    N = 3000
    r = rng.integers(0, N - 50, 40000).astype(np.int32)
    c = rng.integers(0, N, 40000).astype(np.int32)
    v = rng.standard_normal(40000)
    t = oracle.Triplets(N, N, r, c, v)
    A = oracle.triplet_to_crs(t)
    h = pw.CRSHip()
    h.loadFromTriplets(t)
    rs, ci, dv = h.crs_export()
    assert np.array_equal(rs, A.row_start) and np.array_equal(ci, A.col_ind) and np.array_equal(dv, A.data)


def _powerlaw(n, seed=1, cap=50000):
    rs, ci, dv = pw.gen_powerlaw(n, struct_seed=seed, value_seed=seed + 1, cap=cap)
    return oracle.CRSMatrix(n, n, rs, ci, dv)


@pytest.mark.parametrize("values", ["f64", "f32"])
def test_triplets_device_sort_large_shuffled(values):
    """A 1.8M-entry power-law matrix as shuffled COO with 5000 duplicated
    entries (new values, so their order shows): the device sort's CRS arrays
    equal the oracle's TripletToCRS, for every format that keeps CRS."""
    rng = np.random.default_rng(7)
    A = _powerlaw(120000, cap=20000)
    rows = np.repeat(np.arange(A.nor, dtype=np.int32), np.diff(A.row_start))
    dup = rng.integers(0, A.nnz, 5000)
    r = np.concatenate([rows, rows[dup]])
    c = np.concatenate([A.col_ind, A.col_ind[dup]])
    v = np.concatenate([A.data, rng.standard_normal(5000)])
    perm = rng.permutation(len(r))
    t = oracle.Triplets(A.nor, A.noc, r[perm], c[perm], v[perm])
    ref = oracle.triplet_to_crs(t)
    want = ref.data if values == "f64" else ref.data.astype(np.float32).astype(np.float64)
    for cls in (pw.CRSHip, pw.CRSMergeHip):
        h = cls(-1, values)
        h.loadFromTriplets(t)
        rs, ci, dv = h.crs_export()
        assert np.array_equal(rs, ref.row_start) and np.array_equal(ci, ref.col_ind), cls.__name__
        assert np.array_equal(dv, want), cls.__name__
        x = pw.gen_uniform(A.noc, 2).astype(np.float64)
        y = np.zeros(A.nor)
        h.mv(x, y)
        tol = 1e-13 if values == "f64" else 1e-5
        assert _relerr(y, oracle.crs_mv(ref, x)) <= tol


def test_triplets_device_sort_edge_shapes():
    """Empty matrices, 1 x 1, a single column (0-bit column keys), all
    entries in the last row, a wide matrix; rows given out of order."""
    cases = [(5, 7, [], [], []),
             (1, 1, [0], [0], [3.0]),
             (6, 1, [5, 0, 5, 2], [0, 0, 0, 0], [1.0, 2.0, 3.0, 4.0]),
             (9, 4, [8, 8, 8, 8, 8], [3, 1, 2, 1, 0], [1.0, 2.0, 3.0, 4.0, 5.0]),
             (3, 70000, [2, 0, 2, 1, 0], [69999, 5, 0, 40000, 5], [1.0, -2.0, 3.0, 4.0, 0.5])]
    for nor, noc, r, c, v in cases:
        t = oracle.Triplets(nor, noc, r, c, v)
        ref = oracle.triplet_to_crs(t)
        h = pw.CRSHip()
        h.loadFromTriplets(t)
        rs, ci, dv = h.crs_export()
        assert np.array_equal(rs, ref.row_start) and np.array_equal(ci[:ref.nnz], ref.col_ind), (nor, noc)
        assert np.array_equal(dv[:ref.nnz], ref.data), (nor, noc)


def test_triplets_errors_name_the_first_bad_entry():
    h = pw.CRSHip()
    bad = oracle.Triplets(4, 4, [0, 1, 2, 4, 9], [0, 0, 0, 0, 0], [1.0] * 5)
    with pytest.raises(pw.PwmError, match="entry 3"):
        h.loadFromTriplets(bad)
    neg = oracle.Triplets(4, 4, [0, 1, 2, 3], [0, -1, 0, 7], [1.0] * 4)
    with pytest.raises(pw.PwmError, match="entry 1"):
        h.loadFromTriplets(neg)
    hb = pw.CRSHip(-1, "bf16")
    with pytest.raises(pw.PwmError, match="exact in BF16"):
        hb.loadFromTriplets(oracle.Triplets(3, 3, [0, 1, 2], [0, 1, 2], [4.0, -1.0, 0.1]))
    hb.loadFromTriplets(oracle.Triplets(3, 3, [2, 1, 0], [0, 1, 2], [4.0, -1.0, 0.5]))
    assert hb.info()["nnz"] == 3


@pytest.mark.parametrize("threads", [1, 2, 7, 64, 1000, -1])
def test_merge_splits_bit_exact(threads):
    A = _powerlaw(20000)
    m = _load(pw.CRSMergeHip, A, threads)
    a, b = m.merge_splits()
    P = len(a) - 1
    if threads > 0 and -(-(A.nor + A.nnz) // threads) <= 256 * 8:
        assert P == threads
    ra, rb = oracle.merge_splits(A, P)
    assert np.array_equal(a, ra) and np.array_equal(b, rb)


# ---------------------------------------------------------------- FP64 SpMV

@pytest.mark.parametrize("G", [1, 2, 4, 8, 16, 32, 64])
def test_vector_kernel_group_sizes(G):
    A = _powerlaw(30000, cap=3000)
    x = pw.gen_uniform(A.noc, 0)
    ref = oracle.crs_mv(A, x)
    for long_row in (0, 40):  # default threshold, and most rows routed to the chunked path
        m = _load(pw.CRSHip, A, group_size=G, long_row=long_row, long_chunk=256)
        y = np.zeros(A.nor)
        m.mv(x, y)
        assert _relerr(y, ref) <= 1e-13, (G, long_row)


@pytest.mark.parametrize("ipt", [4, 8, 12, 16])
def test_merge_items_per_thread(ipt):
    A = _powerlaw(30000, cap=20000)
    x = pw.gen_uniform(A.noc, 0)
    ref = oracle.crs_mv(A, x)
    m = _load(pw.CRSMergeHip, A, items_per_thread=ipt)
    y = np.zeros(A.nor)
    m.mv(x, y)
    assert _relerr(y, ref) <= 1e-13


@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
def test_edge_shapes(cls):
    cases = []
    # one row, one column
    cases.append(oracle.CRSMatrix(1, 1, [0, 1], [0], [2.5]))
    # all rows empty
    cases.append(oracle.CRSMatrix(5, 5, [0] * 6, [], []))
    # leading / trailing empty rows and a single very long row
    rs = [0, 0, 0, 4000, 4000, 4001, 4001]
    rng = np.random.default_rng(0)
    cases.append(oracle.CRSMatrix(6, 9000, rs, rng.integers(0, 9000, 4001), rng.standard_normal(4001)))
    # rectangular (wide and tall)
    cases.append(oracle.CRSMatrix(3, 700, [0, 300, 300, 700], np.sort(rng.integers(0, 700, 700)),
                                  rng.standard_normal(700)))
    cases.append(oracle.CRSMatrix(900, 2, np.arange(901) // 3 * 1, np.zeros(300, np.int32), np.ones(300)))
    for A in cases:
        x = rng.standard_normal(A.noc)
        ref = oracle.crs_mv(A, x)
        for threads in (1, 4, -1):
            m = _load(cls, A, threads)
            y = np.full(A.nor, np.nan)
            m.mv(x, y)
            assert _relerr(y, ref) <= 1e-13


def test_device_vectors_and_errors():
    A = oracle.poisson_crs(40, 30)
    m = _load(pw.CRSHip, A)
    x = pw.DeviceVector.from_array(np.ones(A.nor))
    y = pw.DeviceVector(A.nor)
    m.mv_device(x, y)
    assert np.array_equal(y.download(), oracle.crs_mv(A, np.ones(A.nor)))
    with pytest.raises(pw.PwmError) as e:
        m.mv_device(x, x)  # aliasing
    assert e.value.code == -1
    # power method on a rectangular matrix
    R = oracle.CRSMatrix(3, 5, [0, 1, 2, 3], [0, 1, 4], [1.0, 2.0, 3.0])
    mr = _load(pw.CRSHip, R)
    with pytest.raises(pw.PwmError) as e:
        mr.powerIteration(np.ones(3), np.zeros(3), 3)
    assert e.value.code == -5
    # out-of-range column index, bad row_start
    with pytest.raises(pw.PwmError):
        _load(pw.CRSHip, oracle.CRSMatrix(2, 2, [0, 1, 2], [0, 2], [1.0, 1.0]))
    with pytest.raises(pw.PwmError):
        pw.CRSHip().loadCRS(2, 2, [0, 2, 1], [0, 1], [1.0, 1.0])
    with pytest.raises(pw.PwmError, match="column index"):
        _load(pw.CRSHip, oracle.CRSMatrix(2, 2, [0, 1, 2], [0, -1], [1.0, 1.0]))
    # one bad index deep inside a large array (the check runs in parallel chunks)
    n = 3_000_000
    ci_big = (np.arange(n) % n).astype(np.int32)
    ci_big[2_500_001] = n
    with pytest.raises(pw.PwmError, match="column index"):
        pw.CRSHip().loadCRS(n, n, np.arange(n + 1, dtype=np.int32), ci_big, np.ones(n))
    # calls before a load
    with pytest.raises(pw.PwmError) as e:
        pw.CRSHip().mv(np.ones(2), np.zeros(2))
    assert e.value.code == -6
    assert pw.select_type(7) is None


# ---------------------------------------------------------------- power method

@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
@pytest.mark.parametrize("it", [0, 1, 2, 3, 10, 101])
def test_power_iteration_state_matches_reference(cls, it):
    """Both x and y are left as the reference leaves them (Matrix/CRS.hpp:118-130)."""
    A = _powerlaw(5000, cap=500)
    A.data[:] = np.abs(A.data)  # nonnegative: a dominant real eigenvalue
    x0 = pw.gen_uniform(A.nor, 3)
    xr, yr = oracle.crs_power_iteration(A, x0, it)
    m = _load(cls, A)
    x, y = x0.copy(), np.full(A.nor, 7.0)
    m.powerIteration(x, y, it)
    if it == 0:
        assert np.array_equal(x, x0) and np.all(y == 7.0)
        return
    yr_cmp = yr if it >= 1 else np.full(A.nor, 7.0)
    assert _relerr(x, xr) <= 1e-12 and _relerr(y, yr_cmp) <= 1e-12


@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
def test_power_method_eigenvalue(cls):
    """Poisson 1024x1024 (config 0): 100 fixed iterations, lambda vs the FP64
    oracle's Rayleigh quotient at the same iteration (1e-10), and a small
    grid run to convergence against the closed form."""
    mm = 1024
    A = oracle.poisson_crs(mm, mm)
    x0 = pw.gen_uniform(A.nor, 0)
    m = cls()
    m.generatePoissonMatrix(mm, mm, 1)
    x = pw.DeviceVector.from_array(x0)
    y = pw.DeviceVector(A.nor)
    lam, ynorm, its = m.power_method_device(x, y, 100, -1.0)
    assert its == 100
    xr, yr = oracle.crs_power_iteration(A, x0, 99)
    v = yr  # iterate 99 is in y (odd)
    rq = float(v @ oracle.crs_mv(A, v) / (v @ v))
    assert abs(lam - rq) <= 1e-10 * abs(rq), (lam, rq)
    assert abs(ynorm - np.linalg.norm(oracle.crs_mv(A, v))) <= 1e-10 * ynorm
    assert _relerr(x.download(), oracle.crs_power_iteration(A, x0, 100)[0]) <= 1e-12
    # converge on 12x9 Poisson (random start: for an even side the all-ones
    # vector is orthogonal to the dominant eigenvector)
    m2 = cls()
    m2.generatePoissonMatrix(12, 9, 1)
    x2 = pw.DeviceVector.from_array(pw.gen_uniform(108, 5))
    y2 = pw.DeviceVector(108)
    lam2, _, its2 = m2.power_method_device(x2, y2, 5000, 1e-13)
    exact = 4 + 2 * np.cos(np.pi / 13) + 2 * np.cos(np.pi / 10)
    assert its2 < 5000 and abs(lam2 - exact) <= 1e-10 * exact


# ---------------------------------------------------------------- mixed precision

@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
@pytest.mark.parametrize("values", ["f32", "bf16"])
def test_mixed_precision_error_reported(cls, values):
    m = cls(-1, values)
    m.generatePoissonMatrix(300, 200, 1)
    A = oracle.poisson_crs(300, 200)
    x0 = pw.gen_uniform(A.nor, 0)
    y = np.zeros(A.nor)
    m.mv(x0, y)
    err = _relerr(y, oracle.crs_mv(A, x0))
    print(f"{cls.__name__} {values} SpMV error vs FP64 oracle: {err:.3e}", flush=True)
    assert err <= 1e-6
    xd = pw.DeviceVector.from_array(x0, np.float32)
    yd = pw.DeviceVector(A.nor, np.float32)
    lam, _, _ = m.power_method_device(xd, yd, 200, -1.0)
    m64 = cls()
    m64.generatePoissonMatrix(300, 200, 1)
    lam64, _, _ = m64.power_method_device(pw.DeviceVector.from_array(x0), pw.DeviceVector(A.nor), 200, -1.0)
    print(f"{cls.__name__} {values} eigenvalue rel. error vs FP64: {abs(lam - lam64) / lam64:.3e}", flush=True)
    assert abs(lam - lam64) <= 1e-6 * lam64


@pytest.mark.parametrize("values", ["f32", "bf16"])
def test_mixed_precision_layout_variants(values):
    """The narrow-storage SpMV and power method of the non-default layouts:
    CSB slab pieces and wide slab streams, BlockCOH built on the host."""
    A = oracle.poisson_crs(150, 120)
    x0 = pw.gen_uniform(A.nor, 2)
    ref = oracle.crs_mv(A, x0)
    lam64 = None
    for cls, opts in ((pw.CSBHip, {"csb_layout": 1}), (pw.CSBHip, {"csb_layout": 3}),
                      (pw.CSBHip, {"csb_layout": 3, "csb_groups": 8}), (pw.BlockCOHHip, {"coh_build": 1})):
        m = cls(-1, values, **opts)
        m.generatePoissonMatrix(150, 120, 1)
        y = np.zeros(A.nor)
        m.mv(x0, y)
        assert _relerr(y, ref) <= 1e-6, (cls.__name__, opts)
        lam, _, _ = m.power_method_device(pw.DeviceVector.from_array(x0, np.float32),
                                          pw.DeviceVector(A.nor, np.float32), 150, -1.0)
        if lam64 is None:
            m64 = pw.CRSHip()
            m64.generatePoissonMatrix(150, 120, 1)
            lam64, _, _ = m64.power_method_device(pw.DeviceVector.from_array(x0), pw.DeviceVector(A.nor), 150, -1.0)
        assert abs(lam - lam64) <= 1e-6 * lam64, (cls.__name__, opts, lam, lam64)


def test_bf16_rejects_inexact_values():
    A = oracle.CRSMatrix(2, 2, [0, 1, 2], [0, 1], [1.0, 0.1])
    with pytest.raises(pw.PwmError):
        _load(pw.CRSHip, A, values="bf16")


def test_cpp_dropin_classes():
    exe = os.path.join(ROOT, "tests", "cpp", "test_classes")
    if not os.path.exists(exe):
        pytest.skip("C++ driver not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


# ---------------------------------------------------------------- CRS kernels
# The CRS format runs either the vector kernel (lanes per row) or the
# row-aligned tiles kernel (tiles of whole short rows, long rows split with
# carries); by default it times both at load, the tiles kernel with plain and
# with non-temporal entry loads.  Each is checked on its own.

@pytest.mark.parametrize("kernel,nt", [(1, -1), (2, 0), (2, 1)])
@pytest.mark.parametrize("idx", range(24))
def test_golden_each_crs_kernel(goldens, kernel, nt, idx):
    case = goldens[idx]
    A, t = load_case_matrix(case)
    m = pw.CRSHip(-1, "f64", crs_kernel=kernel, crs_nt=nt)
    if case["matrix"]["kind"] == "poisson":
        m.generatePoissonMatrix(case["matrix"]["m"], case["matrix"]["n"], 1)
    else:
        m.loadCRS(A.nor, A.noc, A.row_start, A.col_ind, A.data)
    x, y = np.ones(A.nor), np.zeros(A.nor)
    if case["op"] == "mv":
        m.mv(x, y)
        got = y
    else:
        m.powerIteration(x, y, case["iterations"])
        got = x if case["checks"] == "x" else y
    ok = boost_close(got, case["expected"], case["tolerance"])
    assert ok.all(), (case["name"], kernel, np.flatnonzero(~ok)[:5])


def _long_row_matrix(rng, n=16000):
    lens = rng.integers(0, 9, n)
    lens[[3, 1000, 2500, n - 1]] = [9000, 30, 4100, 12000]  # rows longer than a tile, one at the end
    lens[[10, 11, 12]] = 0
    lens = np.minimum(lens, n)
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, int(L), replace=False)) for L in lens]).astype(np.int32)
    return oracle.CRSMatrix(n, n, rs, ci, rng.standard_normal(int(rs[-1])))


@pytest.mark.parametrize("kernel,nt", [(1, -1), (2, 0), (2, 1)])
@pytest.mark.parametrize("values,tol", [("f64", 1e-13), ("f32", 5e-5), ("bf16", 5e-5)])
def test_crs_kernels_long_rows(kernel, nt, values, tol):
    rng = np.random.default_rng(17)
    A = _long_row_matrix(rng)
    if values == "bf16":
        A = oracle.CRSMatrix(A.nor, A.noc, A.row_start, A.col_ind, np.round(A.data * 8) / 8)
    m = _load(pw.CRSHip, A, values=values, crs_kernel=kernel, crs_nt=nt)
    info = m.info()
    assert (info["group_size"] == 0) == (kernel == 2)
    x = rng.uniform(-1, 1, A.nor)
    if values != "f64":
        x = x.astype(np.float32).astype(np.float64)
    y = np.zeros(A.nor)
    m.mv(x, y)
    assert _relerr(y, oracle.crs_mv(A, x)) <= tol


@pytest.mark.parametrize("kernel,nt", [(1, -1), (2, 0), (2, 1)])
def test_crs_kernels_power_method(kernel, nt):
    rng = np.random.default_rng(5)
    A = _long_row_matrix(rng, 4000)
    A = oracle.CRSMatrix(A.nor, A.noc, A.row_start, A.col_ind, np.abs(A.data) + 0.1)
    m = _load(pw.CRSHip, A, crs_kernel=kernel, crs_nt=nt)
    x0 = rng.uniform(0, 1, A.nor)
    xd, yd = pw.DeviceVector.from_array(x0), pw.DeviceVector(A.nor)
    lam, yn, its = m.power_method_device(xd, yd, 40, -1.0)
    xr, yr = oracle.crs_power_iteration(A, x0, 39)
    v = xr if 39 % 2 == 0 else yr
    ref = float(v @ oracle.crs_mv(A, v) / (v @ v))
    assert its == 40 and abs(lam - ref) <= 1e-10 * abs(ref)
    x, y = x0.copy(), np.zeros(A.nor)
    m.powerIteration(x, y, 7)
    xr, yr = oracle.crs_power_iteration(A, x0, 7)
    assert _relerr(y, yr) <= 1e-12 and _relerr(x, xr) <= 1e-12


def test_crs_default_times_both_kernels():
    A = oracle.poisson_crs(300, 300)
    m = _load(pw.CRSHip, A)
    assert m.info()["group_size"] in (0, 2, 4)  # whichever kernel was faster


def test_gather_rate_probe():
    """pwm_gather_rate runs on any FP64 vector and rejects windows that do not fit."""
    x = pw.DeviceVector.from_array(np.ones(1 << 16))
    assert pw.gather_rate(x, 1 << 13, True, 1 << 16) > 0
    assert pw.gather_rate(x, 1 << 16, False, 3 * 8192 + 5) > 0
    assert pw.gather_rate(None, 0, False, 1 << 16) > 0  # the stream alone
    with pytest.raises(pw.PwmError):
        pw.gather_rate(x, 1 << 14, True, 1 << 16)  # eight windows exceed x
    with pytest.raises(pw.PwmError):
        pw.gather_rate(x, 1 << 13, True, 100)  # fewer entries than one workgroup


@pytest.mark.parametrize("cls", [pw.CRSHip, pw.CRSMergeHip, pw.CSBHip, pw.BlockCOHHip], ids=lambda c: c.__name__)
def test_time_mv_back_to_back(cls):
    """pwm_time_mv without the kernel time runs the SpMVs back to back and
    leaves the SpMV's result in y, as with per-step events."""
    A = oracle.poisson_crs(120, 90)
    m = cls()
    m.generatePoissonMatrix(120, 90, 1)
    xh = pw.gen_uniform(A.noc, 3)
    x, y = pw.DeviceVector.from_array(xh), pw.DeviceVector(A.nor)
    ms, kms = m.time_mv(x, y, 7, per_step=False)
    assert ms > 0 and kms is None
    ref = oracle.crs_mv(A, xh)
    assert np.max(np.abs(y.download() - ref)) <= 1e-13 * np.max(np.abs(ref))
    ms2, kms2 = m.time_mv(x, y, 3)
    assert ms2 > 0 and 0 < kms2 <= ms2 * 1.5


def test_time_mv_rotating_results_and_kernel_counts():
    """pwm_time_mv_rotating runs one SpMV of each handle in turn on the first
    handle's stream: every y holds its own matrix's product afterwards, a
    handle may appear twice, and each handle keeps working on its own stream.
    pwm_info's kernel count matches the launches of each format's SpMV."""
    mats, xs, ys, refs = [], [], [], []
    for k, (cls, opt) in enumerate([(pw.CRSHip, {"crs_kernel": 2}), (pw.CRSMergeHip, {}),
                                    (pw.CSBHip, {"csb_layout": 4}), (pw.CSBHip, {"csb_groups": 8}),
                                    (pw.BlockCOHHip, {})]):
        rs, ci, dv = pw.gen_powerlaw(40000, struct_seed=20 + k, value_seed=30 + k, cap=2000)
        A = oracle.CRSMatrix(40000, 40000, rs, ci, dv)
        m = _load(cls, A, **opt)
        xh = pw.gen_uniform(A.noc, 40 + k)
        mats.append(m)
        xs.append(pw.DeviceVector.from_array(xh))
        ys.append(pw.DeviceVector(A.nor))
        refs.append(oracle.crs_mv(A, xh))
    kernels = [m.info()["kernels"] for m in mats]
    print(f"kernels per SpMV: {kernels}", flush=True)
    assert kernels[1] == 2 and kernels[2] == 1 and kernels[3] == 2 and kernels[0] in (1, 2) and kernels[4] in (1, 2)
    for per_step in (True, False):
        for y in ys:
            y.upload(np.full(y.n, np.nan))
        ms, kms = pw.time_mv_rotating(mats + [mats[0]], xs + [xs[0]], ys + [ys[0]], 3, per_step=per_step)
        assert ms > 0 and ((kms is None) != per_step)
        for y, ref in zip(ys, refs):
            assert np.max(np.abs(y.download() - ref)) <= 1e-13 * np.max(np.abs(ref))
    for m, x, y, ref in zip(mats, xs, ys, refs):  # own streams again
        y.upload(np.zeros(y.n))
        m.mv_device(x, y)
        assert np.max(np.abs(y.download() - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_info_reports_tuned_choices():
    """pwm_info reports the kernel variant a handle runs: the options the
    caller fixed, and a valid choice when the load-time tuner picked it."""
    rs, ci, dv = pw.gen_powerlaw(50000, struct_seed=3, value_seed=4, cap=3000)
    A = oracle.CRSMatrix(50000, 50000, rs, ci, dv)  # CSB: 25 block columns of 2^11, slab streams
    for nt in (0, 1):
        inf = _load(pw.CRSHip, A, crs_kernel=2, crs_nt=nt).info()
        assert inf["group_size"] == 0 and inf["non_temporal"] == nt
        inf = _load(pw.CSBHip, A, csb_groups=8, csb_unroll=4, csb_nt=nt).info()
        assert (inf["group_size"], inf["unroll"], inf["non_temporal"]) == (8, 4, nt)
        inf = _load(pw.BlockCOHHip, A, coh_occ8=nt).info()
        assert inf["high_occupancy"] == nt
    inf = _load(pw.CSBHip, A).info()
    assert inf["unroll"] in (2, 4) and inf["non_temporal"] in (0, 1) and inf["group_size"] in (1, 8, 16)
    inf = _load(pw.CRSHip, A, crs_kernel=1).info()
    assert inf["group_size"] > 0 and inf["non_temporal"] == 0


def test_nnz_limit_rejected_before_reading():
    """A matrix over PWM_MAX_NNZ = 2^31 - 65536 nonzeros is refused with
    PWM_ERR_INVALID before any array is read or allocated."""
    max_nnz = (1 << 31) - 65536
    m = pw.CRSHip()
    tiny = np.zeros(4, np.int32)
    rc = pw.lib().pwm_load_triplets(m.h, 10, 10, max_nnz + 1, tiny.ctypes.data, tiny.ctypes.data,
                                    np.zeros(4).ctypes.data, 1)
    assert rc == -1 and b"PWM_MAX_NNZ" in pw.lib().pwm_last_error(m.h)
    assert pw.lib().pwm_generate_poisson(m.h, 20724, 20724, 1) == -1   # 5 m n just above the limit
    m.generatePoissonMatrix(40, 30, 1)                                  # the handle stays usable
    assert m.info()["nnz"] == oracle.poisson_crs(40, 30).nnz


@pytest.mark.parametrize("values", ["f64", "f32"])
@pytest.mark.parametrize("n", [5, 262_147, 1_048_579, 3_000_017])
def test_host_array_mv_equals_device_mv(values, n):
    """The drop-in mv(x, y) on host arrays moves x and y through the pinned
    staging buffers in chunks (a ragged last chunk for these sizes); its y must
    equal, bit for bit, the same kernel run on device vectors, and the
    caller's x must be left as it was."""
    rs, ci, dv = pw.gen_powerlaw(n, alpha=1.8, xm=3.0, cap=200, struct_seed=7, value_seed=8)
    m = pw.CRSHip(-1, values)
    m.loadCRS(n, n, rs, ci, dv)
    x = pw.gen_uniform(n, 3)
    x0 = x.copy()
    y = np.full(n, np.nan)
    m.mv(x, y)
    assert np.array_equal(x, x0)
    dt = np.float64 if values == "f64" else np.float32
    xd = pw.DeviceVector.from_array(x.astype(dt), dtype=dt)
    yd = pw.DeviceVector(n, dtype=dt)
    m.mv_device(xd, yd)
    assert np.array_equal(y, yd.download().astype(np.float64))
    # a power iteration through the host path ends where the device vectors do
    xa, ya = x.copy(), np.zeros(n)
    m.powerIteration(xa, ya, 3)
    assert np.all(np.isfinite(xa)) and np.all(np.isfinite(ya))


# ---------------------------------------------------------------- device pointers

def _hip_runtime():
    """The HIP runtime libpwmhip.so runs on (already loaded in this process)."""
    import ctypes
    pw.lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    assert paths, "libamdhip64 not loaded"
    return ctypes.CDLL(sorted(paths)[0])


class _Ptr:
    def __init__(self, p):
        self.ptr = p


def test_device_entry_points_refuse_host_and_short_buffers():
    """pwm_mv_device / pwm_power_method_device / pwm_time_mv take hipMalloc
    device memory only: pinned host memory (hipHostMalloc), pageable host
    memory and a device buffer shorter than the vector are refused with
    PWM_ERR_INVALID before any kernel runs; y keeps its contents."""
    import ctypes
    A = oracle.poisson_crs(512, 256)
    x = pw.DeviceVector.from_array(np.ones(A.nor))
    y = pw.DeviceVector.from_array(np.full(A.nor, 7.0))
    hip = _hip_runtime()
    pinned = ctypes.c_void_p()
    assert hip.hipHostMalloc(ctypes.byref(pinned), ctypes.c_size_t(A.nor * 8), ctypes.c_uint(0)) == 0
    host = np.zeros(A.nor)
    # the second half of a device buffer of nor doubles: half a vector short
    buf = pw.DeviceVector(A.nor)
    short = _Ptr(ctypes.c_void_p(buf.ptr.value + A.nor * 4))
    try:
        for cls in (pw.CRSHip, pw.CRSMergeHip, pw.CSBHip, pw.BlockCOHHip):
            m = _load(cls, A)
            for bad in (_Ptr(pinned), _Ptr(host.ctypes.data_as(ctypes.c_void_p)), short):
                for call in (lambda: m.mv_device(x, bad), lambda: m.mv_device(bad, y),
                             lambda: m.power_method_device(x, bad, 3),
                             lambda: m.time_mv(x, bad, 1)):
                    with pytest.raises(pw.PwmError) as e:
                        call()
                    assert e.value.code == -1, (cls.__name__, str(e.value))
            assert np.all(y.download() == 7.0)
            # device memory still works
            m.mv_device(x, y)
            assert np.array_equal(y.download(), oracle.crs_mv(A, np.ones(A.nor)))
            y.fill(7.0)
    finally:
        hip.hipHostFree(pinned)


def test_poisson_vector_kernel_long_row_threshold_below_row_length():
    """A generated Poisson matrix on the vector kernel with the long-row
    threshold under the row length: every row goes through the long-row
    path and still matches the oracle."""
    for m_, n_ in ((7, 5), (64, 48), (300, 200)):
        A = oracle.poisson_crs(m_, n_)
        for lr in (1, 3, 4):
            m = pw.CRSHip(-1, "f64", crs_kernel=1, long_row=lr)
            m.generatePoissonMatrix(m_, n_, 1)
            assert m.info()["long_rows"] > 0
            x = pw.gen_uniform(A.nor, 3)
            y = np.full(A.nor, np.nan)
            m.mv(x, y)
            assert _relerr(y, oracle.crs_mv(A, x)) <= 1e-13, (m_, n_, lr)
            xs, ys = np.ones(A.nor), np.zeros(A.nor)
            m.powerIteration(xs, ys, 6)
            xr, yr = oracle.crs_power_iteration(A, np.ones(A.nor), 6)
            assert np.max(np.abs(xs - xr)) <= 1e-12 and np.max(np.abs(ys - yr)) <= 1e-12



# ---------------------------------------------------------------- row partitions

@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
def test_triplet_to_multiple_crs_partitions(cls):
    """TripletToMultipleCRS (Util/TripletToCRS.hpp:309-385) of a loaded
    handle: the floor-rule row plan and every partition's local CRS arrays
    bit-exact with the oracle's, for triplets with duplicates and empty rows,
    and more partitions than rows."""
    rng = np.random.default_rng(12)
    nor, noc, nnz = 3001, 2500, 40000
    r = rng.integers(0, nor, nnz).astype(np.int32)
    r[(r % 11) == 5] = 17
    c = rng.integers(0, noc, nnz).astype(np.int32)
    r[:40], c[:40] = 9, 11
    v = rng.standard_normal(nnz)
    t = oracle.Triplets(nor, noc, r, c, v)
    m = cls()
    m.loadFromTriplets(t)
    for P in (1, 2, 3, 16, 255, nor + 5):
        ref = oracle.triplet_to_multiple_crs(t, P)
        f, rows, z = m.partitions(P)
        fr = np.zeros(P, np.int32)
        tr = np.zeros(P, np.int32)
        oracle.lib().orc_row_partitions(nor, P, fr, tr)
        assert np.array_equal(f, fr) and np.array_equal(rows, tr)
        assert z.tolist() == [A.nnz for _, A in ref]
        for p in sorted({0, 1 % P, P // 2, P - 1}):
            rs, ci, dv = m.partition_crs(P, p)
            A = ref[p][1]
            assert np.array_equal(rs, A.row_start) and np.array_equal(ci, A.col_ind) and np.array_equal(dv, A.data), (P, p)
