"""Pins the CPU oracle (oracle/pwm_oracle.c) before anything is checked
against it: the reference's MATLAB golden vectors, closed forms, published
known answers and brute force on tiny inputs.  CPU only."""
import itertools
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import oracle
from conftest import boost_close, fixture_path, load_case_matrix

# ------------------------------------------------------------------ goldens


def _golden_ids(goldens):
    return [c["name"] + "_" + c["index_type"] for c in goldens]


def test_golden_count(goldens):
    # 12 distinct cases, each duplicated for unsigned indices (Tests/*.cpp)
    assert len(goldens) == 24
    assert sum(c["op"] == "mv" for c in goldens) == 16


@pytest.mark.parametrize("idx", range(24))
def test_crs_matches_matlab_golden(goldens, idx):
    case = goldens[idx]
    A, _ = load_case_matrix(case)
    x = np.ones(A.nor)
    if case["op"] == "mv":
        got = oracle.crs_mv(A, x)
    else:
        xo, yo = oracle.crs_power_iteration(A, x, case["iterations"])
        got = xo if case["checks"] == "x" else yo
    ok = boost_close(got, case["expected"], case["tolerance"])
    assert ok.all(), f"{case['name']}: {np.count_nonzero(~ok)} entries off, e.g. {np.flatnonzero(~ok)[:5]}"


@pytest.mark.parametrize("idx", [i for i in range(24)])
def test_every_format_matches_golden(goldens, idx):
    """The reference runs CRSOMP, CRS_Merge, CSB and BlockCOH with 1..maxT
    threads against the same goldens (Tests/GetMatrices.hpp:28-42)."""
    case = goldens[idx]
    if case["index_type"] != "int32":
        pytest.skip("unsigned duplicate: same data")
    A, t = load_case_matrix(case)
    if t is None:
        t = oracle.poisson_triplets(case["matrix"]["m"], case["matrix"]["n"])
    x = np.ones(A.nor)
    tol = case["tolerance"]

    def run(mv):
        if case["op"] == "mv":
            return mv(x)
        # power iteration restated on top of the format's mv (CRS.hpp:118-130)
        v, w = x.copy(), np.zeros(A.nor)
        for i in range(case["iterations"]):
            if i % 2 == 0:
                w = mv(v); w = w / oracle.norm2(w)
            else:
                v = mv(w); v = v / oracle.norm2(v)
        return v if case["checks"] == "x" else w

    for P in (1, 2, 3, 5, 8, 16):
        got = run(lambda v: oracle.merge_mv(A, v, P))
        assert boost_close(got, case["expected"], tol).all(), ("merge", P)
        got = run(lambda v: oracle.crsomp_mv(A, v, min(P, 8)))
        assert boost_close(got, case["expected"], tol).all(), ("crsomp", P)
    bits, ind, data, blk, bal = oracle.csb_build(t)
    got = run(lambda v: oracle.csb_mv(A.nor, A.noc, bits, ind, data, blk, v))
    assert boost_close(got, case["expected"], tol).all(), "csb"
    for th in (1, 2, 3, 4, 8):
        bands = oracle.blockcoh_build(A, th)
        got = run(lambda v: oracle.blockcoh_mv(bands, A.nor, v))
        assert boost_close(got, case["expected"], tol).all(), ("blockcoh", th)


# ------------------------------------------------------------------ Poisson


@pytest.mark.parametrize("m,n", [(1, 1), (1, 5), (5, 1), (3, 3), (9, 3), (4, 7), (16, 16), (31, 2)])
def test_poisson_closed_form_and_dense(m, n):
    A = oracle.poisson_crs(m, n)
    N = m * n
    assert A.nnz == 5 * m * n - 2 * m - 2 * n  # CRS.hpp:64 simplified
    assert A.row_start[0] == 0 and A.row_start[-1] == A.nnz  # CRS.hpp:72-73
    # dense brute force: T_m kron I_n + I_m kron T_n style 5-point stencil
    D = np.zeros((N, N))
    for r in range(N):
        i, j = divmod(r, m)
        D[r, r] = 4
        for (di, dj) in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            ii, jj = i + di, j + dj
            if 0 <= ii < n and 0 <= jj < m:
                D[r, ii * m + jj] = -1
    S = sp.csr_matrix((A.data, A.col_ind, A.row_start), shape=(N, N)).toarray()
    np.testing.assert_array_equal(S, D)
    assert all(np.all(np.diff(A.col_ind[A.row_start[r]:A.row_start[r + 1]]) > 0) for r in range(N))
    # the COO generator holds the same matrix
    t = oracle.poisson_triplets(m, n)
    B = sp.coo_matrix((t.vals, (t.rows, t.cols)), shape=(N, N)).toarray()
    np.testing.assert_array_equal(B, D)
    assert oracle.triplet_to_crs(t).col_ind.tolist() == A.col_ind.tolist()


@pytest.mark.parametrize("m,n", [(5, 5), (10, 5), (7, 3)])
def test_power_iteration_converges_to_closed_form_eigenpair(m, n):
    A = oracle.poisson_crs(m, n)
    x, y = oracle.crs_power_iteration(A, np.ones(m * n), 4000)
    lam = 4 + 2 * np.cos(np.pi / (m + 1)) + 2 * np.cos(np.pi / (n + 1))
    # eigenvector of lambda_max: alternating-sign product of sines
    i = np.arange(m)
    j = np.arange(n)
    v = np.outer(np.sin((j + 1) * n * np.pi / (n + 1)), np.sin((i + 1) * m * np.pi / (m + 1))).ravel()
    v /= np.linalg.norm(v)
    if v @ x < 0:
        v = -v
    assert np.linalg.norm(x - v) < 1e-6
    rq = x @ oracle.crs_mv(A, x)
    assert abs(rq - lam) < 1e-10 * lam


# ------------------------------------------------------------------ loaders


FIXTURE_SIZES = {
    "arc130.mtx": (130, 130, 1282),
    "gre_1107.mtx": (1107, 1107, 5664),
    "G38.mtx": (2000, 2000, 11779),
    "mycielskian5.mtx": (23, 23, 71),
}


@pytest.mark.parametrize("name", sorted(FIXTURE_SIZES))
def test_mm_loader_against_scipy(name):
    R, Cc, E = FIXTURE_SIZES[name]
    ref = scipy.io.mmread(fixture_path(name)).tocoo()
    symmetric = "symmetric" in open(fixture_path(name)).readline()
    has_data = "pattern" not in open(fixture_path(name)).readline()
    t = oracle.load_mm(fixture_path(name), has_data, symmetric)
    assert (t.nor, t.noc) == (R, Cc)
    diag = 0
    raw = scipy.io.mmread(fixture_path(name))
    if symmetric:
        low = sp.tril(sp.coo_matrix(raw)).tocoo()
        diag = int(np.count_nonzero(low.row == low.col))
        assert t.nnz == 2 * E - diag
    else:
        zeros = int(np.count_nonzero(ref.data == 0))
        assert t.nnz == E - zeros
    mine = sp.coo_matrix((t.vals, (t.rows, t.cols)), shape=(R, Cc)).toarray()
    theirs = ref.toarray()
    if not has_data:
        theirs = (theirs != 0).astype(float)
    np.testing.assert_array_equal(mine, theirs)


def test_bin_loader_sizes():
    t = oracle.load_bin(fixture_path("mycielskian5.bin"), 23, True, False)
    assert os.path.getsize(fixture_path("mycielskian5.bin")) == 568  # 71 pairs
    assert t.nnz == 142 and (t.nor, t.noc) == (23, 23)
    mm = oracle.load_mm(fixture_path("mycielskian5.mtx"), False, True)
    a = sp.coo_matrix((t.vals, (t.rows, t.cols)), shape=(23, 23)).toarray()
    b = sp.coo_matrix((mm.vals, (mm.rows, mm.cols)), shape=(23, 23)).toarray()
    np.testing.assert_array_equal(a, b)
    t = oracle.load_bin(fixture_path("test_mat_8_4.bin"), 256, False, False)
    assert t.nnz == 1024
    raw = np.fromfile(fixture_path("test_mat_8_4.bin"), dtype="<u4").reshape(-1, 2)
    np.testing.assert_array_equal(t.rows, raw[:, 0])
    np.testing.assert_array_equal(t.cols, raw[:, 1])


def test_mm_explicit_zero_dropped_and_blank_line_stops(tmp_path):
    p = tmp_path / "z.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% c\n3 3 4\n1 1 2.5\n2 2 0\n3 1 -1e-3\n\n3 3 9\n")
    t = oracle.load_mm(str(p), True, False)
    assert t.nnz == 2
    assert t.rows.tolist() == [0, 2] and t.cols.tolist() == [0, 0] and t.vals.tolist() == [2.5, -1e-3]


def test_mt19937_known_answer():
    st = np.zeros(625, np.uint32)
    L = oracle.lib()
    L.orc_mt19937_seed(st, 5489)
    out = [L.orc_mt19937_next(st) for _ in range(10000)]
    assert out[0] == 3499211612 and out[-1] == 4123659995


def test_uniform_real_range():
    st = np.zeros(625, np.uint32)
    L = oracle.lib()
    L.orc_mt19937_seed(st, 747846)
    v = np.array([L.orc_uniform_real(st, -100.0, 100.0) for _ in range(5000)])
    assert v.min() >= -100 and v.max() < 100 and abs(v.mean()) < 5


# ------------------------------------------------------------------ COO -> CRS


def test_triplet_to_crs_matches_scipy_without_duplicates():
    rng = np.random.default_rng(7)
    N, Z = 300, 4000
    lin = rng.choice(N * N, Z, replace=False)
    r, c = (lin // N).astype(np.int32), (lin % N).astype(np.int32)
    v = rng.standard_normal(Z)
    A = oracle.triplet_to_crs(oracle.Triplets(N, N, r, c, v))
    S = sp.coo_matrix((v, (r, c)), shape=(N, N)).tocsr()
    S.sort_indices()
    np.testing.assert_array_equal(A.row_start, S.indptr)
    np.testing.assert_array_equal(A.col_ind, S.indices)
    np.testing.assert_array_equal(A.data, S.data)


def test_triplet_to_crs_keeps_duplicates_and_empty_rows():
    r = np.array([5, 0, 5, 5, 2], np.int32)
    c = np.array([1, 3, 1, 0, 2], np.int32)
    v = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    A = oracle.triplet_to_crs(oracle.Triplets(8, 8, r, c, v))
    assert A.row_start.tolist() == [0, 1, 1, 2, 2, 2, 5, 5, 5]
    assert A.col_ind.tolist() == [3, 2, 0, 1, 1]
    assert A.data.tolist() == [2.0, 5.0, 4.0, 1.0, 3.0]
    # long rows take the merge-sort path
    rng = np.random.default_rng(3)
    c = rng.integers(0, 1000, 500).astype(np.int32)
    A = oracle.triplet_to_crs(oracle.Triplets(1, 1000, np.zeros(500, np.int32), c, np.arange(500.0)))
    assert np.all(np.diff(A.col_ind) >= 0)
    for col in np.unique(c):  # stability: input order among equal columns
        d = A.data[A.col_ind == col]
        assert np.all(np.diff(d) > 0)


def test_row_partitions():
    L = oracle.lib()
    f = np.zeros(4, np.int32)
    t = np.zeros(4, np.int32)
    L.orc_row_partitions(10, 4, f, t)
    assert f.tolist() == [0, 2, 4, 6] and t.tolist() == [2, 2, 2, 4]


@pytest.mark.parametrize("P", [1, 2, 3, 7, 64, 130])
def test_triplet_to_multiple_crs(P):
    """TripletToMultipleCRS (Util/TripletToCRS.hpp:309-385) against a brute
    force: each partition's triplets filtered by row range and ordered by a
    stable numpy lexsort on (row, column), row_start counted per local row;
    partition sizes by the floor rule; P = 1 is the whole CRS."""
    rng = np.random.default_rng(P)
    nor, noc, nnz = 100, 80, 700
    r = rng.integers(0, nor, nnz).astype(np.int32)
    r[r % 9 == 4] = 0  # empty rows
    c = rng.integers(0, noc, nnz).astype(np.int32)
    r[:20], c[:20] = 7, 3  # duplicates
    v = rng.standard_normal(nnz)
    t = oracle.Triplets(nor, noc, r, c, v)
    parts = oracle.triplet_to_multiple_crs(t, P)
    am = nor // P
    assert [f for f, _ in parts] == [min(p * am, nor) if p < P else 0 for p in range(P)]
    assert [A.nor for _, A in parts] == [am] * (P - 1) + [nor - (P - 1) * am]
    order = np.lexsort((c, r))  # stable: duplicates keep input order
    for f, A in parts:
        sel = order[(r[order] >= f) & (r[order] < f + A.nor)]
        assert np.array_equal(A.col_ind, c[sel]) and np.array_equal(A.data, v[sel])
        counts = np.bincount(r[sel] - f, minlength=A.nor) if A.nor else np.zeros(0, int)
        assert np.array_equal(A.row_start, np.concatenate([[0], np.cumsum(counts)]))
    if P == 1:
        W = oracle.triplet_to_crs(t)
        A = parts[0][1]
        assert np.array_equal(A.row_start, W.row_start) and np.array_equal(A.col_ind, W.col_ind)


# ------------------------------------------------------------------ merge path


def _brute_merge(row_end, nnz):
    """Linear merge of A=row_end against B=0..nnz-1 (A wins when A[i] <= j
    ... i.e. a row ends before nonzero j is consumed); returns the path
    coordinates for every diagonal."""
    path = [(0, 0)]
    a = b = 0
    while a < len(row_end) or b < nnz:
        if a < len(row_end) and (b >= nnz or row_end[a] <= b):
            a += 1
        else:
            b += 1
        path.append((a, b))
    return path


@pytest.mark.parametrize("seed", range(6))
def test_merge_search_equals_linear_merge(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(1, 40))
    lens = rng.integers(0, 6, N)
    lens[rng.random(N) < 0.2] = 0
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    A = oracle.CRSMatrix(N, N, rs, np.zeros(rs[-1], np.int32), np.zeros(rs[-1]))
    path = _brute_merge(rs[1:], int(rs[-1]))
    for d in range(N + int(rs[-1]) + 1):
        assert oracle.merge_search(A, d) == path[d], d


def test_merge_splits_formula():
    A = oracle.poisson_crs(9, 3)
    total = A.nor + A.nnz
    for P in (1, 2, 7, 64, 200):
        a, b = oracle.merge_splits(A, P)
        items = -(-total // P)
        assert np.all(a + b == np.minimum(items * np.arange(P + 1), total))
        assert np.all(np.diff(a) >= 0) and np.all(np.diff(b) >= 0)
        assert a[-1] == A.nor and b[-1] == A.nnz


def test_merge_mv_long_rows_split_across_threads():
    rng = np.random.default_rng(11)
    N = 50
    lens = np.zeros(N, np.int64)
    lens[3] = 400
    lens[10:20] = 1
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = rng.integers(0, N, rs[-1]).astype(np.int32)
    A = oracle.CRSMatrix(N, N, rs, ci, rng.standard_normal(rs[-1]))
    x = rng.standard_normal(N)
    yref = oracle.crs_mv(A, x)
    for P in (1, 3, 17, 64, 451):
        y = oracle.merge_mv(A, x, P)
        assert np.max(np.abs(y - yref)) <= 1e-13 * np.max(np.abs(yref))
        assert np.array_equal(oracle.merge_mv(A, x, P, omp_threads=4), y)


# ------------------------------------------------------------------ CSB


def test_csb_block_bits_rule():
    # N=1024^2 (FP64) -> 2^13, N=8386816 -> 2^15, N=1 -> 2^3
    assert oracle.csb_block_bits(1 << 20, 1 << 20) == 13
    assert oracle.csb_block_bits(8386816, 8386816) == 15
    assert oracle.csb_block_bits(1, 1) == 3
    assert oracle.csb_block_bits(130, 130) == 7
    # the L2 rule: 8*8*2^16*2 bits > 0.85 MB -> shrink to 2^15 (= lg, stop);
    # 64-byte values at N=2^24 shrink 15 -> 12 (= lg)
    assert oracle.csb_block_bits(1 << 30, 1 << 30, 8) == 15
    assert oracle.csb_block_bits(1 << 24, 1 << 24, 64) == 12
    assert oracle.csb_block_bits(1 << 24, 1 << 24, 8) == 15


def _interleave(r, c, bits):
    k = 0
    for b in range(bits - 1, -1, -1):
        k = (k << 2) | (((r >> b) & 1) << 1) | ((c >> b) & 1)
    return k


def test_zmorton_compare_is_bit_interleave_order():
    L = oracle.lib()
    bits = 3
    pts = list(itertools.product(range(8), range(8)))
    for (r1, c1), (r2, c2) in itertools.product(pts[::3], pts[::2]):
        le = L.orc_zmorton_less_equal(c1, r1, c2, r2, bits)
        assert bool(le) == (_interleave(r1, c1, bits) <= _interleave(r2, c2, bits))


@pytest.mark.parametrize("bits", [0, 2, 3])
def test_csb_layout_invariants(bits):
    A = oracle.poisson_crs(9, 7)
    t = oracle.poisson_triplets(9, 7)
    b, ind, data, blk, bal = oracle.csb_build(t, bits)
    beta = 1 << b
    H = -(-A.noc // beta)
    rows, cols, vals = [], [], []
    for br in range(H):
        for bc in range(H):
            s, e = blk[br * H + bc], blk[br * H + bc + 1]
            keys = [_interleave(int(w >> 16), int(w & 0xFFFF), b) for w in ind[s:e]]
            assert keys == sorted(keys)
            rows += [br * beta + int(w >> 16) for w in ind[s:e]]
            cols += [bc * beta + int(w & 0xFFFF) for w in ind[s:e]]
            vals += list(data[s:e])
    M = sp.coo_matrix((vals, (rows, cols)), shape=(A.nor, A.noc)).toarray()
    S = sp.csr_matrix((A.data, A.col_ind, A.row_start), shape=(A.nor, A.noc)).toarray()
    np.testing.assert_array_equal(M, S)
    assert bal in (0, 1)


def test_csb_balanced_flag():
    # one block row with > 2x the mean block-row nnz -> unbalanced (CSB.hpp:549-559)
    N = 64
    r = np.array([0] * 40 + [63], np.int32)
    c = np.arange(41, dtype=np.int32) % N
    t = oracle.Triplets(N, N, r, c, np.ones(41))
    assert oracle.csb_build(t, 3)[4] == 0
    t = oracle.poisson_triplets(8, 8)
    assert oracle.csb_build(t, 3)[4] == 1


# ------------------------------------------------------------------ BlockCOH


def test_hilbert_known_answers():
    n2 = {(0, 0): 0, (1, 0): 1, (1, 1): 2, (0, 1): 3}
    for (r, c), d in n2.items():
        assert oracle.hilbert(2, r, c) == d
    n4 = [[0, 1, 14, 15], [3, 2, 13, 12], [4, 7, 8, 11], [5, 6, 9, 10]]
    for r in range(4):
        for c in range(4):
            assert oracle.hilbert(4, r, c) == n4[r][c]
    # bijection and unit steps on a larger curve
    pos = {oracle.hilbert(16, r, c): (r, c) for r in range(16) for c in range(16)}
    assert sorted(pos) == list(range(256))
    for d in range(255):
        (r1, c1), (r2, c2) = pos[d], pos[d + 1]
        assert abs(r1 - r2) + abs(c1 - c2) == 1


def _greedy_bands_ok(counts, P, starts, nnz):
    """Characterisation of the band scan: bands are contiguous, cover every
    row, each closed band holds <= share and would exceed it with the next
    band's first row."""
    Z = int(counts.sum())
    share = Z // P
    ends = list(starts[1:]) + [len(counts)]
    assert starts[0] == 0 and int(sum(nnz)) == Z
    for p in range(P):
        assert starts[p] <= ends[p]
        assert nnz[p] == counts[starts[p]:ends[p]].sum()
    opened = [p for p in range(1, P) if starts[p] < len(counts) or nnz[p] > 0]
    for p in range(P - 1):
        if p + 1 in opened:
            rows = ends[p] - starts[p]
            # a band opened on a heavy row may hold it alone: the scan only
            # tests the running sum after the next row is added
            if not (p > 0 and rows == 1):
                assert nnz[p] <= share
            assert nnz[p] + counts[starts[p + 1]] > share


@pytest.mark.parametrize("seed", range(8))
def test_blockcoh_bands_greedy(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(16, 300))
    counts = rng.integers(0, 9, N).astype(np.int32)
    counts[rng.integers(0, N)] += int(rng.integers(0, 200))
    rs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    A = oracle.CRSMatrix(N, N, rs, np.zeros(rs[-1], np.int32), np.zeros(rs[-1]))
    for th in (1, 2, 3, 7):
        P, starts, nnz = oracle.blockcoh_partition(A, th)
        assert P == max(1, th if N // th >= 8 else N // 8)
        _greedy_bands_ok(counts, P, starts, nnz)


def test_blockcoh_thread_clamp():
    L = oracle.lib()
    assert L.orc_blockcoh_threads(9, 4) == 1      # 9/4 < 8 -> 9/8
    assert L.orc_blockcoh_threads(64, 8) == 8
    assert L.orc_blockcoh_threads(63, 8) == 7
    assert L.orc_blockcoh_threads(3, 2) == 1      # reference would pick 0


@pytest.mark.parametrize("m,n,th", [(9, 7, 1), (9, 7, 3), (40, 40, 5), (23, 1, 2)])
def test_blockcoh_layout_roundtrip(m, n, th):
    A = oracle.poisson_crs(m, n)
    bands = oracle.blockcoh_build(A, th)
    S = sp.csr_matrix((A.data, A.col_ind, A.row_start), shape=(A.nor, A.noc)).toarray()
    M = np.zeros_like(S)
    for b in bands:
        beta = 1 << b["beta_bits"]
        hs = oracle.lib().orc_blockcoh_hilbert_size(b["rows"], A.noc)
        # decode block coordinates from the BICRS jumps
        br, bc, rbi = int(b["row_jump_block"][0]), int(b["col_jump_block"][0]), 1
        coords = []
        for k in range(b["blocks"]):
            if k > 0:
                bc += int(b["col_jump_block"][k])
                if bc >= b["horizontal_blocks"]:
                    bc -= b["horizontal_blocks"]
                    br += int(b["row_jump_block"][rbi]); rbi += 1
            coords.append((br, bc))
        keys = [oracle.hilbert(hs // beta, r, c) for r, c in coords]
        assert keys == sorted(keys) and len(set(keys)) == len(keys)
        # decode ICRS inside each block: col = prefix sum mod beta
        ci, ri = 0, 0
        for k, (r0, c0) in enumerate(coords):
            row = int(b["row_jump"][ri]); ri += 1
            col = int(b["col_jump"][ci])
            for i in range(int(b["block_nnz"][k])):
                if i > 0:
                    col += int(b["col_jump"][ci])
                    if col >= beta:
                        col -= beta
                        row += int(b["row_jump"][ri]); ri += 1
                M[b["start"] + r0 * beta + row, c0 * beta + col] += b["data"][ci]
                ci += 1
        assert ci == b["nnz"] and ri == len(b["row_jump"])
    np.testing.assert_array_equal(M, S)


def test_blockcoh_beta_integer_form():
    """ceil(log2(sqrt(M))) equals the smallest k with 4^k >= M."""
    L = oracle.lib()
    for M in list(range(1, 5000)) + [2 ** 20, 2 ** 20 + 1, 8386816, 4194304, 4194305]:
        k = 0
        while 4 ** k < M:
            k += 1
        assert L.orc_blockcoh_beta_bits(M, 1) == min(15, 3 + k)
        assert L.orc_blockcoh_hilbert_size(M, 1) == 1 << (M - 1).bit_length()


# ---------------------------------------------------------------- matrix statistics
def _stats_brute(A):
    """Util_Scripts/calc_stats.cpp:44-105 written out again with numpy block
    counts and Python loops in the tool's order."""
    import math
    N = A.nor
    lens = np.diff(A.row_start).astype(np.int64)
    avg = A.nnz / N
    var = 0.0
    for ln in lens:
        var += (float(ln) - avg) ** 2 / N
    bits = min(16, 3 + int(math.ceil(math.log2(math.sqrt(N)))))
    bs = 1 << bits
    nb = -(-N // bs)
    rows = np.repeat(np.arange(N), lens)
    br, bc = rows // bs, A.col_ind.astype(np.int64) // bs
    keep = bc < nb
    hist = np.zeros((nb, nb), np.int64)
    np.add.at(hist, (br[keep], bc[keep]), 1)
    nbw = nb - 1 if N % bs else nb
    bavg = 0.0
    brow_avg = 0.0
    brow = hist[:nbw, :nbw].sum(axis=1)
    for i in range(nbw):
        for j in range(nbw):
            bavg += float(hist[i, j]) / float(nbw) ** 2
        brow_avg += float(brow[i]) / nbw
    bvar = 0.0
    for i in range(nbw):
        for j in range(nbw):
            bvar += (float(hist[i, j]) - bavg) ** 2 / float(nbw) ** 2
    bal = int(all(brow[i] <= 2 * brow_avg for i in range(nbw)))
    return {"rows_avg": avg, "rows_var": var, "row_max": int(lens.max()) if N else 0, "block_bits": bits,
            "nb_blocks": nbw, "block_avg": bavg, "block_var": bvar, "balanced": bal}


@pytest.mark.parametrize("case", ["poisson", "powerlaw", "rect", "tiny", "empty_rows"])
def test_matrix_stats_brute_force(case):
    rng = np.random.default_rng(11)
    if case == "poisson":
        A = oracle.poisson_crs(70, 53)
    elif case == "powerlaw":
        n = 5000
        lens = np.minimum((3.0 * rng.random(n) ** (-1 / 1.8)).astype(np.int64), 900)
        rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        ci = np.concatenate([np.sort(rng.choice(n, size=int(k), replace=False)) for k in lens]).astype(np.int32)
        A = oracle.CRSMatrix(n, n, rs, ci, np.ones(int(rs[-1])))
    elif case == "rect":  # columns beyond the square's blocks are not counted
        n, nc = 3000, 9000
        lens = rng.integers(0, 12, n)
        rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        ci = np.concatenate([np.sort(rng.choice(nc, size=int(k), replace=False)) for k in lens]).astype(np.int32)
        A = oracle.CRSMatrix(n, nc, rs, ci, np.ones(int(rs[-1])))
    elif case == "tiny":
        A = oracle.CRSMatrix(3, 3, [0, 1, 3, 4], [0, 0, 2, 1], [1.0] * 4)
    else:
        rs = np.array([0, 0, 0, 5, 5, 9] + [9] * 2000, np.int32)
        A = oracle.CRSMatrix(len(rs) - 1, len(rs) - 1, rs, [1, 2, 3, 4, 5, 0, 7, 8, 9], np.ones(9))
    got, want = oracle.matrix_stats(A), _stats_brute(A)
    assert got == want, (got, want)


def test_matrix_stats_poisson_closed_form():
    """Row lengths of the m x n Poisson matrix: 4 corners of 3, 2(m-2)+2(n-2)
    edge rows of 4, (m-2)(n-2) interior rows of 5."""
    for m, n in ((5, 5), (40, 17), (1024, 1024)):
        A = oracle.poisson_crs(m, n)
        N = m * n
        counts = {3: 4, 4: 2 * (m - 2) + 2 * (n - 2), 5: (m - 2) * (n - 2)}
        avg = sum(k * c for k, c in counts.items()) / N
        var = sum(c * (k - avg) ** 2 for k, c in counts.items()) / N
        s = oracle.matrix_stats(A)
        assert s["rows_avg"] == A.nnz / N and s["row_max"] == 5
        assert abs(s["rows_var"] - var) <= 1e-12 * var
        assert s["balanced"] == 1


def test_matrix_stats_balance_flag():
    """A matrix whose first block row holds most entries is unbalanced; the
    same entries spread over every block row are balanced."""
    N = 4096  # bits = min(16, 3 + 6) = 9: 8 whole block rows of 512
    rows = np.concatenate([np.repeat(np.arange(512), 40), np.arange(512, N)])
    cols = np.arange(len(rows)) % N
    order = np.lexsort((cols, rows))
    rs = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=N))]).astype(np.int32)
    A = oracle.CRSMatrix(N, N, rs, cols[order], np.ones(len(rows)))
    s = oracle.matrix_stats(A)
    assert s["block_bits"] == 9 and s["nb_blocks"] == 8 and s["balanced"] == 0
    B = oracle.poisson_crs(64, 64)
    assert oracle.matrix_stats(B)["balanced"] == 1
