"""GPU parity of the blocked formats (CSBHip, BlockCOHHip) against the oracle.

Bars: stored layouts (CSB ind/data/blk_ptr and balanced flag; BlockCOH band
partition, Hilbert block order, ICRS/BICRS arrays) bit-exact with the oracle's
restatement of the reference builders; FP64 y within 1e-13 relative max-norm;
golden vectors at the reference's tolerances."""
import numpy as np
import pytest

import oracle
from conftest import boost_close, load_case_matrix

pw = pytest.importorskip("thesis_spmv_amd")
pytestmark = pytest.mark.gpu

FORMATS = [pw.CSBHip, pw.BlockCOHHip]
# CSB device layouts: 1 = slab pieces over the stored block order, 2 = one
# contiguous stream per slab (the automatic choice when its index packs),
# 3 = the same streams with column and row in slab stored apart (the
# automatic choice past 2^24 columns), 4 = swept slabs (persistent
# nnz-balanced slabs walking x window by window)
# layout options as stored: slab pieces built on the device and on the host,
# packed and wide slab streams, swept slabs with 1 MB windows (runs in column
# order, the default, and in row order) and with 1024-column windows
CSB_VARIANTS = [{"csb_layout": 1}, {"csb_layout": 1, "csb_build": 1}, {"csb_layout": 2}, {"csb_layout": 3},
                {"csb_layout": 4}, {"csb_layout": 4, "csb_sweep_order": 0}, {"csb_layout": 4, "csb_window_bits": 10}]


def _relerr(a, b):
    m = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (m if m > 0 else 1.0))


def _load(cls, A, threads=-1, values="f64", **opt):
    m = cls(threads, values, **opt)
    m.loadCRS(A.nor, A.noc, A.row_start, A.col_ind, A.data)
    return m


def _triplets_of(A):
    rows = np.repeat(np.arange(A.nor, dtype=np.int32), np.diff(A.row_start))
    return oracle.Triplets(A.nor, A.noc, rows, A.col_ind, A.data)


def _powerlaw(n, seed=1, cap=50000):
    rs, ci, dv = pw.gen_powerlaw(n, struct_seed=seed, value_seed=seed + 1, cap=cap)
    return oracle.CRSMatrix(n, n, rs, ci, dv)


def _stencil(side, extra=2, seed=3):
    rs, ci, dv = pw.gen_stencil9(side, extra, seed)
    return oracle.CRSMatrix(side * side, side * side, rs, ci, dv)


@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
@pytest.mark.parametrize("idx", range(24))
def test_golden(goldens, cls, idx):
    case = goldens[idx]
    A, t = load_case_matrix(case)
    print(f"case {case['name']} {case['index_type']} {cls.__name__}", flush=True)
    runs = ([(t, {}) for t in (1, 2, 3, 4, 8, -1)] + [(t, {"coh_build": 1}) for t in (1, 3, -1)]
            + [(t, {"coh_occ8": o}) for t in (1, -1) for o in (0, 1)]
            if cls is pw.BlockCOHHip
            else [(-1, dict(v)) for v in CSB_VARIANTS])
    for threads, opts in runs:
        m = cls(threads, **opts)
        if case["matrix"]["kind"] == "poisson":
            m.generatePoissonMatrix(case["matrix"]["m"], case["matrix"]["n"], 1)
        else:
            m.loadFromTriplets(t, 1)
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


# ---------------------------------------------------------------- layouts

@pytest.mark.parametrize("variant", CSB_VARIANTS, ids=lambda v: "-".join(f"{k}{x}" for k, x in v.items()))
@pytest.mark.parametrize("bits", [0, 2, 3, 5, 9, 13])
def test_csb_layout_bit_exact(bits, variant):
    for A in (oracle.poisson_crs(9, 7), oracle.poisson_crs(40, 31), _powerlaw(3000, cap=400)):
        m = _load(pw.CSBHip, A, block_bits=bits, **variant)
        gb, gind, gdata, gblk = m.csb_layout()
        rb, rind, rdata, rblk, rbal = oracle.csb_build(_triplets_of(A), bits)
        assert gb == rb
        assert np.array_equal(gblk, rblk) and np.array_equal(gind, rind) and np.array_equal(gdata, rdata)
        assert m.info()["balanced"] == rbal


@pytest.mark.parametrize("variant", CSB_VARIANTS, ids=lambda v: "-".join(f"{k}{x}" for k, x in v.items()))
def test_csb_layout_duplicates_and_unbalanced(variant):
    rng = np.random.default_rng(4)
    N = 700
    r = np.concatenate([rng.integers(0, 40, 3000), rng.integers(0, N, 500)]).astype(np.int32)
    c = rng.integers(0, N, 3500).astype(np.int32)
    r[:50], c[:50] = 3, 7  # duplicates
    v = rng.integers(1, 5, 3500).astype(np.float64)
    t = oracle.Triplets(N, N, r, c, v)
    A = oracle.triplet_to_crs(t)
    m = pw.CSBHip(**variant)
    m.loadFromTriplets(t)
    gb, gind, gdata, gblk = m.csb_layout()
    rb, rind, rdata, rblk, rbal = oracle.csb_build(_triplets_of(A), 0)
    assert np.array_equal(gblk, rblk) and np.array_equal(gind, rind) and np.array_equal(gdata, rdata)
    assert m.info()["balanced"] == rbal == 0
    y = np.zeros(N)
    x = rng.standard_normal(N)
    m.mv(x, y)
    assert _relerr(y, oracle.crs_mv(A, x)) <= 1e-13


@pytest.mark.parametrize("values", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("bits", [12, 13, 14, 15])
def test_csb_pieces_device_equals_host(values, bits):
    """Slab pieces with blocks of up to 8 LDS slabs (beta >> k rows, k up to
    3 at these block sizes): the device build and the host build store the
    same arrays, and their pieces give the same y up to the rounding of the
    summation order (a wrong piece would drop or repeat whole entries)."""
    A = oracle.poisson_crs(230, 190) if values == "bf16" else _powerlaw(45000, seed=5, cap=300)
    ms = [_load(pw.CSBHip, A, values=values, block_bits=bits, csb_layout=1, csb_build=b, csb_groups=8)
          for b in (0, 1)]
    assert [m.info()["layout"] for m in ms] == [1, 1]
    e0, e1 = ms[0].csb_layout(), ms[1].csb_layout()
    assert e0[0] == e1[0] == bits
    for a, b in zip(e0[1:], e1[1:]):
        assert np.array_equal(a, b)
    x = np.random.default_rng(bits).standard_normal(A.nor)
    ys = [np.zeros(A.nor), np.zeros(A.nor)]
    for m, y in zip(ms, ys):
        m.mv(x, y)
    assert _relerr(ys[0], ys[1]) <= (1e-13 if values == "f64" else 1e-5)
    if values == "f64":
        assert _relerr(ys[0], oracle.crs_mv(A, x)) <= 1e-13


def _check_bands(m, ref):
    for p in range(len(ref)):
        g, o = m.band(p), ref[p]
        for k in ("nnz", "blocks", "start", "rows", "beta_bits", "horizontal_blocks", "vertical_blocks"):
            assert g[k] == o[k], (p, k)
        for k in ("row_jump", "col_jump", "data", "row_jump_block", "col_jump_block", "block_nnz"):
            assert np.array_equal(g[k], o[k]), (p, k)


# BlockCOH layouts: 0 = built on the device (default), 1 = built on the host
@pytest.mark.parametrize("build", [0, 1])
@pytest.mark.parametrize("threads", [1, 2, 3, 8, 64, -1])
def test_blockcoh_layout_bit_exact(threads, build):
    for A in (oracle.poisson_crs(40, 31), _powerlaw(5000, cap=600), _stencil(60)):
        m = _load(pw.BlockCOHHip, A, threads, coh_build=build)
        P = m.info()["threads"]
        if threads > 0:
            ref = oracle.blockcoh_build(A, threads)
            assert P == len(ref)
        else:
            ref = oracle.blockcoh_build(A, P)
        _check_bands(m, ref)


def test_blockcoh_device_build_load_paths():
    """The device build from every load path (generated Poisson, triplets with
    duplicates and empty rows, a tall matrix whose bands have their own beta,
    a matrix whose first row outweighs a band share) and every value type,
    against the oracle's layout and the host build."""
    rng = np.random.default_rng(5)
    for values in ("f64", "f32", "bf16"):
        for threads in (1, 7, -1):
            m = pw.BlockCOHHip(threads, values)
            m.generatePoissonMatrix(50, 37, 1)
            P = m.info()["threads"]
            _check_bands(m, oracle.blockcoh_build(oracle.poisson_crs(50, 37), P))
    n, nc = 3000, 700
    rows = rng.integers(0, n, 40000).astype(np.int32)
    rows[:5000] = 0
    cols = rng.integers(0, nc, 40000).astype(np.int32)
    vals = rng.integers(-8, 9, 40000).astype(np.float64)
    t = oracle.Triplets(n, nc, rows, cols, vals)
    A = oracle.triplet_to_crs(t)
    for values in ("f64", "bf16"):
        for threads in (1, 5, 40):
            ms = [pw.BlockCOHHip(threads, values, coh_build=b) for b in (0, 1)]
            for m in ms:
                m.loadFromTriplets(t)
            P = ms[0].info()["threads"]
            _check_bands(ms[0], oracle.blockcoh_build(A, P))
            x = rng.standard_normal(nc)
            ref = oracle.crs_mv(A, x)
            # BF16 storage computes in FP32 (x rounded, products and sums in
            # FP32): per row at most (n_i + 2) * 2^-24 * (|A| |x|)_i off
            absA = oracle.CRSMatrix(A.nor, A.noc, A.row_start, A.col_ind, np.abs(A.data))
            bound = (np.diff(A.row_start) + 2) * 2.0 ** -24 * oracle.crs_mv(absA, np.abs(x))
            for m in ms:
                y = np.zeros(n)
                m.mv(x, y)
                if values == "f64":
                    assert _relerr(y, ref) <= 1e-13
                else:
                    assert np.all(np.abs(y - ref) <= bound), np.max(np.abs(y - ref) / np.maximum(bound, 1e-300))


# ---------------------------------------------------------------- SpMV values

@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
def test_spmv_random_and_stencil(cls):
    cases = [_powerlaw(20000, cap=5000), _stencil(150), oracle.poisson_crs(200, 150)]
    for A in cases:
        x = pw.gen_uniform(A.noc, 0)
        ref = oracle.crs_mv(A, x)
        csb_opts = ({}, {"block_bits": 7}, {"csb_groups": 1}, {"csb_groups": 3}, {"csb_groups": 8},
                    {"csb_groups": 16}, {"block_bits": 7, "csb_groups": 16}, {"csb_layout": 1},
                    {"csb_layout": 1, "csb_groups": 8}, {"csb_layout": 1, "csb_build": 1, "csb_groups": 8},
                    {"csb_layout": 1, "block_bits": 7, "csb_groups": 3},
                    {"csb_layout": 2, "block_bits": 3, "csb_groups": 5}, {"csb_unroll": 2, "csb_groups": 1},
                    {"csb_unroll": 4, "csb_groups": 8}, {"csb_nt": 1, "csb_groups": 8}, {"csb_nt": 1, "csb_groups": 1},
                    {"csb_nt": 0, "csb_unroll": 4}, {"csb_layout": 3}, {"csb_layout": 3, "csb_groups": 8},
                    {"csb_layout": 3, "csb_nt": 1, "csb_unroll": 4, "csb_groups": 16})
        for opt in csb_opts if cls is pw.CSBHip else ({}, {"coh_build": 1}, {"coh_occ8": 0}, {"coh_occ8": 1}):
            for threads in ((-1,) if cls is pw.CSBHip else (-1, 1, 2, 5)):
                m = _load(cls, A, threads, **opt)
                y = np.zeros(A.nor)
                m.mv(x, y)
                assert _relerr(y, ref) <= 1e-13, (A.nor, opt, threads)


def test_blockcoh_wide_bands_several_windows():
    """Bands with more rows than one workgroup's LDS window (4096 FP64 rows)."""
    A = _powerlaw(30000, cap=2000)
    x = pw.gen_uniform(A.noc, 1)
    m = _load(pw.BlockCOHHip, A, 2)
    assert max(m.band(p)["rows"] for p in range(2)) > 4096
    y = np.zeros(A.nor)
    m.mv(x, y)
    assert _relerr(y, oracle.crs_mv(A, x)) <= 1e-13


@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
def test_edge_shapes(cls):
    rng = np.random.default_rng(0)
    cases = [oracle.CRSMatrix(1, 1, [0, 1], [0], [2.5]),
             oracle.CRSMatrix(40, 40, [0] * 41, [], []),
             oracle.CRSMatrix(9, 9000, [0, 0, 0, 4000, 4000, 4001, 4001, 4001, 4001, 4001],
                              np.sort(rng.integers(0, 9000, 4001)), rng.standard_normal(4001)),
             oracle.CRSMatrix(900, 2, np.arange(901) // 3, np.zeros(300, np.int32), np.ones(300))]
    for A in cases:
        x = rng.standard_normal(A.noc)
        ref = oracle.crs_mv(A, x)
        for opt in ({"csb_layout": 1}, {"csb_layout": 2}, {"csb_layout": 2, "csb_groups": 8}, {"csb_layout": 3}) \
                if cls is pw.CSBHip else ({},):
            m = _load(cls, A, **opt)
            y = np.full(A.nor, np.nan)
            m.mv(x, y)
            assert _relerr(y, ref) <= 1e-13, (A.nor, A.noc, opt)


def test_config2_full_size():
    """9-point stencil 2896^2 + 2 random off-band per row (~92M nnz), FP64,
    CSB at beta = 2^14 (the workload's setting) and the reference rule (2^15),
    BlockCOH with its default bands; and the FP32-value run with its error."""
    A = _stencil(2896)
    assert A.nnz == (3 * 2896 - 2) ** 2 + 2 * 2896 * 2896
    x = pw.gen_uniform(A.noc, 0)
    ref = oracle.crs_mv(A, x)
    for cls, opt in ((pw.CSBHip, {"block_bits": 14}), (pw.CSBHip, {}), (pw.BlockCOHHip, {})):
        m = _load(cls, A, **opt)
        y = np.zeros(A.nor)
        m.mv(x, y)
        err = _relerr(y, ref)
        print(f"{cls.__name__} {opt} config-2 FP64 rel.err {err:.3e} info {m.info()}", flush=True)
        assert err <= 1e-13
        del m
    for cls in FORMATS:
        m = _load(cls, A, values="f32")
        y = np.zeros(A.nor)
        m.mv(x, y)
        err = _relerr(y, ref)
        print(f"{cls.__name__} config-2 FP32 values/accumulation rel.err vs FP64 oracle {err:.3e}", flush=True)
        assert err <= 1e-5
        del m


# ---------------------------------------------------------------- power method

@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
@pytest.mark.parametrize("it", [1, 2, 7, 50])
def test_power_iteration_state(cls, it):
    A = _powerlaw(6000, cap=700)
    A.data[:] = np.abs(A.data)
    x0 = pw.gen_uniform(A.nor, 3)
    xr, yr = oracle.crs_power_iteration(A, x0, it)
    for opt in (({"csb_groups": 1}, {"csb_groups": 8, "block_bits": 9}, {"csb_groups": 16, "block_bits": 8},
                 {"csb_unroll": 4, "csb_groups": 1}, {"csb_unroll": 2, "csb_groups": 8},
                 {"csb_layout": 1, "csb_groups": 1}, {"csb_layout": 1, "csb_groups": 8, "block_bits": 9},
                 {"csb_layout": 3, "csb_groups": 1}, {"csb_layout": 3, "csb_groups": 8, "block_bits": 9})
                if cls is pw.CSBHip else ({}, {"long_chunk": 256})[:1]):
        m = _load(cls, A, **opt)
        x, y = x0.copy(), np.zeros(A.nor)
        m.powerIteration(x, y, it)
        assert _relerr(x, xr) <= 1e-12 and _relerr(y, yr) <= 1e-12, opt


@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
def test_power_method_eigenvalue_poisson(cls):
    m = cls()
    m.generatePoissonMatrix(20, 17, 1)
    x = pw.DeviceVector.from_array(pw.gen_uniform(340, 5))
    y = pw.DeviceVector(340)
    lam, _, its = m.power_method_device(x, y, 20000, 1e-13)
    exact = 4 + 2 * np.cos(np.pi / 21) + 2 * np.cos(np.pi / 18)
    assert its < 20000 and abs(lam - exact) <= 1e-10 * exact


@pytest.mark.parametrize("values,tol", [("f32", 1e-5), ("bf16", 1e-5)])
def test_csb_device_build_narrow_values(values, tol):
    """The slab streams gather FP32 and BF16 values too (2- and 4-byte
    elements): Poisson (exact in BF16) and, for FP32, a power-law matrix."""
    cases = [oracle.poisson_crs(300, 200)] + ([_powerlaw(20000, cap=3000)] if values == "f32" else [])
    for A in cases:
        x = pw.gen_uniform(A.noc, 4)
        ref = oracle.crs_mv(A, x)
        for opt in ({}, {"csb_groups": 8}, {"csb_layout": 3}):
            m = _load(pw.CSBHip, A, values=values, **opt)
            y = np.zeros(A.nor)
            m.mv(x, y)
            assert _relerr(y, ref) <= tol, (values, A.nor, opt)
            gb, gind, gdata, gblk = m.csb_layout()
            rb, rind, rdata, rblk, _ = oracle.csb_build(_triplets_of(A), 0)
            assert np.array_equal(gind, rind) and np.array_equal(gblk, rblk)
            assert np.array_equal(gdata, rdata.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
def test_empty_triplets(cls):
    """No entries at all, given as triplets: y = 0."""
    m = cls()
    m.loadFromTriplets(oracle.Triplets(50, 50, [], [], []))
    y = np.full(50, np.nan)
    m.mv(np.ones(50), y)
    assert np.all(y == 0)
    assert m.stats()["nnz"] == 0


def test_csb_wide_streams_past_2_24_columns():
    """With more than 2^24 columns the packed (row in slab, column) index would
    leave slabs of under 256 rows, so the automatic choice builds wide slab
    streams on the device (not the host-built slab pieces): stored layout
    bit-exact, y within 1e-13, and the packed layout on the same matrix."""
    rng = np.random.default_rng(11)
    nor, noc = 3000, 1 << 25
    lens = rng.integers(0, 40, nor)
    lens[7] = 3000
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(noc, int(k), replace=False)) for k in lens]).astype(np.int32)
    A = oracle.CRSMatrix(nor, noc, rs, ci, rng.standard_normal(int(rs[-1])))
    x = rng.standard_normal(noc)
    ref = oracle.crs_mv(A, x)
    rb, rind, rdata, rblk, _ = oracle.csb_build(_triplets_of(A), 0)
    for opt, layout in (({}, 3), ({"csb_groups": 8}, 3), ({"csb_layout": 2}, 2)):
        m = _load(pw.CSBHip, A, **opt)
        assert m.info()["layout"] == layout, opt
        y = np.zeros(nor)
        m.mv(x, y)
        assert _relerr(y, ref) <= 1e-13, opt
        gb, gind, gdata, gblk = m.csb_layout()
        assert gb == rb and np.array_equal(gind, rind) and np.array_equal(gblk, rblk) and np.array_equal(gdata, rdata)


# ---------------------------------------------------------------- swept slabs

@pytest.mark.parametrize("values,tol", [("f64", 1e-13), ("f32", 2e-5), ("bf16", 2e-5)])
@pytest.mark.parametrize("opt", [{}, {"csb_window_bits": 12}, {"csb_nt": 0}, {"csb_sweep_unroll": 1},
                                 {"csb_sweep_unroll": 4, "csb_window_bits": 12}, {"csb_sweep_waves": 2},
                                 {"csb_sweep_order": 0}],
                         ids=lambda o: "-".join(f"{k}{v}" for k, v in o.items()) or "default")
def test_csb_swept_slabs(values, tol, opt):
    """Swept slabs on power-law (FP64/FP32) and Poisson (BF16, exact values)
    matrices against the oracle: SpMV, and 9 power iterations from ones."""
    A = oracle.poisson_crs(300, 170) if values == "bf16" else _powerlaw(150000, seed=9, cap=20000)
    m = _load(pw.CSBHip, A, values=values, csb_layout=4, **opt)
    info = m.info()
    assert info["layout"] == 4 and info["threads"] >= 1
    x = np.random.default_rng(3).uniform(-1, 1, A.nor)
    y = np.full(A.nor, np.nan)
    m.mv(x, y)
    err = _relerr(y, oracle.crs_mv(A, x))
    print(f"swept slabs {values} {opt}: {info['threads']} slabs, rel.err {err:.2e}", flush=True)
    assert err <= tol
    if values == "f64":
        xs, ys = np.ones(A.nor), np.zeros(A.nor)
        m.powerIteration(xs, ys, 9)
        xr, yr = oracle.crs_power_iteration(A, np.ones(A.nor), 9)
        assert np.max(np.abs(xs - xr)) <= 1e-12 and np.max(np.abs(ys - yr)) <= 1e-12


def test_csb_swept_slabs_edge_shapes():
    """Empty rows, a row longer than a slab's share, and rectangular matrices
    whose column count is not a multiple of the window."""
    rng = np.random.default_rng(11)
    cases = []
    n = 5000
    lens = rng.integers(0, 6, n)
    lens[:100] = 0
    lens[2500] = 4000
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, int(L), replace=False)) for L in lens]).astype(np.int32)
    cases.append(oracle.CRSMatrix(n, n, rs, ci, rng.standard_normal(int(rs[-1]))))
    for nor, noc in ((700, 5000), (5000, 1300)):
        lens = rng.integers(0, 9, nor)
        rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        ci = np.concatenate([np.sort(rng.choice(noc, int(L), replace=False)) for L in lens]).astype(np.int32)
        cases.append(oracle.CRSMatrix(nor, noc, rs, ci, rng.standard_normal(int(rs[-1]))))
    for A in cases:
        for wbits in (0, 10, 11):
            for order in (1, 0):
                m = _load(pw.CSBHip, A, csb_layout=4, csb_window_bits=wbits, csb_sweep_order=order)
                x = rng.standard_normal(A.noc)
                y = np.full(A.nor, np.nan)
                m.mv(x, y)
                assert _relerr(y, oracle.crs_mv(A, x)) <= 1e-13, (A.nor, A.noc, wbits, order)


def test_csb_swept_padding_adds_nothing():
    """Column-ordered runs are padded to whole blocks of 256 entries; a
    padding entry gathers its window's first column and must add nothing.
    No row of this matrix uses a window's first column and x is infinite
    there, so a padding entry that added 0 * inf would turn a row into NaN."""
    rng = np.random.default_rng(13)
    n, wbits = 30000, 10
    lens = rng.integers(0, 12, n)
    cols = np.arange(n)[np.arange(n) % (1 << wbits) != 0]
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(cols, int(L), replace=False)) for L in lens]).astype(np.int32)
    A = oracle.CRSMatrix(n, n, rs, ci, rng.standard_normal(int(rs[-1])))
    x = rng.standard_normal(n)
    x[:: 1 << wbits] = np.inf
    ref = oracle.crs_mv(A, x)
    assert np.all(np.isfinite(ref))
    for values, tol in (("f64", 1e-13), ("f32", 2e-5)):
        m = _load(pw.CSBHip, A, values=values, csb_layout=4, csb_window_bits=wbits)
        y = np.full(n, np.nan)
        m.mv(x, y)
        assert np.all(np.isfinite(y)) and _relerr(y, ref) <= tol, values

