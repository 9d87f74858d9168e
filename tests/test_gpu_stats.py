"""GPU parity of the matrix statistics (Util_Scripts/calc_stats.cpp:44-105)
against the oracle's restatement: every field equal (the block counts come
from the device, the sums are taken in the tool's order), for every format
and every way of loading a matrix."""
import numpy as np
import pytest

import oracle

pw = pytest.importorskip("thesis_spmv_amd")
pytestmark = pytest.mark.gpu

FORMATS = [pw.CRSHip, pw.CRSMergeHip, pw.CSBHip, pw.BlockCOHHip]
KEYS = ("rows_avg", "rows_var", "row_max", "block_bits", "nb_blocks", "block_avg", "block_var", "balanced")


def _check(m, A):
    got, want = m.stats(), oracle.matrix_stats(A)
    assert got["nor"] == A.nor and got["nnz"] == A.nnz
    for k in KEYS:
        assert got[k] == want[k], (k, got[k], want[k])


@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
def test_stats_poisson_generated(cls):
    for m_, n_ in ((40, 31), (300, 300)):
        m = cls()
        m.generatePoissonMatrix(m_, n_, 1)
        _check(m, oracle.poisson_crs(m_, n_))


@pytest.mark.parametrize("cls", FORMATS, ids=lambda c: c.__name__)
def test_stats_powerlaw_and_rectangular(cls):
    rs, ci, dv = pw.gen_powerlaw(50000, struct_seed=3, value_seed=4, cap=5000)
    A = oracle.CRSMatrix(50000, 50000, rs, ci, dv)
    m = cls()
    m.loadCRS(A.nor, A.noc, A.row_start, A.col_ind, A.data)
    _check(m, A)
    rng = np.random.default_rng(2)
    n, nc = 3000, 9000
    lens = rng.integers(0, 12, n)
    rsr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cir = np.concatenate([np.sort(rng.choice(nc, size=int(k), replace=False)) for k in lens]).astype(np.int32)
    R = oracle.CRSMatrix(n, nc, rsr, cir, np.ones(int(rsr[-1])))
    m = cls()
    m.loadCRS(R.nor, R.noc, R.row_start, R.col_ind, R.data)
    _check(m, R)


def test_stats_from_triplets_and_unbalanced():
    N = 4096
    rows = np.concatenate([np.repeat(np.arange(512), 40), np.arange(512, N)]).astype(np.int32)
    cols = (np.arange(len(rows)) % N).astype(np.int32)
    t = oracle.Triplets(N, N, rows[::-1].copy(), cols[::-1].copy(), np.ones(len(rows)))
    A = oracle.triplet_to_crs(t)
    for cls in FORMATS:
        m = cls()
        m.loadFromTriplets(t)
        _check(m, A)
        assert m.stats()["balanced"] == 0


def test_stats_edge_shapes():
    cases = [oracle.CRSMatrix(1, 1, [0, 1], [0], [2.0]),
             oracle.CRSMatrix(40, 40, [0] * 41, [], []),
             oracle.CRSMatrix(9, 9000, [0, 0, 0, 4000, 4000, 4001, 4001, 4001, 4001, 4001],
                              np.sort(np.random.default_rng(0).integers(0, 9000, 4001)), np.ones(4001))]
    for A in cases:
        m = pw.CRSHip()
        m.loadCRS(A.nor, A.noc, A.row_start, A.col_ind, A.data)
        _check(m, A)
    with pytest.raises(pw.PwmError):
        pw.CRSHip().stats()
