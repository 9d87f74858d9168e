"""CPU tests of the library's matrix readers (pwm_read_mm / pwm_read_bin, the
reference's Triplet::loadFromMM / loadFromBin, Matrix/Triplet.hpp:75-216)
against the oracle's loaders on the reference's fixtures, on a file large
enough to be parsed in several chunks, and on malformed input.  No GPU."""
import os

import numpy as np
import pytest

import oracle
from conftest import fixture_path

pw = pytest.importorskip("thesis_spmv_amd")

MTX = ["arc130.mtx", "gre_1107.mtx", "G38.mtx", "mycielskian5.mtx"]


def _flags(name):
    banner = open(fixture_path(name)).readline()
    return "pattern" not in banner, "symmetric" in banner


def _same(t, o):
    assert (t.row_size, t.col_size) == (o.nor, o.noc)
    np.testing.assert_array_equal(t.row_coord, o.rows)
    np.testing.assert_array_equal(t.col_coord, o.cols)
    np.testing.assert_array_equal(t.data, o.vals)


@pytest.mark.parametrize("name", MTX)
@pytest.mark.parametrize("random_fill", [False, True])
def test_mm_matches_oracle(name, random_fill):
    has_data, symmetric = _flags(name)
    t = pw.read_mm(fixture_path(name), has_data, symmetric, random_fill)
    _same(t, oracle.load_mm(fixture_path(name), has_data, symmetric, random_fill))


@pytest.mark.parametrize("name", MTX)
def test_mm_flags_as_the_caller_passes_them(name):
    # the banner is not parsed: reading a symmetric file as general keeps only
    # the stored triangle, and reading with data where there is none fails
    has_data, symmetric = _flags(name)
    t = pw.read_mm(fixture_path(name), has_data, False)
    _same(t, oracle.load_mm(fixture_path(name), has_data, False))


@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("random_vals", [False, True])
def test_bin_matches_oracle(symmetric, random_vals):
    for name, size in (("mycielskian5.bin", 23), ("test_mat_8_4.bin", 256)):
        t = pw.read_bin(fixture_path(name), size, symmetric, random_vals)
        # (the reference reads past the end of the file when a symmetric
        # .bin has self-loops; the oracle and the library both stop at the
        # end, see the oracle's header)
        _same(t, oracle.load_bin(fixture_path(name), size, symmetric, random_vals))


def test_large_file_parsed_in_chunks(tmp_path):
    rng = np.random.default_rng(3)
    n, e = 5000, 300000
    r = rng.integers(1, n + 1, e)
    c = rng.integers(1, n + 1, e)
    v = np.round(rng.standard_normal(e), 6)
    v[::97] = 0.0  # explicit zeros are dropped
    p = tmp_path / "big.mtx"
    with open(p, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% generated\n")
        f.write(f"{n} {n} {e}\n")
        f.write("".join(f"{a} {b} {x:.6f}\n" for a, b, x in zip(r, c, v)))
    assert os.path.getsize(p) > 4 << 20
    t = pw.read_mm(str(p), True, False)
    _same(t, oracle.load_mm(str(p), True, False))
    assert t.nnz == e - len(v[::97])
    ts = pw.read_mm(str(p), True, True)
    _same(ts, oracle.load_mm(str(p), True, True))


def test_blank_line_stops_and_zero_dropped(tmp_path):
    p = tmp_path / "z.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% c\n3 3 4\n1 1 2.5\n2 2 0\n3 1 -1e-3\n\n3 3 9\n")
    t = pw.read_mm(str(p), True, False)
    assert t.row_coord.tolist() == [0, 2] and t.col_coord.tolist() == [0, 0] and t.data.tolist() == [2.5, -1e-3]
    q = tmp_path / "crlf.mtx"
    q.write_bytes(b"%%MatrixMarket matrix coordinate real general\r\n2 2 2\r\n1 2 3.0\r\n\r\n2 1 4.0\r\n")
    t = pw.read_mm(str(q), True, False)
    assert t.nnz == 1 and t.data.tolist() == [3.0]


def test_loaded_triplets_feed_the_crs_converter():
    has_data, symmetric = _flags("gre_1107.mtx")
    t = pw.read_mm(fixture_path("gre_1107.mtx"), has_data, symmetric)
    o = oracle.load_mm(fixture_path("gre_1107.mtx"), has_data, symmetric)
    A = oracle.triplet_to_crs(o)
    assert A.nnz == t.nnz and A.nor == t.row_size


@pytest.mark.parametrize("text", [
    "%%MatrixMarket\n3 3\n1 1 1\n",              # short size line
    "%%MatrixMarket\n3 3 1\n1 1 1\n2 2 1\n",     # more entries than declared
    "%%MatrixMarket\n3 3 2\n1 4 1\n",            # column outside the matrix
    "%%MatrixMarket\n3 3 2\n1 1\n",              # value missing
    "%%MatrixMarket\n3 3 2\n0 1 1\n",            # 0 is not a 1-based index
])
def test_malformed_mm_is_an_io_error(tmp_path, text):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(pw.PwmError) as e:
        pw.read_mm(str(p), True, False)
    assert e.value.code == -2


def test_missing_file_and_bad_bin(tmp_path):
    with pytest.raises(pw.PwmError) as e:
        pw.read_mm(str(tmp_path / "none.mtx"), True, False)
    assert e.value.code == -2
    p = tmp_path / "b.bin"
    np.array([0, 1, 5, 2], dtype="<u4").tofile(p)
    with pytest.raises(pw.PwmError):
        pw.read_bin(str(p), 4, False)  # row 5 outside a 4x4 matrix
    assert pw.read_bin(str(p), 6, False).nnz == 2
