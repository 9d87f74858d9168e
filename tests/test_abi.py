"""CPU checks of the C ABI: libpwmhip.so loads, exports every entry point
include/*.h declares, and its host-side generators produce what they promise.
No GPU computation happens here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

pw = pytest.importorskip("thesis_spmv_amd")


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            src = open(os.path.join(ROOT, "include", f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(pwm_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    L = pw.lib()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert _declared() <= set(pw._SIGS) | {"pwm_last_error"}


def test_create_without_device_fails_loudly():
    h = C.c_void_p()
    rc = pw.lib().pwm_create(1, -1, 0, C.byref(h))
    if rc == 0:
        pw.lib().pwm_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc == -3
    assert b"device" in pw.lib().pwm_last_error(None)


def test_bad_arguments_are_rejected_before_the_device():
    h = C.c_void_p()
    assert pw.lib().pwm_create(5, -1, 0, C.byref(h)) == -1      # unknown format id
    assert pw.lib().pwm_create(1, -1, 7, C.byref(h)) == -1      # unknown value type
    assert pw.lib().pwm_mv(None, None, None) == -1
    assert pw.select_type(3) is None and pw.select_type(11) is None


def test_gen_uniform_range_and_determinism():
    a = pw.gen_uniform(100000, 0)
    b = pw.gen_uniform(100000, 0)
    assert np.array_equal(a, b) and a.min() >= -1 and a.max() < 1 and abs(a.mean()) < 0.01
    # mt19937_64 seed 0 first output: 0x9d4bc8aeaa39d6d3? -> check via numpy-free restatement
    # (std::mt19937_64 default-seeded 10000th output is 9981545732273789042)
    assert not np.array_equal(a, pw.gen_uniform(100000, 1))


def test_gen_powerlaw_invariants():
    n = 200000
    rs, ci, dv = pw.gen_powerlaw(n)
    lens = np.diff(rs)
    assert rs[0] == 0 and np.all(lens >= 0) and lens.max() <= 50000
    assert abs(lens.mean() - 16.0) < 0.6, lens.mean()
    assert np.all((ci >= 0) & (ci < n))
    # strictly ascending (no repeats) inside every row
    d = np.diff(ci)
    starts = rs[1:-1]
    mask = np.ones(len(d), bool)
    mask[starts[starts < len(d) + 1] - 1] = False
    assert np.all(d[mask[:len(d)]] > 0)
    assert abs(dv.mean()) < 0.01 and abs(dv.std() - 1) < 0.01
    rs2, ci2, dv2 = pw.gen_powerlaw(n)
    assert np.array_equal(ci, ci2) and np.array_equal(dv, dv2)
    # tail: P(L >= k) ~ (xm/k)^alpha
    frac = np.mean(lens >= 100)
    assert 0.5 * (7.3333 / 100) ** 1.8 < frac < 2 * (7.3333 / 100) ** 1.8


def test_gen_spd_powerlaw_is_spd():
    import scipy.sparse as sp
    n = 3000
    rs, ci, dv = pw.gen_spd_powerlaw(n, seed=4)
    A = sp.csr_matrix((dv, ci, rs), shape=(n, n))
    assert abs(A - A.T).max() == 0
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d)
    assert np.all(d > 0) and np.all(d >= off + 1 - 1e-9)
    assert np.all(np.diff(ci[rs[0]:rs[1]]) > 0)
    assert 10 < A.nnz / n < 30
    B = A[:400, :400].toarray()
    np.linalg.cholesky(B)  # a principal submatrix of an SPD matrix is SPD
    rs2, ci2, dv2 = pw.gen_spd_powerlaw(n, seed=4)
    assert np.array_equal(ci, ci2) and np.array_equal(dv, dv2)


def test_gen_stencil9_structure():
    side, extra = 37, 2
    rs, ci, dv = pw.gen_stencil9(side, extra, 3)
    N = side * side
    assert rs[-1] == (3 * side - 2) ** 2 + extra * N
    for r in (0, 1, side, N // 2, N - 1):
        i, j = divmod(r, side)
        row = ci[rs[r]:rs[r + 1]]
        vals = dv[rs[r]:rs[r + 1]]
        assert np.all(np.diff(row) > 0)
        sten = {(ii * side + jj) for ii in range(i - 1, i + 2) for jj in range(j - 1, j + 2)
                if 0 <= ii < side and 0 <= jj < side}
        assert sten <= set(row.tolist())
        assert len(set(row.tolist()) - sten) == extra
        for c, v in zip(row, vals):
            if c == r:
                assert v == 8.0
            elif c in sten:
                assert v == -1.0
            else:
                assert -1 <= v < 1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="forks the test process: run only where no GPU is present")
def test_parallel_loops_after_fork():
    """The library's host thread pool belongs to the process that made it: a
    forked child runs its parallel loops on threads of its own (same result,
    no wait on the parent's workers).  Host code only, no device."""
    a = pw.gen_powerlaw(200_000, struct_seed=1, value_seed=2, cap=500)  # starts the pool
    pid = os.fork()
    if pid == 0:
        try:
            b = pw.gen_powerlaw(200_000, struct_seed=1, value_seed=2, cap=500)
            os._exit(0 if all(np.array_equal(u, v) for u, v in zip(a, b)) else 3)
        except BaseException:
            os._exit(4)
    _, status = os.waitpid(pid, 0)
    assert os.WIFEXITED(status) and os.WEXITSTATUS(status) == 0
