"""numpy front end of the CPU oracle (liboracle.so, built from pwm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker.  The product package
(thesis_spmv_amd) never imports it.

Each wrapper names the reference function it restates (file:line relative to
the reference checkout); the C file holds the restatement and its pinning.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_BASELINE_PATH = os.path.join(_HERE, "liboracle_baseline.so")
_lib = None
_baseline = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
I64 = C.c_int64


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    """The checker (-O2, no FMA)."""
    global _lib
    if _lib is None:
        _lib = _load(_LIB_PATH)
    return _lib


def baseline_lib():
    """The same restatements built at -O3 (the reference's optimisation
    level), for bench.py's CPU baseline timings only."""
    global _baseline
    if _baseline is None:
        _baseline = _load(_BASELINE_PATH)
    return _baseline


def _load(path):
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    sig = {
        "orc_integer_ceil": (I64, [I64, I64]),
        "orc_binary_search_leftmost": (I64, [i32p, C.c_int32, I64]),
        "orc_norm2": (C.c_double, [f64p, I64]),
        "orc_normalize": (None, [f64p, I64]),
        "orc_poisson_nnz": (I64, [I64, I64]),
        "orc_fill_poisson": (None, [I64, I64, i32p, i32p, f64p]),
        "orc_triplet_poisson": (None, [I64, I64, i32p, i32p, f64p]),
        "orc_load_mm": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(I64), C.POINTER(I64),
                                  C.POINTER(I64), C.POINTER(C.POINTER(C.c_int32)),
                                  C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_double))]),
        "orc_load_bin": (C.c_int, [C.c_char_p, I64, C.c_int, C.c_int, C.POINTER(I64), C.POINTER(I64),
                                   C.POINTER(I64), C.POINTER(C.POINTER(C.c_int32)),
                                   C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_double))]),
        "orc_free": (None, [C.c_void_p]),
        "orc_mt19937_seed": (None, [u32p, C.c_uint32]),
        "orc_mt19937_next": (C.c_uint32, [u32p]),
        "orc_uniform_real": (C.c_double, [u32p, C.c_double, C.c_double]),
        "orc_triplet_to_crs": (None, [I64, I64, i32p, i32p, f64p, i32p, i32p, f64p]),
        "orc_row_partitions": (None, [I64, C.c_int, i32p, i32p]),
        "orc_triplet_to_multiple_crs": (None, [I64, I64, i32p, i32p, f64p, C.c_int, i32p, i32p, i64p, i32p, i32p,
                                               f64p]),
        "orc_crs_mv": (None, [I64, i32p, i32p, f64p, f64p, f64p]),
        "orc_crs_power_iteration": (None, [I64, i32p, i32p, f64p, f64p, f64p, I64]),
        "orc_crsomp_mv": (None, [C.c_int, I64, i32p, i32p, f64p, f64p, f64p]),
        "orc_crsomp_power_iteration": (None, [C.c_int, I64, i32p, i32p, f64p, f64p, f64p, I64]),
        "orc_merge_search": (None, [I64, i32p, I64, I64, C.POINTER(I64), C.POINTER(I64)]),
        "orc_merge_splits": (None, [I64, I64, I64, i32p, i64p, i64p]),
        "orc_merge_mv": (None, [I64, C.c_int, I64, I64, i32p, i32p, f64p, f64p, f64p]),
        "orc_merge_power_iteration": (None, [I64, C.c_int, I64, I64, i32p, i32p, f64p, f64p, f64p, I64]),
        "orc_csb_block_bits": (C.c_int, [I64, I64, C.c_int]),
        "orc_zmorton_less_equal": (C.c_int, [C.c_uint32] * 4 + [C.c_int]),
        "orc_csb_build": (C.c_int, [I64, I64, I64, i32p, i32p, f64p, C.c_int, u32p, f64p, i32p]),
        "orc_csb_mv": (None, [I64, I64, C.c_int, u32p, f64p, i32p, f64p, f64p]),
        "orc_hilbert": (I64, [I64, I64, I64]),
        "orc_blockcoh_threads": (C.c_int, [I64, C.c_int]),
        "orc_blockcoh_bands": (None, [I64, I64, i32p, C.c_int, i32p, i64p]),
        "orc_blockcoh_beta_bits": (C.c_int, [I64, I64]),
        "orc_blockcoh_hilbert_size": (I64, [I64, I64]),
        "orc_blockcoh_build_band": (None, [I64, i32p, i32p, f64p, I64, I64, C.c_void_p]),
        "orc_blockcoh_mv_band": (None, [C.c_void_p, I64, I64, f64p, f64p]),
        "orc_matrix_stats": (None, [I64, i32p, i32p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(I64), C.POINTER(C.c_int), C.POINTER(I64),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


# ---------------------------------------------------------------- matrices
class CRSMatrix:
    """Plain CRS arrays (row_start int32[N+1], col_ind int32[Z], data f64[Z])."""

    def __init__(self, nor, noc, row_start, col_ind, data):
        self.nor, self.noc = int(nor), int(noc)
        self.row_start = np.ascontiguousarray(row_start, dtype=np.int32)
        self.col_ind = np.ascontiguousarray(col_ind, dtype=np.int32)
        self.data = np.ascontiguousarray(data, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.row_start[-1])


class Triplets:
    def __init__(self, nor, noc, rows, cols, vals):
        self.nor, self.noc = int(nor), int(noc)
        self.rows = np.ascontiguousarray(rows, dtype=np.int32)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)

    @property
    def nnz(self):
        return len(self.rows)


def poisson_crs(m, n):
    """Util/Poisson.hpp:32-83 via CRS::generatePoissonMatrix (Matrix/CRS.hpp:60-74)."""
    L = lib()
    z = L.orc_poisson_nnz(m, n)
    rs = np.zeros(m * n + 1, np.int32)
    ci = np.zeros(z, np.int32)
    d = np.zeros(z, np.float64)
    L.orc_fill_poisson(m, n, rs, ci, d)
    return CRSMatrix(m * n, m * n, rs, ci, d)


def poisson_triplets(m, n):
    """Matrix/Triplet.hpp:226-278."""
    L = lib()
    z = L.orc_poisson_nnz(m, n)
    r = np.zeros(z, np.int32)
    c = np.zeros(z, np.int32)
    v = np.zeros(z, np.float64)
    L.orc_triplet_poisson(m, n, r, c, v)
    return Triplets(m * n, m * n, r, c, v)


def _take(nr, nc, nz, r, c, v):
    n = nz.value
    rows = np.ctypeslib.as_array(r, shape=(max(n, 1),))[:n].copy()
    cols = np.ctypeslib.as_array(c, shape=(max(n, 1),))[:n].copy()
    vals = np.ctypeslib.as_array(v, shape=(max(n, 1),))[:n].copy()
    L = lib()
    for p in (r, c, v):
        L.orc_free(C.cast(p, C.c_void_p))
    return Triplets(nr.value, nc.value, rows, cols, vals)


def load_mm(path, has_data, symmetric, random_fill=False):
    """Matrix/Triplet.hpp:75-154."""
    L = lib()
    nr, nc, nz = I64(), I64(), I64()
    r, c, v = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)()
    rc = L.orc_load_mm(path.encode(), int(has_data), int(symmetric), int(random_fill),
                       C.byref(nr), C.byref(nc), C.byref(nz), C.byref(r), C.byref(c), C.byref(v))
    if rc != 0:
        raise OSError(f"orc_load_mm({path}) failed: {rc}")
    return _take(nr, nc, nz, r, c, v)


def load_bin(path, mat_size, symmetric, random_vals):
    """Matrix/Triplet.hpp:164-216."""
    L = lib()
    nr, nc, nz = I64(), I64(), I64()
    r, c, v = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)()
    rc = L.orc_load_bin(path.encode(), mat_size, int(symmetric), int(random_vals),
                        C.byref(nr), C.byref(nc), C.byref(nz), C.byref(r), C.byref(c), C.byref(v))
    if rc != 0:
        raise OSError(f"orc_load_bin({path}) failed: {rc}")
    return _take(nr, nc, nz, r, c, v)


def triplet_to_crs(t):
    """Util/TripletToCRS.hpp:135-172."""
    rs = np.zeros(t.nor + 1, np.int32)
    ci = np.zeros(t.nnz, np.int32)
    d = np.zeros(t.nnz, np.float64)
    lib().orc_triplet_to_crs(t.nor, t.nnz, t.rows, t.cols, t.vals, rs, ci, d)
    return CRSMatrix(t.nor, t.noc, rs, ci, d)


# ---------------------------------------------------------------- operators
def triplet_to_multiple_crs(t, partitions):
    """Util/TripletToCRS.hpp:309-385: a list of (first_row, CRSMatrix) per
    partition, each CRS local to its rows."""
    P = int(partitions)
    fr, tr = np.zeros(P, np.int32), np.zeros(P, np.int32)
    off = np.zeros(P + 1, np.int64)
    rs = np.zeros(t.nor + P, np.int32)
    ci = np.zeros(max(t.nnz, 1), np.int32)
    dv = np.zeros(max(t.nnz, 1), np.float64)
    lib().orc_triplet_to_multiple_crs(t.nor, t.nnz, t.rows, t.cols, t.vals, P, fr, tr, off, rs, ci, dv)
    out = []
    for p in range(P):
        f, n = int(fr[p]), int(tr[p])
        out.append((f, CRSMatrix(n, t.noc, rs[f + p:f + p + n + 1], ci[off[p]:off[p + 1]], dv[off[p]:off[p + 1]])))
    return out


def crs_mv(A, x):
    """Matrix/CRS.hpp:99-109."""
    y = np.zeros(A.nor, np.float64)
    lib().orc_crs_mv(A.nor, A.row_start, A.col_ind, A.data, np.ascontiguousarray(x, np.float64), y)
    return y


def crs_power_iteration(A, x, it):
    """Matrix/CRS.hpp:118-130.  Returns (x, y) after the call."""
    x = np.array(x, np.float64, copy=True)
    y = np.zeros(A.nor, np.float64)
    lib().orc_crs_power_iteration(A.nor, A.row_start, A.col_ind, A.data, x, y, it)
    return x, y


def crsomp_mv(A, x, threads):
    """Env_Implementations/CRSOMP.hpp:110-123."""
    y = np.zeros(A.nor, np.float64)
    lib().orc_crsomp_mv(threads, A.nor, A.row_start, A.col_ind, A.data, np.ascontiguousarray(x, np.float64), y)
    return y


def merge_search(A, diagonal):
    """SOA_Implementations/CRS_Merge.hpp:42-62."""
    a, b = I64(), I64()
    lib().orc_merge_search(diagonal, A.row_start[1:].copy(), A.nor, A.nnz, C.byref(a), C.byref(b))
    return a.value, b.value


def merge_splits(A, P):
    """CRS_Merge.hpp:96-112 split points for P threads (P+1 entries)."""
    a = np.zeros(P + 1, np.int64)
    b = np.zeros(P + 1, np.int64)
    lib().orc_merge_splits(P, A.nor, A.nnz, A.row_start, a, b)
    return a, b


def merge_mv(A, x, P, omp_threads=1):
    """CRS_Merge.hpp:95-141."""
    y = np.zeros(A.nor, np.float64)
    lib().orc_merge_mv(P, omp_threads, A.nor, A.nnz, A.row_start, A.col_ind, A.data,
                       np.ascontiguousarray(x, np.float64), y)
    return y


def csb_block_bits(nor, noc, value_size=8):
    """SOA_Implementations/CSB.hpp:96-113."""
    return lib().orc_csb_block_bits(nor, noc, value_size)


def csb_build(t, block_bits=0):
    """CSB.hpp:464-567.  Returns (bits, ind, data, blk_ptr, balanced)."""
    L = lib()
    bits = block_bits if block_bits > 0 else L.orc_csb_block_bits(t.nor, t.noc, 8)
    beta = 1 << bits
    H = -(-t.noc // beta)
    V = -(-t.nor // beta)
    ind = np.zeros(max(t.nnz, 1), np.uint32)
    data = np.zeros(max(t.nnz, 1), np.float64)
    blk = np.zeros(H * V + 1, np.int32)
    bal = L.orc_csb_build(t.nor, t.noc, t.nnz, t.rows, t.cols, t.vals, bits, ind, data, blk)
    return bits, ind[:t.nnz], data[:t.nnz], blk, bal


def csb_mv(nor, noc, bits, ind, data, blk_ptr, x):
    """CSB.hpp:593-605 with sparseRowMult :348-364."""
    y = np.zeros(nor, np.float64)
    lib().orc_csb_mv(nor, noc, bits, np.ascontiguousarray(ind, np.uint32), np.ascontiguousarray(data),
                     np.ascontiguousarray(blk_ptr, np.int32), np.ascontiguousarray(x, np.float64), y)
    return y


def hilbert(size, row, col):
    """SOA_Implementations/BlockCOH.hpp:103-114."""
    return lib().orc_hilbert(size, row, col)


class _CohBand(C.Structure):
    _fields_ = [("nnz", I64), ("blocks", I64), ("row_jumps", I64), ("block_row_jumps", I64),
                ("beta_bits", C.c_int32), ("horizontal_blocks", C.c_int32), ("vertical_blocks", C.c_int32),
                ("row_jump", C.c_void_p), ("col_jump", C.c_void_p), ("data", C.c_void_p),
                ("row_jump_block", C.c_void_p), ("col_jump_block", C.c_void_p), ("block_nnz", C.c_void_p)]


def blockcoh_partition(A, threads):
    """BlockCOH.hpp:325-355: (threads_used, band_row_start[int32], band_nnz[int64])."""
    L = lib()
    P = L.orc_blockcoh_threads(A.nor, threads)
    counts = np.diff(A.row_start).astype(np.int32)
    starts = np.zeros(P, np.int32)
    nnz = np.zeros(P, np.int64)
    L.orc_blockcoh_bands(A.nor, A.nnz, counts, P, starts, nnz)
    return P, starts, nnz


def blockcoh_build(A, threads):
    """BlockCOH.hpp:319-455 on a CRS-ordered matrix.  Returns a list of band dicts."""
    L = lib()
    P, starts, nnzs = blockcoh_partition(A, threads)
    bands = []
    for p in range(P):
        s = int(starts[p])
        e = int(starts[p + 1]) if p + 1 < P else A.nor
        rows = e - s
        z = int(A.row_start[e] - A.row_start[s]) if rows > 0 else 0
        arr = {
            "row_jump": np.zeros(max(z, 1), np.uint16),
            "col_jump": np.zeros(max(z, 1), np.uint16),
            "data": np.zeros(max(z, 1), np.float64),
            "row_jump_block": np.zeros(z + 1, np.int16),
            "col_jump_block": np.zeros(z + 1, np.int16),
            "block_nnz": np.zeros(z + 1, np.int32),
        }
        st = _CohBand()
        for k, v in arr.items():
            setattr(st, k, v.ctypes.data)
        L.orc_blockcoh_build_band(A.noc, A.row_start, A.col_ind, A.data, s, rows, C.byref(st))
        band = {
            "start": s, "rows": rows, "nnz": st.nnz, "blocks": st.blocks, "beta_bits": st.beta_bits,
            "horizontal_blocks": st.horizontal_blocks, "vertical_blocks": st.vertical_blocks,
            "row_jump": arr["row_jump"][:st.row_jumps].copy(),
            "col_jump": arr["col_jump"][:z].copy(),
            "data": arr["data"][:z].copy(),
            "row_jump_block": arr["row_jump_block"][:st.block_row_jumps].copy(),
            "col_jump_block": arr["col_jump_block"][:st.blocks].copy(),
            "block_nnz": arr["block_nnz"][:st.blocks].copy(),
            "_struct": st, "_arrays": arr,
        }
        assert band["nnz"] == nnzs[p]
        bands.append(band)
    return bands


def blockcoh_mv(bands, nor, x):
    """BlockCOH.hpp:479-510 with blockMult :235-256."""
    y = np.zeros(nor, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    for b in bands:
        lib().orc_blockcoh_mv_band(C.byref(b["_struct"]), b["start"], b["rows"], xx, y)
    return y


def norm2(x):
    """Util/VectorUtill.hpp:23-26."""
    return lib().orc_norm2(np.ascontiguousarray(x, np.float64), len(x))


# ---------------------------------------------------------------- statistics
def matrix_stats(A):
    """Util_Scripts/calc_stats.cpp:44-105 on CRS arrays (see pwm_oracle.c)."""
    ra, rv, bav, bvr = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    rmax, nb = I64(), I64()
    bits, bal = C.c_int(), C.c_int()
    lib().orc_matrix_stats(A.nor, A.row_start, A.col_ind, C.byref(ra), C.byref(rv), C.byref(rmax), C.byref(bits),
                           C.byref(nb), C.byref(bav), C.byref(bvr), C.byref(bal))
    return {"rows_avg": ra.value, "rows_var": rv.value, "row_max": rmax.value, "block_bits": bits.value,
            "nb_blocks": nb.value, "block_avg": bav.value, "block_var": bvr.value, "balanced": bal.value}
