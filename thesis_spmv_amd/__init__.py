"""MI355X (gfx950) SpMV and power-method library: Python host mirror.

The classes here mirror the reference's operator interface
pwm::SparseMatrix<T,int_type> (Matrix/SparseMatrix.hpp:13-68): the same four
operations under the same names -- generatePoissonMatrix, loadFromTriplets,
mv, powerIteration -- with `mult` / `powerMethod` as aliases, and a factory
select_type(method, threads) beside the reference's selectType
(Util/DriverUtil.hpp:38-84).  Every call goes through the C ABI of
libpwmhip.so (include/pwm_hip.h); there is no CPU fallback: if the library
is not built, importing a class raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpwmhip.so")

FORMAT_CRS, FORMAT_CRS_MERGE, FORMAT_CSB, FORMAT_BLOCKCOH = 1, 8, 9, 10
VALUES = {"f64": 0, "f32": 1, "bf16": 2}
ERRORS = {-1: "invalid argument", -2: "I/O", -3: "HIP error", -4: "out of device memory",
          -5: "not square", -6: "no matrix loaded", -7: "HBM budget", -8: "unsupported"}

_lib = None
I64, I32, F64 = C.c_int64, C.c_int32, C.c_double
VP = C.c_void_p


class PwmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [("nor", I64), ("nnz", I64), ("rows_avg", F64), ("rows_var", F64), ("row_max", I64),
                ("block_bits", I32), ("balanced", I32), ("nb_blocks", I64), ("block_avg", F64), ("block_var", F64)]


class Info(C.Structure):
    _fields_ = [("nor", I64), ("noc", I64), ("nnz", I64), ("format", I32), ("value_type", I32),
                ("threads", I32), ("group_size", I32), ("device_bytes", I64), ("long_rows", I64),
                ("block_bits", I32), ("balanced", I32), ("unroll", I32), ("non_temporal", I32), ("layout", I32),
                ("kernels", I32), ("high_occupancy", I32)]


_SIGS = {
    "pwm_create": [C.c_int, C.c_int, C.c_int, C.POINTER(VP)],
    "pwm_destroy": [VP],
    "pwm_last_error": [VP],
    "pwm_generate_poisson": [VP, I64, I64, C.c_int],
    "pwm_load_triplets": [VP, I64, I64, I64, VP, VP, VP, C.c_int],
    "pwm_load_crs": [VP, I64, I64, VP, VP, VP, C.c_int],
    "pwm_mv": [VP, VP, VP],
    "pwm_power_iteration": [VP, VP, VP, I64],
    "pwm_info": [VP, C.POINTER(Info)],
    "pwm_set_option": [VP, C.c_char_p, I64],
    "pwm_mv_device": [VP, VP, VP],
    "pwm_power_method_device": [VP, VP, VP, I64, F64, C.POINTER(F64), C.POINTER(F64), C.POINTER(I64)],
    "pwm_power_method_device_noalloc": [VP, VP, VP, I64, F64, C.POINTER(F64), C.POINTER(F64), C.POINTER(I64)],
    "pwm_merge_splits": [VP, C.POINTER(I64), VP, VP],
    "pwm_crs_export": [VP, VP, VP, VP],
    "pwm_crs_partitions": [VP, C.c_int, VP, VP, VP],
    "pwm_crs_partition_export": [VP, C.c_int, C.c_int, VP, VP, VP],
    "pwm_csb_layout": [VP, C.POINTER(I32), C.POINTER(I64), VP, VP, VP],
    "pwm_blockcoh_band": [VP, C.c_int, VP, VP, VP, VP, VP, VP, VP, VP],
    "pwm_device_alloc": [C.c_size_t, C.POINTER(VP)],
    "pwm_device_free": [VP],
    "pwm_memcpy_h2d": [VP, VP, C.c_size_t],
    "pwm_memcpy_d2h": [VP, VP, C.c_size_t],
    "pwm_device_synchronize": [],
    "pwm_device_fill": [VP, I64, C.c_int, F64],
    "pwm_device_copy": [VP, VP, I64],
    "pwm_time_mv": [VP, VP, VP, C.c_int, I64, C.POINTER(F64), C.POINTER(F64)],
    "pwm_time_mv_rotating": [C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), C.c_int, C.c_int, C.POINTER(F64),
                             C.POINTER(F64)],
    "pwm_matrix_stats": [VP, VP],
    "pwm_gather_rate": [VP, I64, I64, C.c_int, I64, C.POINTER(F64)],
    "pwm_gen_powerlaw": [I64, F64, F64, I64, C.c_uint64, C.c_uint64, C.POINTER(VP), C.POINTER(VP),
                         C.POINTER(VP), C.POINTER(I64)],
    "pwm_gen_stencil9": [I64, I64, C.c_uint64, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), C.POINTER(I64)],
    "pwm_gen_uniform": [I64, C.c_uint64, VP],
    "pwm_gen_spd_powerlaw": [I64, C.c_uint64, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), C.POINTER(I64)],
    "pwm_free": [VP],
    "pwm_read_mm": [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64),
                    C.POINTER(VP), C.POINTER(VP), C.POINTER(VP)],
    "pwm_read_bin": [C.c_char_p, I64, C.c_int, C.c_int, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64),
                     C.POINTER(VP), C.POINTER(VP), C.POINTER(VP)],
    "pwm_batch_create": [C.c_int, C.POINTER(VP)],
    "pwm_batch_destroy": [VP],
    "pwm_batch_last_error": [VP],
    "pwm_batch_add_crs": [VP, I64, VP, VP, VP, C.POINTER(I32)],
    "pwm_batch_add_poisson": [VP, I64, I64, C.POINTER(I32)],
    "pwm_batch_solve": [VP, I64, F64, C.c_uint64, VP, VP, C.POINTER(F64)],
    "pwm_batch_start_vector": [VP, I32, C.c_uint64, VP],
    "pwm_batch_vector": [VP, I32, VP, C.POINTER(F64)],
    "pwm_batch_size": [VP, C.POINTER(I32), C.POINTER(I64), C.POINTER(I64)],
    "pwm_batch_set_option": [VP, C.c_char_p, I64],
    "pwm_batch_stats": [VP, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64)],
    "pwm_batch_schedule": [VP, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32), VP],
    "pwm_server_create": [C.c_int, I64, I32, C.POINTER(VP)],
    "pwm_server_destroy": [VP],
    "pwm_server_last_error": [VP],
    "pwm_server_set_option": [VP, C.c_char_p, I64],
    "pwm_server_add_matrix": [VP, I64, VP, VP, VP, I32, C.POINTER(I32)],
    "pwm_server_add_poisson": [VP, I64, I64, C.POINTER(I32)],
    "pwm_server_remove_matrix": [VP, I32],
    "pwm_server_matrix_resident": [VP, I32, C.POINTER(I32)],
    "pwm_server_submit": [VP, I32, I64, F64, C.c_uint64, C.POINTER(I64)],
    "pwm_server_step": [VP, C.POINTER(I32), C.POINTER(I32)],
    "pwm_server_finished": [VP, VP, I32, C.POINTER(I32)],
    "pwm_server_timing": [VP, C.POINTER(F64), C.POINTER(F64), C.POINTER(F64), C.POINTER(I64)],
    "pwm_server_result": [VP, I64, C.POINTER(F64), C.POINTER(F64), C.POINTER(I64), VP],
    "pwm_server_stats": [VP, C.POINTER(I64), C.POINTER(I32), C.POINTER(I32), C.POINTER(I32), C.POINTER(I64)],
}
_CHAR_RESULT = {"pwm_last_error", "pwm_batch_last_error", "pwm_server_last_error"}


def lib():
    """Load libpwmhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name in _CHAR_RESULT else (None if name == "pwm_free" else C.c_int)
        _lib = L
    return _lib


def _check(rc, h=None):
    if rc != 0:
        msg = lib().pwm_last_error(h)
        raise PwmError(rc, msg.decode() if msg else "")


def _ptr(a):
    return a.ctypes.data_as(VP)


# ------------------------------------------------------------ device memory
class DeviceVector:
    """A vector in HBM (double for FP64 handles, float otherwise)."""

    def __init__(self, n, dtype=np.float64):
        self.n, self.dtype = int(n), np.dtype(dtype)
        p = VP()
        _check(lib().pwm_device_alloc(max(self.n, 1) * self.dtype.itemsize, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_array(cls, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        v = cls(len(a), dtype)
        v.upload(a)
        return v

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert len(a) == self.n
        _check(lib().pwm_memcpy_h2d(self.ptr, _ptr(a), a.nbytes))

    def download(self):
        out = np.empty(self.n, self.dtype)
        _check(lib().pwm_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def fill(self, v):
        _check(lib().pwm_device_fill(self.ptr, self.n, int(self.dtype == np.float64), float(v)))

    def free(self):
        if self.ptr:
            lib().pwm_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def time_mv_rotating(mats, xs, ys, reps, per_step=True):
    """(ms per SpMV, ms of its dominant kernel or None) of `reps` rounds of one
    SpMV on each handle in `mats` in turn, back to back on one stream
    (pwm_time_mv_rotating): with enough independent matrices, none of them is
    still in the Infinity Cache when its turn comes."""
    k = len(mats)
    hs = (VP * k)(*[m.h for m in mats])
    xp = (VP * k)(*[x.ptr for x in xs])
    yp = (VP * k)(*[y.ptr for y in ys])
    ms, kms = F64(), F64()
    _check(lib().pwm_time_mv_rotating(hs, xp, yp, k, int(reps), C.byref(ms), C.byref(kms) if per_step else None))
    return ms.value, (kms.value if per_step else None)


def gather_rate(x, window, per_xcd, n):
    """Milliseconds of one pass of the gather-rate probe (pwm_gather_rate):
    n 12-byte (index, value) entries, each gathering one double of the FP64
    DeviceVector x from a window of `window` doubles (per XCD with per_xcd);
    window 0 (x may be None): the entry stream alone."""
    if x is not None and x.dtype != np.float64:
        raise ValueError("gather_rate needs an FP64 DeviceVector")
    ms = F64()
    _check(lib().pwm_gather_rate(x.ptr if x is not None else None, int(x.n) if x is not None else 0, int(window),
                                 int(bool(per_xcd)), int(n), C.byref(ms)))
    return ms.value


def synchronize():
    _check(lib().pwm_device_synchronize())


# ------------------------------------------------------------ operators
class SparseMatrixHip:
    """GPU counterpart of pwm::SparseMatrix<double, int_type>
    (Matrix/SparseMatrix.hpp:13-68).  `threads` is the reference's constructor
    argument (:30); on the GPU it sets the partition count of the formats
    that have one (merge-path tiles, BlockCOH bands)."""

    FORMAT = None

    def __init__(self, threads=-1, values="f64", **options):
        self.h = VP()
        _check(lib().pwm_create(self.FORMAT, int(threads), VALUES[values], C.byref(self.h)))
        self.values = values
        for k, v in options.items():
            self.set_option(k, v)

    def __del__(self):
        try:
            if self.h:
                lib().pwm_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def vector_dtype(self):
        return np.float64 if self.values == "f64" else np.float32

    def set_option(self, key, value):
        _check(lib().pwm_set_option(self.h, key.encode(), int(value)), self.h)

    # SparseMatrix::generatePoissonMatrix (:41)
    def generatePoissonMatrix(self, m, n, partitions=1):
        _check(lib().pwm_generate_poisson(self.h, int(m), int(n), int(partitions)), self.h)

    # SparseMatrix::loadFromTriplets (:48).  `t` has row_coord/col_coord/data
    # (or rows/cols/vals) and row_size/col_size (or nor/noc), as the
    # reference's Triplet (Matrix/Triplet.hpp:26-45).
    def loadFromTriplets(self, t, partitions=1):
        rows = np.ascontiguousarray(getattr(t, "row_coord", getattr(t, "rows", None)), dtype=np.int32)
        cols = np.ascontiguousarray(getattr(t, "col_coord", getattr(t, "cols", None)), dtype=np.int32)
        vals = np.ascontiguousarray(getattr(t, "data", getattr(t, "vals", None)), dtype=np.float64)
        nor = int(getattr(t, "row_size", getattr(t, "nor", 0)))
        noc = int(getattr(t, "col_size", getattr(t, "noc", 0)))
        _check(lib().pwm_load_triplets(self.h, nor, noc, len(rows), _ptr(rows), _ptr(cols), _ptr(vals),
                                       int(partitions)), self.h)

    def loadCRS(self, nor, noc, row_start, col_ind, data, partitions=1):
        rs = np.ascontiguousarray(row_start, dtype=np.int32)
        ci = np.ascontiguousarray(col_ind, dtype=np.int32)
        dv = np.ascontiguousarray(data, dtype=np.float64)
        _check(lib().pwm_load_crs(self.h, int(nor), int(noc), _ptr(rs), _ptr(ci), _ptr(dv), int(partitions)),
               self.h)

    # SparseMatrix::mv (:57): y = A x on host arrays (y is written in place)
    def mv(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert isinstance(y, np.ndarray) and y.dtype == np.float64 and y.flags.c_contiguous
        _check(lib().pwm_mv(self.h, _ptr(x), _ptr(y)), self.h)

    mult = mv

    # SparseMatrix::powerIteration (:66): x and y are updated in place
    def powerIteration(self, x, y, it):
        for a in (x, y):
            assert isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
        _check(lib().pwm_power_iteration(self.h, _ptr(x), _ptr(y), int(it)), self.h)

    powerMethod = powerIteration

    # device-resident additions
    def mv_device(self, x, y):
        _check(lib().pwm_mv_device(self.h, x.ptr, y.ptr), self.h)

    def power_method_device(self, x, y, max_it, tol=-1.0):
        lam, nrm, its = F64(), F64(), I64()
        _check(lib().pwm_power_method_device(self.h, x.ptr, y.ptr, int(max_it), float(tol), C.byref(lam),
                                             C.byref(nrm), C.byref(its)), self.h)
        return lam.value, nrm.value, its.value

    def power_method_device_noalloc(self, x, y, max_it, tol=-1.0):
        """power_method_device with the library's fixed per-GPU workspace:
        allocates no device memory (pwm_power_method_device_noalloc)."""
        lam, nrm, its = F64(), F64(), I64()
        _check(lib().pwm_power_method_device_noalloc(self.h, x.ptr, y.ptr, int(max_it), float(tol), C.byref(lam),
                                                     C.byref(nrm), C.byref(its)), self.h)
        return lam.value, nrm.value, its.value

    def time_mv(self, x, y, reps, flush_bytes=0, per_step=True):
        """(ms per SpMV, ms of its dominant kernel) over `reps` SpMVs timed
        with HIP events per step; per_step=False: the SpMVs run back to back
        between one pair of events and the kernel time is None."""
        ms, kms = F64(), F64()
        _check(lib().pwm_time_mv(self.h, x.ptr, y.ptr, int(reps), int(flush_bytes), C.byref(ms),
                                 C.byref(kms) if per_step else None), self.h)
        return ms.value, (kms.value if per_step else None)

    def crs_export(self):
        """The CRS arrays held on the device (values widened to double)."""
        inf = self.info()
        rs = np.zeros(inf["nor"] + 1, np.int32)
        ci = np.zeros(max(inf["nnz"], 1), np.int32)
        dv = np.zeros(max(inf["nnz"], 1), np.float64)
        _check(lib().pwm_crs_export(self.h, _ptr(rs), _ptr(ci), _ptr(dv)), self.h)
        return rs, ci[:inf["nnz"]], dv[:inf["nnz"]]

    def stats(self):
        """Matrix statistics of the loaded matrix (Util_Scripts/calc_stats.cpp)."""
        st = Stats()
        _check(lib().pwm_matrix_stats(self.h, C.byref(st)), self.h)
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def partitions(self, partitions):
        """CRS and CRS_Merge handles: TripletToMultipleCRS's row plan (Util/TripletToCRS.hpp:320-330):
        (first_rows, thread_rows, nonzeros) per partition."""
        P = int(partitions)
        f, r, z = np.zeros(P, np.int32), np.zeros(P, np.int32), np.zeros(P, np.int64)
        _check(lib().pwm_crs_partitions(self.h, P, _ptr(f), _ptr(r), _ptr(z)), self.h)
        return f, r, z

    def partition_crs(self, partitions, part):
        """Partition `part`'s local CRS arrays (row_start, col_ind, values)."""
        f, r, z = self.partitions(partitions)
        rs = np.zeros(int(r[part]) + 1, np.int32)
        ci = np.zeros(max(int(z[part]), 1), np.int32)
        dv = np.zeros(max(int(z[part]), 1), np.float64)
        _check(lib().pwm_crs_partition_export(self.h, int(partitions), int(part), _ptr(rs), _ptr(ci), _ptr(dv)), self.h)
        return rs, ci[:int(z[part])], dv[:int(z[part])]

    def info(self):
        i = Info()
        _check(lib().pwm_info(self.h, C.byref(i)), self.h)
        return {k: getattr(i, k) for k, _ in Info._fields_}

    @property
    def nor(self):
        return self.info()["nor"]


class CRSHip(SparseMatrixHip):
    """CRS (beside CRS / CRSOMP, Matrix/CRS.hpp, Env_Implementations/CRSOMP.hpp): the
    vector kernel or the row-aligned tiles kernel, whichever timed faster at
    load (option crs_kernel: 1 vector, 2 tiles)."""
    FORMAT = FORMAT_CRS


class CRSMergeHip(SparseMatrixHip):
    """Merge-path CRS (beside CRS_Merge, SOA_Implementations/CRS_Merge.hpp)."""
    FORMAT = FORMAT_CRS_MERGE

    def merge_splits(self):
        P = I64()
        _check(lib().pwm_merge_splits(self.h, C.byref(P), None, None), self.h)
        a = np.zeros(P.value + 1, np.int64)
        b = np.zeros(P.value + 1, np.int64)
        _check(lib().pwm_merge_splits(self.h, C.byref(P), _ptr(a), _ptr(b)), self.h)
        return a, b


class CSBHip(SparseMatrixHip):
    """Compressed sparse blocks (beside CSB, SOA_Implementations/CSB.hpp)."""
    FORMAT = FORMAT_CSB

    def csb_layout(self):
        """(block_bits, ind, data, blk_ptr) as stored on the device."""
        bits, nb = I32(), I64()
        _check(lib().pwm_csb_layout(self.h, C.byref(bits), C.byref(nb), None, None, None), self.h)
        z = self.info()["nnz"]
        ind = np.zeros(max(z, 1), np.uint32)
        data = np.zeros(max(z, 1), np.float64)
        blk = np.zeros(nb.value + 1, np.int32)
        _check(lib().pwm_csb_layout(self.h, C.byref(bits), C.byref(nb), _ptr(ind), _ptr(data), _ptr(blk)), self.h)
        return bits.value, ind[:z], data[:z], blk


class BlockCOHHip(SparseMatrixHip):
    """Row-distributed block CO-H (beside BlockCOH, SOA_Implementations/BlockCOH.hpp)."""
    FORMAT = FORMAT_BLOCKCOH

    def band(self, p):
        """Band p's stored arrays (BlockCOH.hpp:27-67) and sizes."""
        counts = np.zeros(6, np.int64)
        meta = np.zeros(3, np.int32)
        L = lib()
        _check(L.pwm_blockcoh_band(self.h, int(p), _ptr(counts), _ptr(meta), None, None, None, None, None, None),
               self.h)
        nnz, blocks, rj, rjb, start, rows = (int(v) for v in counts)
        a = {"row_jump": np.zeros(max(rj, 1), np.uint16), "col_jump": np.zeros(max(nnz, 1), np.uint16),
             "data": np.zeros(max(nnz, 1), np.float64), "row_jump_block": np.zeros(max(rjb, 1), np.int16),
             "col_jump_block": np.zeros(max(blocks, 1), np.int16), "block_nnz": np.zeros(max(blocks, 1), np.int32)}
        _check(L.pwm_blockcoh_band(self.h, int(p), _ptr(counts), _ptr(meta), _ptr(a["row_jump"]),
                                   _ptr(a["col_jump"]), _ptr(a["data"]), _ptr(a["row_jump_block"]),
                                   _ptr(a["col_jump_block"]), _ptr(a["block_nnz"])), self.h)
        return {"nnz": nnz, "blocks": blocks, "start": start, "rows": rows, "beta_bits": int(meta[0]),
                "horizontal_blocks": int(meta[1]), "vertical_blocks": int(meta[2]),
                "row_jump": a["row_jump"][:rj], "col_jump": a["col_jump"][:nnz], "data": a["data"][:nnz],
                "row_jump_block": a["row_jump_block"][:rjb], "col_jump_block": a["col_jump_block"][:blocks],
                "block_nnz": a["block_nnz"][:blocks]}


GPU_METHODS = {101: CRSHip, 102: CRSHip, 108: CRSMergeHip, 109: CSBHip, 110: BlockCOHHip}


def select_type(method, threads=-1, values="f64"):
    """selectType (Util/DriverUtil.hpp:38-84) for the GPU classes: the
    reference's ids + 100 (101/102 CRS, 108 CRS_Merge, 109 CSB, 110 BlockCOH).
    Unknown ids return None, as the reference returns NULL."""
    cls = GPU_METHODS.get(int(method))
    return cls(threads, values) if cls else None


# ------------------------------------------------------------ generators
def _take_crs(n, rs, ci, dv, nnz):
    L = lib()
    z = nnz.value
    rsa = np.ctypeslib.as_array(C.cast(rs, C.POINTER(C.c_int32)), shape=(n + 1,)).copy()
    cia = np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_int32)), shape=(max(z, 1),))[:z].copy()
    dva = np.ctypeslib.as_array(C.cast(dv, C.POINTER(C.c_double)), shape=(max(z, 1),))[:z].copy()
    for p in (rs, ci, dv):
        L.pwm_free(p)
    return rsa, cia, dva


def gen_powerlaw(n, alpha=1.8, xm=7.3333, cap=50000, struct_seed=1, value_seed=2):
    """Power-law CRS: lengths floor(xm*U^(-1/alpha)) capped, columns uniform
    without repeats in a row, values N(0,1).  Returns (row_start, col_ind, data)."""
    rs, ci, dv, nnz = VP(), VP(), VP(), I64()
    _check(lib().pwm_gen_powerlaw(int(n), float(alpha), float(xm), int(cap), int(struct_seed), int(value_seed),
                                  C.byref(rs), C.byref(ci), C.byref(dv), C.byref(nnz)))
    return _take_crs(int(n), rs, ci, dv, nnz)


def gen_stencil9(side, extra=2, seed=3):
    """9-point stencil (8/-1) on side x side plus `extra` random off-stencil
    columns per row (values U(-1,1))."""
    rs, ci, dv, nnz = VP(), VP(), VP(), I64()
    _check(lib().pwm_gen_stencil9(int(side), int(extra), int(seed), C.byref(rs), C.byref(ci), C.byref(dv),
                                  C.byref(nnz)))
    return _take_crs(int(side) * int(side), rs, ci, dv, nnz)


def gen_spd_powerlaw(n, seed=4):
    """Symmetric positive definite power-law CRS (diagonal = row |sum| + 1)."""
    rs, ci, dv, nnz = VP(), VP(), VP(), I64()
    _check(lib().pwm_gen_spd_powerlaw(int(n), int(seed), C.byref(rs), C.byref(ci), C.byref(dv), C.byref(nnz)))
    return _take_crs(int(n), rs, ci, dv, nnz)


class Batch:
    """Many power methods stepped together (pwm_batch_* in pwm_hip.h).

    Options: items_per_thread (4, 8, 16), chunk (steps per graph replay),
    graph (False: plain launches), persistent, bf16_when_exact (FP32 batches:
    BF16 storage for the instances whose values are exact in BF16), autotune
    (the first solve after an upload times the tile schedules and keeps the
    fastest; schedule() reports the choice)."""

    def __init__(self, values="f32", **options):
        self.h = VP()
        _check(lib().pwm_batch_create(VALUES[values], C.byref(self.h)))
        self.values = values
        self._sizes = []  # rows of each instance, in id order
        for k, v in options.items():
            self.set_option(k, v)

    def set_option(self, key, value):
        self._chk(lib().pwm_batch_set_option(self.h, key.encode(), int(value)))

    def stats(self):
        a, b, c = I64(), I64(), I64()
        self._chk(lib().pwm_batch_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"launches": a.value, "graphs": b.value, "list_builds": c.value}

    def schedule(self):
        """Tile schedule of the next solve; candidate_ms maps (ipt, persistent)
        to the ms per step the autotuner timed (empty when it did not run)."""
        ipt, per, tuned = I32(), I32(), I32()
        ms = (C.c_float * 4)()
        self._chk(lib().pwm_batch_schedule(self.h, C.byref(ipt), C.byref(per), C.byref(tuned), ms))
        cands = {}
        for k, key in enumerate([(8, False), (8, True), (16, False), (16, True)]):
            if tuned.value and ms[k] >= 0:
                cands[key] = float(ms[k])
        return {"items_per_thread": ipt.value, "persistent": bool(per.value), "tuned": bool(tuned.value),
                "candidate_ms": cands}

    def __del__(self):
        try:
            if self.h:
                lib().pwm_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise PwmError(rc, lib().pwm_batch_last_error(self.h).decode())

    def add_crs(self, row_start, col_ind, data):
        rs = np.ascontiguousarray(row_start, np.int32)
        ci = np.ascontiguousarray(col_ind, np.int32)
        dv = np.ascontiguousarray(data, np.float64)
        i = I32()
        self._chk(lib().pwm_batch_add_crs(self.h, len(rs) - 1, _ptr(rs), _ptr(ci), _ptr(dv), C.byref(i)))
        self._sizes.append(len(rs) - 1)
        return i.value

    def add_poisson(self, m, n):
        i = I32()
        self._chk(lib().pwm_batch_add_poisson(self.h, int(m), int(n), C.byref(i)))
        self._sizes.append(int(m) * int(n))
        return i.value

    def _rows(self, i, n):
        """Rows of instance i; a caller's n that disagrees is an error (the
        library writes the instance's own number of entries)."""
        if not 0 <= int(i) < len(self._sizes):
            raise PwmError(-1, f"batch instance {i} does not exist")
        rows = self._sizes[int(i)]
        if n is not None and int(n) != rows:
            raise PwmError(-1, f"batch instance {i} has {rows} rows, not {n}")
        return rows

    def size(self):
        a, b, c = I32(), I64(), I64()
        self._chk(lib().pwm_batch_size(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def solve(self, max_it, tol=-1.0, seed=0):
        n = self.size()[0]
        lam = np.zeros(n)
        its = np.zeros(n, np.int64)
        ms = F64()
        self._chk(lib().pwm_batch_solve(self.h, int(max_it), float(tol), int(seed), _ptr(lam), _ptr(its),
                                        C.byref(ms)))
        return lam, its, ms.value

    def start_vector(self, i, n=None, seed=0):
        out = np.zeros(self._rows(i, n))
        self._chk(lib().pwm_batch_start_vector(self.h, int(i), int(seed), _ptr(out)))
        return out

    def vector(self, i, n=None):
        out = np.zeros(self._rows(i, n))
        yn = F64()
        self._chk(lib().pwm_batch_vector(self.h, int(i), _ptr(out), C.byref(yn)))
        return out, yn.value


class Server:
    """Continuous batching of power-method jobs and SpMV chains under a device
    memory budget (pwm_server_* in pwm_hip.h).

    Borrowed matrices (borrow=True) keep references to the caller's arrays;
    the Python object holds them alive until remove_matrix."""

    def __init__(self, values="f32", budget_bytes=64 << 30, max_jobs=1024, **options):
        self.h = VP()
        _check(lib().pwm_server_create(VALUES[values], int(budget_bytes), int(max_jobs), C.byref(self.h)))
        self.values = values
        self._sizes = {}
        self._held = {}
        self._job_rows = {}  # job id -> rows of its matrix
        for k, v in options.items():
            self.set_option(k, v)

    def __del__(self):
        try:
            if self.h:
                lib().pwm_server_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise PwmError(rc, lib().pwm_server_last_error(self.h).decode())

    def set_option(self, key, value):
        self._chk(lib().pwm_server_set_option(self.h, key.encode(), int(value)))

    def add_matrix(self, row_start, col_ind, data, borrow=False):
        rs = np.ascontiguousarray(row_start, np.int32)
        ci = np.ascontiguousarray(col_ind, np.int32)
        dv = np.ascontiguousarray(data, np.float64)
        m = I32()
        self._chk(lib().pwm_server_add_matrix(self.h, len(rs) - 1, _ptr(rs), _ptr(ci), _ptr(dv), 1 if borrow else 0,
                                              C.byref(m)))
        if borrow:
            self._held[m.value] = (rs, ci, dv)
        self._sizes[m.value] = len(rs) - 1
        return m.value

    def add_poisson(self, m, n):
        k = I32()
        self._chk(lib().pwm_server_add_poisson(self.h, int(m), int(n), C.byref(k)))
        self._sizes[k.value] = int(m) * int(n)
        return k.value

    def remove_matrix(self, mat):
        self._chk(lib().pwm_server_remove_matrix(self.h, int(mat)))
        self._held.pop(mat, None)

    def resident(self, mat):
        r = I32()
        self._chk(lib().pwm_server_matrix_resident(self.h, int(mat), C.byref(r)))
        return bool(r.value)

    def submit(self, mat, max_it, tol=-1.0, seed=0):
        j = I64()
        self._chk(lib().pwm_server_submit(self.h, int(mat), int(max_it), float(tol), int(seed), C.byref(j)))
        self._job_rows[j.value] = self._sizes[int(mat)]
        return j.value

    def step(self):
        """One admission + chunk of steps; returns (running, finished_now)."""
        a, f = I32(), I32()
        self._chk(lib().pwm_server_step(self.h, C.byref(a), C.byref(f)))
        return a.value, f.value

    def finished(self):
        """Ids of the jobs the last step retired."""
        cnt = I32()
        buf = np.zeros(256, np.int64)
        self._chk(lib().pwm_server_finished(self.h, _ptr(buf), len(buf), C.byref(cnt)))
        if cnt.value > len(buf):
            buf = np.zeros(cnt.value, np.int64)
            self._chk(lib().pwm_server_finished(self.h, _ptr(buf), len(buf), C.byref(cnt)))
        return buf[:cnt.value].tolist()

    def result(self, job, n=None):
        """(lambda, ynorm, iterations, vector or None); raises while unfinished.
        The vector is returned when n is given; n must be the rows of the job's
        matrix (the library writes that many entries)."""
        lam, yn, its = F64(), F64(), I64()
        rows = self._job_rows.get(int(job))
        if n is not None and rows is not None and int(n) != rows:
            raise PwmError(-1, f"job {job} has {rows} rows, not {n}")
        vec = None if n is None else np.zeros(int(n))
        self._chk(lib().pwm_server_result(self.h, int(job), C.byref(lam), C.byref(yn), C.byref(its),
                                          None if vec is None else _ptr(vec)))
        self._job_rows.pop(int(job), None)
        return lam.value, yn.value, its.value, vec

    def timing(self):
        """Seconds in step's phases so far: admit (incl. uploads), run, retire."""
        a, r, t, u = F64(), F64(), F64(), I64()
        self._chk(lib().pwm_server_timing(self.h, C.byref(a), C.byref(r), C.byref(t), C.byref(u)))
        return {"admit_s": a.value, "run_s": r.value, "retire_s": t.value, "uploads": u.value}

    def stats(self):
        u, r, q, a, f = I64(), I32(), I32(), I32(), I64()
        self._chk(lib().pwm_server_stats(self.h, C.byref(u), C.byref(r), C.byref(q), C.byref(a), C.byref(f)))
        return {"used_bytes": u.value, "resident": r.value, "queued": q.value, "running": a.value,
                "finished": f.value}


def server_start_vector(n, seed):
    """The start vector of a server job with `seed` (host restatement of the
    device hash with stream id 0; FP32 servers round it to float)."""
    i = np.arange(int(n), dtype=np.uint64)
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) * np.uint64(0x9e3779b97f4a7c15) + i + np.uint64(0x632be59bd9b4e019)) & M
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)) & M
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)) & M
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0


class Triplet:
    """COO matrix with the reference Triplet's members (Matrix/Triplet.hpp:26-45):
    row_coord, col_coord, data, row_size, col_size, nnz -- what
    SparseMatrixHip.loadFromTriplets reads."""

    def __init__(self, row_size, col_size, row_coord, col_coord, data):
        self.row_size, self.col_size = int(row_size), int(col_size)
        self.row_coord = np.ascontiguousarray(row_coord, np.int32)
        self.col_coord = np.ascontiguousarray(col_coord, np.int32)
        self.data = np.ascontiguousarray(data, np.float64)

    @property
    def nnz(self):
        return len(self.row_coord)


def _take_coo(fn, *args):
    nr, nc, nz = I64(), I64(), I64()
    r, c, v = VP(), VP(), VP()
    _check(fn(*args, C.byref(nr), C.byref(nc), C.byref(nz), C.byref(r), C.byref(c), C.byref(v)))
    try:
        n = nz.value
        rows = np.ctypeslib.as_array(C.cast(r, C.POINTER(C.c_int32)), (max(n, 1),))[:n].copy()
        cols = np.ctypeslib.as_array(C.cast(c, C.POINTER(C.c_int32)), (max(n, 1),))[:n].copy()
        vals = np.ctypeslib.as_array(C.cast(v, C.POINTER(C.c_double)), (max(n, 1),))[:n].copy()
    finally:
        for p in (r, c, v):
            lib().pwm_free(p)
    return Triplet(nr.value, nc.value, rows, cols, vals)


def read_mm(path, has_data, symmetric, random_fill=False):
    """Triplet::loadFromMM (Matrix/Triplet.hpp:75-154) into a Triplet."""
    return _take_coo(lib().pwm_read_mm, os.fsencode(path), int(has_data), int(symmetric), int(random_fill))


def read_bin(path, mat_size, symmetric, random_vals=False):
    """Triplet::loadFromBin (Matrix/Triplet.hpp:164-216) into a Triplet."""
    return _take_coo(lib().pwm_read_bin, os.fsencode(path), int(mat_size), int(symmetric), int(random_vals))


def gen_uniform(n, seed=0):
    """x ~ U[-1,1] from mt19937_64(seed): (u >> 11) * 2^-53 * 2 - 1."""
    out = np.empty(int(n), np.float64)
    _check(lib().pwm_gen_uniform(int(n), int(seed), _ptr(out)))
    return out
