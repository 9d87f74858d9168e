/*
 * pwm_hip.h -- C ABI of libpwmhip.so, the MI355X (gfx950) SpMV and power-method
 * library.  Every entry point returns an int status (PWM_OK = 0, negative on
 * error) and pwm_last_error() returns the message of the last failure.
 *
 * The drop-in boundary is the reference's abstract operator
 * pwm::SparseMatrix<T,int_type> (Matrix/SparseMatrix.hpp:13-68) and its
 * factory pwm::selectType(method, threads) (Util/DriverUtil.hpp:38-84).  The
 * C++ classes in include/pwm/SparseMatrixHip.hpp derive from a restatement of
 * that interface and forward each virtual to the entry point named beside it
 * below.  Plain pointers and sizes only: no torch or HIP types cross it.
 *
 * Index arrays are 32-bit.  The reference instantiates int_type as `int` and
 * `unsigned int` (Tests/poisson.cpp:28,146); every index it stores is
 * non-negative and below 2^31, so both have the bits this ABI reads.  A
 * matrix holds at most PWM_MAX_NNZ = 2^31 - 65536 nonzeros (the kernels'
 * 32-bit entry offsets keep that much headroom); loading more fails with
 * PWM_ERR_INVALID.  Host vectors are double (the reference's data_t,
 * Util/Constants.hpp:32).
 *
 * Threading: calls on one handle are not re-entrant (like the reference's
 * objects, whose CRS_Merge carry buffers are mutable members,
 * SOA_Implementations/CRS_Merge.hpp:27-30); different handles are
 * independent.  Each handle owns one HIP stream at default priority.
 */
#ifndef PWM_HIP_H
#define PWM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PWM_OK               0
#define PWM_MAX_NNZ         2147418112LL  /* 2^31 - 65536 */

#define PWM_ERR_INVALID     -1  /* bad argument (sizes, indices, aliasing) */
#define PWM_ERR_IO          -2  /* file cannot be read / is malformed */
#define PWM_ERR_HIP         -3  /* HIP runtime error (no device, launch failure) */
#define PWM_ERR_NOMEM       -4  /* device allocation failed */
#define PWM_ERR_NOT_SQUARE  -5  /* power method on a non-square matrix */
#define PWM_ERR_STATE       -6  /* no matrix loaded yet */
#define PWM_ERR_BUDGET      -7  /* serving admission: HBM budget exhausted */
#define PWM_ERR_UNSUPPORTED -8  /* option or format not available */

/* ---- formats: the GPU classes and the reference classes they stand beside.
 * The ids are the reference's method ids (Util/DriverUtil.hpp:45-74); the
 * GPU factory pwm::selectType adds 100 to them (see SparseMatrixHip.hpp). */
#define PWM_FORMAT_CRS        1   /* CRS vector kernel   ~ CRS / CRSOMP (ids 1, 2) */
#define PWM_FORMAT_CRS_MERGE  8   /* merge-path CRS      ~ CRS_Merge (id 8) */
#define PWM_FORMAT_CSB        9   /* compressed sparse blocks ~ CSB (id 9) */
#define PWM_FORMAT_BLOCKCOH   10  /* row-distributed block CO-H ~ BlockCOH (id 10) */

/* ---- device storage precision of the matrix values ---------------------
 * F64: values and vectors FP64, FP64 accumulation (parity mode).
 * F32: values and vectors FP32, FP32 accumulation per row.
 * BF16: values BF16 (accepted only when every value is exact in BF16, e.g.
 *       the Poisson stencil's 4/-1), vectors FP32, FP32 accumulation.
 * Norms, Rayleigh quotients and convergence tests are FP64 in every mode. */
#define PWM_VALUES_F64  0
#define PWM_VALUES_F32  1
#define PWM_VALUES_BF16 2

typedef struct pwm_matrix pwm_matrix;

typedef struct {
    int64_t nor, noc, nnz;     /* SparseMatrix fields (Matrix/SparseMatrix.hpp:17-23) */
    int32_t format, value_type;
    int32_t threads;           /* partitions actually used (tiles, bands) */
    int32_t group_size;        /* CRS vector kernel: lanes per row; CSB: block-column groups */
    int64_t device_bytes;      /* HBM held by the handle */
    int64_t long_rows;         /* rows routed to the long-row path */
    int32_t block_bits;        /* CSB beta = 2^block_bits (0 when not CSB) */
    int32_t balanced;          /* CSB balanced flag (CSB.hpp:549-559) */
    int32_t unroll;            /* CSB slab-stream kernel: groups of 4 entries per thread per step */
    int32_t non_temporal;      /* CSB slab streams / CRS tiles: 1 = non-temporal entry loads */
    int32_t layout;            /* CSB: 1 slab pieces, 2 slab streams, 3 wide slab streams, 4 swept slabs */
    int32_t kernels;           /* kernel launches of one pwm_mv_device */
    int32_t high_occupancy;    /* BlockCOH: 1 when the kernel instance with registers capped for 8 waves per SIMD runs */
} pwm_info_t;

/* Constructor SparseMatrix(int threads) (Matrix/SparseMatrix.hpp:30) through
 * the factory selectType(method, threads) (Util/DriverUtil.hpp:38-84).
 * threads <= 0 lets the library pick the partition count. */
int pwm_create(int format, int threads, int value_type, pwm_matrix** out);
/* Destructor (e.g. Matrix/CRS.hpp:44-52). */
int pwm_destroy(pwm_matrix* h);
/* Message of the last failure on h (h == NULL: last failure of pwm_create
 * or of a call that had no handle, on this thread). */
const char* pwm_last_error(const pwm_matrix* h);

/* SparseMatrix::generatePoissonMatrix(m, n, partitions)
 * (Matrix/SparseMatrix.hpp:41; CRS: Matrix/CRS.hpp:60-74 via
 * Util/Poisson.hpp:32-83).  The matrix is generated on the device. */
int pwm_generate_poisson(pwm_matrix* h, int64_t m, int64_t n, int partitions);

/* SparseMatrix::loadFromTriplets(Triplet*, partitions)
 * (Matrix/SparseMatrix.hpp:48; CRS: Matrix/CRS.hpp:81-91 via
 * Util/TripletToCRS.hpp:135-172).  The triplets are copied, never modified
 * (the reference sorts the caller's Triplet in place). */
int pwm_load_triplets(pwm_matrix* h, int64_t nor, int64_t noc, int64_t nnz, const int32_t* rows,
                      const int32_t* cols, const double* vals, int partitions);

/* Addition: load a matrix that is already in CRS form (rows ascending,
 * columns ascending inside each row), e.g. from a generator. */
int pwm_load_crs(pwm_matrix* h, int64_t nor, int64_t noc, const int32_t* row_start,
                 const int32_t* col_ind, const double* vals, int partitions);

/* SparseMatrix::mv(const T* x, T* y) (Matrix/SparseMatrix.hpp:57; oracle
 * Matrix/CRS.hpp:99-109).  Host pointers; x and y must not overlap. */
int pwm_mv(pwm_matrix* h, const double* x, double* y);

/* SparseMatrix::powerIteration(T* x, T* y, it) (Matrix/SparseMatrix.hpp:66;
 * Matrix/CRS.hpp:118-130).  Host pointers.  On return x and y hold what the
 * reference leaves in them: the last normalised iterate is in x when `it` is
 * even and in y when it is odd; the other holds the one before it. */
int pwm_power_iteration(pwm_matrix* h, double* x, double* y, int64_t it);

int pwm_info(const pwm_matrix* h, pwm_info_t* out);

/* Matrix statistics (Util_Scripts/calc_stats.cpp:44-105), computed when the
 * matrix is loaded (block counts on the device):
 *   rows_avg = nnz/N, rows_var = sum_i (len_i - rows_avg)^2 / N;
 *   block_bits = min(16, 3 + ceil(log2 sqrt(N))); nb_blocks whole blocks per
 *   side (ceil(N / 2^bits), one fewer when 2^bits does not divide N);
 *   block_avg, block_var over those nb_blocks^2 blocks; balanced = no whole
 *   block row holds more than twice the mean block-row count.
 * Entries whose column block lies beyond the square's blocks are not
 * counted.  row_max (the longest row) is an addition.  The block statistics
 * are left 0 (nb_blocks = -1) above 8192 blocks per side.  The autotuners
 * read row_max.  Needs a loaded matrix. */
typedef struct {
    int64_t nor, nnz;
    double rows_avg, rows_var;
    int64_t row_max;
    int32_t block_bits, balanced;
    int64_t nb_blocks;
    double block_avg, block_var;
} pwm_stats_t;
int pwm_matrix_stats(const pwm_matrix* h, pwm_stats_t* out);

/* Options (addition).  Keys: "group_size" (CRS vector kernel lanes per row,
 * 0 = automatic), "long_row" (CRS: rows longer than this use the long-row
 * path), "items_per_thread" (merge path), "block_bits" (CSB beta override),
 * "csb_groups" (CSB block-column groups pinned to XCDs, 0 = time 1/8/16 at
 * load and keep the fastest), "csb_layout" (CSB device layout: 1 = slab
 * pieces over the stored block order, 2 = one contiguous entry stream per row
 * slab with (row in slab, column) packed in 32 bits, 3 = the same streams
 * with the column (u32) and the row in slab (u16) stored apart, 4 = swept
 * slabs: one nnz-balanced row slab per CU, its y-slice in LDS, the entries
 * stored window by window (1 MB of x per window) and every workgroup
 * walking the windows in order; 0 = automatic: 2 when slabs of at least 256
 * rows pack, else 3, replaced by 4 when x exceeds 2 MB, the slabs fit and 4
 * runs faster in the load-time timing), "csb_window_bits" (swept slabs: log2 of the
 * columns per window, 0 = 1 MB of x), "csb_sync_every" (swept slabs: windows
 * between workgroup barriers, 1, 2, 4 or 8, default 2), "csb_sweep_unroll"
 * (swept slabs: groups of 4 entries per lane per chunk, 1, 2 or 4, default 2),
 * "csb_sweep_order" (swept slabs: 1, the default, stores each window's
 * entries in column order so that the lanes of a gather instruction read
 * consecutive columns and share x lines; 0 keeps CRS order),
 * "csb_sweep_waves" (swept slabs: slabs per CU, 1..8, default 1; above 1 the
 * workgroups run in that many waves, so matrices with more rows than one
 * LDS-sized slab per CU can be swept),
 * "csb_unroll" (slab-stream kernel groups of 4 entries
 * per thread per step, 2 or 4; 0 = time both with each group count),
 * "csb_nt" (slab-stream kernel: 1 = non-temporal entry loads and partial-slab
 * stores, 0 = plain, -1 = time both at load, the default),
 * "csb_build" (CSB slab-piece layout built on the device, 0, or on the host,
 * 1; both give the same arrays; the slab streams are always built on the
 * device), "coh_build" (BlockCOH layout built on the device, 0, or on the
 * host, 1; both give the same arrays), "coh_occ8" (BlockCOH kernel instance
 * with registers capped for 8 waves per SIMD: 1, the compiler's budget: 0,
 * or -1 = time both at load, the default), "crs_kernel" (CRS: 0 = time the vector kernel at
 * three lane counts and the row-aligned tiles kernel at load and keep the
 * fastest, 1 = vector, 2 = tiles), "crs_nt" (CRS tiles kernel: 1 =
 * non-temporal col_ind/value loads, 0 = plain, -1 = time both at load, the
 * default).  Must be set before the matrix is loaded unless noted. */
int pwm_set_option(pwm_matrix* h, const char* key, int64_t value);

/* ---- device-resident additions ----------------------------------------
 * Vectors live in HBM: double for PWM_VALUES_F64, float otherwise.  A device
 * vector must be coarse-grained device memory (hipMalloc, or
 * pwm_device_alloc) holding the whole vector from the pointer on; host
 * memory (pinned or not), managed memory and fine-grained device memory are
 * refused with PWM_ERR_INVALID before anything is launched, since the kernels
 * add into y with hardware FP64 atomics that such memory does not support. */
int pwm_mv_device(pwm_matrix* h, const void* x_dev, void* y_dev);

/* Power method on device vectors with an FP64 eigenvalue estimate and a
 * convergence test (the reference has none, Matrix/CRS.hpp:118-130):
 *   lambda_k = Rayleigh quotient x_k^T A x_k / x_k^T x_k, also |A x_k|;
 * stop when |lambda_k - lambda_{k-1}| <= tol*|lambda_k| (tol < 0: never) or
 * after max_it iterations.  x_dev holds the start vector; on return x_dev and
 * y_dev are left as pwm_power_iteration leaves x and y.  Outputs may be NULL. */
int pwm_power_method_device(pwm_matrix* h, void* x_dev, void* y_dev, int64_t max_it, double tol,
                            double* lambda, double* ynorm, int64_t* iterations);

/* No counterpart in the reference.  pwm_power_method_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same kernels and results, but
 * the power method's workspace (step control, one record per step, FP64
 * partial sums) is a fixed 8 MiB device array of the library, one per GPU,
 * held until the process ends, instead of buffers the handle allocates on
 * first use.  No hipMalloc or hipFree is called; the handle itself must be
 * loaded (loading allocates).  Limits: max_it <= PWM_NOALLOC_MAX_IT and the
 * handle's partial-sum slots (its tiles / workgroups) <=
 * PWM_NOALLOC_MAX_SLOTS, else PWM_ERR_INVALID.  Every call on a GPU uses that
 * GPU's one array: calls that overlap on one GPU (from other threads, on
 * other handles or streams) race and must not be made. */
#define PWM_NOALLOC_MAX_IT    65536
#define PWM_NOALLOC_MAX_SLOTS 262133
int pwm_power_method_device_noalloc(pwm_matrix* h, void* x_dev, void* y_dev, int64_t max_it, double tol,
                                   double* lambda, double* ynorm, int64_t* iterations);

/* CRS_Merge split points actually used on the device (P+1 entries each;
 * SOA_Implementations/CRS_Merge.hpp:96-112).  Call with a == NULL to get P. */
int pwm_merge_splits(pwm_matrix* h, int64_t* P, int64_t* a, int64_t* b);

/* TripletToMultipleCRS (Util/TripletToCRS.hpp:309-385) of a loaded CRS or
 * CRS_MERGE handle: the matrix as `partitions` CRS matrices of consecutive
 * row ranges, the static row plan of the reference's partitioned classes
 * and the row shards a multi-GPU power method would place one per GPU.
 * Partition p holds rows [first_rows[p], first_rows[p] + thread_rows[p]):
 * nor / partitions rows each (integer division), the remainder in the last
 * (:320-330).  pwm_crs_partitions writes the plan (arrays of `partitions`
 * entries, any may be NULL) and each partition's nonzero count;
 * pwm_crs_partition_export writes partition `part`'s local CRS arrays
 * (row_start[thread_rows+1] starting at 0, col_ind/values[part_nnz], values
 * widened to double).  A partition without rows or entries is empty (the
 * reference writes into its zero-length arrays). */
int pwm_crs_partitions(pwm_matrix* h, int partitions, int32_t* first_rows, int32_t* thread_rows, int64_t* part_nnz);
int pwm_crs_partition_export(pwm_matrix* h, int partitions, int part, int32_t* row_start, int32_t* col_ind,
                             double* values);

/* CRS arrays as held on the device (row_start[nor+1], col_ind[nnz], values
 * widened to double), for bit-exact checks of the device-side builders
 * (generatePoissonMatrix: Util/Poisson.hpp:32-83; loadFromTriplets:
 * Util/TripletToCRS.hpp:135-172). */
int pwm_crs_export(pwm_matrix* h, int32_t* row_start, int32_t* col_ind, double* values);

/* Device layout export for bit-exact parity checks.
 * CSB (SOA_Implementations/CSB.hpp:464-567): ind[nnz], data[nnz] (double),
 * blk_ptr[H*V+1] in the stored block order (a slab-stream handle rebuilds it
 * from its streams).  NULL arrays: sizes only. */
int pwm_csb_layout(pwm_matrix* h, int32_t* block_bits, int64_t* nblocks, uint32_t* ind, double* data,
                   int32_t* blk_ptr);
/* BlockCOH band p (SOA_Implementations/BlockCOH.hpp:319-455).  counts[0..5] =
 * nnz, blocks, row_jumps, block_row_jumps, band_start, band_rows; meta[0..2] =
 * beta_bits, horizontal_blocks, vertical_blocks.  NULL arrays: sizes only. */
int pwm_blockcoh_band(pwm_matrix* h, int band, int64_t* counts, int32_t* meta, uint16_t* row_jump,
                      uint16_t* col_jump, double* data, int16_t* row_jump_block, int16_t* col_jump_block,
                      int32_t* block_nnz);

/* ---- device memory helpers (plumbing for callers without a HIP runtime) */
int pwm_device_alloc(size_t bytes, void** ptr);
int pwm_device_free(void* ptr);
int pwm_memcpy_h2d(void* dst, const void* src, size_t bytes);
int pwm_memcpy_d2h(void* dst, const void* src, size_t bytes);
int pwm_device_synchronize(void);
/* Fill n entries of a device vector (double when is_f64, else float) with a
 * constant. */
int pwm_device_fill(void* ptr, int64_t n, int is_f64, double value);
/* Copy `bytes` (a multiple of 16) between device buffers with the library's
 * own streaming kernel: reads and writes exactly `bytes` each, the reference
 * point for converting profiler byte counters. */
int pwm_device_copy(void* dst, const void* src, int64_t bytes);

/* Timing (addition, used by bench.py): run `reps` SpMVs on the handle's
 * stream between two HIP events; with flush_bytes > 0 a buffer of that many
 * bytes is rewritten before each SpMV (outside the events' spans) so the
 * matrix is not served from the 256 MB Infinity Cache.  Writes the mean
 * milliseconds per SpMV of the whole launch sequence and, if kernel_ms is
 * not NULL, of the dominant kernel alone.  With kernel_ms NULL and no flush
 * the reps SpMVs run back to back between one pair of events (no markers
 * between steps). */
int pwm_time_mv(pwm_matrix* h, const void* x_dev, void* y_dev, int reps, int64_t flush_bytes,
                double* ms_per_mv, double* kernel_ms);

/* Timing (addition, used by bench.py): `reps` rounds of one SpMV on each of
 * `count` loaded handles in turn (hs[k] with xs[k], ys[k]), all launched on
 * hs[0]'s stream, between HIP events.  When the handles' arrays together
 * exceed the 256 MB Infinity Cache, each SpMV finds its matrix in HBM: the
 * back-to-back rate of a matrix that alone would fit that cache.  Writes the
 * mean milliseconds per SpMV and, if kernel_ms is not NULL, of the dominant
 * kernel alone (events around every SpMV; NULL: one pair of events around
 * all of them).  Calls on these handles must not overlap it. */
int pwm_time_mv_rotating(pwm_matrix* const* hs, const void* const* xs, void* const* ys, int count, int reps,
                         double* ms_per_mv, double* kernel_ms);

/* Gather-rate probe (addition, used by bench.py): mean milliseconds of one
 * pass over n entries (rounded down to a multiple of 8192) of a stream of
 * (u32 index, f64 value) pairs, 12 bytes each like a CRS nonzero, where every
 * index gathers one double of x_dev (x_len doubles): from [0, window), or with
 * per_xcd from the window [j*window, (j+1)*window) of the XCD j that runs the
 * workgroup.  Indices are hashed, uniform in the window.  The bound a sparse
 * SpMV's gathers meet: n nonzeros cannot take less than n / this rate.
 * window = 0: the stream alone (x_dev may be NULL), its streaming ceiling.
 * per_xcd bits 8..15 choose how the gathers load x: 0 = plain global loads;
 * 1 + c = bounds-checked buffer loads with cache-policy bits c (1 sc0, 2 nt,
 * 16 sc1; c in 0..3 or 16..19), to measure how each policy fetches a line
 * that misses the L2. */
int pwm_gather_rate(const void* x_dev, int64_t x_len, int64_t window, int per_xcd, int64_t n, double* ms);

/* ---- generators for the benchmark workloads (host side, deterministic) -
 * Each writes CRS arrays into malloc'd buffers released with pwm_free.
 * Power-law: N rows, lengths floor(xm * U^(-1/alpha)) capped at cap, columns
 * uniform without repeats inside a row, values N(0,1).
 * Stencil9: 9-point stencil (8 / -1) on a side x side grid plus `extra`
 * uniform off-stencil columns per row with values U(-1,1). */
int pwm_gen_powerlaw(int64_t n, double alpha, double xm, int64_t cap, uint64_t struct_seed,
                     uint64_t value_seed, int32_t** row_start, int32_t** col_ind, double** vals,
                     int64_t* nnz);
int pwm_gen_stencil9(int64_t side, int64_t extra, uint64_t seed, int32_t** row_start, int32_t** col_ind,
                     double** vals, int64_t* nnz);
/* Symmetric positive definite power-law matrix (batch workload): the
 * power-law row law, half of each row drawn off the diagonal and mirrored,
 * values +-U(0,1] (duplicates summed), diagonal = row absolute sum + 1. */
int pwm_gen_spd_powerlaw(int64_t n, uint64_t seed, int32_t** row_start, int32_t** col_ind, double** vals,
                         int64_t* nnz);
/* x ~ U[-1,1] from a 64-bit Mersenne twister with the given seed:
 * (u >> 11) * 2^-53 * 2 - 1. */
int pwm_gen_uniform(int64_t n, uint64_t seed, double* out);
void pwm_free(void* p);

/* Triplet::loadFromMM (Matrix/Triplet.hpp:75-154): '%' lines skipped, the
 * first other line gives rows, cols, entries, entries are 1-based
 * "row col [value]", an explicit 0 is dropped when has_data, pattern entries
 * are 1 or (random_fill) uniform_real(-100,100) from mt19937(747846),
 * symmetric mirrors off-diagonal entries, reading stops at the first empty
 * line.  COO arrays are allocated by the library (free with pwm_free).
 * PWM_ERR_IO for an unreadable or malformed file (more entries than
 * declared, an index outside the declared size). */
int pwm_read_mm(const char* path, int has_data, int symmetric, int random_fill, int64_t* nor, int64_t* noc,
                int64_t* nnz, int32_t** rows, int32_t** cols, double** vals);
/* Triplet::loadFromBin (Matrix/Triplet.hpp:164-216): little-endian uint32
 * (row, col) pairs, 0-based, entries = file size / 8, mat_size x mat_size,
 * values 1 or random as above, symmetric mirrors off-diagonal pairs. */
int pwm_read_bin(const char* path, int64_t mat_size, int symmetric, int random_vals, int64_t* nor, int64_t* noc,
                 int64_t* nnz, int32_t** rows, int32_t** cols, double** vals);

/* ---- batched power method (addition) ----------------------------------
 * Many independent power methods (the reference's powerIteration,
 * Matrix/CRS.hpp:118-130, plus the Rayleigh-quotient stopping rule of
 * pwm_power_method_device) solved together: every step is one launch over
 * the merge-path tiles of all unfinished instances plus one launch that
 * reduces each instance's norms (FP64) and applies its stop test; chunks of
 * steps are replayed as one hipGraph.  Each instance's result is
 * bit-identical to solving it alone.  value_type F32 or BF16 (FP32 vectors
 * and row sums) or F64.  Instances are square.  Not re-entrant per batch. */
typedef struct pwm_batch pwm_batch;
int pwm_batch_create(int value_type, pwm_batch** out);
int pwm_batch_destroy(pwm_batch* b);
const char* pwm_batch_last_error(const pwm_batch* b);
int pwm_batch_add_crs(pwm_batch* b, int64_t n, const int32_t* row_start, const int32_t* col_ind, const double* vals,
                      int32_t* id);
/* 2D Poisson instance generated on the device (Util/Poisson.hpp:32-83). */
int pwm_batch_add_poisson(pwm_batch* b, int64_t m, int64_t n, int32_t* id);
/* Start vectors x0[i] = U[-1,1) hashed from (seed, id, i); stop at relative
 * eigenvalue change <= tol (tol < 0: never) or max_it iterations.  Outputs
 * (per instance, in id order) may be NULL; *ms = device time of the launch. */
int pwm_batch_solve(pwm_batch* b, int64_t max_it, double tol, uint64_t seed, double* lambdas, int64_t* iterations,
                    double* ms);
int pwm_batch_start_vector(pwm_batch* b, int32_t id, uint64_t seed, double* out);
int pwm_batch_vector(pwm_batch* b, int32_t id, double* out, double* ynorm);
int pwm_batch_size(const pwm_batch* b, int32_t* instances, int64_t* rows, int64_t* nnz);
/* "items_per_thread" (4, 8, 16, default 16: tile = 256 x items merge items),
 * "chunk" (steps per graph replay, 1..4096), "graph" (0: plain launches),
 * "persistent" (0: one workgroup per tile instead of resident workgroups
 * that prefetch their next tile), "bf16_when_exact" (FP32 batches: 1 stores
 * each instance whose values are all exact in BF16 as BF16, the others as
 * FP32; vectors stay FP32), "autotune" (1: the first solve after the
 * instances are uploaded first times 6 steps of the whole batch for
 * items_per_thread 8 and 16, each with plain and persistent tile grids
 * (plain only with bf16_when_exact), and keeps the fastest; it overrides
 * "items_per_thread" and "persistent").  Results are bit-identical for a
 * given schedule; schedules differ in the order of FP additions only. */
int pwm_batch_set_option(pwm_batch* b, const char* key, int64_t value);
/* Kernel launches, graph instantiations and work-list builds so far. */
int pwm_batch_stats(const pwm_batch* b, int64_t* launches, int64_t* graphs, int64_t* lists);
/* The tile schedule of the next solve; *tuned = 1 when the autotuner chose
 * it, and then candidate_ms[4] (may be NULL) holds the ms per step it timed
 * for (8, plain), (8, persistent), (16, plain), (16, persistent), -1 for a
 * candidate not timed. */
int pwm_batch_schedule(const pwm_batch* b, int32_t* items_per_thread, int32_t* persistent, int32_t* tuned,
                       float* candidate_ms);

/* ---- request server (addition) ----------------------------------------
 * Continuous batching of power-method jobs (stop at relative eigenvalue
 * change <= tol, or after max_it steps) and SpMV chains (tol < 0: exactly
 * max_it normalised SpMVs) on registered matrices.  All device memory of a
 * server (resident matrices, job vectors) comes from one region of
 * `budget_bytes`; a matrix is made resident when its first job is admitted
 * and stays resident, while unreferenced it may be evicted (least recently
 * used first) to admit another job.  pwm_server_step admits queued jobs in
 * arrival order, runs `chunk` steps of every running job and retires the
 * finished ones; while chunks run it stages further queued jobs (uploads,
 * tile plans, start vectors) on a second stream and starts another chunk with
 * them whenever one finishes, so one call may run several chunks.  A job's
 * result does not depend on what else is in flight.
 * Start vectors x0[i] = U[-1,1) hashed from (seed, i).  Single-threaded: calls
 * on one server must not overlap. */
typedef struct pwm_server pwm_server;
int pwm_server_create(int value_type, int64_t budget_bytes, int32_t max_jobs, pwm_server** out);
int pwm_server_destroy(pwm_server* s);
const char* pwm_server_last_error(const pwm_server* s);
/* "chunk" (steps per pwm_server_step, default 8), "graph" (0: plain
 * launches), "persistent" (as for pwm_batch_set_option), "items_per_thread"
 * (4, 8, 16, default 16; only with nothing resident), "bf16_when_exact"
 * (FP32 servers, before any matrix is added: 1 stores each matrix whose
 * values are all exact in BF16, and every Poisson matrix, as BF16; vectors
 * stay FP32; a job's result is bit-identical to the same job on a pure BF16
 * or pure FP32 server), "overlap" (1, the default: stage queued jobs while
 * chunks run; 0: stage and admit, then run one chunk). */
int pwm_server_set_option(pwm_server* s, const char* key, int64_t value);
/* PWM_SERVER_BORROW: keep the caller's arrays instead of a copy.  They must
 * stay alive and unchanged until the matrix is removed; they are checked
 * again (row_start order, column range) before every upload, and a matrix
 * that no longer passes makes pwm_server_step fail with PWM_ERR_INVALID
 * (remove it to go on). */
#define PWM_SERVER_BORROW 1
int pwm_server_add_matrix(pwm_server* s, int64_t n, const int32_t* row_start, const int32_t* col_ind,
                          const double* vals, int32_t flags, int32_t* mat);
/* 2D Poisson matrix, generated on the device when it becomes resident. */
int pwm_server_add_poisson(pwm_server* s, int64_t m, int64_t n, int32_t* mat);
int pwm_server_remove_matrix(pwm_server* s, int32_t mat);
int pwm_server_matrix_resident(const pwm_server* s, int32_t mat, int32_t* resident);
int pwm_server_submit(pwm_server* s, int32_t mat, int64_t max_it, double tol, uint64_t seed, int64_t* job);
int pwm_server_step(pwm_server* s, int32_t* running, int32_t* finished);
/* Ids of the jobs the last pwm_server_step retired: up to `cap` into ids,
 * their number into *count. */
int pwm_server_finished(const pwm_server* s, int64_t* ids, int32_t cap, int32_t* count);
/* Seconds spent so far in pwm_server_step's phases: admission (including
 * matrix uploads and tile plans, also those made while chunks run), waiting
 * for the engine's chunks, retiring jobs; and the number of matrix uploads. */
int pwm_server_timing(const pwm_server* s, double* admit_s, double* run_s, double* retire_s, int64_t* uploads);
/* Finished job: eigenvalue estimate, |y| of the last step, steps taken and
 * (vec non-NULL, n doubles) the normalised final iterate.  PWM_ERR_STATE
 * while the job is queued or running; afterwards the job is forgotten. */
int pwm_server_result(pwm_server* s, int64_t job, double* lambda, double* ynorm, int64_t* iters, double* vec);
int pwm_server_stats(const pwm_server* s, int64_t* used_bytes, int32_t* resident, int32_t* queued, int32_t* running,
                     int64_t* finished);

#ifdef __cplusplus
}
#endif
#endif /* PWM_HIP_H */
