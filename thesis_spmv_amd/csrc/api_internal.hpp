// Internal state of a pwm_matrix handle and helpers shared by the host-side
// translation units of libpwmhip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pwm_hip.h"
#include "kernels.hpp"

namespace blocked {
struct CohTask;
}

namespace pwm_impl {

// Host-side parallel loop over [0, n) in chunks of `chunk`, on at most 16
// threads (std::thread; the library does not depend on an OpenMP runtime).
// fn(begin, end, worker) with worker < 16, each worker id used by one thread
// at a time.  A persistent pool of 15 threads plus the caller runs it; a call
// made while the pool is busy (another thread, or from inside fn) starts its
// own threads instead.
void parallel_run(int64_t n, int64_t chunk, const std::function<void(int64_t, int64_t, int)>& fn);
template <typename F>
void parallel_for(int64_t n, int64_t chunk, F fn) {
    if (n <= 0) return;
    if ((n + chunk - 1) / chunk <= 1) {
        fn((int64_t)0, n, 0);
        return;
    }
    parallel_run(n, chunk, std::function<void(int64_t, int64_t, int)>(fn));
}

// Thread-local message for calls without a handle.
void set_global_error(const std::string& msg);

// Owned device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    hipError_t alloc(size_t n);
    void release();
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
    template <typename T> T* as() const { return (T*)p; }
};

// col_ind / value buffers that merge-path tiles read end in kPadEntries zero
// entries: tiles load aligned groups of 4 entries.
constexpr int64_t kPadEntries = 4;
// n entries of `elem` bytes plus the zeroed pad.
hipError_t alloc_padded(DevBuf& b, int64_t n, size_t elem);

// CSB device layout (SOA_Implementations/CSB.hpp:34-59).
struct CsbDev {
    int32_t bits = 0;
    int64_t H = 0, V = 0;
    int32_t balanced = 1;
    DevBuf ind, data, blk_ptr;   // u32[Z], value[Z], i32[H*V+1]
    DevBuf work;                 // per-workgroup tasks
    int64_t n_tasks = 0;
    int32_t slab_bits = 0;
    int32_t slabs = 1, pieces_per_block = 1;
    int64_t slab_rows = 0;
    int32_t groups = 1;          // block-column groups (one per XCD)
    DevBuf ypart;                // groups x nor partial sums
    // Slab-stream layout (stream = 1): ind/data hold every slab's entries
    // contiguously, block columns ascending, as (row in slab << col_bits) |
    // column; soff[t*H + bc] is the first entry of slab t in block column bc.
    int32_t stream = 0;
    int32_t col_bits = 0;
    int32_t unroll = 2;          // stream kernel groups of 4 per thread per step
    int32_t nt = 0;              // stream kernel: non-temporal entry loads / partial stores
    int32_t wide = 0;            // wide streams: ind = column, srows = row in slab (u16)
    DevBuf soff, srows;
    // Swept slabs (sweep = 1): T nnz-balanced row slabs [sstart[t],
    // sstart[t+1]), each one persistent workgroup; a slab's entries are
    // stored window by window (W windows of 2^sweep_wb columns, soff[t*W + p]
    // the first entry of window p), packed (row in slab << sweep_wb) |
    // column in window.
    int32_t sweep = 0;
    int32_t sweep_wb = 0;
    int64_t sweep_W = 0;
    int32_t sweep_rows = 0;      // LDS rows of a slab (the largest slab)
    DevBuf sweep_start;
    // sweep_order 1: inside a run the entries are in column order, the run
    // is padded to whole blocks of 256 (index 0xffffffff, value 0) and each
    // block stores sorted entry j*64 + L at 4L + j; sweep_len holds the runs'
    // entry counts (host)
    int32_t sweep_order = 0;
    std::vector<int32_t> sweep_len;
    void clear() {
        bits = 0;
        H = V = 0;
        balanced = 1;
        groups = 1;
        ind.release();
        data.release();
        blk_ptr.release();
        work.release();
        ypart.release();
        soff.release();
        srows.release();
        sweep_start.release();
        sweep = 0;
        sweep_wb = 0;
        sweep_W = 0;
        sweep_rows = 0;
        sweep_order = 0;
        sweep_len.clear();
        wide = 0;
        stream = 0;
        col_bits = 0;
        unroll = 2;
        nt = 0;
        n_tasks = 0;
        slab_bits = 0;
    }
};

// BlockCOH device layout (SOA_Implementations/BlockCOH.hpp:27-67), all bands
// concatenated; offsets say where each band's part of each array starts.
struct CohBandMeta {
    int64_t start, rows, nnz, blocks, row_jumps, block_row_jumps;
    int32_t beta_bits, H, V;
    int64_t off_nz, off_rj, off_blk, off_rjb;
};
struct CohDev {
    std::vector<CohBandMeta> bands;
    DevBuf row_jump, col_jump, data, row_jump_block, col_jump_block, block_nnz;
    DevBuf block_coords;   // decoded (block_row, block_col, entry offset, row_jump offset) per block
    DevBuf band_table;     // per band: start, rows, beta_bits, first block, blocks
    DevBuf tasks;          // per band: chunk -> first block table
    DevBuf task_list;      // CohTask per workgroup
    DevBuf atomic_bands;   // bands accumulated with global atomics
    int64_t n_tasks = 0, total_blocks = 0, n_atomic = 0;
    int32_t lds_rows = 0;  // rows of the largest band accumulated in LDS (the kernel's LDS window)
    int32_t occ8 = 0;      // kernel instance with 8 waves per SIMD (timed at load)
    void clear() {
        bands.clear();
        for (DevBuf* b : {&row_jump, &col_jump, &data, &row_jump_block, &col_jump_block, &block_nnz, &block_coords,
                          &band_table, &tasks, &task_list, &atomic_bands})
            b->release();
        n_tasks = total_blocks = n_atomic = 0;
        lds_rows = 0;
        occ8 = 0;
    }
};

}  // namespace pwm_impl

struct pwm_matrix {
    int format = 0, vtype = 0, threads = 0;
    int64_t nor = 0, noc = 0, nnz = 0;
    bool loaded = false;
    hipStream_t stream = nullptr;
    std::string err;

    // CRS arrays (used by CRS and CRS_MERGE; CSB/BlockCOH keep them for their
    // parity exports only when asked)
    pwm_impl::DevBuf rs, ci, val;
    // host copy of row_start (long-row planning)
    std::vector<int32_t> h_rs;
    pwm_stats_t stats{};  // computed at load (stats.hip)

    // CRS vector kernel plan
    int group_opt = 0;       // option: lanes per row, 0 = autotune
    int vec_blocks_opt = 0;  // option: grid size, 0 = automatic
    int group = 0;           // the plan in use
    int vec_blocks = 0;
    int64_t long_row = 0;   // 0 = auto
    int chunk = 4096;
    pwmk::LongRows lr;
    pwm_impl::DevBuf lr_chunks, lr_rows, lr_begin, lr_partial;
    // CRS kernel: 0 = time the vector and tiles kernels at load and keep the
    // faster, 1 = vector, 2 = tiles (row-aligned tiles, carries through the
    // merge fix-up; the tiles plan lives in mp / split_* / carry_*)
    int crs_kernel = 0;
    int crs_used = 1;
    pwm_impl::DevBuf tiles;
    int crs_nt = -1;          // tiles kernel: non-temporal entry loads (-1 = timed at load)
    int crs_nt_used = 0;
    bool tiles_carry = true;  // some tile is a merge tile (its carry needs the fix-up)

    // merge plan
    int ipt = 8;
    pwmk::MergePlan mp;
    pwm_impl::DevBuf split_a, split_b, carry_row, carry_val;

    // blocked formats
    int32_t csb_bits_override = 0;
    int32_t csb_groups = 0;  // 0 = autotune
    int32_t csb_layout = 0;  // 0 = automatic, 1 = pieces, 2 = streams, 3 = wide streams, 4 = swept slabs
    int32_t csb_window_bits = 0;  // swept slabs: log2 of the columns per window (0 = 1 MB of x)
    int32_t csb_sync_every = 2;   // swept slabs: windows between workgroup barriers
    int32_t csb_sweep_unroll = 2; // swept slabs: groups of 4 entries per lane per chunk (1, 2 or 4)
    int32_t csb_sweep_waves = 1;  // swept slabs: slabs per CU (workgroups run in that many waves)
    int32_t csb_sweep_order = 1;  // swept slabs: 0 = rows inside a window's run, 1 = columns (lanes share x lines)
    int32_t coh_occ8 = -1;        // BlockCOH kernel at 8 waves per SIMD: -1 timed at load, 0 no, 1 yes
    int32_t csb_unroll = 0;  // stream kernel groups of 4 per thread per step (0 = default)
    int32_t csb_nt = -1;     // slab streams: non-temporal entry loads / partial stores (-1 = timed)
    int32_t csb_build = 0;   // CSB slab pieces: 0 = built on the device, 1 = on the host
    int32_t coh_build = 0;   // BlockCOH layout: 0 = built on the device, 1 = on the host
    pwm_impl::CsbDev csb;
    pwm_impl::CohDev coh;

    // power-method workspace
    pwm_impl::DevBuf pw_ctrl, pw_rec, pw_partials;
    int64_t pw_rec_cap = 0;
    // host-API staging vectors
    pwm_impl::DevBuf hx, hy;
    // timing flush buffer
    pwm_impl::DevBuf flush;

};

namespace pwm_impl {

int fail(pwm_matrix* h, int code, const std::string& msg);
// Host -> device copy through pinned staging buffers (double-buffered,
// threads fill one while the other is in flight); returns once the copy has
// completed.  A plain copy when pinned memory cannot be allocated.
// `between` (may be NULL) is called after each chunk's copy is queued, so
// that a caller can do other work between the chunks of a long upload; it
// runs with the staging buffers locked and must not start a staged
// transfer itself.
hipError_t upload_staged(void* dst, const void* src, size_t bytes, hipStream_t s,
                         const std::function<void()>* between = nullptr);
// n FP64 host values uploaded as stored in value type vt (FP64, FP32, or
// BF16 for values exact in BF16), narrowed by host threads into the pinned
// buffers on the way
hipError_t upload_values_staged(void* dst, const double* src, int64_t n, int vt, hipStream_t s,
                                const std::function<void()>* between = nullptr);
// n int32 indices uploaded as they are, each checked on the way against
// [0, bound): *out_of_range tells whether one was not (the device copy is
// then complete but must not be used)
hipError_t upload_indices_staged(void* dst, const int32_t* src, int64_t n, int64_t bound, bool* out_of_range,
                                 hipStream_t s, const std::function<void()>* between = nullptr);
// Device -> host the same way (after the work queued on s).
hipError_t download_staged(void* dst, const void* src, size_t bytes, hipStream_t s);
// CSB slab-stream layout built on the device from the device CRS arrays
// (csb_build.hip); h->csb.bits/H/V set.  csb_key_bits: bits of its sort key.
int csb_key_bits(int64_t cells, int bits);
int csb_stream_device(pwm_matrix* h, int rb, int cb, int64_t S, bool wide);
// CSB slab-piece layout (the stored order, blk_ptr and the pieces of slabs of
// beta >> k rows) built on the device the same way; h->csb.bits/H/V set.
int csb_pieces_device(pwm_matrix* h, int k);
// CSB swept-slab layout built on the device: slabs cut from the host
// row_start `rs` (nnz-balanced towards `target_slabs`, at most `max_rows`
// rows each), windows of 2^wb columns.  Returns PWM_ERR_UNSUPPORTED, with
// nothing built, when the cut needs more than `max_slabs` slabs.
int csb_sweep_device(pwm_matrix* h, const int32_t* rs, int64_t target_slabs, int64_t max_rows, int64_t max_slabs,
                     int wb);
// BlockCOH layout built on the device from the device CRS arrays
// (coh_build.hip); h->coh.bands holds the band partition (start, rows, nnz,
// beta_bits, H, V, off_nz) on entry and the per-band sizes and offsets on
// return.  coh_plan_tasks: the kernel's work list for those bands, decode
// states left zero.
int coh_layout_device(pwm_matrix* h, int64_t lds_rows);
void coh_plan_tasks(const std::vector<CohBandMeta>& B, int64_t lds_rows, std::vector<blocked::CohTask>& tasks,
                    std::vector<int32_t>& atomic_bands, int32_t* max_lds_rows);
// Statistics of the loaded CRS arrays (stats.hip) into h->stats; needs the
// device row_start / col_ind.
int compute_matrix_stats(pwm_matrix* h);
// COO -> CRS on the device (coo.hip): row_start, padded col_ind and padded
// FP64 values of the sorted matrix.  Indices must already be validated.
int coo_to_crs_device(pwm_matrix* h, int64_t nor, int64_t noc, int64_t nnz, const int32_t* rows, const int32_t* cols,
                      const double* vals, DevBuf& rs, DevBuf& ci, DevBuf& dv);
int hip_fail(pwm_matrix* h, hipError_t e, const char* what);
size_t value_size(int vtype);
// True when every value is exact in BF16 (float-exact with a zero low half).
bool exact_in_bf16(const double* v, int64_t n);
size_t vector_size(int vtype);

// per-format planning after the CRS arrays are on the device
int plan_crs(pwm_matrix* h);
int plan_merge(pwm_matrix* h);
int build_csb(pwm_matrix* h, const int32_t* rs, const int32_t* ci, const double* vals);
int build_blockcoh(pwm_matrix* h, const int32_t* rs, const int32_t* ci, const double* vals);
// one SpMV of the loaded format on device vectors; pw != nullptr: power step
int spmv_launch(pwm_matrix* h, const void* x, void* y, const pwmk::PowerArgs* pw, int* n_partials_out);
hipError_t csb_launch(pwm_matrix* h, const void* x, void* y, const pwmk::PowerArgs* pw, int part = 0);
hipError_t coh_launch(pwm_matrix* h, const void* x, void* y, const pwmk::PowerArgs* pw);
int64_t csb_partials(const pwm_matrix* h);
int64_t coh_partials(const pwm_matrix* h);

}  // namespace pwm_impl
