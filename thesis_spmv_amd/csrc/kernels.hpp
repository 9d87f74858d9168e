// Host-side launchers of the gfx950 kernels (defined in the k_*.hip files).
// Only plain pointers cross this header; the C ABI layer (api.cpp) owns all
// device buffers and streams.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "k_common.hpp"

namespace pwmk {

enum ValueType { VT_F64 = 0, VT_F32 = 1, VT_BF16 = 2 };

// Device CRS arrays (row_start[N+1], col_ind[Z], values[Z] of `vtype`).
struct CrsDev {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    const int32_t* rs = nullptr;
    const int32_t* ci = nullptr;
    const void* val = nullptr;
    int vtype = VT_F64;
    int nt = 0;  // tiles kernel: non-temporal col_ind / value loads
};

// A slice of one long row handled by one workgroup: entries [k0, k1) of row
// `row`; `slot` is the long-row index the partial sums are gathered under.
struct LongChunk {
    int32_t row, slot;
    int32_t k0, k1;
};

struct LongRows {
    int32_t n_rows = 0;             // long rows
    int32_t n_chunks = 0;           // chunks over all long rows
    const LongChunk* chunks = nullptr;
    const int32_t* row_ids = nullptr;      // [n_rows]
    const int32_t* chunk_begin = nullptr;  // [n_rows+1] first chunk of each long row
    void* chunk_partial = nullptr;         // accumulator-typed [n_chunks]
    int32_t threshold = 1 << 30;           // rows longer than this are long
};

// ---- CRS vector kernel (k_crs.hip) ----------------------------------------
// Short rows: G lanes per row; long rows: chunked, finished by crs_finalize.
hipError_t launch_crs_vector(const CrsDev& A, int G, int n_vec_blocks, const LongRows& lr, const void* x,
                             void* y, const PowerArgs* pw, hipStream_t s);
// Finishes long rows (sum of chunk partials, in chunk order) and, for the
// power method, reduces the per-block partials into rec[k] and advances k.
hipError_t launch_crs_finalize(const CrsDev& A, int n_partials, const LongRows& lr, void* y,
                               const PowerArgs* pw, hipStream_t s);

// ---- merge path (k_merge.hip) -----------------------------------------------
struct MergePlan {
    int64_t P = 0;                  // tiles (= logical threads of CRS_Merge::mv)
    int64_t items_per_tile = 0;     // ceil((N+Z)/P)
    int ipt = 8;                    // template items per thread (tile capacity 256*ipt)
    int64_t* split_a = nullptr;     // [P+1]
    int64_t* split_b = nullptr;     // [P+1]
    int32_t* carry_row = nullptr;   // [P]
    void* carry_val = nullptr;      // accumulator-typed [P]
};
hipError_t launch_merge_splits(const CrsDev& A, const MergePlan& mp, hipStream_t s);
hipError_t launch_merge_spmv(const CrsDev& A, const MergePlan& mp, const void* x, void* y, const PowerArgs* pw,
                             hipStream_t s);
hipError_t launch_merge_fixup(const CrsDev& A, const MergePlan& mp, void* y, const PowerArgs* pw, hipStream_t s);
// CRS over row-aligned tiles (k_tiles.hip): the tiles are the iteration
// engine's plan (struct ETile, engine.hpp) with tiles_items_per_thread()
// items per thread; carries go through launch_merge_fixup with mp.P = tiles.
struct ETile;
int tiles_items_per_thread();
hipError_t launch_tiles_spmv(const CrsDev& A, const ETile* tiles, const MergePlan& mp, const void* x, void* y,
                             const PowerArgs* pw, hipStream_t s);

// Power-mode carry fix-up runs in this many workgroups (each sums its share
// of the tile partials into a second-level slot after the P tile slots).
inline int merge_fix_blocks(int64_t P) {
    const int64_t b = (P + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 256 ? 256 : b));
}

// ---- utilities (k_util.hip) -------------------------------------------------
// Util/Poisson.hpp:32-83 on the device, one thread per row.
hipError_t launch_fill_poisson(int64_t m, int64_t n, int32_t* rs, int32_t* ci, void* val, int vtype,
                               hipStream_t s);
// double -> storage type conversion of values.
hipError_t launch_convert_values(const double* src, void* dst, int64_t n, int vtype, hipStream_t s);
// storage type -> double widening of values (exports).
hipError_t launch_copy_values_f64(const void* src, int vtype, double* dst, int64_t n, hipStream_t s);
hipError_t launch_fill(void* p, int64_t n, int is_f64, double v, hipStream_t s);
hipError_t launch_copy_vec(const void* src_f64_or_f32, int src_f64, void* dst, int dst_f64, int64_t n,
                           hipStream_t s);
// Final normalisation of the power method (see api.cpp).
hipError_t launch_power_finish(const PowerArgs& pw, int64_t n, int is_f64, int64_t steps, hipStream_t s);
hipError_t launch_power_reset(const PowerArgs& pw, hipStream_t s);
// Rewrite a buffer (Infinity-Cache flush for the roofline timing).
hipError_t launch_flush(void* p, int64_t bytes, hipStream_t s);
// Plain streaming copy (profiler counter calibration).
hipError_t launch_stream_copy(void* dst, const void* src, int64_t bytes, hipStream_t s);
// gather-rate probe: entries rounded down to whole workgroups; out holds
// (entries / 8192) * 256 doubles
int64_t gather_probe_entries(int64_t n);
hipError_t launch_gather_fill(uint32_t* idx, double* val, int64_t n, uint32_t window, hipStream_t s);
hipError_t launch_gather_probe(const uint32_t* idx, const double* val, const double* x, double* out, int64_t n,
                               int64_t window, int per_xcd, hipStream_t s);

}  // namespace pwmk
