// CRS SpMV over row-aligned tiles (the CRS format's "tiles" kernel).
//
// The plan is the iteration engine's (engine_plan_tiles): the merge path cut
// every 3/4 of 256*8 items, each cut moved back to its row's start when the
// row is short, so tiles of short rows hold whole rows ("stream" tiles: a
// thread per row sums the staged products left to right, no carry) and only
// long rows are split across tiles (merge tiles with a carry).  The carries
// are added by the merge path's fix-up kernels (k_merge.hip), which only need
// the tiles' row/nonzero starts (mp.split_a/split_b) and carry arrays.
// One workgroup per tile; XCD j takes the j-th contiguous eighth of the tiles.
#include "engine.hpp"
#include "kernels.hpp"
#include "merge_tile.hpp"

namespace pwmk {

constexpr int kTBS = 256;
constexpr int kTilesIPT = 8;

template <typename VT, bool POWER, bool NT>
__global__ __launch_bounds__(kTBS) void k_tiles_spmv(CrsDev A, const ETile* __restrict__ tiles, MergePlan mp,
                                                     const typename Arith<VT>::X* __restrict__ x_arg,
                                                     typename Arith<VT>::X* __restrict__ y_arg, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    __shared__ MergeTileSmem<Acc, kTBS, kTilesIPT> sm;
    __shared__ double red[48];
    const X* __restrict__ x = x_arg;
    X* __restrict__ y = y_arg;
    double scale = 1.0;
    if constexpr (POWER) {
        if (pw.ctrl[1]) return;
        const int k = pw.ctrl[0];
        x = (const X*)((k & 1) ? pw.buf1 : pw.buf0);
        y = (X*)((k & 1) ? pw.buf0 : pw.buf1);
        scale = step_scale(pw, k);
    }
    const int64_t p = xcd_remap(blockIdx.x, mp.P);
    const ETile T = tiles[p];
    const int na = T.a1 - T.a0;
    TileRegs<VT, kTilesIPT> R;
    X xr[kTilesIPT];
    tile_load<VT, kTBS, kTilesIPT, NT>(R, A.ci, (const VT*)A.val, T.b0, T.b1);
    if (POWER && T.stream) stream_xrows<VT, kTBS, kTilesIPT>(x, T.a0, na, xr);
    tile_stage<VT, kTBS, kTilesIPT>(R, A.rs, x, T.a0, na, T.b0, T.b1, sm);
    __syncthreads();
    double yy = 0.0, xy = 0.0, xx = 0.0;
    Acc c = 0;
    if (T.stream)
        stream_compute<VT, kTBS, kTilesIPT, POWER>(y, xr, T.a0, T.a1, T.b0, scale, sm, yy, xy, xx);
    else
        c = merge_compute<VT, kTBS, kTilesIPT, POWER>(A.rs, x, y, T.a0, T.a1, T.b0, T.b1, scale, sm, yy, xy, xx);
    if (threadIdx.x == kTBS - 1) {
        mp.carry_row[p] = T.a1;
        ((Acc*)mp.carry_val)[p] = c;
    }
    if constexpr (POWER) {
        block_sum3<kTBS>(yy, xy, xx, red);
        if (threadIdx.x == 0) {
            pw.partials[3 * p + 0] = yy;
            pw.partials[3 * p + 1] = xy;
            pw.partials[3 * p + 2] = xx;
        }
    }
}

template <typename VT>
static hipError_t launch_tiles_t(const CrsDev& A, const ETile* tiles, const MergePlan& mp, const void* x, void* y,
                                 const PowerArgs* pw, hipStream_t s) {
    using X = typename Arith<VT>::X;
    if (pw)
        hipLaunchKernelGGL((A.nt ? k_tiles_spmv<VT, true, true> : k_tiles_spmv<VT, true, false>), dim3((unsigned)mp.P),
                           dim3(kTBS), 0, s, A, tiles, mp, (const X*)nullptr, (X*)nullptr, *pw);
    else
        hipLaunchKernelGGL((A.nt ? k_tiles_spmv<VT, false, true> : k_tiles_spmv<VT, false, false>),
                           dim3((unsigned)mp.P), dim3(kTBS), 0, s, A, tiles, mp, (const X*)x, (X*)y, PowerArgs{});
    return hipGetLastError();
}

int tiles_items_per_thread() { return kTilesIPT; }

hipError_t launch_tiles_spmv(const CrsDev& A, const ETile* tiles, const MergePlan& mp, const void* x, void* y,
                             const PowerArgs* pw, hipStream_t s) {
    if (mp.P <= 0) return hipSuccess;
    switch (A.vtype) {
        case VT_F64: return launch_tiles_t<double>(A, tiles, mp, x, y, pw, s);
        case VT_F32: return launch_tiles_t<float>(A, tiles, mp, x, y, pw, s);
        case VT_BF16: return launch_tiles_t<bf16_t>(A, tiles, mp, x, y, pw, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pwmk
