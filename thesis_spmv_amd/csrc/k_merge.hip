// Merge-path CRS SpMV for gfx950 (the GPU counterpart of CRS_Merge,
// SOA_Implementations/CRS_Merge.hpp:95-141).
//
// The merge of list A = row_end_offsets (row_start+1) with list B = 0..Z-1 is
// cut into P tiles with exactly the reference's formulas:
// items = ceil((N+Z)/P), tile p covers diagonals [min(items*p, N+Z),
// min(items*p + items, N+Z)), and each cut is found with searchPathOnDiag
// (k_merge_splits, one thread per cut; the cuts are computed once at load
// time and are bit-identical to the reference's).
//
// One 256-thread workgroup per tile runs merge_tile (merge_tile.hpp): LDS
// staging of the tile's products and row ends, a per-thread merge of ITEMS
// items, a segmented scan for rows that span threads; the tile's trailing
// partial row is left as (carry_row, carry_val).
// k_merge_fixup then adds the carries (the reference's serial fix-up,
// :136-140): each run of equal carry rows is summed in tile order and added
// once, so the result does not depend on scheduling.
#include "kernels.hpp"
#include "merge_tile.hpp"

namespace pwmk {

constexpr int kMBS = 256;

// searchPathOnDiag (CRS_Merge.hpp:42-62) for every cut p = 0..P.
__global__ void k_merge_splits(int64_t nrows, int64_t nnz, const int32_t* __restrict__ rs, int64_t P, int64_t items,
                               int64_t* __restrict__ sa, int64_t* __restrict__ sb) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > P) return;
    const int64_t total = nrows + nnz;
    const int64_t d = items * p < total ? items * p : total;
    const int32_t* row_end = rs + 1;
    int64_t lo = d > nnz ? d - nnz : 0;
    int64_t hi = d < nrows ? d : nrows;
    while (lo < hi) {
        const int64_t piv = (lo + hi) >> 1;
        if ((int64_t)row_end[piv] <= d - piv - 1)
            lo = piv + 1;
        else
            hi = piv;
    }
    sa[p] = lo < nrows ? lo : nrows;
    sb[p] = d - lo;
}

template <typename VT, int ITEMS, bool POWER>
__global__ __launch_bounds__(kMBS) void k_merge_spmv(CrsDev A, MergePlan mp, const typename Arith<VT>::X* __restrict__ x_arg,
                                                     typename Arith<VT>::X* __restrict__ y_arg, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    __shared__ MergeTileSmem<Acc, kMBS, ITEMS> sm;
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
    const int64_t a1 = mp.split_a[p + 1];
    double yy = 0.0, xy = 0.0, xx = 0.0;
    const Acc carry = merge_tile<VT, kMBS, ITEMS, POWER>(A.rs, A.ci, (const VT*)A.val, x, y, mp.split_a[p], a1,
                                                         mp.split_b[p], mp.split_b[p + 1], scale, sm, yy, xy, xx);
    if (threadIdx.x == kMBS - 1) {
        mp.carry_row[p] = (int32_t)a1;
        ((Acc*)mp.carry_val)[p] = carry;
    }
    if constexpr (POWER) {
        block_sum3<kMBS>(yy, xy, xx, red);
        if (threadIdx.x == 0) {
            pw.partials[3 * p + 0] = yy;
            pw.partials[3 * p + 1] = xy;
            pw.partials[3 * p + 2] = xx;
        }
    }
}

// Carry fix-up.  A run of tiles p..q-1 carrying into the same row r is summed
// in tile order and added to y[r] once (by the thread of the run's head).
template <typename VT>
__global__ void k_merge_fixup(int64_t nrows, MergePlan mp, typename Arith<VT>::X* __restrict__ y) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= mp.P) return;
    const int32_t r = mp.carry_row[p];
    if (r >= nrows || (p > 0 && mp.carry_row[p - 1] == r)) return;
    const Acc* cv = (const Acc*)mp.carry_val;
    Acc s = 0;
    for (int64_t q = p; q < mp.P && mp.carry_row[q] == r; ++q) s += cv[q];
    y[r] = (X)((Acc)y[r] + s);
}

constexpr int kFixBS = 256;

// Power mode, pass 1: workgroup b owns tiles [b*per, (b+1)*per): it adds the
// carries of the runs whose head tile it owns (a run of tiles p..q-1 carrying
// into row r is summed in tile order and added to y[r] once, by the thread of
// the run's head, which follows the run past its own range) and sums, in a
// fixed order, its tiles' FP64 partials plus the completed rows'
// contributions, into the second-level slot after the tile partials.
template <typename VT>
__global__ __launch_bounds__(kFixBS) void k_merge_fixup_power1(CrsDev A, MergePlan mp, PowerArgs pw, int64_t per) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    __shared__ double red[48];
    if (pw.ctrl[1]) return;
    const int k = pw.ctrl[0];
    const X* __restrict__ x = (const X*)((k & 1) ? pw.buf1 : pw.buf0);
    X* __restrict__ y = (X*)((k & 1) ? pw.buf0 : pw.buf1);
    const double scale = step_scale(pw, k);
    const Acc* cv = (const Acc*)mp.carry_val;
    const int64_t nrows = A.nrows;
    const int64_t p0 = (int64_t)blockIdx.x * per;
    const int64_t p1 = p0 + per < mp.P ? p0 + per : mp.P;
    double yy = 0.0, xy = 0.0, xx = 0.0;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += kFixBS) {
        yy += pw.partials[3 * p + 0];
        xy += pw.partials[3 * p + 1];
        xx += pw.partials[3 * p + 2];
        const int32_t r = mp.carry_row[p];
        if (r >= nrows || (p > 0 && mp.carry_row[p - 1] == r)) continue;
        Acc s = 0;
        int64_t q = p;
        for (; q < mp.P && mp.carry_row[q] == r; ++q) s += cv[q];
        // Tile q completes row r.  If row r starts at or after that tile's
        // first nonzero, the carries are empty (a tile ended exactly where
        // row r begins) and tile q already counted the row: nothing to do.
        if ((int64_t)A.rs[r] >= mp.split_b[q]) continue;
        const X out = (X)((double)y[r] + (double)s * scale);
        y[r] = out;
        const double yd = (double)out, xd = (double)x[r] * scale;
        yy += yd * yd;
        xy += xd * yd;
        xx += xd * xd;
    }
    block_sum3<kFixBS>(yy, xy, xx, red);
    if (threadIdx.x == 0) {
        double* q2 = pw.partials + 3 * mp.P + 3 * (int64_t)blockIdx.x;
        q2[0] = yy;
        q2[1] = xy;
        q2[2] = xx;
    }
}

// Power mode, pass 2 (one workgroup): the second-level sums in workgroup
// order, then step k's record and k+1.
__global__ __launch_bounds__(kFixBS) void k_merge_fixup_power2(MergePlan mp, PowerArgs pw, int nblk) {
    __shared__ double red[48];
    if (pw.ctrl[1]) return;
    const int k = pw.ctrl[0];
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kFixBS) {
        const double* q2 = pw.partials + 3 * mp.P + 3 * (int64_t)i;
        a += q2[0];
        b += q2[1];
        c += q2[2];
    }
    block_sum3<kFixBS>(a, b, c, red);
    if (threadIdx.x == 0) {
        const double lam = b / c;
        pw.rec[4 * k + 0] = a;
        pw.rec[4 * k + 1] = b;
        pw.rec[4 * k + 2] = c;
        pw.rec[4 * k + 3] = lam;
        int done = (k + 1 >= pw.max_it);
        if (pw.tol >= 0.0 && k >= 1) {
            const double prev = pw.rec[4 * (k - 1) + 3];
            if (fabs(lam - prev) <= pw.tol * fabs(lam)) done = 1;
        }
        pw.ctrl[0] = k + 1;
        pw.ctrl[1] = done;
    }
}

hipError_t launch_merge_splits(const CrsDev& A, const MergePlan& mp, hipStream_t s) {
    const int64_t n = mp.P + 1;
    hipLaunchKernelGGL(k_merge_splits, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A.nrows, A.nnz, A.rs, mp.P,
                       mp.items_per_tile, mp.split_a, mp.split_b);
    return hipGetLastError();
}

template <typename VT, int ITEMS>
static hipError_t launch_tiles(const CrsDev& A, const MergePlan& mp, const void* x, void* y, const PowerArgs* pw,
                               hipStream_t s) {
    using X = typename Arith<VT>::X;
    if (pw)
        hipLaunchKernelGGL((k_merge_spmv<VT, ITEMS, true>), dim3((unsigned)mp.P), dim3(kMBS), 0, s, A, mp,
                           (const X*)nullptr, (X*)nullptr, *pw);
    else
        hipLaunchKernelGGL((k_merge_spmv<VT, ITEMS, false>), dim3((unsigned)mp.P), dim3(kMBS), 0, s, A, mp,
                           (const X*)x, (X*)y, PowerArgs{});
    return hipGetLastError();
}

template <typename VT>
static hipError_t launch_tiles_t(const CrsDev& A, const MergePlan& mp, const void* x, void* y, const PowerArgs* pw,
                                 hipStream_t s) {
    switch (mp.ipt) {
        case 4: return launch_tiles<VT, 4>(A, mp, x, y, pw, s);
        case 8: return launch_tiles<VT, 8>(A, mp, x, y, pw, s);
        case 12: return launch_tiles<VT, 12>(A, mp, x, y, pw, s);
        case 16: return launch_tiles<VT, 16>(A, mp, x, y, pw, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_merge_spmv(const CrsDev& A, const MergePlan& mp, const void* x, void* y, const PowerArgs* pw,
                             hipStream_t s) {
    if (mp.P <= 0) return hipSuccess;
    switch (A.vtype) {
        case VT_F64: return launch_tiles_t<double>(A, mp, x, y, pw, s);
        case VT_F32: return launch_tiles_t<float>(A, mp, x, y, pw, s);
        case VT_BF16: return launch_tiles_t<bf16_t>(A, mp, x, y, pw, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename VT>
static hipError_t launch_fix_t(const CrsDev& A, const MergePlan& mp, void* y, const PowerArgs* pw, hipStream_t s) {
    using X = typename Arith<VT>::X;
    if (pw) {
        const int nblk = merge_fix_blocks(mp.P);
        const int64_t per = (mp.P + nblk - 1) / nblk;
        hipLaunchKernelGGL((k_merge_fixup_power1<VT>), dim3(nblk), dim3(kFixBS), 0, s, A, mp, *pw, per);
        hipLaunchKernelGGL(k_merge_fixup_power2, dim3(1), dim3(kFixBS), 0, s, mp, *pw, nblk);
    } else
        hipLaunchKernelGGL((k_merge_fixup<VT>), dim3((unsigned)((mp.P + 255) / 256)), dim3(256), 0, s, A.nrows, mp,
                           (X*)y);
    return hipGetLastError();
}

hipError_t launch_merge_fixup(const CrsDev& A, const MergePlan& mp, void* y, const PowerArgs* pw, hipStream_t s) {
    if (mp.P <= 0) return hipSuccess;
    switch (A.vtype) {
        case VT_F64: return launch_fix_t<double>(A, mp, y, pw, s);
        case VT_F32: return launch_fix_t<float>(A, mp, y, pw, s);
        case VT_BF16: return launch_fix_t<bf16_t>(A, mp, y, pw, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pwmk
