// One merge-path tile of a CRS SpMV, shared by the per-matrix merge kernel
// (k_merge_spmv) and the batched power method (k_batch_tiles).
//
// A tile owns the merge items of rows [a0,a1) (list A, the row ends) and
// nonzeros [b0,b1) (list B).  The workgroup
//   1. stages the row ends and the products val[k]*x[col_ind[k]] in LDS
//      (coalesced val/col_ind loads, x gathered);
//   2. lets every thread merge ITEMS consecutive items (its own merge-path
//      search inside the tile), writing completed rows to LDS;
//   3. adds the partial sums of rows that span threads with a segmented
//      inclusive scan over the threads' (row, partial) pairs;
//   4. writes the rows completed in the tile to y with coalesced stores.
// The trailing partial sum (row a1, continued by the next tile) is returned
// to every thread; the caller stores it as the tile's carry.
//
// POWER: outputs are scaled by `scale` (1/|x|, the previous iterate's norm)
// and each thread accumulates its share of |y|^2, x.y and x.x (x scaled as
// well) in FP64 over the rows the tile owns.  Row a0 is not owned when it
// began in an earlier tile: its value here is partial and the carry fix-up
// completes it and counts it.
#pragma once

#include "k_common.hpp"

namespace pwmk {

template <typename Acc, int BS, int ITEMS>
struct MergeTileSmem {
    int32_t rowend[BS * ITEMS];
    Acc prod[BS * ITEMS];
    Acc out[BS * ITEMS];
    int32_t key[BS];
    Acc val[BS];
};

template <typename VT, int BS, int ITEMS, bool POWER>
__device__ __forceinline__ typename Arith<VT>::A merge_tile(
    const int32_t* __restrict__ rs, const int32_t* __restrict__ ci, const VT* __restrict__ val,
    const typename Arith<VT>::X* __restrict__ x, typename Arith<VT>::X* __restrict__ y, int64_t a0, int64_t a1,
    int64_t b0, int64_t b1, double scale, MergeTileSmem<typename Arith<VT>::A, BS, ITEMS>& sm, double& yy,
    double& xy, double& xx) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    const int tid = threadIdx.x;
    const int na = (int)(a1 - a0), nb = (int)(b1 - b0);

    for (int i = tid; i < na; i += BS) sm.rowend[i] = rs[a0 + 1 + i];
    {
        int i = tid;
        for (; i + 3 * BS < nb; i += 4 * BS) {
            const int64_t k = b0 + i;
            const int32_t j0 = ci[k], j1 = ci[k + BS], j2 = ci[k + 2 * BS], j3 = ci[k + 3 * BS];
            const Acc v0 = widen(val[k]), v1 = widen(val[k + BS]), v2 = widen(val[k + 2 * BS]),
                      v3 = widen(val[k + 3 * BS]);
            sm.prod[i] = v0 * (Acc)x[j0];
            sm.prod[i + BS] = v1 * (Acc)x[j1];
            sm.prod[i + 2 * BS] = v2 * (Acc)x[j2];
            sm.prod[i + 3 * BS] = v3 * (Acc)x[j3];
        }
        for (; i < nb; i += BS) sm.prod[i] = widen(val[b0 + i]) * (Acc)x[ci[b0 + i]];
    }
    __syncthreads();

    // this thread's sub-range of the tile's merge items
    const int items = na + nb;
    const int d0 = tid * ITEMS < items ? tid * ITEMS : items;
    const int d1 = d0 + ITEMS < items ? d0 + ITEMS : items;
    int lo = d0 > nb ? d0 - nb : 0;
    int hi = d0 < na ? d0 : na;
    while (lo < hi) {
        const int piv = (lo + hi) >> 1;
        if ((int64_t)sm.rowend[piv] <= b0 + d0 - piv - 1)
            lo = piv + 1;
        else
            hi = piv;
    }
    int ai = lo, bi = d0 - lo;
    const int a_start = ai;
    Acc run = 0;
    int first = -1;
    Acc first_val = 0;
    for (int d = d0; d < d1; ++d) {
        if (ai < na && (int64_t)sm.rowend[ai] <= b0 + bi) {
            if (first < 0) {
                first = ai;
                first_val = run;
            } else {
                sm.out[ai] = run;
            }
            run = 0;
            ++ai;
        } else {
            run += sm.prod[bi];
            ++bi;
        }
    }

    // segmented inclusive scan of (row, partial) over threads; keys ascend
    sm.key[tid] = ai;
    sm.val[tid] = run;
    __syncthreads();
#pragma unroll
    for (int off = 1; off < BS; off <<= 1) {
        Acc add = 0;
        const bool take = tid >= off && sm.key[tid - off] == sm.key[tid];
        if (take) add = sm.val[tid - off];
        __syncthreads();
        if (take) sm.val[tid] = add + sm.val[tid];
        __syncthreads();
    }
    if (first >= 0) {
        const Acc carry_in = (tid > 0 && sm.key[tid - 1] == a_start) ? sm.val[tid - 1] : (Acc)0;
        sm.out[first] = carry_in + first_val;
    }
    __syncthreads();

    // row a0 is unfinished here when it began in an earlier tile
    const bool head_open = na > 0 && (int64_t)rs[a0] < b0;
    for (int i = tid; i < na; i += BS) {
        const Acc v = sm.out[i];
        if constexpr (POWER) {
            const X out = (X)((double)v * scale);
            y[a0 + i] = out;
            if (!(i == 0 && head_open)) {
                const double yd = (double)out, xd = (double)x[a0 + i] * scale;
                yy += yd * yd;
                xy += xd * yd;
                xx += xd * xd;
            }
        } else {
            y[a0 + i] = (X)v;
        }
    }
    return sm.val[BS - 1];
}

}  // namespace pwmk
