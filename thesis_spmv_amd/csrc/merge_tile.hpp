// Tiles of a CRS SpMV, shared by the per-matrix merge kernel (k_merge_spmv)
// and the iteration engine (k_eng_tiles).
//
// A tile owns the merge items of rows [a0,a1) (list A, the row ends) and
// nonzeros [b0,b1) (list B), at most BS*ITEMS items.  Processing a tile:
//   load     each thread issues 128-bit loads of its col_ind/value groups
//            (groups of 4 aligned to the array start) into registers;
//   stage    row ends to LDS, x gathered, products val[k]*x[col_ind[k]] to
//            LDS (then a barrier);
//   compute  merge:  every thread merges ITEMS consecutive items (its own
//                    merge-path search inside the tile); a row that starts
//                    and ends inside one thread's items is written at once,
//                    rows that span threads are completed by a segmented
//                    inclusive scan of (row, partial): within each wave by
//                    shuffles, then across waves from their last lanes;
//                    the trailing partial (row a1, continued by the next
//                    tile) is returned in thread BS-1 as the tile's carry;
//            stream: row-aligned tiles of short rows: thread i sums rows
//                    i, i+BS, ... left to right from LDS; no carry.
// Splitting load from stage lets a persistent caller issue the next tile's
// loads before computing the current one.
//
// POWER: outputs are scaled by `scale` (1/|x|, the previous iterate's norm)
// and each thread accumulates its share of |y|^2, x.y and x.x (x scaled as
// well) in FP64 over the rows the tile owns.  Row a0 is not owned when it
// began in an earlier tile: its value here is partial and the carry fix-up
// completes it and counts it.
#pragma once

#include "k_common.hpp"

namespace pwmk {

// prod holds the tile's products at index k - (b0 & ~3) (entry k), so that
// each aligned group of 4 entries is one 16-byte-aligned LDS store; the
// products of [b0, b1) start at offset b0 & 3.
template <typename Acc, int BS, int ITEMS>
struct MergeTileSmem {
    int32_t rowend[BS * ITEMS];
    alignas(16) Acc prod[BS * ITEMS + 8];
    int32_t wkey[BS / 64];
    Acc wval[BS / 64];
};


template <typename T>
__device__ __forceinline__ T shfl_up64(T v, int off) {
    return __shfl_up(v, (unsigned)off, 64);
}

// Four consecutive values loaded with 128-bit (FP64: two 128-bit, BF16: one
// 64-bit) loads; p must be aligned to four values.
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef double vd2 __attribute__((ext_vector_type(2)));
typedef int vi4 __attribute__((ext_vector_type(4)));
typedef unsigned vu2 __attribute__((ext_vector_type(2)));
template <typename Acc>
__device__ __forceinline__ void store4(Acc* p, const Acc (&v)[4]);
template <>
__device__ __forceinline__ void store4<float>(float* p, const float (&v)[4]) {
    *(vf4*)p = vf4{v[0], v[1], v[2], v[3]};
}
template <>
__device__ __forceinline__ void store4<double>(double* p, const double (&v)[4]) {
    ((vd2*)p)[0] = vd2{v[0], v[1]};
    ((vd2*)p)[1] = vd2{v[2], v[3]};
}

template <typename VT> struct Quad;
template <> struct Quad<float> {
    vf4 q;
    __device__ __forceinline__ void load(const float* p) { q = *G((const vf4*)p); }
    __device__ __forceinline__ void load_nt(const float* p) { q = __builtin_nontemporal_load(G((const vf4*)p)); }
    __device__ __forceinline__ void set(const float* t) { q = vf4{t[0], t[1], t[2], t[3]}; }
    __device__ __forceinline__ float get(int e) const { return q[e]; }
};
template <> struct Quad<double> {
    vd2 a, b;
    __device__ __forceinline__ void load(const double* p) {
        a = *G((const vd2*)p);
        b = *G((const vd2*)(p + 2));
    }
    __device__ __forceinline__ void load_nt(const double* p) {
        a = __builtin_nontemporal_load(G((const vd2*)p));
        b = __builtin_nontemporal_load(G((const vd2*)(p + 2)));
    }
    __device__ __forceinline__ void set(const double* t) {
        a = vd2{t[0], t[1]};
        b = vd2{t[2], t[3]};
    }
    __device__ __forceinline__ double get(int e) const { return e < 2 ? a[e] : b[e - 2]; }
};
template <> struct Quad<bf16_t> {
    vu2 q;
    __device__ __forceinline__ void load(const bf16_t* p) { q = *G((const vu2*)p); }
    __device__ __forceinline__ void load_nt(const bf16_t* p) { q = __builtin_nontemporal_load(G((const vu2*)p)); }
    __device__ __forceinline__ void set(const bf16_t* t) {
        q.x = (uint32_t)t[0].u | ((uint32_t)t[1].u << 16);
        q.y = (uint32_t)t[2].u | ((uint32_t)t[3].u << 16);
    }
    __device__ __forceinline__ float get(int e) const {
        const uint32_t w = e < 2 ? q.x : q.y;
        return __uint_as_float((e & 1) ? (w & 0xffff0000u) : (w << 16));
    }
};

// col_ind and value arrays read by tiles extend kTilePad zero entries past
// the last nonzero: a group of 4 aligned to the array start never leaves the
// allocation.
constexpr int kTilePad = 4;

// A thread's col_ind/value groups of one tile (ITEMS/4 + 1 groups: the
// tile's first group may start before b0).
template <typename VT, int ITEMS>
struct TileRegs {
    static_assert(ITEMS % 4 == 0, "items per thread must be a multiple of 4");
    static constexpr int VU = ITEMS / 4 + 1;
    vi4 jq[VU];
    Quad<VT> vq[VU];
};

// Every lane loads a group (lanes past the tile's last group reload that
// group), so the loads need no branch and nothing waits on them until
// tile_stage.  Requires kTilePad zero entries after col_ind/values.
// NT: non-temporal loads (entries read once leave the L2 to the x gathers).
template <typename VT, int BS, int ITEMS, bool NT = false>
__device__ __forceinline__ void tile_load(TileRegs<VT, ITEMS>& R, const int32_t* __restrict__ ci,
                                          const VT* __restrict__ val, int64_t b0, int64_t b1) {
    constexpr int VU = TileRegs<VT, ITEMS>::VU;
    const int tid = threadIdx.x;
    const int64_t g0 = b0 & ~(int64_t)3;
    const int ng = (int)((b1 - g0 + 3) >> 2);
    const int qmax = ng > 0 ? ng - 1 : 0;
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const int q = tid + u * BS < qmax ? tid + u * BS : qmax;
        const int64_t g = g0 + 4 * (int64_t)q;
        if constexpr (NT) {
            R.jq[u] = __builtin_nontemporal_load(G((const vi4*)(ci + g)));
            R.vq[u].load_nt(val + g);
        } else {
            R.jq[u] = *G((const vi4*)(ci + g));
            R.vq[u].load(val + g);
        }
    }
}

// x[a0+r] for this thread's rows r = tid + u*BS (u < ceil(na/BS)); issued
// before tile_stage so that no later global load waits behind them.
template <typename VT, int BS, int ITEMS>
__device__ __forceinline__ void stream_xrows(const typename Arith<VT>::X* __restrict__ x, int64_t a0, int na,
                                             typename Arith<VT>::X* xr) {
    // the first ceil(na/BS) slices only (a workgroup-uniform bound), each
    // index clamped to the last row so that the loads need no per-lane
    // branch; an empty tile (a0 may be n) loads nothing
    if (na == 0) return;
    const int last = na - 1;
    const int nu = (na + BS - 1) / BS;
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
        if (u < nu) {
            const int r = (int)threadIdx.x + u * BS;
            xr[u] = G(x)[a0 + (r < last ? r : last)];
        }
    }
}

// Row ends rs[a0+1+i] (i < na) to sm.rowend[i]; products of [b0,b1) to
// sm.prod[k-(b0&~3)], a whole aligned group of 4 per store.  Every load is
// issued (with clamped, always valid indices)
// before the first one is used, so the thread waits for memory once.  The
// caller places the barrier.
template <typename VT, int BS, int ITEMS>
__device__ __forceinline__ void tile_stage(const TileRegs<VT, ITEMS>& R, const int32_t* __restrict__ rs,
                                           const typename Arith<VT>::X* __restrict__ x, int64_t a0, int na,
                                           int64_t b0, int64_t b1,
                                           MergeTileSmem<typename Arith<VT>::A, BS, ITEMS>& sm) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    constexpr int VU = TileRegs<VT, ITEMS>::VU;
    const int tid = threadIdx.x;
    const int nu = (na + BS - 1) / BS;
    // the first nu slices only (a workgroup-uniform bound), clamped to the
    // last row end so that the loads need no per-lane branch; an empty tile
    // (a0 may be n) loads no row end
    const int last = na > 0 ? na - 1 : 0;
    int32_t re[ITEMS];
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
        if (u < nu) {
            const int i = tid + u * BS;
            re[u] = G(rs)[a0 + 1 + (i < last ? i : last)];
        }
    }
    X xv[VU][4];
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const auto xg = G(x);
        xv[u][0] = xg[R.jq[u].x];
        xv[u][1] = xg[R.jq[u].y];
        xv[u][2] = xg[R.jq[u].z];
        xv[u][3] = xg[R.jq[u].w];
    }
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
        const int i = tid + u * BS;
        if (u < nu && i < na) sm.rowend[i] = re[u];
    }
    const int64_t g0 = b0 & ~(int64_t)3;
    const int ng = (int)((b1 - g0 + 3) >> 2);
#pragma unroll
    for (int u = 0; u < VU; ++u) {
        const int q = tid + u * BS;
        if (q < ng) {
            // the whole group (entries outside [b0, b1) are neighbours' or
            // padding, never read back)
            Acc pr[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) pr[e] = (Acc)R.vq[u].get(e) * (Acc)xv[u][e];
            store4<Acc>(&sm.prod[4 * q], pr);
        }
    }
}

// Stream compute (after the barrier): rows [a0,a1) complete, b0 = rs[a0].
// POWER: outputs scaled and norms accumulated, with xr from stream_xrows.
// Reads only LDS (no global load waits here).
template <typename VT, int BS, int ITEMS, bool POWER>
__device__ __forceinline__ void stream_compute(typename Arith<VT>::X* __restrict__ y, const typename Arith<VT>::X* xr,
                                               int64_t a0, int64_t a1, int64_t b0, double scale,
                                               const MergeTileSmem<typename Arith<VT>::A, BS, ITEMS>& sm, double& yy,
                                               double& xy, double& xx) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    const int tid = threadIdx.x;
    const int na = (int)(a1 - a0);
    const int nu = (na + BS - 1) / BS;
    const int off = (int)(b0 & 3);  // products of entry k at k - (b0 & ~3)
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
        const int r = tid + u * BS;
        if (u < nu && r < na) {
            const int beg = (r == 0 ? 0 : (int)(sm.rowend[r - 1] - b0)) + off;
            const int end = (int)(sm.rowend[r] - b0) + off;
            // left to right, four LDS reads in flight at a time
            Acc s = 0;
            int k = beg;
            for (; k + 4 <= end; k += 4) {
                const Acc p0 = sm.prod[k], p1 = sm.prod[k + 1], p2 = sm.prod[k + 2], p3 = sm.prod[k + 3];
                s += p0;
                s += p1;
                s += p2;
                s += p3;
            }
            for (; k < end; ++k) s += sm.prod[k];
            if constexpr (POWER) {
                const double xd = (double)xr[u] * scale;
                const X out = (X)((double)s * scale);
                G(y)[a0 + r] = out;
                const double yd = (double)out;
                yy += yd * yd;
                xy += xd * yd;
                xx += xd * xd;
            } else {
                G(y)[a0 + r] = (X)s;
            }
        }
    }
}

// Merge compute (after the barrier); returns the carry in thread BS-1.
template <typename VT, int BS, int ITEMS, bool POWER>
__device__ __forceinline__ typename Arith<VT>::A merge_compute(
    const int32_t* __restrict__ rs, const typename Arith<VT>::X* __restrict__ x, typename Arith<VT>::X* __restrict__ y,
    int64_t a0, int64_t a1, int64_t b0, int64_t b1, double scale, MergeTileSmem<typename Arith<VT>::A, BS, ITEMS>& sm,
    double& yy, double& xy, double& xx) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int na = (int)(a1 - a0), nb = (int)(b1 - b0);

    // this thread's sub-range of the tile's merge items
    const bool head_open = na > 0 && (int64_t)G(rs)[a0] < b0;
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
    const int poff = (int)(b0 & 3);  // products of entry k at k - (b0 & ~3)
    const int a_start = ai;
    Acc run = 0;
    int first = -1;
    Acc first_val = 0;
    auto emit = [&](int row, Acc v) {
        if constexpr (POWER) {
            const X out = (X)((double)v * scale);
            G(y)[a0 + row] = out;
            if (!(row == 0 && head_open)) {
                const double yd = (double)out, xd = (double)G(x)[a0 + row] * scale;
                yy += yd * yd;
                xy += xd * yd;
                xx += xd * xd;
            }
        } else {
            G(y)[a0 + row] = (X)v;
        }
    };
    for (int d = d0; d < d1; ++d) {
        if (ai < na && (int64_t)sm.rowend[ai] <= b0 + bi) {
            if (first < 0) {
                first = ai;
                first_val = run;
            } else {
                emit(ai, run);
            }
            run = 0;
            ++ai;
        } else {
            run += sm.prod[bi + poff];
            ++bi;
        }
    }

    // segmented inclusive scan of (key = row, run); keys ascend with tid
    const int key = ai;
    Acc inc = run;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const Acc vu = shfl_up64(inc, off);
        const int ku = shfl_up64(key, off);
        if (lane >= off && ku == key) inc += vu;
    }
    if (lane == 63) {
        sm.wkey[wave] = key;
        sm.wval[wave] = inc;
    }
    const Acc prev_inc = shfl_up64(inc, 1);
    const int prev_key = shfl_up64(key, 1);
    const int key0 = __shfl(key, 0, 64);
    __syncthreads();
    // total of row k over the waves before this one (a run of equal keys
    // crosses a wave boundary only through whole waves of that key)
    auto cross = [&](int k) {
        Acc c = 0;
        for (int w = wave - 1; w >= 0 && sm.wkey[w] == k; --w) c += sm.wval[w];
        return c;
    };
    // first row of this thread: its items plus everything before them
    if (first >= 0) {
        Acc cin = 0;
        if (lane > 0) {
            if (prev_key == a_start) cin = prev_inc + (key0 == a_start ? cross(a_start) : (Acc)0);
        } else if (wave > 0 && sm.wkey[wave - 1] == a_start) {
            cin = cross(a_start);
        }
        emit(first, cin + first_val);
    }
    return inc + (key0 == key ? cross(key) : (Acc)0);
}

// A whole merge tile (load, stage, barrier, compute).
template <typename VT, int BS, int ITEMS, bool POWER>
__device__ __forceinline__ typename Arith<VT>::A merge_tile(
    const int32_t* __restrict__ rs, const int32_t* __restrict__ ci, const VT* __restrict__ val,
    const typename Arith<VT>::X* __restrict__ x, typename Arith<VT>::X* __restrict__ y, int64_t a0, int64_t a1,
    int64_t b0, int64_t b1, double scale, MergeTileSmem<typename Arith<VT>::A, BS, ITEMS>& sm, double& yy,
    double& xy, double& xx) {
    TileRegs<VT, ITEMS> R;
    tile_load<VT, BS, ITEMS>(R, ci, val, b0, b1);
    tile_stage<VT, BS, ITEMS>(R, rs, x, a0, (int)(a1 - a0), b0, b1, sm);
    __syncthreads();
    return merge_compute<VT, BS, ITEMS, POWER>(rs, x, y, a0, a1, b0, b1, scale, sm, yy, xy, xx);
}

}  // namespace pwmk
