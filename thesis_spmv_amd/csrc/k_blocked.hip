// CSB and BlockCOH SpMV kernels for gfx950.
//
// Both formats are processed by workgroups that own a slab of rows and keep
// its y-slice in LDS (FP64 or FP32 atomic adds, ds_add_*): every entry of a
// slab, wherever its block sits in the block row or band, adds into LDS and
// the slab is written to y once with coalesced stores.
//
// x locality: every workgroup walks its blocks in the same order (CSB: block
// columns ascending, CSB.hpp:354-363; BlockCOH: the Hilbert order of the
// blocks, which is the same sequence of block columns for every band whose
// rows fit one block row).  The workgroups resident on one XCD therefore
// gather from the same x-slice at about the same time, and that slice
// (beta * 8 bytes, 128-256 KB) stays in the XCD's 4 MB L2 instead of each
// gather missing to the Infinity Fabric as in a row-ordered CRS sweep over
// random columns.
//
// Swept slabs (k_csb_sweep) carry this to matrices whose x outgrows the L2:
// one persistent 1024-thread workgroup per CU owns an nnz-balanced row slab
// (y-slice in all of the CU's LDS) and walks x in windows of 1 MB, the slab's
// entries stored window by window.  The 32 slabs of an XCD start together and
// sweep at about the same rate, so the XCD fetches each window about once per
// SpMV and the gathers hit its L2; no partial sums are written.  Inside a
// window the entries are stored in column order, four to a lane and
// permuted so that the 64 lanes of one gather instruction read 64
// consecutive entries: with about one entry per x line per slab and window,
// a third of the gathers then share a line, and so an L2 request, with a
// neighbouring lane.
#include "blocked.hpp"
#include "kernels.hpp"
#include "merge_tile.hpp"

namespace blocked {

using pwmk::Arith;
using pwmk::bf16_t;
using pwmk::PowerArgs;
using pwmk::widen;

constexpr int kBS = kBlockedBS;
constexpr int kWaves = kBS / 64;
constexpr int kWin = 2 * kBS;  // CSB pieces staged per window

template <typename Acc>
struct LdsRows;
template <> struct LdsRows<double> { static constexpr int v = kLdsRowsF64; };
template <> struct LdsRows<float> { static constexpr int v = kLdsRowsF32; };

// Block-wide exclusive scan of one int per thread; returns the prefix and
// writes the total to *total.  `tmp` holds kWaves ints.
__device__ __forceinline__ int block_excl_scan(int v, int* tmp, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(incl, d);
        if (lane >= d) incl += u;
    }
    if (lane == 63) tmp[wave] = incl;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const int t = tmp[w];
        if (w < wave) before += t;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

// Block-wide segmented scan.  Each thread brings the sum `v` of its entries
// after its last segment start and `f` = 1 if it holds a segment start.  The
// result is the value that flows into the thread's first entry: the sum over
// earlier threads back to the most recent segment start, plus `carry` if no
// segment starts before this thread.  `tv`/`tf` hold kWaves values each.
__device__ __forceinline__ int block_seg_excl(int v, int f, int carry, int* tv, int* tf) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int iv = v, ifl = f;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int uv = __shfl_up(iv, d);
        const int uf = __shfl_up(ifl, d);
        if (lane >= d) {
            if (!ifl) iv += uv;
            ifl |= uf;
        }
    }
    if (lane == 63) {
        tv[wave] = iv;
        tf[wave] = ifl;
    }
    __syncthreads();
    // prefix of the earlier waves, seeded with the chunk carry
    int pv = carry;
    for (int w = 0; w < wave; ++w) {
        if (tf[w]) {
            pv = tv[w];
        } else {
            pv += tv[w];
        }
    }
    __syncthreads();
    // exclusive value of this lane: inclusive of the previous lane in the wave
    int ev = __shfl_up(iv, 1);
    int ef = __shfl_up(ifl, 1);
    if (lane == 0) {
        ev = 0;
        ef = 0;
    }
    return ef ? ev : ev + pv;
}

// ---------------------------------------------------------------- CSB

// workgroup -> (slab task t, column group g).  With a multiple of 8 groups the
// launch runs in rounds of 8 groups so that XCD (b % 8) keeps one group per
// round: b = round * 8 * slabs_total + slab * 8 + xcd.
__device__ __forceinline__ void csb_task(const CsbArgs& a, int& g, int64_t& t) {
    if (a.groups % 8 == 0) {
        const int64_t per_round = (a.n_tasks / a.groups) * 8;
        const int64_t round = blockIdx.x / per_round, rem = blockIdx.x % per_round;
        t = rem / 8;
        g = (int)(round * 8 + rem % 8);
    } else {
        g = (int)(blockIdx.x % a.groups);
        t = blockIdx.x / a.groups;
    }
}

// Writes a slab's LDS sums: to ypart[g] with several column groups, else to
// y (POWER: scaled, with the slab's |y|^2, x.y, x.x in partial slot t).
template <typename X, typename Acc, int BS, bool POWER, bool NT = false>
__device__ __forceinline__ void csb_write_slab(const CsbArgs& a, const Acc* ylds, int h, int64_t row0, int g, int64_t t,
                                               const X* __restrict__ x, X* __restrict__ y, double scale,
                                               const PowerArgs& pw, double* red) {
    const int tid = threadIdx.x;
    if (a.groups > 1) {  // partial slab sums; k_csb_reduce finishes y
        Acc* yp = (Acc*)a.ypart + (int64_t)g * a.nor;
        for (int i = tid; i < h; i += BS) {
            const int64_t row = row0 + i;
            if (row >= a.nor) break;
            if constexpr (NT)
                __builtin_nontemporal_store(ylds[i], &yp[row]);
            else
                yp[row] = ylds[i];
        }
        return;
    }
    double yy = 0.0, xy = 0.0, xx = 0.0;
    for (int i = tid; i < h; i += BS) {
        const int64_t row = row0 + i;
        if (row >= a.nor) break;
        if constexpr (POWER) {
            const X out = (X)((double)ylds[i] * scale);
            y[row] = out;
            const double yd = (double)out, xd = (double)x[row] * scale;
            yy += yd * yd;
            xy += xd * yd;
            xx += xd * xd;
        } else {
            y[row] = (X)ylds[i];
        }
    }
    if constexpr (POWER) {
        pwmk::block_sum3<BS>(yy, xy, xx, red);
        if (tid == 0) {
            pw.partials[3 * t + 0] = yy;
            pw.partials[3 * t + 1] = xy;
            pw.partials[3 * t + 2] = xx;
        }
    }
}

template <typename VT, bool POWER>
__global__ __launch_bounds__(kBS) void k_csb(CsbArgs a, const typename Arith<VT>::X* __restrict__ x_arg,
                                             typename Arith<VT>::X* __restrict__ y_arg, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    constexpr int LR = LdsRows<Acc>::v;
    __shared__ Acc ylds[LR];
    __shared__ int32_t p_start[kWin];
    __shared__ int32_t p_pre[kWin + 1];
    __shared__ int scan_tmp[kWaves];
    __shared__ double red[48];

    const X* __restrict__ x = x_arg;
    X* __restrict__ y = y_arg;
    double scale = 1.0;
    if constexpr (POWER) {
        if (pw.ctrl[1]) return;
        const int k = pw.ctrl[0];
        x = (const X*)((k & 1) ? pw.buf1 : pw.buf0);
        y = (X*)((k & 1) ? pw.buf0 : pw.buf1);
        scale = pwmk::step_scale(pw, k);
    }
    const VT* __restrict__ val = (const VT*)a.val;
    const uint32_t* __restrict__ ind = a.ind;
    const int tid = threadIdx.x;
    int g;
    int64_t t;
    csb_task(a, g, t);
    const int64_t br = t / a.slabs;
    const int s = (int)(t % a.slabs);
    const int64_t beta = (int64_t)1 << a.bits;
    const int h = (int)a.slab_rows;
    const int slab_base = s * h;
    for (int i = tid; i < h; i += kBS) ylds[i] = 0;

    const int64_t bc_lo = g * a.H / a.groups, bc_hi = (g + 1) * a.H / a.groups;
    const int64_t npieces = (bc_hi - bc_lo) * a.np;
    const Piece* __restrict__ P = a.pieces + t * a.H * a.np + bc_lo * a.np;
    for (int64_t w0 = 0; w0 < npieces; w0 += kWin) {
        const int nw = (int)(npieces - w0 < kWin ? npieces - w0 : kWin);
        __syncthreads();
        const int i0 = 2 * tid, i1 = 2 * tid + 1;
        int len0 = 0, len1 = 0;
        if (i0 < nw) {
            const Piece q = P[w0 + i0];
            p_start[i0] = q.start;
            len0 = q.end - q.start;
        }
        if (i1 < nw) {
            const Piece q = P[w0 + i1];
            p_start[i1] = q.start;
            len1 = q.end - q.start;
        }
        int total;
        const int base = block_excl_scan(len0 + len1, scan_tmp, &total);
        p_pre[i0] = base;
        p_pre[i1] = base + len0;
        if (tid == kBS - 1) p_pre[kWin] = total;
        __syncthreads();
        const int T = total;
        int cur = 0;
        for (int v = tid; v < T; v += 4 * kBS) {
            int kk[4], pc[4];
            bool ok[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int vv = v + u * kBS;
                ok[u] = vv < T;
                if (ok[u]) {
                    while (p_pre[cur + 1] <= vv) ++cur;
                    pc[u] = cur;
                    kk[u] = p_start[cur] + (vv - p_pre[cur]);
                } else {
                    pc[u] = 0;
                    kk[u] = 0;
                }
            }
            uint32_t w[4];
            Acc vv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                w[u] = ok[u] ? ind[kk[u]] : 0u;
                vv[u] = ok[u] ? (Acc)widen(val[kk[u]]) : (Acc)0;
            }
            Acc xv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t bc = bc_lo + ((w0 + pc[u]) >> a.np_shift);
                xv[u] = ok[u] ? (Acc)x[bc * beta + (w[u] & 0xffffu)] : (Acc)0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (ok[u]) atomicAdd(&ylds[(int)(w[u] >> 16) - slab_base], vv[u] * xv[u]);
        }
    }
    __syncthreads();
    csb_write_slab<X, Acc, kBS, POWER>(a, ylds, h, br * beta + slab_base, g, t, x, y, scale, pw, red);
}

// Slab-stream CSB: the workgroup's entries are one contiguous run
// [soff[t*H + bc_lo], soff[t*H + bc_hi]) read as aligned groups of 4 (128-bit
// index loads, 128-bit value loads); every entry adds val * x[col] into the
// slab's LDS rows.  Lanes past the run reload its last group and add nothing,
// so the loads carry no branch.
constexpr int kSBS = 256;

// kSU: groups of 4 entries per thread per step (2 or 4); NT: non-temporal
// entry loads and partial-slab stores; WIDE: ind holds the column and srows
// the row in the slab.
template <typename VT, bool POWER, int kSU, bool NT, bool WIDE>
__global__ __launch_bounds__(kSBS) void k_csb_stream(CsbArgs a, const typename Arith<VT>::X* __restrict__ x_arg,
                                                     typename Arith<VT>::X* __restrict__ y_arg, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    extern __shared__ __align__(16) unsigned char csb_smem[];
    Acc* ylds = (Acc*)csb_smem;
    __shared__ double red[48];

    const X* __restrict__ x = x_arg;
    X* __restrict__ y = y_arg;
    double scale = 1.0;
    if constexpr (POWER) {
        if (pw.ctrl[1]) return;
        const int k = pw.ctrl[0];
        x = (const X*)((k & 1) ? pw.buf1 : pw.buf0);
        y = (X*)((k & 1) ? pw.buf0 : pw.buf1);
        scale = pwmk::step_scale(pw, k);
    }
    const VT* __restrict__ val = (const VT*)a.val;
    const uint32_t* __restrict__ ind = a.ind;
    const int tid = threadIdx.x;
    int g;
    int64_t t;
    csb_task(a, g, t);
    const int h = (int)a.slab_rows;
    for (int i = tid; i < h; i += kSBS) ylds[i] = 0;
    const int64_t bc_lo = g * a.H / a.groups, bc_hi = (g + 1) * a.H / a.groups;
    // entry offsets fit 31 bits (int32 CRS row starts)
    const int32_t e0 = a.soff[t * a.H + bc_lo], e1 = a.soff[t * a.H + bc_hi];
    const int cb = a.col_bits;
    const uint32_t cmask = (uint32_t)((1ull << cb) - 1);
    __syncthreads();
    const int32_t q0 = e0 >> 2, q1 = (int32_t)(((int64_t)e1 + 3) >> 2);
    for (int32_t qb = q0; qb < q1; qb += kSBS * kSU) {
        pwmk::vi4 w[kSU];
        pwmk::vu2 r[kSU];
        pwmk::Quad<VT> v[kSU];
#pragma unroll
        for (int u = 0; u < kSU; ++u) {
            const int32_t q = qb + tid + u * kSBS;
            const int64_t qc = q < q1 ? q : q1 - 1;
            if constexpr (NT) {
                w[u] = __builtin_nontemporal_load(pwmk::G((const pwmk::vi4*)(ind + 4 * qc)));
                if constexpr (WIDE) r[u] = __builtin_nontemporal_load(pwmk::G((const pwmk::vu2*)(a.srows + 4 * qc)));
                v[u].load_nt(val + 4 * qc);
            } else {
                w[u] = *pwmk::G((const pwmk::vi4*)(ind + 4 * qc));
                if constexpr (WIDE) r[u] = *pwmk::G((const pwmk::vu2*)(a.srows + 4 * qc));
                v[u].load(val + 4 * qc);
            }
        }
        X xv[kSU][4];
#pragma unroll
        for (int u = 0; u < kSU; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) xv[u][j] = x[WIDE ? (uint32_t)w[u][j] : ((uint32_t)w[u][j] & cmask)];
#pragma unroll
        for (int u = 0; u < kSU; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t k = 4 * (int64_t)(qb + tid + u * kSBS) + j;
                uint32_t row;
                if constexpr (WIDE)
                    row = ((j < 2 ? r[u].x : r[u].y) >> (16 * (j & 1))) & 0xffffu;
                else
                    row = (uint32_t)w[u][j] >> cb;
                if (k >= e0 && k < e1) atomicAdd(&ylds[row], (Acc)v[u].get(j) * (Acc)xv[u][j]);
            }
    }
    __syncthreads();
    const int64_t beta = (int64_t)1 << a.bits;
    const int64_t row0 = (t / a.slabs) * beta + (t % a.slabs) * (int64_t)h;
    csb_write_slab<X, Acc, kSBS, POWER, NT>(a, ylds, h, row0, g, t, x, y, scale, pw, red);
}

// y[r] = sum over the column groups of ypart[g][r], groups in order (and the
// power-method scale and partial sums).
template <typename VT, bool POWER>
__global__ __launch_bounds__(kBS) void k_csb_reduce(CsbArgs a, typename Arith<VT>::X* __restrict__ y_arg,
                                                    PowerArgs pw) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    __shared__ double red[16];
    X* __restrict__ y = y_arg;
    const X* __restrict__ x = nullptr;
    double scale = 1.0;
    if constexpr (POWER) {
        if (pw.ctrl[1]) return;
        const int k = pw.ctrl[0];
        x = (const X*)((k & 1) ? pw.buf1 : pw.buf0);
        y = (X*)((k & 1) ? pw.buf0 : pw.buf1);
        scale = pwmk::step_scale(pw, k);
    }
    const Acc* __restrict__ yp = (const Acc*)a.ypart;
    double yy = 0.0, xy = 0.0, xx = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * kBS + threadIdx.x; r < a.nor; r += (int64_t)gridDim.x * kBS) {
        Acc s = 0;
        for (int g = 0; g < a.groups; ++g) s += yp[(int64_t)g * a.nor + r];
        if constexpr (POWER) {
            const X out = (X)((double)s * scale);
            y[r] = out;
            const double yd = (double)out, xd = (double)x[r] * scale;
            yy += yd * yd;
            xy += xd * yd;
            xx += xd * xd;
        } else {
            y[r] = (X)s;
        }
    }
    if constexpr (POWER) {
        const double ra = pwmk::block_sum<kBS>(yy, red);
        const double rb = pwmk::block_sum<kBS>(xy, red);
        const double rc = pwmk::block_sum<kBS>(xx, red);
        if (threadIdx.x == 0) {
            pw.partials[3 * blockIdx.x + 0] = ra;
            pw.partials[3 * blockIdx.x + 1] = rb;
            pw.partials[3 * blockIdx.x + 2] = rc;
        }
    }
}

int csb_reduce_blocks(int64_t nor) {
    const int64_t b = (nor + kBS - 1) / kBS;
    return (int)(b < 2048 ? (b > 0 ? b : 1) : 2048);
}

// ---------------------------------------------------------------- CSB swept slabs

// Swept slabs: one persistent workgroup of kSwBS threads per CU, each
// owning an nnz-balanced row slab [sstart[t], sstart[t+1]) whose y-slice
// stays in LDS for the whole kernel.  The slab's entries are stored window
// by window: window p holds the columns [p << wb, (p+1) << wb) (1 MB of x by
// default), and every workgroup walks the windows in the same order, adding
// val * x[col] into its LDS rows (ds_add).  In window p, wave w takes the
// chunks w, w + 16, ... of 64 x kSU groups of 4 entries of the slab's run;
// the first chunk of the next window's entries is loaded while the current
// one is gathered, and a workgroup barrier ends every window.  The barrier
// keeps a workgroup's 16 waves on one window, and the nnz-balanced
// workgroups, all started together, stay close enough that each XCD's L2
// holds the windows its workgroups gather from; no partial sums are written
// and read back.
constexpr int kSwBS = 1024;

template <typename VT, int kSU, bool NT>
struct SweepChunk {
    pwmk::vi4 w[kSU];
    pwmk::Quad<VT> v[kSU];
    // groups qb + lane + 64 u; a group past the run's last one (or of an
    // empty run) reloads a valid group and its entries are masked
    __device__ __forceinline__ void load(const uint32_t* __restrict__ ind, const VT* __restrict__ val, int32_t qb,
                                         int32_t q1, int lane) {
#pragma unroll
        for (int u = 0; u < kSU; ++u) {
            const int32_t q = qb + lane + u * 64;
            const int64_t qc = q < q1 ? q : (q1 > 0 ? q1 - 1 : 0);
            if constexpr (NT) {
                w[u] = __builtin_nontemporal_load(pwmk::G((const pwmk::vi4*)(ind + 4 * qc)));
                v[u].load_nt(val + 4 * qc);
            } else {
                w[u] = *pwmk::G((const pwmk::vi4*)(ind + 4 * qc));
                v[u].load(val + 4 * qc);
            }
        }
    }
};

template <typename VT, int kSU, bool NT, typename X, typename Acc>
__device__ __forceinline__ void sweep_add(const SweepChunk<VT, kSU, NT>& c, int32_t qb, int lane, int32_t e0, int32_t e1,
                                          const X* __restrict__ xw, uint32_t wmask, int wb, Acc* ylds) {
    // an entry outside [e0, e1) belongs to a neighbouring run (its column is
    // in another window), and an index 0xffffffff pads a column-ordered run:
    // such an entry gathers the window's first column, which exists, and
    // adds nothing
    bool ok[kSU][4];
    X xv[kSU][4];
#pragma unroll
    for (int u = 0; u < kSU; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t k = 4 * (int64_t)(qb + lane + u * 64) + j;
            ok[u][j] = k >= e0 && k < e1 && (uint32_t)c.w[u][j] != 0xffffffffu;
            xv[u][j] = xw[ok[u][j] ? ((uint32_t)c.w[u][j] & wmask) : 0u];
        }
#pragma unroll
    for (int u = 0; u < kSU; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (ok[u][j]) atomicAdd(&ylds[(uint32_t)c.w[u][j] >> wb], (Acc)c.v[u].get(j) * (Acc)xv[u][j]);
}

template <typename VT, bool POWER, bool NT, int SU = 2>
__global__ __launch_bounds__(kSwBS) void k_csb_sweep(CsbArgs a, const typename Arith<VT>::X* __restrict__ x_arg,
                                                     typename Arith<VT>::X* __restrict__ y_arg, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    constexpr int kSU = SU;  // groups of 4 entries per lane per chunk
    constexpr int kWv = kSwBS / 64;
    constexpr int kChunk = 64 * kSU;  // groups of 4 entries per wave per chunk
    extern __shared__ __align__(16) unsigned char csb_smem[];
    Acc* ylds = (Acc*)csb_smem;
    __shared__ double red[48];

    const X* __restrict__ x = x_arg;
    X* __restrict__ y = y_arg;
    double scale = 1.0;
    if constexpr (POWER) {
        if (pw.ctrl[1]) return;
        const int k = pw.ctrl[0];
        x = (const X*)((k & 1) ? pw.buf1 : pw.buf0);
        y = (X*)((k & 1) ? pw.buf0 : pw.buf1);
        scale = pwmk::step_scale(pw, k);
    }
    const VT* __restrict__ val = (const VT*)a.val;
    const uint32_t* __restrict__ ind = a.ind;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t t = blockIdx.x;
    const int32_t r0 = a.sstart[t];
    const int nrows = a.sstart[t + 1] - r0;
    for (int i = tid; i < nrows; i += kSwBS) ylds[i] = 0;
    const int wb = a.wb;
    const uint32_t wmask = (uint32_t)(((uint64_t)1 << wb) - 1);
    const int32_t* __restrict__ woff = a.soff + t * a.W;
    SweepChunk<VT, kSU, NT> cur, nxt;
    int32_t e0 = woff[0], e1 = woff[1];
    int32_t q0 = e0 >> 2, q1 = (int32_t)(((int64_t)e1 + 3) >> 2);
    cur.load(ind, val, q0 + wv * kChunk, q1, lane);
    __syncthreads();
    for (int64_t p = 0; p < a.W; ++p) {
        const X* __restrict__ xw = x + (p << wb);
        const int32_t pe0 = e0, pe1 = e1, pq0 = q0, pq1 = q1;
        if (p + 1 < a.W) {  // the next window's run and its first chunk, loaded ahead
            e0 = woff[p + 1];
            e1 = woff[p + 2];
            q0 = e0 >> 2;
            q1 = (int32_t)(((int64_t)e1 + 3) >> 2);
            nxt.load(ind, val, q0 + wv * kChunk, q1, lane);
        }
        sweep_add<VT, kSU, NT>(cur, pq0 + wv * kChunk, lane, pe0, pe1, xw, wmask, wb, ylds);
        // the wave's further chunks of this window, if its share is larger
        for (int32_t qb = pq0 + (wv + kWv) * kChunk; qb < pq1; qb += kWv * kChunk) {
            SweepChunk<VT, kSU, NT> more;
            more.load(ind, val, qb, pq1, lane);
            sweep_add<VT, kSU, NT>(more, qb, lane, pe0, pe1, xw, wmask, wb, ylds);
        }
        if (((p + 1) & (a.sync_every - 1)) == 0) __syncthreads();  // windows up to p done by every wave
        cur = nxt;
    }
    __syncthreads();
    double yy = 0.0, xy = 0.0, xx = 0.0;
    for (int i = tid; i < nrows; i += kSwBS) {
        const int64_t row = (int64_t)r0 + i;
        if constexpr (POWER) {
            const X out = (X)((double)ylds[i] * scale);
            y[row] = out;
            const double yd = (double)out, xd = (double)x[row] * scale;
            yy += yd * yd;
            xy += xd * yd;
            xx += xd * xd;
        } else {
            y[row] = (X)ylds[i];
        }
    }
    if constexpr (POWER) {
        pwmk::block_sum3<kSwBS>(yy, xy, xx, red);
        if (tid == 0) {
            pw.partials[3 * t + 0] = yy;
            pw.partials[3 * t + 1] = xy;
            pw.partials[3 * t + 2] = xx;
        }
    }
}

template <typename VT>
static int sweep_blocks_t(int rows) {
    using Acc = typename Arith<VT>::A;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_csb_sweep<VT, false, true>, kSwBS,
                                                     (size_t)rows * sizeof(Acc)) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

int csb_sweep_blocks_per_cu(int vtype, int rows) {
    switch (vtype) {
        case 0: return sweep_blocks_t<double>(rows);
        case 1: return sweep_blocks_t<float>(rows);
        case 2: return sweep_blocks_t<bf16_t>(rows);
    }
    return 0;
}

// ---------------------------------------------------------------- BlockCOH

// One workgroup per task (CohTask): chunks of kCohChunk consecutive entries
// of one band, decoded (ICRS, TripletToCRS.hpp:402-452 / BlockCOH.hpp:235-256)
// with two segmented scans that restart at every block:
//   S = running sum of col_jump: column = S mod beta, row changes = S div beta
//   (a row change adds exactly beta, so S = col + beta * changes);
//   R = running sum of the row_jump entered at each row change: the row.
// LDS tasks own their band (rows <= LDS window) and write it once; atomic
// tasks add into y (zeroed by k_coh_zero) with global atomics.
// OCC8: registers capped so that 8 waves per SIMD fit (4 workgroups per CU
// when the LDS window allows it) instead of the compiler's 7.
template <typename VT, bool POWER, bool OCC8>
__global__ __launch_bounds__(kBS, OCC8 ? 8 : 1) void k_coh(CohArgs a, const typename Arith<VT>::X* __restrict__ x_arg,
                                             typename Arith<VT>::X* __restrict__ y_arg, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    // the LDS window holds the rows of the largest band accumulated in LDS
    // (at most LdsRows<Acc>::v), sized at launch
    extern __shared__ __align__(16) unsigned char coh_smem[];
    Acc* ylds = (Acc*)coh_smem;
    __shared__ int32_t bstart[kCohChunk + 2];  // up to kCohChunk+1 blocks and a sentinel
    // the records of the chunk's blocks when there are at most kCohStage of
    // them (the usual case), so the decode reads them from LDS
    constexpr int kCohStage = 256;
    __shared__ CohBlock bstage[kCohStage];
    __shared__ int tv[kWaves], tf[kWaves];
    __shared__ int carry_s, carry_r;
    __shared__ double red[16];

    const X* __restrict__ x = x_arg;
    X* __restrict__ y = y_arg;
    double scale = 1.0;
    if constexpr (POWER) {
        if (pw.ctrl[1]) return;
        const int k = pw.ctrl[0];
        x = (const X*)((k & 1) ? pw.buf1 : pw.buf0);
        y = (X*)((k & 1) ? pw.buf0 : pw.buf1);
        scale = pwmk::step_scale(pw, k);
    }
    const int tid = threadIdx.x;
    const CohTask task = a.tasks[blockIdx.x];
    const CohBand band = a.bands[task.band];
    const VT* __restrict__ val = (const VT*)a.val + band.entry;
    const uint16_t* __restrict__ cjmp = a.col_jump + band.entry;
    const uint16_t* __restrict__ rjmp = a.row_jump + band.entry;
    const CohBlock* __restrict__ blocks = a.blocks + band.first_block;
    const int32_t* __restrict__ ctab = a.chunk_block + band.chunk_table;
    const int bits = band.beta_bits;
    const int64_t beta = (int64_t)1 << bits;
    const int nnz = band.nnz;
    const bool lds = task.atomic == 0;

    if (lds)
        for (int i = tid; i < band.rows; i += kBS) ylds[i] = 0;
    if (tid == 0) {
        carry_s = task.carry_s;
        carry_r = task.carry_r;
    }
    for (int c = task.chunk0; c < task.chunk0 + task.nchunks; ++c) {
        const int base = c * kCohChunk;
        const int b_lo = ctab[c], b_hi = ctab[c + 1];  // blocks touching this chunk
        const int nbl = b_hi - b_lo + 1;
        __syncthreads();
        const bool staged = nbl <= kCohStage;
        for (int i = tid; i < nbl; i += kBS) {
            const CohBlock cb = blocks[b_lo + i];
            bstart[i] = cb.entry;
            if (staged) bstage[i] = cb;
        }
        if (tid == 0) bstart[nbl] = 0x7fffffff;
        __syncthreads();
        const int e0 = base + tid * kCohIPT;
        // block of my first entry: last block start <= e0
        int lo = 0, hi = nbl;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (bstart[mid] <= e0) lo = mid;
            else hi = mid;
        }
        int bj[kCohIPT], cj[kCohIPT], rst[kCohIPT];
        Acc vv[kCohIPT];
        int b = lo;
#pragma unroll
        for (int j = 0; j < kCohIPT; ++j) {
            const int e = e0 + j;
            if (e < nnz) {
                while (bstart[b + 1] <= e) ++b;
                bj[j] = b;
                rst[j] = bstart[b] == e;
                cj[j] = cjmp[e];
                vv[j] = (Acc)widen(val[e]);
            } else {
                bj[j] = b;
                rst[j] = 0;
                cj[j] = 0;
                vv[j] = 0;
            }
        }
        // scan 1: S
        int run = 0, anyr = 0;
#pragma unroll
        for (int j = 0; j < kCohIPT; ++j) {
            run = rst[j] ? cj[j] : run + cj[j];
            anyr |= rst[j];
        }
        const int sin = block_seg_excl(run, anyr, carry_s, tv, tf);
        int S[kCohIPT];
        run = sin;
#pragma unroll
        for (int j = 0; j < kCohIPT; ++j) {
            run = rst[j] ? cj[j] : run + cj[j];
            S[j] = run;
        }
        // row deltas at every row change
        CohBlock blk[kCohIPT];
        int dl[kCohIPT];
#pragma unroll
        for (int j = 0; j < kCohIPT; ++j) {
            const int e = e0 + j;
            dl[j] = 0;
            if (e < nnz) {
                blk[j] = staged ? bstage[bj[j]] : blocks[b_lo + bj[j]];
                const int kk = S[j] >> bits;
                const bool newrow = rst[j] || (kk != ((S[j] - cj[j]) >> bits));
                // a band holds at most nnz row segments: the clamp keeps a
                // decode error an error in y, never a read outside the band
                int ri = blk[j].row_jump + kk;
                ri = ri < nnz ? ri : nnz - 1;
                if (newrow) dl[j] = rjmp[ri];
            } else {
                blk[j] = CohBlock{0, 0, 0, 0};
            }
        }
        // scan 2: R
        run = 0;
#pragma unroll
        for (int j = 0; j < kCohIPT; ++j) run = rst[j] ? dl[j] : run + dl[j];
        const int rin = block_seg_excl(run, anyr, carry_r, tv, tf);
        run = rin;
#pragma unroll
        for (int j = 0; j < kCohIPT; ++j) {
            run = rst[j] ? dl[j] : run + dl[j];
            const int e = e0 + j;
            if (e < nnz) {
                const int col = S[j] & (int)(beta - 1);
                const int row = (int)(blk[j].brow * beta) + run;  // row inside the band
                if (row >= 0 && row < band.rows) {
                    int64_t xi = (int64_t)blk[j].bcol * beta + col;
                    xi = xi < a.noc ? xi : a.noc - 1;
                    const Acc p = vv[j] * (Acc)x[xi];
                    if (lds) {
                        atomicAdd(&ylds[row], p);
                    } else {
                        // atomic tasks add the scaled product: y rows of big
                        // bands are final once every task has finished
                        atomicAdd(&y[(int64_t)band.start + row], POWER ? (X)((double)p * scale) : (X)p);
                    }
                }
            }
            // the entry that closes the chunk leaves its S and R for the next chunk
            if (e == base + kCohChunk - 1 || e == nnz - 1) {
                carry_s = S[j];
                carry_r = run;
            }
        }
    }
    __syncthreads();
    if (!lds) return;
    double yy = 0.0, xy = 0.0, xx = 0.0;
    for (int i = tid; i < band.rows; i += kBS) {
        const int64_t row = (int64_t)band.start + i;
        if constexpr (POWER) {
            const X out = (X)((double)ylds[i] * scale);
            y[row] = out;
            const double yd = (double)out, xd = (double)x[row] * scale;
            yy += yd * yd;
            xy += xd * yd;
            xx += xd * xd;
        } else {
            y[row] = (X)ylds[i];
        }
    }
    if constexpr (POWER) {
        const double ra = pwmk::block_sum<kBS>(yy, red);
        const double rb = pwmk::block_sum<kBS>(xy, red);
        const double rc = pwmk::block_sum<kBS>(xx, red);
        if (tid == 0) {
            pw.partials[3 * blockIdx.x + 0] = ra;
            pw.partials[3 * blockIdx.x + 1] = rb;
            pw.partials[3 * blockIdx.x + 2] = rc;
        }
    }
}

// Zero the rows of the bands that atomic tasks accumulate into (one workgroup
// per such band), and in power mode clear the atomic tasks' partial slots.
template <typename VT, bool POWER>
__global__ __launch_bounds__(kBS) void k_coh_zero(CohArgs a, const int32_t* __restrict__ atomic_bands,
                                                  typename Arith<VT>::X* __restrict__ y_arg, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    X* y = y_arg;
    if constexpr (POWER) {
        if (pw.ctrl[1]) return;
        const int k = pw.ctrl[0];
        y = (X*)((k & 1) ? pw.buf0 : pw.buf1);
    }
    const CohBand band = a.bands[atomic_bands[blockIdx.x]];
    for (int i = threadIdx.x; i < band.rows; i += kBS) y[(int64_t)band.start + i] = (X)0;
    if constexpr (POWER)
        for (int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x; t < a.n_tasks; t += (int64_t)gridDim.x * kBS)
            if (a.tasks[t].atomic) {
                pw.partials[3 * t + 0] = 0.0;
                pw.partials[3 * t + 1] = 0.0;
                pw.partials[3 * t + 2] = 0.0;
            }
}

// Power mode: |y|^2, x^T y, x^T x over the rows of atomic bands, into the
// partial slots after the tasks' (slot n_tasks + i for atomic band i).
template <typename VT>
__global__ __launch_bounds__(kBS) void k_coh_stats(CohArgs a, const int32_t* __restrict__ atomic_bands, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    __shared__ double red[16];
    if (pw.ctrl[1]) return;
    const int k = pw.ctrl[0];
    const X* x = (const X*)((k & 1) ? pw.buf1 : pw.buf0);
    const X* y = (const X*)((k & 1) ? pw.buf0 : pw.buf1);
    const double scale = pwmk::step_scale(pw, k);
    const CohBand band = a.bands[atomic_bands[blockIdx.x]];
    double yy = 0.0, xy = 0.0, xx = 0.0;
    for (int i = threadIdx.x; i < band.rows; i += kBS) {
        const int64_t row = (int64_t)band.start + i;
        const double yd = (double)y[row], xd = (double)x[row] * scale;
        yy += yd * yd;
        xy += xd * yd;
        xx += xd * xd;
    }
    const double ra = pwmk::block_sum<kBS>(yy, red);
    const double rb = pwmk::block_sum<kBS>(xy, red);
    const double rc = pwmk::block_sum<kBS>(xx, red);
    if (threadIdx.x == 0) {
        const int64_t slot = a.n_tasks + blockIdx.x;
        pw.partials[3 * slot + 0] = ra;
        pw.partials[3 * slot + 1] = rb;
        pw.partials[3 * slot + 2] = rc;
    }
}

// ---------------------------------------------------------------- launchers

template <typename VT>
static hipError_t csb_t(const CsbArgs& a, const void* x, void* y, const PowerArgs* pw, hipStream_t s, int part) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    if (a.n_tasks <= 0) return hipSuccess;
    if (a.sweep) {
        if (part == 2) return hipSuccess;  // no column-group reduction
        const size_t lds = (size_t)a.max_rows * sizeof(Acc);
        auto pick = [&](auto power) {
            constexpr bool P = decltype(power)::value;
            if (a.unroll == 1) return a.nt ? k_csb_sweep<VT, P, true, 1> : k_csb_sweep<VT, P, false, 1>;
            if (a.unroll == 4) return a.nt ? k_csb_sweep<VT, P, true, 4> : k_csb_sweep<VT, P, false, 4>;
            return a.nt ? k_csb_sweep<VT, P, true, 2> : k_csb_sweep<VT, P, false, 2>;
        };
        auto kp = pick(std::true_type{});
        auto kf = pick(std::false_type{});
        if (pw)
            hipLaunchKernelGGL(kp, dim3((unsigned)a.n_tasks), dim3(kSwBS), lds, s, a, (const X*)nullptr, (X*)nullptr,
                               *pw);
        else
            hipLaunchKernelGGL(kf, dim3((unsigned)a.n_tasks), dim3(kSwBS), lds, s, a, (const X*)x, (X*)y, PowerArgs{});
        return hipGetLastError();
    }
    if (part == 2) {
        // the column-group reduction only
    } else if (a.stream) {
        const size_t lds = (size_t)a.slab_rows * sizeof(Acc);
        const dim3 grid((unsigned)a.n_tasks);
        // kernel instance for (power step, unroll, non-temporal)
        auto pick = [&](auto power, auto wide) {
            constexpr bool P = decltype(power)::value, W = decltype(wide)::value;
            if (a.unroll == 4) return a.nt ? k_csb_stream<VT, P, 4, true, W> : k_csb_stream<VT, P, 4, false, W>;
            return a.nt ? k_csb_stream<VT, P, 2, true, W> : k_csb_stream<VT, P, 2, false, W>;
        };
        if (pw)
            hipLaunchKernelGGL(a.wide ? pick(std::true_type{}, std::true_type{}) : pick(std::true_type{}, std::false_type{}),
                               grid, dim3(kSBS), lds, s, a, (const X*)nullptr, (X*)nullptr, *pw);
        else
            hipLaunchKernelGGL(a.wide ? pick(std::false_type{}, std::true_type{})
                                      : pick(std::false_type{}, std::false_type{}),
                               grid, dim3(kSBS), lds, s, a, (const X*)x, (X*)y, PowerArgs{});
    } else if (pw) {
        hipLaunchKernelGGL((k_csb<VT, true>), dim3((unsigned)a.n_tasks), dim3(kBS), 0, s, a, (const X*)nullptr,
                           (X*)nullptr, *pw);
    } else {
        hipLaunchKernelGGL((k_csb<VT, false>), dim3((unsigned)a.n_tasks), dim3(kBS), 0, s, a, (const X*)x, (X*)y,
                           PowerArgs{});
    }
    if (a.groups > 1 && part != 1) {
        const unsigned rb = (unsigned)csb_reduce_blocks(a.nor);
        if (pw)
            hipLaunchKernelGGL((k_csb_reduce<VT, true>), dim3(rb), dim3(kBS), 0, s, a, (X*)nullptr, *pw);
        else
            hipLaunchKernelGGL((k_csb_reduce<VT, false>), dim3(rb), dim3(kBS), 0, s, a, (X*)y, PowerArgs{});
    }
    return hipGetLastError();
}

hipError_t launch_csb(const CsbArgs& a, const void* x, void* y, const PowerArgs* pw, hipStream_t s, int part) {
    switch (a.vtype) {
        case 0: return csb_t<double>(a, x, y, pw, s, part);
        case 1: return csb_t<float>(a, x, y, pw, s, part);
        case 2: return csb_t<bf16_t>(a, x, y, pw, s, part);
    }
    return hipErrorInvalidValue;
}

template <typename VT>
static hipError_t coh_t(const CohArgs& a, const int32_t* atomic_bands, int n_atomic, const void* x, void* y,
                        const PowerArgs* pw, hipStream_t s) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    if (a.n_tasks <= 0) return hipSuccess;
    const size_t lds = (size_t)std::max<int32_t>(a.lds_rows, 1) * sizeof(Acc);
    if (pw) {
        if (n_atomic > 0)
            hipLaunchKernelGGL((k_coh_zero<VT, true>), dim3(n_atomic), dim3(kBS), 0, s, a, atomic_bands, (X*)nullptr,
                               *pw);
        auto kp = a.occ8 ? k_coh<VT, true, true> : k_coh<VT, true, false>;
        hipLaunchKernelGGL(kp, dim3((unsigned)a.n_tasks), dim3(kBS), lds, s, a, (const X*)nullptr, (X*)nullptr, *pw);
        if (n_atomic > 0)
            hipLaunchKernelGGL((k_coh_stats<VT>), dim3(n_atomic), dim3(kBS), 0, s, a, atomic_bands, *pw);
    } else {
        if (n_atomic > 0)
            hipLaunchKernelGGL((k_coh_zero<VT, false>), dim3(n_atomic), dim3(kBS), 0, s, a, atomic_bands, (X*)y,
                               PowerArgs{});
        auto kf = a.occ8 ? k_coh<VT, false, true> : k_coh<VT, false, false>;
        hipLaunchKernelGGL(kf, dim3((unsigned)a.n_tasks), dim3(kBS), lds, s, a, (const X*)x, (X*)y, PowerArgs{});
    }
    return hipGetLastError();
}

hipError_t launch_coh(const CohArgs& a, const int32_t* atomic_bands, int n_atomic, const void* x, void* y,
                      const PowerArgs* pw, hipStream_t s) {
    switch (a.vtype) {
        case 0: return coh_t<double>(a, atomic_bands, n_atomic, x, y, pw, s);
        case 1: return coh_t<float>(a, atomic_bands, n_atomic, x, y, pw, s);
        case 2: return coh_t<bf16_t>(a, atomic_bands, n_atomic, x, y, pw, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_blocked_finalize(int vtype, int n_partials, const PowerArgs& pw, hipStream_t s) {
    pwmk::CrsDev A;
    A.vtype = vtype;
    pwmk::LongRows none;
    return pwmk::launch_crs_finalize(A, n_partials, none, nullptr, &pw, s);
}

}  // namespace blocked
