// CRS vector SpMV for gfx950.
//
// Rows of length <= lr.threshold: a group of G lanes (G = 1..64, a power of
// two) per row; each lane strides over the row by G, the partial sums are
// combined with DPP/ds_swizzle (group_sum), lane 0 writes y.  The loads of a
// row's val/col_ind are coalesced across the group, x is gathered.  A lane's
// first four entries are loaded and gathered unconditionally (clamped to the
// row), so a row of at most 4G entries costs one round of independent loads
// instead of a loop of dependent ones; longer rows continue in the loop.
// Rows longer than the threshold are cut into chunks; each chunk is summed by
// one workgroup and crs_finalize adds a row's chunk sums in chunk order, so a
// 50 000-entry row of a power-law matrix does not hold one wave for its whole
// length.
//
// Work distribution: units of 64/G rows (one wave); block b is renumbered by
// xcd_remap so each XCD sweeps a contiguous eighth of the rows (neighbouring
// rows share x entries through that XCD's L2).
//
// Power-method mode (POWER): step k reads buf[k&1] and writes buf[(k+1)&1]
// scaled by 1/|y_{k-1}| (the reference normalises y in a separate pass,
// Util/VectorUtill.hpp:34-41, and then multiplies it; here the division is
// folded into the next SpMV's epilogue), and every block leaves FP64 partial
// sums of |y|^2, x^T y and x^T x for crs_finalize.
#include "kernels.hpp"

namespace pwmk {

constexpr int kBS = 256;

template <int G, typename VT, bool POWER>
__global__ __launch_bounds__(kBS) void k_crs_vector(CrsDev A, int n_vec_blocks, LongRows lr,
                                                    const typename Arith<VT>::X* __restrict__ x_arg,
                                                    typename Arith<VT>::X* __restrict__ y_arg, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    __shared__ double red[16];

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
    const int32_t* __restrict__ rs = A.rs;
    const int32_t* __restrict__ ci = A.ci;
    const VT* __restrict__ val = (const VT*)A.val;

    if ((int)blockIdx.x >= n_vec_blocks) {
        // one chunk of one long row
        const LongChunk c = lr.chunks[blockIdx.x - n_vec_blocks];
        Acc sum = 0;
        int32_t k = c.k0 + threadIdx.x;
        for (; k + 3 * kBS < c.k1; k += 4 * kBS) {
            const int32_t j0 = ci[k], j1 = ci[k + kBS], j2 = ci[k + 2 * kBS], j3 = ci[k + 3 * kBS];
            const Acc v0 = widen(val[k]), v1 = widen(val[k + kBS]), v2 = widen(val[k + 2 * kBS]),
                      v3 = widen(val[k + 3 * kBS]);
            sum += v0 * (Acc)x[j0];
            sum += v1 * (Acc)x[j1];
            sum += v2 * (Acc)x[j2];
            sum += v3 * (Acc)x[j3];
        }
        for (; k < c.k1; k += kBS) sum += widen(val[k]) * (Acc)x[ci[k]];
        const double t = block_sum<kBS>((double)sum, red);
        if (threadIdx.x == 0) ((double*)lr.chunk_partial)[blockIdx.x - n_vec_blocks] = t;
        return;
    }

    constexpr int RPW = 64 / G;  // rows per wave step
    const int lane64 = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lane = lane64 & (G - 1);
    const int grp = lane64 / G;
    const int64_t nrows = A.nrows;
    const int64_t U = (nrows + RPW - 1) / RPW;
    const int64_t vb = xcd_remap(blockIdx.x, n_vec_blocks);
    const int64_t upb = (U + n_vec_blocks - 1) / n_vec_blocks;
    const int64_t u0 = vb * upb;
    const int64_t u1 = u0 + upb < U ? u0 + upb : U;

    double yy = 0.0, xy = 0.0, xx = 0.0;
    for (int64_t u = u0 + wave; u < u1; u += kBS / 64) {
        const int64_t row = u * RPW + grp;
        int32_t s = 0, e = 0;
        bool valid = row < nrows;
        if (valid) {
            s = rs[row];
            e = rs[row + 1];
            if (e - s > lr.threshold) {
                e = s;
                valid = false;
            }
        }
        Acc sum = 0;
        // a lane's first 4 entries (all of a row of <= 4G) as independent
        // loads; masked entries are pointed at the row's first entry and not
        // added
        if (e > s) {
            int32_t kk[4];
            bool ok[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int32_t k = s + lane + j * G;
                ok[j] = k < e;
                kk[j] = ok[j] ? k : s;
            }
            int32_t jj[4];
            Acc vv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                jj[j] = ci[kk[j]];
                vv[j] = widen(val[kk[j]]);
            }
            Acc xv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) xv[j] = (Acc)x[jj[j]];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (ok[j]) sum += vv[j] * xv[j];
        }
        int32_t k = s + lane + 4 * G;
        for (; k + 3 * G < e; k += 4 * G) {
            const int32_t j0 = ci[k], j1 = ci[k + G], j2 = ci[k + 2 * G], j3 = ci[k + 3 * G];
            const Acc v0 = widen(val[k]), v1 = widen(val[k + G]), v2 = widen(val[k + 2 * G]),
                      v3 = widen(val[k + 3 * G]);
            sum += v0 * (Acc)x[j0];
            sum += v1 * (Acc)x[j1];
            sum += v2 * (Acc)x[j2];
            sum += v3 * (Acc)x[j3];
        }
        for (; k < e; k += G) sum += widen(val[k]) * (Acc)x[ci[k]];
        sum = group_sum<G>(sum);
        if (lane == 0 && valid) {
            if constexpr (POWER) {
                const X out = (X)((double)sum * scale);
                y[row] = out;
                const double yd = (double)out, xd = (double)x[row] * scale;
                yy += yd * yd;
                xy += xd * yd;
                xx += xd * xd;
            } else {
                y[row] = (X)sum;
            }
        }
    }
    if constexpr (POWER) {
        const double a = block_sum<kBS>(yy, red);
        const double b = block_sum<kBS>(xy, red);
        const double c = block_sum<kBS>(xx, red);
        if (threadIdx.x == 0) {
            pw.partials[3 * blockIdx.x + 0] = a;
            pw.partials[3 * blockIdx.x + 1] = b;
            pw.partials[3 * blockIdx.x + 2] = c;
        }
    }
}

constexpr int kFinBS = 1024;

// Long rows: y[row] = sum of the row's chunk partials in chunk order.  Power
// mode: FP64 reduction of the block partials (fixed order), then step k's
// record, the convergence test and k+1.
template <typename VT, bool POWER>
__global__ __launch_bounds__(kFinBS) void k_crs_finalize(CrsDev A, int n_partials, LongRows lr,
                                                         typename Arith<VT>::X* __restrict__ y_arg, PowerArgs pw) {
    using X = typename Arith<VT>::X;
    __shared__ double red[16];
    X* __restrict__ y = y_arg;
    const X* __restrict__ x = nullptr;
    double scale = 1.0;
    int k = 0;
    if constexpr (POWER) {
        if (pw.ctrl[1]) return;
        k = pw.ctrl[0];
        x = (const X*)((k & 1) ? pw.buf1 : pw.buf0);
        y = (X*)((k & 1) ? pw.buf0 : pw.buf1);
        scale = step_scale(pw, k);
    }
    double yy = 0.0, xy = 0.0, xx = 0.0;
    const double* part = (const double*)lr.chunk_partial;
    for (int r = threadIdx.x; r < lr.n_rows; r += kFinBS) {
        double s = 0.0;
        for (int c = lr.chunk_begin[r]; c < lr.chunk_begin[r + 1]; ++c) s += part[c];
        const int32_t row = lr.row_ids[r];
        const X out = (X)(s * scale);
        y[row] = out;
        if constexpr (POWER) {
            const double yd = (double)out, xd = (double)x[row] * scale;
            yy += yd * yd;
            xy += xd * yd;
            xx += xd * xd;
        }
    }
    if constexpr (POWER) {
        for (int i = threadIdx.x; i < n_partials; i += kFinBS) {
            yy += pw.partials[3 * i + 0];
            xy += pw.partials[3 * i + 1];
            xx += pw.partials[3 * i + 2];
        }
        const double a = block_sum<kFinBS>(yy, red);
        const double b = block_sum<kFinBS>(xy, red);
        const double c = block_sum<kFinBS>(xx, red);
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
}

template <int G, typename VT>
static hipError_t launch_vec_g(const CrsDev& A, int n_vec_blocks, const LongRows& lr, const void* x, void* y,
                               const PowerArgs* pw, hipStream_t s) {
    using X = typename Arith<VT>::X;
    const int grid = n_vec_blocks + lr.n_chunks;
    if (grid <= 0) return hipSuccess;
    if (pw)
        hipLaunchKernelGGL((k_crs_vector<G, VT, true>), dim3(grid), dim3(kBS), 0, s, A, n_vec_blocks, lr,
                           (const X*)nullptr, (X*)nullptr, *pw);
    else
        hipLaunchKernelGGL((k_crs_vector<G, VT, false>), dim3(grid), dim3(kBS), 0, s, A, n_vec_blocks, lr,
                           (const X*)x, (X*)y, PowerArgs{});
    return hipGetLastError();
}

template <typename VT>
static hipError_t launch_vec_t(const CrsDev& A, int G, int nb, const LongRows& lr, const void* x, void* y,
                               const PowerArgs* pw, hipStream_t s) {
    switch (G) {
        case 1: return launch_vec_g<1, VT>(A, nb, lr, x, y, pw, s);
        case 2: return launch_vec_g<2, VT>(A, nb, lr, x, y, pw, s);
        case 4: return launch_vec_g<4, VT>(A, nb, lr, x, y, pw, s);
        case 8: return launch_vec_g<8, VT>(A, nb, lr, x, y, pw, s);
        case 16: return launch_vec_g<16, VT>(A, nb, lr, x, y, pw, s);
        case 32: return launch_vec_g<32, VT>(A, nb, lr, x, y, pw, s);
        case 64: return launch_vec_g<64, VT>(A, nb, lr, x, y, pw, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_crs_vector(const CrsDev& A, int G, int n_vec_blocks, const LongRows& lr, const void* x, void* y,
                             const PowerArgs* pw, hipStream_t s) {
    switch (A.vtype) {
        case VT_F64: return launch_vec_t<double>(A, G, n_vec_blocks, lr, x, y, pw, s);
        case VT_F32: return launch_vec_t<float>(A, G, n_vec_blocks, lr, x, y, pw, s);
        case VT_BF16: return launch_vec_t<bf16_t>(A, G, n_vec_blocks, lr, x, y, pw, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename VT>
static hipError_t launch_fin_t(const CrsDev& A, int n_partials, const LongRows& lr, void* y, const PowerArgs* pw,
                               hipStream_t s) {
    using X = typename Arith<VT>::X;
    if (pw)
        hipLaunchKernelGGL((k_crs_finalize<VT, true>), dim3(1), dim3(kFinBS), 0, s, A, n_partials, lr, (X*)nullptr,
                           *pw);
    else if (lr.n_rows > 0)
        hipLaunchKernelGGL((k_crs_finalize<VT, false>), dim3(1), dim3(kFinBS), 0, s, A, n_partials, lr, (X*)y,
                           PowerArgs{});
    return hipGetLastError();
}

hipError_t launch_crs_finalize(const CrsDev& A, int n_partials, const LongRows& lr, void* y, const PowerArgs* pw,
                               hipStream_t s) {
    switch (A.vtype) {
        case VT_F64: return launch_fin_t<double>(A, n_partials, lr, y, pw, s);
        case VT_F32: return launch_fin_t<float>(A, n_partials, lr, y, pw, s);
        case VT_BF16: return launch_fin_t<bf16_t>(A, n_partials, lr, y, pw, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pwmk
