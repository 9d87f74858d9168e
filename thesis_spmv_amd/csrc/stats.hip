// Matrix statistics (Util_Scripts/calc_stats.cpp:44-105), computed when a
// matrix is loaded: the per-block nonzero counts on the device (one pass over
// the entries, global atomic adds into an nb x nb table), the row-length
// statistics and every floating-point sum on the host in the tool's order.
#include <math.h>

#include "api_internal.hpp"
#include "kernels.hpp"

namespace pwmk {

// Entry k of row r adds one to block (r >> bits, col >> bits); r is the last
// row with row_start[r] <= k.  Column blocks past the nb x nb table are not
// counted.
__global__ void k_block_hist(const int32_t* __restrict__ rs, const int32_t* __restrict__ ci, int64_t nor, int64_t nnz,
                             int bits, int64_t nb, uint32_t* __restrict__ hist) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = nor - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (rs[mid] <= k)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int64_t bc = (int64_t)ci[k] >> bits;
        if (bc < nb) atomicAdd(&hist[(lo >> bits) * nb + bc], 1u);
    }
}

// The same counts a workgroup of 256 consecutive rows at a time: with blocks
// of at least 256 rows those rows share one block row, so the workgroup's
// threads stride over the rows' (contiguous) entries and count into an LDS
// histogram of that block row (nb bins); the workgroup then adds its
// non-zero bins to the table once.
constexpr int kHistRows = 256;
__global__ __launch_bounds__(kHistRows) void k_block_hist_rows(const int32_t* __restrict__ rs,
                                                               const int32_t* __restrict__ ci, int64_t nor, int bits,
                                                               int64_t nb, uint32_t* __restrict__ hist) {
    extern __shared__ uint32_t lh[];
    const int tid = threadIdx.x;
    for (int64_t i = tid; i < nb; i += kHistRows) lh[i] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * kHistRows;
    const int64_t r1 = r0 + kHistRows < nor ? r0 + kHistRows : nor;
    for (int64_t e = rs[r0] + tid, e1 = rs[r1]; e < e1; e += kHistRows) {
        const int64_t bc = (int64_t)ci[e] >> bits;
        if (bc < nb) atomicAdd(&lh[bc], 1u);
    }
    __syncthreads();
    const int64_t br = r0 >> bits;
    if (br < nb)
        for (int64_t i = tid; i < nb; i += kHistRows)
            if (lh[i]) atomicAdd(&hist[br * nb + i], lh[i]);
}

}  // namespace pwmk

namespace pwm_impl {

int compute_matrix_stats(pwm_matrix* h) {
#pragma clang fp contract(off)
    const int64_t N = h->nor, Z = h->nnz;
    pwm_stats_t st{};
    st.nor = N;
    st.nnz = Z;
    const double avg = (double)Z / (double)N;
    st.rows_avg = avg;
    std::vector<int32_t> down;
    const int32_t* rs = h->h_rs.data();
    hipError_t e;
    if ((int64_t)h->h_rs.size() != N + 1) {
        down.resize(N + 1);
        if ((e = hipMemcpy(down.data(), h->rs.p, (N + 1) * 4, hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(h, e, "statistics: row_start download");
        rs = down.data();
    }
    double var = 0.0;
    int64_t mx = 0;
    for (int64_t i = 0; i < N; ++i) {
        const int64_t len = rs[i + 1] - rs[i];
        var += pow((double)len - avg, 2) / (double)N;
        if (len > mx) mx = len;
    }
    st.rows_var = var;
    st.row_max = mx;
    const int lg = (int)ceil(log2(sqrt((double)N)));
    const int bits = std::min(16, 3 + lg);
    const int64_t bs = (int64_t)1 << bits;
    const int64_t nb = (N + bs - 1) / bs;
    st.block_bits = bits;
    st.balanced = 1;
    if (nb > 8192) {
        st.nb_blocks = -1;
        h->stats = st;
        return PWM_OK;
    }
    std::vector<uint32_t> hist(nb * nb, 0u);
    if (Z > 0) {
        DevBuf d;
        if ((e = d.alloc(nb * nb * 4)) != hipSuccess) return hip_fail(h, e, "statistics table");
        e = hipMemsetAsync(d.p, 0, nb * nb * 4, h->stream);
        if (e == hipSuccess) {
            if (bits >= 8) {  // 256-row workgroups stay inside one block row
                const int64_t wgs = (N + pwmk::kHistRows - 1) / pwmk::kHistRows;
                hipLaunchKernelGGL(pwmk::k_block_hist_rows, dim3((unsigned)wgs), dim3(pwmk::kHistRows), nb * 4,
                                   h->stream, h->rs.as<int32_t>(), h->ci.as<int32_t>(), N, bits, nb, d.as<uint32_t>());
            } else {
                const int64_t blocks = std::min<int64_t>((Z + 255) / 256, 8192);
                hipLaunchKernelGGL(pwmk::k_block_hist, dim3((unsigned)blocks), dim3(256), 0, h->stream,
                                   h->rs.as<int32_t>(), h->ci.as<int32_t>(), N, Z, bits, nb, d.as<uint32_t>());
            }
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hist.data(), d.p, nb * nb * 4, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "statistics: block counts");
    }
    // whole blocks only (calc_stats.cpp:72-75), sums in the tool's order
    const int64_t nbw = N % bs != 0 ? nb - 1 : nb;
    double bavg = 0.0, brow_avg = 0.0, bvar = 0.0;
    std::vector<int64_t> brow(std::max<int64_t>(nbw, 1), 0);
    for (int64_t i = 0; i < nbw; ++i) {
        for (int64_t j = 0; j < nbw; ++j) {
            brow[i] += hist[i * nb + j];
            bavg += (double)hist[i * nb + j] / pow((double)nbw, 2);
        }
        brow_avg += (double)brow[i] / (double)nbw;
    }
    for (int64_t i = 0; i < nbw; ++i) {
        if ((double)brow[i] > 2 * brow_avg) st.balanced = 0;
        for (int64_t j = 0; j < nbw; ++j) bvar += pow((double)hist[i * nb + j] - bavg, 2) / pow((double)nbw, 2);
    }
    st.nb_blocks = nbw;
    st.block_avg = bavg;
    st.block_var = bvar;
    h->stats = st;
    return PWM_OK;
}

}  // namespace pwm_impl

using namespace pwm_impl;

extern "C" int pwm_matrix_stats(const pwm_matrix* h, pwm_stats_t* out) {
    if (!h || !out) return fail(const_cast<pwm_matrix*>(h), PWM_ERR_INVALID, "pwm_matrix_stats: NULL argument");
    if (!h->loaded) return fail(const_cast<pwm_matrix*>(h), PWM_ERR_STATE, "pwm_matrix_stats: no matrix loaded");
    *out = h->stats;
    return PWM_OK;
}
