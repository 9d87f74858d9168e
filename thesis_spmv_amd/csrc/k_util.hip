// Utility kernels: device Poisson generator, value conversion, vector fills
// and copies, the power method's final normalisation, cache flush.
#include <algorithm>

#include "kernels.hpp"

namespace pwmk {

__device__ __forceinline__ void store_val(void* val, int64_t k, double v, int vtype) {
    if (vtype == VT_F64)
        ((double*)val)[k] = v;
    else if (vtype == VT_F32)
        ((float*)val)[k] = (float)v;
    else
        ((uint16_t*)val)[k] = (uint16_t)(__float_as_uint((float)v) >> 16);  // exact for 4 and -1
}

// Util/Poisson.hpp:32-83, one thread per row.  The row's first entry offset
// is the closed-form count of entries in rows q < r (one per row, plus
// q >= m, q % m != 0, q % m != m-1 and q < N-m), the same decomposition the
// reference's fillPoissonOMP uses to fill rows independently (:111-122).
__global__ void k_fill_poisson(int64_t m, int64_t n, int32_t* __restrict__ rs, int32_t* __restrict__ ci,
                               void* __restrict__ val, int vtype) {
    const int64_t N = m * n;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int64_t below_m = r > m ? r - m : 0;              // q < r with q >= m
    const int64_t not_first = r - (r + m - 1) / m;          // q < r with q % m != 0
    const int64_t not_last = r - r / m;                     // q < r with q % m != m-1
    const int64_t before_last_block = r < N - m ? r : N - m;  // q < r with q < N-m
    int64_t k = r + below_m + not_first + not_last + before_last_block;
    if (r == 0) rs[0] = 0;
    if (r >= m) { ci[k] = (int32_t)(r - m); store_val(val, k++, -1.0, vtype); }
    if (r % m != 0) { ci[k] = (int32_t)(r - 1); store_val(val, k++, -1.0, vtype); }
    ci[k] = (int32_t)r; store_val(val, k++, 4.0, vtype);
    if (r % m != m - 1) { ci[k] = (int32_t)(r + 1); store_val(val, k++, -1.0, vtype); }
    if (r < N - m) { ci[k] = (int32_t)(r + m); store_val(val, k++, -1.0, vtype); }
    rs[r + 1] = (int32_t)k;
}

hipError_t launch_fill_poisson(int64_t m, int64_t n, int32_t* rs, int32_t* ci, void* val, int vtype, hipStream_t s) {
    const int64_t N = m * n;
    hipLaunchKernelGGL(k_fill_poisson, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, m, n, rs, ci, val, vtype);
    return hipGetLastError();
}

__global__ void k_convert(const double* __restrict__ src, void* __restrict__ dst, int64_t n, int vtype) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        store_val(dst, i, src[i], vtype);
}

static unsigned grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (unsigned)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

hipError_t launch_convert_values(const double* src, void* dst, int64_t n, int vtype, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_convert, dim3(grid_for(n)), dim3(256), 0, s, src, dst, n, vtype);
    return hipGetLastError();
}

__global__ void k_widen(const void* __restrict__ src, int vtype, double* __restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (vtype == VT_F64)
            dst[i] = ((const double*)src)[i];
        else if (vtype == VT_F32)
            dst[i] = (double)((const float*)src)[i];
        else
            dst[i] = (double)__uint_as_float(((uint32_t)((const uint16_t*)src)[i]) << 16);
    }
}

hipError_t launch_copy_values_f64(const void* src, int vtype, double* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_widen, dim3(grid_for(n)), dim3(256), 0, s, src, vtype, dst, n);
    return hipGetLastError();
}

__global__ void k_fill(void* p, int64_t n, int is_f64, double v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (is_f64)
            ((double*)p)[i] = v;
        else
            ((float*)p)[i] = (float)v;
    }
}

hipError_t launch_fill(void* p, int64_t n, int is_f64, double v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, p, n, is_f64, v);
    return hipGetLastError();
}

__global__ void k_copy_vec(const void* src, int src_f64, void* dst, int dst_f64, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = src_f64 ? ((const double*)src)[i] : (double)((const float*)src)[i];
        if (dst_f64)
            ((double*)dst)[i] = v;
        else
            ((float*)dst)[i] = (float)v;
    }
}

hipError_t launch_copy_vec(const void* src, int src_f64, void* dst, int dst_f64, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_vec, dim3(grid_for(n)), dim3(256), 0, s, src, src_f64, dst, dst_f64, n);
    return hipGetLastError();
}

// After K = ctrl[0] steps, buf[K&1] holds A*x_{K-1} (not yet divided by its
// norm) and buf[(K-1)&1] holds A*x_{K-2} (or the caller's start vector when
// K == 1).  Divide each by its own norm so both buffers hold what the
// reference's powerIteration leaves in x and y (Matrix/CRS.hpp:118-130).
__global__ void k_power_finish(PowerArgs pw, int64_t n, int is_f64) {
    const int K = pw.ctrl[0];
    if (K <= 0) return;
    void* last = (K & 1) ? pw.buf1 : pw.buf0;
    void* prev = (K & 1) ? pw.buf0 : pw.buf1;
    const double s_last = 1.0 / sqrt(pw.rec[4 * (K - 1)]);
    const double s_prev = K >= 2 ? 1.0 / sqrt(pw.rec[4 * (K - 2)]) : 1.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (is_f64) {
            ((double*)last)[i] = ((double*)last)[i] * s_last;
            if (K >= 2) ((double*)prev)[i] = ((double*)prev)[i] * s_prev;
        } else {
            ((float*)last)[i] = (float)((double)((float*)last)[i] * s_last);
            if (K >= 2) ((float*)prev)[i] = (float)((double)((float*)prev)[i] * s_prev);
        }
    }
}

hipError_t launch_power_finish(const PowerArgs& pw, int64_t n, int is_f64, int64_t, hipStream_t s) {
    hipLaunchKernelGGL(k_power_finish, dim3(grid_for(n)), dim3(256), 0, s, pw, n, is_f64);
    return hipGetLastError();
}

__global__ void k_power_reset(int* ctrl) {
    if (threadIdx.x == 0) {
        ctrl[0] = 0;
        ctrl[1] = 0;
    }
}

hipError_t launch_power_reset(const PowerArgs& pw, hipStream_t s) {
    hipLaunchKernelGGL(k_power_reset, dim3(1), dim3(64), 0, s, pw.ctrl);
    return hipGetLastError();
}

__global__ void k_stream_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_stream_copy(void* dst, const void* src, int64_t bytes, hipStream_t s) {
    const int64_t n16 = bytes / 16;
    if (n16 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stream_copy, dim3(8192), dim3(256), 0, s, (uint4*)dst, (const uint4*)src, n16);
    return hipGetLastError();
}

__global__ void k_flush(uint4* p, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x) {
        uint4 v = p[i];
        v.x += 1u;
        p[i] = v;
    }
}

hipError_t launch_flush(void* p, int64_t bytes, hipStream_t s) {
    const int64_t n16 = bytes / 16;
    if (n16 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_flush, dim3(4096), dim3(256), 0, s, (uint4*)p, n16);
    return hipGetLastError();
}

// Gather-rate probe: what one SpMV's x[col] gathers can reach at best.  A
// stream of (u32 index, f64 value) entries, 12 bytes each like a CRS
// nonzero, is read with coalesced loads and every index gathers one double
// of x: from all of x, or (per_xcd) from the window [xcd*window,
// (xcd+1)*window) of the XCD the workgroup runs on (b % 8), the way a CSB
// column group keeps its eighth of x.  Indices are a hash of the entry
// number modulo the window, so the gathers have no locality beyond it.
// window = 0: the stream alone, no gathers (the streaming ceiling).
__global__ void k_gather_fill(uint32_t* __restrict__ idx, double* __restrict__ val, int64_t n, uint32_t window) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        uint32_t v = (uint32_t)k * 0x9e3779b9u + (uint32_t)(k >> 32);
        v ^= v >> 16;
        v *= 0x7feb352du;
        v ^= v >> 15;
        v *= 0x846ca68bu;
        v ^= v >> 16;
        idx[k] = v % window;
        val[k] = 1.0;
    }
}

constexpr int kProbeU = 8;
constexpr int kProbePerBlock = 8192;  // entries per workgroup, a multiple of 256 * kProbeU

// AUX < 0: the gathers are plain global loads; else buffer loads (bounds
// checked against the x window) with cache-policy bits AUX: 1 sc0, 2 nt,
// 16 sc1.
template <bool PER_XCD, bool GATHER, int AUX = -1>
__global__ __launch_bounds__(256) void k_gather_probe(const uint32_t* __restrict__ idx, const double* __restrict__ val,
                                                      const double* __restrict__ x, double* __restrict__ out,
                                                      int64_t window) {
    const double* xw = PER_XCD ? x + (int64_t)(blockIdx.x % 8) * window : x;
    __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)xw, (short)0, (int)std::min<int64_t>(window * 8, INT32_MAX), 0x00020000);
    const int64_t base = (int64_t)blockIdx.x * kProbePerBlock;
    double acc = 0.0;
    for (int i = threadIdx.x; i < kProbePerBlock; i += kProbeU * 256) {
        uint32_t j[kProbeU];
        double v[kProbeU];
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) {
            j[u] = idx[base + i + u * 256];
            v[u] = val[base + i + u * 256];
        }
        double xv[kProbeU];
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) {
            if constexpr (!GATHER) {
                xv[u] = (double)j[u];
            } else if constexpr (AUX < 0) {
                xv[u] = xw[j[u]];
            } else {
                const auto w = __builtin_amdgcn_raw_buffer_load_b64(rsrc, (int)(j[u] * 8u), 0, AUX);
                __builtin_memcpy(&xv[u], &w, 8);
            }
        }
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) acc += v[u] * xv[u];
    }
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = acc;
}

int64_t gather_probe_entries(int64_t n) { return n / kProbePerBlock * kProbePerBlock; }

hipError_t launch_gather_fill(uint32_t* idx, double* val, int64_t n, uint32_t window, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_fill, dim3(4096), dim3(256), 0, s, idx, val, n, window);
    return hipGetLastError();
}

template <int AUX>
static void launch_probe_aux(unsigned blocks, const uint32_t* idx, const double* val, const double* x, double* out,
                             int64_t window, bool per_xcd, hipStream_t s) {
    if (per_xcd)
        hipLaunchKernelGGL((k_gather_probe<true, true, AUX>), dim3(blocks), dim3(256), 0, s, idx, val, x, out, window);
    else
        hipLaunchKernelGGL((k_gather_probe<false, true, AUX>), dim3(blocks), dim3(256), 0, s, idx, val, x, out, window);
}

hipError_t launch_gather_probe(const uint32_t* idx, const double* val, const double* x, double* out, int64_t n,
                               int64_t window, int flags, hipStream_t s) {
    const unsigned blocks = (unsigned)(n / kProbePerBlock);
    if (blocks == 0) return hipSuccess;
    const bool per_xcd = flags & 1;
    if (window == 0) {
        hipLaunchKernelGGL((k_gather_probe<false, false>), dim3(blocks), dim3(256), 0, s, idx, val, x, out, window);
        return hipGetLastError();
    }
    switch ((flags >> 8) & 0xff) {  // gather load kind
        case 0: launch_probe_aux<-1>(blocks, idx, val, x, out, window, per_xcd, s); break;
        case 1: launch_probe_aux<0>(blocks, idx, val, x, out, window, per_xcd, s); break;
        case 2: launch_probe_aux<1>(blocks, idx, val, x, out, window, per_xcd, s); break;
        case 3: launch_probe_aux<2>(blocks, idx, val, x, out, window, per_xcd, s); break;
        case 4: launch_probe_aux<3>(blocks, idx, val, x, out, window, per_xcd, s); break;
        case 17: launch_probe_aux<16>(blocks, idx, val, x, out, window, per_xcd, s); break;
        case 18: launch_probe_aux<17>(blocks, idx, val, x, out, window, per_xcd, s); break;
        case 19: launch_probe_aux<18>(blocks, idx, val, x, out, window, per_xcd, s); break;
        case 20: launch_probe_aux<19>(blocks, idx, val, x, out, window, per_xcd, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pwmk
