// Device-side helpers shared by the gfx950 kernels: storage/arithmetic types,
// wavefront reductions with DPP, and the XCD-aware block-to-work mapping.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pwmk {

// Storage type for BF16 matrix values (upper half of an IEEE float).
struct bf16_t {
    uint16_t u;
};

// Vector (x, y) and per-row accumulator types for each value storage type:
// FP64 values use FP64 vectors and FP64 accumulation; FP32 and BF16 values
// use FP32 vectors and FP32 accumulation.
template <typename VT> struct Arith;
template <> struct Arith<double> { using X = double; using A = double; };
template <> struct Arith<float> { using X = float; using A = float; };
template <> struct Arith<bf16_t> { using X = float; using A = float; };

__device__ __forceinline__ double widen(double v) { return v; }
__device__ __forceinline__ float widen(float v) { return v; }
__device__ __forceinline__ float widen(bf16_t v) { return __uint_as_float(((uint32_t)v.u) << 16); }

// ---- DPP lane exchange ---------------------------------------------------
// dpp_ctrl encodings (GCN3/CDNA): quad_perm = 0x00..0xFF,
// row_mirror = 0x140, row_half_mirror = 0x141.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ int dpp(int v) {
    return dpp_i32<CTRL>(v);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(dpp_i32<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
    long long b = __double_as_longlong(v);
    int lo = dpp_i32<CTRL>((int)(b & 0xffffffffll));
    int hi = dpp_i32<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ds_swizzle in bit-mask mode: lane i reads lane (i ^ 16) inside each
// 32-lane half (and_mask 0x1f, or_mask 0, xor_mask 0x10).
__device__ __forceinline__ int swz16(int v) { return __builtin_amdgcn_ds_swizzle(v, 0x401F); }
__device__ __forceinline__ float swz16(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
}
__device__ __forceinline__ double swz16(double v) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffll), 0x401F);
    int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x401F);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over aligned groups of G lanes (G = 1..64); every lane of a group ends
// with the group total.  Steps: xor 1 and xor 2 inside a quad (quad_perm),
// 4 via row_half_mirror, 8 via row_mirror (after the previous steps all lanes
// of a sub-group hold the same value, so a mirror is as good as an xor),
// 16 via ds_swizzle and 32 via ds_bpermute.  All 64 lanes must be active.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
    if constexpr (G >= 2) v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    if constexpr (G >= 4) v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
    if constexpr (G >= 8) v += dpp<0x141>(v);  // row_half_mirror
    if constexpr (G >= 16) v += dpp<0x140>(v); // row_mirror
    if constexpr (G >= 32) v += swz16(v);
    if constexpr (G >= 64) v += __shfl_xor(v, 32);
    return v;
}

// Block-wide FP64 sum of one value per thread in a fixed order (wave sums,
// then the waves' sums in wave order).  `scratch` holds >= 16 doubles.
// Result valid in every thread.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* scratch) {
    v = group_sum<64>(v);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) scratch[wave] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) t += scratch[w];
    __syncthreads();
    return t;
}

// Global-address-space view of a device pointer.  Pointers read from memory
// (job descriptors) are otherwise generic, and generic (flat) accesses count
// against the LDS/scalar wait counter as well, so every LDS wait would also
// drain all outstanding global loads.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> G(T* p) {
    return (gptr<T>)p;
}

// Three block sums in one pass (same order as three block_sum calls).
// `scratch` holds >= 48 doubles.
template <int BS>
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double* scratch) {
    a = group_sum<64>(a);
    b = group_sum<64>(b);
    c = group_sum<64>(c);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        scratch[wave] = a;
        scratch[16 + wave] = b;
        scratch[32 + wave] = c;
    }
    __syncthreads();
    double ta = 0.0, tb = 0.0, tc = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) {
        ta += scratch[w];
        tb += scratch[16 + w];
        tc += scratch[32 + w];
    }
    __syncthreads();
    a = ta;
    b = tb;
    c = tc;
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with its own 4 MB L2.  Renumber blocks so that XCD j works on the j-th
// contiguous eighth of the work items: neighbouring rows (which share x
// entries, e.g. the +-m neighbours of a stencil) then meet in the same L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nblocks) {
    constexpr int64_t kXcd = 8;
    if (nblocks < kXcd) return b;
    const int64_t per = nblocks / kXcd, extra = nblocks % kXcd;
    const int64_t xcd = b % kXcd, slot = b / kXcd;
    // XCD j owns per (+1 for the first `extra` XCDs) consecutive items
    return xcd * per + (xcd < extra ? xcd : extra) + slot;
}

// Per-iteration control block of the device-side power method.
struct PowerArgs {
    void* buf0;        // ping-pong vectors: step k reads buf[k&1], writes buf[(k+1)&1]
    void* buf1;
    int* ctrl;         // ctrl[0] = step k, ctrl[1] = done flag
    double* rec;       // rec[4k..4k+3] = |y|^2, x^T y, x^T x, lambda of step k
    double* partials;  // per-block FP64 partial sums (3 per block)
    int64_t max_it;
    double tol;        // < 0: never stop early
};

// FP64 scale applied to step k's output: 1/|y_{k-1}| (1 for the first step,
// which multiplies the caller's unnormalised start vector).
__device__ __forceinline__ double step_scale(const PowerArgs& p, int k) {
    return k == 0 ? 1.0 : 1.0 / sqrt(p.rec[4 * (k - 1)]);
}

}  // namespace pwmk
