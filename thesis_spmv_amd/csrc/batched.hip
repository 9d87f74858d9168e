// Batched power method: many instances of uneven size solved in ONE
// persistent launch (BASELINE workload "batched mixed-precision power method").
//
// Each workgroup takes instances from a global queue (largest first, one
// atomic increment per instance, no waiting) and runs the whole power method
// of that instance by itself, so no synchronisation between workgroups is
// ever needed: per iteration
//   y = (A x) / |x|   (x is the previous, not yet normalised, iterate)
//   |y|^2, x^T y, x^T x in FP64; lambda = Rayleigh quotient;
//   stop when |lambda_k - lambda_{k-1}| <= tol |lambda_k| or after max_it.
// The SpMV walks the instance's merge path (row ends against nonzeros) in
// tiles of CAP items: a tile's end is found by counting, over a coalesced
// window of row ends, the rows whose end item lies before the tile's last
// diagonal (no binary search), its products are staged in LDS, every thread
// merges its items, and a segmented scan adds the partial sums of rows that
// span threads; the partial row at the tile's end carries into the next tile
// in registers.  Long rows and empty rows need no special case.
//
// Values: FP32 or BF16 (exact values only) with FP32 vectors and per-row FP32
// accumulation, or FP64 throughout; norms and Rayleigh quotients in FP64.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "api_internal.hpp"

namespace pwmk {

struct BatchInst {
    int64_t n, nnz;
    int64_t rs_base;  // into the concatenated row_start arrays (n+1 entries each, relative offsets)
    int64_t e_base;   // into col_ind / values
    int64_t v_base;   // into the vector pool: x at v_base, y at v_base + n
};

struct BatchArgs {
    const BatchInst* inst = nullptr;
    const int32_t* order = nullptr;  // instance ids, largest first
    int32_t n_inst = 0;
    int* queue = nullptr;            // next position in `order`
    const int32_t* rs = nullptr;
    const int32_t* ci = nullptr;
    const void* val = nullptr;
    void* vec = nullptr;
    int64_t max_it = 0;
    double tol = -1.0;
    double* lambda = nullptr;        // per instance
    double* ynorm = nullptr;
    int64_t* iters = nullptr;
    int32_t* result_in_y = nullptr;  // 1: final normalised vector is at v_base + n
    int vtype = 0;
};

constexpr int kBB = 1024;  // threads per workgroup

template <typename X>
struct BatchCfg;
template <> struct BatchCfg<float> { static constexpr int ipt = 4; };
template <> struct BatchCfg<double> { static constexpr int ipt = 2; };

template <typename VT>
__global__ __launch_bounds__(kBB) void k_batch_power(BatchArgs a) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    constexpr int IPT = BatchCfg<X>::ipt;
    constexpr int CAP = kBB * IPT;
    __shared__ int32_t s_rowend[CAP];
    __shared__ Acc s_prod[CAP];
    __shared__ Acc s_out[CAP];
    __shared__ int32_t s_key[kBB];
    __shared__ Acc s_val[kBB];
    __shared__ double red[16];
    __shared__ int s_cnt[kBB / 64];
    __shared__ int s_inst;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const VT* __restrict__ val = (const VT*)a.val;

    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(a.queue, 1);
            s_inst = q < a.n_inst ? a.order[q] : -1;
        }
        __syncthreads();
        const int id = s_inst;
        __syncthreads();
        if (id < 0) return;
        const BatchInst I = a.inst[id];
        const int64_t n = I.n, nnz = I.nnz, total = n + nnz;
        const int32_t* __restrict__ rs = a.rs + I.rs_base;
        const int32_t* __restrict__ ci = a.ci + I.e_base;
        const VT* __restrict__ va = val + I.e_base;
        X* bufs[2] = {(X*)a.vec + I.v_base, (X*)a.vec + I.v_base + n};
        double scale = 1.0, lam_prev = 0.0, lam = 0.0, yy_last = 1.0;
        int64_t it = 0;
        bool done = nnz == 0 && n == 0;
        while (!done) {
            const X* __restrict__ x = bufs[it & 1];
            X* __restrict__ y = bufs[(it + 1) & 1];
            double yy = 0.0, xy = 0.0, xx = 0.0;
            int64_t d0 = 0, a0 = 0, b0 = 0;
            Acc carry = 0;
            while (d0 < total) {
                const int64_t d1 = d0 + CAP < total ? d0 + CAP : total;
                const int win = (int)(n - a0 < CAP ? n - a0 : CAP);
                // row ends of the window; count rows whose end item precedes d1
                int cnt = 0;
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int i = tid + j * kBB;
                    if (i < win) {
                        const int32_t re = rs[a0 + 1 + i];
                        s_rowend[i] = re;
                        cnt += (a0 + i + (int64_t)re < d1);
                    }
                }
                cnt = group_sum<64>(cnt);
                if (lane == 0) s_cnt[wave] = cnt;
                __syncthreads();
                int na = 0;
#pragma unroll
                for (int w = 0; w < kBB / 64; ++w) na += s_cnt[w];
                const int64_t a1 = a0 + na, b1 = d1 - a1;
                const int nb = (int)(b1 - b0);
                for (int i = tid; i < nb; i += kBB) s_prod[i] = widen(va[b0 + i]) * (Acc)x[ci[b0 + i]];
                __syncthreads();
                // thread-level merge of ITEMS items (as in k_merge_spmv)
                const int items = na + nb;
                const int t0 = tid * IPT < items ? tid * IPT : items;
                const int t1 = t0 + IPT < items ? t0 + IPT : items;
                int lo = t0 > nb ? t0 - nb : 0, hi = t0 < na ? t0 : na;
                while (lo < hi) {
                    const int piv = (lo + hi) >> 1;
                    if ((int64_t)s_rowend[piv] <= b0 + t0 - piv - 1) lo = piv + 1;
                    else hi = piv;
                }
                int ai = lo, bi = t0 - lo;
                const int a_start = ai;
                Acc run = tid == 0 ? carry : (Acc)0;  // row a0's sum from earlier tiles
                int first = -1;
                Acc first_val = 0;
                for (int d = t0; d < t1; ++d) {
                    if (ai < na && (int64_t)s_rowend[ai] <= b0 + bi) {
                        if (first < 0) {
                            first = ai;
                            first_val = run;
                        } else {
                            s_out[ai] = run;
                        }
                        run = 0;
                        ++ai;
                    } else {
                        run += s_prod[bi];
                        ++bi;
                    }
                }
                s_key[tid] = ai;
                s_val[tid] = run;
                __syncthreads();
                for (int off = 1; off < kBB; off <<= 1) {
                    Acc add = 0;
                    const bool take = tid >= off && s_key[tid - off] == s_key[tid];
                    if (take) add = s_val[tid - off];
                    __syncthreads();
                    if (take) s_val[tid] = add + s_val[tid];
                    __syncthreads();
                }
                if (first >= 0) {
                    const Acc cin = (tid > 0 && s_key[tid - 1] == a_start) ? s_val[tid - 1] : (Acc)0;
                    s_out[first] = cin + first_val;
                }
                __syncthreads();
                for (int i = tid; i < na; i += kBB) {
                    const X out = (X)((double)s_out[i] * scale);
                    y[a0 + i] = out;
                    const double yd = (double)out, xd = (double)x[a0 + i] * scale;
                    yy += yd * yd;
                    xy += xd * yd;
                    xx += xd * xd;
                }
                carry = s_val[kBB - 1];
                __syncthreads();
                d0 = d1;
                a0 = a1;
                b0 = b1;
            }
            const double YY = block_sum<kBB>(yy, red);
            const double XY = block_sum<kBB>(xy, red);
            const double XX = block_sum<kBB>(xx, red);
            lam = XY / XX;
            yy_last = YY;
            ++it;
            done = it >= a.max_it || (a.tol >= 0.0 && it >= 2 && fabs(lam - lam_prev) <= a.tol * fabs(lam));
            lam_prev = lam;
            scale = 1.0 / sqrt(YY);
        }
        // normalise the last iterate in place
        X* last = bufs[it & 1];
        const double s = 1.0 / sqrt(yy_last);
        for (int64_t i = tid; i < n; i += kBB) last[i] = (X)((double)last[i] * s);
        if (tid == 0) {
            a.lambda[id] = lam;
            a.ynorm[id] = sqrt(yy_last);
            a.iters[id] = it;
            a.result_in_y[id] = (int)(it & 1);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Start vectors: x0[i] = U[-1,1) from a hash of (seed, instance, i).
template <typename X>
__global__ void k_batch_start(const BatchInst* inst, int n_inst, X* vec, uint64_t seed) {
    const int id = blockIdx.y;
    if (id >= n_inst) return;
    const BatchInst I = inst[id];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < I.n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t h = mix64(seed * 0x9e3779b97f4a7c15ull + ((uint64_t)id << 32) + (uint64_t)i + 0x632be59bd9b4e019ull);
        vec[I.v_base + i] = (X)((double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    }
}

}  // namespace pwmk

// ---------------------------------------------------------------- host side

namespace {

inline uint64_t mix64h(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

struct pwm_batch {
    int vtype = PWM_VALUES_F32;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<pwmk::BatchInst> inst;
    // host staging of CRS instances until the first solve uploads them
    struct Staged {
        std::vector<int32_t> rs, ci;
        std::vector<double> val;
    };
    std::vector<Staged> staged;  // one per instance (empty for Poisson instances)
    std::vector<std::pair<int64_t, int64_t>> poisson;  // (m, n) of generated instances, -1 for CRS
    std::vector<int32_t> kind;                          // 0 = CRS from host, 1 = Poisson
    int64_t rs_total = 0, e_total = 0, v_total = 0;
    bool uploaded = false;
    pwm_impl::DevBuf d_inst, d_order, d_queue, d_rs, d_ci, d_val, d_vec, d_lambda, d_ynorm, d_iters, d_where;
    std::vector<double> lambda, ynorm;
    std::vector<int64_t> iters;
    std::vector<int32_t> where;
};

using pwm_impl::DevBuf;

static int bfail(pwm_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    else pwm_impl::set_global_error(m);
    return code;
}
static int bhip(pwm_batch* b, hipError_t e, const char* what) {
    return bfail(b, e == hipErrorOutOfMemory ? PWM_ERR_NOMEM : PWM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

extern "C" int pwm_batch_create(int value_type, pwm_batch** out) {
    if (!out) return bfail(nullptr, PWM_ERR_INVALID, "pwm_batch_create: out is NULL");
    *out = nullptr;
    if (value_type < PWM_VALUES_F64 || value_type > PWM_VALUES_BF16)
        return bfail(nullptr, PWM_ERR_INVALID, "pwm_batch_create: unknown value type");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return bfail(nullptr, PWM_ERR_HIP, "pwm_batch_create: no HIP device available");
    pwm_batch* b = new pwm_batch();
    b->vtype = value_type;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return bfail(nullptr, PWM_ERR_HIP, "pwm_batch_create: stream");
    }
    *out = b;
    return PWM_OK;
}

extern "C" int pwm_batch_destroy(pwm_batch* b) {
    if (!b) return PWM_OK;
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipStreamDestroy(b->stream);
    }
    delete b;
    return PWM_OK;
}

extern "C" const char* pwm_batch_last_error(const pwm_batch* b) { return b ? b->err.c_str() : pwm_last_error(nullptr); }

static void add_instance(pwm_batch* b, int64_t n, int64_t nnz) {
    pwmk::BatchInst I;
    I.n = n;
    I.nnz = nnz;
    I.rs_base = b->rs_total;
    I.e_base = b->e_total;
    I.v_base = b->v_total;
    b->rs_total += n + 1;
    b->e_total += nnz;
    b->v_total += 2 * n;
    b->inst.push_back(I);
    b->uploaded = false;
}

extern "C" int pwm_batch_add_crs(pwm_batch* b, int64_t n, const int32_t* rs, const int32_t* ci, const double* vals,
                                 int32_t* id) {
    if (!b || !rs || n <= 0 || n > INT32_MAX - 1) return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_crs: bad argument");
    const int64_t nnz = rs[n];
    if (rs[0] != 0 || nnz < 0 || (nnz > 0 && (!ci || !vals)))
        return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_crs: bad row_start or NULL arrays");
    for (int64_t r = 0; r < n; ++r)
        if (rs[r + 1] < rs[r]) return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_crs: row_start must be non-decreasing");
    for (int64_t k = 0; k < nnz; ++k) {
        if (ci[k] < 0 || ci[k] >= n) return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_crs: column index out of range");
        if (b->vtype == PWM_VALUES_BF16) {
            const float f = (float)vals[k];
            uint32_t bits;
            memcpy(&bits, &f, 4);
            if ((double)f != vals[k] || (bits & 0xffffu))
                return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_crs: BF16 storage needs values exact in BF16");
        }
    }
    if (id) *id = (int32_t)b->inst.size();
    pwm_batch::Staged st;
    st.rs.assign(rs, rs + n + 1);
    st.ci.assign(ci, ci + nnz);
    st.val.assign(vals, vals + nnz);
    b->staged.push_back(std::move(st));
    b->kind.push_back(0);
    b->poisson.push_back({-1, -1});
    add_instance(b, n, nnz);
    return PWM_OK;
}

// Poisson instances are generated on the device at upload (Util/Poisson.hpp:32-83).
extern "C" int pwm_batch_add_poisson(pwm_batch* b, int64_t m, int64_t n, int32_t* id) {
    if (!b || m <= 0 || n <= 0 || 5 * m * n > INT32_MAX) return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_poisson: bad size");
    if (id) *id = (int32_t)b->inst.size();
    const int64_t N = m * n, Z = n * (m + 2 * (m - 1)) + 2 * (n - 1) * m;
    b->kind.push_back(1);
    b->poisson.push_back({m, n});
    b->staged.emplace_back();
    add_instance(b, N, Z);
    return PWM_OK;
}

static int batch_upload(pwm_batch* b) {
    if (b->uploaded) return PWM_OK;
    const int ni = (int)b->inst.size();
    const size_t vs = pwm_impl::value_size(b->vtype), xs = pwm_impl::vector_size(b->vtype);
    hipError_t e;
    if ((e = b->d_inst.alloc(ni * sizeof(pwmk::BatchInst))) != hipSuccess || (e = b->d_order.alloc(ni * 4)) != hipSuccess ||
        (e = b->d_queue.alloc(64)) != hipSuccess || (e = b->d_rs.alloc(b->rs_total * 4)) != hipSuccess ||
        (e = b->d_ci.alloc(std::max<int64_t>(b->e_total, 1) * 4)) != hipSuccess ||
        (e = b->d_val.alloc(std::max<int64_t>(b->e_total, 1) * vs)) != hipSuccess ||
        (e = b->d_vec.alloc(std::max<int64_t>(b->v_total, 1) * xs)) != hipSuccess ||
        (e = b->d_lambda.alloc(ni * 8)) != hipSuccess || (e = b->d_ynorm.alloc(ni * 8)) != hipSuccess ||
        (e = b->d_iters.alloc(ni * 8)) != hipSuccess || (e = b->d_where.alloc(ni * 4)) != hipSuccess)
        return bhip(b, e, "batch allocation");
    // instances largest first: the queue hands them out in this order
    std::vector<int32_t> order(ni);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t p, int32_t q) { return b->inst[p].nnz + b->inst[p].n > b->inst[q].nnz + b->inst[q].n; });
    (void)hipMemcpy(b->d_inst.p, b->inst.data(), ni * sizeof(pwmk::BatchInst), hipMemcpyHostToDevice);
    (void)hipMemcpy(b->d_order.p, order.data(), ni * 4, hipMemcpyHostToDevice);
    int64_t max_stage = 0;
    for (int i = 0; i < ni; ++i) max_stage = std::max<int64_t>(max_stage, (int64_t)b->staged[i].val.size());
    DevBuf tmp;
    if (max_stage > 0 && (e = tmp.alloc(max_stage * 8)) != hipSuccess) return bhip(b, e, "batch staging");
    for (int i = 0; i < ni; ++i)
        if (b->kind[i] == 0) {
            const pwmk::BatchInst& I = b->inst[i];
            const pwm_batch::Staged& st = b->staged[i];
            e = hipMemcpy(b->d_rs.as<int32_t>() + I.rs_base, st.rs.data(), st.rs.size() * 4, hipMemcpyHostToDevice);
            if (e == hipSuccess && I.nnz) e = hipMemcpy(b->d_ci.as<int32_t>() + I.e_base, st.ci.data(), I.nnz * 4, hipMemcpyHostToDevice);
            if (e == hipSuccess && I.nnz) e = hipMemcpy(tmp.p, st.val.data(), I.nnz * 8, hipMemcpyHostToDevice);
            if (e == hipSuccess && I.nnz)
                e = pwmk::launch_convert_values(tmp.as<double>(), (char*)b->d_val.p + I.e_base * vs, I.nnz, b->vtype, b->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
            if (e != hipSuccess) return bhip(b, e, "batch upload");
        }
    for (int i = 0; i < ni; ++i)
        if (b->kind[i] == 1) {
            const pwmk::BatchInst& I = b->inst[i];
            e = pwmk::launch_fill_poisson(b->poisson[i].first, b->poisson[i].second, b->d_rs.as<int32_t>() + I.rs_base,
                                          b->d_ci.as<int32_t>() + I.e_base, (char*)b->d_val.p + I.e_base * vs, b->vtype,
                                          b->stream);
            if (e != hipSuccess) return bhip(b, e, "batch Poisson fill");
        }
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return bhip(b, e, "batch upload");
    b->uploaded = true;
    return PWM_OK;
}

template <typename VT>
static hipError_t launch_batch_t(const pwmk::BatchArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((pwmk::k_batch_power<VT>), dim3(grid), dim3(pwmk::kBB), 0, s, a);
    return hipGetLastError();
}

extern "C" int pwm_batch_solve(pwm_batch* b, int64_t max_it, double tol, uint64_t seed, double* lambdas,
                               int64_t* iterations, double* ms) {
    if (!b || max_it <= 0) return bfail(b, PWM_ERR_INVALID, "pwm_batch_solve: bad argument");
    if (b->inst.empty()) return bfail(b, PWM_ERR_STATE, "pwm_batch_solve: no instances");
    int rc = batch_upload(b);
    if (rc != PWM_OK) return rc;
    const int ni = (int)b->inst.size();
    int64_t maxn = 0;
    for (auto& I : b->inst) maxn = std::max(maxn, I.n);
    hipError_t e;
    if (b->vtype == PWM_VALUES_F64)
        hipLaunchKernelGGL(pwmk::k_batch_start<double>, dim3((unsigned)std::min<int64_t>((maxn + 255) / 256, 1024), ni), dim3(256), 0,
                           b->stream, b->d_inst.as<pwmk::BatchInst>(), ni, b->d_vec.as<double>(), seed);
    else
        hipLaunchKernelGGL(pwmk::k_batch_start<float>, dim3((unsigned)std::min<int64_t>((maxn + 255) / 256, 1024), ni), dim3(256), 0,
                           b->stream, b->d_inst.as<pwmk::BatchInst>(), ni, b->d_vec.as<float>(), seed);
    {
        pwmk::PowerArgs q;  // reset the queue head with a kernel (no memset node)
        q.ctrl = b->d_queue.as<int>();
        (void)pwmk::launch_power_reset(q, b->stream);
    }
    pwmk::BatchArgs a;
    a.inst = b->d_inst.as<pwmk::BatchInst>();
    a.order = b->d_order.as<int32_t>();
    a.n_inst = ni;
    a.queue = b->d_queue.as<int>();
    a.rs = b->d_rs.as<int32_t>();
    a.ci = b->d_ci.as<int32_t>();
    a.val = b->d_val.p;
    a.vec = b->d_vec.p;
    a.max_it = max_it;
    a.tol = tol;
    a.lambda = b->d_lambda.as<double>();
    a.ynorm = b->d_ynorm.as<double>();
    a.iters = b->d_iters.as<int64_t>();
    a.result_in_y = b->d_where.as<int32_t>();
    a.vtype = b->vtype;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = std::min(ni, 2 * cus);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, b->stream);
    switch (b->vtype) {
        case PWM_VALUES_F64: e = launch_batch_t<double>(a, grid, b->stream); break;
        case PWM_VALUES_F32: e = launch_batch_t<float>(a, grid, b->stream); break;
        default: e = launch_batch_t<pwmk::bf16_t>(a, grid, b->stream); break;
    }
    (void)hipEventRecord(e1, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return bhip(b, e, "batched power method");
    b->lambda.resize(ni);
    b->ynorm.resize(ni);
    b->iters.resize(ni);
    b->where.resize(ni);
    (void)hipMemcpy(b->lambda.data(), b->d_lambda.p, ni * 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(b->ynorm.data(), b->d_ynorm.p, ni * 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(b->iters.data(), b->d_iters.p, ni * 8, hipMemcpyDeviceToHost);
    if ((e = hipMemcpy(b->where.data(), b->d_where.p, ni * 4, hipMemcpyDeviceToHost)) != hipSuccess)
        return bhip(b, e, "batch results");
    if (lambdas) memcpy(lambdas, b->lambda.data(), ni * 8);
    if (iterations) memcpy(iterations, b->iters.data(), ni * 8);
    if (ms) *ms = t;
    return PWM_OK;
}

// The start vector the solver used for instance `id` with `seed` (host
// restatement of k_batch_start), for running the same problem elsewhere.
extern "C" int pwm_batch_start_vector(pwm_batch* b, int32_t id, uint64_t seed, double* out) {
    if (!b || !out || id < 0 || id >= (int32_t)b->inst.size()) return bfail(b, PWM_ERR_INVALID, "pwm_batch_start_vector");
    const int64_t n = b->inst[id].n;
    const bool f64 = b->vtype == PWM_VALUES_F64;
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t h = mix64h(seed * 0x9e3779b97f4a7c15ull + ((uint64_t)id << 32) + (uint64_t)i + 0x632be59bd9b4e019ull);
        const double v = (double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        out[i] = f64 ? v : (double)(float)v;
    }
    return PWM_OK;
}

// Normalised final iterate of instance `id` after pwm_batch_solve.
extern "C" int pwm_batch_vector(pwm_batch* b, int32_t id, double* out, double* ynorm) {
    if (!b || !out || id < 0 || id >= (int32_t)b->inst.size()) return bfail(b, PWM_ERR_INVALID, "pwm_batch_vector");
    if (b->where.empty()) return bfail(b, PWM_ERR_STATE, "pwm_batch_vector: solve first");
    const pwmk::BatchInst& I = b->inst[id];
    const int64_t off = I.v_base + (b->where[id] ? I.n : 0);
    hipError_t e;
    if (b->vtype == PWM_VALUES_F64) {
        e = hipMemcpy(out, b->d_vec.as<double>() + off, I.n * 8, hipMemcpyDeviceToHost);
    } else {
        std::vector<float> f(I.n);
        e = hipMemcpy(f.data(), b->d_vec.as<float>() + off, I.n * 4, hipMemcpyDeviceToHost);
        for (int64_t i = 0; i < I.n; ++i) out[i] = f[i];
    }
    if (e != hipSuccess) return bhip(b, e, "pwm_batch_vector");
    if (ynorm) *ynorm = b->ynorm[id];
    return PWM_OK;
}

extern "C" int pwm_batch_size(const pwm_batch* b, int32_t* instances, int64_t* rows, int64_t* nnz) {
    if (!b) return bfail(nullptr, PWM_ERR_INVALID, "pwm_batch_size: NULL");
    if (instances) *instances = (int32_t)b->inst.size();
    if (rows) *rows = b->v_total / 2;
    if (nnz) *nnz = b->e_total;
    return PWM_OK;
}
