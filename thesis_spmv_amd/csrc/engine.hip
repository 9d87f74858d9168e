// Iteration engine: kernels and host-side slot, list and graph management.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "api_internal.hpp"
#include "engine.hpp"
#include "merge_tile.hpp"

namespace pwmk {

static int64_t plan_step(int ipt) {
    const int64_t cap = (int64_t)kEngBS * ipt;
    return cap - cap / 4;
}

int32_t engine_tiles_for(int64_t n, int64_t nnz, int ipt) {
    const int64_t step = plan_step(ipt);
    return (int32_t)((n + nnz + step - 1) / step);
}

// searchPathOnDiag (SOA_Implementations/CRS_Merge.hpp:42-62) at diagonal d.
__device__ __forceinline__ void eng_search(const int32_t* __restrict__ rs, int64_t n, int64_t nnz, int64_t d,
                                           int32_t& a, int32_t& b) {
    int64_t lo = d > nnz ? d - nnz : 0;
    int64_t hi = d < n ? d : n;
    while (lo < hi) {
        const int64_t piv = (lo + hi) >> 1;
        if ((int64_t)rs[piv + 1] <= d - piv - 1)
            lo = piv + 1;
        else
            hi = piv;
    }
    a = (int32_t)(lo < n ? lo : n);
    b = (int32_t)(d - lo);
}

// Cut p of the merge path: the path point at diagonal min(p*step, n+nnz),
// moved back to the start of its row when that row has <= snap nonzeros.
__device__ __forceinline__ void eng_cut(const int32_t* __restrict__ rs, int64_t n, int64_t nnz, int64_t p,
                                        int64_t step, int64_t snap, int32_t& a, int32_t& b) {
    const int64_t total = n + nnz;
    eng_search(rs, n, nnz, p * step < total ? p * step : total, a, b);
    if (a < n && (int64_t)rs[a + 1] - rs[a] <= snap) b = rs[a];
}

// Tile t runs from cut t to cut t+1 (at most step + snap = 256*ipt items).
__global__ void k_eng_plan(const int32_t* __restrict__ rs, int64_t n, int64_t nnz, int64_t step, int64_t snap,
                           int32_t nt, ETile* __restrict__ tiles) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    ETile T;
    eng_cut(rs, n, nnz, t, step, snap, T.a0, T.b0);
    eng_cut(rs, n, nnz, t + 1, step, snap, T.a1, T.b1);
    int stream = T.b0 == rs[T.a0] && T.b1 == rs[T.a1];
    for (int32_t r = T.a0; stream && r < T.a1; ++r) stream = rs[r + 1] - rs[r] <= kStreamRow;
    T.stream = stream;
    T.rs_end = rs[T.a1];
    T.pad[0] = T.pad[1] = 0;
    tiles[t] = T;
}

hipError_t engine_plan_tiles(const int32_t* rs, int64_t n, int64_t nnz, int ipt, ETile* tiles, hipStream_t s) {
    const int32_t nt = engine_tiles_for(n, nnz, ipt);
    const int64_t step = plan_step(ipt), snap = (int64_t)kEngBS * ipt / 4;
    hipLaunchKernelGGL(k_eng_plan, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, rs, n, nnz, step, snap, nt,
                       tiles);
    return hipGetLastError();
}

__device__ __forceinline__ uint64_t eng_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// x0[i] = U[-1,1) from a hash of (seed, stream_id, i).
template <typename X>
__global__ void k_eng_start(X* __restrict__ x, int64_t n, uint64_t seed, uint64_t sid) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t h = eng_mix64(seed * 0x9e3779b97f4a7c15ull + (sid << 32) + (uint64_t)i + 0x632be59bd9b4e019ull);
        x[i] = (X)((double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    }
}

template <typename X>
__global__ void k_eng_scale(X* __restrict__ v, int64_t n, double s) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = (X)((double)v[i] * s);
}

static unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

hipError_t engine_start_vector(void* x, int64_t n, int vtype, uint64_t seed, uint64_t sid, hipStream_t s) {
    if (vtype == VT_F64)
        hipLaunchKernelGGL(k_eng_start<double>, dim3(grid_for(n)), dim3(256), 0, s, (double*)x, n, seed, sid);
    else
        hipLaunchKernelGGL(k_eng_start<float>, dim3(grid_for(n)), dim3(256), 0, s, (float*)x, n, seed, sid);
    return hipGetLastError();
}

hipError_t engine_normalize(void* v, int64_t n, int vtype, double inv_norm, hipStream_t s) {
    if (vtype == VT_F64)
        hipLaunchKernelGGL(k_eng_scale<double>, dim3(grid_for(n)), dim3(256), 0, s, (double*)v, n, inv_norm);
    else
        hipLaunchKernelGGL(k_eng_scale<float>, dim3(grid_for(n)), dim3(256), 0, s, (float*)v, n, inv_norm);
    return hipGetLastError();
}

// The step-dependent view of a job: x, y, scale (x and y swap every step).
template <typename X>
struct EView {
    const X* x;
    X* y;
    double scale;
    double* partials;
    void* carry;
};

template <typename X>
__device__ __forceinline__ bool eng_view(const EJob* __restrict__ jobs, const EState* __restrict__ st, int s,
                                         EView<X>& v) {
    if (s < 0 || st[s].done) return false;
    const EJob& J = jobs[s];
    const int64_t it = st[s].it;
    v.x = (const X*)((it & 1) ? J.buf1 : J.buf0);
    v.y = (X*)((it & 1) ? J.buf0 : J.buf1);
    v.scale = st[s].scale;
    v.partials = J.partials;
    v.carry = J.carry;
    return true;
}

// Stage (after tile_load), barrier, compute, carry and FP64 partials of one
// tile.  `after_stage` runs right after the staging barrier.
template <typename VT, int ITEMS, typename F>
__device__ __forceinline__ void eng_tile(const EWork& W, const EView<typename Arith<VT>::X>& v,
                                         const TileRegs<VT, ITEMS>& R,
                                         MergeTileSmem<typename Arith<VT>::A, kEngBS, ITEMS>& sm, double* red,
                                         F&& after_stage) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    const int na = W.a1 - W.a0;
    X xr[ITEMS];
    if (W.stream) stream_xrows<VT, kEngBS, ITEMS>(v.x, W.a0, na, xr);
    tile_stage<VT, kEngBS, ITEMS>(R, W.rs, v.x, W.a0, na, W.b0, W.b1, sm);
    __syncthreads();
    after_stage();
    double yy = 0.0, xy = 0.0, xx = 0.0;
    Acc c = 0;
    if (W.stream)
        stream_compute<VT, kEngBS, ITEMS, true>(v.y, xr, W.a0, W.a1, W.b0, v.scale, sm, yy, xy, xx);
    else
        c = merge_compute<VT, kEngBS, ITEMS, true>(W.rs, v.x, v.y, W.a0, W.a1, W.b0, W.b1, v.scale, sm, yy, xy, xx);
    if (threadIdx.x == kEngBS - 1) G((Acc*)v.carry)[W.tile] = c;
    block_sum3<kEngBS>(yy, xy, xx, red);  // its closing barrier also frees the LDS tile
    if (threadIdx.x == 0) {
        const auto p = G(v.partials + 3 * (int64_t)W.tile);
        p[0] = yy;
        p[1] = xy;
        p[2] = xx;
    }
}

// One workgroup per entry of the work list.  Entry b runs on XCD b % 8
// (workgroups are dispatched round-robin over the XCDs); the host laid the
// list out for that.
// The tile kernels take their read-only arrays as restrict parameters, so the
// compiler reads the uniform work entry, job and state with scalar loads.
// They are compiled for 8 waves per SIMD (256 x 8-item FP32 tiles: 62
// instead of 68 VGPRs), which keeps more tiles in flight when few jobs are
// left in a solve.
template <typename VT, int ITEMS>
__global__ __launch_bounds__(kEngBS, 8) void k_eng_tiles(const EWork* __restrict__ work, const EJob* __restrict__ jobs,
                                                      const EState* __restrict__ st, const ECtl* __restrict__ ctl) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    __shared__ MergeTileSmem<Acc, kEngBS, ITEMS> sm;
    __shared__ double red[48];
    if ((int)blockIdx.x >= ctl->n_work) return;
    const EWork W = work[blockIdx.x];
    EView<X> v;
    if (!eng_view(jobs, st, W.slot, v)) return;
    TileRegs<VT, ITEMS> R;
    tile_load<VT, kEngBS, ITEMS>(R, W.ci, (const VT*)W.val, W.b0, W.b1);
    eng_tile<VT, ITEMS>(W, v, R, sm, red, [] {});
}

// FP32 vectors, each tile's values FP32 or BF16 by its job (a branch that is
// uniform over the workgroup); the accumulator and LDS tile are FP32 either
// way.
template <int ITEMS>
__global__ __launch_bounds__(kEngBS, 8) void k_eng_tiles_mixed(const EWork* __restrict__ work,
                                                            const EJob* __restrict__ jobs,
                                                            const EState* __restrict__ st,
                                                            const ECtl* __restrict__ ctl) {
    __shared__ MergeTileSmem<float, kEngBS, ITEMS> sm;
    __shared__ double red[48];
    if ((int)blockIdx.x >= ctl->n_work) return;
    const EWork W = work[blockIdx.x];
    EView<float> v;
    if (!eng_view(jobs, st, W.slot, v)) return;
    if (W.vt16) {
        TileRegs<bf16_t, ITEMS> R;
        tile_load<bf16_t, kEngBS, ITEMS>(R, W.ci, (const bf16_t*)W.val, W.b0, W.b1);
        eng_tile<bf16_t, ITEMS>(W, v, R, sm, red, [] {});
    } else {
        TileRegs<float, ITEMS> R;
        tile_load<float, kEngBS, ITEMS>(R, W.ci, (const float*)W.val, W.b0, W.b1);
        eng_tile<float, ITEMS>(W, v, R, sm, red, [] {});
    }
}

// Persistent variant: gridDim.x (a multiple of 8, at most the resident
// capacity) workgroups; workgroup b takes entries b, b+G, b+2G, ... (all on
// its own XCD).  The next entry is read while the current tile is staged,
// and its col_ind/value loads are issued right after the staging barrier, so
// they are in flight while the current tile is computed.
template <typename VT, int ITEMS>
__global__ __launch_bounds__(kEngBS) void k_eng_tiles_p(const EWork* __restrict__ work, const EJob* __restrict__ jobs,
                                                        const EState* __restrict__ st, const ECtl* __restrict__ ctl) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    __shared__ MergeTileSmem<Acc, kEngBS, ITEMS> sm;
    __shared__ double red[48];
    const int nw = ctl->n_work;
    const int Gs = gridDim.x;
    int e = blockIdx.x;
    if (e >= nw) return;
    EWork W = work[e];
    TileRegs<VT, ITEMS> R;
    if (W.slot >= 0) tile_load<VT, kEngBS, ITEMS>(R, W.ci, (const VT*)W.val, W.b0, W.b1);
    for (;;) {
        const int en = e + Gs;
        EWork Wn;
        Wn.slot = -1;
        if (en < nw) Wn = work[en];
        EView<X> v;
        if (eng_view(jobs, st, W.slot, v)) {
            eng_tile<VT, ITEMS>(W, v, R, sm, red, [&] {
                if (Wn.slot >= 0) tile_load<VT, kEngBS, ITEMS>(R, Wn.ci, (const VT*)Wn.val, Wn.b0, Wn.b1);
            });
        } else if (Wn.slot >= 0) {
            tile_load<VT, kEngBS, ITEMS>(R, Wn.ci, (const VT*)Wn.val, Wn.b0, Wn.b1);
        }
        if (en >= nw) break;
        e = en;
        W = Wn;
    }
}

// One workgroup per active slot: carries, norms, lambda and the stop test.
// A run of tiles t..q-1 carrying into the same row r is summed in tile
// order and added to y[r] once; the tiles' partials are summed in a fixed
// thread/tile order, so the result does not depend on scheduling.
// One 1024-thread workgroup per job: a job of T tiles takes T / 4096 rounds
// of dependent loads, so the largest jobs of a step do not hold it up.
constexpr int kFinBS = 1024;

template <typename VT>
__global__ __launch_bounds__(kFinBS) void k_eng_finish(EArgs a) {
    using X = typename Arith<VT>::X;
    using Acc = typename Arith<VT>::A;
    __shared__ double red[16];
    if ((int)blockIdx.x >= a.ctl->n_slots) return;
    const int s = a.slots[blockIdx.x];
    if (a.st[s].done) return;
    const EJob& J = a.jobs[s];
    const int64_t it = a.st[s].it;
    const double scale = a.st[s].scale;
    const X* __restrict__ x = (const X*)((it & 1) ? J.buf1 : J.buf0);
    X* __restrict__ y = (X*)((it & 1) ? J.buf0 : J.buf1);
    const Acc* __restrict__ cv = (const Acc*)J.carry;
    const ETile* __restrict__ tiles = J.tiles;
    const int nt = J.n_tiles;
    double yy = 0.0, xy = 0.0, xx = 0.0;
    // Each thread takes tiles t = tid, tid + 1024, ... in order, kU at a time:
    // the loads of kU tiles (their partials, their end row and its start from
    // the plan, the carry and the next tile's bounds) are issued together,
    // then the tiles are summed in order.  Tile t ends in row r = a1; when the tiles
    // after t share that end row the row spans them, and the carries of
    // tiles t, t+1, ... (in tile order) complete it in the first tile that
    // ends past it, unless that tile starts at row r's start (then tile t's
    // carry is empty and that tile counted the row).
    constexpr int kU = 4;
    for (int t0 = threadIdx.x; t0 < nt; t0 += kU * kFinBS) {
        double p0[kU], p1[kU], p2[kU];
        int32_t r[kU], rprev[kU], nx_a1[kU], nx_b0[kU], rs_r[kU];
        Acc c[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const int t = t0 + k * kFinBS;
            if (t >= nt) break;
            p0[k] = J.partials[3 * (int64_t)t + 0];
            p1[k] = J.partials[3 * (int64_t)t + 1];
            p2[k] = J.partials[3 * (int64_t)t + 2];
            r[k] = tiles[t].a1;
            rs_r[k] = tiles[t].rs_end;
            rprev[k] = t > 0 ? tiles[t - 1].a1 : -1;
            c[k] = cv[t];
            nx_a1[k] = t + 1 < nt ? tiles[t + 1].a1 : -1;
            nx_b0[k] = t + 1 < nt ? tiles[t + 1].b0 : 0;
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const int t = t0 + k * kFinBS;
            if (t >= nt) break;
            yy += p0[k];
            xy += p1[k];
            xx += p2[k];
            const int32_t row = r[k];
            if (row >= J.n || (t > 0 && rprev[k] == row)) continue;
            Acc sum = 0;
            sum += c[k];
            int q = t + 1;
            int32_t b0 = nx_b0[k];
            if (q < nt && nx_a1[k] == row) {  // the row spans further tiles: rare
                for (; q < nt && tiles[q].a1 == row; ++q) sum += cv[q];
                b0 = q < nt ? tiles[q].b0 : 0;
            }
            if (q >= nt || rs_r[k] >= b0) continue;
            const X out = (X)((double)y[row] + (double)sum * scale);
            y[row] = out;
            const double yd = (double)out, xd = (double)x[row] * scale;
            yy += yd * yd;
            xy += xd * yd;
            xx += xd * xd;
        }
    }
    const double YY = block_sum<kFinBS>(yy, red);
    const double XY = block_sum<kFinBS>(xy, red);
    const double XX = block_sum<kFinBS>(xx, red);
    if (threadIdx.x == 0) {
        EState& S = a.st[s];
        const double lam = XY / XX;
        const int64_t it1 = it + 1;
        int done = it1 >= J.max_it;
        if (J.tol >= 0.0 && it1 >= 2 && fabs(lam - S.lambda) <= J.tol * fabs(lam)) done = 1;
        S.it = it1;
        S.lambda = lam;
        S.yy = YY;
        S.scale = 1.0 / sqrt(YY);
        S.done = done;
    }
}

template <typename VT, int ITEMS>
static hipError_t step_t(const EArgs& a, int gw, int gp, int gs, hipStream_t s) {
    if (gp > 0)
        hipLaunchKernelGGL((k_eng_tiles_p<VT, ITEMS>), dim3(gp), dim3(kEngBS), 0, s, a.work, a.jobs,
                           (const EState*)a.st, a.ctl);
    else
        hipLaunchKernelGGL((k_eng_tiles<VT, ITEMS>), dim3(gw), dim3(kEngBS), 0, s, a.work, a.jobs,
                           (const EState*)a.st, a.ctl);
    hipLaunchKernelGGL((k_eng_finish<VT>), dim3(gs), dim3(kFinBS), 0, s, a);
    return hipGetLastError();
}

template <typename VT, int ITEMS>
static int resident_t() {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_eng_tiles_p<VT, ITEMS>, kEngBS, 0) != hipSuccess)
        return 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return per_cu * cus / 8 * 8;
}

template <typename VT>
static hipError_t step_v(const EArgs& a, int ipt, int gw, int gp, int gs, hipStream_t s) {
    switch (ipt) {
        case 4: return step_t<VT, 4>(a, gw, gp, gs, s);
        case 8: return step_t<VT, 8>(a, gw, gp, gs, s);
        case 16: return step_t<VT, 16>(a, gw, gp, gs, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename VT>
static int resident_v(int ipt) {
    switch (ipt) {
        case 4: return resident_t<VT, 4>();
        case 8: return resident_t<VT, 8>();
        case 16: return resident_t<VT, 16>();
        default: return 0;
    }
}

template <int ITEMS>
static hipError_t step_mixed(const EArgs& a, int gw, int gs, hipStream_t s) {
    hipLaunchKernelGGL((k_eng_tiles_mixed<ITEMS>), dim3(gw), dim3(kEngBS), 0, s, a.work, a.jobs, (const EState*)a.st,
                       a.ctl);
    hipLaunchKernelGGL((k_eng_finish<float>), dim3(gs), dim3(kFinBS), 0, s, a);
    return hipGetLastError();
}

hipError_t engine_step(const EArgs& a, int vtype, int ipt, int gw, int gp, int gs, hipStream_t s) {
    if (vtype == VT_MIXED16) {  // no persistent variant
        switch (ipt) {
            case 4: return step_mixed<4>(a, gw, gs, s);
            case 8: return step_mixed<8>(a, gw, gs, s);
            case 16: return step_mixed<16>(a, gw, gs, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (vtype) {
        case VT_F64: return step_v<double>(a, ipt, gw, gp, gs, s);
        case VT_F32: return step_v<float>(a, ipt, gw, gp, gs, s);
        case VT_BF16: return step_v<bf16_t>(a, ipt, gw, gp, gs, s);
        default: return hipErrorInvalidValue;
    }
}

int engine_resident_blocks(int vtype, int ipt) {
    switch (vtype) {
        case VT_F64: return resident_v<double>(ipt);
        case VT_F32: return resident_v<float>(ipt);
        case VT_BF16: return resident_v<bf16_t>(ipt);
        default: return 0;
    }
}

}  // namespace pwmk

namespace pwm_impl {

// ---------------------------------------------------------------- arena

Arena::~Arena() {
    if (base_) (void)hipFree(base_);
}

hipError_t Arena::reserve(size_t bytes) {
    bytes = round(std::max<size_t>(bytes, 256));
    if (base_ && cap_ >= bytes) {
        reset();
        return hipSuccess;
    }
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    cap_ = 0;
    const hipError_t e = hipMalloc(&base_, bytes);
    if (e != hipSuccess) {
        base_ = nullptr;
        free_.clear();
        used_ = 0;
        return e;
    }
    cap_ = bytes;
    reset();
    return hipSuccess;
}

void Arena::reset() {
    free_.clear();
    if (cap_) free_[0] = cap_;
    used_ = 0;
}

bool Arena::alloc(size_t bytes, size_t* off) {
    bytes = round(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        *off = it->first;
        const size_t rest = it->second - bytes, at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used_ += bytes;
        return true;
    }
    return false;
}

void Arena::release(size_t off, size_t bytes) {
    bytes = round(std::max<size_t>(bytes, 1));
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

// ---------------------------------------------------------------- engine

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_) (void)hipGraphExecDestroy(graph_);
    for (void* p : {d_jobs_, d_state_, d_ctl_, d_work_, d_slots_})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)h_jobs_, (void*)h_state_, (void*)h_ctl_, (void*)h_work_, (void*)h_slots_})
        if (p) (void)hipHostFree(p);
}

int Engine::init(int vtype, int max_slots, hipStream_t stream, std::string& err) {
    vtype_ = vtype;
    max_slots_ = std::max(max_slots, 1);
    stream_ = stream;
    hipError_t e;
    const size_t ns = max_slots_;
    if ((e = hipHostMalloc((void**)&h_jobs_, ns * sizeof(pwmk::EJob), hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc((void**)&h_state_, ns * sizeof(pwmk::EState), hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc((void**)&h_ctl_, sizeof(pwmk::ECtl), hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc((void**)&h_slots_, ns * 4, hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc(&d_jobs_, ns * sizeof(pwmk::EJob))) != hipSuccess ||
        (e = hipMalloc(&d_state_, ns * sizeof(pwmk::EState))) != hipSuccess ||
        (e = hipMalloc(&d_ctl_, sizeof(pwmk::ECtl))) != hipSuccess || (e = hipMalloc(&d_slots_, ns * 4)) != hipSuccess) {
        err = std::string("engine allocation: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? PWM_ERR_NOMEM : PWM_ERR_HIP;
    }
    memset(h_jobs_, 0, ns * sizeof(pwmk::EJob));
    memset(h_state_, 0, ns * sizeof(pwmk::EState));
    memset(h_ctl_, 0, sizeof(pwmk::ECtl));
    in_use_.assign(ns, 0);
    h_tiles_.assign(ns, nullptr);
    running_.assign(ns, 0);
    fresh_.assign(ns, 0);
    return PWM_OK;
}

int Engine::set_ipt(int v) {
    if (v != 4 && v != 8 && v != 16) return PWM_ERR_INVALID;
    for (int s = 0; s < max_slots_; ++s)
        if (in_use_[s]) return PWM_ERR_STATE;
    if (v != ipt_ && graph_) {  // the captured steps launch the old tile kernel
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    ipt_ = v;
    return PWM_OK;
}

int Engine::add(const pwmk::EJob& j, const pwmk::ETile* host_tiles) {
    for (int s = 0; s < max_slots_; ++s) {
        if (in_use_[s]) continue;
        h_jobs_[s] = j;
        h_tiles_[s] = host_tiles;
        pwmk::EState st;
        memset(&st, 0, sizeof(st));
        st.scale = 1.0;
        h_state_[s] = st;
        in_use_[s] = 1;
        running_[s] = 1;
        fresh_[s] = 1;
        ++n_active_;
        lists_dirty_ = true;
        return s;
    }
    return -1;
}

void Engine::release(int slot) {
    if (slot < 0 || slot >= max_slots_ || !in_use_[slot]) return;
    if (running_[slot]) {
        running_[slot] = 0;
        --n_active_;
        lists_dirty_ = true;
    }
    in_use_[slot] = 0;
}

void* Engine::last_buffer(int slot) const {
    const pwmk::EJob& j = h_jobs_[slot];
    return (h_state_[slot].it & 1) ? j.buf1 : j.buf0;
}

// Work list layout: entry b runs on XCD b % 8, so XCD j's entries are
// j, j+8, j+16, ...  Every job with at least 8 tiles gives each XCD one
// contiguous eighth of its tiles (neighbouring tiles share x in that XCD's
// L2, and every XCD gets the same share of every kind of job, so no XCD
// waits on the others); smaller jobs go whole to the least loaded XCD.
// Shorter XCD sequences are padded with (-1, -1) entries.
static void xcd_sequences(const pwmk::EJob* jobs, const std::vector<char>& running, int max_slots,
                          std::vector<std::vector<int2>>& seq) {  // (slot, tile) per XCD
    constexpr int kX = 8;
    seq.assign(kX, {});
    int64_t load[kX] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < max_slots; ++s) {
        if (!running[s]) continue;
        const int32_t T = jobs[s].n_tiles;
        if (T >= kX) {
            for (int j = 0; j < kX; ++j) {
                const int32_t lo = (int32_t)((int64_t)T * j / kX), hi = (int32_t)((int64_t)T * (j + 1) / kX);
                for (int32_t t = lo; t < hi; ++t) seq[j].push_back(make_int2(s, t));
                load[j] += hi - lo;
            }
        } else {
            int j = 0;
            for (int k = 1; k < kX; ++k)
                if (load[k] < load[j]) j = k;
            for (int32_t t = 0; t < T; ++t) seq[j].push_back(make_int2(s, t));
            load[j] += T;
        }
    }
}

int Engine::build_lists(std::string& err) {
    std::vector<std::vector<int2>> seq;
    xcd_sequences(h_jobs_, running_, max_slots_, seq);
    size_t L = 0;
    for (const auto& q : seq) L = std::max(L, q.size());
    const int64_t nw = (int64_t)L * (int64_t)seq.size();
    int ns = 0;
    for (int s = 0; s < max_slots_; ++s) ns += running_[s] ? 1 : 0;
    if (nw > INT32_MAX / 2) {
        err = "engine: too many tiles in flight";
        return PWM_ERR_INVALID;
    }
    if ((size_t)nw > work_cap_alloc_) {
        // grow the work list buffers (device and pinned mirror)
        const size_t cap = std::max<size_t>((size_t)(nw + nw / 4), 4096);
        if (d_work_) (void)hipFree(d_work_);
        if (h_work_) (void)hipHostFree(h_work_);
        d_work_ = nullptr;
        h_work_ = nullptr;
        work_cap_alloc_ = 0;
        hipError_t e;
        if ((e = hipMalloc(&d_work_, cap * sizeof(pwmk::EWork))) != hipSuccess ||
            (e = hipHostMalloc((void**)&h_work_, cap * sizeof(pwmk::EWork), hipHostMallocDefault)) != hipSuccess) {
            err = std::string("engine work list: ") + hipGetErrorString(e);
            return e == hipErrorOutOfMemory ? PWM_ERR_NOMEM : PWM_ERR_HIP;
        }
        work_cap_alloc_ = cap;
        if (graph_) {
            (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
        }
    }
    const int nx = (int)seq.size();
    for (size_t i = 0; i < L; ++i)
        for (int j = 0; j < nx; ++j) {
            pwmk::EWork& W = h_work_[i * nx + j];
            memset(&W, 0, sizeof(W));
            W.slot = -1;
            if (i >= seq[j].size()) continue;
            const int s = seq[j][i].x, t = seq[j][i].y;
            const pwmk::EJob& J = h_jobs_[s];
            const pwmk::ETile& T = h_tiles_[s][t];
            W.rs = J.rs;
            W.ci = J.ci;
            W.val = J.val;
            W.slot = s;
            W.tile = t;
            W.a0 = T.a0;
            W.a1 = T.a1;
            W.b0 = T.b0;
            W.b1 = T.b1;
            W.stream = T.stream;
            W.vt16 = J.vt16;
        }
    int j = 0;
    for (int s = 0; s < max_slots_; ++s)
        if (running_[s]) h_slots_[j++] = s;
    n_work_ = (int)nw;
    n_slots_ = ns;
    n_stale_work_ = 0;
    h_ctl_->n_work = n_work_;
    h_ctl_->n_slots = n_slots_;
    // grids: resized (and the graph rebuilt) only when a list outgrows its
    // grid or falls below a quarter of it; spare blocks return at once
    const int gw = std::max(1, (int)std::min<int64_t>((int64_t)work_cap_alloc_, ((int64_t)n_work_ * 5 / 4 + 1023) / 1024 * 1024));
    if (n_work_ > grid_work_ || n_work_ < grid_work_ / 4) grid_work_ = std::max(gw, n_work_);
    const int gs = std::min(max_slots_, std::max(64, (n_slots_ * 5 / 4 + 63) / 64 * 64));
    if (n_slots_ > grid_slots_ || n_slots_ < grid_slots_ / 4) grid_slots_ = std::max(gs, n_slots_);
    hipError_t e = hipSuccess;
    if (n_work_ > 0)
        e = hipMemcpyAsync(d_work_, h_work_, (size_t)n_work_ * sizeof(pwmk::EWork), hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess && n_slots_ > 0)
        e = hipMemcpyAsync(d_slots_, h_slots_, (size_t)n_slots_ * 4, hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ctl_, h_ctl_, sizeof(pwmk::ECtl), hipMemcpyHostToDevice, stream_);
    if (e != hipSuccess) {
        err = std::string("engine list upload: ") + hipGetErrorString(e);
        return PWM_ERR_HIP;
    }
    ++list_builds;
    lists_dirty_ = false;
    return PWM_OK;
}

int Engine::run(int steps, std::vector<int>& finished, std::string& err) {
    finished.clear();
    const int rc = launch(steps, err);
    if (rc != PWM_OK) {
        if (pending_) {  // let what was queued finish before the caller goes on
            (void)hipStreamSynchronize(stream_);
            pending_ = false;
        }
        return rc;
    }
    return collect(finished, err);
}

int Engine::launch(int steps, std::string& err) {
    if (n_active_ == 0) return PWM_OK;
    // no more steps than any running job can still take (max_it bounds it)
    int64_t most = 1;
    for (int s = 0; s < max_slots_; ++s)
        if (running_[s]) most = std::max<int64_t>(most, h_jobs_[s].max_it - h_state_[s].it);
    if (steps > most) steps = (int)most;
    if (steps < 1) steps = 1;
    hipError_t e = hipSuccess;
    // new jobs: descriptor and step-0 state
    for (int s = 0; s < max_slots_ && e == hipSuccess; ++s)
        if (fresh_[s]) {
            e = hipMemcpyAsync((pwmk::EJob*)d_jobs_ + s, h_jobs_ + s, sizeof(pwmk::EJob), hipMemcpyHostToDevice, stream_);
            if (e == hipSuccess)
                e = hipMemcpyAsync((pwmk::EState*)d_state_ + s, h_state_ + s, sizeof(pwmk::EState), hipMemcpyHostToDevice,
                                   stream_);
            fresh_[s] = 0;
        }
    if (e != hipSuccess) {
        err = std::string("engine job upload: ") + hipGetErrorString(e);
        return PWM_ERR_HIP;
    }
    if (lists_dirty_) {
        const int rc = build_lists(err);
        if (rc != PWM_OK) return rc;
    }
    pwmk::EArgs a;
    a.jobs = (const pwmk::EJob*)d_jobs_;
    a.st = (pwmk::EState*)d_state_;
    a.work = (const pwmk::EWork*)d_work_;
    a.slots = (const int32_t*)d_slots_;
    a.ctl = (const pwmk::ECtl*)d_ctl_;
    // persistent tile grid: the resident capacity, at most the list length
    int gp = 0;
    if (persistent) {
        if (resident_ipt_ != ipt_) {
            resident_ = pwmk::engine_resident_blocks(vtype_, ipt_);
            resident_ipt_ = ipt_;
        }
        if (resident_ >= 8) gp = std::min(resident_, (n_work_ + 7) / 8 * 8);
    }
    if (use_graph) {
        if (graph_ && (graph_steps_ != steps || graph_gw_ != grid_work_ || graph_gs_ != grid_slots_ || graph_gp_ != gp)) {
            (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
        }
        if (!graph_) {
            hipGraph_t g = nullptr;
            e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
            hipError_t le = hipSuccess;
            for (int i = 0; i < steps && e == hipSuccess && le == hipSuccess; ++i)
                le = pwmk::engine_step(a, vtype_, ipt_, grid_work_, gp, grid_slots_, stream_);
            if (e == hipSuccess) e = hipStreamEndCapture(stream_, &g);
            if (e == hipSuccess) e = le;
            if (e == hipSuccess) e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                graph_ = nullptr;
                err = std::string("engine graph capture: ") + hipGetErrorString(e);
                return PWM_ERR_HIP;
            }
            graph_steps_ = steps;
            graph_gw_ = grid_work_;
            graph_gs_ = grid_slots_;
            graph_gp_ = gp;
            ++graph_builds;
        }
        e = hipGraphLaunch(graph_, stream_);
    } else {
        for (int i = 0; i < steps && e == hipSuccess; ++i)
            e = pwmk::engine_step(a, vtype_, ipt_, grid_work_, gp, grid_slots_, stream_);
    }
    launches += 2 * steps;
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_state_, d_state_, (size_t)max_slots_ * sizeof(pwmk::EState), hipMemcpyDeviceToHost, stream_);
    pending_ = true;  // collect() waits for the stream even after an error
    if (e != hipSuccess) {
        err = std::string("engine step: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? PWM_ERR_NOMEM : PWM_ERR_HIP;
    }
    return PWM_OK;
}

int Engine::collect(std::vector<int>& finished, std::string& err) {
    finished.clear();
    if (!pending_) return PWM_OK;
    pending_ = false;
    const hipError_t e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) {
        err = std::string("engine step: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? PWM_ERR_NOMEM : PWM_ERR_HIP;
    }
    for (int s = 0; s < max_slots_; ++s)
        if (running_[s] && h_state_[s].done) {
            finished.push_back(s);
            running_[s] = 0;
            --n_active_;
            n_stale_work_ += h_jobs_[s].n_tiles;
        }
    // drop finished jobs' tiles once they make up a tenth of the list
    if ((int64_t)n_stale_work_ * 10 > (int64_t)n_work_) lists_dirty_ = true;
    return PWM_OK;
}

}  // namespace pwm_impl
