// C ABI of libpwmhip.so (declared in include/pwm_hip.h): handle lifetime,
// matrix loading, planning, and the SpMV / power-method entry points.
#include <math.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>
#include <numeric>
#include <vector>

#include "api_internal.hpp"
#include "engine.hpp"

using namespace pwm_impl;
using pwmk::CrsDev;
using pwmk::PowerArgs;

namespace pwm_impl {

static thread_local std::string g_err;
void set_global_error(const std::string& msg) { g_err = msg; }

hipError_t DevBuf::alloc(size_t n) {
    release();
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        return e;
    }
    bytes = n;
    return hipSuccess;
}
hipError_t alloc_padded(DevBuf& b, int64_t n, size_t elem) {
    hipError_t e = b.alloc((size_t)(n + kPadEntries) * elem);
    if (e == hipSuccess) e = hipMemset((char*)b.p + (size_t)n * elem, 0, (size_t)kPadEntries * elem);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);  // before any handle stream reads it
    return e;
}
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

// ---- host thread pool behind parallel_for
namespace {
constexpr int kMaxWorkers = 16;  // the GPU host grants 16 cores per GPU
using LoopFn = std::function<void(int64_t, int64_t, int)>;

void run_chunks(std::atomic<int64_t>& next, int64_t n, int64_t chunk, const LoopFn& fn, int worker) {
    for (;;) {
        const int64_t b = next.fetch_add(chunk);
        if (b >= n) break;
        fn(b, std::min(n, b + chunk), worker);
    }
}

// one job at a time: workers 1..nworkers join the caller (worker 0)
struct HostPool {
    std::mutex run_mu;  // held by the caller whose job the pool runs
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    const LoopFn* fn = nullptr;
    int64_t n = 0, chunk = 0;
    std::atomic<int64_t> next{0};
    int nworkers = 0, active = 0, pending = 0;
    uint64_t epoch = 0;
    pid_t pid = 0;  // the process whose threads these are (a forked child has none of them)
};
HostPool* g_pool = nullptr;  // created once, never destroyed (its threads are detached)
std::once_flag g_pool_once;

void pool_worker(HostPool* p, int id) {
    uint64_t seen = 0;
    for (;;) {
        const LoopFn* fn;
        int64_t n, chunk;
        {
            std::unique_lock<std::mutex> lk(p->mu);
            p->cv_work.wait(lk, [&] { return p->epoch != seen; });
            seen = p->epoch;
            if (id > p->active) continue;  // not needed for this job
            fn = p->fn;
            n = p->n;
            chunk = p->chunk;
        }
        run_chunks(p->next, n, chunk, *fn, id);
        std::lock_guard<std::mutex> lk(p->mu);
        if (--p->pending == 0) p->cv_done.notify_all();
    }
}

void spawn_run(int64_t n, int64_t chunk, const LoopFn& fn, int nt) {
    std::atomic<int64_t> next(0);
    std::vector<std::thread> ts;
    for (int t = 1; t < nt; ++t) ts.emplace_back([&, t]() { run_chunks(next, n, chunk, fn, t); });
    run_chunks(next, n, chunk, fn, 0);
    for (auto& th : ts) th.join();
}
}  // namespace

void parallel_run(int64_t n, int64_t chunk, const LoopFn& fn) {
    const unsigned hw = std::thread::hardware_concurrency();
    const int64_t pieces = (n + chunk - 1) / chunk;
    const int nt = (int)std::min<int64_t>(std::min<int64_t>(hw ? hw : 4, kMaxWorkers), pieces);
    if (nt <= 1) {
        fn(0, n, 0);
        return;
    }
    std::call_once(g_pool_once, [&] {
        HostPool* p = new HostPool();
        p->pid = getpid();
        const int want = (int)std::min<unsigned>(hw ? hw : 4, kMaxWorkers) - 1;
        for (int i = 1; i <= want; ++i) {
            try {
                std::thread(pool_worker, p, i).detach();
                p->nworkers = i;
            } catch (...) {
                break;
            }
        }
        g_pool = p;
    });
    HostPool* p = g_pool;
    std::unique_lock<std::mutex> run(p->run_mu, std::try_to_lock);
    if (!run.owns_lock() || p->nworkers == 0 || p->pid != getpid()) {
        spawn_run(n, chunk, fn, nt);
        return;
    }
    {
        std::lock_guard<std::mutex> lk(p->mu);
        p->fn = &fn;
        p->n = n;
        p->chunk = chunk;
        p->next = 0;
        p->active = std::min(p->nworkers, nt - 1);
        p->pending = p->active;
        ++p->epoch;
    }
    p->cv_work.notify_all();
    run_chunks(p->next, n, chunk, fn, 0);
    std::unique_lock<std::mutex> lk(p->mu);
    p->cv_done.wait(lk, [&] { return p->pending == 0; });
}

// Pinned staging for host -> device uploads: two process-wide buffers of
// kStageBytes, held until the process ends; one upload at a time uses them.
namespace {
constexpr size_t kStageBytes = (size_t)64 << 20;
std::mutex g_stage_mu;
void* g_stage[2] = {nullptr, nullptr};
hipEvent_t g_stage_ev[2] = {nullptr, nullptr};
bool g_stage_failed = false;
}  // namespace

// Chunk of one staged transfer: about an eighth of it (whole MB, at least
// 2 MB, at most a buffer), so that the host copy into one buffer overlaps the
// DMA of the other from the second chunk on; host threads copy pieces of at
// least 256 KB.
static size_t stage_chunk(size_t bytes) {
    constexpr size_t MB = (size_t)1 << 20;
    const size_t c = (bytes / 8 + MB - 1) & ~(MB - 1);
    return std::min(kStageBytes, std::max(c, 2 * MB));
}
static int64_t stage_piece(size_t chunk) { return std::max<int64_t>((int64_t)(chunk / 16), (int64_t)256 << 10); }

// Allocates the staging pair once (the caller holds g_stage_mu); false when
// pinned memory cannot be had.
static bool stage_ready() {
    if (!g_stage[0] && !g_stage_failed) {
        for (int i = 0; i < 2 && !g_stage_failed; ++i)
            if (hipHostMalloc(&g_stage[i], kStageBytes, hipHostMallocDefault) != hipSuccess ||
                hipEventCreateWithFlags(&g_stage_ev[i], hipEventDisableTiming) != hipSuccess)
                g_stage_failed = true;
        if (g_stage_failed) {
            for (int i = 0; i < 2; ++i) {
                if (g_stage[i]) (void)hipHostFree(g_stage[i]);
                if (g_stage_ev[i]) (void)hipEventDestroy(g_stage_ev[i]);
                g_stage[i] = nullptr;
                g_stage_ev[i] = nullptr;
            }
            (void)hipGetLastError();
        }
    }
    return !g_stage_failed;
}

// Host -> device through the pinned pair: fill(stage, a0, a1) writes bytes
// [a0, a1) of the transfer into `stage` (called by host threads on pieces of
// a chunk); host threads fill one buffer while the other is in flight.  The
// caller holds g_stage_mu and stage_ready() is true.
template <typename Fill>
static hipError_t staged_upload(void* dst, size_t bytes, hipStream_t s, Fill fill,
                                const std::function<void()>* between) {
    hipError_t e = hipSuccess;
    bool busy[2] = {false, false};
    int b = 0;
    const size_t chunk = stage_chunk(bytes);
    const int64_t piece = stage_piece(chunk);
    for (size_t off = 0; off < bytes && e == hipSuccess; off += chunk, b ^= 1) {
        const size_t n = std::min(chunk, bytes - off);
        if (busy[b] && (e = hipEventSynchronize(g_stage_ev[b])) != hipSuccess) break;
        char* stage = (char*)g_stage[b];
        parallel_for((int64_t)n, piece, [&](int64_t a0, int64_t a1, int) { fill(stage + a0, off + a0, off + a1); });
        e = hipMemcpyAsync((char*)dst + off, stage, n, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(g_stage_ev[b], s);
        busy[b] = true;
        if (between && *between) (*between)();
    }
    // the buffers are the next upload's: wait until the copies have read them
    const hipError_t e2 = hipStreamSynchronize(s);
    return e != hipSuccess ? e : e2;
}

hipError_t upload_staged(void* dst, const void* src, size_t bytes, hipStream_t s,
                         const std::function<void()>* between) {
    if (bytes == 0) return hipSuccess;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (!stage_ready()) {  // no pinned memory: a plain copy from pageable memory
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
        return e == hipSuccess ? hipStreamSynchronize(s) : e;
    }
    const char* from = (const char*)src;
    return staged_upload(
        dst, bytes, s, [from](char* to, size_t a0, size_t a1) { memcpy(to, from + a0, a1 - a0); }, between);
}

hipError_t upload_indices_staged(void* dst, const int32_t* src, int64_t n, int64_t bound, bool* out_of_range,
                                 hipStream_t s, const std::function<void()>* between) {
    std::atomic<bool> bad{false};
    *out_of_range = false;
    if (n <= 0) return hipSuccess;
    const uint64_t ub = (uint64_t)bound;
    auto scan = [&bad, ub](const int32_t* p, int64_t m) {
        bool b = false;
        for (int64_t k = 0; k < m; ++k) b |= (uint64_t)(uint32_t)p[k] >= ub;  // negative values wrap above bound
        if (b) bad = true;
    };
    hipError_t e;
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        if (!stage_ready()) {  // no pinned memory: check, then a plain copy from pageable memory
            parallel_for(n, 1 << 20, [&](int64_t a0, int64_t a1, int) { scan(src + a0, a1 - a0); });
            if (bad) {
                *out_of_range = true;
                return hipSuccess;
            }
            e = hipMemcpyAsync(dst, src, (size_t)n * 4, hipMemcpyHostToDevice, s);
            return e == hipSuccess ? hipStreamSynchronize(s) : e;
        }
        // pieces start and end on whole indices (multiples of 64 KB)
        const char* from = (const char*)src;
        e = staged_upload(
            dst, (size_t)n * 4, s,
            [from, &scan](char* to, size_t a0, size_t a1) {
                memcpy(to, from + a0, a1 - a0);
                scan((const int32_t*)to, (int64_t)((a1 - a0) / 4));
            },
            between);
    }
    *out_of_range = bad;
    return e;
}

// FP64 value k as stored: FP32, or BF16 as the upper half of the FP32 value
// (exact for values exact in BF16, which the callers check).
static inline void narrow_value(const double* v, int vt, int64_t k, char* out) {
    const float f = (float)v[k];
    if (vt == PWM_VALUES_F32) {
        memcpy(out, &f, 4);
    } else {
        uint32_t u;
        memcpy(&u, &f, 4);
        const uint16_t h = (uint16_t)(u >> 16);
        memcpy(out, &h, 2);
    }
}

hipError_t upload_values_staged(void* dst, const double* src, int64_t n, int vt, hipStream_t s,
                                const std::function<void()>* between) {
    if (vt == PWM_VALUES_F64) return upload_staged(dst, src, (size_t)n * 8, s, between);
    if (n <= 0) return hipSuccess;
    const size_t es = vt == PWM_VALUES_F32 ? 4 : 2;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (!stage_ready()) {  // no pinned memory: narrow into host memory, then a plain copy
        std::vector<char> tmp((size_t)n * es);
        char* o = tmp.data();
        parallel_for(n, 1 << 20, [&](int64_t a0, int64_t a1, int) {
            for (int64_t k = a0; k < a1; ++k) narrow_value(src, vt, k, o + k * es);
        });
        hipError_t e = hipMemcpyAsync(dst, o, tmp.size(), hipMemcpyHostToDevice, s);
        return e == hipSuccess ? hipStreamSynchronize(s) : e;
    }
    // pieces of a chunk start and end on whole values (chunks and pieces are
    // multiples of 256 KB, and every range below is split at multiples of es)
    return staged_upload(
        dst, (size_t)n * es, s,
        [src, vt, es](char* to, size_t a0, size_t a1) {
            for (size_t k = a0 / es; k < a1 / es; ++k) narrow_value(src, vt, (int64_t)k, to + (k - a0 / es) * es);
        },
        between);
}

hipError_t download_staged(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (!stage_ready()) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
        return e == hipSuccess ? hipStreamSynchronize(s) : e;
    }
    // the copy of chunk k+1 into one buffer runs while host threads empty
    // chunk k from the other
    const size_t chunk = stage_chunk(bytes);
    const int64_t piece = stage_piece(chunk);
    const size_t nch = (bytes + chunk - 1) / chunk;
    auto issue = [&](size_t k) {
        const size_t off = k * chunk, n = std::min(chunk, bytes - off);
        hipError_t e = hipMemcpyAsync(g_stage[k & 1], (const char*)src + off, n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(g_stage_ev[k & 1], s);
        return e;
    };
    hipError_t e = issue(0);
    for (size_t k = 0; k < nch && e == hipSuccess; ++k) {
        if (k + 1 < nch && (e = issue(k + 1)) != hipSuccess) break;
        if ((e = hipEventSynchronize(g_stage_ev[k & 1])) != hipSuccess) break;
        const size_t off = k * chunk, n = std::min(chunk, bytes - off);
        const char* stage = (const char*)g_stage[k & 1];
        char* to = (char*)dst + off;
        parallel_for((int64_t)n, piece,
                     [&](int64_t a0, int64_t a1, int) { memcpy(to + a0, stage + a0, (size_t)(a1 - a0)); });
    }
    const hipError_t e2 = hipStreamSynchronize(s);
    return e != hipSuccess ? e : e2;
}

int fail(pwm_matrix* h, int code, const std::string& msg) {
    if (h)
        h->err = msg;
    else
        g_err = msg;
    return code;
}
int hip_fail(pwm_matrix* h, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(h, e == hipErrorOutOfMemory ? PWM_ERR_NOMEM : PWM_ERR_HIP, m);
}
size_t value_size(int vtype) { return vtype == PWM_VALUES_F64 ? 8 : (vtype == PWM_VALUES_F32 ? 4 : 2); }

bool exact_in_bf16(const double* v, int64_t n) {
    std::atomic<bool> ok{true};
    parallel_for(n, 1 << 20, [&](int64_t b, int64_t e, int) {
        for (int64_t k = b; k < e && ok.load(std::memory_order_relaxed); ++k) {
            const float f = (float)v[k];
            uint32_t bits;
            memcpy(&bits, &f, 4);
            if ((double)f != v[k] || (bits & 0xffffu)) ok = false;
        }
    });
    return ok;
}
size_t vector_size(int vtype) { return vtype == PWM_VALUES_F64 ? 8 : 4; }

static CrsDev crs_of(const pwm_matrix* h) {
    CrsDev A;
    A.nrows = h->nor;
    A.ncols = h->noc;
    A.nnz = h->nnz;
    A.rs = h->rs.as<int32_t>();
    A.ci = h->ci.as<int32_t>();
    A.val = h->val.p;
    A.vtype = h->vtype;
    A.nt = h->crs_nt_used;
    return A;
}

static int pow2_ceil(int64_t v) {
    int g = 1;
    while (g < v && g < 64) g <<= 1;
    return g;
}

// CRS tiles kernel: the engine's row-aligned tile plan, kept as merge-path
// splits (split_a/split_b = each tile's first row / nonzero) for the fix-up.
static int plan_crs_tiles(pwm_matrix* h) {
    const int ipt = pwmk::tiles_items_per_thread();
    const int32_t nt = pwmk::engine_tiles_for(h->nor, h->nnz, ipt);
    hipError_t e;
    if ((e = h->tiles.alloc((size_t)nt * sizeof(pwmk::ETile))) != hipSuccess ||
        (e = h->split_a.alloc((nt + 1) * 8)) != hipSuccess || (e = h->split_b.alloc((nt + 1) * 8)) != hipSuccess ||
        (e = h->carry_row.alloc(nt * 4)) != hipSuccess || (e = h->carry_val.alloc(nt * 8)) != hipSuccess)
        return hip_fail(h, e, "tiles plan allocation");
    e = pwmk::engine_plan_tiles(h->rs.as<int32_t>(), h->nor, h->nnz, ipt, h->tiles.as<pwmk::ETile>(), h->stream);
    std::vector<pwmk::ETile> t(nt);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t.data(), h->tiles.p, (size_t)nt * sizeof(pwmk::ETile), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "tiles plan");
    std::vector<int64_t> sa(nt + 1), sb(nt + 1);
    h->tiles_carry = false;
    for (int32_t i = 0; i < nt; ++i) {
        sa[i] = t[i].a0;
        sb[i] = t[i].b0;
        if (!t[i].stream) h->tiles_carry = true;
    }
    sa[nt] = h->nor;
    sb[nt] = h->nnz;
    if ((e = hipMemcpy(h->split_a.p, sa.data(), (nt + 1) * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(h->split_b.p, sb.data(), (nt + 1) * 8, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(h, e, "tiles plan upload");
    h->mp = pwmk::MergePlan{};
    h->mp.P = nt;
    h->mp.ipt = ipt;
    h->mp.items_per_tile = 0;
    h->mp.split_a = h->split_a.as<int64_t>();
    h->mp.split_b = h->split_b.as<int64_t>();
    h->mp.carry_row = h->carry_row.as<int32_t>();
    h->mp.carry_val = h->carry_val.p;
    return PWM_OK;
}

static hipError_t crs_kernel_launch(pwm_matrix* h, int which, const void* x, void* y, const PowerArgs* pw) {
    const CrsDev A = crs_of(h);
    hipError_t e;
    if (which == 2) {
        // a plain SpMV whose tiles are all row-aligned stream tiles has no
        // carries to add; power steps always finish in the fix-up (norms)
        e = pwmk::launch_tiles_spmv(A, h->tiles.as<pwmk::ETile>(), h->mp, x, y, pw, h->stream);
        if (e == hipSuccess && (pw || h->tiles_carry)) e = pwmk::launch_merge_fixup(A, h->mp, y, pw, h->stream);
        return e;
    }
    e = pwmk::launch_crs_vector(A, h->group, h->vec_blocks, h->lr, x, y, pw, h->stream);
    if (e == hipSuccess && (pw || h->lr.n_rows > 0)) e = pwmk::launch_crs_finalize(A, h->vec_blocks, h->lr, y, pw, h->stream);
    return e;
}

// Average time of 5 SpMVs (after 2 warm-up ones) of CRS kernel `which` on
// the device vectors x, y.
static hipError_t time_crs_kernel(pwm_matrix* h, int which, const void* x, void* y, float* ms) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = crs_kernel_launch(h, which, x, y, nullptr);
    (void)hipEventRecord(a, h->stream);
    for (int i = 0; i < 5 && e == hipSuccess; ++i) e = crs_kernel_launch(h, which, x, y, nullptr);
    (void)hipEventRecord(b, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    *ms = 0.0f;
    (void)hipEventElapsedTime(ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return e;
}

// CRS vector kernel with G lanes per row: grid size and the long-row chunk
// plan (rows longer than max(1024, 128 G) are cut into chunks).
static int plan_crs_vector(pwm_matrix* h, int G) {
    h->group = G;
    const int64_t rpw = 64 / G;
    const int64_t units = (h->nor + rpw - 1) / rpw;
    int64_t blocks = h->vec_blocks_opt > 0 ? h->vec_blocks_opt : std::min<int64_t>((units + 3) / 4, 2048);
    h->vec_blocks = (int)std::max<int64_t>(blocks, 1);

    int64_t thr = h->long_row > 0 ? h->long_row : std::max<int64_t>(1024, 128 * G);
    h->lr = pwmk::LongRows{};
    h->lr.threshold = (int32_t)std::min<int64_t>(thr, INT32_MAX);
    std::vector<pwmk::LongChunk> chunks;
    std::vector<int32_t> ids, begin;
    // the row offsets: the host copy when the load kept one, else read back
    // from the device (every row longer than the threshold must be planned,
    // since the vector kernel skips them)
    std::vector<int32_t> rs_dev;
    const int32_t* rsh = h->h_rs.empty() ? nullptr : h->h_rs.data();
    if (!rsh && h->nor > 0) {
        rs_dev.resize(h->nor + 1);
        hipError_t e = hipMemcpy(rs_dev.data(), h->rs.p, (h->nor + 1) * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(h, e, "long-row plan: row offsets");
        rsh = rs_dev.data();
    }
    if (rsh) {
        for (int64_t r = 0; r < h->nor; ++r) {
            const int64_t s = rsh[r], e = rsh[r + 1];
            if (e - s <= thr) continue;
            begin.push_back((int32_t)chunks.size());
            for (int64_t k = s; k < e; k += h->chunk) {
                pwmk::LongChunk c;
                c.row = (int32_t)r;
                c.slot = (int32_t)ids.size();
                c.k0 = (int32_t)k;
                c.k1 = (int32_t)std::min<int64_t>(e, k + h->chunk);
                chunks.push_back(c);
            }
            ids.push_back((int32_t)r);
        }
    }
    begin.push_back((int32_t)chunks.size());
    h->lr.n_rows = (int32_t)ids.size();
    h->lr.n_chunks = (int32_t)chunks.size();
    if (!ids.empty()) {
        hipError_t e;
        if ((e = h->lr_chunks.alloc(chunks.size() * sizeof(pwmk::LongChunk))) != hipSuccess ||
            (e = h->lr_rows.alloc(ids.size() * 4)) != hipSuccess ||
            (e = h->lr_begin.alloc(begin.size() * 4)) != hipSuccess ||
            (e = h->lr_partial.alloc(chunks.size() * 8)) != hipSuccess)
            return hip_fail(h, e, "long-row plan allocation");
        if ((e = hipMemcpy(h->lr_chunks.p, chunks.data(), chunks.size() * sizeof(pwmk::LongChunk),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->lr_rows.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->lr_begin.p, begin.data(), begin.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(h, e, "long-row plan upload");
        h->lr.chunks = h->lr_chunks.as<pwmk::LongChunk>();
        h->lr.row_ids = h->lr_rows.as<int32_t>();
        h->lr.chunk_begin = h->lr_begin.as<int32_t>();
        h->lr.chunk_partial = h->lr_partial.p;
    }
    return PWM_OK;
}

// CRS planning and the autotuner.  Candidates: the vector kernel with the
// rows-per-wave heuristic G0 = pow2(avg row length / 2) in [2, 32] and with
// G0/2 and 2 G0 when the longest row fills them (only G0 when the caller
// fixed group_size), and the tiles kernel (unless crs_kernel excludes
// either).  With more than one candidate,
// each is timed on x = 1 (2 warm-up + 5 timed SpMVs) and the fastest is kept.
int plan_crs(pwm_matrix* h) {
    const double avg = h->nor > 0 ? (double)h->nnz / (double)h->nor : 0.0;
    const int G0 = h->group_opt > 0 ? h->group_opt : std::min(32, std::max(2, pow2_ceil((int64_t)ceil(avg / 2.0))));
    // more lanes per row than the longest row has entries leave lanes idle
    // in every row: such candidates (other than G0) are not timed
    std::vector<int> gs;
    if (h->crs_kernel != 2) {
        if (h->group_opt > 0) {
            gs.push_back(G0);
        } else {
            for (int g : {G0 / 2, G0, 2 * G0})
                if (g >= 1 && g <= 64 && (g == G0 || g <= std::max<int64_t>(1, h->stats.row_max))) gs.push_back(g);
        }
    }
    const bool tiles = h->crs_kernel != 1;
    int rc;
    if (tiles && (rc = plan_crs_tiles(h)) != PWM_OK) return rc;
    h->crs_nt_used = h->crs_nt > 0 ? 1 : 0;
    if ((int)gs.size() + (tiles ? 1 : 0) == 1 && (!tiles || h->crs_nt >= 0)) {
        h->crs_used = tiles ? 2 : 1;
        return plan_crs_vector(h, tiles ? G0 : gs[0]);
    }
    const size_t vs = vector_size(h->vtype);
    DevBuf x, y;
    hipError_t e;
    if ((e = x.alloc(std::max<int64_t>(h->noc, 1) * vs)) != hipSuccess ||
        (e = y.alloc(std::max<int64_t>(h->nor, 1) * vs)) != hipSuccess)
        return hip_fail(h, e, "CRS autotune vectors");
    e = pwmk::launch_fill(x.p, h->noc, h->vtype == PWM_VALUES_F64 ? 1 : 0, 1.0, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "CRS autotune vectors");
    float best = -1.0f;
    int best_kernel = 1, best_g = G0, best_nt = 0;
    if (tiles) {
        // entry loads plain or non-temporal: the caller's choice, else both
        for (int nt : {0, 1}) {
            if (h->crs_nt >= 0 && nt != h->crs_nt) continue;
            h->crs_nt_used = nt;
            float ms = 0.0f;
            if ((e = time_crs_kernel(h, 2, x.p, y.p, &ms)) != hipSuccess) return hip_fail(h, e, "CRS autotune");
            if (best < 0.0f || ms < best) {
                best = ms;
                best_kernel = 2;
                best_nt = nt;
            }
        }
    }
    h->crs_nt_used = best_nt;
    for (int g : gs) {
        if ((rc = plan_crs_vector(h, g)) != PWM_OK) return rc;
        float ms = 0.0f;
        if ((e = time_crs_kernel(h, 1, x.p, y.p, &ms)) != hipSuccess) return hip_fail(h, e, "CRS autotune");
        if (best < 0.0f || ms < best) {
            best = ms;
            best_kernel = 1;
            best_g = g;
        }
    }
    h->crs_used = best_kernel;
    return plan_crs_vector(h, best_kernel == 1 ? best_g : G0);
}

// Merge path: P tiles of at most 256*ipt merge items (the threads count
// given to the constructor is used when its tiles fit that capacity).
int plan_merge(pwm_matrix* h) {
    const int64_t total = h->nor + h->nnz;
    const int64_t cap = 256 * (int64_t)h->ipt;
    int64_t P = (total + cap - 1) / cap;
    if (P < 1) P = 1;
    if (h->threads > 0) {
        const int64_t want = h->threads;
        const int64_t items = (total + want - 1) / want;
        if (items <= cap) P = want;
    }
    h->mp = pwmk::MergePlan{};
    h->mp.P = P;
    h->mp.items_per_tile = (total + P - 1) / P;
    h->mp.ipt = h->ipt;
    hipError_t e;
    if ((e = h->split_a.alloc((P + 1) * 8)) != hipSuccess || (e = h->split_b.alloc((P + 1) * 8)) != hipSuccess ||
        (e = h->carry_row.alloc(P * 4)) != hipSuccess || (e = h->carry_val.alloc(P * 8)) != hipSuccess)
        return hip_fail(h, e, "merge plan allocation");
    h->mp.split_a = h->split_a.as<int64_t>();
    h->mp.split_b = h->split_b.as<int64_t>();
    h->mp.carry_row = h->carry_row.as<int32_t>();
    h->mp.carry_val = h->carry_val.p;
    e = pwmk::launch_merge_splits(crs_of(h), h->mp, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "merge split search");
    return PWM_OK;
}

int spmv_launch(pwm_matrix* h, const void* x, void* y, const PowerArgs* pw, int* n_partials_out) {
    const CrsDev A = crs_of(h);
    hipError_t e = hipSuccess;
    int np = 0;
    switch (h->format) {
        case PWM_FORMAT_CRS:
            e = crs_kernel_launch(h, h->crs_used, x, y, pw);
            np = h->crs_used == 2 ? (int)h->mp.P : h->vec_blocks;
            break;
        case PWM_FORMAT_CRS_MERGE:
            e = pwmk::launch_merge_spmv(A, h->mp, x, y, pw, h->stream);
            if (e == hipSuccess) e = pwmk::launch_merge_fixup(A, h->mp, y, pw, h->stream);
            np = (int)h->mp.P;
            break;
        case PWM_FORMAT_CSB:
            e = csb_launch(h, x, y, pw);
            np = (int)csb_partials(h);
            break;
        case PWM_FORMAT_BLOCKCOH:
            e = coh_launch(h, x, y, pw);
            np = (int)coh_partials(h);
            break;
        default:
            return fail(h, PWM_ERR_INVALID, "unknown format");
    }
    if (n_partials_out) *n_partials_out = np;
    if (e != hipSuccess) return hip_fail(h, e, "SpMV launch");
    return PWM_OK;
}

static int64_t partials_needed(const pwm_matrix* h) {
    switch (h->format) {
        case PWM_FORMAT_CRS:
            return std::max<int64_t>(h->vec_blocks, h->tiles.p ? h->mp.P + pwmk::merge_fix_blocks(h->mp.P) : 0);
        case PWM_FORMAT_CRS_MERGE: return h->mp.P + pwmk::merge_fix_blocks(h->mp.P);
        case PWM_FORMAT_CSB: return csb_partials(h);
        case PWM_FORMAT_BLOCKCOH: return coh_partials(h);
    }
    return 0;
}

}  // namespace pwm_impl

// ---------------------------------------------------------------- lifetime

static int noalloc_scratch(pwm_matrix* h, char** out);

extern "C" int pwm_create(int format, int threads, int value_type, pwm_matrix** out) {
    if (!out) return fail(nullptr, PWM_ERR_INVALID, "pwm_create: out is NULL");
    *out = nullptr;
    if (format != PWM_FORMAT_CRS && format != PWM_FORMAT_CRS_MERGE && format != PWM_FORMAT_CSB &&
        format != PWM_FORMAT_BLOCKCOH)
        return fail(nullptr, PWM_ERR_INVALID, "pwm_create: unknown format id " + std::to_string(format));
    if (value_type < PWM_VALUES_F64 || value_type > PWM_VALUES_BF16)
        return fail(nullptr, PWM_ERR_INVALID, "pwm_create: unknown value type");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1)
        return fail(nullptr, PWM_ERR_HIP, "pwm_create: no HIP device available");
    pwm_matrix* h = new pwm_matrix();
    h->format = format;
    h->threads = threads;
    h->vtype = value_type;
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(nullptr, e, "pwm_create: stream");
    }
    // set-up of the no-allocation power method: the per-GPU workspace array
    // is made resident on this device and its address looked up, once per
    // device, so that pwm_power_method_device_noalloc itself allocates nothing
    char* ws = nullptr;
    if (int rc = noalloc_scratch(nullptr, &ws)) {
        (void)hipStreamDestroy(h->stream);
        delete h;
        return rc;
    }
    *out = h;
    return PWM_OK;
}

extern "C" int pwm_destroy(pwm_matrix* h) {
    if (!h) return PWM_OK;
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
    }
    delete h;
    return PWM_OK;
}

extern "C" const char* pwm_last_error(const pwm_matrix* h) { return h ? h->err.c_str() : g_err.c_str(); }

extern "C" int pwm_set_option(pwm_matrix* h, const char* key, int64_t value) {
    if (!h || !key) return fail(h, PWM_ERR_INVALID, "pwm_set_option: NULL argument");
    const std::string k(key);
    if (k == "group_size") {
        if (value != 0 && (value < 1 || value > 64 || (value & (value - 1))))
            return fail(h, PWM_ERR_INVALID, "group_size must be 0 or a power of two in 1..64");
        h->group_opt = (int)value;
    } else if (k == "vec_blocks") {
        h->vec_blocks_opt = (int)std::max<int64_t>(0, value);
    } else if (k == "long_row") {
        h->long_row = std::max<int64_t>(0, value);
    } else if (k == "long_chunk") {
        if (value < 256) return fail(h, PWM_ERR_INVALID, "long_chunk must be >= 256");
        h->chunk = (int)value;
    } else if (k == "items_per_thread") {
        if (value != 4 && value != 8 && value != 12 && value != 16)
            return fail(h, PWM_ERR_INVALID, "items_per_thread must be 4, 8, 12 or 16");
        h->ipt = (int)value;
    } else if (k == "block_bits") {
        if (value < 0 || value > 16) return fail(h, PWM_ERR_INVALID, "block_bits must be 0..16");
        h->csb_bits_override = (int32_t)value;
    } else if (k == "csb_groups") {
        if (value < 0 || value > 64) return fail(h, PWM_ERR_INVALID, "csb_groups must be 0..64");
        h->csb_groups = (int32_t)value;
    } else if (k == "csb_layout") {
        if (value < 0 || value > 4)
            return fail(h, PWM_ERR_INVALID,
                        "csb_layout must be 0 (automatic), 1 (slab pieces), 2 (slab streams), 3 (wide slab streams) "
                        "or 4 (swept slabs)");
        h->csb_layout = (int32_t)value;
    } else if (k == "csb_window_bits") {
        if (value != 0 && (value < 10 || value > 24))
            return fail(h, PWM_ERR_INVALID, "csb_window_bits must be 0 (1 MB of x) or in [10, 24]");
        h->csb_window_bits = (int32_t)value;
    } else if (k == "csb_sync_every") {
        if (value != 1 && value != 2 && value != 4 && value != 8)
            return fail(h, PWM_ERR_INVALID, "csb_sync_every must be 1, 2, 4 or 8");
        h->csb_sync_every = (int32_t)value;
    } else if (k == "csb_sweep_unroll") {
        if (value != 1 && value != 2 && value != 4) return fail(h, PWM_ERR_INVALID, "csb_sweep_unroll must be 1, 2 or 4");
        h->csb_sweep_unroll = (int32_t)value;
    } else if (k == "csb_sweep_order") {
        if (value < 0 || value > 1) return fail(h, PWM_ERR_INVALID, "csb_sweep_order must be 0 (rows) or 1 (columns)");
        h->csb_sweep_order = (int32_t)value;
    } else if (k == "csb_sweep_waves") {
        if (value < 1 || value > 8) return fail(h, PWM_ERR_INVALID, "csb_sweep_waves must be 1..8");
        h->csb_sweep_waves = (int32_t)value;
    } else if (k == "coh_occ8") {
        if (value < -1 || value > 1) return fail(h, PWM_ERR_INVALID, "coh_occ8 must be -1 (timed), 0 or 1");
        h->coh_occ8 = (int32_t)value;
    } else if (k == "csb_unroll") {
        if (value != 0 && value != 2 && value != 4) return fail(h, PWM_ERR_INVALID, "csb_unroll must be 0, 2 or 4");
        h->csb_unroll = (int32_t)value;
    } else if (k == "crs_nt") {
        if (value < -1 || value > 1) return fail(h, PWM_ERR_INVALID, "crs_nt must be -1 (timed), 0 or 1");
        h->crs_nt = (int)value;
    } else if (k == "csb_nt") {
        if (value < -1 || value > 1) return fail(h, PWM_ERR_INVALID, "csb_nt must be -1 (timed), 0 or 1");
        h->csb_nt = (int32_t)value;
    } else if (k == "csb_build") {
        if (value < 0 || value > 1) return fail(h, PWM_ERR_INVALID, "csb_build must be 0 (device) or 1 (host)");
        h->csb_build = (int32_t)value;
    } else if (k == "coh_build") {
        if (value < 0 || value > 1) return fail(h, PWM_ERR_INVALID, "coh_build must be 0 (device) or 1 (host)");
        h->coh_build = (int32_t)value;
    } else if (k == "crs_kernel") {
        if (value < 0 || value > 2) return fail(h, PWM_ERR_INVALID, "crs_kernel must be 0 (time both), 1 (vector) or 2 (tiles)");
        h->crs_kernel = (int)value;
    } else {
        return fail(h, PWM_ERR_UNSUPPORTED, "unknown option " + k);
    }
    if (h->loaded && (k == "group_size" || k == "vec_blocks" || k == "long_row" || k == "long_chunk" || k == "crs_kernel" ||
                      k == "crs_nt") &&
        h->format == PWM_FORMAT_CRS)
        return plan_crs(h);
    if (h->loaded && k == "items_per_thread" && h->format == PWM_FORMAT_CRS_MERGE) return plan_merge(h);
    return PWM_OK;
}

// ---------------------------------------------------------------- loading

static int upload_values(pwm_matrix* h, const double* vals, int64_t nnz) {
    hipError_t e = alloc_padded(h->val, nnz, value_size(h->vtype));
    if (e != hipSuccess) return hip_fail(h, e, "value allocation");
    if (nnz == 0) return PWM_OK;
    if (h->vtype == PWM_VALUES_BF16 && !exact_in_bf16(vals, nnz)) {
        for (int64_t i = 0; i < nnz; ++i) {  // name the first value that is not
            const float f = (float)vals[i];
            uint32_t b;
            memcpy(&b, &f, 4);
            if ((double)f != vals[i] || (b & 0xffffu))
                return fail(h, PWM_ERR_INVALID, "BF16 storage needs values exact in BF16 (entry " +
                                                    std::to_string(i) + ")");
        }
    }
    e = upload_values_staged(h->val.p, vals, nnz, h->vtype, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "value upload");
    return PWM_OK;
}

static void reset_matrix(pwm_matrix* h) {
    h->loaded = false;
    h->rs.release();
    h->ci.release();
    h->val.release();
    h->h_rs.clear();
    h->lr = pwmk::LongRows{};
    h->mp = pwmk::MergePlan{};
    h->tiles.release();
    h->crs_used = 1;
    h->csb.clear();
    h->coh.clear();
}

static int finish_load(pwm_matrix* h, const int32_t* hrs, const int32_t* hci, const double* hval) {
    int rc = compute_matrix_stats(h);
    if (rc != PWM_OK) return rc;
    switch (h->format) {
        case PWM_FORMAT_CRS: rc = plan_crs(h); break;
        case PWM_FORMAT_CRS_MERGE: rc = plan_merge(h); break;
        case PWM_FORMAT_CSB: rc = build_csb(h, hrs, hci, hval); break;
        case PWM_FORMAT_BLOCKCOH: rc = build_blockcoh(h, hrs, hci, hval); break;
    }
    if (rc == PWM_OK) h->loaded = true;
    return rc;
}

// Validated host CRS -> device.
static int load_crs_host(pwm_matrix* h, int64_t nor, int64_t noc, const int32_t* rs, const int32_t* ci,
                         const double* vals) {
    const int64_t nnz = rs[nor];
    h->nor = nor;
    h->noc = noc;
    h->nnz = nnz;
    hipError_t e;
    if ((e = h->rs.alloc((nor + 1) * 4)) != hipSuccess || (e = alloc_padded(h->ci, nnz, 4)) != hipSuccess)
        return hip_fail(h, e, "CRS allocation");
    if ((e = upload_staged(h->rs.p, rs, (nor + 1) * 4, h->stream)) != hipSuccess ||
        (e = upload_staged(h->ci.p, ci, nnz * 4, h->stream)) != hipSuccess)
        return hip_fail(h, e, "CRS upload");
    int rc = upload_values(h, vals, nnz);
    if (rc != PWM_OK) return rc;
    h->h_rs.assign(rs, rs + nor + 1);
    return finish_load(h, rs, ci, vals);
}

extern "C" int pwm_load_crs(pwm_matrix* h, int64_t nor, int64_t noc, const int32_t* rs, const int32_t* ci,
                            const double* vals, int partitions) {
    (void)partitions;
    if (!h) return fail(nullptr, PWM_ERR_INVALID, "pwm_load_crs: NULL handle");
    if (nor <= 0 || noc <= 0 || nor > INT32_MAX - 1 || noc > INT32_MAX || !rs)
        return fail(h, PWM_ERR_INVALID, "pwm_load_crs: need 0 < nor, noc < 2^31 and row_start");
    if (rs[0] != 0) return fail(h, PWM_ERR_INVALID, "pwm_load_crs: row_start[0] must be 0");
    for (int64_t r = 0; r < nor; ++r)
        if (rs[r + 1] < rs[r]) return fail(h, PWM_ERR_INVALID, "pwm_load_crs: row_start must be non-decreasing");
    const int64_t nnz = rs[nor];
    if (nnz > PWM_MAX_NNZ) return fail(h, PWM_ERR_INVALID, "pwm_load_crs: more than PWM_MAX_NNZ nonzeros");
    if (nnz > 0 && (!ci || !vals)) return fail(h, PWM_ERR_INVALID, "pwm_load_crs: NULL col_ind/values");
    std::atomic<bool> bad{false};
    parallel_for(nnz, 1 << 20, [&](int64_t k0, int64_t k1, int) {
        bool b = false;
        for (int64_t k = k0; k < k1; ++k) b |= (ci[k] < 0) | (ci[k] >= noc);
        if (b) bad = true;
    });
    if (bad) return fail(h, PWM_ERR_INVALID, "pwm_load_crs: column index out of range");
    reset_matrix(h);
    return load_crs_host(h, nor, noc, rs, ci, vals);
}

// COO -> CRS (Util/TripletToCRS.hpp:135-172): rows ascending, columns
// ascending inside a row, duplicates kept in input order.  The indices are
// validated on the host (in parallel); the sort runs on the device
// (coo_to_crs_device).  A host-built BlockCOH then takes the sorted arrays
// back; the other builds need only row_start on the host.
extern "C" int pwm_load_triplets(pwm_matrix* h, int64_t nor, int64_t noc, int64_t nnz, const int32_t* rows,
                                 const int32_t* cols, const double* vals, int partitions) {
    (void)partitions;
    if (!h) return fail(nullptr, PWM_ERR_INVALID, "pwm_load_triplets: NULL handle");
    if (nor <= 0 || noc <= 0 || nnz < 0 || nor > INT32_MAX - 1 || noc > INT32_MAX || nnz > PWM_MAX_NNZ)
        return fail(h, PWM_ERR_INVALID, "pwm_load_triplets: need 0 < nor, noc < 2^31 and 0 <= nnz <= PWM_MAX_NNZ");
    if (nnz > 0 && (!rows || !cols || !vals)) return fail(h, PWM_ERR_INVALID, "pwm_load_triplets: NULL arrays");
    // first out-of-range entry and first value not exact in BF16 (when BF16 storage)
    std::atomic<int64_t> bad_idx{INT64_MAX}, bad_val{INT64_MAX};
    const bool bf16 = h->vtype == PWM_VALUES_BF16;
    parallel_for(nnz, 1 << 16, [&](int64_t i0, int64_t i1, int) {
        for (int64_t i = i0; i < i1; ++i) {
            if (rows[i] < 0 || rows[i] >= nor || cols[i] < 0 || cols[i] >= noc) {
                int64_t cur = bad_idx.load();
                while (i < cur && !bad_idx.compare_exchange_weak(cur, i)) {
                }
                break;
            }
        }
        if (!bf16) return;
        for (int64_t i = i0; i < i1; ++i) {
            const float f = (float)vals[i];
            uint32_t b;
            memcpy(&b, &f, 4);
            if ((double)f != vals[i] || (b & 0xffffu)) {
                int64_t cur = bad_val.load();
                while (i < cur && !bad_val.compare_exchange_weak(cur, i)) {
                }
                break;
            }
        }
    });
    if (bad_idx.load() != INT64_MAX)
        return fail(h, PWM_ERR_INVALID, "pwm_load_triplets: index out of range at entry " + std::to_string(bad_idx.load()));
    if (bad_val.load() != INT64_MAX)
        return fail(h, PWM_ERR_INVALID,
                    "BF16 storage needs values exact in BF16 (entry " + std::to_string(bad_val.load()) + ")");
    reset_matrix(h);
    h->nor = nor;
    h->noc = noc;
    h->nnz = nnz;
    const bool f64 = h->vtype == PWM_VALUES_F64;
    DevBuf dtmp;
    int rc = coo_to_crs_device(h, nor, noc, nnz, rows, cols, vals, h->rs, h->ci, f64 ? h->val : dtmp);
    if (rc != PWM_OK) return rc;
    hipError_t e;
    if (!f64) {
        if ((e = alloc_padded(h->val, nnz, value_size(h->vtype))) != hipSuccess)
            return hip_fail(h, e, "value allocation");
        if (nnz > 0) {
            e = pwmk::launch_convert_values(dtmp.as<double>(), h->val.p, nnz, h->vtype, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) return hip_fail(h, e, "value conversion");
        }
    }
    std::vector<int32_t> hrs(nor + 1);
    if ((e = hipMemcpy(hrs.data(), h->rs.p, (nor + 1) * 4, hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(h, e, "row_start download");
    h->h_rs = hrs;
    if (h->format == PWM_FORMAT_CSB || (h->format == PWM_FORMAT_BLOCKCOH && h->coh_build != 1))
        return finish_load(h, h->h_rs.data(), nullptr, nullptr);
    if (h->format == PWM_FORMAT_BLOCKCOH) {
        std::vector<int32_t> hci(std::max<int64_t>(nnz, 1));
        std::vector<double> hdv(std::max<int64_t>(nnz, 1));
        const DevBuf& dsrc = f64 ? h->val : dtmp;
        if (nnz && (e = hipMemcpy(hci.data(), h->ci.p, nnz * 4, hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(h, e, "CRS download");
        if (nnz && (e = hipMemcpy(hdv.data(), dsrc.p, nnz * 8, hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(h, e, "CRS download");
        return finish_load(h, hrs.data(), hci.data(), hdv.data());
    }
    return finish_load(h, nullptr, nullptr, nullptr);
}

// SparseMatrix::generatePoissonMatrix: CRS filled on the device
// (Util/Poisson.hpp:32-83).  The blocked formats build from the same matrix.
extern "C" int pwm_generate_poisson(pwm_matrix* h, int64_t m, int64_t n, int partitions) {
    (void)partitions;
    if (!h) return fail(nullptr, PWM_ERR_INVALID, "pwm_generate_poisson: NULL handle");
    if (m <= 0 || n <= 0 || m * n > INT32_MAX - 1 || 5 * m * n > PWM_MAX_NNZ)
        return fail(h, PWM_ERR_INVALID, "pwm_generate_poisson: need m, n > 0 and 5*m*n <= PWM_MAX_NNZ");
    reset_matrix(h);
    const int64_t N = m * n, Z = n * (m + 2 * (m - 1)) + 2 * (n - 1) * m;
    h->nor = h->noc = N;
    h->nnz = Z;
    hipError_t e;
    if ((e = h->rs.alloc((N + 1) * 4)) != hipSuccess || (e = alloc_padded(h->ci, Z, 4)) != hipSuccess ||
        (e = alloc_padded(h->val, Z, value_size(h->vtype))) != hipSuccess)
        return hip_fail(h, e, "Poisson allocation");
    e = pwmk::launch_fill_poisson(m, n, h->rs.as<int32_t>(), h->ci.as<int32_t>(), h->val.p, h->vtype, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "Poisson fill");
    if (h->format == PWM_FORMAT_CSB || (h->format == PWM_FORMAT_BLOCKCOH && h->coh_build != 1)) {
        // the device builds need only row_start on the host (block-row sums,
        // band cuts); CSB's host-built fallback downloads the rest itself
        h->h_rs.resize(N + 1);
        if ((e = hipMemcpy(h->h_rs.data(), h->rs.p, (N + 1) * 4, hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(h, e, "Poisson download");
        return finish_load(h, h->h_rs.data(), nullptr, nullptr);
    }
    if (h->format == PWM_FORMAT_BLOCKCOH) {
        // the host builder works from host CRS arrays
        std::vector<int32_t> rs(N + 1), ci(Z);
        std::vector<double> dv(Z);
        DevBuf tmp;
        if ((e = tmp.alloc(Z * 8)) != hipSuccess) return hip_fail(h, e, "staging allocation");
        e = pwmk::launch_fill_poisson(m, n, h->rs.as<int32_t>(), h->ci.as<int32_t>(), tmp.p, pwmk::VT_F64, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) e = hipMemcpy(rs.data(), h->rs.p, (N + 1) * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(ci.data(), h->ci.p, Z * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(dv.data(), tmp.p, Z * 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(h, e, "Poisson download");
        h->h_rs = rs;
        return finish_load(h, rs.data(), ci.data(), dv.data());
    }
    return finish_load(h, nullptr, nullptr, nullptr);
}

extern "C" int pwm_info(const pwm_matrix* h, pwm_info_t* out) {
    if (!h || !out) return fail(const_cast<pwm_matrix*>(h), PWM_ERR_INVALID, "pwm_info: NULL argument");
    memset(out, 0, sizeof(*out));
    out->nor = h->nor;
    out->noc = h->noc;
    out->nnz = h->nnz;
    out->format = h->format;
    out->value_type = h->vtype;
    // CRS: the lanes per row of the vector kernel, 0 when the tiles kernel runs
    out->group_size = h->format == PWM_FORMAT_CSB ? h->csb.groups
                      : (h->format == PWM_FORMAT_CRS && h->crs_used == 2) ? 0
                                                                           : h->group;
    out->long_rows = h->lr.n_rows;
    out->threads = (h->format == PWM_FORMAT_CRS_MERGE || (h->format == PWM_FORMAT_CRS && h->crs_used == 2))
                       ? (int32_t)h->mp.P
                   : h->format == PWM_FORMAT_BLOCKCOH ? (int32_t)h->coh.bands.size()
                   : h->format == PWM_FORMAT_CSB      ? (int32_t)h->csb.n_tasks
                                                      : h->vec_blocks;
    out->block_bits = h->format == PWM_FORMAT_CSB ? h->csb.bits : 0;
    out->balanced = h->format == PWM_FORMAT_CSB ? h->csb.balanced : 0;
    out->unroll = h->format == PWM_FORMAT_CSB && h->csb.stream ? h->csb.unroll : 0;
    out->layout = h->format != PWM_FORMAT_CSB ? 0 : h->csb.sweep ? 4 : !h->csb.stream ? 1 : h->csb.wide ? 3 : 2;
    switch (h->format) {
        case PWM_FORMAT_CRS:
            out->kernels = 1 + (h->crs_used == 2 ? (h->tiles_carry ? 1 : 0) : (h->lr.n_rows > 0 ? 1 : 0));
            break;
        case PWM_FORMAT_CRS_MERGE: out->kernels = 2; break;
        case PWM_FORMAT_CSB: out->kernels = h->csb.sweep ? 1 : 1 + (h->csb.groups > 1 ? 1 : 0); break;
        case PWM_FORMAT_BLOCKCOH: out->kernels = 1 + (h->coh.n_atomic > 0 ? 1 : 0); break;
    }
    if (h->format == PWM_FORMAT_CSB && h->csb.n_tasks <= 0) out->kernels = 0;
    out->high_occupancy = h->format == PWM_FORMAT_BLOCKCOH ? h->coh.occ8 : 0;
    out->non_temporal = h->format == PWM_FORMAT_CSB                         ? h->csb.nt
                        : (h->format == PWM_FORMAT_CRS && h->crs_used == 2) ? h->crs_nt_used
                                                                           : 0;
    const pwm_impl::DevBuf* bufs[] = {&h->rs, &h->ci, &h->val, &h->lr_chunks, &h->lr_rows, &h->lr_begin,
                                      &h->lr_partial, &h->split_a, &h->split_b, &h->carry_row, &h->carry_val, &h->tiles,
                                      &h->csb.ind, &h->csb.data, &h->csb.blk_ptr, &h->csb.work, &h->csb.ypart,
                                      &h->csb.soff, &h->csb.srows, &h->csb.sweep_start,
                                      &h->coh.row_jump, &h->coh.col_jump, &h->coh.data, &h->coh.row_jump_block,
                                      &h->coh.col_jump_block, &h->coh.block_nnz, &h->coh.block_coords,
                                      &h->coh.band_table, &h->coh.tasks, &h->pw_ctrl, &h->pw_rec, &h->pw_partials,
                                      &h->hx, &h->hy};
    int64_t b = 0;
    for (auto* d : bufs) b += (int64_t)d->bytes;
    out->device_bytes = b;
    return PWM_OK;
}

// ---------------------------------------------------------------- SpMV

// A vector handed to a _device entry point must be coarse-grained device
// memory (hipMalloc) that holds `bytes` from p.  Host memory, pinned
// (hipHostMalloc) or not, managed memory and fine-grained device allocations
// are refused: the kernels add into y with hardware FP64 atomics (the library
// is built with unsafe FP atomics), which are not performed on fine-grained
// or host memory.
static int check_device_vec(pwm_matrix* h, const void* p, size_t bytes, const char* who, const char* name) {
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // a refused pointer must not show up as a later launch error
        return fail(h, PWM_ERR_INVALID, std::string(who) + ": " + name + " is not device memory");
    }
    if (a.type != hipMemoryTypeDevice || a.isManaged)
        return fail(h, PWM_ERR_INVALID, std::string(who) + ": " + name + " is host or managed memory, not hipMalloc device memory");
    if (a.allocationFlags & hipDeviceMallocFinegrained)
        return fail(h, PWM_ERR_INVALID, std::string(who) + ": " + name + " is fine-grained device memory");
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    e = hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(h, PWM_ERR_INVALID, std::string(who) + ": " + name + " has no device allocation");
    }
    if ((const char*)p + bytes > (const char*)base + size)
        return fail(h, PWM_ERR_INVALID, std::string(who) + ": " + name + " is shorter than the vector");
    return PWM_OK;
}

static int check_device_xy(pwm_matrix* h, const void* x, size_t x_elems, const void* y, size_t y_elems,
                           const char* who) {
    const size_t vs = vector_size(h->vtype);
    int rc = check_device_vec(h, x, x_elems * vs, who, "x");
    if (rc == PWM_OK) rc = check_device_vec(h, y, y_elems * vs, who, "y");
    return rc;
}

extern "C" int pwm_mv_device(pwm_matrix* h, const void* x, void* y) {
    if (!h) return fail(nullptr, PWM_ERR_INVALID, "pwm_mv_device: NULL handle");
    if (!h->loaded) return fail(h, PWM_ERR_STATE, "pwm_mv_device: no matrix loaded");
    if (!x || !y) return fail(h, PWM_ERR_INVALID, "pwm_mv_device: NULL vector");
    const size_t vs = vector_size(h->vtype);
    const char* xb = (const char*)x;
    const char* yb = (const char*)y;
    if (xb < yb + h->nor * vs && yb < xb + h->noc * vs)
        return fail(h, PWM_ERR_INVALID, "pwm_mv_device: x and y overlap");
    if (int rc0 = check_device_xy(h, x, h->noc, y, h->nor, "pwm_mv_device")) return rc0;
    int rc = spmv_launch(h, x, y, nullptr, nullptr);
    if (rc != PWM_OK) return rc;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "pwm_mv_device");
    return PWM_OK;
}

static int stage_vectors(pwm_matrix* h) {
    const size_t vs = vector_size(h->vtype);
    const size_t need = (size_t)std::max(h->nor, h->noc) * vs;
    hipError_t e;
    if (h->hx.bytes < need && (e = h->hx.alloc(need)) != hipSuccess) return hip_fail(h, e, "staging allocation");
    if (h->hy.bytes < need && (e = h->hy.alloc(need)) != hipSuccess) return hip_fail(h, e, "staging allocation");
    return PWM_OK;
}

// Host double vector -> device vector of the handle's type.
static int upload_vec(pwm_matrix* h, const double* src, void* dst, int64_t n) {
    hipError_t e;
    if (h->vtype == PWM_VALUES_F64) {
        e = upload_staged(dst, src, n * 8, h->stream);
    } else {
        std::unique_ptr<float[]> f(new float[std::max<int64_t>(n, 1)]);
        parallel_for(n, (int64_t)1 << 18, [&](int64_t i0, int64_t i1, int) {
            for (int64_t i = i0; i < i1; ++i) f[i] = (float)src[i];
        });
        e = upload_staged(dst, f.get(), n * 4, h->stream);
    }
    if (e != hipSuccess) return hip_fail(h, e, "vector upload");
    return PWM_OK;
}
static int download_vec(pwm_matrix* h, const void* src, double* dst, int64_t n) {
    hipError_t e;
    if (h->vtype == PWM_VALUES_F64) {
        e = download_staged(dst, src, n * 8, h->stream);
    } else {
        std::unique_ptr<float[]> f(new float[std::max<int64_t>(n, 1)]);
        e = download_staged(f.get(), src, n * 4, h->stream);
        if (e == hipSuccess)
            parallel_for(n, (int64_t)1 << 18, [&](int64_t i0, int64_t i1, int) {
                for (int64_t i = i0; i < i1; ++i) dst[i] = f[i];
            });
    }
    if (e != hipSuccess) return hip_fail(h, e, "vector download");
    return PWM_OK;
}

extern "C" int pwm_mv(pwm_matrix* h, const double* x, double* y) {
    if (!h) return fail(nullptr, PWM_ERR_INVALID, "pwm_mv: NULL handle");
    if (!h->loaded) return fail(h, PWM_ERR_STATE, "pwm_mv: no matrix loaded");
    if (!x || !y) return fail(h, PWM_ERR_INVALID, "pwm_mv: NULL vector");
    if (x < y + h->nor && y < x + h->noc) return fail(h, PWM_ERR_INVALID, "pwm_mv: x and y overlap");
    int rc = stage_vectors(h);
    if (rc == PWM_OK) rc = upload_vec(h, x, h->hx.p, h->noc);
    if (rc == PWM_OK) rc = spmv_launch(h, h->hx.p, h->hy.p, nullptr, nullptr);
    if (rc == PWM_OK) rc = download_vec(h, h->hy.p, y, h->nor);
    return rc;
}

// ---------------------------------------------------------------- power method

static int power_workspace(pwm_matrix* h, int64_t max_it) {
    hipError_t e;
    if (!h->pw_ctrl.p && (e = h->pw_ctrl.alloc(64)) != hipSuccess) return hip_fail(h, e, "power workspace");
    const int64_t np = std::max<int64_t>(partials_needed(h), 1);
    if ((int64_t)h->pw_partials.bytes < np * 24 && (e = h->pw_partials.alloc(np * 24)) != hipSuccess)
        return hip_fail(h, e, "power workspace");
    if (h->pw_rec_cap < max_it) {
        if ((e = h->pw_rec.alloc(std::max<int64_t>(max_it, 1) * 32)) != hipSuccess)
            return hip_fail(h, e, "power workspace");
        h->pw_rec_cap = max_it;
    }
    return PWM_OK;
}

// Device workspace of one power method: step control (2 ints), one record
// of 4 doubles per step, 3 FP64 partial sums per partial slot.
struct PowerWs {
    int* ctrl;
    double* rec;
    double* partials;
};

// Runs the device power loop on buf0 (start vector) / buf1 with workspace ws
// (NULL: the handle's own, allocated on first use); returns steps done.
static int run_power(pwm_matrix* h, void* buf0, void* buf1, int64_t max_it, double tol, int64_t* steps,
                     double* lambda, double* ynorm, const PowerWs* ws = nullptr) {
    if (h->nor != h->noc) return fail(h, PWM_ERR_NOT_SQUARE, "power method needs a square matrix");
    if (max_it < 0) return fail(h, PWM_ERR_INVALID, "negative iteration count");
    if (max_it == 0) {
        if (steps) *steps = 0;
        return PWM_OK;
    }
    int rc;
    PowerWs own;
    if (!ws) {
        if ((rc = power_workspace(h, max_it)) != PWM_OK) return rc;
        own = PowerWs{h->pw_ctrl.as<int>(), h->pw_rec.as<double>(), h->pw_partials.as<double>()};
        ws = &own;
    }
    PowerArgs pw;
    pw.buf0 = buf0;
    pw.buf1 = buf1;
    pw.ctrl = ws->ctrl;
    pw.rec = ws->rec;
    pw.partials = ws->partials;
    pw.max_it = max_it;
    pw.tol = tol;
    hipError_t e = pwmk::launch_power_reset(pw, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "power reset");
    const int64_t chunk = tol < 0 ? max_it : 32;
    int64_t issued = 0;
    int ctrl[2] = {0, 0};
    while (issued < max_it) {
        const int64_t n = std::min(chunk, max_it - issued);
        for (int64_t i = 0; i < n; ++i) {
            rc = spmv_launch(h, nullptr, nullptr, &pw, nullptr);
            if (rc != PWM_OK) return rc;
        }
        issued += n;
        if (tol >= 0) {
            e = hipMemcpyAsync(ctrl, ws->ctrl, 8, hipMemcpyDeviceToHost, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) return hip_fail(h, e, "power loop");
            if (ctrl[1]) break;
        }
    }
    e = pwmk::launch_power_finish(pw, h->nor, h->vtype == PWM_VALUES_F64, 0, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ctrl, ws->ctrl, 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "power finish");
    if (steps) *steps = ctrl[0];
    if ((lambda || ynorm) && ctrl[0] > 0) {
        double rec[4];
        e = hipMemcpyAsync(rec, ws->rec + 4 * (ctrl[0] - 1), 32, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "power record");
        if (lambda) *lambda = rec[3];
        if (ynorm) *ynorm = sqrt(rec[0]);
    }
    return PWM_OK;
}

extern "C" int pwm_power_iteration(pwm_matrix* h, double* x, double* y, int64_t it) {
    if (!h) return fail(nullptr, PWM_ERR_INVALID, "pwm_power_iteration: NULL handle");
    if (!h->loaded) return fail(h, PWM_ERR_STATE, "pwm_power_iteration: no matrix loaded");
    if (!x || !y) return fail(h, PWM_ERR_INVALID, "pwm_power_iteration: NULL vector");
    if (h->nor != h->noc) return fail(h, PWM_ERR_NOT_SQUARE, "power method needs a square matrix");
    if (x < y + h->nor && y < x + h->nor) return fail(h, PWM_ERR_INVALID, "pwm_power_iteration: x and y overlap");
    if (it <= 0) return it < 0 ? fail(h, PWM_ERR_INVALID, "negative iteration count") : PWM_OK;
    int rc = stage_vectors(h);
    if (rc == PWM_OK) rc = upload_vec(h, x, h->hx.p, h->nor);
    int64_t steps = 0;
    if (rc == PWM_OK) rc = run_power(h, h->hx.p, h->hy.p, it, -1.0, &steps, nullptr, nullptr);
    if (rc == PWM_OK) rc = download_vec(h, h->hx.p, x, h->nor);
    if (rc == PWM_OK) rc = download_vec(h, h->hy.p, y, h->nor);
    return rc;
}

extern "C" int pwm_power_method_device(pwm_matrix* h, void* x, void* y, int64_t max_it, double tol, double* lambda,
                                       double* ynorm, int64_t* iterations) {
    if (!h) return fail(nullptr, PWM_ERR_INVALID, "pwm_power_method_device: NULL handle");
    if (!h->loaded) return fail(h, PWM_ERR_STATE, "pwm_power_method_device: no matrix loaded");
    if (!x || !y) return fail(h, PWM_ERR_INVALID, "pwm_power_method_device: NULL vector");
    const size_t vs = vector_size(h->vtype);
    if ((const char*)x < (const char*)y + h->nor * vs && (const char*)y < (const char*)x + h->nor * vs)
        return fail(h, PWM_ERR_INVALID, "pwm_power_method_device: x and y overlap");
    if (int rc0 = check_device_xy(h, x, h->nor, y, h->nor, "pwm_power_method_device")) return rc0;
    int64_t steps = 0;
    int rc = run_power(h, x, y, max_it, tol, &steps, lambda, ynorm);
    if (iterations) *iterations = steps;
    return rc;
}

// ---------------------------------------------------------------- row partitions

// TripletToMultipleCRS (Util/TripletToCRS.hpp:309-385) over the device CRS:
// partition p holds the rows [first_rows[p], first_rows[p] + thread_rows[p])
// with am_rows = nor / partitions rows each (integer division, :320) and the
// remainder in the last partition (:325-326).
static void row_partitions(int64_t nor, int partitions, std::vector<int64_t>& first, std::vector<int64_t>& rows) {
    first.assign(partitions, 0);
    rows.assign(partitions, 0);
    const int64_t am_rows = nor / partitions;
    int64_t last = 0;
    for (int p = 0; p < partitions; ++p) {
        first[p] = last;
        last = p == partitions - 1 ? nor : first[p] + am_rows;
        rows[p] = last - first[p];
    }
}

static int partition_args(pwm_matrix* h, int partitions, const char* who) {
    if (!h) return fail(nullptr, PWM_ERR_INVALID, std::string(who) + ": NULL handle");
    if (!h->loaded || !h->rs.p) return fail(h, PWM_ERR_STATE, std::string(who) + ": needs a loaded CRS or CRS_MERGE handle");
    if (partitions < 1) return fail(h, PWM_ERR_INVALID, std::string(who) + ": partitions must be at least 1");
    return PWM_OK;
}

extern "C" int pwm_crs_partitions(pwm_matrix* h, int partitions, int32_t* first_rows, int32_t* thread_rows,
                                  int64_t* part_nnz) {
    if (int rc = partition_args(h, partitions, "pwm_crs_partitions")) return rc;
    std::vector<int64_t> first, rows;
    row_partitions(h->nor, partitions, first, rows);
    for (int p = 0; p < partitions; ++p) {
        if (first_rows) first_rows[p] = (int32_t)first[p];
        if (thread_rows) thread_rows[p] = (int32_t)rows[p];
        if (part_nnz) {
            int32_t b[2] = {0, 0};
            hipError_t e = hipMemcpy(&b[0], h->rs.as<int32_t>() + first[p], 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(&b[1], h->rs.as<int32_t>() + first[p] + rows[p], 4, hipMemcpyDeviceToHost);
            if (e != hipSuccess) return hip_fail(h, e, "pwm_crs_partitions");
            part_nnz[p] = (int64_t)b[1] - b[0];
        }
    }
    return PWM_OK;
}

extern "C" int pwm_crs_partition_export(pwm_matrix* h, int partitions, int part, int32_t* row_start, int32_t* col_ind,
                                        double* values) {
    if (int rc = partition_args(h, partitions, "pwm_crs_partition_export")) return rc;
    if (part < 0 || part >= partitions || !row_start)
        return fail(h, PWM_ERR_INVALID, "pwm_crs_partition_export: bad partition or NULL row_start");
    std::vector<int64_t> first, rows;
    row_partitions(h->nor, partitions, first, rows);
    const int64_t f = first[part], n = rows[part];
    hipError_t e = hipMemcpy(row_start, h->rs.as<int32_t>() + f, (n + 1) * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(h, e, "pwm_crs_partition_export");
    const int32_t base = row_start[0];
    for (int64_t j = 0; j <= n; ++j) row_start[j] -= base;
    const int64_t nz = row_start[n];
    if (nz > 0 && (!col_ind || !values)) return fail(h, PWM_ERR_INVALID, "pwm_crs_partition_export: NULL arrays");
    if (nz > 0) {
        e = hipMemcpy(col_ind, h->ci.as<int32_t>() + base, nz * 4, hipMemcpyDeviceToHost);
        DevBuf tmp;
        if (e == hipSuccess) e = tmp.alloc(nz * 8);
        if (e == hipSuccess)
            e = pwmk::launch_copy_values_f64((const char*)h->val.p + (size_t)base * value_size(h->vtype), h->vtype,
                                             tmp.as<double>(), nz, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) e = hipMemcpy(values, tmp.p, nz * 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(h, e, "pwm_crs_partition_export");
    }
    return PWM_OK;
}

// ---------------------------------------------------------------- no-allocation power method

// Workspace of pwm_power_method_device_noalloc: one fixed device array per
// GPU (each device holds its own copy until the process ends).  Layout:
// step control at 0, kNoAllocMaxIt step records (32 B) from 256, then the
// FP64 partial sums (24 B per slot) up to the end.
constexpr int64_t kNoAllocBytes = (int64_t)8 << 20;
constexpr int64_t kNoAllocMaxIt = 65536;
constexpr int64_t kNoAllocRecOff = 256;
constexpr int64_t kNoAllocPartOff = kNoAllocRecOff + 32 * kNoAllocMaxIt;
constexpr int64_t kNoAllocMaxSlots = (kNoAllocBytes - kNoAllocPartOff) / 24;
static_assert(PWM_NOALLOC_MAX_IT == kNoAllocMaxIt, "header and library disagree on the iteration limit");
static_assert(PWM_NOALLOC_MAX_SLOTS == kNoAllocMaxSlots, "header and library disagree on the slot limit");
static __device__ __align__(256) char scratch[kNoAllocBytes];

// Device address of `scratch` on the current device, looked up once per
// device with that device current (pwm_create does it for every device it
// runs on; the lookup loads the array onto the device).  The host-side name
// `scratch` is only a stand-in: host code passes this address to kernels,
// never the name.
static int noalloc_scratch(pwm_matrix* h, char** out) {
    static std::mutex mu;
    static std::vector<char*> addr;  // per device ordinal
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(h, e, "pwm_power_method_device_noalloc: hipGetDevice");
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return hip_fail(h, e != hipSuccess ? e : hipErrorInvalidSymbol,
                            "pwm_power_method_device_noalloc: workspace address");
        }
        addr[dev] = (char*)p;
    }
    *out = addr[dev];
    return PWM_OK;
}

extern "C" int pwm_power_method_device_noalloc(pwm_matrix* h, void* x, void* y, int64_t max_it, double tol,
                                               double* lambda, double* ynorm, int64_t* iterations) {
    if (!h) return fail(nullptr, PWM_ERR_INVALID, "pwm_power_method_device_noalloc: NULL handle");
    if (!h->loaded) return fail(h, PWM_ERR_STATE, "pwm_power_method_device_noalloc: no matrix loaded");
    if (!x || !y) return fail(h, PWM_ERR_INVALID, "pwm_power_method_device_noalloc: NULL vector");
    const size_t vs = vector_size(h->vtype);
    if ((const char*)x < (const char*)y + h->nor * vs && (const char*)y < (const char*)x + h->nor * vs)
        return fail(h, PWM_ERR_INVALID, "pwm_power_method_device_noalloc: x and y overlap");
    if (max_it > kNoAllocMaxIt)
        return fail(h, PWM_ERR_INVALID, "pwm_power_method_device_noalloc: max_it above " +
                                            std::to_string(kNoAllocMaxIt) + ", the workspace's step records");
    const int64_t slots = std::max<int64_t>(partials_needed(h), 1);
    if (slots > kNoAllocMaxSlots)
        return fail(h, PWM_ERR_INVALID, "pwm_power_method_device_noalloc: the matrix's plan needs " +
                                            std::to_string(slots) + " partial-sum slots, above the workspace's " +
                                            std::to_string(kNoAllocMaxSlots));
    if (int rc0 = check_device_xy(h, x, h->nor, y, h->nor, "pwm_power_method_device_noalloc")) return rc0;
    char* base = nullptr;
    if (int rc0 = noalloc_scratch(h, &base)) return rc0;
    // the step control is reset by the power loop's first kernel; records
    // and partial sums are written before they are read
    const PowerWs ws{(int*)base, (double*)(base + kNoAllocRecOff), (double*)(base + kNoAllocPartOff)};
    int64_t steps = 0;
    int rc = run_power(h, x, y, max_it, tol, &steps, lambda, ynorm, &ws);
    if (iterations) *iterations = steps;
    return rc;
}

// ---------------------------------------------------------------- exports

extern "C" int pwm_merge_splits(pwm_matrix* h, int64_t* P, int64_t* a, int64_t* b) {
    if (!h || !P) return fail(h, PWM_ERR_INVALID, "pwm_merge_splits: NULL argument");
    if (h->format != PWM_FORMAT_CRS_MERGE || !h->loaded)
        return fail(h, PWM_ERR_STATE, "pwm_merge_splits: not a loaded CRS_MERGE handle");
    *P = h->mp.P;
    if (!a || !b) return PWM_OK;
    hipError_t e = hipMemcpy(a, h->split_a.p, (h->mp.P + 1) * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(b, h->split_b.p, (h->mp.P + 1) * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(h, e, "pwm_merge_splits");
    return PWM_OK;
}

extern "C" int pwm_crs_export(pwm_matrix* h, int32_t* row_start, int32_t* col_ind, double* values) {
    if (!h || !row_start || !col_ind || !values) return fail(h, PWM_ERR_INVALID, "pwm_crs_export: NULL argument");
    if (!h->loaded || !h->rs.p) return fail(h, PWM_ERR_STATE, "pwm_crs_export: no CRS arrays on the device");
    hipError_t e = hipMemcpy(row_start, h->rs.p, (h->nor + 1) * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && h->nnz) e = hipMemcpy(col_ind, h->ci.p, h->nnz * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && h->nnz) {
        DevBuf tmp;
        e = tmp.alloc(h->nnz * 8);
        if (e == hipSuccess) e = pwmk::launch_copy_values_f64(h->val.p, h->vtype, tmp.as<double>(), h->nnz, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) e = hipMemcpy(values, tmp.p, h->nnz * 8, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return hip_fail(h, e, "pwm_crs_export");
    return PWM_OK;
}

// ---------------------------------------------------------------- device helpers

extern "C" int pwm_device_alloc(size_t bytes, void** ptr) {
    if (!ptr) return fail(nullptr, PWM_ERR_INVALID, "pwm_device_alloc: NULL out");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return hip_fail(nullptr, e, "pwm_device_alloc");
    return PWM_OK;
}
extern "C" int pwm_device_free(void* ptr) {
    if (ptr) {
        hipError_t e = hipFree(ptr);
        if (e != hipSuccess) return hip_fail(nullptr, e, "pwm_device_free");
    }
    return PWM_OK;
}
extern "C" int pwm_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? PWM_OK : hip_fail(nullptr, e, "pwm_memcpy_h2d");
}
extern "C" int pwm_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? PWM_OK : hip_fail(nullptr, e, "pwm_memcpy_d2h");
}
extern "C" int pwm_device_synchronize(void) {
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? PWM_OK : hip_fail(nullptr, e, "pwm_device_synchronize");
}
extern "C" int pwm_device_fill(void* ptr, int64_t n, int is_f64, double value) {
    hipError_t e = pwmk::launch_fill(ptr, n, is_f64, value, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? PWM_OK : hip_fail(nullptr, e, "pwm_device_fill");
}

extern "C" int pwm_device_copy(void* dst, const void* src, int64_t bytes) {
    if (!dst || !src || bytes < 0 || bytes % 16) return fail(nullptr, PWM_ERR_INVALID, "pwm_device_copy: bad argument");
    hipError_t e = pwmk::launch_stream_copy(dst, src, bytes, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? PWM_OK : hip_fail(nullptr, e, "pwm_device_copy");
}

// ---------------------------------------------------------------- timing

extern "C" int pwm_gather_rate(const void* x_dev, int64_t x_len, int64_t window, int per_xcd, int64_t n,
                               double* ms) {
    const int kind = (per_xcd >> 8) & 0xff;
    if (kind != 0 && (kind < 1 || (kind > 4 && kind < 17) || kind > 20 || window * 8 > INT32_MAX))
        return fail(nullptr, PWM_ERR_INVALID, "pwm_gather_rate: unknown gather load kind (or window over 2 GB)");
    if (!ms || window < 0 || window > UINT32_MAX ||
        (window > 0 && (!x_dev || ((per_xcd & 1) ? 8 * window : window) > x_len)))
        return fail(nullptr, PWM_ERR_INVALID, "pwm_gather_rate: window must fit x (8 windows when per_xcd)");
    const int64_t m = pwmk::gather_probe_entries(n);
    if (m <= 0) return fail(nullptr, PWM_ERR_INVALID, "pwm_gather_rate: n must be at least 8192");
    DevBuf idx, val, out;
    hipError_t e;
    if ((e = idx.alloc(m * 4)) != hipSuccess || (e = val.alloc(m * 8)) != hipSuccess ||
        (e = out.alloc(m / 8192 * 256 * 8)) != hipSuccess)
        return hip_fail(nullptr, e, "pwm_gather_rate buffers");
    e = pwmk::launch_gather_fill(idx.as<uint32_t>(), val.as<double>(), m, (uint32_t)std::max<int64_t>(window, 1), nullptr);
    hipEvent_t a = nullptr, b = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    const double* x = (const double*)x_dev;
    for (int r = 0; r < 2 && e == hipSuccess; ++r)
        e = pwmk::launch_gather_probe(idx.as<uint32_t>(), val.as<double>(), x, out.as<double>(), m, window, per_xcd, nullptr);
    if (e == hipSuccess) e = hipEventRecord(a, nullptr);
    for (int r = 0; r < 10 && e == hipSuccess; ++r)
        e = pwmk::launch_gather_probe(idx.as<uint32_t>(), val.as<double>(), x, out.as<double>(), m, window, per_xcd, nullptr);
    if (e == hipSuccess) e = hipEventRecord(b, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (e != hipSuccess) return hip_fail(nullptr, e, "pwm_gather_rate");
    *ms = t / 10.0;
    return PWM_OK;
}

// One SpMV of h on h->stream with `mid` recorded after its dominant kernel
// (the tiles / vector / merge / CSB slab kernel; the whole SpMV for BlockCOH).
static int launch_marked(pwm_matrix* h, const void* x, void* y, hipEvent_t mid) {
    const CrsDev A = crs_of(h);
    hipError_t e = hipSuccess;
    switch (h->format) {
        case PWM_FORMAT_CRS:
            if (h->crs_used == 2) {
                e = pwmk::launch_tiles_spmv(A, h->tiles.as<pwmk::ETile>(), h->mp, x, y, nullptr, h->stream);
                (void)hipEventRecord(mid, h->stream);
                if (e == hipSuccess && h->tiles_carry) e = pwmk::launch_merge_fixup(A, h->mp, y, nullptr, h->stream);
                break;
            }
            e = pwmk::launch_crs_vector(A, h->group, h->vec_blocks, h->lr, x, y, nullptr, h->stream);
            (void)hipEventRecord(mid, h->stream);
            if (e == hipSuccess && h->lr.n_rows > 0)
                e = pwmk::launch_crs_finalize(A, h->vec_blocks, h->lr, y, nullptr, h->stream);
            break;
        case PWM_FORMAT_CRS_MERGE:
            e = pwmk::launch_merge_spmv(A, h->mp, x, y, nullptr, h->stream);
            (void)hipEventRecord(mid, h->stream);
            if (e == hipSuccess) e = pwmk::launch_merge_fixup(A, h->mp, y, nullptr, h->stream);
            break;
        case PWM_FORMAT_CSB:
            e = csb_launch(h, x, y, nullptr, 1);
            (void)hipEventRecord(mid, h->stream);
            if (e == hipSuccess) e = csb_launch(h, x, y, nullptr, 2);
            break;
        default: {
            const int rc = spmv_launch(h, x, y, nullptr, nullptr);
            (void)hipEventRecord(mid, h->stream);
            return rc;
        }
    }
    if (e != hipSuccess) return hip_fail(h, e, "timed SpMV");
    return PWM_OK;
}

extern "C" int pwm_time_mv(pwm_matrix* h, const void* x, void* y, int reps, int64_t flush_bytes, double* ms_per_mv,
                           double* kernel_ms) {
    if (!h || !x || !y || reps <= 0) return fail(h, PWM_ERR_INVALID, "pwm_time_mv: bad argument");
    if (!h->loaded) return fail(h, PWM_ERR_STATE, "pwm_time_mv: no matrix loaded");
    if (int rc0 = check_device_xy(h, x, h->noc, y, h->nor, "pwm_time_mv")) return rc0;
    hipError_t e;
    if (flush_bytes > 0 && (int64_t)h->flush.bytes < flush_bytes && (e = h->flush.alloc(flush_bytes)) != hipSuccess)
        return hip_fail(h, e, "flush buffer");
    int rc = PWM_OK;
    if (!kernel_ms && flush_bytes <= 0) {
        // back-to-back SpMVs (the pwm_mv_device launch sequence) between one
        // pair of events: no markers between the steps
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
            if (e0) (void)hipEventDestroy(e0);
            return hip_fail(h, e, "event create");
        }
        (void)hipEventRecord(e0, h->stream);
        for (int r = 0; r < reps && rc == PWM_OK; ++r) rc = spmv_launch(h, x, y, nullptr, nullptr);
        (void)hipEventRecord(e1, h->stream);
        e = hipStreamSynchronize(h->stream);
        if (rc == PWM_OK && e != hipSuccess) rc = hip_fail(h, e, "timed SpMV");
        float t = 0;
        if (rc == PWM_OK) (void)hipEventElapsedTime(&t, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc == PWM_OK && ms_per_mv) *ms_per_mv = t / reps;
        return rc;
    }
    std::vector<hipEvent_t> ev(3 * reps);
    for (auto& v : ev)
        if ((e = hipEventCreate(&v)) != hipSuccess) return hip_fail(h, e, "event create");
    for (int r = 0; r < reps && rc == PWM_OK; ++r) {
        if (flush_bytes > 0) (void)pwmk::launch_flush(h->flush.p, flush_bytes, h->stream);
        (void)hipEventRecord(ev[3 * r], h->stream);
        rc = launch_marked(h, x, y, ev[3 * r + 1]);
        (void)hipEventRecord(ev[3 * r + 2], h->stream);
    }
    e = hipStreamSynchronize(h->stream);
    if (rc == PWM_OK && e != hipSuccess) rc = hip_fail(h, e, "timed SpMV");
    double tot = 0, ker = 0;
    for (int r = 0; r < reps && rc == PWM_OK; ++r) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, ev[3 * r], ev[3 * r + 2]);
        (void)hipEventElapsedTime(&b, ev[3 * r], ev[3 * r + 1]);
        tot += a;
        ker += b;
    }
    for (auto& v : ev) (void)hipEventDestroy(v);
    if (rc != PWM_OK) return rc;
    if (ms_per_mv) *ms_per_mv = tot / reps;
    if (kernel_ms) *kernel_ms = ker / reps;
    return PWM_OK;
}

// SpMVs of `count` handles in turn, reps rounds, all on the first handle's
// stream: with count independent matrices whose arrays together exceed the
// 256 MB Infinity Cache, every SpMV reads its matrix from HBM, back to back.
extern "C" int pwm_time_mv_rotating(pwm_matrix* const* hs, const void* const* xs, void* const* ys, int count,
                                    int reps, double* ms_per_mv, double* kernel_ms) {
    if (!hs || !xs || !ys || count <= 0 || reps <= 0)
        return fail(nullptr, PWM_ERR_INVALID, "pwm_time_mv_rotating: bad argument");
    for (int k = 0; k < count; ++k) {
        pwm_matrix* h = hs[k];
        if (!h) return fail(nullptr, PWM_ERR_INVALID, "pwm_time_mv_rotating: NULL handle");
        if (!h->loaded) return fail(h, PWM_ERR_STATE, "pwm_time_mv_rotating: no matrix loaded");
        if (!xs[k] || !ys[k]) return fail(h, PWM_ERR_INVALID, "pwm_time_mv_rotating: NULL vector");
        if (int rc0 = check_device_xy(h, xs[k], h->noc, ys[k], h->nor, "pwm_time_mv_rotating")) return rc0;
    }
    pwm_matrix* h0 = hs[0];
    const hipStream_t s = h0->stream;
    // work still queued on a handle's own stream finishes before its SpMVs
    // run on the first handle's stream
    for (int k = 1; k < count; ++k) {
        const hipError_t es = hipStreamSynchronize(hs[k]->stream);
        if (es != hipSuccess) return hip_fail(hs[k], es, "pwm_time_mv_rotating");
    }
    std::vector<hipStream_t> saved(count);
    for (int k = 0; k < count; ++k) {
        saved[k] = hs[k]->stream;
        hs[k]->stream = s;
    }
    const int n = count * reps;
    std::vector<hipEvent_t> ev(kernel_ms ? 3 * n : 2, nullptr);
    int rc = PWM_OK;
    hipError_t e = hipSuccess;
    for (auto& v : ev)
        if (e == hipSuccess) e = hipEventCreate(&v);
    if (e != hipSuccess) rc = hip_fail(h0, e, "event create");
    if (rc == PWM_OK && !kernel_ms) {
        (void)hipEventRecord(ev[0], s);
        for (int i = 0; i < n && rc == PWM_OK; ++i) rc = spmv_launch(hs[i % count], xs[i % count], ys[i % count], nullptr, nullptr);
        (void)hipEventRecord(ev[1], s);
    } else if (rc == PWM_OK) {
        for (int i = 0; i < n && rc == PWM_OK; ++i) {
            (void)hipEventRecord(ev[3 * i], s);
            rc = launch_marked(hs[i % count], xs[i % count], ys[i % count], ev[3 * i + 1]);
            (void)hipEventRecord(ev[3 * i + 2], s);
        }
    }
    e = hipStreamSynchronize(s);
    if (rc == PWM_OK && e != hipSuccess) rc = hip_fail(h0, e, "timed SpMV");
    for (int k = count - 1; k >= 0; --k) hs[k]->stream = saved[k];  // reverse order: a handle may repeat
    if (rc == PWM_OK) {
        double tot = 0, ker = 0;
        if (!kernel_ms) {
            float t = 0;
            (void)hipEventElapsedTime(&t, ev[0], ev[1]);
            tot = t;
        } else {
            for (int i = 0; i < n; ++i) {
                float a = 0, b = 0;
                (void)hipEventElapsedTime(&a, ev[3 * i], ev[3 * i + 2]);
                (void)hipEventElapsedTime(&b, ev[3 * i], ev[3 * i + 1]);
                tot += a;
                ker += b;
            }
            *kernel_ms = ker / n;
        }
        if (ms_per_mv) *ms_per_mv = tot / n;
    }
    for (auto& v : ev)
        if (v) (void)hipEventDestroy(v);
    return rc;
}
