// Request server: SpMV-chain and power-method jobs on resident matrices,
// batched continuously on the iteration engine (engine.hpp).
//
// pwm_server_step() admits queued jobs in arrival order while an engine slot
// and device memory are free (memory comes from one arena of `budget`
// bytes; matrices no running job uses are evicted least recently used first
// to make room), runs one chunk of steps of every running job, and retires
// the jobs that stopped.  While chunks run, the queued jobs that free slots
// can take are staged one at a time on a second stream (matrix upload or
// generation, tile plan, job block and start vector); whenever the chunk in
// flight has finished, the next one starts with the jobs staged so far, so
// uploads overlap the running jobs (option overlap = 0: stage, then run one
// chunk).  A job's arithmetic depends only on its own matrix tiles, so its
// result is bit-identical whatever else is in flight.  With the option
// bf16_when_exact an FP32 server stores Poisson matrices and matrices whose
// values are exact in BF16 as BF16 (the engine's tiles branch on their job's
// storage).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <vector>

#include "api_internal.hpp"
#include "engine.hpp"

namespace {

struct SrvMatrix {
    int64_t n = 0, nnz = 0;
    int kind = 0;  // 0: CRS from the host, 1: Poisson generated on the device
    int64_t pm = 0, pn = 0;
    bool borrowed = false;
    const int32_t* rs_p = nullptr;  // host arrays (owned copies or the caller's)
    const int32_t* ci_p = nullptr;
    const double* val_p = nullptr;
    std::vector<int32_t> rs, ci;
    std::vector<double> val;
    bool alive = true;
    bool resident = false;
    size_t off = 0, bytes = 0;
    size_t o_ci = 0, o_val = 0, o_tiles = 0;  // inside the block
    int32_t n_tiles = 0;
    bool vt16 = false;  // BF16 values (bf16_when_exact and every value exact in BF16)
    std::vector<pwmk::ETile> tiles;  // host copy of the plan while resident
    int refs = 0;
    uint64_t last_use = 0;
};

struct SrvJob {
    int32_t mat = 0;
    int64_t max_it = 0;
    double tol = -1.0;
    uint64_t seed = 0;
    int state = 0;  // 0 queued, 1 running, 2 finished
    int slot = -1;
    size_t off = 0, bytes = 0;
    bool has_block = false;  // staged: job block allocated, start vector written, matrix referenced
    double lambda = 0.0, ynorm = 0.0;
    int64_t iters = 0;
    void* last = nullptr;
};

}  // namespace

struct pwm_server {
    int vtype = PWM_VALUES_F32;
    int ipt = 16;
    int chunk = 8;
    int use_graph = 1;
    hipStream_t stream = nullptr;   // the engine's steps, retirement, result downloads
    hipStream_t ustream = nullptr;  // staging: uploads, Poisson generation, tile plans, start vectors
    int overlap = 1;                // stage queued jobs while a chunk runs
    // called between the chunks of a long upload while chunks run: starts the
    // next chunk as soon as the one in flight has finished
    std::function<void()> pump;
    int pump_rc = PWM_OK;
    std::string err;
    pwm_impl::Arena arena;
    std::unique_ptr<pwm_impl::Engine> eng;
    int bf16_exact = 0;  // FP32 server: BF16 storage for matrices exact in BF16
    int persistent = 0;
    std::vector<SrvMatrix> mats;
    std::map<int64_t, SrvJob> jobs;
    std::deque<int64_t> queue;
    size_t staged_front = 0;  // the first staged_front queued jobs are staged (staging goes in order)
    std::vector<int64_t> slot_job;
    std::vector<int64_t> last_finished;  // jobs retired by the last step
    int64_t next_job = 0, finished_total = 0;
    uint64_t clock = 0;
    double t_admit = 0.0, t_run = 0.0, t_retire = 0.0;  // seconds in each phase of pwm_server_step
    int64_t uploads = 0;
};

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int sfail(pwm_server* s, int code, const std::string& m) {
    if (s)
        s->err = m;
    else
        pwm_impl::set_global_error(m);
    return code;
}
static int shp(pwm_server* s, hipError_t e, const char* what) {
    return sfail(s, e == hipErrorOutOfMemory ? PWM_ERR_NOMEM : PWM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static size_t acc_size(int vtype) { return vtype == PWM_VALUES_F64 ? 8 : 4; }

static size_t matrix_value_size(const pwm_server* s, const SrvMatrix& M) {
    return M.vt16 ? 2 : pwm_impl::value_size(s->vtype);
}

static void matrix_layout(const pwm_server* s, SrvMatrix& M) {
    using pwm_impl::Arena;
    const size_t vs = matrix_value_size(s, M);
    M.n_tiles = pwmk::engine_tiles_for(M.n, M.nnz, s->ipt);
    const int64_t zp = M.nnz + pwm_impl::kPadEntries;  // tiles read aligned groups of 4
    M.o_ci = Arena::round((M.n + 1) * 4);
    M.o_val = M.o_ci + Arena::round(zp * 4);
    M.o_tiles = M.o_val + Arena::round(zp * vs);
    M.bytes = M.o_tiles + Arena::round((size_t)M.n_tiles * sizeof(pwmk::ETile));
}

static size_t job_bytes(const pwm_server* s, const SrvMatrix& M) {
    using pwm_impl::Arena;
    return Arena::round(2 * M.n * pwm_impl::vector_size(s->vtype)) + Arena::round((size_t)M.n_tiles * 24) +
           Arena::round((size_t)M.n_tiles * acc_size(s->vtype));
}

extern "C" int pwm_server_create(int value_type, int64_t budget_bytes, int32_t max_jobs, pwm_server** out) {
    if (!out) return sfail(nullptr, PWM_ERR_INVALID, "pwm_server_create: out is NULL");
    *out = nullptr;
    if (value_type < PWM_VALUES_F64 || value_type > PWM_VALUES_BF16)
        return sfail(nullptr, PWM_ERR_INVALID, "pwm_server_create: unknown value type");
    if (budget_bytes < (1 << 20) || max_jobs < 1 || max_jobs > 65536)
        return sfail(nullptr, PWM_ERR_INVALID, "pwm_server_create: budget must be >= 1 MiB and max_jobs in [1, 65536]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return sfail(nullptr, PWM_ERR_HIP, "pwm_server_create: no HIP device available");
    pwm_server* s = new pwm_server();
    s->vtype = value_type;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->ustream, hipStreamNonBlocking);
    if (e == hipSuccess) e = s->arena.reserve((size_t)budget_bytes);
    if (e != hipSuccess) {
        const int rc = e == hipErrorOutOfMemory ? PWM_ERR_NOMEM : PWM_ERR_HIP;
        if (s->stream) (void)hipStreamDestroy(s->stream);
        if (s->ustream) (void)hipStreamDestroy(s->ustream);
        delete s;
        return sfail(nullptr, rc, std::string("pwm_server_create: ") + hipGetErrorString(e));
    }
    std::string err;
    s->eng.reset(new pwm_impl::Engine());
    const int rc = s->eng->init(value_type, max_jobs, s->stream, err);
    if (rc != PWM_OK) {
        s->eng.reset();
        (void)hipStreamDestroy(s->stream);
        (void)hipStreamDestroy(s->ustream);
        delete s;
        return sfail(nullptr, rc, err);
    }
    s->eng->chunk = s->chunk;
    s->slot_job.assign(max_jobs, -1);
    *out = s;
    return PWM_OK;
}

extern "C" int pwm_server_destroy(pwm_server* s) {
    if (!s) return PWM_OK;
    hipStream_t st = s->stream, ust = s->ustream;
    if (st) (void)hipStreamSynchronize(st);
    if (ust) (void)hipStreamSynchronize(ust);
    delete s;  // engine and arena free their device memory here
    if (st) (void)hipStreamDestroy(st);
    if (ust) (void)hipStreamDestroy(ust);
    return PWM_OK;
}

extern "C" const char* pwm_server_last_error(const pwm_server* s) { return s ? s->err.c_str() : pwm_last_error(nullptr); }

extern "C" int pwm_server_set_option(pwm_server* s, const char* key, int64_t value) {
    if (!s || !key) return sfail(s, PWM_ERR_INVALID, "pwm_server_set_option: NULL");
    const std::string k(key);
    if (k == "chunk") {
        if (value < 1 || value > 4096) return sfail(s, PWM_ERR_INVALID, "pwm_server_set_option: chunk must be in [1, 4096]");
        s->chunk = (int)value;
        s->eng->chunk = s->chunk;
        return PWM_OK;
    }
    if (k == "graph") {
        s->use_graph = value != 0;
        s->eng->use_graph = s->use_graph != 0;
        return PWM_OK;
    }
    if (k == "overlap") {
        s->overlap = value != 0;
        return PWM_OK;
    }
    if (k == "persistent") {
        s->persistent = value != 0;
        s->eng->persistent = s->persistent != 0;
        return PWM_OK;
    }
    if (k == "bf16_when_exact") {
        if (value != 0 && s->vtype != PWM_VALUES_F32)
            return sfail(s, PWM_ERR_INVALID, "pwm_server_set_option: bf16_when_exact needs an FP32 server");
        if (!s->mats.empty() || !s->jobs.empty())
            return sfail(s, PWM_ERR_STATE, "pwm_server_set_option: bf16_when_exact is set before any matrix is added");
        if ((value != 0) == (s->bf16_exact != 0)) return PWM_OK;
        // the engine of a mixed server branches per tile on its job's storage
        std::unique_ptr<pwm_impl::Engine> eng(new pwm_impl::Engine());
        std::string err;
        const int rc = eng->init(value != 0 ? pwmk::VT_MIXED16 : s->vtype, (int)s->slot_job.size(), s->stream, err);
        if (rc != PWM_OK) return sfail(s, rc, err);
        if (s->ipt != eng->ipt() && eng->set_ipt(s->ipt) != PWM_OK)
            return sfail(s, PWM_ERR_STATE, "pwm_server_set_option: items_per_thread");
        eng->chunk = s->chunk;
        eng->use_graph = s->use_graph != 0;
        eng->persistent = s->persistent != 0;
        s->eng.swap(eng);
        s->bf16_exact = value != 0;
        return PWM_OK;
    }
    if (k == "items_per_thread") {
        if (value != 4 && value != 8 && value != 16)
            return sfail(s, PWM_ERR_INVALID, "pwm_server_set_option: items_per_thread must be 4, 8 or 16");
        for (const SrvMatrix& M : s->mats)
            if (M.alive && M.resident) return sfail(s, PWM_ERR_STATE, "pwm_server_set_option: matrices are resident");
        if (s->eng->set_ipt((int)value) != PWM_OK) return sfail(s, PWM_ERR_STATE, "pwm_server_set_option: jobs in flight");
        s->ipt = (int)value;
        for (SrvMatrix& M : s->mats)
            if (M.alive) matrix_layout(s, M);
        return PWM_OK;
    }
    return sfail(s, PWM_ERR_INVALID, "pwm_server_set_option: unknown option " + k);
}

static int add_matrix_common(pwm_server* s, SrvMatrix&& M, int32_t* mat) {
    matrix_layout(s, M);
    if (M.bytes + job_bytes(s, M) > s->arena.capacity())
        return sfail(s, PWM_ERR_BUDGET, "pwm_server_add_matrix: matrix and one job exceed the device budget");
    if (mat) *mat = (int32_t)s->mats.size();
    s->mats.push_back(std::move(M));
    return PWM_OK;
}

// Checks a host CRS matrix of n rows with `nnz` entries before it is
// uploaded: non-decreasing row_start from 0 to nnz, columns in [0, n) (unless
// check_cols is false: the upload checks them), and for BF16 storage values
// exact in BF16.  Runs on the host thread pool.
static int check_crs(pwm_server* s, int64_t n, int64_t nnz, const int32_t* rs, const int32_t* ci, const double* vals,
                     const char* who, bool need_bf16, bool check_cols = true) {
    if (rs[0] != 0 || rs[n] != nnz) return sfail(s, PWM_ERR_INVALID, std::string(who) + ": bad row_start");
    std::atomic<int> bad{0};  // 1 row_start, 2 column, 3 BF16 value
    pwm_impl::parallel_for(n, 1 << 18, [&](int64_t b, int64_t e, int) {
        for (int64_t r = b; r < e; ++r)
            if (rs[r + 1] < rs[r]) bad = 1;
    });
    if (bad == 0 && (check_cols || need_bf16))
        pwm_impl::parallel_for(nnz, 1 << 20, [&](int64_t b, int64_t e, int) {
            for (int64_t k = b; k < e; ++k) {
                if (check_cols && (ci[k] < 0 || ci[k] >= n)) {
                    bad = 2;
                    return;
                }
                if (need_bf16) {
                    const float f = (float)vals[k];
                    uint32_t bits;
                    memcpy(&bits, &f, 4);
                    if ((double)f != vals[k] || (bits & 0xffffu)) bad = 3;
                }
            }
        });
    if (bad == 1) return sfail(s, PWM_ERR_INVALID, std::string(who) + ": row_start must be non-decreasing");
    if (bad == 2) return sfail(s, PWM_ERR_INVALID, std::string(who) + ": column index out of range");
    if (bad == 3) return sfail(s, PWM_ERR_INVALID, std::string(who) + ": BF16 storage needs values exact in BF16");
    return PWM_OK;
}

// flags bit 0 (PWM_SERVER_BORROW): keep the caller's arrays instead of a copy.
extern "C" int pwm_server_add_matrix(pwm_server* s, int64_t n, const int32_t* rs, const int32_t* ci, const double* vals,
                                     int32_t flags, int32_t* mat) {
    if (!s || !rs || n <= 0 || n > INT32_MAX - 1) return sfail(s, PWM_ERR_INVALID, "pwm_server_add_matrix: bad argument");
    const int64_t nnz = rs[n];
    if (rs[0] != 0 || nnz < 0 || (nnz > 0 && (!ci || !vals)) || n + nnz > INT32_MAX)
        return sfail(s, PWM_ERR_INVALID, "pwm_server_add_matrix: bad row_start or NULL arrays");
    int rc = check_crs(s, n, nnz, rs, ci, vals, "pwm_server_add_matrix", s->vtype == PWM_VALUES_BF16);
    if (rc != PWM_OK) return rc;
    SrvMatrix M;
    M.n = n;
    M.nnz = nnz;
    M.vt16 = s->bf16_exact && pwm_impl::exact_in_bf16(vals, nnz);
    if (flags & 1) {
        M.borrowed = true;
        M.rs_p = rs;
        M.ci_p = ci;
        M.val_p = vals;
    } else {
        M.rs.assign(rs, rs + n + 1);
        M.ci.assign(ci, ci + nnz);
        M.val.assign(vals, vals + nnz);
        M.rs_p = M.rs.data();
        M.ci_p = M.ci.data();
        M.val_p = M.val.data();
    }
    return add_matrix_common(s, std::move(M), mat);
}

// 2D Poisson matrix generated on the device when it becomes resident
// (Util/Poisson.hpp:32-83).
extern "C" int pwm_server_add_poisson(pwm_server* s, int64_t m, int64_t n, int32_t* mat) {
    if (!s || m <= 0 || n <= 0 || 6 * m * n > INT32_MAX) return sfail(s, PWM_ERR_INVALID, "pwm_server_add_poisson: bad size");
    SrvMatrix M;
    M.kind = 1;
    M.pm = m;
    M.pn = n;
    M.n = m * n;
    M.nnz = n * (m + 2 * (m - 1)) + 2 * (n - 1) * m;
    M.vt16 = s->bf16_exact != 0;  // 4 and -1
    return add_matrix_common(s, std::move(M), mat);
}

static void evict(pwm_server* s, SrvMatrix& M) {
    if (!M.resident) return;
    s->arena.release(M.off, M.bytes);
    M.resident = false;
}

// Remove a matrix: its device copy is freed and its id is never reused.
extern "C" int pwm_server_remove_matrix(pwm_server* s, int32_t mat) {
    if (!s || mat < 0 || mat >= (int32_t)s->mats.size() || !s->mats[mat].alive)
        return sfail(s, PWM_ERR_INVALID, "pwm_server_remove_matrix: unknown matrix");
    SrvMatrix& M = s->mats[mat];
    for (const auto& kv : s->jobs)
        if (kv.second.mat == mat && kv.second.state != 2)
            return sfail(s, PWM_ERR_STATE, "pwm_server_remove_matrix: jobs on this matrix are queued or running");
    evict(s, M);
    M.alive = false;
    M.rs = {};
    M.ci = {};
    M.val = {};
    M.rs_p = nullptr;
    M.ci_p = nullptr;
    M.val_p = nullptr;
    return PWM_OK;
}

// Allocate from the arena, evicting unreferenced resident matrices (least
// recently used first, never `keep`) until the request fits.
static bool alloc_evicting(pwm_server* s, size_t bytes, size_t* off, int keep) {
    for (;;) {
        if (s->arena.alloc(bytes, off)) return true;
        int victim = -1;
        for (int i = 0; i < (int)s->mats.size(); ++i) {
            const SrvMatrix& M = s->mats[i];
            if (i == keep || !M.resident || M.refs > 0) continue;
            if (victim < 0 || M.last_use < s->mats[victim].last_use) victim = i;
        }
        if (victim < 0) return false;
        evict(s, s->mats[victim]);
    }
}

// Upload (or generate) matrix `mi` into its arena block and plan its tiles,
// on the staging stream; returns once they are done.
static int make_resident(pwm_server* s, int mi) {
    SrvMatrix& M = s->mats[mi];
    char* base = (char*)s->arena.at(M.off);
    int32_t* rs = (int32_t*)base;
    const size_t vs = matrix_value_size(s, M);
    const int vt = M.vt16 ? PWM_VALUES_BF16 : s->vtype;  // storage of this matrix
    hipError_t e = hipMemsetAsync(base + M.o_ci + M.nnz * 4, 0, pwm_impl::kPadEntries * 4, s->ustream);
    if (e == hipSuccess) e = hipMemsetAsync(base + M.o_val + M.nnz * vs, 0, pwm_impl::kPadEntries * vs, s->ustream);
    if (e == hipSuccess && M.kind == 1) {
        e = pwmk::launch_fill_poisson(M.pm, M.pn, rs, (int32_t*)(base + M.o_ci), base + M.o_val, vt, s->ustream);
    } else if (e == hipSuccess) {
        if (M.borrowed) {
            // borrowed arrays may have changed since registration: check them
            // again before the kernels index x with them (and, for BF16
            // storage, before their values are narrowed); the columns are
            // checked as they are copied
            const int rc = check_crs(s, M.n, M.nnz, M.rs_p, M.ci_p, M.val_p, "server: borrowed matrix",
                                     vt == PWM_VALUES_BF16, false);
            if (rc != PWM_OK) return rc;
        }
        bool bad_col = false;
        const std::function<void()>* pump = s->pump ? &s->pump : nullptr;
        e = pwm_impl::upload_staged(rs, M.rs_p, (M.n + 1) * 4, s->ustream, pump);
        if (e == hipSuccess)
            e = pwm_impl::upload_indices_staged(base + M.o_ci, M.ci_p, M.nnz, M.n, &bad_col, s->ustream, pump);
        if (e == hipSuccess && bad_col)
            return sfail(s, PWM_ERR_INVALID, "server: borrowed matrix: column index out of range");
        if (e == hipSuccess) e = pwm_impl::upload_values_staged(base + M.o_val, M.val_p, M.nnz, vt, s->ustream, pump);
    }
    if (e == hipSuccess)
        e = pwmk::engine_plan_tiles(rs, M.n, M.nnz, s->ipt, (pwmk::ETile*)(base + M.o_tiles), s->ustream);
    M.tiles.resize(M.n_tiles);
    if (e == hipSuccess)
        e = hipMemcpyAsync(M.tiles.data(), base + M.o_tiles, (size_t)M.n_tiles * sizeof(pwmk::ETile),
                           hipMemcpyDeviceToHost, s->ustream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ustream);
    if (e != hipSuccess) return shp(s, e, "matrix upload");
    M.resident = true;
    return PWM_OK;
}

extern "C" int pwm_server_submit(pwm_server* s, int32_t mat, int64_t max_it, double tol, uint64_t seed, int64_t* job) {
    if (!s || mat < 0 || mat >= (int32_t)s->mats.size() || !s->mats[mat].alive)
        return sfail(s, PWM_ERR_INVALID, "pwm_server_submit: unknown matrix");
    if (max_it < 1) return sfail(s, PWM_ERR_INVALID, "pwm_server_submit: max_it must be >= 1");
    SrvJob J;
    J.mat = mat;
    J.max_it = max_it;
    J.tol = tol;
    J.seed = seed;
    const int64_t id = s->next_job++;
    s->jobs[id] = J;
    s->queue.push_back(id);
    if (job) *job = id;
    return PWM_OK;
}

// Stage at most `max_new` queued jobs among the first `max_jobs` in arrival
// order (jobs staged before count and are skipped): make each one's matrix
// resident and give it a job block with its start vector, on the staging
// stream.  A staged job holds a reference to its matrix, so the matrix is not
// evicted before the job ends.  Stops at the first job that does not fit;
// returns once the staging stream has finished.  *staged = jobs staged.
static int stage(pwm_server* s, size_t max_jobs, size_t max_new, int* staged) {
    int rc = PWM_OK;
    *staged = 0;
    for (size_t k = s->staged_front; k < s->queue.size() && k < max_jobs && (size_t)*staged < max_new; ++k) {
        SrvJob& J = s->jobs[s->queue[k]];
        SrvMatrix& M = s->mats[J.mat];
        if (!M.resident) {
            if (!alloc_evicting(s, M.bytes, &M.off, J.mat)) break;
            rc = make_resident(s, J.mat);
            ++s->uploads;
            if (rc != PWM_OK) {
                s->arena.release(M.off, M.bytes);
                break;
            }
        }
        M.refs++;
        M.last_use = ++s->clock;
        const size_t jb = job_bytes(s, M);
        if (!alloc_evicting(s, jb, &J.off, J.mat)) {
            M.refs--;
            break;
        }
        J.bytes = jb;
        J.has_block = true;
        ++s->staged_front;
        ++*staged;
        const hipError_t e = pwmk::engine_start_vector(s->arena.at(J.off), M.n, s->vtype, J.seed, 0, s->ustream);
        if (e != hipSuccess) {
            rc = shp(s, e, "job start vector");
            break;
        }
    }
    if (*staged > 0) {
        const hipError_t e = hipStreamSynchronize(s->ustream);
        if (e != hipSuccess && rc == PWM_OK) rc = shp(s, e, "job staging");
    }
    return rc;
}

static size_t free_slots(const pwm_server* s) { return (size_t)std::max(0, s->eng->max_slots() - s->eng->active()); }

// Hand the staged jobs at the front of the queue to the engine while slots
// are free.  Only while no steps are in flight (Engine::add).
static void add_staged(pwm_server* s) {
    const size_t xs = pwm_impl::vector_size(s->vtype);
    while (!s->queue.empty() && s->eng->active() < s->eng->max_slots()) {
        const int64_t id = s->queue.front();
        SrvJob& J = s->jobs[id];
        if (!J.has_block) break;
        SrvMatrix& M = s->mats[J.mat];
        char* base = (char*)s->arena.at(J.off);
        const char* mb = (const char*)s->arena.at(M.off);
        const size_t o_part = pwm_impl::Arena::round(2 * M.n * xs);
        const size_t o_carry = o_part + pwm_impl::Arena::round((size_t)M.n_tiles * 24);
        pwmk::EJob j;
        memset(&j, 0, sizeof(j));
        j.rs = (const int32_t*)mb;
        j.ci = (const int32_t*)(mb + M.o_ci);
        j.val = mb + M.o_val;
        j.buf0 = base;
        j.buf1 = base + M.n * xs;
        j.tiles = (const pwmk::ETile*)(mb + M.o_tiles);
        j.partials = (double*)(base + o_part);
        j.carry = base + o_carry;
        j.n = M.n;
        j.nnz = M.nnz;
        j.max_it = J.max_it;
        j.tol = J.tol;
        j.n_tiles = M.n_tiles;
        j.vt16 = M.vt16;
        const int slot = s->eng->add(j, M.tiles.data());
        if (slot < 0) break;  // cannot happen: a slot was free above
        J.slot = slot;
        J.state = 1;
        s->slot_job[slot] = id;
        s->queue.pop_front();
        --s->staged_front;
    }
}

// Admit queued jobs in arrival order while engine slots are free, staging
// those not staged yet; stop at the first that does not fit.
static int admit(pwm_server* s) {
    int n = 0;
    const int rc = stage(s, free_slots(s), SIZE_MAX, &n);
    add_staged(s);
    return rc;
}

// Record the finished slots' results (appended to last_finished), normalise
// their final iterates and free their engine slots.
static int retire(pwm_server* s, const std::vector<int>& fin) {
    hipError_t e = hipSuccess;
    for (int slot : fin) {
        const int64_t id = s->slot_job[slot];
        s->last_finished.push_back(id);
        SrvJob& J = s->jobs[id];
        const pwmk::EState& S = s->eng->state(slot);
        J.lambda = S.lambda;
        J.ynorm = sqrt(S.yy);
        J.iters = S.it;
        J.last = s->eng->last_buffer(slot);
        J.state = 2;
        if (e == hipSuccess) e = pwmk::engine_normalize(J.last, s->mats[J.mat].n, s->vtype, 1.0 / J.ynorm, s->stream);
        s->eng->release(slot);
        s->slot_job[slot] = -1;
        s->mats[J.mat].refs--;
        ++s->finished_total;
    }
    if (e == hipSuccess && !fin.empty()) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return shp(s, e, "job retirement");
    return PWM_OK;
}

// Wait for the steps in flight (if any), retire the jobs that stopped, hand
// the staged jobs to the engine and, with `relaunch`, start the next chunk.
static int cycle(pwm_server* s, bool relaunch) {
    std::vector<int> fin;
    std::string err;
    const double t0 = now_s();
    int rc = s->eng->collect(fin, err);
    const double t1 = now_s();
    s->t_run += t1 - t0;
    if (rc != PWM_OK) return sfail(s, rc, err);
    rc = retire(s, fin);
    s->t_retire += now_s() - t1;
    if (rc != PWM_OK) return rc;
    add_staged(s);
    if (relaunch && s->eng->active() > 0 && (rc = s->eng->launch(s->chunk, err)) != PWM_OK) {
        std::vector<int> none;
        std::string e2;
        (void)s->eng->collect(none, e2);  // let what was queued finish
        return sfail(s, rc, err);
    }
    return PWM_OK;
}

extern "C" int pwm_server_step(pwm_server* s, int32_t* active, int32_t* finished) {
    if (!s) return sfail(nullptr, PWM_ERR_INVALID, "pwm_server_step: NULL");
    s->last_finished.clear();
    std::string err;
    int rc = PWM_OK;
    if (!s->overlap) {
        // stage and admit, then one chunk of every running job
        const double t0 = now_s();
        rc = admit(s);
        s->t_admit += now_s() - t0;
        if (rc != PWM_OK) return rc;
        if (s->eng->active() == 0 && !s->queue.empty())
            return sfail(s, PWM_ERR_BUDGET,
                         "pwm_server_step: the next job does not fit the budget with nothing running "
                         "(collect finished results to free their memory)");
        if (s->eng->active() > 0 && (rc = s->eng->launch(s->chunk, err)) != PWM_OK) {
            std::vector<int> none;
            std::string e2;
            (void)s->eng->collect(none, e2);
            return sfail(s, rc, err);
        }
        if ((rc = cycle(s, false)) != PWM_OK) return rc;
    } else {
        // staged jobs join now; with nothing running, the first queued job is
        // staged at once so that the GPU has work
        const double t0 = now_s();
        add_staged(s);
        if (s->eng->active() == 0 && !s->queue.empty()) {
            int n = 0;
            rc = stage(s, 1, 1, &n);
            add_staged(s);
            if (rc != PWM_OK) return rc;
            if (s->eng->active() == 0)
                return sfail(s, PWM_ERR_BUDGET,
                             "pwm_server_step: the next job does not fit the budget with nothing running "
                             "(collect finished results to free their memory)");
        }
        s->t_admit += now_s() - t0;
        if (s->eng->active() > 0 && (rc = s->eng->launch(s->chunk, err)) != PWM_OK) {
            std::vector<int> none;
            std::string e2;
            (void)s->eng->collect(none, e2);
            return sfail(s, rc, err);
        }
        // the other queued jobs that free slots can take are staged one at a
        // time while chunks run; whenever the chunk in flight has finished
        // (checked after each staged job and between the chunks of a long
        // upload), the next one starts with the jobs staged so far
        auto chunk_done = [s] {
            if (!s->eng->pending()) return true;
            const bool done = hipStreamQuery(s->stream) == hipSuccess;
            (void)hipGetLastError();  // a "not ready" answer is no error for later checks
            return done;
        };
        s->pump_rc = PWM_OK;
        double pumped = 0.0;  // seconds in cycle() from inside uploads (counted as run and retire time)
        s->pump = [s, chunk_done, &pumped] {
            if (s->pump_rc == PWM_OK && s->eng->pending() && chunk_done()) {
                const double tp = now_s();
                s->pump_rc = cycle(s, true);
                pumped += now_s() - tp;
            }
        };
        int stage_rc = PWM_OK;
        for (;;) {
            int n = 0;
            const double ts = now_s(), p0 = pumped;
            stage_rc = stage(s, free_slots(s), 1, &n);
            s->t_admit += now_s() - ts - (pumped - p0);
            if (stage_rc != PWM_OK || n == 0 || s->pump_rc != PWM_OK) break;
            if (chunk_done() && (rc = cycle(s, true)) != PWM_OK) break;
        }
        s->pump = nullptr;
        if (rc == PWM_OK) rc = s->pump_rc;
        const int crc = cycle(s, false);  // the chunk in flight
        if (rc == PWM_OK) rc = crc;
        if (rc != PWM_OK) return rc;
        if (stage_rc != PWM_OK) return stage_rc;
    }
    if (active) *active = s->eng->active();
    if (finished) *finished = (int32_t)s->last_finished.size();
    return PWM_OK;
}

// Results of a finished job; `vec` (n doubles, may be NULL) receives the
// normalised final iterate.  The job and its device memory are released.
extern "C" int pwm_server_result(pwm_server* s, int64_t job, double* lambda, double* ynorm, int64_t* iters, double* vec) {
    if (!s) return sfail(nullptr, PWM_ERR_INVALID, "pwm_server_result: NULL");
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return sfail(s, PWM_ERR_INVALID, "pwm_server_result: unknown job");
    SrvJob& J = it->second;
    if (J.state != 2) return sfail(s, PWM_ERR_STATE, "pwm_server_result: job has not finished");
    if (lambda) *lambda = J.lambda;
    if (ynorm) *ynorm = J.ynorm;
    if (iters) *iters = J.iters;
    hipError_t e = hipSuccess;
    if (vec) {
        const int64_t n = s->mats[J.mat].n;
        if (s->vtype == PWM_VALUES_F64) {
            e = hipMemcpyAsync(vec, J.last, n * 8, hipMemcpyDeviceToHost, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        } else {
            std::vector<float> f(n);
            e = hipMemcpyAsync(f.data(), J.last, n * 4, hipMemcpyDeviceToHost, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
            for (int64_t i = 0; i < n; ++i) vec[i] = f[i];
        }
    }
    if (J.has_block) s->arena.release(J.off, J.bytes);
    s->jobs.erase(it);
    if (e != hipSuccess) return shp(s, e, "pwm_server_result");
    return PWM_OK;
}

extern "C" int pwm_server_stats(const pwm_server* s, int64_t* used_bytes, int32_t* resident, int32_t* queued,
                                int32_t* running, int64_t* finished) {
    if (!s) return sfail(nullptr, PWM_ERR_INVALID, "pwm_server_stats: NULL");
    if (used_bytes) *used_bytes = (int64_t)s->arena.used();
    if (resident) {
        int r = 0;
        for (const SrvMatrix& M : s->mats) r += M.resident;
        *resident = r;
    }
    if (queued) *queued = (int32_t)s->queue.size();
    if (running) *running = s->eng->active();
    if (finished) *finished = s->finished_total;
    return PWM_OK;
}

// Seconds spent in pwm_server_step's phases so far (admission including
// matrix uploads, waiting for the engine's chunks, retiring jobs) and the
// uploads made.
extern "C" int pwm_server_timing(const pwm_server* s, double* admit_s, double* run_s, double* retire_s,
                                 int64_t* uploads) {
    if (!s) return sfail(nullptr, PWM_ERR_INVALID, "pwm_server_timing: NULL");
    if (admit_s) *admit_s = s->t_admit;
    if (run_s) *run_s = s->t_run;
    if (retire_s) *retire_s = s->t_retire;
    if (uploads) *uploads = s->uploads;
    return PWM_OK;
}

// Ids of the jobs the last pwm_server_step retired (up to `cap`).
extern "C" int pwm_server_finished(const pwm_server* s, int64_t* ids, int32_t cap, int32_t* count) {
    if (!s || !count || (cap > 0 && !ids)) return sfail(nullptr, PWM_ERR_INVALID, "pwm_server_finished: bad argument");
    const int32_t n = (int32_t)std::min<size_t>(s->last_finished.size(), cap > 0 ? (size_t)cap : 0);
    for (int32_t i = 0; i < n; ++i) ids[i] = s->last_finished[i];
    *count = (int32_t)s->last_finished.size();
    return PWM_OK;
}

extern "C" int pwm_server_matrix_resident(const pwm_server* s, int32_t mat, int32_t* resident) {
    if (!s || mat < 0 || mat >= (int32_t)s->mats.size() || !resident)
        return sfail(nullptr, PWM_ERR_INVALID, "pwm_server_matrix_resident: bad argument");
    *resident = s->mats[mat].resident ? 1 : 0;
    return PWM_OK;
}
