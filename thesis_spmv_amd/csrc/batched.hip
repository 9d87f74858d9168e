// Batched power method: many instances of uneven size solved together on
// the iteration engine (engine.hpp).  All instances live in one device
// arena; every instance is one engine job, so a step of the batch is one
// tile launch over the tiles of every unfinished instance plus one finish
// launch, and chunks of steps replay as one hipGraph.  Instances are
// independent: each result is bit-identical to solving that instance alone.
//
// Values: FP32 or BF16 (exact values only) with FP32 vectors and per-row FP32
// accumulation, or FP64 throughout; norms and Rayleigh quotients in FP64.
// With the option bf16_when_exact an FP32 batch stores each instance whose
// values are all exact in BF16 (every Poisson instance: 4 and -1) as BF16
// and the others as FP32, in one launch per step.
//
// Tile schedule: 256 x items_per_thread merge items per tile, plain or
// persistent tile grid; with the option autotune the first solve after an
// upload times 6 steps of the whole batch for each candidate (re-planning the
// tiles in place) and keeps the fastest.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "api_internal.hpp"
#include "engine.hpp"

namespace {

inline uint64_t mix64h(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct BatchInst {
    int64_t n = 0, nnz = 0;
    int kind = 0;           // 0: CRS from the host, 1: Poisson generated on the device
    int vt16 = 0;           // BF16 values in a bf16_when_exact batch
    int64_t pm = 0, pn = 0;  // Poisson grid
    std::vector<int32_t> rs, ci;
    std::vector<double> val;
    // arena offsets
    size_t o_rs = 0, o_ci = 0, o_val = 0, o_tiles = 0, o_part = 0, o_carry = 0, o_vec = 0;
    int32_t n_tiles = 0;
    int32_t cap_tiles = 0;  // tile slots in the arena (the autotuner re-plans within them)
    std::vector<pwmk::ETile> tiles;  // host copy of the plan
    int slot = -1;
};

// The autotuner's tile sizes: 256 x ipt merge items for ipt in kTuneIpts.
constexpr int kTuneIpts[] = {8, 16};
constexpr int kTuneMinIpt = 8;
// Steps of every instance the autotuner times per candidate schedule.
constexpr int64_t kTuneSteps = 6;

}  // namespace

struct pwm_batch {
    int vtype = PWM_VALUES_F32;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<BatchInst> inst;
    int64_t rows = 0, nnz = 0;
    bool uploaded = false;
    int planned_ipt = 0;
    int ipt = 16;
    int chunk = 32;
    int use_graph = 1;
    int persistent = 0;
    int bf16_exact = 0;     // FP32 batch: BF16 storage for instances exact in BF16
    int autotune = 0;       // time the tile schedules at the first solve after an upload
    bool tuned = false;     // the schedule in ipt / persistent was chosen by the autotuner
    std::vector<float> tune_ms;  // ms per step of (ipt 8, 16) x (plain, persistent); -1 where not timed
    pwm_impl::Arena arena;
    pwm_impl::Engine* eng = nullptr;
    std::vector<double> lambda, ynorm;
    std::vector<int64_t> iters;
    std::vector<int32_t> where;
};

static int bfail(pwm_batch* b, int code, const std::string& m) {
    if (b)
        b->err = m;
    else
        pwm_impl::set_global_error(m);
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
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    delete b->eng;
    b->eng = nullptr;
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PWM_OK;
}

extern "C" const char* pwm_batch_last_error(const pwm_batch* b) { return b ? b->err.c_str() : pwm_last_error(nullptr); }

extern "C" int pwm_batch_add_crs(pwm_batch* b, int64_t n, const int32_t* rs, const int32_t* ci, const double* vals,
                                 int32_t* id) {
    if (!b || !rs || n <= 0 || n > INT32_MAX - 1) return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_crs: bad argument");
    const int64_t nnz = rs[n];
    if (rs[0] != 0 || nnz < 0 || (nnz > 0 && (!ci || !vals)))
        return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_crs: bad row_start or NULL arrays");
    if (n + nnz > INT32_MAX) return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_crs: rows + nonzeros must fit int32");
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
    BatchInst I;
    I.n = n;
    I.nnz = nnz;
    I.rs.assign(rs, rs + n + 1);
    I.ci.assign(ci, ci + nnz);
    I.val.assign(vals, vals + nnz);
    b->inst.push_back(std::move(I));
    b->where.clear();  // results of an earlier solve do not cover the new instance
    b->rows += n;
    b->nnz += nnz;
    b->uploaded = false;
    b->tuned = false;  // the next solve uploads and, with autotune, times the schedules again
    return PWM_OK;
}

// Poisson instances are generated on the device at upload (Util/Poisson.hpp:32-83).
extern "C" int pwm_batch_add_poisson(pwm_batch* b, int64_t m, int64_t n, int32_t* id) {
    if (!b || m <= 0 || n <= 0 || 6 * m * n > INT32_MAX) return bfail(b, PWM_ERR_INVALID, "pwm_batch_add_poisson: bad size");
    if (id) *id = (int32_t)b->inst.size();
    BatchInst I;
    I.kind = 1;
    I.pm = m;
    I.pn = n;
    I.n = m * n;
    I.nnz = n * (m + 2 * (m - 1)) + 2 * (n - 1) * m;
    b->rows += I.n;
    b->nnz += I.nnz;
    b->inst.push_back(std::move(I));
    b->where.clear();  // results of an earlier solve do not cover the new instance
    b->uploaded = false;
    b->tuned = false;
    return PWM_OK;
}

extern "C" int pwm_batch_set_option(pwm_batch* b, const char* key, int64_t value) {
    if (!b || !key) return bfail(b, PWM_ERR_INVALID, "pwm_batch_set_option: NULL");
    const std::string k(key);
    if (k == "items_per_thread") {
        if (value != 4 && value != 8 && value != 16)
            return bfail(b, PWM_ERR_INVALID, "pwm_batch_set_option: items_per_thread must be 4, 8 or 16");
        b->ipt = (int)value;
        return PWM_OK;
    }
    if (k == "chunk") {
        if (value < 1 || value > 4096) return bfail(b, PWM_ERR_INVALID, "pwm_batch_set_option: chunk must be in [1, 4096]");
        b->chunk = (int)value;
        return PWM_OK;
    }
    if (k == "graph") {
        b->use_graph = value != 0;
        return PWM_OK;
    }
    if (k == "persistent") {
        b->persistent = value != 0;
        return PWM_OK;
    }
    if (k == "autotune") {
        if (b->eng && b->eng->active() != 0) return bfail(b, PWM_ERR_STATE, "pwm_batch_set_option: solve in flight");
        b->autotune = value != 0;
        b->tuned = false;
        b->uploaded = false;  // the arena then keeps room for the smallest tiles
        return PWM_OK;
    }
    if (k == "bf16_when_exact") {
        if (value != 0 && b->vtype != PWM_VALUES_F32)
            return bfail(b, PWM_ERR_INVALID, "pwm_batch_set_option: bf16_when_exact needs an FP32 batch");
        if (b->eng && b->eng->active() != 0) return bfail(b, PWM_ERR_STATE, "pwm_batch_set_option: solve in flight");
        b->bf16_exact = value != 0;
        b->uploaded = false;
        delete b->eng;  // the engine's value type changes with it
        b->eng = nullptr;
        return PWM_OK;
    }
    return bfail(b, PWM_ERR_INVALID, "pwm_batch_set_option: unknown option " + k);
}

static int batch_upload(pwm_batch* b) {
    if (b->uploaded && b->planned_ipt == b->ipt) return PWM_OK;
    const size_t xs = pwm_impl::vector_size(b->vtype);
    const size_t acc = b->vtype == PWM_VALUES_F64 ? 8 : 4;
    using pwm_impl::Arena;
    size_t total = 0;
    for (BatchInst& I : b->inst) {
        I.vt16 = b->bf16_exact && (I.kind == 1 || pwm_impl::exact_in_bf16(I.val.data(), (int64_t)I.val.size()));
        const size_t vs = I.vt16 ? 2 : pwm_impl::value_size(b->vtype);
        I.n_tiles = pwmk::engine_tiles_for(I.n, I.nnz, b->ipt);
        I.cap_tiles = b->autotune ? std::max(I.n_tiles, pwmk::engine_tiles_for(I.n, I.nnz, kTuneMinIpt)) : I.n_tiles;
        const int64_t zp = I.nnz + pwm_impl::kPadEntries;  // tiles read aligned groups of 4
        total += Arena::round((I.n + 1) * 4) + Arena::round(zp * 4) + Arena::round(zp * vs) +
                 Arena::round((size_t)I.cap_tiles * sizeof(pwmk::ETile)) + Arena::round((size_t)I.cap_tiles * 24) +
                 Arena::round((size_t)I.cap_tiles * acc) + Arena::round(2 * I.n * xs);
    }
    hipError_t e = b->arena.reserve(total);
    if (e != hipSuccess) return bhip(b, e, "batch allocation");
    Arena& A = b->arena;
    for (BatchInst& I : b->inst) {
        const int vt = I.vt16 ? PWM_VALUES_BF16 : b->vtype;
        const size_t vs = pwm_impl::value_size(vt);
        const int64_t zp = I.nnz + pwm_impl::kPadEntries;
        bool ok = A.alloc((I.n + 1) * 4, &I.o_rs) && A.alloc(zp * 4, &I.o_ci) && A.alloc(zp * vs, &I.o_val) &&
                  A.alloc((size_t)I.cap_tiles * sizeof(pwmk::ETile), &I.o_tiles) &&
                  A.alloc((size_t)I.cap_tiles * 24, &I.o_part) && A.alloc((size_t)I.cap_tiles * acc, &I.o_carry) &&
                  A.alloc(2 * I.n * xs, &I.o_vec);
        if (!ok) return bfail(b, PWM_ERR_NOMEM, "batch arena exhausted");
        int32_t* rs = (int32_t*)A.at(I.o_rs);
        e = hipMemsetAsync((char*)A.at(I.o_ci) + I.nnz * 4, 0, pwm_impl::kPadEntries * 4, b->stream);
        if (e == hipSuccess)
            e = hipMemsetAsync((char*)A.at(I.o_val) + I.nnz * vs, 0, pwm_impl::kPadEntries * vs, b->stream);
        if (e != hipSuccess) return bhip(b, e, "batch padding");
        if (I.kind == 0) {
            e = pwm_impl::upload_staged(rs, I.rs.data(), (I.n + 1) * 4, b->stream);
            if (e == hipSuccess) e = pwm_impl::upload_staged(A.at(I.o_ci), I.ci.data(), I.nnz * 4, b->stream);
            if (e == hipSuccess) e = pwm_impl::upload_values_staged(A.at(I.o_val), I.val.data(), I.nnz, vt, b->stream);
        } else {
            e = pwmk::launch_fill_poisson(I.pm, I.pn, rs, (int32_t*)A.at(I.o_ci), A.at(I.o_val), vt, b->stream);
        }
        if (e == hipSuccess)
            e = pwmk::engine_plan_tiles(rs, I.n, I.nnz, b->ipt, (pwmk::ETile*)A.at(I.o_tiles), b->stream);
        I.tiles.resize(I.n_tiles);
        if (e == hipSuccess)
            e = hipMemcpyAsync(I.tiles.data(), A.at(I.o_tiles), (size_t)I.n_tiles * sizeof(pwmk::ETile),
                               hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) return bhip(b, e, "batch upload");
    }
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return bhip(b, e, "batch upload");
    b->uploaded = true;
    b->planned_ipt = b->ipt;
    b->tuned = false;
    return PWM_OK;
}

// Re-plans every instance's tiles for `ipt` items per thread inside the tile
// slots the upload reserved (the autotuner's candidates).
static int batch_replan(pwm_batch* b, int ipt) {
    hipError_t e = hipSuccess;
    for (BatchInst& I : b->inst) {
        const int32_t nt = pwmk::engine_tiles_for(I.n, I.nnz, ipt);
        if (nt > I.cap_tiles) return bfail(b, PWM_ERR_STATE, "batch re-plan: tiles exceed the reserved slots");
        I.n_tiles = nt;
        e = pwmk::engine_plan_tiles((const int32_t*)b->arena.at(I.o_rs), I.n, I.nnz, ipt,
                                    (pwmk::ETile*)b->arena.at(I.o_tiles), b->stream);
        I.tiles.resize(nt);
        if (e == hipSuccess)
            e = hipMemcpyAsync(I.tiles.data(), b->arena.at(I.o_tiles), (size_t)nt * sizeof(pwmk::ETile),
                               hipMemcpyDeviceToHost, b->stream);
        if (e != hipSuccess) break;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bhip(b, e, "batch re-plan");
    b->planned_ipt = ipt;
    return PWM_OK;
}

// Hands every instance to the engine as a job of `max_it` steps (tol < 0: no
// stop test).
static void batch_add_jobs(pwm_batch* b, int64_t max_it, double tol) {
    const size_t xs = pwm_impl::vector_size(b->vtype);
    pwm_impl::Arena& A = b->arena;
    for (BatchInst& I : b->inst) {
        pwmk::EJob j;
        memset(&j, 0, sizeof(j));
        j.rs = (const int32_t*)A.at(I.o_rs);
        j.ci = (const int32_t*)A.at(I.o_ci);
        j.val = A.at(I.o_val);
        j.buf0 = A.at(I.o_vec);
        j.buf1 = (char*)A.at(I.o_vec) + I.n * xs;
        j.tiles = (const pwmk::ETile*)A.at(I.o_tiles);
        j.partials = (double*)A.at(I.o_part);
        j.carry = A.at(I.o_carry);
        j.n = I.n;
        j.nnz = I.nnz;
        j.max_it = max_it;
        j.tol = tol;
        j.n_tiles = I.n_tiles;
        j.vt16 = I.vt16;
        I.slot = b->eng->add(j, I.tiles.data());
    }
}

static void batch_release_jobs(pwm_batch* b) {
    for (BatchInst& I : b->inst) {
        b->eng->release(I.slot);
        I.slot = -1;
    }
}

// Chooses the tile schedule: for 256 x ipt merge items per tile (ipt in
// kTuneIpts) and plain or persistent tile grids, runs kTuneSteps steps of
// every instance twice (the first run builds the lists and the graph) and
// times the second; the fastest candidate's ipt and persistent flag are kept
// in the batch and its tiles stay planned.  The timed steps start from the
// seed-0 start vectors; the solve that follows resets them.
static int batch_autotune(pwm_batch* b) {
    pwm_impl::Engine& E = *b->eng;
    struct Cand {
        int ipt, persistent;
    };
    std::vector<Cand> cands;
    for (int ipt : kTuneIpts) {
        cands.push_back({ipt, 0});
        if (!b->bf16_exact) cands.push_back({ipt, 1});  // the mixed-storage engine has no persistent grid
    }
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < b->inst.size() && e == hipSuccess; ++i)
        e = pwmk::engine_start_vector(b->arena.at(b->inst[i].o_vec), b->inst[i].n, b->vtype, 0, (uint64_t)i, b->stream);
    if (e != hipSuccess) return bhip(b, e, "batch autotune start vectors");
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return bhip(b, e, "batch autotune");
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return bhip(b, e, "batch autotune");
    }
    b->tune_ms.assign(2 * (sizeof(kTuneIpts) / sizeof(kTuneIpts[0])), -1.0f);
    std::vector<float> ms_of(cands.size(), 0.0f);
    int rc = PWM_OK, best = -1;
    std::vector<int> fin;
    std::string err;
    for (size_t c = 0; c < cands.size() && rc == PWM_OK; ++c) {
        if (b->planned_ipt != cands[c].ipt && (rc = batch_replan(b, cands[c].ipt)) != PWM_OK) break;
        if ((rc = E.set_ipt(cands[c].ipt)) != PWM_OK) {
            rc = bfail(b, rc, "batch autotune: items_per_thread");
            break;
        }
        E.persistent = cands[c].persistent != 0;
        batch_add_jobs(b, 2 * kTuneSteps, -1.0);
        float ms = 0;
        if ((rc = E.run((int)kTuneSteps, fin, err)) == PWM_OK) {
            (void)hipEventRecord(e0, b->stream);
            rc = E.run((int)kTuneSteps, fin, err);
            (void)hipEventRecord(e1, b->stream);
            if (rc == PWM_OK && (e = hipEventSynchronize(e1)) == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (rc == PWM_OK && e != hipSuccess) rc = bhip(b, e, "batch autotune timing");
        } else {
            rc = bfail(b, rc, err);
        }
        batch_release_jobs(b);
        if (rc != PWM_OK) break;
        ms_of[c] = ms / (float)kTuneSteps;
        b->tune_ms[(c / (b->bf16_exact ? 1 : 2)) * 2 + cands[c].persistent] = ms_of[c];
        if (best < 0 || ms_of[c] < ms_of[best]) best = (int)c;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != PWM_OK) return rc;
    b->ipt = cands[best].ipt;
    b->persistent = cands[best].persistent;
    if (b->planned_ipt != b->ipt && (rc = batch_replan(b, b->ipt)) != PWM_OK) return rc;
    b->tuned = true;
    return PWM_OK;
}

extern "C" int pwm_batch_solve(pwm_batch* b, int64_t max_it, double tol, uint64_t seed, double* lambdas,
                               int64_t* iterations, double* ms) {
    if (!b || max_it <= 0) return bfail(b, PWM_ERR_INVALID, "pwm_batch_solve: bad argument");
    if (b->inst.empty()) return bfail(b, PWM_ERR_STATE, "pwm_batch_solve: no instances");
    int rc = batch_upload(b);
    if (rc != PWM_OK) return rc;
    const int ni = (int)b->inst.size();
    if (b->eng && b->eng->max_slots() < ni) {
        delete b->eng;
        b->eng = nullptr;
    }
    if (!b->eng) {
        b->eng = new pwm_impl::Engine();
        std::string err;
        const int evt = b->bf16_exact ? pwmk::VT_MIXED16 : b->vtype;
        if ((rc = b->eng->init(evt, ni, b->stream, err)) != PWM_OK) {
            delete b->eng;
            b->eng = nullptr;
            return bfail(b, rc, err);
        }
    } else if (b->eng->active() != 0) {
        return bfail(b, PWM_ERR_STATE, "pwm_batch_solve: engine busy");
    }
    pwm_impl::Engine& E = *b->eng;
    E.chunk = b->chunk;
    E.use_graph = b->use_graph != 0;
    if (b->autotune && !b->tuned && (rc = batch_autotune(b)) != PWM_OK) return rc;
    if (E.ipt() != b->ipt && (rc = E.set_ipt(b->ipt)) != PWM_OK) return bfail(b, rc, "pwm_batch_solve: items_per_thread");
    E.persistent = b->persistent != 0;
    pwm_impl::Arena& A = b->arena;
    hipError_t e = hipSuccess;
    for (int i = 0; i < ni && e == hipSuccess; ++i) {
        BatchInst& I = b->inst[i];
        e = pwmk::engine_start_vector(A.at(I.o_vec), I.n, b->vtype, seed, (uint64_t)i, b->stream);
    }
    if (e != hipSuccess) return bhip(b, e, "batch start vectors");
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, b->stream);
    batch_add_jobs(b, max_it, tol);
    std::vector<int> fin;
    std::string err;
    const int64_t max_runs = (max_it + E.chunk - 1) / E.chunk + 1;
    for (int64_t r = 0; E.active() > 0; ++r) {
        if (r > max_runs) {
            rc = bfail(b, PWM_ERR_STATE, "pwm_batch_solve: jobs did not stop after max_it steps");
            break;
        }
        if ((rc = E.run(E.chunk, fin, err)) != PWM_OK) {
            bfail(b, rc, err);
            break;
        }
    }
    if (rc == PWM_OK)
        for (int i = 0; i < ni && e == hipSuccess; ++i) {
            const BatchInst& I = b->inst[i];
            e = pwmk::engine_normalize(E.last_buffer(I.slot), I.n, b->vtype, 1.0 / sqrt(E.state(I.slot).yy), b->stream);
        }
    (void)hipEventRecord(e1, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->lambda.assign(ni, 0.0);
    b->ynorm.assign(ni, 0.0);
    b->iters.assign(ni, 0);
    b->where.assign(ni, 0);
    for (int i = 0; i < ni; ++i) {
        const BatchInst& I = b->inst[i];
        if (I.slot < 0) continue;
        const pwmk::EState& S = E.state(I.slot);
        b->lambda[i] = S.lambda;
        b->ynorm[i] = sqrt(S.yy);
        b->iters[i] = S.it;
        b->where[i] = (int32_t)(S.it & 1);
        E.release(I.slot);
    }
    if (rc != PWM_OK) {
        b->where.clear();
        return rc;
    }
    if (e != hipSuccess) {
        b->where.clear();
        return bhip(b, e, "batched power method");
    }
    if (lambdas) memcpy(lambdas, b->lambda.data(), ni * 8);
    if (iterations) memcpy(iterations, b->iters.data(), ni * 8);
    if (ms) *ms = t;
    return PWM_OK;
}

// The start vector the solver used for instance `id` with `seed` (host
// restatement of k_eng_start with stream id = instance id).
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
    if (b->where.empty() || id >= (int32_t)b->where.size())
        return bfail(b, PWM_ERR_STATE, "pwm_batch_vector: solve first");
    const BatchInst& I = b->inst[id];
    const size_t xs = pwm_impl::vector_size(b->vtype);
    const char* src = (const char*)b->arena.at(I.o_vec) + (b->where[id] ? I.n * xs : 0);
    hipError_t e;
    if (b->vtype == PWM_VALUES_F64) {
        e = hipMemcpy(out, src, I.n * 8, hipMemcpyDeviceToHost);
    } else {
        std::vector<float> f(I.n);
        e = hipMemcpy(f.data(), src, I.n * 4, hipMemcpyDeviceToHost);
        for (int64_t i = 0; i < I.n; ++i) out[i] = f[i];
    }
    if (e != hipSuccess) return bhip(b, e, "pwm_batch_vector");
    if (ynorm) *ynorm = b->ynorm[id];
    return PWM_OK;
}

extern "C" int pwm_batch_size(const pwm_batch* b, int32_t* instances, int64_t* rows, int64_t* nnz) {
    if (!b) return bfail(nullptr, PWM_ERR_INVALID, "pwm_batch_size: NULL");
    if (instances) *instances = (int32_t)b->inst.size();
    if (rows) *rows = b->rows;
    if (nnz) *nnz = b->nnz;
    return PWM_OK;
}

// Engine counters of the last solve: launches, graph instantiations, work
// list rebuilds.
extern "C" int pwm_batch_stats(const pwm_batch* b, int64_t* launches, int64_t* graphs, int64_t* lists) {
    if (!b) return bfail(nullptr, PWM_ERR_INVALID, "pwm_batch_stats: NULL");
    if (launches) *launches = b->eng ? b->eng->launches : 0;
    if (graphs) *graphs = b->eng ? b->eng->graph_builds : 0;
    if (lists) *lists = b->eng ? b->eng->list_builds : 0;
    return PWM_OK;
}

// The tile schedule the next solve uses and, when the autotuner chose it, the
// ms per step it timed for each candidate.
extern "C" int pwm_batch_schedule(const pwm_batch* b, int32_t* items_per_thread, int32_t* persistent, int32_t* tuned,
                                  float* candidate_ms) {
    if (!b) return bfail(nullptr, PWM_ERR_INVALID, "pwm_batch_schedule: NULL");
    if (items_per_thread) *items_per_thread = b->ipt;
    if (persistent) *persistent = b->persistent;
    if (tuned) *tuned = b->tuned ? 1 : 0;
    if (candidate_ms)
        for (int c = 0; c < 4; ++c) candidate_ms[c] = b->tuned ? b->tune_ms[c] : -1.0f;
    return PWM_OK;
}
