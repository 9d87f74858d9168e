// Iteration engine shared by the batched power method (pwm_batch_*) and the
// request server (pwm_server_*).
//
// A job is one power iteration (the reference's powerIteration,
// Matrix/CRS.hpp:118-130, with the Rayleigh-quotient stop test of
// pwm_power_method_device) on one matrix: step k computes
//   y = (A x) / |x|      (x is the previous, not yet normalised, iterate)
//   |y|^2, x.y, x.x in FP64;  lambda_k = x.y / x.x
// and the job stops when |lambda_k - lambda_{k-1}| <= tol |lambda_k| (k >= 1,
// tol >= 0) or after max_it steps.  A fixed-length SpMV chain is a job with
// tol < 0.
//
// Matrices are cut once into merge-path tiles of 256*ipt items aligned to
// the matrix's own start, so a job's arithmetic never depends on the other
// jobs in flight.  Active jobs sit in slots; one step of every active job is
// two launches:
//   k_eng_tiles   one workgroup per (slot, tile) of the work list;
//   k_eng_finish  one workgroup per active slot: carry fix-up, FP64
//                 reduction of the tile partials in a fixed order, lambda,
//                 stop test, 1/|y| for the next step.
// The per-job state (step, stop flag, scale, lambda, |y|^2) lives on the
// device and both kernels read the list lengths from device memory, so every
// step of a chunk has the same launch arguments: a chunk of steps is
// captured once in a hipGraph (grids sized to the list capacities) and
// replayed until a list outgrows its capacity or shrinks below a quarter.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "k_common.hpp"

namespace pwmk {

struct ETile {  // rows [a0,a1), nonzeros [b0,b1), matrix-relative
    int32_t a0, a1, b0, b1;
    int32_t stream;  // 1: row-aligned, every row <= kStreamRow nonzeros
    int32_t rs_end;  // row_start[a1]
    int32_t pad[2];
};

// Tiles hold at most 256*ipt merge items.  The planner cuts the merge path
// every 3/4 of that and moves a cut back to the start of its row when the row
// has at most 1/4 of it, so short rows are never split.
constexpr int kStreamRow = 64;

struct EJob {  // device view of one slot
    const int32_t* rs;
    const int32_t* ci;
    const void* val;
    void* buf0;  // step k reads buf[k&1] and writes buf[(k+1)&1]
    void* buf1;
    const ETile* tiles;
    double* partials;  // 3 per tile
    void* carry;       // one accumulator per tile
    int64_t n;
    int64_t nnz;
    int64_t max_it;
    double tol;
    int32_t n_tiles;
    int32_t vt16;  // mixed engines: 1 when the values are BF16 (else FP32)
};

struct EState {  // per slot, written by k_eng_finish
    int64_t it;
    double scale;  // 1/|x| of the step's input (1 at step 0)
    double lambda;
    double yy;     // |y|^2 of the last step
    int32_t done;
    int32_t pad;
};

// One entry of the work list: everything a tile needs to start loading
// (64 B, one vector load).  slot < 0: padding.
struct EWork {
    const int32_t* rs;
    const int32_t* ci;
    const void* val;
    int32_t slot, tile;
    int32_t a0, a1, b0, b1;
    int32_t stream;
    int32_t vt16;  // the job's vt16
    int32_t pad[2];
};

struct ECtl {  // list lengths, read by the kernels
    int32_t n_work;
    int32_t n_slots;
    int32_t pad[2];
};

struct EArgs {
    const EJob* jobs;
    EState* st;
    const EWork* work;
    const int32_t* slots;  // active slots
    const ECtl* ctl;
};

constexpr int kEngBS = 256;
// Engine value type for FP32 vectors with per-job FP32 or BF16 values
// (EJob::vt16), beside VT_F64 / VT_F32 / VT_BF16.
constexpr int VT_MIXED16 = 3;

int32_t engine_tiles_for(int64_t n, int64_t nnz, int ipt);
hipError_t engine_plan_tiles(const int32_t* rs, int64_t n, int64_t nnz, int ipt, ETile* tiles, hipStream_t s);
hipError_t engine_start_vector(void* x, int64_t n, int vtype, uint64_t seed, uint64_t stream_id, hipStream_t s);
hipError_t engine_normalize(void* v, int64_t n, int vtype, double inv_norm, hipStream_t s);
// grid_persist > 0: persistent tile kernel with that many workgroups.
hipError_t engine_step(const EArgs& a, int vtype, int ipt, int grid_work, int grid_persist, int grid_slots,
                       hipStream_t s);
int engine_resident_blocks(int vtype, int ipt);

}  // namespace pwmk

namespace pwm_impl {

// First-fit device sub-allocator over one hipMalloc'd region (256-byte
// granules, neighbouring free ranges merged on release).
class Arena {
   public:
    ~Arena();
    hipError_t reserve(size_t bytes);
    void reset();  // forget every allocation (the region stays)
    bool alloc(size_t bytes, size_t* off);
    void release(size_t off, size_t bytes);
    void* at(size_t off) const { return (char*)base_ + off; }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    static size_t round(size_t b) { return (b + 255) & ~(size_t)255; }

   private:
    void* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
};

// Slots, work lists and chunked execution.
class Engine {
   public:
    ~Engine();
    int init(int vtype, int max_slots, hipStream_t stream, std::string& err);
    int vtype() const { return vtype_; }
    int ipt() const { return ipt_; }
    int max_slots() const { return max_slots_; }
    int set_ipt(int v);  // only while no job is active
    int chunk = 32;      // steps per replay
    bool use_graph = true;
    bool persistent = false;  // persistent, prefetching tile kernel (slower here: fewer waves)

    // Claim a free slot for job `j` (state reset to step 0); -1 if none.
    // `host_tiles` (j.n_tiles entries, the host copy of j.tiles) must stay
    // valid until the job is released.
    int add(const pwmk::EJob& j, const pwmk::ETile* host_tiles);
    int active() const { return n_active_; }
    // Run `steps` steps of every active job.  On return `finished` lists
    // the slots whose jobs stopped; their state is in state(slot) and the
    // final iterate (not yet normalised) in buffer last_buffer(slot).
    int run(int steps, std::vector<int>& finished, std::string& err);
    // run() in two halves: launch() queues the steps and the state download
    // and returns at once; collect() waits for them and lists the finished
    // slots.  Between the two the host may do other work, but must not call
    // add(), release(), set_ipt() or launch() (the download is writing the
    // state mirror).
    int launch(int steps, std::string& err);
    int collect(std::vector<int>& finished, std::string& err);
    bool pending() const { return pending_; }
    const pwmk::EState& state(int slot) const { return h_state_[slot]; }
    const pwmk::EJob& job(int slot) const { return h_jobs_[slot]; }
    void* last_buffer(int slot) const;
    void release(int slot);  // after the caller has read the results
    int64_t launches = 0, graph_builds = 0, list_builds = 0;

   private:
    int build_lists(std::string& err);
    int ensure_capacity(std::string& err);
    int vtype_ = 0, ipt_ = 16, max_slots_ = 0;
    hipStream_t stream_ = nullptr;
    pwmk::EJob* h_jobs_ = nullptr;      // pinned mirrors
    pwmk::EState* h_state_ = nullptr;
    pwmk::ECtl* h_ctl_ = nullptr;
    pwmk::EWork* h_work_ = nullptr;
    std::vector<const pwmk::ETile*> h_tiles_;
    int32_t* h_slots_ = nullptr;
    void* d_jobs_ = nullptr;
    void* d_state_ = nullptr;
    void* d_ctl_ = nullptr;
    void* d_work_ = nullptr;
    void* d_slots_ = nullptr;
    size_t work_cap_alloc_ = 0;
    std::vector<char> in_use_, running_, fresh_;
    int n_active_ = 0;
    bool pending_ = false;  // launched steps not yet collected
    bool lists_dirty_ = true;
    int n_work_ = 0, n_slots_ = 0, n_stale_work_ = 0;
    int grid_work_ = 0, grid_slots_ = 0;
    hipGraphExec_t graph_ = nullptr;
    int graph_steps_ = 0, graph_gw_ = 0, graph_gs_ = 0, graph_gp_ = 0;
    int resident_ = 0, resident_ipt_ = 0;
};

}  // namespace pwm_impl
