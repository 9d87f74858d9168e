// Device-side descriptions of the CSB and BlockCOH layouts and the launchers
// of their kernels (k_blocked.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "k_common.hpp"

namespace blocked {

// Rows of y one workgroup accumulates in LDS (FP64: 32 KB, FP32: 32 KB).
inline int lds_rows(int vtype) { return vtype == 0 ? 4096 : 8192; }
constexpr int kLdsRowsF64 = 4096;
constexpr int kLdsRowsF32 = 8192;

// A contiguous run [start, end) of CSB entries: the part of one block that
// falls into one workgroup's row slab.
struct Piece {
    int32_t start, end;
};

struct CsbArgs {
    int64_t nor = 0;
    int32_t bits = 0;        // beta = 2^bits
    int64_t H = 0;           // block columns
    int32_t slabs = 1;       // slabs per block row
    int64_t slab_rows = 0;   // beta / slabs
    int32_t np = 1;          // pieces per (block, slab), a power of two
    int32_t np_shift = 0;    // log2(np)
    const uint32_t* ind = nullptr;
    const void* val = nullptr;
    const Piece* pieces = nullptr;  // [slab task][bc][piece]
    int64_t n_tasks = 0;            // slab tasks x column groups
    // Column groups: task t works on slab t / groups and block columns
    // [g*H/groups, (g+1)*H/groups) with g = t % groups.  MI355X dispatches
    // workgroup t to XCD t % 8, so with 8 groups each XCD gathers only from
    // its own eighth of x.  With groups > 1 the slab sums go to
    // ypart[g * nor + row] and k_csb_reduce adds the groups in order.
    int32_t groups = 1;
    void* ypart = nullptr;          // accumulator type, groups * nor
    int vtype = 0;
    // Slab-stream layout: entries of slab task t in block columns [bc0, bc1)
    // are ind/val[soff[t*H + bc0] .. soff[t*H + bc1]), each index packed as
    // (row in slab << col_bits) | column; ind/val extend 4 zero entries past
    // the last one.  slabs/slab_rows describe these slabs; pieces is unused.
    int32_t stream = 0;
    int32_t nt = 0;          // slab streams: non-temporal entry loads and partial-slab stores
    int32_t col_bits = 0;
    const int32_t* soff = nullptr;
    int32_t unroll = 2;  // groups of 4 entries per thread per step: 2 or 4
    // Wide slab streams (wide = 1): ind holds the column itself and srows the
    // row in the slab (u16), for column indices too wide to pack.
    int32_t wide = 0;
    const uint16_t* srows = nullptr;
    // Swept slabs (sweep = 1): n_tasks slabs [sstart[t], sstart[t+1]), one
    // persistent workgroup each with max_rows LDS rows; entries of slab t in
    // column window p are ind/val[soff[t*W + p] .. soff[t*W + p + 1]),
    // packed (row in slab << wb) | column in window.
    int32_t sweep = 0;
    int32_t wb = 0;
    int64_t W = 0;
    int32_t max_rows = 0;
    int32_t sync_every = 1;  // windows between workgroup barriers (a power of two)
    const int32_t* sstart = nullptr;
};

// Swept slabs: workgroups of the sweep kernel one CU holds at once with
// `rows` LDS rows of the value type's accumulator (0 if none fit).
int csb_sweep_blocks_per_cu(int vtype, int rows);


// One BlockCOH block (offsets are relative to its band's arrays).
struct CohBlock {
    int32_t entry;      // first entry (col_jump / data index)
    int32_t row_jump;   // first row_jump index
    int32_t brow, bcol; // block coordinates inside the band
};

// Workgroup size and entries per thread of the blocked kernels; a BlockCOH
// workgroup decodes kCohChunk consecutive entries of its band per step.
constexpr int kBlockedBS = 512;
constexpr int kCohIPT = 4;
constexpr int kCohChunk = kBlockedBS * kCohIPT;

struct CohBand {
    int32_t start, rows, beta_bits, first_block, blocks, entry, nnz;
    int32_t chunk_table;  // offset of the band's chunk -> first-block table
};

// Unit of work of the BlockCOH kernel: chunks [chunk0, chunk0+nchunks) of one
// band.  A band whose rows fit the LDS window and whose entries fit one task
// accumulates in LDS (atomic = 0); larger bands are cut into several tasks
// that add into y with global atomics (atomic = 1), each starting from the
// decode state (carry_s, carry_r) the build computed for its first entry.
constexpr int kCohTaskChunks = 16;
struct CohTask {
    int32_t band, chunk0, nchunks, atomic;
    int32_t carry_s, carry_r;
};

struct CohArgs {
    int64_t nor = 0, noc = 0;
    int64_t n_bands = 0;
    int64_t n_tasks = 0;
    const CohTask* tasks = nullptr;
    const CohBand* bands = nullptr;
    const CohBlock* blocks = nullptr;
    const int32_t* chunk_block = nullptr;  // per band: block holding entry c*kCohChunk, c = 0..chunks
    const uint16_t* col_jump = nullptr;
    const uint16_t* row_jump = nullptr;
    const void* val = nullptr;
    int vtype = 0;
    int32_t lds_rows = 0;  // LDS window of the kernel: rows of the largest LDS-accumulated band
    int32_t occ8 = 0;      // 1: the kernel instance capped to 8 waves per SIMD
};

// part: 0 = the whole SpMV, 1 = the slab kernel only, 2 = the column-group
// reduction only (nothing when groups = 1)
hipError_t launch_csb(const CsbArgs& a, const void* x, void* y, const pwmk::PowerArgs* pw, hipStream_t s,
                      int part = 0);
// workgroups of the CSB column-group reduction (= its power-method partials)
int csb_reduce_blocks(int64_t nor);
hipError_t launch_coh(const CohArgs& a, const int32_t* atomic_bands, int n_atomic, const void* x, void* y,
                      const pwmk::PowerArgs* pw, hipStream_t s);
// FP64 reduction of the per-workgroup power-method partials (one workgroup).
hipError_t launch_blocked_finalize(int vtype, int n_partials, const pwmk::PowerArgs& pw, hipStream_t s);

}  // namespace blocked
