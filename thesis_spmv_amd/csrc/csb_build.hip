// CSB slab-stream layout built on the device (the stored order of
// SOA_Implementations/CSB.hpp:464-567, regrouped by row slab).
//
// Entry k of row r, column c gets the key
//     cell << 2b | zkey(r mod beta, c mod beta),   cell = (br*S + s)*H + bc,
// with br = r >> b, bc = c >> b, s = (r mod beta) >> rb the slab inside the
// block row, and zkey the Z-Morton interleave (row bit above column bit,
// CSB.hpp:147-173).  A stable radix sort of (key, k) puts every slab's entries
// together, block columns ascending, each block's entries in Z-Morton order
// and duplicates in CRS order -- the order the host build gives them.  The
// packed indices and values are then gathered through the sorted positions,
// each (slab, block column) run start is found by binary search, and the
// reference's blk_ptr follows from the run lengths.
//
// The slab-piece layout (the stored order itself, CSB.hpp:464-567) sorts the
// same keys with one slab per block row (cell = br*H + bc), packs each index
// as (row in block << 16) | column in block, and finds each block's pieces by
// binary search for its Z-order prefixes.
//
// The swept-slab layout (csb_sweep_device) cuts the rows into nnz-balanced
// slabs on the host (one per CU), keys entry k of slab t, column c with
// cell = t*W + (c >> w) (W windows of 2^w columns) and, in column order (the
// default), cell << w | (c mod 2^w); sorts stably (equal keys keep CRS
// order) and packs each index as (row in slab << w) | (c mod 2^w).  In
// column order every (slab, window) run is padded to whole blocks of 256
// entries (index 0xffffffff, value 0) and each block stores its sorted entry
// j*64 + L at 4L + j, so that lane L, reading 4 consecutive stored entries,
// holds sorted entries L, 64 + L, 128 + L and 192 + L: each gather
// instruction of a wave reads 64 consecutive columns.  soff gives every
// run's stored start.
#include <rocprim/device/device_radix_sort.hpp>

#include "api_internal.hpp"
#include "blocked.hpp"
#include "kernels.hpp"

namespace pwmk {

// bit i of a 16-bit value to bit 2i, and back
__device__ __forceinline__ uint32_t spread16(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
__device__ __forceinline__ uint32_t compact16(uint32_t v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return v;
}

struct CsbKeyArgs {
    int64_t nor, nnz, S, H;
    int bits, rb, cb;
    int local;  // 1: index = (row in block << 16) | column in block
};

__global__ void k_csb_keys(CsbKeyArgs a, const int32_t* __restrict__ rs, const int32_t* __restrict__ ci,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ pos) {
    const uint32_t bmask = (uint32_t)(((uint64_t)1 << a.bits) - 1);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.nnz; k += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = a.nor - 1;  // the last row starting at or before k
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (rs[mid] <= k)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint32_t c = (uint32_t)ci[k];
        const uint32_t lrow = (uint32_t)lo & bmask, lcol = c & bmask;
        const uint64_t cell = ((uint64_t)(lo >> a.bits) * a.S + (lrow >> a.rb)) * a.H + (c >> a.bits);
        keys[k] = (cell << (2 * a.bits)) | (((uint64_t)spread16(lrow) << 1) | spread16(lcol));
        pos[k] = (uint32_t)k;
    }
}

// srow != NULL: wide streams (sind = column, srow = row in slab)
template <typename E>
__global__ void k_csb_gather(CsbKeyArgs a, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ pos,
                             const E* __restrict__ val, uint32_t* __restrict__ sind, uint16_t* __restrict__ srow,
                             E* __restrict__ sdat) {
    const uint64_t zmask = (a.bits == 0) ? 0 : ((((uint64_t)1 << (2 * a.bits)) - 1));
    const uint32_t smask = (uint32_t)(((uint64_t)1 << a.rb) - 1);
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[p];
        const uint32_t zk = (uint32_t)(key & zmask);
        const uint64_t bc = (key >> (2 * a.bits)) % (uint64_t)a.H;
        const uint32_t lrow = compact16(zk >> 1), lcol = compact16(zk);
        const uint32_t col = (uint32_t)(bc << a.bits) + lcol;
        if (a.local) {
            sind[p] = (lrow << 16) | lcol;
        } else if (srow) {
            sind[p] = col;
            srow[p] = (uint16_t)(lrow & smask);
        } else {
            sind[p] = ((lrow & smask) << a.cb) | col;
        }
        sdat[p] = val[pos[p]];
    }
}

// soff[cell] = number of sorted keys below cell << shift, cell < cells
__global__ void k_csb_soff(const uint64_t* __restrict__ keys, int64_t nnz, int64_t cells, int shift,
                           int32_t* __restrict__ soff) {
    for (int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; cell < cells;
         cell += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t t = (uint64_t)cell << shift;
        int64_t lo = 0, hi = nnz;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < t)
                lo = mid + 1;
            else
                hi = mid;
        }
        soff[cell] = (int32_t)lo;
    }
}

// Swept slabs: entry k of row r, column c gets the key t * W + (c >> wb),
// t the slab holding row r; a stable sort keeps CRS order (rows, then
// columns, ascending) inside every (slab, window) run.
struct SweepKeyArgs {
    int64_t nor, nnz, T, W;
    int wb;
    int order;  // 1: the column in the window joins the key (column order inside a run)
    const int32_t* roff = nullptr;  // order 1: sorted run starts
    const int32_t* poff = nullptr;  // order 1: stored (padded) run starts
};

__device__ __forceinline__ int64_t last_at_or_below(const int32_t* __restrict__ a, int64_t n, int64_t v) {
    int64_t lo = 0, hi = n - 1;  // the last i in [0, n) with a[i] <= v (a[0] <= v)
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (a[mid] <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void k_sweep_keys(SweepKeyArgs a, const int32_t* __restrict__ rs, const int32_t* __restrict__ ci,
                             const int32_t* __restrict__ sstart, uint64_t* __restrict__ keys,
                             uint32_t* __restrict__ pos) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.nnz; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = last_at_or_below(rs, a.nor, k);
        const int64_t t = last_at_or_below(sstart, a.T, row);
        const uint64_t cell = (uint64_t)t * (uint64_t)a.W + (uint64_t)((uint32_t)ci[k] >> a.wb);
        keys[k] = a.order ? (cell << a.wb) | ((uint32_t)ci[k] & (((uint32_t)1 << a.wb) - 1)) : cell;
        pos[k] = (uint32_t)k;
    }
}

template <typename E>
__global__ void k_sweep_gather(SweepKeyArgs a, const int32_t* __restrict__ rs, const int32_t* __restrict__ ci,
                               const int32_t* __restrict__ sstart, const uint64_t* __restrict__ keys,
                               const uint32_t* __restrict__ pos, const E* __restrict__ val,
                               uint32_t* __restrict__ sind, E* __restrict__ sdat) {
    const uint32_t wmask = (uint32_t)(((uint64_t)1 << a.wb) - 1);
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = pos[p];
        const uint64_t cell = a.order ? keys[p] >> a.wb : keys[p];
        const int64_t t = (int64_t)(cell / (uint64_t)a.W);
        const int64_t row = last_at_or_below(rs, a.nor, k);
        // order 1: sorted entry i of a run goes to block i / 256 of the run's
        // padded storage, at 4 * (i mod 64) + (i / 64 mod 4): lane L of a
        // wave reading 4 consecutive stored entries per lane then holds
        // sorted entries L, 64 + L, 128 + L, 192 + L of the block
        int64_t q = p;
        if (a.order) {
            const int64_t i = p - a.roff[cell];
            q = a.poff[cell] + (i & ~(int64_t)255) + 4 * (i & 63) + ((i >> 6) & 3);
        }
        sind[q] = ((uint32_t)(row - sstart[t]) << a.wb) | ((uint32_t)ci[k] & wmask);
        sdat[q] = val[k];
    }
}

static unsigned grid_of(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace pwmk

namespace pwm_impl {

int csb_key_bits(int64_t cells, int bits) {
    int cb = 0;
    while (((int64_t)1 << cb) < cells) ++cb;
    return cb + 2 * bits;
}

// Keys, one stable sort and the gather of indices and values into c.ind /
// c.data (and c.srows for wide streams); the sorted keys are left in keys_out.
static int csb_sort_entries(pwm_matrix* h, const pwmk::CsbKeyArgs& a, unsigned end_bit, DevBuf& keys_out) {
    CsbDev& c = h->csb;
    const int64_t nnz = h->nnz;
    const size_t vs = value_size(h->vtype);
    hipError_t e;
    DevBuf keys, pos, pos_out, tmp;
    if ((e = keys.alloc(nnz * 8)) != hipSuccess || (e = keys_out.alloc(nnz * 8)) != hipSuccess ||
        (e = pos.alloc(nnz * 4)) != hipSuccess || (e = pos_out.alloc(nnz * 4)) != hipSuccess)
        return hip_fail(h, e, "CSB build allocation");
    hipLaunchKernelGGL(pwmk::k_csb_keys, dim3(pwmk::grid_of(nnz)), dim3(256), 0, h->stream, a,
                       h->rs.as<int32_t>(), h->ci.as<int32_t>(), keys.as<uint64_t>(), pos.as<uint32_t>());
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(h, e, "CSB keys");
    size_t tb = 0;
    e = rocprim::radix_sort_pairs(nullptr, tb, keys.as<uint64_t>(), keys_out.as<uint64_t>(), pos.as<uint32_t>(),
                                  pos_out.as<uint32_t>(), (size_t)nnz, 0u, end_bit, h->stream);
    if (e == hipSuccess) e = tmp.alloc(std::max<size_t>(tb, 4));
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(tmp.p, tb, keys.as<uint64_t>(), keys_out.as<uint64_t>(),
                                      pos.as<uint32_t>(), pos_out.as<uint32_t>(), (size_t)nnz, 0u, end_bit,
                                      h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "CSB sort");
    const dim3 g(pwmk::grid_of(nnz));
    if (vs == 8)
        hipLaunchKernelGGL(pwmk::k_csb_gather<uint64_t>, g, dim3(256), 0, h->stream, a, keys_out.as<uint64_t>(),
                           pos_out.as<uint32_t>(), h->val.as<uint64_t>(), c.ind.as<uint32_t>(), c.srows.as<uint16_t>(),
                           c.data.as<uint64_t>());
    else if (vs == 4)
        hipLaunchKernelGGL(pwmk::k_csb_gather<uint32_t>, g, dim3(256), 0, h->stream, a, keys_out.as<uint64_t>(),
                           pos_out.as<uint32_t>(), h->val.as<uint32_t>(), c.ind.as<uint32_t>(), c.srows.as<uint16_t>(),
                           c.data.as<uint32_t>());
    else
        hipLaunchKernelGGL(pwmk::k_csb_gather<uint16_t>, g, dim3(256), 0, h->stream, a, keys_out.as<uint64_t>(),
                           pos_out.as<uint32_t>(), h->val.as<uint16_t>(), c.ind.as<uint32_t>(), c.srows.as<uint16_t>(),
                           c.data.as<uint16_t>());
    if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(h->stream);  // before the temporaries go
    if (e != hipSuccess) return hip_fail(h, e, "CSB gather");
    return PWM_OK;
}

int csb_stream_device(pwm_matrix* h, int rb, int cb, int64_t S, bool wide) {
    CsbDev& c = h->csb;
    const int bits = c.bits;
    const int64_t H = c.H, V = c.V, nnz = h->nnz, nor = h->nor;
    const int64_t cells = V * S * H;
    const int64_t zpad = ((nnz + 3) & ~(int64_t)3) + 4;
    const size_t vs = value_size(h->vtype);
    hipError_t e;
    if ((e = c.ind.alloc(zpad * 4)) != hipSuccess || (e = c.data.alloc(zpad * vs)) != hipSuccess ||
        (e = c.soff.alloc((cells + 1) * 4)) != hipSuccess || (wide && (e = c.srows.alloc(zpad * 2)) != hipSuccess))
        return hip_fail(h, e, "CSB stream allocation");
    if ((e = hipMemsetAsync(c.ind.p, 0, zpad * 4, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(c.data.p, 0, zpad * vs, h->stream)) != hipSuccess ||
        (wide && (e = hipMemsetAsync(c.srows.p, 0, zpad * 2, h->stream)) != hipSuccess))
        return hip_fail(h, e, "CSB stream allocation");
    if (!wide) c.srows.release();
    pwmk::CsbKeyArgs a{nor, nnz, S, H, bits, rb, cb, 0};
    std::vector<int32_t> soff(cells + 1, 0);
    if (nnz > 0) {
        DevBuf keys_out;
        const int rc = csb_sort_entries(h, a, (unsigned)std::max(csb_key_bits(cells, bits), 1), keys_out);
        if (rc != PWM_OK) return rc;
        hipLaunchKernelGGL(pwmk::k_csb_soff, dim3(pwmk::grid_of(cells)), dim3(256), 0, h->stream,
                           keys_out.as<uint64_t>(), nnz, cells, 2 * bits, c.soff.as<int32_t>());
        if ((e = hipGetLastError()) == hipSuccess)
            e = hipMemcpyAsync(soff.data(), c.soff.p, cells * 4, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "CSB build");
    }
    soff[cells] = (int32_t)nnz;
    // the stored order's blk_ptr: block (br, bc) holds the runs of its S slabs
    std::vector<int32_t> blk_ptr(H * V + 1, 0);
    for (int64_t br = 0; br < V; ++br)
        for (int64_t bc = 0; bc < H; ++bc) {
            int64_t n = 0;
            for (int64_t sl = 0; sl < S; ++sl) {
                const int64_t cell = (br * S + sl) * H + bc;
                n += soff[cell + 1] - soff[cell];
            }
            blk_ptr[br * H + bc + 1] = (int32_t)(blk_ptr[br * H + bc] + n);
        }
    if ((e = hipMemcpy((char*)c.soff.p + cells * 4, &soff[cells], 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = c.blk_ptr.alloc(blk_ptr.size() * 4)) != hipSuccess ||
        (e = hipMemcpy(c.blk_ptr.p, blk_ptr.data(), blk_ptr.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(h, e, "CSB build");
    c.stream = 1;
    c.wide = wide ? 1 : 0;
    c.col_bits = cb;
    c.slab_bits = rb;
    c.slabs = (int32_t)S;
    c.slab_rows = (int64_t)1 << rb;
    c.pieces_per_block = 1;
    return PWM_OK;
}

int csb_pieces_device(pwm_matrix* h, int k) {
    CsbDev& c = h->csb;
    const int bits = c.bits;
    const int64_t H = c.H, V = c.V, nnz = h->nnz, nor = h->nor;
    const int64_t cells = V * H;
    // a slab is the top k row bits of a block; its entries in one block form
    // 2^(k-1) Z-order runs, one per value of the top k-1 column bits (k = 0:
    // the whole block).  Prefix q of a key's top 2k-1 Z bits is r c r c ... r.
    const int64_t S = (int64_t)1 << k;
    const int64_t np = k == 0 ? 1 : ((int64_t)1 << (k - 1));
    const int64_t nprefix = k == 0 ? 1 : ((int64_t)1 << (2 * k - 1));
    const int shift = 2 * bits - (k == 0 ? 0 : 2 * k - 1);
    const int64_t sub = cells * nprefix;
    const int64_t zpad = ((nnz + 3) & ~(int64_t)3) + 4;
    const size_t vs = value_size(h->vtype);
    hipError_t e;
    DevBuf soff;
    if ((e = c.ind.alloc(zpad * 4)) != hipSuccess || (e = c.data.alloc(zpad * vs)) != hipSuccess ||
        (e = soff.alloc((sub + 1) * 4)) != hipSuccess)
        return hip_fail(h, e, "CSB allocation");
    if ((e = hipMemsetAsync(c.ind.p, 0, zpad * 4, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(c.data.p, 0, zpad * vs, h->stream)) != hipSuccess)
        return hip_fail(h, e, "CSB allocation");
    c.srows.release();
    c.soff.release();
    const pwmk::CsbKeyArgs a{nor, nnz, 1, H, bits, bits, 16, 1};
    std::vector<int32_t> off(sub + 1, 0);
    if (nnz > 0) {
        DevBuf keys_out;
        const int rc = csb_sort_entries(h, a, (unsigned)std::max(csb_key_bits(cells, bits), 1), keys_out);
        if (rc != PWM_OK) return rc;
        hipLaunchKernelGGL(pwmk::k_csb_soff, dim3(pwmk::grid_of(sub)), dim3(256), 0, h->stream,
                           keys_out.as<uint64_t>(), nnz, sub, shift, soff.as<int32_t>());
        if ((e = hipGetLastError()) == hipSuccess)
            e = hipMemcpyAsync(off.data(), soff.p, sub * 4, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "CSB build");
    }
    off[sub] = (int32_t)nnz;
    std::vector<int32_t> blk_ptr(cells + 1);
    for (int64_t cell = 0; cell <= cells; ++cell) blk_ptr[cell] = off[cell * nprefix];
    // task t = br*S + s; pieces[t][bc][j] = (start, end)
    std::vector<blocked::Piece> pieces(V * S * H * np);
    parallel_for(cells, 4096, [&](int64_t c0, int64_t c1, int) {
        for (int64_t cell = c0; cell < c1; ++cell) {
            const int64_t br = cell / H, bc = cell % H;
            for (int64_t q = 0; q < nprefix; ++q) {
                int64_t sr = 0, pc = 0;
                for (int lvl = 0; lvl < k; ++lvl) {
                    sr = (sr << 1) | ((q >> (2 * k - 2 - 2 * lvl)) & 1);
                    if (lvl < k - 1) pc = (pc << 1) | ((q >> (2 * k - 3 - 2 * lvl)) & 1);
                }
                pieces[((br * S + sr) * H + bc) * np + pc] = {off[cell * nprefix + q], off[cell * nprefix + q + 1]};
            }
        }
    });
    if ((e = c.blk_ptr.alloc(blk_ptr.size() * 4)) != hipSuccess ||
        (e = hipMemcpy(c.blk_ptr.p, blk_ptr.data(), blk_ptr.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = c.work.alloc(pieces.size() * sizeof(blocked::Piece))) != hipSuccess ||
        (e = hipMemcpy(c.work.p, pieces.data(), pieces.size() * sizeof(blocked::Piece), hipMemcpyHostToDevice)) !=
            hipSuccess)
        return hip_fail(h, e, "CSB build");
    c.stream = 0;
    c.wide = 0;
    c.col_bits = 0;
    c.slab_bits = bits - k;
    c.slabs = (int32_t)S;
    c.slab_rows = ((int64_t)1 << bits) >> k;
    c.pieces_per_block = (int32_t)np;
    return PWM_OK;
}

int csb_sweep_device(pwm_matrix* h, const int32_t* rs, int64_t target_slabs, int64_t max_rows, int64_t max_slabs,
                     int wb) {
    CsbDev& c = h->csb;
    const int64_t nnz = h->nnz, nor = h->nor, noc = h->noc;
    // nnz-balanced row slabs: a slab closes before the row that would take it
    // past `target` entries (a longer row stands alone), or at max_rows rows.
    // target starts at nnz / target_slabs (at least 4096 entries: a smaller
    // slab would leave its workgroup mostly idle) and grows by 1/64 while the
    // cut needs more than max_slabs slabs
    int64_t target = std::max<int64_t>(4096, (nnz + target_slabs - 1) / std::max<int64_t>(target_slabs, 1));
    std::vector<int32_t> sstart;
    for (;;) {
        sstart.assign(1, 0);
        bool over = false;
        for (int64_t r = 0; r < nor && !over;) {
            const int64_t base = rs[r];
            int64_t lo = r + 1, hi = std::min(nor, r + max_rows);  // the last end in [r+1, hi] within target
            while (lo < hi) {
                const int64_t mid = (lo + hi + 1) >> 1;
                if ((int64_t)rs[mid] - base <= target)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            r = lo;
            sstart.push_back((int32_t)r);
            over = (int64_t)sstart.size() - 1 > max_slabs;
        }
        if (!over) break;
        if (target >= nnz) return PWM_ERR_UNSUPPORTED;  // the row cap alone needs too many slabs
        target += target / 64 + 1;
    }
    const int64_t T = (int64_t)sstart.size() - 1;
    int32_t rows_max = 0;
    for (int64_t t = 0; t < T; ++t) rows_max = std::max(rows_max, sstart[t + 1] - sstart[t]);
    const int64_t W = std::max<int64_t>(1, (noc + ((int64_t)1 << wb) - 1) >> wb);
    const int64_t cells = T * W;
    const int order = h->csb_sweep_order;
    const size_t vs = value_size(h->vtype);
    hipError_t e;
    if ((e = c.soff.alloc((cells + 1) * 4)) != hipSuccess ||
        (e = c.sweep_start.alloc(sstart.size() * 4)) != hipSuccess)
        return hip_fail(h, e, "CSB sweep allocation");
    if ((e = hipMemcpyAsync(c.sweep_start.p, sstart.data(), sstart.size() * 4, hipMemcpyHostToDevice, h->stream)) !=
        hipSuccess)
        return hip_fail(h, e, "CSB sweep allocation");
    c.srows.release();
    c.blk_ptr.release();
    c.work.release();
    c.sweep_order = order;
    c.sweep_len.clear();
    int64_t stored = nnz;  // entries stored (order 1: every run padded to whole blocks of 256)
    if (nnz > 0) {
        pwmk::SweepKeyArgs a{nor, nnz, T, W, wb, order};
        DevBuf keys, keys_out, pos, pos_out, tmp, roff;
        if ((e = keys.alloc(nnz * 8)) != hipSuccess || (e = keys_out.alloc(nnz * 8)) != hipSuccess ||
            (e = pos.alloc(nnz * 4)) != hipSuccess || (e = pos_out.alloc(nnz * 4)) != hipSuccess)
            return hip_fail(h, e, "CSB sweep build allocation");
        const dim3 g(pwmk::grid_of(nnz));
        hipLaunchKernelGGL(pwmk::k_sweep_keys, g, dim3(256), 0, h->stream, a, h->rs.as<int32_t>(),
                           h->ci.as<int32_t>(), c.sweep_start.as<int32_t>(), keys.as<uint64_t>(), pos.as<uint32_t>());
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(h, e, "CSB sweep keys");
        unsigned end_bit = 1;
        while (end_bit < 64 && ((uint64_t)1 << end_bit) < (uint64_t)cells) ++end_bit;
        if (order) end_bit = std::min(64u, end_bit + (unsigned)wb);
        size_t tb = 0;
        e = rocprim::radix_sort_pairs(nullptr, tb, keys.as<uint64_t>(), keys_out.as<uint64_t>(), pos.as<uint32_t>(),
                                      pos_out.as<uint32_t>(), (size_t)nnz, 0u, end_bit, h->stream);
        if (e == hipSuccess) e = tmp.alloc(std::max<size_t>(tb, 4));
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(tmp.p, tb, keys.as<uint64_t>(), keys_out.as<uint64_t>(),
                                          pos.as<uint32_t>(), pos_out.as<uint32_t>(), (size_t)nnz, 0u, end_bit,
                                          h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "CSB sweep sort");
        // run starts; order 1 pads every run to whole blocks of 256 entries
        DevBuf& runs = order ? roff : c.soff;
        if (order && (e = roff.alloc((cells + 1) * 4)) != hipSuccess) return hip_fail(h, e, "CSB sweep allocation");
        hipLaunchKernelGGL(pwmk::k_csb_soff, dim3(pwmk::grid_of(cells)), dim3(256), 0, h->stream,
                           keys_out.as<uint64_t>(), nnz, cells, order ? wb : 0, runs.as<int32_t>());
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(h, e, "CSB sweep runs");
        if (order) {
            std::vector<int32_t> r(cells + 1), pr(cells + 1);
            if ((e = hipMemcpyAsync(r.data(), roff.p, cells * 4, hipMemcpyDeviceToHost, h->stream)) != hipSuccess ||
                (e = hipStreamSynchronize(h->stream)) != hipSuccess)
                return hip_fail(h, e, "CSB sweep runs");
            r[cells] = (int32_t)nnz;
            c.sweep_len.resize(cells);
            int64_t acc = 0;
            for (int64_t i = 0; i < cells; ++i) {
                pr[i] = (int32_t)acc;
                c.sweep_len[i] = r[i + 1] - r[i];
                acc += ((int64_t)c.sweep_len[i] + 255) & ~(int64_t)255;
                if (acc > INT32_MAX - 8) return fail(h, PWM_ERR_UNSUPPORTED, "CSB swept slabs: padded runs exceed int32");
            }
            pr[cells] = (int32_t)acc;
            stored = acc;
            if ((e = hipMemcpyAsync(roff.p, r.data(), (cells + 1) * 4, hipMemcpyHostToDevice, h->stream)) !=
                    hipSuccess ||
                (e = hipMemcpyAsync(c.soff.p, pr.data(), (cells + 1) * 4, hipMemcpyHostToDevice, h->stream)) !=
                    hipSuccess ||
                (e = hipStreamSynchronize(h->stream)) != hipSuccess)
                return hip_fail(h, e, "CSB sweep runs");
            a.roff = roff.as<int32_t>();
            a.poff = c.soff.as<int32_t>();
        }
        const int64_t zpad = ((stored + 3) & ~(int64_t)3) + 4;
        if ((e = c.ind.alloc(zpad * 4)) != hipSuccess || (e = c.data.alloc(zpad * vs)) != hipSuccess)
            return hip_fail(h, e, "CSB sweep allocation");
        // padding entries: index 0xffffffff (skipped by the kernel), value 0
        if ((e = hipMemsetAsync(c.ind.p, order ? 0xff : 0, zpad * 4, h->stream)) != hipSuccess ||
            (e = hipMemsetAsync(c.data.p, 0, zpad * vs, h->stream)) != hipSuccess)
            return hip_fail(h, e, "CSB sweep allocation");
        if (vs == 8)
            hipLaunchKernelGGL(pwmk::k_sweep_gather<uint64_t>, g, dim3(256), 0, h->stream, a, h->rs.as<int32_t>(),
                               h->ci.as<int32_t>(), c.sweep_start.as<int32_t>(), keys_out.as<uint64_t>(),
                               pos_out.as<uint32_t>(), h->val.as<uint64_t>(), c.ind.as<uint32_t>(),
                               c.data.as<uint64_t>());
        else if (vs == 4)
            hipLaunchKernelGGL(pwmk::k_sweep_gather<uint32_t>, g, dim3(256), 0, h->stream, a, h->rs.as<int32_t>(),
                               h->ci.as<int32_t>(), c.sweep_start.as<int32_t>(), keys_out.as<uint64_t>(),
                               pos_out.as<uint32_t>(), h->val.as<uint32_t>(), c.ind.as<uint32_t>(),
                               c.data.as<uint32_t>());
        else
            hipLaunchKernelGGL(pwmk::k_sweep_gather<uint16_t>, g, dim3(256), 0, h->stream, a, h->rs.as<int32_t>(),
                               h->ci.as<int32_t>(), c.sweep_start.as<int32_t>(), keys_out.as<uint64_t>(),
                               pos_out.as<uint32_t>(), h->val.as<uint16_t>(), c.ind.as<uint32_t>(),
                               c.data.as<uint16_t>());
        if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(h->stream);  // before the temporaries go
        if (e != hipSuccess) return hip_fail(h, e, "CSB sweep build");
    } else {
        if ((e = c.ind.alloc(16)) != hipSuccess || (e = c.data.alloc(4 * vs)) != hipSuccess ||
            (e = hipMemsetAsync(c.ind.p, 0, 16, h->stream)) != hipSuccess ||
            (e = hipMemsetAsync(c.data.p, 0, 4 * vs, h->stream)) != hipSuccess)
            return hip_fail(h, e, "CSB sweep allocation");
        hipLaunchKernelGGL(pwmk::k_csb_soff, dim3(pwmk::grid_of(cells)), dim3(256), 0, h->stream, (const uint64_t*)nullptr,
                           (int64_t)0, cells, 0, c.soff.as<int32_t>());
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(h, e, "CSB sweep build");
        c.sweep_len.assign(order ? cells : 0, 0);
    }
    const int32_t last = (int32_t)stored;
    if ((e = hipMemcpy((char*)c.soff.p + cells * 4, &last, 4, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(h, e, "CSB sweep build");
    c.stream = 0;
    c.wide = 0;
    c.sweep = 1;
    c.sweep_wb = wb;
    c.sweep_W = W;
    c.sweep_rows = std::max<int32_t>(rows_max, 1);
    c.slabs = 1;
    c.slab_rows = c.sweep_rows;
    c.n_tasks = T;
    c.groups = 1;
    c.unroll = 2;
    return PWM_OK;
}

}  // namespace pwm_impl
