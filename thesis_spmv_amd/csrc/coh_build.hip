// BlockCOH layout built on the device (SOA_Implementations/BlockCOH.hpp:319-455,
// ICRS inside each block as in Util/TripletToCRS.hpp:402-452).
//
// The bands (a short sequential cut over the row lengths) and their block
// sizes come from the host; everything per entry runs here:
//  1. entry k of row r in band p gets the key  p << 2h | hilbert(r, c)  (the
//     block's Hilbert index on the band's grid, 2h bits enough for every
//     band) and the payload  k << 32 | lrow << 16 | lcol  (row and column
//     inside the block, both below beta <= 2^15);
//  2. a stable radix sort of (key, payload) puts each band's blocks in
//     Hilbert order with the entries of a block in CRS order -- the stable
//     counting sort of the host build;
//  3. per sorted entry, a block-start flag and a row-start flag (new block,
//     or another row than the entry before it in the same block); col_jump
//     and the gathered value follow from the entry and its predecessor;
//  4. one inclusive scan of the packed flags numbers the row jumps and the
//     blocks, so every row start writes its row_jump and every block start
//     its block record in place;
//  5. per block, the block-level BICRS jumps (BlockCOH.hpp:433-439), a flag
//     for each block that starts a block row, and a second scan that places
//     row_jump_block;
//  6. per band, the chunk -> block table, and per atomic task the decode
//     state (S, R) of the entry before its first one.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "api_internal.hpp"
#include "blocked.hpp"
#include "kernels.hpp"

namespace pwmk {

// per band, what the build kernels need
struct CohBuildBand {
    int32_t start, rows, beta_bits, hs;  // hs = Hilbert grid side (hsize >> beta_bits, 0 when hsize < beta)
    int32_t H, pad;
    int64_t off_nz;
};

// BlockCOH.hpp:79-114 (Wikipedia xy2d with x = column, y = row; the flip uses
// the full grid side).
__device__ __forceinline__ uint64_t coh_hilbert(int64_t size, int64_t row, int64_t col) {
    uint64_t index = 0;
    for (int64_t s = size / 2; s > 0; s /= 2) {
        const int64_t rr = (row & s) > 0, cr = (col & s) > 0;
        index += (uint64_t)(s * s) * (uint64_t)((3 * cr) ^ rr);
        if (rr == 0) {
            if (cr == 1) {
                col = size - 1 - col;
                row = size - 1 - row;
            }
            const int64_t t = row;
            row = col;
            col = t;
        }
    }
    return index;
}

// the last row r with rs[r] <= k
__device__ __forceinline__ int64_t coh_row_of(const int32_t* __restrict__ rs, int64_t nor, int64_t k) {
    int64_t lo = 0, hi = nor - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (rs[mid] <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

struct CohKeyArgs {
    int64_t nor, nnz, n_bands;
    int hb2;  // 2h: bits of the Hilbert part of a key
};

__global__ void k_coh_keys(CohKeyArgs a, const CohBuildBand* __restrict__ bands, const int32_t* __restrict__ rs,
                           const int32_t* __restrict__ ci, uint64_t* __restrict__ keys, uint64_t* __restrict__ pay) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.nnz; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = coh_row_of(rs, a.nor, k);
        // the last band starting at or before r (an empty band shares its
        // start with the next one, which is the one holding the row)
        int64_t lo = 0, hi = a.n_bands - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (bands[mid].start <= r)
                lo = mid;
            else
                hi = mid - 1;
        }
        const CohBuildBand b = bands[lo];
        const int64_t lrow = r - b.start, c = ci[k];
        const int64_t mask = ((int64_t)1 << b.beta_bits) - 1;
        const uint64_t h = coh_hilbert(b.hs, lrow >> b.beta_bits, c >> b.beta_bits);
        keys[k] = ((uint64_t)lo << a.hb2) | h;
        pay[k] = ((uint64_t)k << 32) | ((uint64_t)(lrow & mask) << 16) | (uint64_t)(c & mask);
    }
}

__device__ __forceinline__ uint32_t pay_row(uint64_t p) { return (uint32_t)(p >> 16) & 0xffffu; }
__device__ __forceinline__ uint32_t pay_col(uint64_t p) { return (uint32_t)p & 0xffffu; }

// step 3: flags (row start << 32 | block start), col_jump and the value
template <typename E>
__global__ void k_coh_entries(CohKeyArgs a, const CohBuildBand* __restrict__ bands, const uint64_t* __restrict__ keys,
                              const uint64_t* __restrict__ pay, const E* __restrict__ val,
                              uint64_t* __restrict__ flags, uint16_t* __restrict__ col_jump, E* __restrict__ data) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < a.nnz; q += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[q], p = pay[q];
        const bool bs = q == 0 || keys[q - 1] != key;
        const uint64_t pp = bs ? 0 : pay[q - 1];
        const bool rsf = bs || pay_row(pp) != pay_row(p);
        const uint32_t lc = pay_col(p);
        uint32_t cj = lc;
        if (!bs) {
            const uint32_t beta = 1u << bands[key >> a.hb2].beta_bits;
            cj = rsf ? lc + beta - pay_col(pp) : lc - pay_col(pp);
        }
        col_jump[q] = (uint16_t)cj;
        data[q] = val[p >> 32];
        flags[q] = ((uint64_t)rsf << 32) | (uint64_t)bs;
    }
}

// step 4: row jumps and block records.  With s = inclusive scan of the flags,
// the row segment of entry q is number (s[q] >> 32) - 1 overall and
// (s[q] >> 32) - (s[q0] >> 32) inside its band (q0 = the band's first entry);
// its block is number (s[q] & 0xffffffff) - 1.
__global__ void k_coh_scatter(CohKeyArgs a, const CohBuildBand* __restrict__ bands, const int32_t* __restrict__ rs,
                              const int32_t* __restrict__ ci, const uint64_t* __restrict__ keys,
                              const uint64_t* __restrict__ pay, const uint64_t* __restrict__ flags,
                              const uint64_t* __restrict__ scan, uint16_t* __restrict__ row_jump,
                              blocked::CohBlock* __restrict__ blocks, int32_t* __restrict__ block_q,
                              int32_t* __restrict__ block_band) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < a.nnz; q += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t f = flags[q];
        if (!(f >> 32)) continue;  // block starts are row starts too
        const uint64_t key = keys[q], p = pay[q];
        const int64_t band = (int64_t)(key >> a.hb2);
        const CohBuildBand b = bands[band];
        const uint64_t s = scan[q], s0 = scan[b.off_nz];
        const int64_t rj = (int64_t)(s >> 32) - (int64_t)(s0 >> 32);
        if (f & 1) {
            row_jump[b.off_nz + rj] = (uint16_t)pay_row(p);
            const int64_t e = (int64_t)(p >> 32);
            const int64_t r = coh_row_of(rs, a.nor, e);
            const int64_t blk = (int64_t)(s & 0xffffffffu) - 1;
            blocked::CohBlock cb;
            cb.entry = (int32_t)(q - b.off_nz);
            cb.row_jump = (int32_t)rj;
            cb.brow = (int32_t)((r - b.start) >> b.beta_bits);
            cb.bcol = (int32_t)(ci[e] >> b.beta_bits);
            blocks[blk] = cb;
            block_q[blk] = (int32_t)q;
            block_band[blk] = (int32_t)band;
        } else {
            row_jump[b.off_nz + rj] = (uint16_t)(pay_row(p) - pay_row(pay[q - 1]));
        }
    }
}

// step 5: block-level BICRS jumps, block sizes and the block-row-start flags
__global__ void k_coh_blocks(int64_t total, int64_t nnz, const CohBuildBand* __restrict__ bands,
                             const uint64_t* __restrict__ scan, const blocked::CohBlock* __restrict__ blocks,
                             const int32_t* __restrict__ block_q, const int32_t* __restrict__ block_band,
                             int16_t* __restrict__ col_jump_block, int16_t* __restrict__ rjb_val,
                             uint32_t* __restrict__ row_flag, int32_t* __restrict__ block_nnz) {
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < total; b += (int64_t)gridDim.x * blockDim.x) {
        const CohBuildBand bd = bands[block_band[b]];
        const int64_t first = (int64_t)(scan[bd.off_nz] & 0xffffffffu) - 1;
        const blocked::CohBlock cur = blocks[b];
        int64_t cj, rj = 0;
        uint32_t fl = 1;
        if (b == first) {
            rj = cur.brow;
            cj = cur.bcol;
        } else {
            const blocked::CohBlock prev = blocks[b - 1];
            if (cur.brow == prev.brow) {
                fl = 0;
                cj = (int64_t)cur.bcol - prev.bcol;
            } else {
                rj = (int64_t)cur.brow - prev.brow;
                cj = (int64_t)cur.bcol - prev.bcol + bd.H;
            }
        }
        col_jump_block[b] = (int16_t)cj;
        rjb_val[b] = (int16_t)rj;
        row_flag[b] = fl;
        block_nnz[b] = (b + 1 < total ? block_q[b + 1] : (int32_t)nnz) - block_q[b];
    }
}

__global__ void k_coh_rjb(int64_t total, const uint32_t* __restrict__ row_flag, const uint32_t* __restrict__ row_scan,
                          const int16_t* __restrict__ rjb_val, int16_t* __restrict__ row_jump_block) {
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < total; b += (int64_t)gridDim.x * blockDim.x)
        if (row_flag[b]) row_jump_block[row_scan[b] - 1] = rjb_val[b];
}

// per band: blocks, row jumps and block-row jumps (zero for an empty band)
__global__ void k_coh_band_counts(int64_t n_bands, const CohBuildBand* __restrict__ bands,
                                  const int64_t* __restrict__ band_nnz, const uint64_t* __restrict__ scan,
                                  const uint32_t* __restrict__ row_scan, int32_t* __restrict__ out) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n_bands; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = band_nnz[p];
        int32_t nb = 0, nrj = 0, nrjb = 0;
        if (n > 0) {
            const uint64_t s0 = scan[bands[p].off_nz], s1 = scan[bands[p].off_nz + n - 1];
            const int64_t b0 = (int64_t)(s0 & 0xffffffffu) - 1, b1 = (int64_t)(s1 & 0xffffffffu) - 1;
            nb = (int32_t)(b1 - b0 + 1);
            nrj = (int32_t)((s1 >> 32) - (s0 >> 32) + 1);
            nrjb = (int32_t)(row_scan[b1] - row_scan[b0] + 1);
        }
        out[3 * p] = nb;
        out[3 * p + 1] = nrj;
        out[3 * p + 2] = nrjb;
    }
}

// step 6: chunk c of band p starts in the band's last block whose first entry
// is <= c * kCohChunk; entry `chunks` is the band's last block (0 if none)
__global__ void k_coh_chunks(int64_t n_bands, const blocked::CohBand* __restrict__ table,
                             const blocked::CohBlock* __restrict__ blocks, int32_t* __restrict__ ctab) {
    for (int64_t p = blockIdx.x; p < n_bands; p += gridDim.x) {
        const blocked::CohBand bd = table[p];
        const int64_t nch = ((int64_t)bd.nnz + blocked::kCohChunk - 1) / blocked::kCohChunk;
        const blocked::CohBlock* __restrict__ bl = blocks + bd.first_block;
        for (int64_t c = threadIdx.x; c <= nch; c += blockDim.x) {
            int32_t v;
            if (c == nch) {
                v = bd.blocks > 0 ? bd.blocks - 1 : 0;
            } else {
                const int64_t e = c * blocked::kCohChunk;
                int64_t lo = 0, hi = bd.blocks - 1;
                while (lo < hi) {
                    const int64_t mid = (lo + hi + 1) >> 1;
                    if (bl[mid].entry <= e)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                v = (int32_t)lo;
            }
            ctab[bd.chunk_table + c] = v;
        }
    }
}

// decode state after the entries before an atomic task's first one: the row
// inside the block (R) and S = column + beta * (row changes since the block
// started)
__global__ void k_coh_carries(int64_t n_tasks, const blocked::CohBand* __restrict__ table,
                              const uint64_t* __restrict__ pay, const uint64_t* __restrict__ scan,
                              const int32_t* __restrict__ block_q, blocked::CohTask* __restrict__ tasks) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_tasks; t += (int64_t)gridDim.x * blockDim.x) {
        blocked::CohTask tk = tasks[t];
        if (!tk.atomic || tk.chunk0 == 0) continue;
        const blocked::CohBand bd = table[tk.band];
        const int64_t q = (int64_t)bd.entry + (int64_t)tk.chunk0 * blocked::kCohChunk - 1;
        const uint64_t s = scan[q];
        const int64_t qb = block_q[(int64_t)(s & 0xffffffffu) - 1];
        const int64_t changes = (int64_t)(s >> 32) - (int64_t)(scan[qb] >> 32);
        tk.carry_s = (int32_t)(pay_col(pay[q]) + (changes << bd.beta_bits));
        tk.carry_r = (int32_t)pay_row(pay[q]);
        tasks[t] = tk;
    }
}

static unsigned grid_of(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace pwmk

namespace pwm_impl {

int coh_layout_device(pwm_matrix* h, int64_t lds_rows) {
    CohDev& c = h->coh;
    std::vector<CohBandMeta>& B = c.bands;
    const int64_t P = (int64_t)B.size(), nnz = h->nnz, nor = h->nor, noc = h->noc;
    const size_t vs = value_size(h->vtype);
    hipError_t e;
    // host side of step 1: the band table and the width of the Hilbert part
    std::vector<pwmk::CohBuildBand> bb(P);
    std::vector<int64_t> band_nnz(P);
    int hbits = 0;
    for (int64_t p = 0; p < P; ++p) {
        const CohBandMeta& m = B[p];
        int lg = 0;
        while (((int64_t)1 << lg) < std::max(m.rows, noc)) ++lg;  // hsize = 2^lg (BlockCOH.hpp:401)
        const int hs_bits = lg >= m.beta_bits ? lg - m.beta_bits : -1;
        bb[p] = pwmk::CohBuildBand{(int32_t)m.start, (int32_t)m.rows, m.beta_bits, hs_bits >= 0 ? (1 << hs_bits) : 0,
                                   m.H, 0, m.off_nz};
        band_nnz[p] = m.nnz;
        hbits = std::max(hbits, hs_bits);
    }
    int pbits = 0;
    while (((int64_t)1 << pbits) < P) ++pbits;
    pwmk::CohKeyArgs ka{nor, nnz, P, 2 * hbits};

    const int64_t zpad = std::max<int64_t>(nnz, 1) + 64;
    if ((e = c.col_jump.alloc(zpad * 2)) != hipSuccess || (e = c.row_jump.alloc(zpad * 2)) != hipSuccess ||
        (e = c.data.alloc(zpad * vs)) != hipSuccess)
        return hip_fail(h, e, "BlockCOH allocation");
    if ((e = hipMemsetAsync(c.col_jump.p, 0, zpad * 2, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(c.row_jump.p, 0, zpad * 2, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(c.data.p, 0, zpad * vs, h->stream)) != hipSuccess)
        return hip_fail(h, e, "BlockCOH allocation");

    DevBuf d_bands, d_bnnz, keys, keys_out, pay, pay_out, tmp, block_q, block_band, rjb_val, row_flag, row_scan,
        d_counts;
    std::vector<int32_t> counts(3 * std::max<int64_t>(P, 1), 0);
    int64_t total = 0;
    if ((e = d_bands.alloc(P * sizeof(pwmk::CohBuildBand))) != hipSuccess ||
        (e = hipMemcpy(d_bands.p, bb.data(), P * sizeof(pwmk::CohBuildBand), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(h, e, "BlockCOH band table");
    if (nnz > 0) {
        if ((e = keys.alloc(nnz * 8)) != hipSuccess || (e = keys_out.alloc(nnz * 8)) != hipSuccess ||
            (e = pay.alloc(nnz * 8)) != hipSuccess || (e = pay_out.alloc(nnz * 8)) != hipSuccess)
            return hip_fail(h, e, "BlockCOH build allocation");
        const dim3 g(pwmk::grid_of(nnz)), t(256);
        hipLaunchKernelGGL(pwmk::k_coh_keys, g, t, 0, h->stream, ka, d_bands.as<pwmk::CohBuildBand>(),
                           h->rs.as<int32_t>(), h->ci.as<int32_t>(), keys.as<uint64_t>(), pay.as<uint64_t>());
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(h, e, "BlockCOH keys");
        const unsigned end_bit = (unsigned)std::max(pbits + 2 * hbits, 1);
        size_t tb = 0;
        e = rocprim::radix_sort_pairs(nullptr, tb, keys.as<uint64_t>(), keys_out.as<uint64_t>(), pay.as<uint64_t>(),
                                      pay_out.as<uint64_t>(), (size_t)nnz, 0u, end_bit, h->stream);
        if (e == hipSuccess) e = tmp.alloc(std::max<size_t>(tb, 4));
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(tmp.p, tb, keys.as<uint64_t>(), keys_out.as<uint64_t>(), pay.as<uint64_t>(),
                                          pay_out.as<uint64_t>(), (size_t)nnz, 0u, end_bit, h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "BlockCOH sort");
        // keys / pay now hold the flags and their scan
        uint64_t* flags = keys.as<uint64_t>();
        uint64_t* scan = pay.as<uint64_t>();
        if (vs == 8)
            hipLaunchKernelGGL(pwmk::k_coh_entries<uint64_t>, g, t, 0, h->stream, ka,
                               d_bands.as<pwmk::CohBuildBand>(), keys_out.as<uint64_t>(), pay_out.as<uint64_t>(),
                               h->val.as<uint64_t>(), flags, c.col_jump.as<uint16_t>(), c.data.as<uint64_t>());
        else if (vs == 4)
            hipLaunchKernelGGL(pwmk::k_coh_entries<uint32_t>, g, t, 0, h->stream, ka,
                               d_bands.as<pwmk::CohBuildBand>(), keys_out.as<uint64_t>(), pay_out.as<uint64_t>(),
                               h->val.as<uint32_t>(), flags, c.col_jump.as<uint16_t>(), c.data.as<uint32_t>());
        else
            hipLaunchKernelGGL(pwmk::k_coh_entries<uint16_t>, g, t, 0, h->stream, ka,
                               d_bands.as<pwmk::CohBuildBand>(), keys_out.as<uint64_t>(), pay_out.as<uint64_t>(),
                               h->val.as<uint16_t>(), flags, c.col_jump.as<uint16_t>(), c.data.as<uint16_t>());
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(h, e, "BlockCOH entries");
        tb = 0;
        e = rocprim::inclusive_scan(nullptr, tb, flags, scan, (size_t)nnz, rocprim::plus<uint64_t>(), h->stream);
        if (e == hipSuccess && tb > tmp.bytes) {
            tmp.release();
            e = tmp.alloc(tb);
        }
        if (e == hipSuccess)
            e = rocprim::inclusive_scan(tmp.p, tb, flags, scan, (size_t)nnz, rocprim::plus<uint64_t>(), h->stream);
        uint64_t last = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&last, scan + nnz - 1, 8, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "BlockCOH scan");
        total = (int64_t)(last & 0xffffffffu);
        c.total_blocks = total;
        if ((e = c.block_coords.alloc(total * sizeof(blocked::CohBlock))) != hipSuccess ||
            (e = c.block_nnz.alloc(total * 4)) != hipSuccess || (e = c.col_jump_block.alloc(total * 2)) != hipSuccess ||
            (e = c.row_jump_block.alloc(total * 2)) != hipSuccess || (e = block_q.alloc(total * 4)) != hipSuccess ||
            (e = block_band.alloc(total * 4)) != hipSuccess || (e = rjb_val.alloc(total * 2)) != hipSuccess ||
            (e = row_flag.alloc(total * 4)) != hipSuccess || (e = row_scan.alloc(total * 4)) != hipSuccess ||
            (e = d_bnnz.alloc(P * 8)) != hipSuccess || (e = d_counts.alloc(3 * P * 4)) != hipSuccess)
            return hip_fail(h, e, "BlockCOH block allocation");
        hipLaunchKernelGGL(pwmk::k_coh_scatter, g, t, 0, h->stream, ka, d_bands.as<pwmk::CohBuildBand>(),
                           h->rs.as<int32_t>(), h->ci.as<int32_t>(), keys_out.as<uint64_t>(), pay_out.as<uint64_t>(),
                           flags, scan, c.row_jump.as<uint16_t>(), c.block_coords.as<blocked::CohBlock>(),
                           block_q.as<int32_t>(), block_band.as<int32_t>());
        const dim3 gb(pwmk::grid_of(total));
        hipLaunchKernelGGL(pwmk::k_coh_blocks, gb, t, 0, h->stream, total, nnz, d_bands.as<pwmk::CohBuildBand>(),
                           scan, c.block_coords.as<blocked::CohBlock>(), block_q.as<int32_t>(),
                           block_band.as<int32_t>(), c.col_jump_block.as<int16_t>(), rjb_val.as<int16_t>(),
                           row_flag.as<uint32_t>(), c.block_nnz.as<int32_t>());
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(h, e, "BlockCOH blocks");
        tb = 0;
        e = rocprim::inclusive_scan(nullptr, tb, row_flag.as<uint32_t>(), row_scan.as<uint32_t>(), (size_t)total,
                                    rocprim::plus<uint32_t>(), h->stream);
        if (e == hipSuccess && tb > tmp.bytes) {
            tmp.release();
            e = tmp.alloc(tb);
        }
        if (e == hipSuccess)
            e = rocprim::inclusive_scan(tmp.p, tb, row_flag.as<uint32_t>(), row_scan.as<uint32_t>(), (size_t)total,
                                        rocprim::plus<uint32_t>(), h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "BlockCOH block scan");
        hipLaunchKernelGGL(pwmk::k_coh_rjb, gb, t, 0, h->stream, total, row_flag.as<uint32_t>(),
                           row_scan.as<uint32_t>(), rjb_val.as<int16_t>(), c.row_jump_block.as<int16_t>());
        if ((e = hipMemcpyAsync(d_bnnz.p, band_nnz.data(), P * 8, hipMemcpyHostToDevice, h->stream)) != hipSuccess)
            return hip_fail(h, e, "BlockCOH band counts");
        hipLaunchKernelGGL(pwmk::k_coh_band_counts, dim3(pwmk::grid_of(P)), t, 0, h->stream, P,
                           d_bands.as<pwmk::CohBuildBand>(), d_bnnz.as<int64_t>(), scan, row_scan.as<uint32_t>(),
                           d_counts.as<int32_t>());
        if ((e = hipGetLastError()) == hipSuccess)
            e = hipMemcpyAsync(counts.data(), d_counts.p, 3 * P * 4, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "BlockCOH band counts");
    }
    // band metadata, the kernel's band table and the chunk-table offsets
    std::vector<blocked::CohBand> table(P);
    int64_t off_blk = 0, off_rjb = 0, off_ctab = 0;
    for (int64_t p = 0; p < P; ++p) {
        CohBandMeta& m = B[p];
        m.blocks = counts[3 * p];
        m.row_jumps = counts[3 * p + 1];
        m.block_row_jumps = counts[3 * p + 2];
        m.off_rj = m.off_nz;
        m.off_blk = off_blk;
        m.off_rjb = off_rjb;
        off_blk += m.blocks;
        off_rjb += m.block_row_jumps;
        blocked::CohBand& t = table[p];
        t.start = (int32_t)m.start;
        t.rows = (int32_t)m.rows;
        t.beta_bits = m.beta_bits;
        t.first_block = (int32_t)m.off_blk;
        t.blocks = (int32_t)m.blocks;
        t.entry = (int32_t)m.off_nz;
        t.nnz = (int32_t)m.nnz;
        t.chunk_table = (int32_t)off_ctab;
        off_ctab += (m.nnz + blocked::kCohChunk - 1) / blocked::kCohChunk + 1;
    }
    if (off_blk != total) return fail(h, PWM_ERR_HIP, "BlockCOH build: band block counts disagree");
    std::vector<blocked::CohTask> tasks;
    std::vector<int32_t> atomic_bands;
    coh_plan_tasks(B, lds_rows, tasks, atomic_bands, &c.lds_rows);
    c.n_atomic = (int64_t)atomic_bands.size();
    c.n_tasks = (int64_t)tasks.size();
    if ((e = c.band_table.alloc(P * sizeof(blocked::CohBand))) != hipSuccess ||
        (e = hipMemcpy(c.band_table.p, table.data(), P * sizeof(blocked::CohBand), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = c.tasks.alloc(off_ctab * 4)) != hipSuccess ||
        (e = c.task_list.alloc(tasks.size() * sizeof(blocked::CohTask))) != hipSuccess ||
        (!tasks.empty() && (e = hipMemcpy(c.task_list.p, tasks.data(), tasks.size() * sizeof(blocked::CohTask),
                                          hipMemcpyHostToDevice)) != hipSuccess) ||
        (e = c.atomic_bands.alloc(atomic_bands.size() * 4)) != hipSuccess ||
        (!atomic_bands.empty() && (e = hipMemcpy(c.atomic_bands.p, atomic_bands.data(), atomic_bands.size() * 4,
                                                 hipMemcpyHostToDevice)) != hipSuccess))
        return hip_fail(h, e, "BlockCOH tables");
    hipLaunchKernelGGL(pwmk::k_coh_chunks, dim3((unsigned)std::min<int64_t>(std::max<int64_t>(P, 1), 8192)),
                       dim3(256), 0, h->stream, P, c.band_table.as<blocked::CohBand>(),
                       c.block_coords.as<blocked::CohBlock>(), c.tasks.as<int32_t>());
    if (nnz > 0 && !tasks.empty())
        hipLaunchKernelGGL(pwmk::k_coh_carries, dim3(pwmk::grid_of((int64_t)tasks.size())), dim3(256), 0, h->stream,
                           (int64_t)tasks.size(), c.band_table.as<blocked::CohBand>(), pay_out.as<uint64_t>(),
                           pay.as<uint64_t>(), block_q.as<int32_t>(), c.task_list.as<blocked::CohTask>());
    if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "BlockCOH chunk tables");
    h->rs.release();
    h->ci.release();
    h->val.release();
    return PWM_OK;
}

}  // namespace pwm_impl
