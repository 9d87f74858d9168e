// Host side of the CSB and BlockCOH builds and their parity exports.  The
// stored arrays follow the reference's layouts exactly
// (SOA_Implementations/CSB.hpp:464-567, BlockCOH.hpp:319-455).  Here: the CSB
// block size, balanced flag, layout choice and load-time tuning, the BlockCOH
// band cuts and task list, and the host-built layouts (CSB slab pieces when
// csb_build = 1, the BlockCOH layout when coh_build = 1).  The default
// layouts are sorted on the device (csb_build.hip, coh_build.hip).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "api_internal.hpp"
#include "blocked.hpp"

namespace pwm_impl {

namespace {

// bit i of v moves to bit 2i
inline uint64_t spread16(uint64_t v) {
    v &= 0xffff;
    v = (v | (v << 8)) & 0x00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0full;
    v = (v | (v << 2)) & 0x33333333ull;
    v = (v | (v << 1)) & 0x55555555ull;
    return v;
}
// CSB.hpp:147-173: at each bit level the row bit ranks above the column bit.
inline uint32_t zkey(uint32_t r, uint32_t c) { return (uint32_t)((spread16(r) << 1) | spread16(c)); }

// ceil(log2(sqrt(M))) as the smallest k with 4^k >= M (equal to the
// reference's floating-point form for every M < 2^52).
inline int ceil_log4(int64_t M) {
    int k = 0;
    while (((int64_t)1 << (2 * k)) < M) ++k;
    return k;
}

// SOA_Implementations/CSB.hpp:96-113 with sizeof(T) = 8 (the reference's
// data_t): bits = min(16, 3 + lg), shrunk while bits > lg and the x and y
// slices of a block exceed 0.85 MB.
int csb_reference_bits(int64_t nor, int64_t noc) {
    const int lg = ceil_log4(std::min(nor, noc));
    int bits = std::min(16, 3 + lg);
    while (bits > lg) {
        const double beta = (double)((int64_t)1 << bits);
        if (8.0 * 8 * beta * 2 > 1 * 0.85 * 1e6 * 8)
            --bits;
        else
            break;
    }
    return bits;
}

hipError_t upload(DevBuf& d, const void* src, size_t bytes) {
    hipError_t e = d.alloc(std::max<size_t>(bytes, 4));
    if (e != hipSuccess || bytes == 0) return e;
    return hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice);
}

}  // namespace

// Values of the handle's storage type: uploaded as double, converted on the
// device (same path as the CRS formats).
// A swept-slab layout set aside while the other layouts are built and timed.
struct SweptSave {
    DevBuf ind, data, soff, start;
    int32_t wb = 0, rows = 0, nt = 1, order = 0;
    int64_t W = 0, T = 0;
    std::vector<int32_t> len;
    void take(CsbDev& c) {
        ind.swap(c.ind);
        data.swap(c.data);
        soff.swap(c.soff);
        start.swap(c.sweep_start);
        len.swap(c.sweep_len);
        order = c.sweep_order;
        c.sweep_order = 0;
        wb = c.sweep_wb;
        W = c.sweep_W;
        rows = c.sweep_rows;
        T = c.n_tasks;
        nt = c.nt;
        c.sweep = 0;
        c.sweep_wb = 0;
        c.sweep_W = 0;
        c.sweep_rows = 0;
    }
    void put(CsbDev& c) {
        c.ind.swap(ind);
        c.data.swap(data);
        c.soff.swap(soff);
        c.sweep_start.swap(start);
        c.sweep_len.swap(len);
        c.sweep_order = order;
        c.sweep = 1;
        c.sweep_wb = wb;
        c.sweep_W = W;
        c.sweep_rows = rows;
        c.n_tasks = T;
        c.nt = nt;
        c.stream = 0;
        c.wide = 0;
        c.groups = 1;
        c.unroll = 2;
        c.slabs = 1;
        c.slab_rows = rows;
    }
};

static int upload_blocked_values(pwm_matrix* h, DevBuf& dst, const std::vector<double>& v) {
    hipError_t e = dst.alloc(std::max<size_t>(v.size(), 1) * value_size(h->vtype));
    if (e != hipSuccess) return hip_fail(h, e, "blocked values");
    if (v.empty()) return PWM_OK;
    if (h->vtype == PWM_VALUES_F64) {
        e = hipMemcpy(dst.p, v.data(), v.size() * 8, hipMemcpyHostToDevice);
    } else {
        DevBuf tmp;
        e = tmp.alloc(v.size() * 8);
        if (e == hipSuccess) e = hipMemcpy(tmp.p, v.data(), v.size() * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = pwmk::launch_convert_values(tmp.as<double>(), dst.p, (int64_t)v.size(), h->vtype, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) return hip_fail(h, e, "blocked values upload");
    return PWM_OK;
}

static void check_bf16(pwm_matrix* h, const double* vals, int64_t n, int* rc) {
    if (h->vtype != PWM_VALUES_BF16) return;
    for (int64_t i = 0; i < n; ++i) {
        const float f = (float)vals[i];
        uint32_t b;
        memcpy(&b, &f, 4);
        if ((double)f != vals[i] || (b & 0xffffu)) {
            *rc = fail(h, PWM_ERR_INVALID, "BF16 storage needs values exact in BF16");
            return;
        }
    }
}

// ---------------------------------------------------------------- CSB

// Host copies of the device CRS col_ind and values (widened to double), for
// the host-built layouts when the caller has none.
static int download_crs(pwm_matrix* h, std::vector<int32_t>& ci, std::vector<double>& vals) {
    const int64_t nnz = h->nnz;
    ci.resize(std::max<int64_t>(nnz, 1));
    vals.resize(std::max<int64_t>(nnz, 1));
    if (nnz == 0) return PWM_OK;
    DevBuf tmp;
    hipError_t e = hipMemcpy(ci.data(), h->ci.p, nnz * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = tmp.alloc(nnz * 8);
    if (e == hipSuccess) e = pwmk::launch_copy_values_f64(h->val.p, h->vtype, tmp.as<double>(), nnz, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(vals.data(), tmp.p, nnz * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(h, e, "CRS download");
    return PWM_OK;
}

// Mean milliseconds of 5 SpMVs (after 2 warm-up ones) of the handle's CSB
// layout as it stands, on x = 1.
static int time_csb(pwm_matrix* h, float* ms) {
    DevBuf tx, ty;
    const size_t vs = vector_size(h->vtype);
    hipError_t e;
    if ((e = tx.alloc(std::max<int64_t>(h->noc, 1) * vs)) != hipSuccess ||
        (e = ty.alloc(std::max<int64_t>(h->nor, 1) * vs)) != hipSuccess)
        return hip_fail(h, e, "CSB autotune vectors");
    e = pwmk::launch_fill(tx.p, h->noc, h->vtype == PWM_VALUES_F64, 1.0, h->stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int w = 0; w < 2 && e == hipSuccess; ++w) e = csb_launch(h, tx.p, ty.p, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
    for (int r = 0; r < 5 && e == hipSuccess; ++r) e = csb_launch(h, tx.p, ty.p, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(h, e, "CSB autotune");
    *ms = t / 5.0f;
    return PWM_OK;
}

// Swept-slab layout: as many nnz-balanced slabs as the chip has CUs (one
// persistent workgroup of 1024 threads each, with all of the CU's LDS for its
// y-slice), or csb_sweep_waves times as many (the workgroups then run in that
// many waves), windows of 1 MB of x unless csb_window_bits says otherwise.
// PWM_ERR_UNSUPPORTED when the slabs' rows cannot fit.
static int build_swept(pwm_matrix* h, const int32_t* rs) {
    CsbDev& c = h->csb;
    int cus = 0, lds_cu = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, 0);
    if (e != hipSuccess) return hip_fail(h, e, "CSB swept slabs: device attributes");
    const int acc = h->vtype == PWM_VALUES_F64 ? 8 : 4;
    const int64_t slabs = (int64_t)std::max(1, cus) * h->csb_sweep_waves;
    int wb = h->csb_window_bits;
    if (wb == 0) {
        wb = 10;
        while (((int64_t)acc << (wb + 1)) <= ((int64_t)1 << 20)) ++wb;
        // the packed index keeps 32 - wb bits for the row in the slab: narrow
        // the window while the slabs' expected rows (a fifth of slack) need more
        const int64_t rows_needed = (h->nor + slabs - 1) / slabs * 6 / 5;
        while (wb > 12 && (((int64_t)1 << (32 - wb)) - 1) < rows_needed) --wb;
    }
    const int64_t rows_cap = std::min<int64_t>(((lds_cu - 1024) / acc) & ~(int64_t)63, ((int64_t)1 << (32 - wb)) - 1);
    if (rows_cap <= 0 || blocked::csb_sweep_blocks_per_cu(h->vtype, (int)rows_cap) < 1)
        return fail(h, PWM_ERR_UNSUPPORTED, "CSB swept slabs: no workgroup with LDS rows fits a CU");
    const int rc = csb_sweep_device(h, rs, slabs, rows_cap, slabs, wb);
    if (rc == PWM_ERR_UNSUPPORTED)
        return fail(h, rc, "CSB swept slabs: the rows do not fit " + std::to_string(slabs) + " slabs of at most " +
                               std::to_string(rows_cap) + " rows (the LDS of one workgroup per CU)");
    if (rc != PWM_OK) return rc;
    c.nt = h->csb_nt >= 0 ? h->csb_nt : 1;
    return PWM_OK;
}

// rs: host row_start; ci / vals: host CRS arrays, or NULL (values already
// validated for the storage type) -- they are then downloaded only if the
// host-built slab-piece layout is used.
int build_csb(pwm_matrix* h, const int32_t* rs, const int32_t* ci, const double* vals) {
    const int64_t nor = h->nor, noc = h->noc, nnz = h->nnz;
    int rc = PWM_OK;
    if (vals) check_bf16(h, vals, nnz, &rc);
    if (rc != PWM_OK) return rc;
    std::vector<int32_t> ci_dl;
    std::vector<double> vals_dl;
    const int bits = h->csb_bits_override > 0 ? h->csb_bits_override : csb_reference_bits(nor, noc);
    const int64_t beta = (int64_t)1 << bits;
    const int64_t H = (noc + beta - 1) / beta, V = (nor + beta - 1) / beta;
    CsbDev& c = h->csb;
    c.bits = bits;
    c.H = H;
    c.V = V;

    // CSB.hpp:549-559: unbalanced when a block row holds more than twice the mean
    c.balanced = 1;
    {
        std::vector<double> brn(V, 0.0);
        for (int64_t br = 0; br < V; ++br) brn[br] = (double)(rs[std::min(nor, (br + 1) * beta)] - rs[br * beta]);
        const double avg = std::accumulate(brn.begin(), brn.end(), 0.0) / (double)V;
        for (int64_t i = 0; i < V; ++i)
            if (brn[i] > 2 * avg) {
                c.balanced = 0;
                break;
            }
    }

    // Slab-stream layout: slabs of 2^rb rows, rb <= the LDS rows and beta.
    // Packed streams store (row in slab) << cb | column in 32 bits, so rb <=
    // 32 - cb (cb = bits of a column index); when that leaves slabs of fewer
    // than 256 rows (more than 2^24 columns) the automatic choice takes wide
    // streams, which store the column and the row in slab (u16) apart.  The
    // slab pieces are used when the caller asks for them or the stream sort
    // key (slab, block column, Z-Morton code) would need more than 64 bits;
    // they are sorted on the device too (key: block, Z-Morton code), or
    // built on the host with csb_build = 1.
    const int lds_bits = blocked::lds_rows(h->vtype) == 8192 ? 13 : 12;
    int cb = 0;
    while (((int64_t)1 << cb) < noc) ++cb;
    // no longer than needed for about 4096 slabs (16 workgroups per CU),
    // but at least 256 rows
    int fill_bits = 8;
    while (fill_bits < lds_bits && (nor >> (fill_bits + 1)) >= 4096) ++fill_bits;
    const int rb_fit = std::min(std::min(lds_bits, fill_bits), bits);
    const int rb_packed = std::min(rb_fit, 32 - cb);
    const bool wide = h->csb_layout == 3 || (h->csb_layout == 0 && rb_packed < std::min(8, bits));
    const bool use_stream = h->csb_layout != 1;
    const int rb = wide ? rb_fit : rb_packed;
    const int64_t S_stream = beta >> rb;
    // Swept slabs (one persistent workgroup per CU, nnz-balanced row slabs,
    // x walked in 1 MB windows): built when asked (csb_layout = 4) or, by
    // default, when x outgrows one XCD's L2 (2 MB), every slab's rows fit
    // the LDS of one workgroup per CU, and the caller fixed no slab-stream
    // kernel parameter (csb_groups, csb_unroll).  By default they are then
    // timed against the slab-stream candidates below and the faster is kept.
    SweptSave swept;
    float swept_ms = -1.0f;
    const bool try_sweep = h->csb_layout == 4 ||
                           (h->csb_layout == 0 && h->csb_groups == 0 && h->csb_unroll == 0 && nnz > 0 &&
                            noc * (int64_t)vector_size(h->vtype) > ((int64_t)2 << 20));
    if (try_sweep) {
        rc = build_swept(h, rs);
        if (rc == PWM_OK) {
            if (h->csb_layout == 4) {
                // the CRS arrays are not needed by the CSB kernels
                h->rs.release();
                h->ci.release();
                h->val.release();
                return PWM_OK;
            }
            if ((rc = time_csb(h, &swept_ms)) != PWM_OK) return rc;
            swept.take(c);
        } else if (rc != PWM_ERR_UNSUPPORTED || h->csb_layout == 4) {
            return rc;  // with the message of build_swept
        }
        rc = PWM_OK;
    }
    if (use_stream && csb_key_bits(V * S_stream * H, bits) <= 64) {
        // the slab streams, sorted on the device from the device CRS arrays
        rc = csb_stream_device(h, rb, cb, S_stream, wide);
        if (rc != PWM_OK) return rc;
    } else if (h->csb_build != 1 && csb_key_bits(V * H, bits) <= 64) {
        // slab pieces over the stored order, sorted on the device
        const int64_t lds_rows = (int64_t)blocked::lds_rows(h->vtype);
        int k = 0;
        while ((beta >> k) > lds_rows) ++k;
        rc = csb_pieces_device(h, k);
        if (rc != PWM_OK) return rc;
    } else {
        // slab pieces over the stored order, built on the host
        if (!ci || !vals) {
            if ((rc = download_crs(h, ci_dl, vals_dl)) != PWM_OK) return rc;
            ci = ci_dl.data();
            vals = vals_dl.data();
        }
        std::vector<uint32_t> ind(std::max<int64_t>(nnz, 1));
        std::vector<double> data(std::max<int64_t>(nnz, 1));
        std::vector<int32_t> blk_ptr(H * V + 1, 0);
        // Per block row: bucket by block column (stable), then sort each block by
        // Z-Morton key (stable, so duplicates keep their CRS order).
        parallel_for(V, 1, [&](int64_t b0, int64_t b1, int) {
            std::vector<int64_t> cnt;
            struct KP {
                uint32_t key;
                uint32_t lrow;
                int64_t e;
            };
            std::vector<KP> kp;
            for (int64_t br = b0; br < b1; ++br) {
                const int64_t r0 = br * beta, r1 = std::min(nor, r0 + beta);
                const int64_t e0 = rs[r0], e1 = rs[r1];
                cnt.assign(H + 1, 0);
                for (int64_t e = e0; e < e1; ++e) cnt[ci[e] / beta + 1]++;
                for (int64_t bc = 0; bc < H; ++bc) cnt[bc + 1] += cnt[bc];
                for (int64_t bc = 0; bc < H; ++bc) blk_ptr[br * H + bc + 1] = (int32_t)(e0 + cnt[bc + 1]);
                kp.resize(e1 - e0);
                std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
                for (int64_t r = r0; r < r1; ++r)
                    for (int64_t e = rs[r]; e < rs[r + 1]; ++e) {
                        const int64_t bc = ci[e] / beta;
                        kp[pos[bc]++] = {zkey((uint32_t)(r - r0), (uint32_t)(ci[e] - bc * beta)), (uint32_t)(r - r0), e};
                    }
                for (int64_t bc = 0; bc < H; ++bc)
                    std::stable_sort(kp.begin() + cnt[bc], kp.begin() + cnt[bc + 1],
                                     [](const KP& a, const KP& b) { return a.key < b.key; });
                for (int64_t i = 0; i < e1 - e0; ++i) {
                    const int64_t e = kp[i].e;
                    ind[e0 + i] = (kp[i].lrow << 16) | (uint32_t)(ci[e] % beta);
                    data[e0 + i] = vals[e];
                }
            }
        });
        blk_ptr[0] = 0;
        // Slab pieces: a workgroup owns `slab` rows of a block row (its y-slice
        // lives in LDS).  With slab = beta >> k, the entries of one block whose
        // top k row bits equal s form 2^(k-1) contiguous Z-order runs (one per
        // value of the top k-1 column bits), or the whole block when k = 0.
        const int64_t lds_rows = (int64_t)blocked::lds_rows(h->vtype);
        int k = 0;
        while ((beta >> k) > lds_rows) ++k;
        const int64_t slab = beta >> k;
        const int64_t np = k == 0 ? 1 : ((int64_t)1 << (k - 1));
        const int64_t S = (int64_t)1 << k;  // slabs per block row
        c.slab_bits = bits - k;
        // task t = br*S + s; pieces[t][bc][j] = (start, end)
        std::vector<blocked::Piece> pieces(V * S * H * np);
        const int shift = 2 * bits - (k == 0 ? 0 : 2 * k - 1);  // key bits below the piece prefix
        parallel_for(V, 1, [&](int64_t b0, int64_t b1, int) {
            std::vector<int64_t> cntp;
            for (int64_t br = b0; br < b1; ++br)
                for (int64_t bc = 0; bc < H; ++bc) {
                    const int64_t s0 = blk_ptr[br * H + bc], s1 = blk_ptr[br * H + bc + 1];
                    if (k == 0) {
                        pieces[((br * S) * H + bc) * np] = {(int32_t)s0, (int32_t)s1};
                        continue;
                    }
                    // prefixes in key order: 2^(2k-1) runs; prefix bits are
                    // r_{b-1} c_{b-1} ... r_{b-k}: slab = its k row bits, piece = its k-1 column bits
                    const int64_t nprefix = (int64_t)1 << (2 * k - 1);
                    cntp.assign(nprefix + 1, 0);
                    for (int64_t e = s0; e < s1; ++e) {
                        const uint32_t key = zkey(ind[e] >> 16, ind[e] & 0xffffu);
                        cntp[(key >> shift) + 1]++;
                    }
                    for (int64_t q = 0; q < nprefix; ++q) cntp[q + 1] += cntp[q];
                    for (int64_t q = 0; q < nprefix; ++q) {
                        // decode prefix q: bit 2k-2 is r_{b-1}, then c_{b-1}, r_{b-2}, ...
                        int64_t sr = 0, pc = 0;
                        for (int lvl = 0; lvl < k; ++lvl) {
                            const int rbit = (int)((q >> (2 * k - 2 - 2 * lvl)) & 1);
                            sr = (sr << 1) | rbit;
                            if (lvl < k - 1) pc = (pc << 1) | ((q >> (2 * k - 3 - 2 * lvl)) & 1);
                        }
                        pieces[((br * S + sr) * H + bc) * np + pc] = {(int32_t)(s0 + cntp[q]), (int32_t)(s0 + cntp[q + 1])};
                    }
                }
        });
        hipError_t e;
        if ((e = upload(c.ind, ind.data(), nnz * 4)) != hipSuccess ||
            (e = upload(c.blk_ptr, blk_ptr.data(), blk_ptr.size() * 4)) != hipSuccess ||
            (e = upload(c.work, pieces.data(), pieces.size() * sizeof(blocked::Piece))) != hipSuccess)
            return hip_fail(h, e, "CSB upload");
        data.resize(nnz);
        rc = upload_blocked_values(h, c.data, data);
        if (rc != PWM_OK) return rc;
        c.pieces_per_block = (int32_t)np;
        c.slabs = (int32_t)S;
        c.slab_rows = slab;
    }
    const int64_t S = c.slabs;
    const int64_t units = V * S;  // workgroups per column group
    // block-column groups: the caller's count, else the fastest of 1, 8 and
    // 16 (timed below, once the arrays are on the device); only 1 when x
    // (2 MB or less) already fits one XCD's L2, where groups cannot make the
    // gathers cheaper and only add partial sums
    std::vector<int32_t> cands;
    if (h->csb_groups > 0)
        cands.push_back((int32_t)std::min<int64_t>(h->csb_groups, H));
    else if (noc * (int64_t)vector_size(h->vtype) <= ((int64_t)2 << 20))
        cands.push_back(1);
    else
        for (int64_t gcount : {(int64_t)1, (int64_t)8, (int64_t)16})
            if (std::find(cands.begin(), cands.end(), (int32_t)std::min(gcount, H)) == cands.end())
                cands.push_back((int32_t)std::min(gcount, H));
    const int32_t gmax = *std::max_element(cands.begin(), cands.end());
    // stream kernel groups of 4 entries per thread per step: the caller's
    // choice, else 2 and 4 are timed with each group count
    std::vector<int32_t> unrolls;
    if (!c.stream)
        unrolls.push_back(2);
    else if (h->csb_unroll > 0)
        unrolls.push_back(h->csb_unroll);
    else
        unrolls = {2, 4};
    // stream kernel loads and partial stores: non-temporal (the entries are
    // read once; keeps more of the x window in L2) or plain (keeps a matrix
    // that fits the Infinity Cache there); the caller's choice, else both
    std::vector<int32_t> nts;
    if (!c.stream)
        nts.push_back(0);
    else if (h->csb_nt >= 0)
        nts.push_back(h->csb_nt);
    else
        nts = {0, 1};
    c.groups = cands[0];
    c.unroll = unrolls[0];
    c.nt = nts[0];
    c.n_tasks = units * c.groups;
    hipError_t e;
    if (gmax > 1 && (e = c.ypart.alloc((size_t)gmax * nor * vector_size(h->vtype))) != hipSuccess)
        return hip_fail(h, e, "CSB partial sums");
    if (cands.size() * unrolls.size() * nts.size() > 1) {
        // autotune: a few timed SpMVs per candidate on x = 1
        DevBuf tx, ty;
        const size_t vs = vector_size(h->vtype);
        if ((e = tx.alloc(noc * vs)) != hipSuccess || (e = ty.alloc(nor * vs)) != hipSuccess)
            return hip_fail(h, e, "CSB autotune vectors");
        (void)pwmk::launch_fill(tx.p, noc, h->vtype == PWM_VALUES_F64, 1.0, h->stream);
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        float best = 1e30f;
        int32_t best_g = cands[0], best_u = unrolls[0], best_nt = nts[0];
        for (int32_t gc : cands)
            for (int32_t uc : unrolls)
                for (int32_t ntc : nts) {
                    c.groups = gc;
                    c.unroll = uc;
                    c.nt = ntc;
                    c.n_tasks = units * gc;
                    for (int w = 0; w < 2; ++w) (void)csb_launch(h, tx.p, ty.p, nullptr);
                    (void)hipEventRecord(e0, h->stream);
                    for (int r = 0; r < 5; ++r) (void)csb_launch(h, tx.p, ty.p, nullptr);
                    (void)hipEventRecord(e1, h->stream);
                    (void)hipEventSynchronize(e1);
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (ms < best) {
                        best = ms;
                        best_g = gc;
                        best_u = uc;
                        best_nt = ntc;
                    }
                }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c.groups = best_g;
        c.unroll = best_u;
        c.nt = best_nt;
        c.n_tasks = units * best_g;
        if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(h, e, "CSB autotune");
    }
    if (swept_ms >= 0.0f) {
        // swept slabs against the tuned slab streams / pieces: keep the faster
        float ms = 0.0f;
        if ((rc = time_csb(h, &ms)) != PWM_OK) return rc;
        if (swept_ms < ms) {
            c.ypart.release();
            c.soff.release();
            c.srows.release();
            c.work.release();
            c.blk_ptr.release();
            swept.put(c);
        }
    }
    // the CRS arrays are not needed by the CSB kernels
    h->rs.release();
    h->ci.release();
    h->val.release();
    return PWM_OK;
}

// ---------------------------------------------------------------- BlockCOH

// BlockCOH.hpp:79-114 (Wikipedia xy2d, x = col, y = row, flips with the full size).
static int64_t hilbert(int64_t size, int64_t row, int64_t col) {
    int64_t index = 0;
    for (int64_t s = size / 2; s > 0; s /= 2) {
        const int64_t rr = (row & s) > 0, cr = (col & s) > 0;
        index += s * s * ((3 * cr) ^ rr);
        if (rr == 0) {
            if (cr == 1) {
                col = size - 1 - col;
                row = size - 1 - row;
            }
            std::swap(row, col);
        }
    }
    return index;
}

// BlockCOH.hpp:325-355: thread clamp and nnz-balanced bands.
static void coh_bands(int64_t nor, int64_t nnz, const int32_t* rs, int threads, std::vector<int64_t>& start,
                      std::vector<int64_t>& bnnz) {
    start.assign(threads, nor);
    bnnz.assign(threads, 0);
    const int64_t share = nnz / threads;
    int64_t sum = 0;
    int pid = 0;
    start[0] = 0;
    for (int64_t i = 0; i < nor; ++i) {
        const int64_t cnt = rs[i + 1] - rs[i];
        sum += cnt;
        if (sum > share && pid != threads - 1) {
            bnnz[pid++] = sum - cnt;
            start[pid] = i;
            sum = cnt;
        }
    }
    bnnz[pid] = sum;
}

// Work list: one LDS task per band whose rows fit the LDS window and whose
// entries fit one task, atomic tasks of at most kCohTaskChunks chunks
// otherwise (decode states left zero for the caller to fill).
void coh_plan_tasks(const std::vector<CohBandMeta>& B, int64_t lds_rows, std::vector<blocked::CohTask>& tasks,
                    std::vector<int32_t>& atomic_bands, int32_t* max_lds_rows) {
    tasks.clear();
    atomic_bands.clear();
    *max_lds_rows = 0;
    for (int p = 0; p < (int)B.size(); ++p) {
        const CohBandMeta& m = B[p];
        const int64_t nch = (m.nnz + blocked::kCohChunk - 1) / blocked::kCohChunk;
        if (m.rows <= lds_rows && nch <= blocked::kCohTaskChunks) {
            if (m.rows > 0) tasks.push_back({p, 0, (int32_t)nch, 0, 0, 0});
            *max_lds_rows = std::max(*max_lds_rows, (int32_t)m.rows);
            continue;
        }
        atomic_bands.push_back(p);
        for (int64_t c0 = 0; c0 < nch; c0 += blocked::kCohTaskChunks)
            tasks.push_back({p, (int32_t)c0, (int32_t)std::min<int64_t>(blocked::kCohTaskChunks, nch - c0), 1, 0, 0});
    }
}

// rs: host row_start; ci / vals: host CRS arrays, needed only by the host
// layout build (option coh_build = 1; the device build reads the device CRS).
static int build_blockcoh_layout(pwm_matrix* h, const int32_t* rs, const int32_t* ci, const double* vals) {
    const int64_t nor = h->nor, noc = h->noc, nnz = h->nnz;
    const bool device = h->coh_build != 1;
    int rc = PWM_OK;
    if (!device) check_bf16(h, vals, nnz, &rc);
    if (rc != PWM_OK) return rc;
    const int64_t lds_rows = blocked::lds_rows(h->vtype);
    std::vector<int64_t> bstart, bnnz;
    int threads = h->threads;
    auto clamp = [&](int t) {  // BlockCOH.hpp:325-328 (kept >= 1)
        if (t < 1) t = 1;
        if (nor / t < 8) t = (int)(nor / 8);
        return t < 1 ? 1 : t;
    };
    // the caller's thread count, else about 16K entries per band and at
    // least 640 bands (2.5 per CU: a 5M-entry Poisson matrix runs in 40 us
    // with 639 bands against 57 us with 319); bands the LDS window or one
    // task cannot hold are split into atomic tasks below
    threads = threads > 0 ? clamp(threads)
                          : clamp((int)std::min<int64_t>(INT32_MAX, std::max<int64_t>(640, nnz / 16384)));
    coh_bands(nor, nnz, rs, threads, bstart, bnnz);
    CohDev& c = h->coh;
    c.bands.assign(threads, CohBandMeta{});
    int64_t tot_nz = 0;
    std::vector<CohBandMeta>& B = c.bands;
    for (int p = 0; p < threads; ++p) {
        CohBandMeta& m = B[p];
        m.start = bstart[p];
        m.rows = (p + 1 < threads ? bstart[p + 1] : nor) - bstart[p];
        if (m.rows < 0) m.rows = 0;
        m.nnz = m.rows > 0 ? rs[m.start + m.rows] - rs[m.start] : 0;
        const int64_t mx = std::max(m.rows, noc);
        m.beta_bits = std::min(15, 3 + ceil_log4(mx));               // BlockCOH.hpp:119-132
        const int64_t beta = (int64_t)1 << m.beta_bits;
        m.H = (int32_t)((noc + beta - 1) / beta);
        m.V = (int32_t)((m.rows + beta - 1) / beta);
        m.off_nz = tot_nz;
        tot_nz += m.nnz;
    }
    if (device) return coh_layout_device(h, lds_rows);
    // Per band: Hilbert order of blocks (hilbert_size/beta grid), ICRS inside
    // each block, BICRS jumps between blocks.
    std::vector<uint16_t> col_jump(std::max<int64_t>(nnz, 1)), row_jump(std::max<int64_t>(nnz, 1));
    std::vector<double> data(std::max<int64_t>(nnz, 1));
    std::vector<std::vector<int16_t>> rjb(threads), cjb(threads);
    std::vector<std::vector<int32_t>> bnz(threads);
    std::vector<std::vector<blocked::CohBlock>> blocks(threads);
    std::vector<int64_t> nrj(threads, 0);
    parallel_for(threads, 1, [&](int64_t p0, int64_t p1, int) {
        std::vector<int64_t> key_of_block, order, rank, cnt;
        std::vector<int64_t> perm;
        for (int64_t p = p0; p < p1; ++p) {
            CohBandMeta& m = B[p];
            if (m.nnz == 0) continue;
            const int64_t beta = (int64_t)1 << m.beta_bits;
            const int64_t mx = std::max(m.rows, noc);
            int64_t hsize = 1;
            while (hsize < mx) hsize <<= 1;                               // BlockCOH.hpp:401
            const int64_t H = m.H, V = m.V, nb = H * V;
            key_of_block.resize(nb);
            for (int64_t br = 0; br < V; ++br)
                for (int64_t bc = 0; bc < H; ++bc) key_of_block[br * H + bc] = hilbert(hsize / beta, br, bc);
            order.resize(nb);
            std::iota(order.begin(), order.end(), 0);
            std::sort(order.begin(), order.end(),
                      [&](int64_t a, int64_t b) { return key_of_block[a] < key_of_block[b]; });
            rank.resize(nb);
            for (int64_t i = 0; i < nb; ++i) rank[order[i]] = i;
            // stable counting sort of the band's entries by block rank
            const int64_t e0 = rs[m.start];
            cnt.assign(nb + 1, 0);
            for (int64_t r = m.start; r < m.start + m.rows; ++r)
                for (int64_t e = rs[r]; e < rs[r + 1]; ++e)
                    cnt[rank[((r - m.start) / beta) * H + ci[e] / beta] + 1]++;
            for (int64_t i = 0; i < nb; ++i) cnt[i + 1] += cnt[i];
            perm.resize(m.nnz);
            std::vector<int64_t> lrow(m.nnz);
            {
                std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
                for (int64_t r = m.start; r < m.start + m.rows; ++r)
                    for (int64_t e = rs[r]; e < rs[r + 1]; ++e) {
                        const int64_t q = pos[rank[((r - m.start) / beta) * H + ci[e] / beta]]++;
                        perm[q] = e;
                        lrow[q] = r - m.start;
                    }
            }
            int64_t rj = 0;
            int64_t prev_br = -1, prev_bc = -1;
            for (int64_t q = 0; q < nb; ++q) {
                const int64_t s = cnt[q], t = cnt[q + 1];
                if (s == t) continue;
                const int64_t blk = order[q], br = blk / H, bc = blk % H;
                blocked::CohBlock cb;
                cb.entry = (int32_t)s;
                cb.row_jump = (int32_t)rj;
                cb.brow = (int32_t)br;
                cb.bcol = (int32_t)bc;
                // ICRS (TripletToCRS.hpp:402-452)
                int64_t pr = -1, pc = -1;
                for (int64_t i = s; i < t; ++i) {
                    const int64_t e = perm[i];
                    const int64_t r = lrow[i] % beta, cc = ci[e] % beta;
                    data[e0 + i] = vals[e];
                    if (i == s) {
                        row_jump[e0 + rj++] = (uint16_t)r;
                        col_jump[e0 + i] = (uint16_t)cc;
                    } else if (r != pr) {
                        col_jump[e0 + i] = (uint16_t)(cc + beta - pc);
                        row_jump[e0 + rj++] = (uint16_t)(r - pr);
                    } else {
                        col_jump[e0 + i] = (uint16_t)(cc - pc);
                    }
                    pr = r;
                    pc = cc;
                }
                // block-level BICRS (BlockCOH.hpp:433-439)
                if (blocks[p].empty()) {
                    rjb[p].push_back((int16_t)br);
                    cjb[p].push_back((int16_t)bc);
                } else if (br == prev_br) {
                    cjb[p].push_back((int16_t)(bc - prev_bc));
                } else {
                    rjb[p].push_back((int16_t)(br - prev_br));
                    cjb[p].push_back((int16_t)(bc - prev_bc + H));
                }
                bnz[p].push_back((int32_t)(t - s));
                blocks[p].push_back(cb);
                prev_br = br;
                prev_bc = bc;
            }
            nrj[p] = rj;
        }
    });
    // concatenate: row_jump entries of band p start at its entry offset (a
    // band has at most nnz row segments), block tables band after band
    std::vector<blocked::CohBlock> allb;
    std::vector<blocked::CohBand> table(threads);
    std::vector<int16_t> all_rjb, all_cjb;
    std::vector<int32_t> all_bnz, ctab;
    for (int p = 0; p < threads; ++p) {
        // chunk c of the kernel decodes entries [c*kCohChunk, (c+1)*kCohChunk):
        // ctab[c] = block holding its first entry, ctab[chunks] = last block
        table[p].chunk_table = (int32_t)ctab.size();
        const int64_t nch = (B[p].nnz + blocked::kCohChunk - 1) / blocked::kCohChunk;
        int64_t b = 0;
        for (int64_t c = 0; c < nch; ++c) {
            const int64_t e = c * blocked::kCohChunk;
            while (b + 1 < (int64_t)blocks[p].size() && blocks[p][b + 1].entry <= e) ++b;
            ctab.push_back((int32_t)b);
        }
        ctab.push_back((int32_t)std::max<int64_t>(0, (int64_t)blocks[p].size() - 1));
        CohBandMeta& m = B[p];
        m.blocks = (int64_t)blocks[p].size();
        m.row_jumps = nrj[p];
        m.block_row_jumps = (int64_t)rjb[p].size();
        m.off_rj = m.off_nz;
        m.off_blk = (int64_t)allb.size();
        m.off_rjb = (int64_t)all_rjb.size();
        table[p].start = (int32_t)m.start;
        table[p].rows = (int32_t)m.rows;
        table[p].beta_bits = m.beta_bits;
        table[p].first_block = (int32_t)m.off_blk;
        table[p].blocks = (int32_t)m.blocks;
        table[p].entry = (int32_t)m.off_nz;
        table[p].nnz = (int32_t)m.nnz;
        allb.insert(allb.end(), blocks[p].begin(), blocks[p].end());
        all_rjb.insert(all_rjb.end(), rjb[p].begin(), rjb[p].end());
        all_cjb.insert(all_cjb.end(), cjb[p].begin(), cjb[p].end());
        all_bnz.insert(all_bnz.end(), bnz[p].begin(), bnz[p].end());
    }
    c.total_blocks = (int64_t)allb.size();
    // work list; each atomic task starts from the decode state (S, R) of the
    // entry before its first one
    std::vector<blocked::CohTask> tasks;
    std::vector<int32_t> atomic_bands;
    coh_plan_tasks(B, lds_rows, tasks, atomic_bands, &c.lds_rows);
    for (size_t t0 = 0; t0 < tasks.size();) {
        size_t t1 = t0 + 1;
        while (t1 < tasks.size() && tasks[t1].band == tasks[t0].band) ++t1;
        if (tasks[t0].atomic) {
            const int p = tasks[t0].band;
            const CohBandMeta& m = B[p];
            const uint16_t* cj = col_jump.data() + m.off_nz;
            const uint16_t* rj = row_jump.data() + m.off_nz;
            const std::vector<blocked::CohBlock>& bl = blocks[p];
            int64_t S = 0, R = 0, b = 0, e = 0;
            for (size_t t = t0; t < t1; ++t) {
                // decode up to the entry before the task's first one
                const int64_t stop = (int64_t)tasks[t].chunk0 * blocked::kCohChunk;
                for (; e < stop; ++e) {
                    while (b + 1 < (int64_t)bl.size() && bl[b + 1].entry <= e) ++b;
                    if (bl[b].entry == e) {
                        S = cj[e];
                        R = rj[bl[b].row_jump];
                    } else {
                        S += cj[e];
                        if ((S >> m.beta_bits) != ((S - cj[e]) >> m.beta_bits)) R += rj[bl[b].row_jump + (S >> m.beta_bits)];
                    }
                }
                tasks[t].carry_s = (int32_t)S;
                tasks[t].carry_r = (int32_t)R;
            }
        }
        t0 = t1;
    }
    c.n_atomic = (int64_t)atomic_bands.size();
    hipError_t e;
    if ((e = upload(c.col_jump, col_jump.data(), nnz * 2)) != hipSuccess ||
        (e = upload(c.row_jump, row_jump.data(), nnz * 2)) != hipSuccess ||
        (e = upload(c.row_jump_block, all_rjb.data(), all_rjb.size() * 2)) != hipSuccess ||
        (e = upload(c.col_jump_block, all_cjb.data(), all_cjb.size() * 2)) != hipSuccess ||
        (e = upload(c.block_nnz, all_bnz.data(), all_bnz.size() * 4)) != hipSuccess ||
        (e = upload(c.block_coords, allb.data(), allb.size() * sizeof(blocked::CohBlock))) != hipSuccess ||
        (e = upload(c.band_table, table.data(), table.size() * sizeof(blocked::CohBand))) != hipSuccess ||
        (e = upload(c.tasks, ctab.data(), ctab.size() * 4)) != hipSuccess ||
        (e = upload(c.task_list, tasks.data(), tasks.size() * sizeof(blocked::CohTask))) != hipSuccess ||
        (e = upload(c.atomic_bands, atomic_bands.data(), atomic_bands.size() * 4)) != hipSuccess)
        return hip_fail(h, e, "BlockCOH upload");
    data.resize(nnz);
    rc = upload_blocked_values(h, c.data, data);
    if (rc != PWM_OK) return rc;
    c.n_tasks = (int64_t)tasks.size();
    h->rs.release();
    h->ci.release();
    h->val.release();
    return PWM_OK;
}

int64_t csb_partials(const pwm_matrix* h) {
    return h->csb.groups > 1 && !h->csb.sweep ? blocked::csb_reduce_blocks(h->nor) : h->csb.n_tasks;
}
int64_t coh_partials(const pwm_matrix* h) { return h->coh.n_tasks + h->coh.n_atomic; }

hipError_t csb_launch(pwm_matrix* h, const void* x, void* y, const pwmk::PowerArgs* pw, int part) {
    blocked::CsbArgs a;
    a.nor = h->nor;
    a.bits = h->csb.bits;
    a.H = h->csb.H;
    a.slabs = h->csb.slabs;
    a.slab_rows = h->csb.slab_rows;
    a.np = h->csb.pieces_per_block;
    a.np_shift = 0;
    while ((1 << a.np_shift) < a.np) ++a.np_shift;
    a.ind = h->csb.ind.as<uint32_t>();
    a.val = h->csb.data.p;
    a.pieces = h->csb.work.as<blocked::Piece>();
    a.n_tasks = h->csb.n_tasks;
    a.groups = h->csb.groups;
    a.ypart = h->csb.ypart.p;
    a.vtype = h->vtype;
    a.stream = h->csb.stream;
    a.col_bits = h->csb.col_bits;
    a.soff = h->csb.soff.as<int32_t>();
    a.unroll = h->csb.unroll;
    a.nt = h->csb.nt;
    a.wide = h->csb.wide;
    a.srows = h->csb.srows.as<uint16_t>();
    a.sweep = h->csb.sweep;
    a.wb = h->csb.sweep_wb;
    a.W = h->csb.sweep_W;
    a.max_rows = h->csb.sweep_rows;
    a.sync_every = h->csb_sync_every;
    if (a.sweep) a.unroll = h->csb_sweep_unroll;
    a.sstart = h->csb.sweep_start.as<int32_t>();
    hipError_t e = blocked::launch_csb(a, x, y, pw, h->stream, part);
    if (e == hipSuccess && pw && part != 1)
        e = blocked::launch_blocked_finalize(h->vtype, (int)csb_partials(h), *pw, h->stream);
    return e;
}

hipError_t coh_launch(pwm_matrix* h, const void* x, void* y, const pwmk::PowerArgs* pw) {
    blocked::CohArgs a;
    a.nor = h->nor;
    a.noc = h->noc;
    a.n_bands = (int64_t)h->coh.bands.size();
    a.n_tasks = h->coh.n_tasks;
    a.tasks = h->coh.task_list.as<blocked::CohTask>();
    a.bands = h->coh.band_table.as<blocked::CohBand>();
    a.blocks = h->coh.block_coords.as<blocked::CohBlock>();
    a.chunk_block = h->coh.tasks.as<int32_t>();
    a.col_jump = h->coh.col_jump.as<uint16_t>();
    a.row_jump = h->coh.row_jump.as<uint16_t>();
    a.val = h->coh.data.p;
    a.vtype = h->vtype;
    a.lds_rows = h->coh.lds_rows;
    a.occ8 = h->coh.occ8;
    hipError_t e = blocked::launch_coh(a, h->coh.atomic_bands.as<int32_t>(), (int)h->coh.n_atomic, x, y, pw, h->stream);
    if (e == hipSuccess && pw)
        e = blocked::launch_blocked_finalize(h->vtype, (int)coh_partials(h), *pw, h->stream);
    return e;
}

// Mean milliseconds of 5 BlockCOH SpMVs (after 2 warm-up ones) on x = 1.
static int time_coh(pwm_matrix* h, float* ms) {
    DevBuf tx, ty;
    const size_t vs = vector_size(h->vtype);
    hipError_t e;
    if ((e = tx.alloc(std::max<int64_t>(h->noc, 1) * vs)) != hipSuccess ||
        (e = ty.alloc(std::max<int64_t>(h->nor, 1) * vs)) != hipSuccess)
        return hip_fail(h, e, "BlockCOH autotune vectors");
    e = pwmk::launch_fill(tx.p, h->noc, h->vtype == PWM_VALUES_F64, 1.0, h->stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int w = 0; w < 2 && e == hipSuccess; ++w) e = coh_launch(h, tx.p, ty.p, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
    for (int r = 0; r < 5 && e == hipSuccess; ++r) e = coh_launch(h, tx.p, ty.p, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(h, e, "BlockCOH autotune");
    *ms = t / 5.0f;
    return PWM_OK;
}

// The BlockCOH layout, then the kernel instance: the compiler's register
// budget (FP64: 66 VGPRs, 7 waves per SIMD, 3 workgroups of 8 waves per CU)
// or one capped for 8 waves per SIMD (4 workgroups when the LDS window
// allows), the faster of the two on x = 1 unless the option coh_occ8 fixes it.
int build_blockcoh(pwm_matrix* h, const int32_t* rs, const int32_t* ci, const double* vals) {
    int rc = build_blockcoh_layout(h, rs, ci, vals);
    if (rc != PWM_OK) return rc;
    if (h->coh_occ8 >= 0 || h->coh.n_tasks <= 0) {
        h->coh.occ8 = h->coh_occ8 > 0 ? 1 : 0;
        return PWM_OK;
    }
    float ms0 = 0.0f, ms1 = 0.0f;
    h->coh.occ8 = 0;
    if ((rc = time_coh(h, &ms0)) != PWM_OK) return rc;
    h->coh.occ8 = 1;
    if ((rc = time_coh(h, &ms1)) != PWM_OK) return rc;
    h->coh.occ8 = ms1 < ms0 ? 1 : 0;
    return PWM_OK;
}

}  // namespace pwm_impl

using namespace pwm_impl;

// The stored CSB arrays of a slab-stream handle: every block's entries are
// collected from its slabs' runs and put back in Z-Morton order (stable: a
// duplicate (row, column) pair lies in one slab, in its stored order), which
// is the order the build gave them before they were split into slabs.
static int csb_stream_layout(pwm_matrix* h, uint32_t* ind, double* data, int32_t* blk_ptr) {
    const CsbDev& c = h->csb;
    const int64_t H = c.H, V = c.V, S = c.slabs, nnz = h->nnz;
    const int rb = c.slab_bits, cb = c.col_bits;
    const int64_t beta = (int64_t)1 << c.bits;
    const uint32_t cmask = (uint32_t)(((uint64_t)1 << cb) - 1);
    std::vector<int32_t> off(V * S * H + 1);
    std::vector<uint32_t> sind(std::max<int64_t>(nnz, 1));
    std::vector<uint16_t> srow(c.wide ? std::max<int64_t>(nnz, 1) : 0);
    std::vector<double> sdat(std::max<int64_t>(nnz, 1));
    hipError_t e = hipMemcpy(blk_ptr, c.blk_ptr.p, (H * V + 1) * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(off.data(), c.soff.p, off.size() * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && nnz) e = hipMemcpy(sind.data(), c.ind.p, nnz * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && nnz && c.wide) e = hipMemcpy(srow.data(), c.srows.p, nnz * 2, hipMemcpyDeviceToHost);
    if (e == hipSuccess && nnz) {
        DevBuf tmp;
        e = tmp.alloc(nnz * 8);
        if (e == hipSuccess) e = pwmk::launch_copy_values_f64(c.data.p, h->vtype, tmp.as<double>(), nnz, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) e = hipMemcpy(sdat.data(), tmp.p, nnz * 8, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return hip_fail(h, e, "pwm_csb_layout");
    parallel_for(V, 1, [&](int64_t b0, int64_t b1, int) {
        struct Ent {
            uint32_t key, ind;
            double v;
        };
        std::vector<Ent> tmp;
        for (int64_t br = b0; br < b1; ++br)
            for (int64_t bc = 0; bc < H; ++bc) {
                tmp.clear();
                for (int64_t sl = 0; sl < S; ++sl) {
                    const int64_t cell = (br * S + sl) * H + bc;
                    for (int64_t p = off[cell]; p < off[cell + 1]; ++p) {
                        const uint32_t w = sind[p];
                        const uint32_t lrow = (uint32_t)(sl << rb) + (c.wide ? srow[p] : (cb < 32 ? (w >> cb) : 0u));
                        const uint32_t col = (uint32_t)((c.wide ? w : (w & cmask)) - bc * beta);
                        tmp.push_back({zkey(lrow, col), (lrow << 16) | col, sdat[p]});
                    }
                }
                std::stable_sort(tmp.begin(), tmp.end(), [](const Ent& a, const Ent& b) { return a.key < b.key; });
                const int64_t base = blk_ptr[br * H + bc];
                for (size_t i = 0; i < tmp.size(); ++i) {
                    ind[base + i] = tmp[i].ind;
                    data[base + i] = tmp[i].v;
                }
            }
    });
    return PWM_OK;
}

// The stored CSB arrays of a swept-slab handle: every entry is decoded to
// (row, column), bucketed by block (stably, so a duplicate pair keeps its
// CRS order: both copies lie in one (slab, window) run) and each block put in
// Z-Morton order; blk_ptr counts the blocks in block-row-major order.
static int csb_sweep_layout(pwm_matrix* h, uint32_t* ind, double* data, int32_t* blk_ptr) {
    const CsbDev& c = h->csb;
    const int64_t T = c.n_tasks, W = c.sweep_W, nnz = h->nnz, H = c.H, V = c.V;
    const int wb = c.sweep_wb, bits = c.bits;
    const int64_t beta = (int64_t)1 << bits;
    const uint32_t wmask = (uint32_t)(((uint64_t)1 << wb) - 1);
    std::vector<int32_t> sstart(T + 1), woff(T * W + 1);
    hipError_t e = hipMemcpy(sstart.data(), c.sweep_start.p, (T + 1) * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(woff.data(), c.soff.p, woff.size() * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(h, e, "pwm_csb_layout");
    const int64_t stored = woff[T * W];  // nnz, or with column order the padded runs
    std::vector<uint32_t> sind(std::max<int64_t>(stored, 1));
    std::vector<double> sdat(std::max<int64_t>(stored, 1));
    if (stored) e = hipMemcpy(sind.data(), c.ind.p, stored * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && stored) {
        DevBuf tmp;
        e = tmp.alloc(stored * 8);
        if (e == hipSuccess) e = pwmk::launch_copy_values_f64(c.data.p, h->vtype, tmp.as<double>(), stored, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) e = hipMemcpy(sdat.data(), tmp.p, stored * 8, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return hip_fail(h, e, "pwm_csb_layout");
    // entry n of the runs' sorted order (runs in slab and window order) and
    // where it is stored: run r holds sorted entries [sorted[r], sorted[r+1])
    const bool by_col = c.sweep_order != 0;
    std::vector<int64_t> sorted(T * W + 1, 0);
    for (int64_t r = 0; r < T * W; ++r)
        sorted[r + 1] = sorted[r] + (by_col ? (int64_t)c.sweep_len[r] : (int64_t)woff[r + 1] - woff[r]);
    if (sorted[T * W] != nnz) return fail(h, PWM_ERR_STATE, "pwm_csb_layout: swept runs do not hold nnz entries");
    // decode in sorted order, then count per block
    std::vector<int64_t> blk(std::max<int64_t>(nnz, 1)), at(std::max<int64_t>(nnz, 1));
    std::vector<uint32_t> lr(std::max<int64_t>(nnz, 1)), lc(std::max<int64_t>(nnz, 1));
    parallel_for(T, 1, [&](int64_t t0, int64_t t1, int) {
        for (int64_t t = t0; t < t1; ++t)
            for (int64_t p = 0; p < W; ++p) {
                const int64_t r = t * W + p;
                for (int64_t n = sorted[r]; n < sorted[r + 1]; ++n) {
                    const int64_t i = n - sorted[r];
                    const int64_t q =
                        by_col ? woff[r] + (i & ~(int64_t)255) + 4 * (i & 63) + ((i >> 6) & 3) : woff[r] + i;
                    const int64_t row = sstart[t] + (int64_t)(sind[q] >> wb);
                    const int64_t col = (p << wb) + (int64_t)(sind[q] & wmask);
                    blk[n] = (row / beta) * H + col / beta;
                    lr[n] = (uint32_t)(row % beta);
                    lc[n] = (uint32_t)(col % beta);
                    at[n] = q;
                }
            }
    });
    std::vector<int64_t> cnt(H * V + 1, 0);
    for (int64_t q = 0; q < nnz; ++q) cnt[blk[q] + 1]++;
    for (int64_t b = 0; b < H * V; ++b) cnt[b + 1] += cnt[b];
    for (int64_t b = 0; b <= H * V; ++b) blk_ptr[b] = (int32_t)cnt[b];
    std::vector<int64_t> order(std::max<int64_t>(nnz, 1));
    {
        std::vector<int64_t> next(cnt.begin(), cnt.end() - 1);
        for (int64_t q = 0; q < nnz; ++q) order[next[blk[q]]++] = q;
    }
    parallel_for(H * V, 64, [&](int64_t b0, int64_t b1, int) {
        for (int64_t b = b0; b < b1; ++b) {
            std::stable_sort(order.begin() + cnt[b], order.begin() + cnt[b + 1], [&](int64_t u, int64_t v) {
                return zkey(lr[u], lc[u]) < zkey(lr[v], lc[v]);
            });
            for (int64_t i = cnt[b]; i < cnt[b + 1]; ++i) {
                const int64_t q = order[i];
                ind[i] = (lr[q] << 16) | lc[q];
                data[i] = sdat[at[q]];
            }
        }
    });
    return PWM_OK;
}

extern "C" int pwm_csb_layout(pwm_matrix* h, int32_t* block_bits, int64_t* nblocks, uint32_t* ind, double* data,
                              int32_t* blk_ptr) {
    if (!h || !block_bits || !nblocks) return fail(h, PWM_ERR_INVALID, "pwm_csb_layout: NULL argument");
    if (h->format != PWM_FORMAT_CSB || !h->loaded) return fail(h, PWM_ERR_STATE, "pwm_csb_layout: not a loaded CSB handle");
    *block_bits = h->csb.bits;
    *nblocks = h->csb.H * h->csb.V;
    if (!ind || !data || !blk_ptr) return PWM_OK;
    if (h->csb.sweep) return csb_sweep_layout(h, ind, data, blk_ptr);
    if (h->csb.stream) return csb_stream_layout(h, ind, data, blk_ptr);
    hipError_t e = hipMemcpy(blk_ptr, h->csb.blk_ptr.p, (*nblocks + 1) * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && h->nnz) e = hipMemcpy(ind, h->csb.ind.p, h->nnz * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && h->nnz) {
        DevBuf tmp;
        e = tmp.alloc(h->nnz * 8);
        if (e == hipSuccess) e = pwmk::launch_copy_values_f64(h->csb.data.p, h->vtype, tmp.as<double>(), h->nnz, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) e = hipMemcpy(data, tmp.p, h->nnz * 8, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return hip_fail(h, e, "pwm_csb_layout");
    return PWM_OK;
}

extern "C" int pwm_blockcoh_band(pwm_matrix* h, int band, int64_t* counts, int32_t* meta, uint16_t* row_jump,
                                 uint16_t* col_jump, double* data, int16_t* row_jump_block, int16_t* col_jump_block,
                                 int32_t* block_nnz) {
    if (!h || !counts || !meta) return fail(h, PWM_ERR_INVALID, "pwm_blockcoh_band: NULL argument");
    if (h->format != PWM_FORMAT_BLOCKCOH || !h->loaded)
        return fail(h, PWM_ERR_STATE, "pwm_blockcoh_band: not a loaded BlockCOH handle");
    if (band < 0 || band >= (int)h->coh.bands.size()) return fail(h, PWM_ERR_INVALID, "pwm_blockcoh_band: no such band");
    const CohBandMeta& m = h->coh.bands[band];
    counts[0] = m.nnz;
    counts[1] = m.blocks;
    counts[2] = m.row_jumps;
    counts[3] = m.block_row_jumps;
    counts[4] = m.start;
    counts[5] = m.rows;
    meta[0] = m.beta_bits;
    meta[1] = m.H;
    meta[2] = m.V;
    if (!row_jump || !col_jump || !data || !row_jump_block || !col_jump_block || !block_nnz) return PWM_OK;
    hipError_t e = hipSuccess;
    if (m.row_jumps) e = hipMemcpy(row_jump, h->coh.row_jump.as<uint16_t>() + m.off_rj, m.row_jumps * 2, hipMemcpyDeviceToHost);
    if (e == hipSuccess && m.nnz) e = hipMemcpy(col_jump, h->coh.col_jump.as<uint16_t>() + m.off_nz, m.nnz * 2, hipMemcpyDeviceToHost);
    if (e == hipSuccess && m.block_row_jumps)
        e = hipMemcpy(row_jump_block, h->coh.row_jump_block.as<int16_t>() + m.off_rjb, m.block_row_jumps * 2, hipMemcpyDeviceToHost);
    if (e == hipSuccess && m.blocks) {
        e = hipMemcpy(col_jump_block, h->coh.col_jump_block.as<int16_t>() + m.off_blk, m.blocks * 2, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(block_nnz, h->coh.block_nnz.as<int32_t>() + m.off_blk, m.blocks * 4, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess && m.nnz) {
        DevBuf tmp;
        const size_t vs = value_size(h->vtype);
        e = tmp.alloc(m.nnz * 8);
        if (e == hipSuccess)
            e = pwmk::launch_copy_values_f64((const char*)h->coh.data.p + m.off_nz * vs, h->vtype, tmp.as<double>(), m.nnz,
                                             h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) e = hipMemcpy(data, tmp.p, m.nnz * 8, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return hip_fail(h, e, "pwm_blockcoh_band");
    return PWM_OK;
}
