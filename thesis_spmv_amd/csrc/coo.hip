// COO -> CRS on the device (Util/TripletToCRS.hpp:135-172): rows ascending,
// columns ascending inside a row, duplicate (row, column) pairs kept in
// input order.
//
// Every entry gets the key (row << cb) | column (cb = bits of a column index)
// and its input position; a stable LSD radix sort (rocPRIM) over the key's
// bits orders the positions, so equal keys keep input order, as the
// reference's row sort followed by a per-row column sort leaves distinct
// pairs (its quicksort leaves duplicates in an unspecified order; input
// order is one such order).  Then col_ind and the values are gathered
// through the sorted positions, and row_start[r] is the first sorted key
// not below r << cb (a binary search per row).
#include <rocprim/device/device_radix_sort.hpp>

#include "api_internal.hpp"
#include "kernels.hpp"

namespace pwmk {

__global__ void k_coo_keys(const int32_t* __restrict__ rows, const int32_t* __restrict__ cols, int64_t n, int cb,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ pos) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = ((uint64_t)(uint32_t)rows[i] << cb) | (uint32_t)cols[i];
        pos[i] = (uint32_t)i;
    }
}

__global__ void k_coo_gather(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ pos,
                             const double* __restrict__ vals, int64_t n, uint64_t cmask, int32_t* __restrict__ ci,
                             double* __restrict__ dv) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        ci[k] = (int32_t)(keys[k] & cmask);
        dv[k] = vals[pos[k]];
    }
}

// row_start[r] = number of sorted keys below r << cb, r = 0..nor
__global__ void k_coo_rowstart(const uint64_t* __restrict__ keys, int64_t n, int64_t nor, int cb,
                               int32_t* __restrict__ rs) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= nor; r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t t = (uint64_t)r << cb;
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < t)
                lo = mid + 1;
            else
                hi = mid;
        }
        rs[r] = (int32_t)lo;
    }
}

static unsigned grid_of(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace pwmk

namespace pwm_impl {

int coo_to_crs_device(pwm_matrix* h, int64_t nor, int64_t noc, int64_t nnz, const int32_t* rows, const int32_t* cols,
                      const double* vals, DevBuf& rs, DevBuf& ci, DevBuf& dv) {
    int cb = 0, rb = 0;
    while (((int64_t)1 << cb) < noc) ++cb;
    while (((int64_t)1 << rb) < nor) ++rb;
    hipError_t e;
    if ((e = rs.alloc((nor + 1) * 4)) != hipSuccess || (e = alloc_padded(ci, nnz, 4)) != hipSuccess ||
        (e = alloc_padded(dv, nnz, 8)) != hipSuccess)
        return hip_fail(h, e, "COO conversion allocation");
    if (nnz == 0) {
        e = hipMemsetAsync(rs.p, 0, (nor + 1) * 4, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        return e == hipSuccess ? PWM_OK : hip_fail(h, e, "COO conversion");
    }
    DevBuf d_rows, d_cols, d_vals, keys, keys_out, pos, pos_out, tmp;
    if ((e = d_rows.alloc(nnz * 4)) != hipSuccess || (e = d_cols.alloc(nnz * 4)) != hipSuccess ||
        (e = d_vals.alloc(nnz * 8)) != hipSuccess || (e = keys.alloc(nnz * 8)) != hipSuccess ||
        (e = keys_out.alloc(nnz * 8)) != hipSuccess || (e = pos.alloc(nnz * 4)) != hipSuccess ||
        (e = pos_out.alloc(nnz * 4)) != hipSuccess)
        return hip_fail(h, e, "COO conversion allocation");
    if ((e = upload_staged(d_rows.p, rows, nnz * 4, h->stream)) != hipSuccess ||
        (e = upload_staged(d_cols.p, cols, nnz * 4, h->stream)) != hipSuccess)
        return hip_fail(h, e, "COO upload");
    hipLaunchKernelGGL(pwmk::k_coo_keys, dim3(pwmk::grid_of(nnz)), dim3(256), 0, h->stream, d_rows.as<int32_t>(),
                       d_cols.as<int32_t>(), nnz, cb, keys.as<uint64_t>(), pos.as<uint32_t>());
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(h, e, "COO keys");
    // at least one bit, so the sort always runs (a 1 x 1 matrix has 0-bit keys)
    const unsigned end_bit = (unsigned)std::max(cb + rb, 1);
    size_t tb = 0;
    e = rocprim::radix_sort_pairs(nullptr, tb, keys.as<uint64_t>(), keys_out.as<uint64_t>(), pos.as<uint32_t>(),
                                  pos_out.as<uint32_t>(), (size_t)nnz, 0u, end_bit, h->stream);
    if (e == hipSuccess) e = tmp.alloc(std::max<size_t>(tb, 4));
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(tmp.p, tb, keys.as<uint64_t>(), keys_out.as<uint64_t>(), pos.as<uint32_t>(),
                                      pos_out.as<uint32_t>(), (size_t)nnz, 0u, end_bit, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "COO sort");
    // the values are needed only by the gather: their upload runs on a stream
    // of its own while the sort runs, and has completed when this returns
    {
        hipStream_t vs = nullptr;
        e = hipStreamCreateWithFlags(&vs, hipStreamNonBlocking);
        if (e == hipSuccess) e = upload_staged(d_vals.p, vals, nnz * 8, vs);
        if (vs) {
            const hipError_t e2 = hipStreamDestroy(vs);
            if (e == hipSuccess) e = e2;
        }
        if (e != hipSuccess) return hip_fail(h, e, "COO upload");
    }
    hipLaunchKernelGGL(pwmk::k_coo_gather, dim3(pwmk::grid_of(nnz)), dim3(256), 0, h->stream,
                       keys_out.as<uint64_t>(), pos_out.as<uint32_t>(), d_vals.as<double>(), nnz,
                       ((uint64_t)1 << cb) - 1, ci.as<int32_t>(), dv.as<double>());
    hipLaunchKernelGGL(pwmk::k_coo_rowstart, dim3(pwmk::grid_of(nor + 1)), dim3(256), 0, h->stream,
                       keys_out.as<uint64_t>(), nnz, nor, cb, rs.as<int32_t>());
    if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "COO conversion");
    return PWM_OK;
}

}  // namespace pwm_impl
