/*
 * GPU drop-ins for the reference's CPU operator classes.  Each class derives
 * from pwm::SparseMatrix<T,int_type> (Matrix/SparseMatrix.hpp:13-68) and
 * forwards the four virtuals to libpwmhip.so's C ABI (include/pwm_hip.h):
 *
 *   generatePoissonMatrix -> pwm_generate_poisson
 *   loadFromTriplets      -> pwm_load_triplets
 *   mv                    -> pwm_mv
 *   powerIteration        -> pwm_power_iteration
 *
 * plus `mult` / `powerMethod` aliases.  Errors throw pwm::HipError (the
 * reference has no error channel; it asserts, Matrix/CRS.hpp:72-73,119).
 * T is double (the reference's data_t) or float (converted at the boundary);
 * int_type is int or unsigned int, as in the reference's tests.
 *
 * selectTypeHip(method, threads) mirrors selectType (Util/DriverUtil.hpp:38-84)
 * with the GPU ids 101/102 (CRS), 108 (CRS_Merge), 109 (CSB), 110 (BlockCOH)
 * and returns nullptr for unknown ids, as the reference returns NULL.
 */
#ifndef PWM_SPARSEMATRIX_HIP_HPP
#define PWM_SPARSEMATRIX_HIP_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../pwm_hip.h"
#include "SparseMatrix.hpp"

namespace pwm {

class HipError : public std::runtime_error {
   public:
    int code;
    HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <typename T, typename int_type, int FORMAT>
class SparseMatrixHipBase : public pwm::SparseMatrix<T, int_type> {
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value, "T must be double or float");
    static_assert(sizeof(int_type) == 4, "int_type must be a 32-bit integer");

   protected:
    pwm_matrix* h = nullptr;

    void check(int rc) const {
        if (rc != PWM_OK) throw HipError(rc, pwm_last_error(h));
    }
    void sync_sizes() {
        pwm_info_t i;
        check(pwm_info(h, &i));
        this->nor = (int_type)i.nor;
        this->noc = (int_type)i.noc;
        this->nnz = (int_type)i.nnz;
    }

   public:
    explicit SparseMatrixHipBase(int threads = -1, int value_type = PWM_VALUES_F64)
        : pwm::SparseMatrix<T, int_type>(threads) {
        const int rc = pwm_create(FORMAT, threads, value_type, &h);
        if (rc != PWM_OK) throw HipError(rc, pwm_last_error(nullptr));
    }
    SparseMatrixHipBase(const SparseMatrixHipBase&) = delete;
    SparseMatrixHipBase& operator=(const SparseMatrixHipBase&) = delete;
    ~SparseMatrixHipBase() override { pwm_destroy(h); }

    void setOption(const char* key, int64_t value) { check(pwm_set_option(h, key, value)); }
    pwm_matrix* handle() const { return h; }

    void generatePoissonMatrix(const int_type m, const int_type n, const int partitions) override {
        check(pwm_generate_poisson(h, (int64_t)m, (int64_t)n, partitions));
        sync_sizes();
    }

    void loadFromTriplets(pwm::Triplet<T, int_type>* input, const int partitions_am) override {
        const int64_t z = (int64_t)input->nnz;
        std::vector<double> v(input->data, input->data + z);
        check(pwm_load_triplets(h, (int64_t)input->row_size, (int64_t)input->col_size, z,
                                reinterpret_cast<const int32_t*>(input->row_coord),
                                reinterpret_cast<const int32_t*>(input->col_coord), v.data(), partitions_am));
        sync_sizes();
    }

    void mv(const T* x, T* y) override {
        if constexpr (std::is_same<T, double>::value) {
            check(pwm_mv(h, x, y));
        } else {
            std::vector<double> xd(x, x + (size_t)this->noc), yd((size_t)this->nor);
            check(pwm_mv(h, xd.data(), yd.data()));
            for (size_t i = 0; i < yd.size(); ++i) y[i] = (T)yd[i];
        }
    }

    void powerIteration(T* x, T* y, const int_type it) override {
        if constexpr (std::is_same<T, double>::value) {
            check(pwm_power_iteration(h, x, y, (int64_t)it));
        } else {
            std::vector<double> xd(x, x + (size_t)this->nor), yd(y, y + (size_t)this->nor);
            check(pwm_power_iteration(h, xd.data(), yd.data(), (int64_t)it));
            for (size_t i = 0; i < xd.size(); ++i) {
                x[i] = (T)xd[i];
                y[i] = (T)yd[i];
            }
        }
    }

    // the same operations under the names SpMV users often call them
    void mult(const T* x, T* y) { mv(x, y); }
    void powerMethod(T* x, T* y, const int_type it) { powerIteration(x, y, it); }
};

template <typename T, typename I>
class CRSHip : public SparseMatrixHipBase<T, I, PWM_FORMAT_CRS> {
   public:
    using SparseMatrixHipBase<T, I, PWM_FORMAT_CRS>::SparseMatrixHipBase;
};
template <typename T, typename I>
class CRSMergeHip : public SparseMatrixHipBase<T, I, PWM_FORMAT_CRS_MERGE> {
   public:
    using SparseMatrixHipBase<T, I, PWM_FORMAT_CRS_MERGE>::SparseMatrixHipBase;
};
template <typename T, typename I>
class CSBHip : public SparseMatrixHipBase<T, I, PWM_FORMAT_CSB> {
   public:
    using SparseMatrixHipBase<T, I, PWM_FORMAT_CSB>::SparseMatrixHipBase;
};
template <typename T, typename I>
class BlockCOHHip : public SparseMatrixHipBase<T, I, PWM_FORMAT_BLOCKCOH> {
   public:
    using SparseMatrixHipBase<T, I, PWM_FORMAT_BLOCKCOH>::SparseMatrixHipBase;
};

// Fill a Triplet from a Matrix-Market or .bin file with the library's
// readers (Triplet::loadFromMM / loadFromBin, Matrix/Triplet.hpp:75-216; the
// same flags as the reference's loaders).  The Triplet's arrays are new[]
// allocations, as the reference's loaders leave them; the caller owns them.
namespace detail {
template <typename T, typename I>
void take_coo(int rc, int64_t nor, int64_t noc, int64_t nnz, int32_t* r, int32_t* c, double* v,
              pwm::Triplet<T, I>& t) {
    if (rc != PWM_OK) throw HipError(rc, pwm_last_error(nullptr));
    t.row_coord = new I[nnz ? nnz : 1];
    t.col_coord = new I[nnz ? nnz : 1];
    t.data = new T[nnz ? nnz : 1];
    for (int64_t k = 0; k < nnz; ++k) {
        t.row_coord[k] = (I)r[k];
        t.col_coord[k] = (I)c[k];
        t.data[k] = (T)v[k];
    }
    t.row_size = (I)nor;
    t.col_size = (I)noc;
    t.nnz = (I)nnz;
    pwm_free(r);
    pwm_free(c);
    pwm_free(v);
}
}  // namespace detail

template <typename T, typename I>
void readMatrixMarket(const char* path, bool has_data, bool symmetric, bool random_fill, pwm::Triplet<T, I>& t) {
    int64_t nor = 0, noc = 0, nnz = 0;
    int32_t *r = nullptr, *c = nullptr;
    double* v = nullptr;
    const int rc = pwm_read_mm(path, has_data, symmetric, random_fill, &nor, &noc, &nnz, &r, &c, &v);
    detail::take_coo(rc, nor, noc, nnz, r, c, v, t);
}

template <typename T, typename I>
void readBin(const char* path, int64_t mat_size, bool symmetric, bool random_vals, pwm::Triplet<T, I>& t) {
    int64_t nor = 0, noc = 0, nnz = 0;
    int32_t *r = nullptr, *c = nullptr;
    double* v = nullptr;
    const int rc = pwm_read_bin(path, mat_size, symmetric, random_vals, &nor, &noc, &nnz, &r, &c, &v);
    detail::take_coo(rc, nor, noc, nnz, r, c, v, t);
}

template <typename T, typename I>
pwm::SparseMatrix<T, I>* selectTypeHip(int method, int threads) {
    switch (method) {
        case 101:
        case 102: return new CRSHip<T, I>(threads);
        case 108: return new CRSMergeHip<T, I>(threads);
        case 109: return new CSBHip<T, I>(threads);
        case 110: return new BlockCOHHip<T, I>(threads);
        default: return nullptr;
    }
}

}  // namespace pwm

#endif
