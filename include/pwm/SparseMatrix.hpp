/*
 * Restatement of the reference's operator interface, for building the GPU
 * classes without the reference tree.  The include guards are the
 * reference's own (Matrix/SparseMatrix.hpp:7, Matrix/Triplet.hpp:10), so when
 * the reference's headers are included first these definitions are skipped
 * and the GPU classes derive from the reference's pwm::SparseMatrix itself.
 *
 * pwm::SparseMatrix: Matrix/SparseMatrix.hpp:13-68 (same members, same four
 * pure virtuals with the same signatures).
 * pwm::Triplet: the data members of Matrix/Triplet.hpp:26-45 (the loaders
 * stay in the reference; the GPU classes only read these members).
 */
#ifndef PWM_TRIPLET_HPP
#define PWM_TRIPLET_HPP
namespace pwm {
template <typename T, typename int_type>
class Triplet {
   public:
    int_type* row_coord = nullptr;
    int_type* col_coord = nullptr;
    T* data = nullptr;
    int_type row_size = 0;
    int_type col_size = 0;
    int_type nnz = 0;
};
}  // namespace pwm
#endif

#ifndef PWM_SPARSEMATRIX_HPP
#define PWM_SPARSEMATRIX_HPP
namespace pwm {
template <typename T, typename int_type>
class SparseMatrix {
   protected:
    int_type nor;
    int_type noc;
    int_type nnz;

   public:
    SparseMatrix() {}
    SparseMatrix(int threads) { (void)threads; }
    virtual ~SparseMatrix() {}
    virtual void generatePoissonMatrix(const int_type m, const int_type n, const int partitions) = 0;
    virtual void loadFromTriplets(pwm::Triplet<T, int_type>* input, const int partitions_am) = 0;
    virtual void mv(const T* x, T* y) = 0;
    virtual void powerIteration(T* x, T* y, const int_type it) = 0;
};
}  // namespace pwm
#endif
