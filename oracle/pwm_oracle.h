/*
 * pwm_oracle.h -- CPU restatement of the reference's SpMV / power-method path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the tests, the
 * smoke entry and the bench's CPU-baseline leg call it.  The product
 * (thesis_spmv_amd/libpwmhip.so) never links or calls it.
 *
 * Every function restates one reference function; the citation is given at
 * each declaration (paths relative to the reference checkout).  Indices are
 * 32-bit (the reference's int_type is `int` or `unsigned int`; every value it
 * stores is non-negative and < 2^31, so both have the same bits).  Sizes are
 * int64_t.  Values are double (the reference's data_t, Util/Constants.hpp:32).
 *
 * Build: make -C oracle   (gcc -O2 -ffp-contract=off -fopenmp; no -march, so
 * like the reference's own build no FMA is emitted and CRS::mv's
 * "round after each multiply and each add" order is reproduced exactly).
 *
 * Pinning (what checks each function) is listed in the header of
 * pwm_oracle.c.  Functions that no golden vector, closed form or brute force
 * pins are marked "parity unpinned" there.
 */
#ifndef PWM_ORACLE_H
#define PWM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- L0 support: Util/Math.hpp, Util/VectorUtill.hpp ------------------ */
int64_t orc_integer_ceil(int64_t x, int64_t y);                       /* Math.hpp:18-21 */
int64_t orc_binary_search_leftmost(const int32_t* a, int32_t v, int64_t n); /* Math.hpp:33-47 */
double  orc_norm2(const double* x, int64_t n);                        /* VectorUtill.hpp:23-26 */
void    orc_normalize(double* x, int64_t n);                          /* VectorUtill.hpp:34-41 */

/* ---- L1 generators / loaders / converters ----------------------------- */
int64_t orc_poisson_nnz(int64_t m, int64_t n);                        /* CRS.hpp:64 */
void    orc_fill_poisson(int64_t m, int64_t n, int32_t* row_start,
                         int32_t* col_ind, double* data);             /* Poisson.hpp:32-83 */
void    orc_triplet_poisson(int64_t m, int64_t n, int32_t* rows,
                            int32_t* cols, double* vals);             /* Triplet.hpp:226-278 */

/* Matrix-Market loader (Triplet.hpp:75-154).  Allocates the rows, cols and vals arrays
 * with malloc; free with orc_free.  Returns 0, or -1 if the file cannot be
 * opened, -2 on a malformed size line / entry. */
int orc_load_mm(const char* path, int has_data, int symmetric, int random_fill,
                int64_t* nrows, int64_t* ncols, int64_t* nnz,
                int32_t** rows, int32_t** cols, double** vals);
/* Kronecker .bin loader (Triplet.hpp:164-216). Same conventions. */
int orc_load_bin(const char* path, int64_t mat_size, int symmetric, int random_vals,
                 int64_t* nrows, int64_t* ncols, int64_t* nnz,
                 int32_t** rows, int32_t** cols, double** vals);
void orc_free(void* p);

/* Boost mt19937 (seed 747846) + uniform_real_distribution(-100,100), the
 * reference loaders' random fill (Triplet.hpp:47-54). */
uint32_t orc_mt19937_next(uint32_t* state /* 625 words */);
void     orc_mt19937_seed(uint32_t* state, uint32_t seed);
double   orc_uniform_real(uint32_t* state, double lo, double hi);

/* COO -> CRS (TripletToCRS.hpp:135-172).  Rows ascending, columns ascending
 * inside a row, duplicates kept (in input order).  Input is not modified. */
void orc_triplet_to_crs(int64_t nor, int64_t nnz, const int32_t* rows, const int32_t* cols,
                        const double* vals, int32_t* row_start, int32_t* col_ind, double* data);

/* Static row partitions of TripletToMultipleCRS (TripletToCRS.hpp:320-330):
 * first_rows[p], thread_rows[p] for p < partitions. */
void orc_row_partitions(int64_t nor, int partitions, int32_t* first_rows, int32_t* thread_rows);

/* TripletToMultipleCRS (TripletToCRS.hpp:309-385): the COO matrix as
 * `partitions` CRS matrices of consecutive row ranges (orc_row_partitions).
 * Partition p's arrays are row_start[first_rows[p] + p ..] (thread_rows[p]+1
 * entries, starting at 0) and col_ind/data[part_off[p] .. part_off[p+1]).
 * Input is not modified. */
void orc_triplet_to_multiple_crs(int64_t nor, int64_t nnz, const int32_t* rows, const int32_t* cols,
                                 const double* vals, int partitions, int32_t* first_rows, int32_t* thread_rows,
                                 int64_t* part_off, int32_t* row_start, int32_t* col_ind, double* data);

/* ---- L3 operators: CRS (the oracle) ----------------------------------- */
void orc_crs_mv(int64_t nor, const int32_t* row_start, const int32_t* col_ind,
                const double* data, const double* x, double* y);      /* CRS.hpp:99-109 */
void orc_crs_power_iteration(int64_t nor, const int32_t* row_start, const int32_t* col_ind,
                             const double* data, double* x, double* y, int64_t it); /* CRS.hpp:118-130 */

/* CRSOMP (Env_Implementations/CRSOMP.hpp:110-123, 135-157): OpenMP dynamic,512. */
void orc_crsomp_mv(int threads, int64_t nor, const int32_t* row_start, const int32_t* col_ind,
                   const double* data, const double* x, double* y);
void orc_crsomp_power_iteration(int threads, int64_t nor, const int32_t* row_start,
                                const int32_t* col_ind, const double* data,
                                double* x, double* y, int64_t it);

/* ---- CRS_Merge (SOA_Implementations/CRS_Merge.hpp) -------------------- */
/* searchPathOnDiag (:42-62): list A = row_end_offsets = row_start+1. */
void orc_merge_search(int64_t diagonal, const int32_t* row_end_offsets, int64_t nor, int64_t nnz,
                      int64_t* a_coord, int64_t* b_coord);
/* The split points of CRS_Merge::mv for P threads (:96-112): a[p], b[p] for
 * p = 0..P (entry P is the end of the last range). */
void orc_merge_splits(int64_t P, int64_t nor, int64_t nnz, const int32_t* row_start,
                      int64_t* a, int64_t* b);
/* CRS_Merge::mv (:95-141), run with P logical threads (sequentially or with
 * OpenMP when omp_threads > 1; both give the same bits). */
void orc_merge_mv(int64_t P, int omp_threads, int64_t nor, int64_t nnz, const int32_t* row_start,
                  const int32_t* col_ind, const double* data, const double* x, double* y);
void orc_merge_power_iteration(int64_t P, int omp_threads, int64_t nor, int64_t nnz,
                               const int32_t* row_start, const int32_t* col_ind,
                               const double* data, double* x, double* y, int64_t it);

/* ---- CSB (SOA_Implementations/CSB.hpp) -------------------------------- */
int  orc_csb_block_bits(int64_t nor, int64_t noc, int value_size);   /* :96-113 */
int  orc_zmorton_less_equal(uint32_t col1, uint32_t row1, uint32_t col2, uint32_t row2,
                            int block_bits);                           /* :147-173 */
/* Build (:464-567).  blk_ptr has H*V+1 entries; returns the balanced flag
 * (:549-559) or -1 on bad arguments.  block_bits <= 0 selects the reference rule. */
int  orc_csb_build(int64_t nor, int64_t noc, int64_t nnz, const int32_t* rows, const int32_t* cols,
                   const double* vals, int block_bits, uint32_t* ind, double* data, int32_t* blk_ptr);
void orc_csb_mv(int64_t nor, int64_t noc, int block_bits, const uint32_t* ind, const double* data,
                const int32_t* blk_ptr, const double* x, double* y);    /* :348-364, 593-605 */

/* ---- BlockCOH (SOA_Implementations/BlockCOH.hpp) ---------------------- */
int64_t orc_hilbert(int64_t size, int64_t row, int64_t col);          /* :79-114 */
/* Thread-count clamp (:325-328) and nnz-balanced bands (:332-355). */
int  orc_blockcoh_threads(int64_t nor, int threads);
void orc_blockcoh_bands(int64_t nor, int64_t nnz, const int32_t* row_counts, int threads,
                        int32_t* band_row_start, int64_t* band_nnz);
int  orc_blockcoh_beta_bits(int64_t band_rows, int64_t noc);          /* :119-132 */
int64_t orc_blockcoh_hilbert_size(int64_t band_rows, int64_t noc);    /* :401 */

/* Layout of one band, as the reference stores it (:380-452): arrays sized by
 * the caller from band_nnz (col_jump, data), band_nnz (row_jump, upper bound)
 * and band_nnz+1 (block arrays, upper bound). */
typedef struct {
    int64_t  nnz;            /* entries in the band */
    int64_t  blocks;         /* thread_blocks[pid] */
    int64_t  row_jumps;      /* used entries of row_jump */
    int64_t  block_row_jumps;/* used entries of row_jump_block */
    int32_t  beta_bits;
    int32_t  horizontal_blocks, vertical_blocks;
    uint16_t* row_jump;
    uint16_t* col_jump;
    double*   data;
    int16_t*  row_jump_block;
    int16_t*  col_jump_block;
    int32_t*  block_nnz;
} orc_coh_band;

/* Build band `pid` from a row-sorted (CRS) matrix. */
void orc_blockcoh_build_band(int64_t noc, const int32_t* row_start, const int32_t* col_ind,
                             const double* vals, int64_t band_start, int64_t band_rows,
                             orc_coh_band* out);
void orc_blockcoh_mv_band(const orc_coh_band* b, int64_t band_start, int64_t band_rows,
                          const double* x, double* y);                 /* :235-256, 479-510 */

/* ---- matrix statistics: Util_Scripts/calc_stats.cpp:44-105 ------------ */
void orc_matrix_stats(int64_t nor, const int32_t* row_start, const int32_t* col_ind, double* rows_avg,
                      double* rows_var, int64_t* row_max, int* block_bits, int64_t* nb_blocks, double* block_avg,
                      double* block_var, int* balanced);

#ifdef __cplusplus
}
#endif
#endif
