/*
 * pwm_oracle.c -- CPU restatement of the reference's SpMV / power-method path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, the smoke entry and
 * bench.py's cpu_baseline leg.  The product library never links it.
 *
 * What pins each function (checks live in tests/test_oracle_*.py):
 *   fill_poisson, triplet_poisson  MATLAB mv goldens (Tests/poisson.cpp:24,80-81),
 *                                  nnz closed form (CRS.hpp:64), dense brute force,
 *                                  Poisson eigenvalue closed form.
 *   load_mm, load_bin              fixture size lines / byte counts, scipy.io.mmread,
 *                                  MATLAB mv goldens through crs_mv.
 *   mt19937                        published known answer (seed 5489 -> 10000th
 *                                  output 4123659995).
 *   uniform_real                   parity unpinned: no reference test uses the
 *                                  loaders' random fill; restated from Boost's
 *                                  published generate_uniform_real.
 *   triplet_to_crs                 scipy COO->CSR on duplicate-free inputs, invariants.
 *   crs_mv, crsomp_mv              MATLAB mv goldens (1e-14), dense brute force.
 *   norm2, normalize, power        MATLAB power goldens (1e-12), closed-form Poisson
 *                                  eigenpair.
 *   merge_search, merge_splits     brute-force linear merge on tiny inputs; a+b=d.
 *   merge_mv                       mv goldens (split rows summed as suffix+prefix).
 *   csb_block_bits                 closed-form values quoted for the reference rule.
 *   zmorton / csb_build / csb_mv   bit-interleave brute force, mv goldens, invariants.
 *   hilbert                        Wikipedia xy2d known answers for n = 2 and 4.
 *   blockcoh_*                     brute-force band scan, mv goldens, ICRS decode
 *                                  round trip.
 *   matrix_stats                   numpy brute force of every statistic, the
 *                                  Poisson row-length closed form, hand-made
 *                                  balanced / unbalanced block rows.
 */
#include "pwm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================= L0 support ==================================== */

/* Util/Math.hpp:18-21 */
int64_t orc_integer_ceil(int64_t x, int64_t y) { return x / y + (x % y != 0); }

/* Util/Math.hpp:33-47: leftmost index with a[i] >= v. */
int64_t orc_binary_search_leftmost(const int32_t* a, int32_t v, int64_t n) {
    int64_t left = 0, right = n;
    while (left < right) {
        int64_t mid = (left + right) / 2;
        if (a[mid] < v) left = mid + 1;
        else right = mid;
    }
    return left;
}

/* Util/VectorUtill.hpp:23-26: sqrt(inner_product(x,x,0.)) -- serial, FP64,
 * left to right, one rounding per multiply and per add. */
double orc_norm2(const double* x, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s = s + x[i] * x[i];
    return sqrt(s);
}

/* Util/VectorUtill.hpp:34-41: x[i] /= norm (a divide, not a reciprocal). */
void orc_normalize(double* x, int64_t n) {
    double nrm = orc_norm2(x, n);
    for (int64_t i = 0; i < n; ++i) x[i] /= nrm;
}

/* ======================= L1 generators ================================= */

/* Matrix/CRS.hpp:64 */
int64_t orc_poisson_nnz(int64_t m, int64_t n) { return n * (m + 2 * (m - 1)) + 2 * (n - 1) * m; }

/* Util/Poisson.hpp:32-83: row r holds, in this order, (r-m) if r>=m,
 * (r-1) if r%m!=0, r (value 4), (r+1) if r%m!=m-1, (r+m) if r<mn-m. */
void orc_fill_poisson(int64_t m, int64_t n, int32_t* row_start, int32_t* col_ind, double* data) {
    int64_t N = m * n, k = 0;
    row_start[0] = 0;
    for (int64_t r = 0; r < N; ++r) {
        if (r >= m) { data[k] = -1; col_ind[k++] = (int32_t)(r - m); }
        if (r % m != 0) { data[k] = -1; col_ind[k++] = (int32_t)(r - 1); }
        data[k] = 4; col_ind[k++] = (int32_t)r;
        if (r % m != m - 1) { data[k] = -1; col_ind[k++] = (int32_t)(r + 1); }
        if (r < N - m) { data[k] = -1; col_ind[k++] = (int32_t)(r + m); }
        row_start[r + 1] = (int32_t)k;
    }
}

/* Matrix/Triplet.hpp:226-278: the same matrix as unsorted COO, emitted per
 * i as (i,i,4), (i,i+1), (i+1,i), (i+m,i), (i,i+m). */
void orc_triplet_poisson(int64_t m, int64_t n, int32_t* rows, int32_t* cols, double* vals) {
    int64_t N = m * n, k = 0;
    for (int64_t i = 0; i < N; ++i) {
        rows[k] = (int32_t)i; cols[k] = (int32_t)i; vals[k++] = 4.;
        if (i < N - 1) {
            if (i % m != m - 1) { rows[k] = (int32_t)i; cols[k] = (int32_t)(i + 1); vals[k++] = -1.; }
            if ((i + 1) % m != 0) { rows[k] = (int32_t)(i + 1); cols[k] = (int32_t)i; vals[k++] = -1.; }
        }
        if (i < N - m) {
            rows[k] = (int32_t)(i + m); cols[k] = (int32_t)i; vals[k++] = -1.;
            rows[k] = (int32_t)i; cols[k] = (int32_t)(i + m); vals[k++] = -1.;
        }
    }
}

/* ======================= random fill ================================== */

/* MT19937 (Matsumoto & Nishimura 1998), the engine behind Triplet.hpp:48. */
void orc_mt19937_seed(uint32_t* st, uint32_t seed) {
    st[0] = seed;
    for (int i = 1; i < 624; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
    st[624] = 624;
}

uint32_t orc_mt19937_next(uint32_t* st) {
    if (st[624] >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (st[i] & 0x80000000u) | (st[(i + 1) % 624] & 0x7fffffffu);
            st[i] = st[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        st[624] = 0;
    }
    uint32_t y = st[st[624]++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* uniform_real_distribution(lo,hi) over a 32-bit engine, as Boost's
 * generate_uniform_real publishes it: u / 2^32 * (hi-lo) + lo, redrawn if it
 * reaches hi.  (Triplet.hpp:51,54).  Parity unpinned. */
double orc_uniform_real(uint32_t* st, double lo, double hi) {
    for (;;) {
        double num = (double)orc_mt19937_next(st);
        double den = 4294967295.0 + 1.0;
        double r = num / den * (hi - lo) + lo;
        if (r < hi) return r;
    }
}

/* ======================= loaders ====================================== */

void orc_free(void* p) { free(p); }

static int is_blank(const char* s) { return s[0] == '\0' || s[0] == '\n' || (s[0] == '\r' && s[1] == '\n'); }

/* Matrix/Triplet.hpp:75-154.  '%' lines are skipped; the size line gives
 * rows, cols, entries (entries doubled as capacity when symmetric); each entry
 * is 1-based; explicit zeros are dropped when the file has data; pattern
 * entries become 1 (or a random value); off-diagonal entries are mirrored when
 * symmetric; reading stops at the first empty line.  The banner is not parsed:
 * the caller passes the flags, as in the reference. */
int orc_load_mm(const char* path, int has_data, int symmetric, int random_fill,
                int64_t* nrows, int64_t* ncols, int64_t* nnz,
                int32_t** rows, int32_t** cols, double** vals) {
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    size_t cap_line = 1 << 16;
    char* line = (char*)malloc(cap_line);
    long long R = 0, C = 0, E = 0;
    int got_size = 0;
    while (fgets(line, (int)cap_line, f)) {
        if (line[0] == '%') continue;
        if (sscanf(line, "%lld %lld %lld", &R, &C, &E) != 3) { free(line); fclose(f); return -2; }
        got_size = 1;
        break;
    }
    if (!got_size || R < 0 || C < 0 || E < 0) { free(line); fclose(f); return -2; }
    int64_t cap = symmetric ? 2 * E : E;
    int32_t* r = (int32_t*)malloc(sizeof(int32_t) * (cap ? cap : 1));
    int32_t* c = (int32_t*)malloc(sizeof(int32_t) * (cap ? cap : 1));
    double* v = (double*)malloc(sizeof(double) * (cap ? cap : 1));
    uint32_t mt[625];
    orc_mt19937_seed(mt, 747846u);
    int64_t k = 0;
    while (fgets(line, (int)cap_line, f)) {
        if (is_blank(line)) break;
        long long ri, ci;
        double val = 1.0;
        char* p = line;
        char* end;
        ri = strtoll(p, &end, 10); if (end == p) break; p = end;
        ci = strtoll(p, &end, 10); if (end == p) break; p = end;
        if (k >= cap) { free(line); fclose(f); free(r); free(c); free(v); return -2; }
        if (has_data) {
            val = strtod(p, &end);
            if (end == p) { free(line); fclose(f); free(r); free(c); free(v); return -2; }
            if (val == 0.0) continue;
        } else if (random_fill) {
            val = orc_uniform_real(mt, -100., 100.);
        }
        r[k] = (int32_t)(ri - 1); c[k] = (int32_t)(ci - 1); v[k] = val; ++k;
        if (symmetric && r[k - 1] != c[k - 1]) {
            r[k] = c[k - 1]; c[k] = r[k - 1]; v[k] = v[k - 1]; ++k;
        }
    }
    free(line);
    fclose(f);
    *nrows = R; *ncols = C; *nnz = k; *rows = r; *cols = c; *vals = v;
    return 0;
}

/* Matrix/Triplet.hpp:164-216: raw little-endian uint32 (row, col) pairs,
 * 0-based, no header; Z = filesize/8; values 1 or random; symmetric mirrors
 * off-diagonal pairs.  The reference keeps reading until it has stored 2Z
 * entries when symmetric, which re-reads past the end of a file that holds
 * self-loops; this restatement stops at the end of the file instead (the
 * same result for every file without self-loops, e.g. mycielskian5.bin). */
int orc_load_bin(const char* path, int64_t mat_size, int symmetric, int random_vals,
                 int64_t* nrows, int64_t* ncols, int64_t* nnz,
                 int32_t** rows, int32_t** cols, double** vals) {
    struct stat sb;
    if (stat(path, &sb) != 0) return -1;
    int64_t pairs = (int64_t)sb.st_size / 8;
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    int64_t cap = symmetric ? 2 * pairs : pairs;
    int32_t* r = (int32_t*)malloc(sizeof(int32_t) * (cap ? cap : 1));
    int32_t* c = (int32_t*)malloc(sizeof(int32_t) * (cap ? cap : 1));
    double* v = (double*)malloc(sizeof(double) * (cap ? cap : 1));
    uint32_t mt[625];
    orc_mt19937_seed(mt, 747846u);
    int64_t k = 0;
    for (int64_t p = 0; p < pairs; ++p) {
        uint32_t rc[2];
        if (fread(rc, sizeof(uint32_t), 2, f) != 2) break;
        r[k] = (int32_t)rc[0]; c[k] = (int32_t)rc[1];
        v[k] = random_vals ? orc_uniform_real(mt, -100., 100.) : 1.0;
        ++k;
        if (symmetric && r[k - 1] != c[k - 1]) {
            r[k] = c[k - 1]; c[k] = r[k - 1]; v[k] = v[k - 1]; ++k;
        }
    }
    fclose(f);
    *nrows = mat_size; *ncols = mat_size; *nnz = k; *rows = r; *cols = c; *vals = v;
    return 0;
}

/* ======================= COO -> CRS ==================================== */

/* Util/TripletToCRS.hpp:135-172: sort by row, fill row_start (trailing empty
 * rows get nnz), then sort each row by column.  The reference's quicksort is
 * not stable, so the order of duplicate (row,col) entries is unspecified
 * there; this restatement keeps them in input order (stable sorts). */
void orc_triplet_to_crs(int64_t nor, int64_t nnz, const int32_t* rows, const int32_t* cols,
                        const double* vals, int32_t* row_start, int32_t* col_ind, double* data) {
    int64_t* cnt = (int64_t*)calloc((size_t)nor + 1, sizeof(int64_t));
    for (int64_t i = 0; i < nnz; ++i) cnt[rows[i] + 1]++;
    for (int64_t r = 0; r < nor; ++r) cnt[r + 1] += cnt[r];
    for (int64_t r = 0; r <= nor; ++r) row_start[r] = (int32_t)cnt[r];
    for (int64_t i = 0; i < nnz; ++i) {
        int64_t pos = cnt[rows[i]]++;
        col_ind[pos] = cols[i];
        data[pos] = vals[i];
    }
    free(cnt);
    /* stable insertion sort by column inside each row (rows are short in the
     * fixtures; long rows fall back to a stable merge sort) */
    int32_t* tc = NULL; double* tv = NULL; int64_t tcap = 0;
    for (int64_t r = 0; r < nor; ++r) {
        int64_t s = row_start[r], e = row_start[r + 1], len = e - s;
        if (len <= 64) {
            for (int64_t i = s + 1; i < e; ++i) {
                int32_t ck = col_ind[i]; double dv = data[i];
                int64_t j = i - 1;
                while (j >= s && col_ind[j] > ck) { col_ind[j + 1] = col_ind[j]; data[j + 1] = data[j]; --j; }
                col_ind[j + 1] = ck; data[j + 1] = dv;
            }
        } else {
            if (len > tcap) {
                tcap = len;
                tc = (int32_t*)realloc(tc, sizeof(int32_t) * tcap);
                tv = (double*)realloc(tv, sizeof(double) * tcap);
            }
            for (int64_t w = 1; w < len; w *= 2) {
                for (int64_t lo = 0; lo < len; lo += 2 * w) {
                    int64_t mid = lo + w < len ? lo + w : len, hi = lo + 2 * w < len ? lo + 2 * w : len;
                    int64_t i = lo, j = mid, o = lo;
                    while (i < mid && j < hi) {
                        if (col_ind[s + j] < col_ind[s + i]) { tc[o] = col_ind[s + j]; tv[o++] = data[s + j++]; }
                        else { tc[o] = col_ind[s + i]; tv[o++] = data[s + i++]; }
                    }
                    while (i < mid) { tc[o] = col_ind[s + i]; tv[o++] = data[s + i++]; }
                    while (j < hi) { tc[o] = col_ind[s + j]; tv[o++] = data[s + j++]; }
                }
                memcpy(col_ind + s, tc, sizeof(int32_t) * len);
                memcpy(data + s, tv, sizeof(double) * len);
            }
        }
    }
    free(tc); free(tv);
}

/* Util/TripletToCRS.hpp:320-330: am_rows = round(nor/partitions) evaluated on
 * integers (so it is floor); the last partition takes the remainder. */
void orc_row_partitions(int64_t nor, int partitions, int32_t* first_rows, int32_t* thread_rows) {
    int64_t am_rows = nor / partitions, last = 0;
    for (int i = 0; i < partitions; ++i) {
        first_rows[i] = (int32_t)last;
        last = (i == partitions - 1) ? nor : first_rows[i] + am_rows;
        thread_rows[i] = (int32_t)(last - first_rows[i]);
    }
}

/* Util/TripletToCRS.hpp:309-385.  The reference sorts all triplets by row,
 * then for partition i (rows [first_rows[i], first_rows[i] + thread_rows[i]))
 * collects its entries into local arrays: row_start[0] = 0, local row
 * offsets counted from the partition's first entry, trailing empty rows
 * filled with the partition's entry count, and each row sorted by column.
 * That is the whole matrix's CRS cut at the partition boundaries with
 * row_start rebased, which is how it is computed here (duplicates in input
 * order, as orc_triplet_to_crs keeps them).  A partition without entries is
 * empty here; the reference writes one entry into its zero-length arrays
 * (:350-353), so it has no defined result there. */
void orc_triplet_to_multiple_crs(int64_t nor, int64_t nnz, const int32_t* rows, const int32_t* cols,
                                 const double* vals, int partitions, int32_t* first_rows, int32_t* thread_rows,
                                 int64_t* part_off, int32_t* row_start, int32_t* col_ind, double* data) {
    int32_t* rs = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nor + 1));
    orc_triplet_to_crs(nor, nnz, rows, cols, vals, rs, col_ind, data);
    orc_row_partitions(nor, partitions, first_rows, thread_rows);
    for (int p = 0; p < partitions; ++p) {
        const int64_t f = first_rows[p], n = thread_rows[p];
        part_off[p] = rs[f];
        for (int64_t j = 0; j <= n; ++j) row_start[f + p + j] = rs[f + j] - rs[f];
    }
    part_off[partitions] = nnz;
    free(rs);
}

/* ======================= CRS operators ================================= */

/* Matrix/CRS.hpp:99-109: y = 0, then y[i] = y[i] + data[k]*x[col_ind[k]]
 * for k ascending inside each row. */
void orc_crs_mv(int64_t nor, const int32_t* row_start, const int32_t* col_ind,
                const double* data, const double* x, double* y) {
    for (int64_t i = 0; i < nor; ++i) y[i] = 0.;
    for (int64_t i = 0; i < nor; ++i)
        for (int64_t k = row_start[i]; k < row_start[i + 1]; ++k) y[i] = y[i] + data[k] * x[col_ind[k]];
}

/* Matrix/CRS.hpp:118-130: ping-pong mv + normalize; the result is in x when
 * `it` is even and in y when it is odd. */
void orc_crs_power_iteration(int64_t nor, const int32_t* row_start, const int32_t* col_ind,
                             const double* data, double* x, double* y, int64_t it) {
    for (int64_t i = 0; i < it; ++i) {
        if (i % 2 == 0) { orc_crs_mv(nor, row_start, col_ind, data, x, y); orc_normalize(y, nor); }
        else { orc_crs_mv(nor, row_start, col_ind, data, y, x); orc_normalize(x, nor); }
    }
}

/* Env_Implementations/CRSOMP.hpp:110-123: sum = 0; sum += data*x; y[i] = sum
 * per row, rows scheduled dynamic,512 (Util/Constants.hpp:13).  Per row this
 * is the same arithmetic as CRS::mv. */
void orc_crsomp_mv(int threads, int64_t nor, const int32_t* row_start, const int32_t* col_ind,
                   const double* data, const double* x, double* y) {
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads)
    for (int64_t i = 0; i < nor; ++i) {
        double sum = 0.;
        for (int64_t k = row_start[i]; k < row_start[i + 1]; ++k) sum += data[k] * x[col_ind[k]];
        y[i] = sum;
    }
}

/* Env_Implementations/CRSOMP.hpp:135-157: serial norm2, parallel divide. */
void orc_crsomp_power_iteration(int threads, int64_t nor, const int32_t* row_start,
                                const int32_t* col_ind, const double* data,
                                double* x, double* y, int64_t it) {
    for (int64_t i = 0; i < it; ++i) {
        double* in = (i % 2 == 0) ? x : y;
        double* out = (i % 2 == 0) ? y : x;
        orc_crsomp_mv(threads, nor, row_start, col_ind, data, in, out);
        double nrm = orc_norm2(out, nor);
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t k = 0; k < nor; ++k) out[k] /= nrm;
    }
}

/* ======================= CRS_Merge ===================================== */

/* SOA_Implementations/CRS_Merge.hpp:42-62 */
void orc_merge_search(int64_t diagonal, const int32_t* row_end_offsets, int64_t nor, int64_t nnz,
                      int64_t* a_coord, int64_t* b_coord) {
    int64_t lo = diagonal > nnz ? diagonal - nnz : 0;
    int64_t hi = diagonal < nor ? diagonal : nor;
    while (lo < hi) {
        int64_t pivot = (lo + hi) >> 1;
        if (row_end_offsets[pivot] <= diagonal - pivot - 1) lo = pivot + 1;
        else hi = pivot;
    }
    *a_coord = lo < nor ? lo : nor;
    *b_coord = diagonal - lo;
}

/* CRS_Merge.hpp:96-112: items = ceil((N+Z)/P); range p is
 * [min(items*p, N+Z), min(that+items, N+Z)). */
void orc_merge_splits(int64_t P, int64_t nor, int64_t nnz, const int32_t* row_start,
                      int64_t* a, int64_t* b) {
    int64_t total = nor + nnz, items = (total + P - 1) / P;
    for (int64_t p = 0; p <= P; ++p) {
        int64_t d = items * p < total ? items * p : total;
        orc_merge_search(d, row_start + 1, nor, nnz, &a[p], &b[p]);
    }
}

/* CRS_Merge.hpp:95-141.  Each logical thread writes y for the rows it
 * finishes (the first may be only a suffix of its row) and a carry
 * (row a_end, partial sum); the carries are then added serially in thread
 * order.  Threads write disjoint y entries, so running them in parallel or
 * in sequence gives the same bits. */
void orc_merge_mv(int64_t P, int omp_threads, int64_t nor, int64_t nnz, const int32_t* row_start,
                  const int32_t* col_ind, const double* data, const double* x, double* y) {
    const int32_t* row_end = row_start + 1;
    int64_t total = nor + nnz, items = (total + P - 1) / P;
    int64_t* carry_row = (int64_t*)malloc(sizeof(int64_t) * P);
    double* carry_val = (double*)malloc(sizeof(double) * P);
#pragma omp parallel for schedule(static) num_threads(omp_threads > 0 ? omp_threads : 1)
    for (int64_t tid = 0; tid < P; ++tid) {
        int64_t d0 = items * tid < total ? items * tid : total;
        int64_t d1 = d0 + items < total ? d0 + items : total;
        int64_t a, b, a_end, b_end;
        orc_merge_search(d0, row_end, nor, nnz, &a, &b);
        orc_merge_search(d1, row_end, nor, nnz, &a_end, &b_end);
        double run = 0.;
        for (; a < a_end; ++a) {
            for (; b < row_end[a]; ++b) run += data[b] * x[col_ind[b]];
            y[a] = run;
            run = 0.;
        }
        for (; b < b_end; ++b) run += data[b] * x[col_ind[b]];
        carry_row[tid] = a_end;
        carry_val[tid] = run;
    }
    for (int64_t tid = 0; tid < P; ++tid)
        if (carry_row[tid] < nor) y[carry_row[tid]] += carry_val[tid];
    free(carry_row);
    free(carry_val);
}

/* CRS_Merge inherits CRSOMP::powerIteration (CRSOMP.hpp:135-157). */
void orc_merge_power_iteration(int64_t P, int omp_threads, int64_t nor, int64_t nnz,
                               const int32_t* row_start, const int32_t* col_ind,
                               const double* data, double* x, double* y, int64_t it) {
    for (int64_t i = 0; i < it; ++i) {
        double* in = (i % 2 == 0) ? x : y;
        double* out = (i % 2 == 0) ? y : x;
        orc_merge_mv(P, omp_threads, nor, nnz, row_start, col_ind, data, in, out);
        double nrm = orc_norm2(out, nor);
        for (int64_t k = 0; k < nor; ++k) out[k] /= nrm;
    }
}

/* ======================= CSB =========================================== */

/* SOA_Implementations/CSB.hpp:96-113: bits = min(16, 3+ceil(log2 sqrt(min(N,noc)))),
 * shrunk while bits > lg and sizeof(T)*8*beta*2 > 1 MB*0.85*8 bits. */
int orc_csb_block_bits(int64_t nor, int64_t noc, int value_size) {
    int64_t mn = nor < noc ? nor : noc;
    int lg = (int)ceil(log2(sqrt((double)mn)));
    int bits = 3 + lg < 16 ? 3 + lg : 16;
    while (bits > lg) {
        double beta = pow(2, bits);
        if ((double)value_size * 8 * beta * 2 > 1 * 0.85 * 1e6 * 8) bits -= 1;
        else break;
    }
    return bits;
}

static uint64_t zkey(uint32_t row, uint32_t col, int bits) {
    uint64_t k = 0;
    for (int b = bits - 1; b >= 0; --b) k = (k << 2) | (((row >> b) & 1u) << 1) | ((col >> b) & 1u);
    return k;
}

/* CSB.hpp:147-173: at each bit from high to low, the row bit ranks above the
 * column bit; true when (col1,row1) <= (col2,row2). */
int orc_zmorton_less_equal(uint32_t col1, uint32_t row1, uint32_t col2, uint32_t row2, int bits) {
    for (int b = bits - 1; b >= 0; --b) {
        int s1 = (row1 >> b) & 1, s2 = (row2 >> b) & 1;
        if (s2 && !s1) return 1;
        if (s1 && !s2) return 0;
        s1 = (col1 >> b) & 1; s2 = (col2 >> b) & 1;
        if (s2 && !s1) return 1;
        if (s1 && !s2) return 0;
    }
    return 1;
}

typedef struct { uint64_t key; int64_t pos; } kp_t;
static int kp_cmp(const void* a, const void* b) {
    const kp_t* x = (const kp_t*)a; const kp_t* y = (const kp_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* CSB.hpp:464-567: sort by row; per block row, bucket by col/beta, pack
 * (r%beta)<<16 | (c%beta), sort each block in Z-Morton order, write blocks
 * in row-major block order, blk_ptr[br*H+bc+1] = running count.  Balanced
 * means no block row holds more than twice the mean block-row nnz. */
int orc_csb_build(int64_t nor, int64_t noc, int64_t nnz, const int32_t* rows, const int32_t* cols,
                  const double* vals, int block_bits, uint32_t* ind, double* data, int32_t* blk_ptr) {
    if (nor <= 0 || noc <= 0) return -1;
    int bits = block_bits > 0 ? block_bits : orc_csb_block_bits(nor, noc, (int)sizeof(double));
    int64_t beta = (int64_t)1 << bits;
    int64_t H = orc_integer_ceil(noc, beta), V = orc_integer_ceil(nor, beta);
    kp_t* kp = (kp_t*)malloc(sizeof(kp_t) * (nnz ? nnz : 1));
    for (int64_t i = 0; i < nnz; ++i) {
        uint64_t br = (uint64_t)rows[i] / beta, bc = (uint64_t)cols[i] / beta;
        uint64_t blk = br * H + bc;
        kp[i].key = (blk << (2 * bits)) | zkey((uint32_t)(rows[i] % beta), (uint32_t)(cols[i] % beta), bits);
        kp[i].pos = i;
    }
    qsort(kp, nnz, sizeof(kp_t), kp_cmp);
    for (int64_t b = 0; b <= H * V; ++b) blk_ptr[b] = 0;
    double* brn = (double*)calloc(V, sizeof(double));
    for (int64_t i = 0; i < nnz; ++i) {
        int64_t p = kp[i].pos;
        ind[i] = ((uint32_t)(rows[p] % beta) << 16) | (uint32_t)(cols[p] % beta);
        data[i] = vals[p];
        blk_ptr[(kp[i].key >> (2 * bits)) + 1]++;
        brn[rows[p] / beta] += 1.;
    }
    for (int64_t b = 0; b < H * V; ++b) blk_ptr[b + 1] += blk_ptr[b];
    free(kp);
    double sum = 0.;
    for (int64_t i = 0; i < V; ++i) sum += brn[i];
    double avg = sum / (double)V;
    int balanced = 1;
    for (int64_t i = 0; i < V; ++i)
        if (brn[i] > 2 * avg) { balanced = 0; break; }
    free(brn);
    return balanced;
}

/* CSB.hpp:593-605 with sparseRowMult (:348-364): per block row, zero the
 * y-slice, then y[r] = y[r] + v*x[bc*beta + c] over blocks in column order
 * and entries in stored (Z-Morton) order.  The unbalanced path computes the
 * same terms in a schedule-dependent order. */
void orc_csb_mv(int64_t nor, int64_t noc, int bits, const uint32_t* ind, const double* data,
                const int32_t* blk_ptr, const double* x, double* y) {
    int64_t beta = (int64_t)1 << bits;
    int64_t H = orc_integer_ceil(noc, beta), V = orc_integer_ceil(nor, beta);
    for (int64_t br = 0; br < V; ++br) {
        double* yb = y + br * beta;
        int64_t rows_here = (br + 1) * beta < nor ? beta : nor - br * beta;
        for (int64_t r = 0; r < rows_here; ++r) yb[r] = 0.;
        for (int64_t bc = 0; bc < H; ++bc) {
            int64_t blk = br * H + bc;
            for (int64_t i = blk_ptr[blk]; i < blk_ptr[blk + 1]; ++i) {
                uint32_t r = ind[i] >> 16, c = ind[i] & 0xffffu;
                yb[r] = yb[r] + data[i] * x[bc * beta + c];
            }
        }
    }
}

/* ======================= BlockCOH ====================================== */

/* SOA_Implementations/BlockCOH.hpp:79-114 (Wikipedia xy2d with x = col,
 * y = row; the flip uses the full size). */
int64_t orc_hilbert(int64_t size, int64_t row, int64_t col) {
    int64_t index = 0;
    for (int64_t s = size / 2; s > 0; s /= 2) {
        int64_t rr = (row & s) > 0, cr = (col & s) > 0;
        index += s * s * ((3 * cr) ^ rr);
        if (rr == 0) {
            if (cr == 1) { col = size - 1 - col; row = size - 1 - row; }
            int64_t t = col; col = row; row = t;
        }
    }
    return index;
}

/* BlockCOH.hpp:325-328: fewer than 8 rows per thread lowers the thread
 * count to nor/8.  (The reference would then divide by zero for nor < 8;
 * this restatement keeps at least one band.) */
int orc_blockcoh_threads(int64_t nor, int threads) {
    if (threads < 1) threads = 1;
    if (nor / threads < 8) threads = (int)(nor / 8);
    return threads < 1 ? 1 : threads;
}

/* BlockCOH.hpp:332-355: open band pid+1 at row i when the running sum
 * (including row i) exceeds nnz/threads (integer division); the closed band
 * gets the sum without row i; the last band takes the rest.  Bands the scan
 * never opens are left empty here (start = nor, nnz = 0); the reference
 * leaves them uninitialised. */
void orc_blockcoh_bands(int64_t nor, int64_t nnz, const int32_t* row_counts, int threads,
                        int32_t* band_row_start, int64_t* band_nnz) {
    int64_t share = nnz / threads, sum = 0;
    int pid = 0;
    band_row_start[0] = 0;
    for (int64_t i = 0; i < nor; ++i) {
        sum += row_counts[i];
        if (sum > share && pid != threads - 1) {
            band_nnz[pid++] = sum - row_counts[i];
            band_row_start[pid] = (int32_t)i;
            sum = row_counts[i];
        }
    }
    band_nnz[pid] = sum;
    for (int p = pid + 1; p < threads; ++p) { band_row_start[p] = (int32_t)nor; band_nnz[p] = 0; }
}

/* BlockCOH.hpp:119-132 */
int orc_blockcoh_beta_bits(int64_t band_rows, int64_t noc) {
    int64_t mx = band_rows > noc ? band_rows : noc;
    int lg = (int)ceil(log2(sqrt((double)mx)));
    return 3 + lg < 15 ? 3 + lg : 15;
}

/* BlockCOH.hpp:401 */
int64_t orc_blockcoh_hilbert_size(int64_t band_rows, int64_t noc) {
    int64_t mx = band_rows > noc ? band_rows : noc;
    return (int64_t)pow(2., ceil(log2((double)mx)));
}

/* BlockCOH.hpp:380-452 for one band, starting from row-sorted, column-sorted
 * entries: order the band's entries by the Hilbert index of their block
 * (rowColToHilbert(hilbert_size/beta, local_row/beta, col/beta), :158-192),
 * keep (row, col) order inside a block (TripletToICRS, TripletToCRS.hpp:402-452),
 * and emit ICRS per block plus the block-level BICRS jumps (:433-439). */
void orc_blockcoh_build_band(int64_t noc, const int32_t* row_start, const int32_t* col_ind,
                             const double* vals, int64_t band_start, int64_t band_rows,
                             orc_coh_band* out) {
    int bits = orc_blockcoh_beta_bits(band_rows, noc);
    int64_t beta = (int64_t)1 << bits;
    int64_t hsize = orc_blockcoh_hilbert_size(band_rows, noc);
    int64_t H = orc_integer_ceil(noc, beta), V = orc_integer_ceil(band_rows, beta);
    int64_t e0 = row_start[band_start], e1 = row_start[band_start + band_rows], nnz = e1 - e0;
    out->nnz = nnz; out->beta_bits = bits;
    out->horizontal_blocks = (int32_t)H; out->vertical_blocks = (int32_t)V;
    out->blocks = 0; out->row_jumps = 0; out->block_row_jumps = 0;
    if (nnz == 0) return;
    kp_t* kp = (kp_t*)malloc(sizeof(kp_t) * nnz);
    int64_t* lrow = (int64_t*)malloc(sizeof(int64_t) * nnz);
    for (int64_t r = band_start; r < band_start + band_rows; ++r)
        for (int64_t k = row_start[r]; k < row_start[r + 1]; ++k) lrow[k - e0] = r - band_start;
    for (int64_t i = 0; i < nnz; ++i) {
        kp[i].key = (uint64_t)orc_hilbert(hsize / beta, lrow[i] / beta, col_ind[e0 + i] / beta);
        kp[i].pos = i;
    }
    qsort(kp, nnz, sizeof(kp_t), kp_cmp);
    int64_t rj = 0, nb = 0, rjb = 0;
    int64_t prev_br = -1, prev_bc = -1;
    int64_t blk_start = 0;
    for (int64_t i = 0; i <= nnz; ++i) {
        int end_block = (i == nnz) || (i > 0 && kp[i].key != kp[i - 1].key);
        if (end_block) {
            /* ICRS of entries [blk_start, i) */
            int64_t prev_r = -1, prev_c = -1;
            for (int64_t j = blk_start; j < i; ++j) {
                int64_t p = kp[j].pos;
                int64_t r = lrow[p] % beta, c = col_ind[e0 + p] % beta;
                out->data[j] = vals[e0 + p];
                if (j == blk_start) { out->row_jump[rj++] = (uint16_t)r; out->col_jump[j] = (uint16_t)c; }
                else if (r != prev_r) { out->col_jump[j] = (uint16_t)(c + beta - prev_c); out->row_jump[rj++] = (uint16_t)(r - prev_r); }
                else out->col_jump[j] = (uint16_t)(c - prev_c);
                prev_r = r; prev_c = c;
            }
            int64_t p0 = kp[blk_start].pos;
            int64_t br = lrow[p0] / beta, bc = col_ind[e0 + p0] / beta;
            if (nb == 0) { out->row_jump_block[rjb++] = (int16_t)br; out->col_jump_block[0] = (int16_t)bc; }
            else if (br == prev_br) out->col_jump_block[nb] = (int16_t)(bc - prev_bc);
            else { out->row_jump_block[rjb++] = (int16_t)(br - prev_br); out->col_jump_block[nb] = (int16_t)(bc - prev_bc + H); }
            out->block_nnz[nb] = (int32_t)(i - blk_start);
            prev_br = br; prev_bc = bc;
            nb++;
            blk_start = i;
        }
    }
    out->blocks = nb; out->row_jumps = rj; out->block_row_jumps = rjb;
    free(kp); free(lrow);
}

/* BlockCOH.hpp:479-510 (band loop) and :235-256 (ICRS block multiply). */
void orc_blockcoh_mv_band(const orc_coh_band* b, int64_t band_start, int64_t band_rows,
                          const double* x, double* y) {
    for (int64_t r = 0; r < band_rows; ++r) y[band_start + r] = 0.;
    if (b->blocks == 0) return;
    int64_t beta = (int64_t)1 << b->beta_bits, H = b->horizontal_blocks;
    int64_t block_row = b->row_jump_block[0], block_col = b->col_jump_block[0];
    int64_t rbi = 1, ri = 0, ci = 0;
    for (int64_t blk = 0; blk < b->blocks; ++blk) {
        const double* xb = x + block_col * beta;
        double* yb = y + band_start + block_row * beta;
        int64_t row = b->row_jump[ri++], col = b->col_jump[ci];
        for (int64_t i = 0; i < b->block_nnz[blk] - 1; ++i) {
            yb[row] = yb[row] + b->data[ci] * xb[col];
            ci++;
            col += b->col_jump[ci];
            if (col >= beta) { col -= beta; row += b->row_jump[ri++]; }
        }
        yb[row] = yb[row] + b->data[ci] * xb[col];
        ci++;
        if (blk + 1 < b->blocks) {
            block_col += b->col_jump_block[blk + 1];
            if (block_col >= H) { block_col -= H; block_row += b->row_jump_block[rbi++]; }
        }
    }
}

/* ======================= matrix statistics ============================= */

/* Util_Scripts/calc_stats.cpp:44-105 on a CRS matrix (the same multiset of
 * (row, col) pairs as the Triplet it reads):
 *   rows_avg = nnz / N; rows_var = sum_i (len_i - rows_avg)^2 / N, summed
 *   over the rows in order;
 *   block_bits = min(16, 3 + ceil(log2(sqrt(N)))), nb = ceil(N / 2^bits)
 *   block rows and columns (square, as the tool assumes), one fewer when
 *   2^bits does not divide N (whole blocks only);
 *   block_avg = sum_ij nnz_ij / nb^2 and block_var = sum_ij (nnz_ij -
 *   block_avg)^2 / nb^2 over the whole blocks, i-major;
 *   balanced = no whole block row holds more than twice the mean block-row
 *   count (sum_i row_i / nb).
 * Entries whose column block lies beyond the square's blocks are not counted
 * (the tool would index past its block table there).  row_max (the longest
 * row) is an addition. */
void orc_matrix_stats(int64_t nor, const int32_t* row_start, const int32_t* col_ind, double* rows_avg,
                      double* rows_var, int64_t* row_max, int* block_bits, int64_t* nb_blocks, double* block_avg,
                      double* block_var, int* balanced) {
    const int64_t nnz = row_start[nor];
    const double avg = (double)nnz / (double)nor;
    int lg = (int)ceil(log2(sqrt((double)nor)));
    int bits = 16 < 3 + lg ? 16 : 3 + lg;
    int64_t bs = (int64_t)1 << bits;
    int64_t nb = orc_integer_ceil(nor, bs);
    int64_t* blk = (int64_t*)calloc((size_t)(nb * nb > 0 ? nb * nb : 1), sizeof(int64_t));
    double var = 0.0;
    int64_t mx = 0;
    for (int64_t i = 0; i < nor; ++i) {
        const int64_t len = row_start[i + 1] - row_start[i];
        var += pow((double)len - avg, 2) / (double)nor;
        if (len > mx) mx = len;
        for (int64_t k = row_start[i]; k < row_start[i + 1]; ++k) {
            const int64_t bc = col_ind[k] / bs;
            if (bc < nb) blk[(i / bs) * nb + bc]++;
        }
    }
    const int64_t nbw = nor % bs != 0 ? nb - 1 : nb;
    double bavg = 0.0, brow_avg = 0.0, bvar = 0.0;
    int bal = 1;
    int64_t* brow = (int64_t*)calloc((size_t)(nbw > 0 ? nbw : 1), sizeof(int64_t));
    for (int64_t i = 0; i < nbw; ++i) {
        for (int64_t j = 0; j < nbw; ++j) {
            brow[i] += blk[i * nb + j];
            bavg += (double)blk[i * nb + j] / pow((double)nbw, 2);
        }
        brow_avg += (double)brow[i] / (double)nbw;
    }
    for (int64_t i = 0; i < nbw; ++i) {
        if ((double)brow[i] > 2 * brow_avg) bal = 0;
        for (int64_t j = 0; j < nbw; ++j) bvar += pow((double)blk[i * nb + j] - bavg, 2) / pow((double)nbw, 2);
    }
    free(brow);
    free(blk);
    *rows_avg = avg;
    *rows_var = var;
    *row_max = mx;
    *block_bits = bits;
    *nb_blocks = nbw;
    *block_avg = bavg;
    *block_var = bvar;
    *balanced = bal;
}
