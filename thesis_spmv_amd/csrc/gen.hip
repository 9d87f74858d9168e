// Deterministic generators of the benchmark matrices (host side).
//
// Every row draws from its own counter-based stream (splitmix64 of the seed
// and the row index), so the output does not depend on the thread count.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "api_internal.hpp"

namespace {

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline double u01_open(uint64_t& s) {  // (0, 1]
    return ((splitmix64(s) >> 11) + 1) * (1.0 / 9007199254740992.0);
}
inline uint64_t row_stream(uint64_t seed, int64_t row, uint64_t salt) {
    uint64_t s = seed * 0x632be59bd9b4e019ull + salt;
    uint64_t t = s ^ ((uint64_t)row * 0xd6e8feb86659fd93ull);
    splitmix64(t);
    return t;
}
inline double normal01(uint64_t& s) {  // Box-Muller
    const double u1 = u01_open(s), u2 = u01_open(s);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// `count` distinct sorted columns in [0, n) not in `exclude` (sorted).
void sample_columns(uint64_t& st, int64_t n, int64_t count, const std::vector<int32_t>& exclude,
                    std::vector<int32_t>& out) {
    out.clear();
    const int64_t avail = n - (int64_t)exclude.size();
    if (count > avail) count = avail;
    if (count <= 0) return;
    if (count * 4 > avail) {  // dense: partial Fisher-Yates over the allowed set
        std::vector<int32_t> all;
        all.reserve(avail);
        size_t e = 0;
        for (int64_t c = 0; c < n; ++c) {
            while (e < exclude.size() && exclude[e] < c) ++e;
            if (e < exclude.size() && exclude[e] == c) continue;
            all.push_back((int32_t)c);
        }
        for (int64_t i = 0; i < count; ++i) {
            const int64_t j = i + (int64_t)(splitmix64(st) % (uint64_t)(avail - i));
            std::swap(all[i], all[j]);
        }
        out.assign(all.begin(), all.begin() + count);
        std::sort(out.begin(), out.end());
        return;
    }
    while ((int64_t)out.size() < count) {
        const int64_t need = count - (int64_t)out.size();
        for (int64_t i = 0; i < need; ++i) out.push_back((int32_t)(splitmix64(st) % (uint64_t)n));
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
        if (!exclude.empty()) {
            std::vector<int32_t> keep;
            keep.reserve(out.size());
            std::set_difference(out.begin(), out.end(), exclude.begin(), exclude.end(), std::back_inserter(keep));
            out.swap(keep);
        }
    }
}

template <typename T>
T* dup_out(const std::vector<T>& v) {
    T* p = (T*)malloc(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (p && !v.empty()) memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace

extern "C" void pwm_free(void* p) { free(p); }

extern "C" int pwm_gen_uniform(int64_t n, uint64_t seed, double* out) {
    if (n < 0 || (n > 0 && !out)) return pwm_impl::fail(nullptr, PWM_ERR_INVALID, "pwm_gen_uniform: bad argument");
    std::mt19937_64 g(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = (double)(g() >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    return PWM_OK;
}

extern "C" int pwm_gen_powerlaw(int64_t n, double alpha, double xm, int64_t cap, uint64_t struct_seed,
                                uint64_t value_seed, int32_t** row_start, int32_t** col_ind, double** vals,
                                int64_t* nnz) {
    if (n <= 0 || n > INT32_MAX - 1 || alpha <= 0 || xm <= 0 || cap < 0 || !row_start || !col_ind || !vals || !nnz)
        return pwm_impl::fail(nullptr, PWM_ERR_INVALID, "pwm_gen_powerlaw: bad argument");
    std::vector<int64_t> len(n);
    pwm_impl::parallel_for(n, 1 << 16, [&](int64_t b, int64_t e, int) {
        for (int64_t r = b; r < e; ++r) {
            uint64_t st = row_stream(struct_seed, r, 1);
            const double u = u01_open(st);
            double L = floor(xm * pow(u, -1.0 / alpha));
            if (L > (double)cap) L = (double)cap;
            if (L > (double)n) L = (double)n;
            len[r] = (int64_t)L;
        }
    });
    std::vector<int32_t> rs(n + 1, 0);
    int64_t acc = 0;
    for (int64_t r = 0; r < n; ++r) {
        rs[r] = (int32_t)acc;
        acc += len[r];
        if (acc > INT32_MAX) return pwm_impl::fail(nullptr, PWM_ERR_INVALID, "pwm_gen_powerlaw: nnz >= 2^31");
    }
    rs[n] = (int32_t)acc;
    int32_t* ci = (int32_t*)malloc(sizeof(int32_t) * std::max<int64_t>(acc, 1));
    double* dv = (double*)malloc(sizeof(double) * std::max<int64_t>(acc, 1));
    if (!ci || !dv) {
        free(ci);
        free(dv);
        return pwm_impl::fail(nullptr, PWM_ERR_NOMEM, "pwm_gen_powerlaw: host allocation");
    }
    pwm_impl::parallel_for(n, 1024, [&](int64_t b, int64_t e, int) {
        std::vector<int32_t> cols;
        const std::vector<int32_t> none;
        for (int64_t r = b; r < e; ++r) {
            uint64_t st = row_stream(struct_seed, r, 2);
            sample_columns(st, n, len[r], none, cols);
            uint64_t sv = row_stream(value_seed, r, 3);
            for (size_t i = 0; i < cols.size(); ++i) {
                ci[rs[r] + i] = cols[i];
                dv[rs[r] + i] = normal01(sv);
            }
        }
    });
    *row_start = dup_out(rs);
    *col_ind = ci;
    *vals = dv;
    *nnz = acc;
    return PWM_OK;
}

// Symmetric positive definite power-law matrix: Pareto row lengths (as
// pwm_gen_powerlaw, capped at n-1), half of each row's entries drawn
// uniformly off the diagonal and mirrored, values +-U(0,1] (duplicates
// summed), then the diagonal set to the row's absolute sum + 1 (strict
// diagonal dominance with a positive diagonal, hence SPD).
extern "C" int pwm_gen_spd_powerlaw(int64_t n, uint64_t seed, int32_t** row_start, int32_t** col_ind, double** vals,
                                    int64_t* nnz) {
    if (n <= 1 || n > (1 << 26) || !row_start || !col_ind || !vals || !nnz)
        return pwm_impl::fail(nullptr, PWM_ERR_INVALID, "pwm_gen_spd_powerlaw: bad argument");
    struct E {
        int32_t r, c;
        double v;
    };
    std::vector<std::vector<E>> local(16);  // one list per parallel_for worker
    // draw the upper/lower halves per row, keep mirrors per worker
    pwm_impl::parallel_for(n, 2048, [&](int64_t b, int64_t e, int w) {
        std::vector<int32_t> cols;
        const std::vector<int32_t> none;
        for (int64_t r = b; r < e; ++r) {
            uint64_t st = row_stream(seed, r, 5);
            double L = floor(7.3333 * pow(u01_open(st), -1.0 / 1.8));
            L = std::min(L, (double)std::min<int64_t>(50000, n - 1));
            const int64_t half = (int64_t)(L + 1) / 2;
            std::vector<int32_t> ex(1, (int32_t)r);
            sample_columns(st, n, half, ex, cols);
            for (int32_t c : cols) {
                const double mag = u01_open(st);
                const double v = (splitmix64(st) & 1) ? mag : -mag;
                local[w].push_back({(int32_t)r, c, v});
                local[w].push_back({c, (int32_t)r, v});
            }
        }
    });
    std::vector<int64_t> cnt(n + 1, 0);
    for (auto& L : local)
        for (const E& x : L) cnt[x.r + 1]++;
    for (int64_t r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
    std::vector<E> all(cnt[n]);
    {
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (auto& L : local)
            for (const E& x : L) all[pos[x.r]++] = x;
    }
    local.clear();
    // per row: sort by column, merge duplicates, add the diagonal
    std::vector<int64_t> rowlen(n, 0);
    pwm_impl::parallel_for(n, 2048, [&](int64_t b, int64_t e, int) {
        for (int64_t r = b; r < e; ++r) {
            E* s = all.data() + cnt[r];
            E* t = all.data() + cnt[r + 1];
            std::sort(s, t, [](const E& p, const E& q) { return p.c < q.c; });
            int64_t m = 0;
            for (E* p = s; p < t; ++p) {
                if (m > 0 && s[m - 1].c == p->c) s[m - 1].v += p->v;
                else s[m++] = *p;
            }
            rowlen[r] = m + 1;  // + diagonal
        }
    });
    std::vector<int32_t> rs(n + 1, 0);
    for (int64_t r = 0; r < n; ++r) {
        if ((int64_t)rs[r] + rowlen[r] > INT32_MAX) return pwm_impl::fail(nullptr, PWM_ERR_INVALID, "spd_powerlaw: nnz >= 2^31");
        rs[r + 1] = rs[r] + (int32_t)rowlen[r];
    }
    const int64_t Z = rs[n];
    int32_t* ci = (int32_t*)malloc(sizeof(int32_t) * Z);
    double* dv = (double*)malloc(sizeof(double) * Z);
    if (!ci || !dv) {
        free(ci);
        free(dv);
        return pwm_impl::fail(nullptr, PWM_ERR_NOMEM, "pwm_gen_spd_powerlaw: host allocation");
    }
    pwm_impl::parallel_for(n, 2048, [&](int64_t b, int64_t e, int) {
        for (int64_t r = b; r < e; ++r) {
            const E* s = all.data() + cnt[r];
            const int64_t m = rowlen[r] - 1;
            double asum = 0.0;
            for (int64_t k = 0; k < m; ++k) asum += fabs(s[k].v);
            int64_t o = rs[r];
            bool placed = false;
            for (int64_t k = 0; k < m; ++k) {
                if (!placed && s[k].c > r) {
                    ci[o] = (int32_t)r;
                    dv[o++] = asum + 1.0;
                    placed = true;
                }
                ci[o] = s[k].c;
                dv[o++] = s[k].v;
            }
            if (!placed) {
                ci[o] = (int32_t)r;
                dv[o++] = asum + 1.0;
            }
        }
    });
    *row_start = dup_out(rs);
    *col_ind = ci;
    *vals = dv;
    *nnz = Z;
    return PWM_OK;
}

extern "C" int pwm_gen_stencil9(int64_t side, int64_t extra, uint64_t seed, int32_t** row_start, int32_t** col_ind,
                                double** vals, int64_t* nnz) {
    if (side <= 0 || extra < 0 || side * side > INT32_MAX - 1 || !row_start || !col_ind || !vals || !nnz)
        return pwm_impl::fail(nullptr, PWM_ERR_INVALID, "pwm_gen_stencil9: bad argument");
    const int64_t N = side * side;
    std::vector<int32_t> rs(N + 1, 0);
    for (int64_t r = 0; r < N; ++r) {
        const int64_t i = r / side, j = r % side;
        const int64_t ni = 1 + (i > 0) + (i < side - 1), nj = 1 + (j > 0) + (j < side - 1);
        const int64_t ex = std::min<int64_t>(extra, N - ni * nj);
        rs[r + 1] = rs[r] + (int32_t)(ni * nj + ex);
        if ((int64_t)rs[r + 1] > INT32_MAX - 64) return pwm_impl::fail(nullptr, PWM_ERR_INVALID, "stencil9: nnz >= 2^31");
    }
    const int64_t Z = rs[N];
    int32_t* ci = (int32_t*)malloc(sizeof(int32_t) * Z);
    double* dv = (double*)malloc(sizeof(double) * Z);
    if (!ci || !dv) {
        free(ci);
        free(dv);
        return pwm_impl::fail(nullptr, PWM_ERR_NOMEM, "pwm_gen_stencil9: host allocation");
    }
    pwm_impl::parallel_for(N, 4096, [&](int64_t b, int64_t e, int) {
        std::vector<int32_t> sten, off;
        std::vector<std::pair<int32_t, double>> row;
        for (int64_t r = b; r < e; ++r) {
            const int64_t i = r / side, j = r % side;
            sten.clear();
            for (int di = -1; di <= 1; ++di)
                for (int dj = -1; dj <= 1; ++dj) {
                    const int64_t ii = i + di, jj = j + dj;
                    if (ii < 0 || ii >= side || jj < 0 || jj >= side) continue;
                    sten.push_back((int32_t)(ii * side + jj));
                }
            std::sort(sten.begin(), sten.end());
            const int64_t ex = rs[r + 1] - rs[r] - (int64_t)sten.size();
            uint64_t st = row_stream(seed, r, 4);
            sample_columns(st, N, ex, sten, off);
            row.clear();
            for (int32_t c : sten) row.push_back({c, c == (int32_t)r ? 8.0 : -1.0});
            for (int32_t c : off) row.push_back({c, u01_open(st) * 2.0 - 1.0});
            std::sort(row.begin(), row.end(),
                      [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
                          return a.first < b.first;
                      });
            for (size_t k = 0; k < row.size(); ++k) {
                ci[rs[r] + k] = row[k].first;
                dv[rs[r] + k] = row[k].second;
            }
        }
    });
    *row_start = dup_out(rs);
    *col_ind = ci;
    *vals = dv;
    *nnz = Z;
    return PWM_OK;
}
