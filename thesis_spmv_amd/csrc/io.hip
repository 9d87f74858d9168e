// Host-side matrix readers: Matrix-Market and the Kronecker .bin format, into
// COO arrays for pwm_load_triplets (Triplet::loadFromMM / loadFromBin,
// Matrix/Triplet.hpp:75-216).
//
// Matrix-Market (Triplet.hpp:75-154): lines starting with '%' are skipped;
// the first other line gives rows, cols, entries; every entry line is
// "row col [value]" with 1-based indices; with has_data an explicit 0 is
// dropped; without it the value is 1, or (random_fill) a draw of
// uniform_real(-100, 100) from mt19937 seeded 747846 (Triplet.hpp:47-54); with
// symmetric every off-diagonal entry is followed by its mirror; reading stops
// at the first empty line.  The banner is not parsed: the caller passes the
// flags, as the reference's callers do.
//
// .bin (Triplet.hpp:164-216): the file is a sequence of little-endian uint32
// (row, col) pairs, 0-based, no header (entries = file size / 8); the matrix
// is mat_size x mat_size; values 1 or random as above; symmetric mirrors
// off-diagonal pairs.  Duplicates are kept in both formats.
//
// The text is parsed in parallel: the data section (cut at its first empty
// line) is split at line boundaries, every chunk is parsed into its own
// arrays and the chunks are concatenated in file order, so the result is
// the sequential one.  Random values are drawn afterwards in entry order.
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <charconv>
#include <random>
#include <vector>

#include "api_internal.hpp"

namespace {

struct Coo {
    std::vector<int32_t> r, c;
    std::vector<double> v;
    std::vector<char> mirrored;  // 1: entry is the mirror of the previous one
};

// Boost's uniform_real_distribution over a 32-bit engine: u / 2^32 scaled to
// [lo, hi), redrawn when rounding lands on hi.
double uniform_real(std::mt19937& g, double lo, double hi) {
    for (;;) {
        const double u = (double)g() / 4294967296.0;
        const double r = u * (hi - lo) + lo;
        if (r < hi) return r;
    }
}

// Values of the pattern entries: one draw per stored entry that is not a
// mirror (a mirror copies the value before it), in entry order.
void fill_values(Coo& m, bool random) {
    if (!random) return;
    std::mt19937 g(747846u);
    for (size_t k = 0; k < m.v.size(); ++k) m.v[k] = m.mirrored[k] ? m.v[k - 1] : uniform_real(g, -100.0, 100.0);
}

bool read_file(const char* path, std::vector<char>& buf) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return false;
    }
    const long sz = ftell(f);
    if (sz < 0 || fseek(f, 0, SEEK_SET) != 0) {
        fclose(f);
        return false;
    }
    buf.resize((size_t)sz + 1);
    const size_t got = sz ? fread(buf.data(), 1, (size_t)sz, f) : 0;
    fclose(f);
    if (got != (size_t)sz) return false;
    buf[(size_t)sz] = '\0';
    buf.resize((size_t)sz);
    return true;
}

const char* skip_spaces(const char* p, const char* e) {
    while (p < e && (*p == ' ' || *p == '\t')) ++p;
    return p;
}

// Parse lines [p, e) (whole lines).  Returns false on a malformed entry.
// Stops (without error) at a line whose indices do not parse.
bool parse_entries(const char* p, const char* e, bool has_data, bool symmetric, Coo& out, bool& stopped) {
    stopped = false;
    while (p < e) {
        const char* eol = (const char*)memchr(p, '\n', (size_t)(e - p));
        if (!eol) eol = e;
        long long ri = 0, ci = 0;
        const char* q = skip_spaces(p, eol);
        auto a = std::from_chars(q, eol, ri);
        if (a.ec != std::errc()) {
            stopped = true;
            return true;
        }
        q = skip_spaces(a.ptr, eol);
        auto b = std::from_chars(q, eol, ci);
        if (b.ec != std::errc()) {
            stopped = true;
            return true;
        }
        double v = 1.0;
        if (has_data) {
            q = skip_spaces(b.ptr, eol);
            auto c = std::from_chars(q, eol, v);
            if (c.ec != std::errc()) return false;
        }
        p = eol < e ? eol + 1 : e;
        if (has_data && v == 0.0) continue;
        if (ri < 1 || ci < 1 || ri > INT32_MAX || ci > INT32_MAX) return false;
        out.r.push_back((int32_t)(ri - 1));
        out.c.push_back((int32_t)(ci - 1));
        out.v.push_back(v);
        out.mirrored.push_back(0);
        if (symmetric && ri != ci) {
            out.r.push_back((int32_t)(ci - 1));
            out.c.push_back((int32_t)(ri - 1));
            out.v.push_back(v);
            out.mirrored.push_back(1);
        }
    }
    return true;
}

int hand_out(const Coo& m, int32_t** rows, int32_t** cols, double** vals, int64_t* nnz) {
    const size_t n = m.r.size();
    *rows = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
    *cols = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
    *vals = (double*)malloc(sizeof(double) * (n ? n : 1));
    if (!*rows || !*cols || !*vals) {
        free(*rows);
        free(*cols);
        free(*vals);
        *rows = nullptr;
        *cols = nullptr;
        *vals = nullptr;
        return PWM_ERR_NOMEM;
    }
    if (n) {
        memcpy(*rows, m.r.data(), n * 4);
        memcpy(*cols, m.c.data(), n * 4);
        memcpy(*vals, m.v.data(), n * 8);
    }
    *nnz = (int64_t)n;
    return PWM_OK;
}

int io_fail(int code, const std::string& msg) {
    pwm_impl::set_global_error(msg);
    return code;
}

}  // namespace

extern "C" int pwm_read_mm(const char* path, int has_data, int symmetric, int random_fill, int64_t* nor, int64_t* noc,
                           int64_t* nnz, int32_t** rows, int32_t** cols, double** vals) {
    if (!path || !nor || !noc || !nnz || !rows || !cols || !vals) return io_fail(PWM_ERR_INVALID, "pwm_read_mm: NULL argument");
    std::vector<char> buf;
    if (!read_file(path, buf)) return io_fail(PWM_ERR_IO, std::string("pwm_read_mm: cannot read ") + path);
    const char* p = buf.data();
    const char* e = p + buf.size();
    // header: '%' lines, then the size line
    long long R = -1, C = -1, E = -1;
    for (;;) {
        if (p >= e) return io_fail(PWM_ERR_IO, "pwm_read_mm: no size line");
        const char* eol = (const char*)memchr(p, '\n', (size_t)(e - p));
        if (!eol) eol = e;
        if (*p != '%') {
            const char* q = skip_spaces(p, eol);
            auto a = std::from_chars(q, eol, R);
            q = skip_spaces(a.ptr, eol);
            auto b = std::from_chars(q, eol, C);
            q = skip_spaces(b.ptr, eol);
            auto c = std::from_chars(q, eol, E);
            if (a.ec != std::errc() || b.ec != std::errc() || c.ec != std::errc() || R < 0 || C < 0 || E < 0 ||
                R > INT32_MAX || C > INT32_MAX)
                return io_fail(PWM_ERR_IO, "pwm_read_mm: malformed size line");
            p = eol < e ? eol + 1 : e;
            break;
        }
        p = eol < e ? eol + 1 : e;
    }
    // the data section ends at the first empty line
    const char* stop = e;
    for (const char* q = p; q < e;) {
        if (*q == '\n' || (*q == '\r' && q + 1 < e && q[1] == '\n')) {
            stop = q;
            break;
        }
        const char* eol = (const char*)memchr(q, '\n', (size_t)(e - q));
        if (!eol) break;
        q = eol + 1;
    }
    // chunks at line boundaries
    const int64_t bytes = stop - p;
    const int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(64, bytes / (1 << 20)));
    std::vector<const char*> cut(nchunk + 1);
    cut[0] = p;
    cut[nchunk] = stop;
    for (int i = 1; i < nchunk; ++i) {
        const char* q = p + bytes * i / nchunk;
        if (q < cut[i - 1]) q = cut[i - 1];
        const char* eol = (const char*)memchr(q, '\n', (size_t)(stop - q));
        cut[i] = eol ? eol + 1 : stop;
    }
    std::vector<Coo> part(nchunk);
    std::vector<char> ok(nchunk, 1), stopped(nchunk, 0);
    pwm_impl::parallel_for(nchunk, 1, [&](int64_t b, int64_t en, int) {
        for (int64_t i = b; i < en; ++i) {
            bool st = false;
            ok[i] = parse_entries(cut[i], cut[i + 1], has_data != 0, symmetric != 0, part[i], st);
            stopped[i] = st;
        }
    });
    Coo all;
    int64_t lines = 0;
    for (int i = 0; i < nchunk; ++i) {
        if (!ok[i]) return io_fail(PWM_ERR_IO, "pwm_read_mm: malformed entry");
        all.r.insert(all.r.end(), part[i].r.begin(), part[i].r.end());
        all.c.insert(all.c.end(), part[i].c.begin(), part[i].c.end());
        all.v.insert(all.v.end(), part[i].v.begin(), part[i].v.end());
        all.mirrored.insert(all.mirrored.end(), part[i].mirrored.begin(), part[i].mirrored.end());
        for (char m : part[i].mirrored) lines += m ? 0 : 1;
        if (stopped[i]) break;  // a line without indices ends the data, as an empty line does
    }
    if (lines > E) return io_fail(PWM_ERR_IO, "pwm_read_mm: more entries than the size line declares");
    for (size_t k = 0; k < all.r.size(); ++k)
        if (all.r[k] >= R || all.c[k] >= C) return io_fail(PWM_ERR_IO, "pwm_read_mm: entry outside the declared size");
    fill_values(all, !has_data && random_fill);
    *nor = R;
    *noc = C;
    const int rc = hand_out(all, rows, cols, vals, nnz);
    return rc == PWM_OK ? rc : io_fail(rc, "pwm_read_mm: out of host memory");
}

extern "C" int pwm_read_bin(const char* path, int64_t mat_size, int symmetric, int random_vals, int64_t* nor,
                            int64_t* noc, int64_t* nnz, int32_t** rows, int32_t** cols, double** vals) {
    if (!path || !nor || !noc || !nnz || !rows || !cols || !vals || mat_size < 0 || mat_size > INT32_MAX)
        return io_fail(PWM_ERR_INVALID, "pwm_read_bin: bad argument");
    std::vector<char> buf;
    if (!read_file(path, buf)) return io_fail(PWM_ERR_IO, std::string("pwm_read_bin: cannot read ") + path);
    const int64_t pairs = (int64_t)buf.size() / 8;
    Coo m;
    m.r.reserve(symmetric ? 2 * pairs : pairs);
    m.c.reserve(m.r.capacity());
    m.v.reserve(m.r.capacity());
    m.mirrored.reserve(m.r.capacity());
    for (int64_t k = 0; k < pairs; ++k) {
        uint32_t rc[2];
        memcpy(rc, buf.data() + 8 * k, 8);
        if ((int64_t)rc[0] >= mat_size || (int64_t)rc[1] >= mat_size)
            return io_fail(PWM_ERR_IO, "pwm_read_bin: pair outside the matrix");
        m.r.push_back((int32_t)rc[0]);
        m.c.push_back((int32_t)rc[1]);
        m.v.push_back(1.0);
        m.mirrored.push_back(0);
        if (symmetric && rc[0] != rc[1]) {
            m.r.push_back((int32_t)rc[1]);
            m.c.push_back((int32_t)rc[0]);
            m.v.push_back(1.0);
            m.mirrored.push_back(1);
        }
    }
    fill_values(m, random_vals != 0);
    *nor = mat_size;
    *noc = mat_size;
    const int rc = hand_out(m, rows, cols, vals, nnz);
    return rc == PWM_OK ? rc : io_fail(rc, "pwm_read_bin: out of host memory");
}
