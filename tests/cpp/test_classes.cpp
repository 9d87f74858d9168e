// C++ parity driver for the drop-in classes (include/pwm/SparseMatrixHip.hpp),
// written the way the reference's Boost tests run (Tests/poisson.cpp,
// Tests/input.cpp): every GPU class, int and unsigned int indices, x = ones,
// mv and powerIteration, checked against the oracle (oracle/pwm_oracle.h).
// Exit status 0 = all checks passed.
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <vector>

#include "../../include/pwm/SparseMatrixHip.hpp"
#include "../../oracle/pwm_oracle.h"

static int failures = 0;

static void expect_close(const char* what, const std::vector<double>& got, const std::vector<double>& ref,
                         double tol) {
    double mx = 0, err = 0;
    for (size_t i = 0; i < ref.size(); ++i) {
        mx = std::fmax(mx, std::fabs(ref[i]));
        err = std::fmax(err, std::fabs(got[i] - ref[i]));
    }
    const bool ok = err <= tol * (mx > 0 ? mx : 1.0);
    if (!ok) ++failures;
    std::printf("%-58s rel.err %.3e %s\n", what, mx > 0 ? err / mx : err, ok ? "ok" : "FAIL");
}

template <typename I>
static void run_all(const char* itype) {
    // Poisson 9x3 generated on the device vs fillPoisson + CRS::mv
    const int m = 9, n = 3, N = m * n;
    const int64_t Z = orc_poisson_nnz(m, n);
    std::vector<int32_t> rs(N + 1), ci(Z);
    std::vector<double> dv(Z);
    orc_fill_poisson(m, n, rs.data(), ci.data(), dv.data());

    // random unsorted triplets with duplicates and empty rows
    const int R = 300;
    std::mt19937 g(5);
    std::vector<I> tr, tc;
    std::vector<double> tv;
    for (int k = 0; k < 2500; ++k) {
        tr.push_back((I)(g() % (R - 20)));  // the last 20 rows stay empty
        tc.push_back((I)(g() % R));
        tv.push_back((double)(int)(g() % 2001) / 1000.0 - 1.0);
    }
    pwm::Triplet<double, I> t;
    t.row_coord = tr.data();
    t.col_coord = tc.data();
    t.data = tv.data();
    t.row_size = R;
    t.col_size = R;
    t.nnz = (I)tr.size();
    std::vector<int32_t> trs(R + 1), tci(tr.size());
    std::vector<double> tdv(tr.size());
    std::vector<int32_t> r32(tr.begin(), tr.end()), c32(tc.begin(), tc.end());
    orc_triplet_to_crs(R, (int64_t)tr.size(), r32.data(), c32.data(), tv.data(), trs.data(), tci.data(), tdv.data());

    for (int method : {101, 102, 108, 109, 110}) {
        for (int threads : {1, 3, -1}) {
            std::unique_ptr<pwm::SparseMatrix<double, I>> mat(pwm::selectTypeHip<double, I>(method, threads));
            char name[128];
            std::vector<double> x(N, 1.0), y(N, 0.0), ref(N);
            mat->generatePoissonMatrix(m, n, 1);
            mat->mv(x.data(), y.data());
            orc_crs_mv(N, rs.data(), ci.data(), dv.data(), x.data(), ref.data());
            std::snprintf(name, sizeof name, "[%s] id %d threads %d poisson 9x3 mv", itype, method, threads);
            expect_close(name, y, ref, 1e-14);

            std::vector<double> px(N, 1.0), py(N, 0.0), rx(N, 1.0), ry(N, 0.0);
            mat->powerIteration(px.data(), py.data(), (I)7);
            orc_crs_power_iteration(N, rs.data(), ci.data(), dv.data(), rx.data(), ry.data(), 7);
            std::snprintf(name, sizeof name, "[%s] id %d threads %d poisson 9x3 power(7) x", itype, method, threads);
            expect_close(name, px, rx, 1e-12);
            std::snprintf(name, sizeof name, "[%s] id %d threads %d poisson 9x3 power(7) y", itype, method, threads);
            expect_close(name, py, ry, 1e-12);

            mat->loadFromTriplets(&t, 1);
            std::vector<double> xr(R), yr(R, 0.0), refr(R);
            for (int i = 0; i < R; ++i) xr[i] = std::sin(0.37 * i);
            mat->mv(xr.data(), yr.data());
            orc_crs_mv(R, trs.data(), tci.data(), tdv.data(), xr.data(), refr.data());
            std::snprintf(name, sizeof name, "[%s] id %d threads %d triplets mv", itype, method, threads);
            expect_close(name, yr, refr, 1e-13);
        }
    }
    // Matrix-Market file read by the library's reader, loaded through
    // loadFromTriplets, against the oracle's loader + TripletToCRS + CRS::mv
    {
        const char* path = PWM_FIXTURE_DIR "/gre_1107.mtx";
        pwm::Triplet<double, I> tm;
        pwm::readMatrixMarket(path, true, false, false, tm);
        int64_t nr = 0, nc = 0, nz = 0;
        int32_t *orr = nullptr, *occ = nullptr;
        double* ov = nullptr;
        if (orc_load_mm(path, 1, 0, 0, &nr, &nc, &nz, &orr, &occ, &ov) != 0 || nz != (int64_t)tm.nnz) {
            ++failures;
            std::printf("[%s] reader: entry count FAIL\n", itype);
        } else {
            std::vector<int32_t> mrs(nr + 1), mci(nz);
            std::vector<double> mdv(nz);
            orc_triplet_to_crs(nr, nz, orr, occ, ov, mrs.data(), mci.data(), mdv.data());
            std::unique_ptr<pwm::SparseMatrix<double, I>> mat(pwm::selectTypeHip<double, I>(108, -1));
            mat->loadFromTriplets(&tm, 1);
            std::vector<double> xm(nr), ym(nr, 0.0), refm(nr);
            for (int64_t i = 0; i < nr; ++i) xm[i] = std::cos(0.11 * (double)i);
            mat->mv(xm.data(), ym.data());
            orc_crs_mv(nr, mrs.data(), mci.data(), mdv.data(), xm.data(), refm.data());
            char name[128];
            std::snprintf(name, sizeof name, "[%s] readMatrixMarket gre_1107 + loadFromTriplets mv", itype);
            expect_close(name, ym, refm, 1e-13);
        }
        orc_free(orr);
        orc_free(occ);
        orc_free(ov);
        delete[] tm.row_coord;
        delete[] tm.col_coord;
        delete[] tm.data;
    }
    // unknown id -> nullptr, like selectType's NULL
    if (pwm::selectTypeHip<double, I>(7, 1) != nullptr) {
        ++failures;
        std::printf("selectTypeHip(7) should be nullptr FAIL\n");
    }
}

int main() {
    try {
        run_all<int>("int");
        run_all<unsigned int>("unsigned");
    } catch (const pwm::HipError& e) {
        std::printf("HipError %d: %s\n", e.code, e.what());
        return 2;
    }
    std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
