// hlanc_demo.cpp -- the four scenarios of the reference's demo/HLanc_demo.cu,
// built against the drop-in header: a Matrix Market file, a pp file, the pp
// matrix through a bin round trip, and a plain function as the operator.
// Each solve is checked (exit status 0 only when all four agree with the
// closed form / dense eigenvalues), so the program doubles as an integration test.
//   usage: hlanc_demo <golden_dir> [scratch_dir]
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "HLanc.hpp"

static void op_fun_simple(double const *x, double *y)
{
    for (int i = 0; i < 16; ++i) {
        double s = 0.0;
        for (int j = 0; j < 16; ++j) s += ((i < j ? i : j) + 1) * x[j];
        y[i] = s;
    }
}

static int check(char const *what, int ret, double const *values, double const *ref, int nev)
{
    double err = 0.0;
    for (int i = 0; i < nev; ++i) err = std::fmax(err, std::fabs(values[i] - ref[i]) / std::fabs(ref[i]));
    const bool ok = ret == 0 && err <= 1e-12;
    HLanc::warn("%-28s ret=%d max rel err %.2e %s\n", what, ret, err, ok ? "ok" : "FAILED");
    for (int i = 0; i < nev; ++i) HLanc::warn("%23.15lE\n", values[i]);
    return ok ? 0 : 1;
}

int main(int argc, char *argv[])
{
    if (argc < 2) {
        HLanc::warn("Usage: %s <golden_dir> [scratch_dir]\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const std::string scratch = argc > 2 ? argv[2] : ".";
    // top-4 ascending: Journals by dense eigh; simple.pp closed form 1/(4 sin^2((2k-1) pi/66))
    const double journals[4] = {19795.37857462678, 23432.29299233234, 25008.48027847430, 68542.71718825979};
    double simple[4];
    for (int k = 4; k >= 1; --k) {
        const double s = std::sin((2 * k - 1) * M_PI / 66.0);
        simple[4 - k] = 1.0 / (4.0 * s * s);
    }
    int bad = 0;
    const int nev = 4, ncv = 12;
    std::vector<double> values(ncv), vectors;

    HLanc::dcsr_matrix_t mm("mm", (dir + "/Journals.mtx").c_str());
    HLanc::dcsrmv_operator_gpu op_mm(mm);
    vectors.assign((size_t)mm.get_num_rows() * ncv, 0.0);
    int ret = HLanc::ds_solver_gpu(mm.get_num_rows(), nev, ncv, values.data(), vectors.data(), mm.get_num_rows(), op_mm);
    bad += check("Journals.mtx (mm)", ret, values.data(), journals, nev);

    HLanc::dcsr_matrix_t pp("pp", (dir + "/simple.pp").c_str());
    HLanc::dcsrmv_operator_gpu op_pp(pp);
    vectors.assign((size_t)16 * ncv, 0.0);
    ret = HLanc::ds_solver_gpu(16, nev, ncv, values.data(), vectors.data(), 16, op_pp);
    bad += check("simple.pp (pp)", ret, values.data(), simple, nev);

    const std::string bin = scratch + "/simple.bin";
    pp.save_bin(bin.c_str());
    HLanc::dcsr_matrix_t bm("bin", bin.c_str());
    HLanc::dcsrmv_operator_gpu op_bin(bm);
    ret = HLanc::ds_solver_gpu(16, nev, ncv, values.data(), vectors.data(), 16, op_bin);
    bad += check("simple.bin (bin round trip)", ret, values.data(), simple, nev);

    ret = HLanc::ds_solver_gpu(16, nev, ncv, values.data(), vectors.data(), 16, op_fun_simple);
    bad += check("op_fun_simple (function)", ret, values.data(), simple, nev);

    hl_options o;
    hl_options_default(&o);
    o.tol = 1e-10;
    hl_info info;
    ret = HLanc::ds_solver_gpu(mm.get_num_rows(), nev, ncv, values.data(), vectors.data(), mm.get_num_rows(), op_mm,
                               0, o, &info);
    bad += check("Journals.mtx, tol 1e-10", ret, values.data(), journals, nev);
    HLanc::warn("restarts %d, SpMV %lld, CGS passes %lld, %.3f ms\n", info.restarts, info.spmv, info.cgs_passes,
                info.seconds * 1e3);
    return bad;
}
