// hlanc_solver_gpu.cpp -- command-line solver with the reference's interface
// (demo/HLanc_solver_gpu.cu:23-72): reads a matrix, solves for nev eigenpairs
// with ncv basis vectors, times the solve alone with gettimeofday, and prints
// nev, n, the values and the n x nev vectors in %23.15lE.
//   usage: hlanc_solver_gpu <bin|mm|pp> <matrix_file> <nev> <ncv> [dev_csrmv=0] [dev_solver=0] [tol]
#include <cstdio>
#include <cstdlib>
#include <sys/time.h>
#include <vector>

#include "HLanc.hpp"

template <typename op_t>
struct dcsrmv_counter_t { // counts operator calls like dcsrmv_wrapper_t (demo/HLanc_solver_gpu.cu:6-15)
    explicit dcsrmv_counter_t(op_t const &op) : op(op), seq(0) {}
    void operator()(double const *x, double *y) const
    {
        ++seq;
        op(x, y);
    }
    op_t const &op;
    mutable long seq;
};

int main(int argc, char *argv[])
{
    if (argc < 5 || argc > 8) {
        HLanc::warn("Usage: %s <bin|mm|pp> <matrix_file> <nev> <ncv> [dev_csrmv=0] [dev_solver=0] [tol]\n", argv[0]);
        return 1;
    }
    const char *fmt = argv[1], *file = argv[2];
    const int nev = std::atoi(argv[3]), ncv = std::atoi(argv[4]);
    const int dev_csrmv = argc > 5 ? std::atoi(argv[5]) : 0;
    const int dev_solver = argc > 6 ? std::atoi(argv[6]) : 0;
    HLanc::dcsr_matrix_t m(fmt, file);
    HLanc::warn("%10s: %6d\n%10s: %6d\n%10s: %6d\n", "rows", m.get_num_rows(), "columns", m.get_num_columns(), "nnz",
                m.get_nnz());
    HLanc::dcsrmv_operator_gpu op(m, dev_csrmv);
    const int n = m.get_num_rows();
    std::vector<double> values(ncv), vectors((size_t)n * ncv);
    hl_options o;
    hl_options_default(&o);
    if (argc > 7) o.tol = std::atof(argv[7]);
    hl_info info;
    struct timeval tb, te;
    gettimeofday(&tb, nullptr);
    const int ret = HLanc::ds_solver_gpu(n, nev, ncv, values.data(), vectors.data(), n, op, dev_solver, o, &info);
    gettimeofday(&te, nullptr);
    HLanc::warn("%s\n", ret == 0 ? "Success." : "Failes.");
    HLanc::warn("Time usage for the solver: %.6lf\n", te.tv_sec - tb.tv_sec + 1e-6 * (te.tv_usec - tb.tv_usec));
    HLanc::warn("restarts %d, SpMV %lld, CGS passes %lld\n", info.restarts, info.spmv, info.cgs_passes);
    std::printf("%d\n%d\n", nev, n);
    for (int i = 0; i < nev; ++i) std::printf("%23.15lE\n", values[i]);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < nev; ++j) std::printf("%23.15lE%c", vectors[i + (size_t)j * n], j + 1 == nev ? '\n' : ' ');
    return ret;
}
