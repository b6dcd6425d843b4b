// hlanc_batch_solver.cpp -- many eigensolves in one batched call through the
// drop-in header's extension HLanc::ds_solver_gpu_batch, checked against one
// ds_solver_gpu call per matrix: every value and vector must be bit-identical.
// Prints, per matrix, nev, n and the ascending eigenvalues in %23.15lE.
//   usage: hlanc_batch_solver <nev> <ncv> <bin|mm|pp> <matrix_file> [<fmt> <file> ...]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sys/time.h>
#include <vector>

#include "HLanc.hpp"

static double seconds_since(const timeval &t0)
{
    timeval t;
    gettimeofday(&t, nullptr);
    return (t.tv_sec - t0.tv_sec) + 1e-6 * (t.tv_usec - t0.tv_usec);
}

int main(int argc, char *argv[])
{
    if (argc < 5 || (argc - 3) % 2 != 0) {
        HLanc::warn("Usage: %s <nev> <ncv> <bin|mm|pp> <matrix_file> [<fmt> <file> ...]\n", argv[0]);
        return 1;
    }
    const int nev = std::atoi(argv[1]), ncv = std::atoi(argv[2]);
    const int count = (argc - 3) / 2;
    std::vector<std::unique_ptr<HLanc::dcsr_matrix_t>> mats;
    std::vector<std::unique_ptr<HLanc::dcsrmv_operator_gpu>> ops;
    for (int i = 0; i < count; ++i) {
        mats.emplace_back(new HLanc::dcsr_matrix_t(argv[3 + 2 * i], argv[4 + 2 * i]));
        ops.emplace_back(new HLanc::dcsrmv_operator_gpu(*mats.back(), 0));
    }
    std::vector<std::vector<double>> values(count, std::vector<double>(ncv)), vectors(count);
    std::vector<HLanc::batch_solve> batch(count);
    for (int i = 0; i < count; ++i) {
        const int n = mats[i]->get_num_rows();
        vectors[i].assign((size_t)n * ncv, 0.0);
        batch[i].op = ops[i].get();
        batch[i].nev = nev;
        batch[i].ncv = ncv;
        batch[i].values = values[i].data();
        batch[i].vectors = vectors[i].data();
        batch[i].ldv = n;
    }
    timeval t0;
    gettimeofday(&t0, nullptr);
    HLanc::ds_solver_gpu_batch(batch);
    const double t_batch = seconds_since(t0);

    int mismatches = 0;
    gettimeofday(&t0, nullptr);
    for (int i = 0; i < count; ++i) {
        const int n = mats[i]->get_num_rows();
        std::vector<double> v1(ncv), x1((size_t)n * ncv);
        const int ret = HLanc::ds_solver_gpu(n, nev, ncv, v1.data(), x1.data(), n, *ops[i]);
        if (ret != batch[i].status) ++mismatches;
        if (ret == 0 && (std::memcmp(v1.data(), values[i].data(), sizeof(double) * nev) != 0 ||
                         std::memcmp(x1.data(), vectors[i].data(), sizeof(double) * x1.size()) != 0))
            ++mismatches;
    }
    const double t_single = seconds_since(t0);
    HLanc::warn("%d solves: batched %.6lf s, one at a time %.6lf s, %d mismatches\n", count, t_batch, t_single,
                mismatches);
    for (int i = 0; i < count; ++i) {
        std::printf("%d\n%d\n", nev, mats[i]->get_num_rows());
        for (int k = 0; k < nev; ++k) std::printf("%23.15lE\n", values[i][k]);
    }
    return mismatches == 0 ? 0 : 2;
}
