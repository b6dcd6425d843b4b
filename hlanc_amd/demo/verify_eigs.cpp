// verify_eigs.cpp -- residual check of a solver output file, as the reference's
// demo/verify_eigs.cu:52-98: for each pair, max|Av - lv| / max(|Av|_inf, |lv|_inf)
// with A v computed by dcsrmv_operator_gpu.  Exit status 1 if any pair exceeds
// the optional threshold.
//   usage: verify_eigs <bin|mm|pp> <matrix_file> <eigs_file> [dev=0] [threshold]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "HLanc.hpp"

int main(int argc, char *argv[])
{
    if (argc < 4 || argc > 6) {
        HLanc::warn("Usage: %s <bin|mm|pp> <matrix_file> <eigs_file> [dev=0] [threshold]\n", argv[0]);
        return 1;
    }
    HLanc::dcsr_matrix_t m(argv[1], argv[2]);
    FILE *fp = std::fopen(argv[3], "r");
    if (!fp) HLanc::error_exit("Cannot open file \"%s\"\n", argv[3]);
    int nev, n;
    if (std::fscanf(fp, "%d", &nev) != 1 || std::fscanf(fp, "%d", &n) != 1) HLanc::error_exit("Error reading nev/n\n");
    std::vector<double> value(nev), vector((size_t)n * nev);
    for (int i = 0; i < nev; ++i)
        if (std::fscanf(fp, "%lE", &value[i]) != 1) HLanc::error_exit("Error reading eig value %d\n", i);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < nev; ++c)
            if (std::fscanf(fp, "%lE", &vector[(size_t)c * n + r]) != 1) HLanc::error_exit("Error reading <%d,%d>\n", r, c);
    std::fclose(fp);
    const int dev = argc > 4 ? std::atoi(argv[4]) : 0;
    const double thr = argc > 5 ? std::atof(argv[5]) : -1.0;
    HLanc::dcsrmv_operator_gpu op(m, dev);
    std::vector<double> x(n), y(n);
    int bad = 0;
    for (int c = 0; c < nev; ++c) {
        double const *v = vector.data() + (size_t)c * n;
        op(v, x.data());
        double maxdiff = 0.0, maxabs = 0.0;
        for (int r = 0; r < n; ++r) {
            y[r] = v[r] * value[c];
            maxdiff = std::fmax(maxdiff, std::fabs(x[r] - y[r]));
            maxabs = std::fmax(maxabs, std::fmax(std::fabs(x[r]), std::fabs(y[r])));
        }
        if (maxabs > 0.0) maxdiff /= maxabs;
        HLanc::warn("max r-diff of vector[%3d]: %.6lE\n", c, maxdiff);
        if (thr > 0.0 && !(maxdiff <= thr)) bad = 1;
    }
    return bad;
}
