// tridiag.cpp -- host FP64 dense kernels of the restart (ncv <= 128).
//
// The reference calls LAPACK for these (dsteqr HLanc.hpp:1691,1805; dlartg
// 1514,1536; dgeqr2 1837; dorm2r 940).  They are O(ncv^3) on tiny matrices, so
// they run on the host in FP64, written here from the textbook methods:
//   * tridiag_eig: implicit QL with Wilkinson shifts on a symmetric tridiagonal
//     (Bowdler, Martin, Reinsch & Wilkinson, "The QR and QL algorithms for
//     symmetric matrices", Numer. Math. 11, 1968), accumulating rotations.
//   * dlartg: Givens rotation with the sign convention of the MKL dlartg the
//     reference links.
//   * householder_q: Householder QR with LAPACK's reflector convention, then
//     Q = H(0) ... H(k-1) formed explicitly so that V Q runs as one GEMM.
#include <cmath>
#include <vector>

#include "hl_internal.hpp"

namespace hl {

namespace {
// sqrt(a^2 + b^2): the squares cannot overflow or underflow while the larger
// magnitude lies in [1e-150, 1e150]; outside that range, std::hypot.
inline double norm2(double a, double b)
{
    const double m = std::fmax(std::fabs(a), std::fabs(b));
    return (m > 1e-150 && m < 1e150) ? std::sqrt(a * a + b * b) : std::hypot(a, b);
}

} // namespace

void dlartg(double f, double g, double &c, double &s, double &r)
{
    if (g == 0.0) {
        c = 1.0;
        s = 0.0;
        r = f;
        return;
    }
    if (f == 0.0) {
        c = 0.0;
        s = 1.0;
        r = g;
        return;
    }
    const double h = norm2(f, g);
    c = f / h;
    s = g / h;
    r = h;
    if (std::fabs(f) > std::fabs(g) && c < 0.0) {
        c = -c;
        s = -s;
        r = -r;
    }
}

namespace {
// Implicit QL.  FULL: z is n x n (ldz), the eigenvectors.  !FULL: z holds only
// the last row of the eigenvector matrix (z[i] = Z(n-1, i)); the rotations and
// the eigenvalues are the same, so both forms agree bit for bit there.
template <bool FULL>
int tridiag_ql(int n, double *d, double *e, double *z, int ldz)
{
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(d[i]) || (i + 1 < n && !std::isfinite(e[i]))) return 1;
        if (FULL)
            for (int k = 0; k < n; ++k) z[(size_t)i * ldz + k] = (i == k) ? 1.0 : 0.0;
        else
            z[i] = (i == n - 1) ? 1.0 : 0.0;
    }
    if (n == 0) return 0;
    std::vector<double> off(n, 0.0); // off[i] couples rows i and i+1
    for (int i = 0; i + 1 < n; ++i) off[i] = e[i];
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(off[m]) <= kEps * dd) break;
            }
            if (m == l) break;
            if (++iter > 100) return 1;
            // shift: eigenvalue of the leading 2x2 block nearer to d[l]
            double g = (d[l + 1] - d[l]) / (2.0 * off[l]);
            double r = norm2(g, 1.0);
            g = d[m] - d[l] + off[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool underflow = false;
            for (; i >= l; --i) {
                const double f = s * off[i];
                const double b = c * off[i];
                r = norm2(f, g);
                off[i + 1] = r;
                if (r == 0.0) { // split: recover and restart the sweep
                    d[i + 1] -= p;
                    off[m] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (FULL) {
                    double *zi = z + (size_t)i * ldz, *zj = z + (size_t)(i + 1) * ldz;
                    for (int k = 0; k < n; ++k) {
                        const double t = zj[k];
                        zj[k] = s * zi[k] + c * t;
                        zi[k] = c * zi[k] - s * t;
                    }
                } else {
                    const double t = z[i + 1];
                    z[i + 1] = s * z[i] + c * t;
                    z[i] = c * z[i] - s * t;
                }
            }
            if (underflow) continue;
            d[l] -= p;
            off[l] = g;
            off[m] = 0.0;
        }
    }
    // ascending order, eigenvectors alongside (stable selection)
    for (int i = 0; i < n - 1; ++i) {
        int b = i;
        for (int k = i + 1; k < n; ++k)
            if (d[k] < d[b]) b = k;
        if (b != i) {
            std::swap(d[i], d[b]);
            if (FULL)
                for (int k = 0; k < n; ++k) std::swap(z[(size_t)i * ldz + k], z[(size_t)b * ldz + k]);
            else
                std::swap(z[i], z[b]);
        }
    }
    for (int i = 0; i + 1 < n; ++i) e[i] = 0.0;
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(d[i])) return 1;
    return 0;
}
} // namespace

int tridiag_eig(int n, double *d, double *e, double *z, int ldz) { return tridiag_ql<true>(n, d, e, z, ldz); }

int tridiag_eig_last(int n, double *d, double *e, double *zlast) { return tridiag_ql<false>(n, d, e, zlast, 1); }

void householder_q(int m, int k, double *a, int lda, double *tau, double *q, int ldq)
{
    // QR: reflector i has v = (1, a[i+1:m, i]), H(i) = I - tau_i v v^T
    std::vector<double> w(static_cast<size_t>(k > m ? k : m) + 1);
    for (int i = 0; i < k && i < m; ++i) {
        double *ai = a + (size_t)i * lda;
        double xs = 0.0;
        for (int r = i + 1; r < m; ++r) xs += ai[r] * ai[r];
        if (xs == 0.0) {
            tau[i] = 0.0;
        } else {
            const double alpha = ai[i];
            const double beta = -std::copysign(std::hypot(alpha, std::sqrt(xs)), alpha);
            tau[i] = (beta - alpha) / beta;
            const double sc = 1.0 / (alpha - beta);
            for (int r = i + 1; r < m; ++r) ai[r] *= sc;
            ai[i] = beta;
        }
        if (tau[i] != 0.0) { // A(i:, i+1:) := H(i) A(i:, i+1:)
            for (int c = i + 1; c < k; ++c) {
                double *ac = a + (size_t)c * lda;
                double sdot = ac[i];
                for (int r = i + 1; r < m; ++r) sdot += ac[r] * ai[r];
                const double f = tau[i] * sdot;
                ac[i] -= f;
                for (int r = i + 1; r < m; ++r) ac[r] -= f * ai[r];
            }
        }
    }
    // Q = I H(0) H(1) ... H(k-1), applied from the right in forward order
    for (int c = 0; c < m; ++c)
        for (int r = 0; r < m; ++r) q[(size_t)c * ldq + r] = (r == c) ? 1.0 : 0.0;
    for (int i = 0; i < k && i < m; ++i) {
        if (tau[i] == 0.0) continue;
        const double *ai = a + (size_t)i * lda;
        for (int r = 0; r < m; ++r) { // w = Q(:, i:) v
            double sdot = q[(size_t)i * ldq + r];
            for (int t = i + 1; t < m; ++t) sdot += q[(size_t)t * ldq + r] * ai[t];
            w[r] = sdot;
        }
        for (int r = 0; r < m; ++r) q[(size_t)i * ldq + r] -= tau[i] * w[r];
        for (int t = i + 1; t < m; ++t) {
            const double f = tau[i] * ai[t];
            double *qt = q + (size_t)t * ldq;
            for (int r = 0; r < m; ++r) qt[r] -= w[r] * f;
        }
    }
}

}  // namespace hl
