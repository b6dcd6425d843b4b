// tridiag.cpp -- host FP64 dense kernels of the restart.
//
// The reference calls LAPACK for these (dsteqr HLanc.hpp:1691,1805; dlartg
// 1514,1536; dgeqr2 1837; dorm2r 940).  They are O(ncv^3) on small matrices, so
// they run on the host in FP64, written here from the textbook methods:
//   * tridiag_eig: the symmetric QR algorithm on a tridiagonal, implicit steps
//     with the Wilkinson shift (Golub & Van Loan, "Matrix Computations", 4th
//     ed., Algorithms 8.3.2 and 8.3.3), accumulating the rotations.
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
// Symmetric tridiagonal QR (Golub & Van Loan, Alg. 8.3.3).  d[0..n) is the
// diagonal, e[0..n-1) the off-diagonal (e[i] couples rows i and i+1).  Each
// pass deflates negligible off-diagonals (|e[i]| <= eps (|d[i]| + |d[i+1]|))
// from the bottom, then runs one implicit QR step with the Wilkinson shift
// (Alg. 8.3.2) on the lowest unreduced block [l, m]: a rotation R_k of rows
// and columns (k, k+1), R_k = [c -s; s c], set by (x, z) = (d[l] - mu, e[l])
// for k = l and by (e[k-1], bulge) after that, so that T := R_k^T T R_k
// removes the bulge at (k+1, k-1) and pushes it to (k+2, k).  The rotations
// accumulate into the eigenvector matrix, Z := Z R_k.
//   FULL: z is n x n (ldz), the eigenvectors.  !FULL: z holds only the last
// row of the eigenvector matrix (z[i] = Z(n-1, i)); every row of Z is updated
// by the same rotations on its own, so both forms agree bit for bit there.
template <bool FULL>
int tridiag_qr(int n, double *d, double *e, double *z, int ldz)
{
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(d[i]) || (i + 1 < n && !std::isfinite(e[i]))) return 1;
        if (FULL)
            for (int k = 0; k < n; ++k) z[(size_t)i * ldz + k] = (i == k) ? 1.0 : 0.0;
        else
            z[i] = (i == n - 1) ? 1.0 : 0.0;
    }
    if (n <= 1) return 0;
    std::vector<double> off(e, e + (n - 1));
    auto small = [&](int i) { return std::fabs(off[i]) <= kEps * (std::fabs(d[i]) + std::fabs(d[i + 1])); };
    auto rotate = [&](int k, double c, double s) { // Z := Z R_k on columns k, k+1
        if (FULL) {
            double *za = z + (size_t)k * ldz, *zb = za + ldz;
            for (int r = 0; r < n; ++r) {
                const double a = za[r], b = zb[r];
                za[r] = c * a + s * b;
                zb[r] = c * b - s * a;
            }
        } else {
            const double a = z[k], b = z[k + 1];
            z[k] = c * a + s * b;
            z[k + 1] = c * b - s * a;
        }
    };
    const long long max_steps = 30LL * n; // the sweep budget LAPACK's dsteqr allows
    long long steps = 0;
    int m = n - 1;
    while (m > 0) {
        if (small(m - 1)) { // d[m] has converged
            off[m - 1] = 0.0;
            --m;
            continue;
        }
        int l = m - 1; // [l, m]: the lowest unreduced block
        while (l > 0 && !small(l - 1)) --l;
        if (l > 0) off[l - 1] = 0.0;
        if (++steps > max_steps) return 1;
        // Wilkinson shift: the eigenvalue of the trailing 2 x 2 block nearer d[m]
        const double delta = 0.5 * (d[m - 1] - d[m]);
        const double b = off[m - 1];
        const double den = delta + std::copysign(norm2(delta, b), delta);
        const double mu = d[m] - b * (b / den);
        double x = d[l] - mu, zz = off[l], bulge = 0.0;
        for (int k = l; k < m; ++k) {
            const double r = norm2(x, zz);
            const double c = r > 0.0 ? x / r : 1.0;
            const double s = r > 0.0 ? zz / r : 0.0;
            if (k > l) off[k - 1] = r; // R^T (e[k-1], bulge) = (r, 0)
            const double p = d[k], q = off[k], w = d[k + 1];
            const double cs = c * s, cc = c * c, ss = s * s;
            d[k] = cc * p + 2.0 * cs * q + ss * w;
            d[k + 1] = ss * p - 2.0 * cs * q + cc * w;
            off[k] = cs * (w - p) + (cc - ss) * q;
            if (k + 1 < m) { // row k+2 meets the rotated column k: the new bulge
                bulge = s * off[k + 1];
                off[k + 1] *= c;
            }
            rotate(k, c, s);
            x = off[k];
            zz = bulge;
        }
    }
    // ascending order, eigenvectors alongside (selection, stable for ties)
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

int tridiag_eig(int n, double *d, double *e, double *z, int ldz) { return tridiag_qr<true>(n, d, e, z, ldz); }

int tridiag_eig_last(int n, double *d, double *e, double *zlast) { return tridiag_qr<false>(n, d, e, zlast, 1); }

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
