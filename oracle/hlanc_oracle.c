/*
 * hlanc_oracle.c -- CPU restatement of HLanc's implicitly restarted Lanczos path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the CPU baseline
 * timed by bench.py.  Nothing under hlanc_amd/ or include/ links, loads or calls it.
 *
 * It restates, in plain C, the reference's host path ds_solver_mkl +
 * dcsrmv_operator_naive/mkl + mat_v_mkl_t (HLanc.hpp:399-469, 900-948, 1353-1861),
 * step for step.  LAPACK/BLAS calls of the reference are replaced by code written
 * here from the textbook methods named at each function.
 *
 * What pins each part (tests/test_oracle.py, fixtures in tests/golden/):
 *   hor_csr_spmv       MKL mkl_dcsrmv ("GLNC"/"SLNC") outputs on data/simple.pp,
 *                      data/Journals.mtx and random CSR; scipy.sparse products.
 *   hor_dlartg         MKL dlartg outputs (sign convention included).
 *   hor_tridiag_eig    MKL dsteqr and scipy dstev eigenvalues; T z = z diag(w),
 *                      z^T z = I on random and degenerate tridiagonals.
 *   hor_dgeqr2 /       numpy.linalg.qr (same dgeqrf reflector convention) and MKL
 *   hor_dorm2r_rn      dgeqr2 + dorm2r outputs.
 *   hor_start_vector   fixed golden values (own generator; no reference counterpart).
 *   hor_ds_solver      closed-form spectra (data/simple.pp, 2-D Laplacian), dense
 *                      eigh (data/Journals.mtx, random SPD), Ritz residuals and
 *                      orthogonality, the reference's failure returns.
 *   Trajectory details pinned only by running the reference (exact restart count,
 *   shift order under exact ties, eigenvector signs) are "parity unpinned"; the
 *   reference cannot be built in this image (it includes cublas_v2.h,
 *   cusparse_v2.h and mkl.h unconditionally, HLanc.hpp:9-12).
 */
#include "hlanc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define HOR_MIN(a, b) ((a) < (b) ? (a) : (b))
#define HOR_MAX(a, b) ((a) > (b) ? (a) : (b))

static const double kEps = DBL_EPSILON / 2.0; /* dlamch("E") = 2^-53 */
static const double kSafmin = DBL_MIN;        /* dlamch("S") */
static const double kDgks = 0.717;            /* HLanc.hpp:1444,1451 */

/* ------------------------------------------------------------------ SpMV */

/* HLanc.hpp:403-431: symmetric-lower storage means y = (L + L^T - D) x,
 * accumulated row by row; general storage is a plain row dot product. */
void hor_csr_spmv(const hor_csr *A, const double *x, double *y)
{
    const int n = A->nrows;
    if (A->sym) {
        for (int i = 0; i < n; ++i) y[i] = 0.0;
        for (int r = 0; r < n; ++r) {
            for (int p = A->row_ptr[r]; p < A->row_ptr[r + 1]; ++p) {
                const int c = A->col[p];
                const double v = A->val[p];
                y[r] += x[c] * v;
                if (r != c) y[c] += x[r] * v;
            }
        }
    } else {
#pragma omp parallel for schedule(static) if (n > 32768)
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int p = A->row_ptr[r]; p < A->row_ptr[r + 1]; ++p) s += x[A->col[p]] * A->val[p];
            y[r] = s;
        }
    }
}

void hor_csr_op(void *ctx, const double *x, double *y) { hor_csr_spmv((const hor_csr *)ctx, x, y); }

/* ------------------------------------------------------------ BLAS-1/2 */

/* Deterministic 2-norm: fixed 4096-element chunks summed in order, so the
 * value does not depend on the OpenMP thread count. */
static double hor_nrm2(int n, const double *x)
{
    const int chunk = 4096;
    const int nch = (n + chunk - 1) / chunk;
    double *part = (double *)malloc(sizeof(double) * (size_t)HOR_MAX(nch, 1));
#pragma omp parallel for schedule(static) if (n > 65536)
    for (int b = 0; b < nch; ++b) {
        const int lo = b * chunk, hi = HOR_MIN(n, lo + chunk);
        double s = 0.0;
        for (int i = lo; i < hi; ++i) s += x[i] * x[i];
        part[b] = s;
    }
    double s = 0.0;
    for (int b = 0; b < nch; ++b) s += part[b];
    free(part);
    return sqrt(s);
}

/* h = V(:, 0:m)^T w (cblas_dgemv Trans, HLanc.hpp:911). */
static void hor_gemv_t(int n, int m, const double *V, size_t ldv, const double *w, double *h)
{
#pragma omp parallel for schedule(static) if ((long long)n * m > 65536)
    for (int i = 0; i < m; ++i) {
        const double *vi = V + (size_t)i * ldv;
        double s = 0.0;
        for (int r = 0; r < n; ++r) s += vi[r] * w[r];
        h[i] = s;
    }
}

/* y = alpha V(:, 0:m) x + beta y (cblas_dgemv NoTrans, HLanc.hpp:912,921,924). */
static void hor_gemv_n(int n, int m, double alpha, const double *V, size_t ldv, const double *x,
                       double beta, double *y)
{
#pragma omp parallel for schedule(static) if ((long long)n * m > 65536)
    for (int r = 0; r < n; ++r) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += V[(size_t)i * ldv + r] * x[i];
        y[r] = (beta == 0.0 ? 0.0 : beta * y[r]) + alpha * s;
    }
}

/* ------------------------------------------------------------- dlartg */

/* Givens rotation [c s; -s c] [f; g] = [r; 0] with the sign convention of the
 * MKL dlartg the reference links (LAPACK 3.2-3.9 behaviour): g = 0 gives
 * (1, 0, f); f = 0 gives (0, 1, g); otherwise r = hypot(f, g), and when
 * |f| > |g| with c < 0 all three change sign.  Pinned by MKL fixtures. */
void hor_dlartg(double f, double g, double *c, double *s, double *r)
{
    if (g == 0.0) {
        *c = 1.0;
        *s = 0.0;
        *r = f;
        return;
    }
    if (f == 0.0) {
        *c = 0.0;
        *s = 1.0;
        *r = g;
        return;
    }
    const double rr = hypot(f, g);
    double cc = f / rr, ss = g / rr, r0 = rr;
    if (fabs(f) > fabs(g) && cc < 0.0) {
        cc = -cc;
        ss = -ss;
        r0 = -r0;
    }
    *c = cc;
    *s = ss;
    *r = r0;
}

/* ----------------------------------------------------- tridiagonal eigen */

/* Cyclic Jacobi on the dense form of T (Golub & Van Loan, Matrix Computations,
 * 4th ed., Alg. 8.5.3 with the sym.schur2 rotation of Alg. 8.5.1).  Chosen to be
 * independent of the implicit-shift QL the product uses; eigenvalues are
 * returned ascending with their eigenvectors, as dsteqr("I") does. */
int hor_tridiag_eig(int n, double *d, double *e, double *z, int ldz)
{
    for (int i = 0; i < n; ++i)
        if (!isfinite(d[i]) || (i + 1 < n && !isfinite(e[i]))) return 1;
    double *a = (double *)calloc((size_t)n * n, sizeof(double));
    double *v = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; ++i) {
        a[i * n + i] = d[i];
        v[i * n + i] = 1.0;
        if (i + 1 < n) a[i * n + i + 1] = a[(i + 1) * n + i] = e[i];
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < n; ++p) {
            diag += a[p * n + p] * a[p * n + p];
            for (int q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
        }
        if (off == 0.0 || off <= 1e-36 * diag) break;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[p * n + q];
                if (apq == 0.0) continue;
                const double app = a[p * n + p], aqq = a[q * n + q];
                if (fabs(apq) <= 1e-18 * sqrt(fabs(app * aqq)) && sweep > 3) {
                    a[p * n + q] = a[q * n + p] = 0.0;
                    continue;
                }
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) { /* A := A J (columns p, q) */
                    const double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) { /* A := J^T A (rows p, q) */
                    const double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                a[p * n + q] = a[q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) { /* V := V J; v is row-major V */
                    const double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - s * vkq;
                    v[k * n + q] = s * vkp + c * vkq;
                }
            }
        }
    }
    /* selection sort ascending, stable on ties */
    int *idx = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    for (int i = 0; i < n; ++i) {
        int b = i;
        for (int k = i + 1; k < n; ++k)
            if (a[idx[k] * n + idx[k]] < a[idx[b] * n + idx[b]]) b = k;
        const int t = idx[i];
        idx[i] = idx[b];
        idx[b] = t;
    }
    for (int i = 0; i < n; ++i) {
        d[i] = a[idx[i] * n + idx[i]];
        for (int k = 0; k < n; ++k) z[(size_t)i * ldz + k] = v[k * n + idx[i]];
    }
    free(idx);
    free(a);
    free(v);
    return 0;
}

/* ------------------------------------------------------ Householder QR */

/* LAPACK dlarfg convention: beta = -sign(alpha) * ||(alpha, x)||,
 * tau = (beta - alpha) / beta, x := x / (alpha - beta); tau = 0 when x = 0. */
static void hor_dlarfg(int n, double *alpha, double *x, double *tau)
{
    if (n <= 1) {
        *tau = 0.0;
        return;
    }
    double xs = 0.0;
    for (int i = 0; i < n - 1; ++i) xs += x[i] * x[i];
    const double xnorm = sqrt(xs);
    if (xnorm == 0.0) {
        *tau = 0.0;
        return;
    }
    const double beta = -copysign(hypot(*alpha, xnorm), *alpha);
    *tau = (beta - *alpha) / beta;
    const double scal = 1.0 / (*alpha - beta);
    for (int i = 0; i < n - 1; ++i) x[i] *= scal;
    *alpha = beta;
}

int hor_dgeqr2(int m, int k, double *a, int lda, double *tau)
{
    const int kk = HOR_MIN(m, k);
    double *w = (double *)malloc(sizeof(double) * (size_t)HOR_MAX(k, 1));
    for (int i = 0; i < kk; ++i) {
        double *ai = a + (size_t)i * lda;
        hor_dlarfg(m - i, &ai[i], &ai[i + 1], &tau[i]);
        if (i + 1 < k && tau[i] != 0.0) { /* A(i:, i+1:) := H(i) A(i:, i+1:) */
            const double aii = ai[i];
            ai[i] = 1.0;
            for (int c = i + 1; c < k; ++c) {
                const double *ac = a + (size_t)c * lda;
                double s = 0.0;
                for (int r = i; r < m; ++r) s += ac[r] * ai[r];
                w[c] = s;
            }
            for (int c = i + 1; c < k; ++c) {
                double *ac = a + (size_t)c * lda;
                for (int r = i; r < m; ++r) ac[r] -= tau[i] * ai[r] * w[c];
            }
            ai[i] = aii;
        }
    }
    free(w);
    return 0;
}

void hor_dorm2r_rn(int nr, int nc, int k, const double *a, int lda, const double *tau, double *c,
                   int ldc)
{
    double *w = (double *)malloc(sizeof(double) * (size_t)HOR_MAX(nr, 1));
    double *v = (double *)malloc(sizeof(double) * (size_t)HOR_MAX(nc, 1));
    for (int i = 0; i < k; ++i) { /* forward order for Right + NoTrans */
        if (tau[i] == 0.0) continue;
        const int ni = nc - i;
        v[0] = 1.0;
        for (int t = 1; t < ni; ++t) v[t] = a[(size_t)i * lda + i + t];
        /* w = C(:, i:) v ; C(:, i:) -= tau w v^T */
#pragma omp parallel for schedule(static) if ((long long)nr * ni > 65536)
        for (int r = 0; r < nr; ++r) {
            double s = 0.0;
            for (int t = 0; t < ni; ++t) s += c[(size_t)(i + t) * ldc + r] * v[t];
            w[r] = s;
        }
        for (int t = 0; t < ni; ++t) {
            double *ct = c + (size_t)(i + t) * ldc;
            const double f = tau[i] * v[t];
#pragma omp parallel for schedule(static) if (nr > 65536)
            for (int r = 0; r < nr; ++r) ct[r] -= w[r] * f;
        }
    }
    free(w);
    free(v);
}

/* ----------------------------------------------------------- start vector */

uint64_t hor_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void hor_start_vector(int n, uint64_t seed, double *x)
{
    const uint64_t base = hor_mix64(seed ^ 0x6A09E667F3BCC909ULL);
    for (int i = 0; i < n; ++i) {
        const uint64_t z = hor_mix64(base + 0x9E3779B97F4A7C15ULL * (uint64_t)(i + 1));
        x[i] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
    }
}

/* ------------------------------------------------------------------ solver */

typedef struct {
    int n, ncv;
    double *V;
    size_t ldv;
    hor_op_fn op;
    void *ctx;
    double *ipj, *irj;
    long long spmv, cgs;
} hor_state;

/* One CGS pass, mat_v_mkl_t::op (HLanc.hpp:910-913):
 * irj = V(:, 0:m)^T ipj ; resid -= V(:, 0:m) irj. */
static void hor_cgs(hor_state *st, int m, double *resid)
{
    hor_gemv_t(st->n, m, st->V, st->ldv, st->ipj, st->irj);
    hor_gemv_n(st->n, m, -1.0, st->V, st->ldv, st->irj, 1.0, resid);
    st->cgs++;
}

/* dsaitr (HLanc.hpp:1403-1482): extend the Lanczos factorization from k to k+np
 * steps with full CGS reorthogonalization and at most two DGKS corrections. */
static void hor_dsaitr(hor_state *st, int k, int np, double *resid, double *rnorm, double *h, int ldh)
{
    const int n = st->n;
    double *V = st->V;
    for (int j = k; j < k + np; ++j) {
        double *vj = V + (size_t)j * st->ldv;
        if (*rnorm > kSafmin) { /* STEP 2, HLanc.hpp:1412-1421 */
            const double t = 1.0 / *rnorm;
            for (int i = 0; i < n; ++i) vj[i] = resid[i] * t;
        } else {
            memcpy(vj, resid, sizeof(double) * (size_t)n);
        }
        st->op(st->ctx, vj, resid); /* STEP 3 */
        st->spmv++;
        memcpy(st->ipj, resid, sizeof(double) * (size_t)n); /* STEP 4 */
        const double wnorm = hor_nrm2(n, resid);
        hor_cgs(st, j + 1, resid);
        h[ldh + j] = st->irj[j];
        h[j] = (j == 0) ? 0.0 : *rnorm;
        memcpy(st->ipj, resid, sizeof(double) * (size_t)n);
        *rnorm = hor_nrm2(n, resid);

        int iter = 0; /* STEP 5, DGKS (HLanc.hpp:1443-1465) */
        while (*rnorm < wnorm * kDgks) {
            hor_cgs(st, j + 1, resid);
            h[ldh + j] += st->irj[j];
            memcpy(st->ipj, resid, sizeof(double) * (size_t)n);
            const double rnorm1 = hor_nrm2(n, resid);
            if (rnorm1 > *rnorm * kDgks) {
                *rnorm = rnorm1;
            } else {
                *rnorm = rnorm1;
                ++iter;
                if (iter <= 1) continue;
                for (int i = 0; i < n; ++i) resid[i] = 0.0;
                *rnorm = 0.0;
            }
            break;
        }
        /* HLanc.hpp:1467-1476 flips signs when h[j] < 0; h[j] is a norm, so the
         * branch never runs and is not restated. */
    }
}

/* Ascending sort of (ritz, bound) pairs by ritz, then the first np by bound
 * descending: dsgets for "largest algebraic" (HLanc.hpp:1386-1401). */
typedef struct {
    double d1, d2;
} hor_pair;

static int hor_cmp_d1_asc(const void *a, const void *b)
{
    const double x = ((const hor_pair *)a)->d1, y = ((const hor_pair *)b)->d1;
    return (x > y) - (x < y);
}

static int hor_cmp_d2_desc(const void *a, const void *b)
{
    const double x = ((const hor_pair *)a)->d2, y = ((const hor_pair *)b)->d2;
    return (x < y) - (x > y);
}

static void hor_dsgets(int nev, int np, double *ritz, double *bounds)
{
    const int ncv = nev + np;
    hor_pair *p = (hor_pair *)malloc(sizeof(hor_pair) * (size_t)ncv);
    for (int i = 0; i < ncv; ++i) {
        p[i].d1 = ritz[i];
        p[i].d2 = bounds[i];
    }
    qsort(p, (size_t)ncv, sizeof(hor_pair), hor_cmp_d1_asc);
    qsort(p, (size_t)np, sizeof(hor_pair), hor_cmp_d2_desc);
    for (int i = 0; i < ncv; ++i) {
        ritz[i] = p[i].d1;
        bounds[i] = p[i].d2;
    }
    free(p);
}

/* dsconv (HLanc.hpp:1376-1384). */
static int hor_dsconv(int n, const double *ritz, const double *bounds, double tol)
{
    const double eps23 = pow(kEps, 2.0 / 3.0);
    int nconv = 0;
    for (int i = 0; i < n; ++i) nconv += bounds[i] <= tol * HOR_MAX(eps23, fabs(ritz[i]));
    return nconv;
}

/* One Givens rotation on columns (c0, c0+1) of Q, rows 0..rows-1
 * (HLanc.hpp:1525-1529, 1554-1558). */
static void hor_rot_q(double *q, int ldq, int c0, int rows, double c, double s)
{
    double *qa = q + (size_t)c0 * ldq, *qb = q + (size_t)(c0 + 1) * ldq;
    for (int r = 0; r < rows; ++r) {
        const double a1 = c * qa[r] + s * qb[r];
        qb[r] = -s * qa[r] + c * qb[r];
        qa[r] = a1;
    }
}

/* The shift sweeps of dsapps (HLanc.hpp:1492-1589): np implicit exact-shift QR
 * sweeps on T = (dia, sub) by bulge chasing; q (ncv x ncv, ld ncv) receives Q
 * and dia/sub become Q^T T Q. */
void hor_dsapps_chase(int nev, int np, const double *shift, double *sub, double *dia, double *q)
{
    const int ncv = nev + np, ldq = ncv;
    memset(q, 0, sizeof(double) * (size_t)ncv * ncv);
    for (int i = 0; i < ncv; ++i) q[(size_t)i * ldq + i] = 1.0;
    int itop = 1;
    double c = 0.0, s = 0.0, r = 0.0;
    for (int jj = 1; jj <= np; ++jj) {
        int istart = itop;
        for (;;) {
            int iend = ncv;
            for (int i = istart; i <= ncv - 1; ++i) { /* deflation check */
                const double big = fabs(dia[i - 1]) + fabs(dia[i]);
                if (sub[i] <= kEps * big) {
                    sub[i] = 0.0;
                    iend = i;
                    break;
                }
            }
            if (istart < iend) {
                /* first rotation of the sweep, shift applied (HLanc.hpp:1512-1529) */
                double f = dia[istart - 1] - shift[jj - 1];
                double g = sub[istart];
                hor_dlartg(f, g, &c, &s, &r);
                double a1 = c * dia[istart - 1] + s * sub[istart];
                double a2 = c * sub[istart] + s * dia[istart];
                double a4 = c * dia[istart] - s * sub[istart];
                double a3 = c * sub[istart] - s * dia[istart - 1];
                dia[istart - 1] = c * a1 + s * a2;
                dia[istart] = c * a4 - s * a3;
                sub[istart] = c * a3 + s * a4;
                hor_rot_q(q, ldq, istart - 1, HOR_MIN(istart + jj, ncv), c, s);
                /* chase the bulge down to iend (HLanc.hpp:1531-1559) */
                for (int i = istart + 1; i <= iend - 1; ++i) {
                    f = sub[i - 1];
                    g = s * sub[i];
                    sub[i] = c * sub[i];
                    hor_dlartg(f, g, &c, &s, &r);
                    if (r < 0.0) {
                        r = -r;
                        c = -c;
                        s = -s;
                    }
                    sub[i - 1] = r;
                    a1 = c * dia[i - 1] + s * sub[i];
                    a2 = c * sub[i] + s * dia[i];
                    a3 = c * sub[i] - s * dia[i - 1];
                    a4 = c * dia[i] - s * sub[i];
                    dia[i - 1] = c * a1 + s * a2;
                    dia[i] = c * a4 - s * a3;
                    sub[i] = c * a3 + s * a4;
                    hor_rot_q(q, ldq, i - 1, HOR_MIN(i + jj, ncv), c, s);
                }
            }
            istart = iend + 1;
            if (sub[iend - 1] < 0.0) { /* HLanc.hpp:1563-1568 */
                sub[iend - 1] = -sub[iend - 1];
                for (int i = 0; i < ncv; ++i) q[(size_t)(iend - 1) * ldq + i] *= -1.0;
            }
            if (iend < ncv) continue;
            for (int i = itop; i <= ncv - 1; ++i) {
                if (sub[i] > 0.0) break;
                ++itop;
            }
            break;
        }
    }
    for (int i = itop; i <= ncv - 1; ++i) { /* final deflation sweep */
        const double big = fabs(dia[i - 1]) + fabs(dia[i]);
        if (sub[i] <= kEps * big) sub[i] = 0.0;
    }
}

/* dsapps (HLanc.hpp:1484-1618): the shift sweeps, then V := V Q(:, 0:nev) and
 * the new residual. */
static void hor_dsapps(hor_state *st, int nev, int np, const double *shift, double *h, int ldh,
                       double *resid)
{
    const int n = st->n, ncv = nev + np, ldq = ncv;
    double *q = (double *)calloc((size_t)ncv * ncv, sizeof(double));
    double *sub = h;         /* sub[i] = T(i, i-1), i >= 1 */
    double *dia = h + ldh;   /* dia[i] = T(i, i) */
    hor_dsapps_chase(nev, np, shift, sub, dia, q);

    /* Basis update (HLanc.hpp:1591-1614) written against full columns of Q:
     * entries of Q outside the Givens band are exact zeros, so
     * V(:, ncv-i) = V(:, 0:ncv-i+1) Q(0:ncv-i+1, nev-i) of the reference is the
     * same sum as V Q(:, nev-i). */
    double *V = st->V;
    const size_t ldv = st->ldv;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
    double *vk = NULL;
    if (sub[nev] > 0.0) {
        vk = (double *)malloc(sizeof(double) * (size_t)n);
        hor_gemv_n(n, ncv, 1.0, V, ldv, q + (size_t)nev * ldq, 0.0, vk);
    }
    for (int i = 1; i <= nev; ++i) {
        hor_gemv_n(n, ncv - i + 1, 1.0, V, ldv, q + (size_t)(nev - i) * ldq, 0.0, tmp);
        memcpy(V + (size_t)(ncv - i) * ldv, tmp, sizeof(double) * (size_t)n);
    }
    memmove(V, V + (size_t)np * ldv, sizeof(double) * (size_t)nev * ldv); /* move_columns */
    if (vk) memcpy(V + (size_t)nev * ldv, vk, sizeof(double) * (size_t)n);
    const double qlast = q[(size_t)(nev - 1) * ldq + ncv - 1];
    for (int i = 0; i < n; ++i) resid[i] *= qlast;
    if (vk) {
        const double b = sub[nev];
        for (int i = 0; i < n; ++i) resid[i] += b * vk[i];
    }
    free(vk);
    free(tmp);
    free(q);
}

/* dsesrt (HLanc.hpp:1626-1647): sort values ascending, permuting Q's columns. */
static void hor_dsesrt(int nconv, double *values, int ncv, double *iq, int ldq)
{
    int *idx = (int *)malloc(sizeof(int) * (size_t)nconv);
    for (int i = 0; i < nconv; ++i) idx[i] = i;
    for (int i = 1; i < nconv; ++i) { /* insertion sort on indices */
        const int t = idx[i];
        int k = i - 1;
        while (k >= 0 && values[idx[k]] > values[t]) {
            idx[k + 1] = idx[k];
            --k;
        }
        idx[k + 1] = t;
    }
    double *vv = (double *)malloc(sizeof(double) * (size_t)nconv);
    double *qq = (double *)malloc(sizeof(double) * (size_t)ldq * nconv);
    memcpy(vv, values, sizeof(double) * (size_t)nconv);
    memcpy(qq, iq, sizeof(double) * (size_t)ldq * nconv);
    for (int i = 0; i < nconv; ++i) {
        values[i] = vv[idx[i]];
        memcpy(iq + (size_t)i * ldq, qq + (size_t)idx[i] * ldq, sizeof(double) * (size_t)ncv);
    }
    free(qq);
    free(vv);
    free(idx);
}

int hor_ds_solver(int n, int nev, int ncv, double *values, double *vectors, int ldv, hor_op_fn op,
                  void *ctx, const hor_opts *opts, hor_info *info)
{
    hor_opts o = {0};
    if (opts) o = *opts;
    const double tol = o.tol > 0.0 ? o.tol : kEps;
    const int maxiter = o.maxiter > 0 ? o.maxiter : 200;
    const double eps23 = pow(kEps, 2.0 / 3.0);
#ifdef _OPENMP
    omp_set_num_threads(o.nthreads > 0 ? o.nthreads : 1);
#endif
    hor_info inf = {0};
    int ret = 1;
    if (n <= 0 || nev <= 0 || ncv <= nev || ncv > n || ldv < n) {
        inf.ret = -1;
        if (info) *info = inf;
        return -1;
    }

    hor_state st;
    st.n = n;
    st.ncv = ncv;
    st.V = vectors;
    st.ldv = (size_t)ldv;
    st.op = op;
    st.ctx = ctx;
    st.ipj = (double *)malloc(sizeof(double) * (size_t)n);
    st.irj = (double *)malloc(sizeof(double) * (size_t)HOR_MAX(n, ncv));
    st.spmv = st.cgs = 0;

    int np = ncv - nev, nconv = 0;
    const int nev0 = nev, np0 = np;
    double *resid = (double *)malloc(sizeof(double) * (size_t)n);
    if (o.v0)
        memcpy(resid, o.v0, sizeof(double) * (size_t)n);
    else
        hor_start_vector(n, o.seed, resid); /* dgetv0, HLanc.hpp:1368-1374 */
    double rnorm = hor_nrm2(n, resid);

    /* workl layout of HLanc.hpp:1666-1676, kept as separate arrays here */
    double *h = (double *)calloc((size_t)2 * ncv, sizeof(double));
    const int ldh = ncv;
    double *ritz = (double *)calloc((size_t)ncv, sizeof(double));
    double *bounds = (double *)calloc((size_t)ncv, sizeof(double));
    double *q = (double *)calloc((size_t)ncv * ncv, sizeof(double));
    double *work = (double *)calloc((size_t)ncv, sizeof(double));
    double *ritz_saved = (double *)calloc((size_t)ncv, sizeof(double));   /* irz */
    double *bounds_saved = (double *)calloc((size_t)ncv, sizeof(double)); /* ibd */
    double *e = (double *)calloc((size_t)ncv, sizeof(double));

    hor_dsaitr(&st, 0, nev0, resid, &rnorm, h, ldh);

    int iter;
    for (iter = 0; iter < maxiter; ++iter) {
        hor_dsaitr(&st, nev, np, resid, &rnorm, h, ldh);

        memcpy(ritz, h + ldh, sizeof(double) * (size_t)ncv); /* HLanc.hpp:1687-1698 */
        memcpy(e, h + 1, sizeof(double) * (size_t)(ncv - 1));
        if (hor_tridiag_eig(ncv, ritz, e, q, ncv)) goto done;
        for (int i = 0; i < ncv; ++i) bounds[i] = rnorm * fabs(q[(size_t)i * ncv + ncv - 1]);
        memcpy(ritz_saved, ritz, sizeof(double) * (size_t)ncv);
        memcpy(bounds_saved, bounds, sizeof(double) * (size_t)ncv);

        nev = nev0;
        np = np0;
        hor_dsgets(nev, np, ritz, bounds);
        memcpy(work + np, bounds + np, sizeof(double) * (size_t)nev);
        nconv = hor_dsconv(nev, ritz + np, work + np, tol);

        for (int j = 0; j < np; ++j) { /* HLanc.hpp:1708-1713, np re-read each trip */
            if (bounds[j] == 0.0) {
                --np;
                ++nev;
            }
        }
        if (nconv >= nev0 || np == 0) break;

        const int nevbef = nev; /* HLanc.hpp:1719-1730 */
        nev += HOR_MIN(nconv, np / 2);
        if (nev == 1 && ncv >= 6)
            nev = ncv / 2;
        else if (nev == 1 && ncv > 2)
            nev = 2;
        np = ncv - nev;
        if (nevbef < nev) hor_dsgets(nev, np, ritz, bounds);

        hor_dsapps(&st, nev, np, ritz, h, ldh, resid);
        rnorm = hor_nrm2(n, resid);
        if (o.max_restarts_then_stop > 0 && iter + 1 >= o.max_restarts_then_stop) {
            ++iter;
            inf.restarts = iter;
            ret = 2;
            goto done;
        }
    }
    inf.restarts = iter;
    /* HLanc.hpp:1738-1758 reorders ritz/bounds for display only; nothing
     * downstream reads them, so that block is not restated. */
    h[0] = rnorm;
    if (nconv < nev0) goto done; /* HLanc.hpp:1762 */

    {   /* Ritz vectors, HLanc.hpp:1765-1843 */
        nev = nev0;
        np = np0;
        char *select = (char *)calloc((size_t)ncv, 1);
        double *idxv = (double *)malloc(sizeof(double) * (size_t)ncv);
        for (int j = 0; j < ncv; ++j) idxv[j] = j + 1;
        hor_dsgets(nev, np, ritz_saved, idxv);
        int reord = 0, numcnv = 0;
        for (int j = 0; j < ncv; ++j) {
            const double temp1 = HOR_MAX(eps23, fabs(ritz_saved[ncv - j - 1]));
            const int jj = (int)idxv[ncv - j - 1];
            if (numcnv < nconv && bounds_saved[jj - 1] <= tol * temp1) {
                select[jj - 1] = 1;
                ++numcnv;
                if (jj > nev) reord = 1;
            }
        }
        free(idxv);
        if (numcnv != nconv) {
            free(select);
            goto done;
        }
        double *ihd = (double *)malloc(sizeof(double) * (size_t)ncv);
        double *ihb = (double *)malloc(sizeof(double) * (size_t)ncv);
        double *iq = (double *)malloc(sizeof(double) * (size_t)ncv * ncv);
        double *tau = (double *)malloc(sizeof(double) * (size_t)ncv);
        memcpy(ihb, h + 1, sizeof(double) * (size_t)(ncv - 1));
        memcpy(ihd, h + ldh, sizeof(double) * (size_t)ncv);
        int bad = hor_tridiag_eig(ncv, ihd, ihb, iq, ncv);
        if (!bad && reord) { /* HLanc.hpp:1808-1831 */
            int left = 0, right = ncv - 1;
            for (;;) {
                while (select[left]) ++left;
                while (!select[right]) --right;
                if (left > right) break;
                double t = ihd[left];
                ihd[left] = ihd[right];
                ihd[right] = t;
                for (int i = 0; i < ncv; ++i) {
                    t = iq[(size_t)left * ncv + i];
                    iq[(size_t)left * ncv + i] = iq[(size_t)right * ncv + i];
                    iq[(size_t)right * ncv + i] = t;
                }
                ++left;
                --right;
            }
        }
        if (!bad) {
            memcpy(values, ihd, sizeof(double) * (size_t)nconv);
            hor_dsesrt(nconv, values, ncv, iq, ncv);
            hor_dgeqr2(ncv, nconv, iq, ncv, tau);
            hor_dorm2r_rn(n, ncv, nconv, iq, ncv, tau, vectors, ldv);
            ret = 0;
        }
        free(tau);
        free(iq);
        free(ihb);
        free(ihd);
        free(select);
    }

done:
    inf.ret = ret;
    inf.nconv = nconv;
    inf.spmv = st.spmv;
    inf.cgs = st.cgs;
    inf.rnorm = rnorm;
    if (info) *info = inf;
    free(e);
    free(bounds_saved);
    free(ritz_saved);
    free(work);
    free(q);
    free(bounds);
    free(ritz);
    free(h);
    free(resid);
    free(st.irj);
    free(st.ipj);
    return ret;
}
