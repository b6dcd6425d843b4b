/*
 * hlanc_oracle.h -- CPU restatement of HLanc's implicitly restarted Lanczos path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product is
 * compared against, and the CPU baseline that bench.py times beside it.  Product
 * code (hlanc_amd/, include/) never links, loads or calls it.
 *
 * Every function follows a named part of the reference header HLanc.hpp; the
 * citation sits on each declaration below.  Parity status of each part (what pins
 * it) is stated in the comment block at the top of hlanc_oracle.c.
 */
#ifndef HLANC_ORACLE_H
#define HLANC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y = A x callback: the reference's operator concept (HLanc.hpp:1403-1404,1424). */
typedef void (*hor_op_fn)(void *ctx, const double *x, double *y);

/* CSR matrix handed to hor_csr_op (HLanc.hpp:95-99 layout; sym = lower-stored). */
typedef struct {
    int nrows;
    int ncols;
    const int *row_ptr;
    const int *col;
    const double *val;
    int sym;
} hor_csr;

/* Naive CSR SpMV, general or symmetric-lower semantics (HLanc.hpp:403-431). */
void hor_csr_spmv(const hor_csr *A, const double *x, double *y);
/* Same, with the hor_op_fn signature (ctx = hor_csr*). */
void hor_csr_op(void *ctx, const double *x, double *y);

/* Givens rotation with MKL's dlartg sign convention (called at HLanc.hpp:1514,1536). */
void hor_dlartg(double f, double g, double *c, double *s, double *r);
/* The shift sweeps of dsapps (HLanc.hpp:1492-1589) on T = (dia, sub), sub[i] =
 * T(i, i-1): dia/sub become Q^T T Q, q (ncv x ncv) receives Q. */
void hor_dsapps_chase(int nev, int np, const double *shift, double *sub, double *dia, double *q);

/* Symmetric tridiagonal eigen-decomposition, eigenvalues ascending, eigenvectors
 * in z (ldz >= n) -- the role of dsteqr("I") at HLanc.hpp:1691,1805.
 * d[n] diagonal (overwritten by eigenvalues), e[n-1] off-diagonal (destroyed).
 * Returns 0, or 1 if a non-finite entry is present (dsteqr info != 0). */
int hor_tridiag_eig(int n, double *d, double *e, double *z, int ldz);

/* Householder QR of an m x k block, LAPACK dgeqr2 storage (HLanc.hpp:1837). */
int hor_dgeqr2(int m, int k, double *a, int lda, double *tau);
/* C(nr x nc) := C * H(0) H(1) ... H(k-1), LAPACK dorm2r("Right","Notranspose")
 * (HLanc.hpp:935-942, GPU form 1075-1107). */
void hor_dorm2r_rn(int nr, int nc, int k, const double *a, int lda, const double *tau,
                   double *c, int ldc);

/* Counter-based start-vector generator: x_i = 2u_i - 1, u_i in [0,1) from a
 * splitmix64 hash of (seed, i).  Replaces dlarnv in dgetv0 (HLanc.hpp:1368-1374). */
void hor_start_vector(int n, uint64_t seed, double *x);
/* The splitmix64 output mixer used by hor_start_vector (Steele, Lea & Flood,
 * "Fast splittable pseudorandom number generators", OOPSLA 2014). */
uint64_t hor_mix64(uint64_t z);

typedef struct {
    double tol;         /* <= 0 -> machine epsilon, the reference default (HLanc.hpp:1651) */
    int maxiter;        /* <= 0 -> 200, the reference MAX_ITER (HLanc.hpp:1682) */
    const double *v0;   /* start vector, or NULL for hor_start_vector(seed) */
    uint64_t seed;
    int nthreads;       /* OpenMP threads for the n-sized loops; <= 0 -> 1 */
    int max_restarts_then_stop; /* > 0: stop after this many restarts (timing samples) */
} hor_opts;

typedef struct {
    int ret;            /* ds_solver return code */
    int restarts;       /* completed restart iterations */
    int nconv;
    long long spmv;     /* operator applications */
    long long cgs;      /* CGS passes (1 per step + DGKS corrections) */
    double rnorm;       /* final residual norm */
} hor_info;

/* ds_solver with the host (mat_v_mkl_t) basis: V lives in `vectors`
 * (HLanc.hpp:1649-1850, 1857-1861). */
int hor_ds_solver(int n, int nev, int ncv, double *values, double *vectors, int ldv,
                  hor_op_fn op, void *ctx, const hor_opts *opts, hor_info *info);

#ifdef __cplusplus
}
#endif
#endif
