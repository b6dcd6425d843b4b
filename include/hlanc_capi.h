/*
 * hlanc_capi.h -- C ABI of libhlanc_hip.so, the MI355X (gfx950) implementation of
 * HLanc's implicitly restarted Lanczos hot path.
 *
 * Plain C: pointers, sizes and int status codes; no C++ or torch types.  The
 * drop-in C++ header include/HLanc.hpp is written on top of these entry points;
 * the Python package hlanc_amd binds them with ctypes.  Each entry point names
 * the reference interface it replaces (HLanc.hpp file:line).
 *
 * Status codes: HL_OK (0); HL_NOT_CONVERGED (1) is the reference's numerical
 * "return 1" (HLanc.hpp:1692,1762,1791,1806,1838,1841); negative values are
 * argument or runtime errors; HL_ERR_HIP_BASE + hipError_t for HIP failures.
 * No entry point calls exit().
 *
 * Threading: every entry point is reentrant for distinct handles.  Calls on one
 * operator handle from several host threads are serialized internally.  Device
 * work runs on the stream passed in (0 = the default stream) or, for the solver
 * entry points, on a stream owned by the solve.  The caller's current device is
 * restored on return.
 */
#ifndef HLANC_CAPI_H
#define HLANC_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HL_OK 0
#define HL_NOT_CONVERGED 1
#define HL_ERR_ARG (-1)
#define HL_ERR_NO_DEVICE (-2)
#define HL_ERR_OOM (-3)
#define HL_ERR_IO (-4)
#define HL_ERR_FORMAT (-5)
#define HL_ERR_STATE (-6)
#define HL_ERR_HIP_BASE 1000

/* Storage precision of matrix values or of the Krylov basis. */
#define HL_PREC_F64 0
#define HL_PREC_F32 1
/* Basis only: the FP64 basis plus an FP32 copy of it that the projection-only
 * Gram-Schmidt passes read (see hl_options.basis_prec). */
#define HL_PREC_F64_PROJ32 2

/* Widest Krylov basis the solvers take (ncv); HL_PREC_F32 solves take at most
 * HL_MAX_NCV_F32.  The reference takes any ncv <= n (HLanc.hpp:1649-1676). */
#define HL_MAX_NCV 512
#define HL_MAX_NCV_F32 128

typedef struct hl_csr_s *hl_csr_t;

/* Text of the last error seen on this host thread (never NULL). */
const char *hl_last_error(void);
/* Library version "major.minor"; mirrors HLANC_VERSION_MAJOR/MINOR (HLanc.hpp:15-16). */
const char *hl_version(void);
/* Number of visible HIP devices (0 when none). */
int hl_device_count(int *count);
/* Page-locked host memory, e.g. for the `vectors` output of the solvers: the
 * final device-to-host copy of the basis then runs at full link rate instead of
 * through the runtime's staging buffers.  No counterpart in the reference
 * (its outputs are caller-allocated host memory, as these may also be). */
int hl_host_alloc(size_t bytes, void **ptr);
int hl_host_free(void *ptr);

/* ---------------------------------------------------------------- matrices
 * Host CSR in dcsr_matrix_t's layout (HLanc.hpp:74-100): zero-based int32
 * row_ptr[nrows+1], col[nnz], double val[nnz]; sym = lower triangle stored.
 * Memory returned by hl_matrix_* is owned by the library and released with
 * hl_matrix_free. */
typedef struct {
    int nrows;
    int ncols;
    int nnz;
    int sym;
    int *row_ptr;
    int *col;
    double *val;
} hl_matrix;

/* dcsr_matrix_t(format, filename, is_symmetric) readers (HLanc.hpp:103-285):
 * format "pp", "mm" or "bin".  Returns HL_ERR_IO / HL_ERR_FORMAT with a message
 * where the reference prints and exits (HLanc.hpp:52-59). */
int hl_matrix_read(const char *format, const char *filename, int is_symmetric, hl_matrix *out);
/* save_pp / save_mm / save_bin (HLanc.hpp:287-392); "bin" expands symmetric
 * storage to general as the reference does (HLanc.hpp:339-364). */
int hl_matrix_write(const char *format, const char *filename, const hl_matrix *m);
void hl_matrix_free(hl_matrix *m);

/* ---------------------------------------------------------------- operator
 * dcsrmv_operator_gpu(m, dev) (HLanc.hpp:493-514): uploads the CSR to device
 * `dev`; symmetric-lower input is expanded to general storage on upload, which
 * gives the (L + L^T - D) x semantics of HLanc.hpp:408-419 and 450-451.
 * val_prec = HL_PREC_F32 additionally keeps an FP32 copy of the values for the
 * mixed-precision solver.  Host arrays are only read during the call. */
int hl_csr_create(int dev, int nrows, int ncols, int nnz, const int *row_ptr, const int *col,
                  const double *val, int sym_lower, int val_prec, hl_csr_t *out);
/* ~dcsrmv_operator_gpu (HLanc.hpp:516-528). */
int hl_csr_destroy(hl_csr_t h);
/* Shape queries. */
int hl_csr_info(hl_csr_t h, int *dev, int *nrows, int *ncols, long long *nnz_general);

/* dcsrmv_operator_gpu::operator()(x, y) (HLanc.hpp:530-572): y = A x.  x and y
 * may each be host memory or device memory of the operator's device; host y
 * means the result is complete on return, device y means the call is ordered
 * on the default stream.  FP64 values. */
int hl_csr_apply(hl_csr_t h, const double *x, double *y);
/* Device-pointer form on an explicit stream (no pointer probing, no sync).
 * Rows longer than the work-item capacity are summed through the operator's
 * own partial-sum slots, so calls on one operator must not overlap on
 * different streams; solves keep their own slots and may share an operator. */
int hl_csr_spmv(hl_csr_t h, const double *x_dev, double *y_dev, void *stream);
/* The FP32-value kernel (requires val_prec = HL_PREC_F32 at creation). */
int hl_csr_spmv_f32val(hl_csr_t h, const double *x_dev, double *y_dev, void *stream);
/* Algorithmic bytes of one SpMV: (s_A + 4) nnz + 4 (nrows + 1) + 8 ncols + 8 nrows. */
double hl_csr_spmv_bytes(hl_csr_t h, int val_prec);
/* Work-item shape of the operator's SpMV: stream items hold <= rows_cap whole
 * rows (1..2048) with <= nnz_cap nonzeros (256..2048); longer rows are split
 * into chunks.  The default is 2048 / 256.  Changing it changes the order of
 * the row sums (results move at rounding level) and must not overlap solves
 * that use the operator.  No counterpart in the reference (cuSPARSE picks its
 * own schedule). */
int hl_csr_set_items(hl_csr_t h, int nnz_cap, int rows_cap);
/* Column-sliced SpMV: width > 0 regroups the FP64 matrix into virtual rows
 * (row, slice of `width` columns), slice-major, so that the x gathers of a
 * stretch of work items stay inside an L2-sized slice of x; a second kernel
 * adds each row's partials in slice order.  width = 0 drops the form.  Used by
 * every FP64-value SpMV of the operator (standalone calls and FP64-basis
 * solves); costs 12 nnz + 8 virtual rows bytes on the device and 8 bytes per
 * virtual row per solve.  Same rules as hl_csr_set_items.  No counterpart in
 * the reference. */
int hl_csr_set_slices(hl_csr_t h, int width);
int hl_csr_get_slices(hl_csr_t h, int *width, int *npairs); /* outputs may be NULL */
/* Sliced-ELL SpMV: on != 0 stores the FP64 matrix again as chunks of 64
 * consecutive rows, each column-major and as wide as its longest row; one lane
 * sums one row in its column order, with no partial sums.  Suits matrices of
 * even row lengths (stencils).  It drops the column slices (and they drop it).
 * Refused (HL_ERR_ARG) when the chunks would hold more than twice nnz entries
 * or a row holds more than 2048 nonzeros (one wave would walk it alone).  Used by every FP64-value SpMV of the operator; costs 12 bytes per
 * entry on the device, 10 when the operator is banded (every |col - row| <
 * 32768: columns are then stored as 16-bit offsets from the row).  Same rules
 * as hl_csr_set_items.  No counterpart in the reference. */
int hl_csr_set_sell(hl_csr_t h, int on);
int hl_csr_get_sell(hl_csr_t h, int *on, long long *entries); /* outputs may be NULL */
/* Bytes per stored column index of the sliced-ELL form: 2 (banded, offsets
 * from the row), 4 (absolute columns), 0 when the form is off.  No
 * counterpart in the reference. */
int hl_csr_get_sell_col_bytes(hl_csr_t h, int *bytes);
/* SpMV forms (hl_csr_pick_form). */
#define HL_FORM_ITEMS 0   /* stream items over the rows (hl_csr_set_items) */
#define HL_FORM_SLICES 1  /* column slices (hl_csr_set_slices) */
#define HL_FORM_SELL 2    /* sliced-ELL (hl_csr_set_sell) */
#define HL_FORM_STAGED 3  /* staged slices (hl_csr_set_staged) */
/* The SpMV form chosen from the matrix's structure, without timing anything
 * (for callers that cannot time forms, for instance while other solves
 * run): sliced-ELL where its chunks hold at most 1.1 nnz entries, else
 * staged slices where x spans at least 786,432 columns (6 MB, more than one
 * XCD's L2 holds), else stream items; a form that refuses the matrix passes
 * to the next.  The chosen form is built (the others dropped) and *form
 * names it.  The staged form's host planning takes about 45 ns per nonzero
 * (3.3 s at 70 M), more than its faster SpMV saves in one solve of a few
 * hundred steps: it pays off for operators that serve many solves.
 * hl_csr_autotune times the forms instead.  No counterpart in the reference. */
int hl_csr_pick_form(hl_csr_t h, int *form);
/* Device bytes the operator holds: its CSR (and FP32 values), staging, work
 * items and the arrays of its SpMV form.  No counterpart in the reference. */
int hl_csr_device_bytes(hl_csr_t h, double *bytes);
/* Staged-slice SpMV: on != 0 cuts the FP64 matrix into column slices of 16384
 * columns, few enough that a block holds a slice's x (128 KB) in LDS, so that
 * no x gather leaves the CU.  Each row's entries of a slice (cut every 64
 * entries) form a piece; pieces are stored slice-major in ELL chunks of 64, a
 * lane per piece (16-bit slice-local columns), and a first kernel writes one
 * partial per piece.  A second kernel takes a bin of rows holding at most 3584
 * pieces, copies the bin's partials into LDS and adds each row's partials in
 * slice order.  Suits graphs whose gathers miss the caches (power-law and
 * random graphs).  on = 0 drops the form; building it drops the column slices
 * and the sliced-ELL form (and they drop it).  Refused (HL_ERR_ARG) beyond
 * 4000 slices (65.5 M columns) or for a row of more than 3584 pieces.  Used
 * by every FP64-value SpMV of the operator; costs 10 bytes per stored entry
 * (entries are padded to their chunk's longest piece) plus about 12 per piece
 * on the device and 8 bytes per piece per solve.  Same rules as
 * hl_csr_set_items.  No counterpart in the reference. */
int hl_csr_set_staged(hl_csr_t h, int on);
/* on, pieces, stored entries (nnz plus padding), row bins; outputs may be NULL. */
int hl_csr_get_staged(hl_csr_t h, int *on, long long *pieces, long long *entries, int *bins);
/* Autotuner: times `reps` SpMV launches for every shape of a small grid
 * (nnz_cap in {512, 1024, 2048} x rows_cap in {64 .. 2048}), and, for
 * operators of >= 2^20 columns, every shape again under column slices of 2^19,
 * 2^18 and 2^17 columns (widest first; a narrower width must win by 2 %), then
 * the sliced-ELL form where it holds at most 1.5 nnz entries, then, for
 * operators of >= 65536 nonzeros, the staged-slice form (which must win by
 * 2 %), with HIP events on the null stream; keeps the fastest and reports its
 * item shape (outputs may be NULL; hl_csr_get_slices, hl_csr_get_sell and
 * hl_csr_get_staged tell the form).  Same rules as hl_csr_set_items. */
int hl_csr_autotune(hl_csr_t h, int reps, int *nnz_cap, int *rows_cap, double *seconds);

/* ------------------------------------------------------ basis operations
 * Device-pointer forms of mat_v_gpu_t's dense work, exposed for testing and for
 * callers that keep their own basis.  V is column-major with leading dimension
 * ldv (elements), FP64. */
/* One classical Gram-Schmidt pass, mat_v_gpu_t::op (HLanc.hpp:1000-1024), fused:
 *   w_out = w_in - V(:,0:m) h_in      (skipped when h_in == NULL: w_out = w_in)
 *   h_out = V(:,0:m)^T w_out,  *nrm2_out = ||w_out||^2 (both device pointers).
 * w_out may equal w_in or be NULL (not written).  Asynchronous on `stream`;
 * calls from one host thread share a reduction workspace, so calls on
 * different streams of one thread must not overlap. */
int hl_cgs_pass(int dev, int n, int m, const double *V, long long ldv, const double *w_in,
                double *w_out, const double *h_in, double *h_out, double *nrm2_out, void *stream);
/* hl_cgs_pass without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Same operands, kernels
 * and results as hl_cgs_pass in its default split form (bitwise), but its
 * block partial sums live in a fixed device array of the library (4 MiB per
 * GPU, kept until the process ends) instead of a workspace allocated on the
 * first call, and h_out / *nrm2_out (either may be NULL) are written by the
 * finalize kernel directly.  m <= 128 (HL_ERR_ARG otherwise).  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams,
 * other host threads) race and must be serialized by the caller.  Calls no
 * hipMalloc, hipFree or allocating routine. */
int hl_cgs_pass_noalloc(int dev, int n, int m, const double *V, long long ldv, const double *w_in,
                        double *w_out, const double *h_in, double *h_out, double *nrm2_out, void *stream);
/* Out(:, 0:N) = V(:, 0:K) Q(0:K, 0:N): the V.Q product of dsapps and of the
 * Ritz-vector step (HLanc.hpp:1029-1058, 1075-1107).  Out must not alias V.
 * use_mfma = 1 selects the v_mfma_f64_16x16x4 kernel, 0 the FMA kernel. */
int hl_vq(int dev, int n, int K, int N, const double *V, long long ldv, const double *Q, int ldq,
          double *Out, long long ldo, int use_mfma, void *stream);

/* ------------------------------------------------ host math of the restart
 * The FP64 ncv x ncv kernels the solver runs on the host, exported so that
 * they can be checked on their own (no device work; callable without a GPU).
 * All matrices are column-major. */
/* dsteqr("I") (HLanc.hpp:1691,1805): eigenvalues of the symmetric tridiagonal
 * (d[0..n), e[0..n-1)) ascending in d, eigenvectors in z (n x n, ldz >= n); e
 * is overwritten.  Returns 0, or 1 on non-finite input or no convergence. */
int hl_host_tridiag_eig(int n, double *d, double *e, double *z, int ldz);
/* The same eigenvalues, and only the last row of the eigenvector matrix
 * (zlast[i] = Z(n-1, i)), bitwise those of hl_host_tridiag_eig: what the
 * restart's Ritz estimates need (HLanc.hpp:1691-1696). */
int hl_host_tridiag_eig_last(int n, double *d, double *e, double *zlast);
/* dlartg (HLanc.hpp:1514,1536) with MKL's sign convention: c f + s g = r,
 * -s f + c g = 0. */
void hl_host_dlartg(double f, double g, double *c, double *s, double *r);
/* dgeqr2 on a (m x k, lda) in LAPACK's storage (HLanc.hpp:1837), then
 * q = H(0) ... H(k-1) formed explicitly (m x m, ldq): the Q that dorm2r("R",
 * "N") applies (HLanc.hpp:1840). */
int hl_host_householder_q(int m, int k, double *a, int lda, double *tau, double *q, int ldq);
/* The shift sweeps of dsapps (HLanc.hpp:1492-1589) on T = (dia, sub) of order
 * ncv = nev + np (sub[i] = T(i, i-1), sub[0] unused): np implicit QR steps with
 * exact shifts shift[0..np) by bulge chasing with the reference's deflation
 * test; dia/sub become Q^T T Q and q (ncv x ncv, ld ncv) receives Q. */
int hl_host_restart_chase(int nev, int np, const double *shift, double *sub, double *dia, double *q);

/* ---------------------------------------------------------------- tuning
 * Process-wide kernel parameters, set before solves start (not thread-safe
 * against running solves).  Keys: "cgs_blocks_max" (grid cap of the CGS
 * passes, 1..2048), "cgs_flags" (bit 0 non-temporal basis loads, bit 1
 * prefetch of the next tile, bit 3 finish each pass in a second kernel
 * instead of a last-block reduction; bit 2 is refused), "f32_phase_tol_exp"
 * (mixed precision: the FP32-basis phase ends at relative Ritz estimates of
 * 10^-value, 3..9, default 6), "cgs_f32_tiles" (16-row tiles per trip of the
 * FP32-basis CGS passes, 1 or 2, default 1), "dgks_skip_exp" (HL_PREC_F64_PROJ32
 * solves skip a step's DGKS correction when its exact coefficients h satisfy
 * ||h|| <= 10^-value ||r||; 0..16, default 13, 0 = never skip), "batch_engines"
 * (hl_ds_solve_batch: engines, each on its own stream and host thread, that
 * share the batch's queue and caps; 1..8, default 3; the results are the same
 * bit for bit).
 * The solvers always run the split pass over a 16-row tiled basis; cgs_flags
 * selects the variant of hl_cgs_pass and of the CGS probe (hlanc_probes.h).  No counterpart in the
 * reference (its kernels are fixed). */
int hl_tune_set(const char *key, int value);
int hl_tune_get(const char *key, int *value);
/* The measurement probes (hl_bench_*) are declared in hlanc_probes.h. */

/* ------------------------------------------------------------------ solver */

/* y = A x on host memory, for arbitrary operators (HLanc.hpp:1424: the solver
 * hands the operator host pointers of length n). */
typedef void (*hl_host_op_fn)(void *ctx, const double *x, double *y);

typedef struct {
    double tol;        /* convergence tolerance; <= 0 -> machine epsilon (HLanc.hpp:1651) */
    int maxiter;       /* restart cap; <= 0 -> 200 (HLanc.hpp:1682); a two-phase
                          (HL_PREC_F32) solve counts the restarts of both phases */
    const double *v0;  /* host start vector of length n, or NULL: seeded generator */
    uint64_t seed;     /* generator seed when v0 == NULL */
    int basis_prec;    /* HL_PREC_F64 (hl_options_default): the FP64 basis with the
                          reference's DGKS rule exactly as written (HLanc.hpp:1442-1465:
                          correct when ||r|| < 0.717 ||w||, at most twice);
                          HL_PREC_F64_PROJ32 (opt-in, the fastest mode measured): the FP64
                          basis plus an FP32 copy.  It departs from the reference's
                          Gram-Schmidt in four ways -- the first projection is the Lanczos
                          three-term one (alpha_j from FP64 column j, beta_j known); the
                          subtracting pass reads the FP64 basis and projects exactly; the
                          first DGKS correction is pending after every step and reads the
                          FP32 copy (the FP64 basis when the step cancelled strongly,
                          ||r|| < 1e-2 ||w||); and it is skipped when those exact
                          coefficients are negligible (||h|| <= 1e-13 ||r||).  The basis
                          stays orthogonal to FP64 rounding while reading at most 12
                          instead of 24 bytes per entry; host operators run plain FP64;
                          HL_PREC_F32: two phases for device CSR operators -- the basis in
                          FP32 (and the operator's FP32 values when it was created with
                          them) to a relative Ritz estimate of 1e-6, then an FP64 restart
                          from the sum of that phase's wanted Ritz vectors.  Every mode
                          returns FP64 results to tol. */
    int use_mfma;      /* 1 (default): restart and Ritz-vector V.Q on v_mfma_f64_16x16x4f64
                          (a restart whose K x N block of Q exceeds 48 KB runs on the FMA
                          kernel); 0: FMA kernel.  The results differ in rounding only. */
    int flags;         /* HL_FLAG_* bits */
    int refine;        /* HL_REFINE_NONE (default): the returned values are the Ritz
                          values of T, as the reference returns them (HLanc.hpp:1833);
                          HL_REFINE_RQ: each returned value is replaced by the Rayleigh
                          quotient v^T A v / v^T v of its returned FP64 vector, computed
                          with the FP64 SpMV of the residual check (the order of the
                          values is kept).  Needs the residual check (no HL_FLAG_NO_VERIFY). */
    double *residuals; /* host, >= nev, or NULL: on success receives each returned pair's
                          ||A v_i - l_i v_i||_2 / (||A||_est ||v_i||_2) (see hl_info) */
    int reserved[4];
} hl_options;

/* hl_options.refine values. */
#define HL_REFINE_NONE 0
#define HL_REFINE_RQ 1

typedef struct {
    int ret;                /* status as returned */
    int restarts;           /* restart iterations run */
    int nconv;              /* converged Ritz values */
    int graph_sweeps;       /* dsaitr sweeps replayed from a captured HIP graph */
    long long spmv;         /* operator applications (the residual check's excluded) */
    long long cgs_passes;   /* CGS passes incl. DGKS corrections */
    double rnorm;           /* final residual norm */
    double seconds;         /* wall time of the call */
    double max_rel_resid;   /* residual check: max over the returned pairs of
                               ||A v - l v||_2 / (||A||_est ||v||_2), with A v formed by the
                               FP64 SpMV (or the host operator) after the Ritz-vector step and
                               l the returned value; -1 when no check ran (failure, or
                               HL_FLAG_NO_VERIFY).  The convergence test itself is a Ritz
                               estimate (HLanc.hpp:1376-1384); this is the true residual. */
    double refine_shift;    /* mixed precision: largest relative change of a returned value
                               from the FP32-basis phase to the FP64 result */
    long long spmv_f32;     /* mixed precision: operator applications of the FP32-basis phase
                               (included in spmv) */
    int restarts_f32;       /* mixed precision: restarts of the FP32-basis phase (included in
                               restarts) */
    int proj64_steps;       /* FP32-copy projections: Lanczos steps whose DGKS correction read
                               the FP64 basis (strong cancellation) */
    double anorm_est;       /* ||A||_est: the largest |eigenvalue| of the final ncv x ncv
                               tridiagonal, a lower bound of ||A||_2 (so max_rel_resid errs
                               high); 0 when no check ran */
    double rq_shift;        /* residual check: max over the returned pairs of
                               |v^T A v / v^T v - theta| / |theta|, the Rayleigh quotient's
                               distance from the Ritz value (applied with HL_REFINE_RQ) */
    int verified;           /* pairs checked (nev on success, 0 otherwise) */
    int pad;
} hl_info;

/* hl_options.flags bits. */
#define HL_FLAG_NO_GRAPH 1  /* launch every sweep's kernels one by one; by default a sweep
                               shape met a second time is captured as a HIP graph and
                               replayed (same kernels, same results) */
#define HL_FLAG_PROFILE 2   /* time every SpMV, CGS and V.Q launch with HIP events */
#define HL_FLAG_NO_VERIFY 4 /* skip the residual check: no SpMVs after the Ritz vectors;
                               values-only batch items then skip the Ritz-vector GEMM too */

/* Per-kernel-class totals of the last profiled solve (or batch) on the calling
 * thread: index 0 SpMV, 1 the two CGS passes of every Lanczos step (partials
 * kernels), 2 V.Q GEMM, 3 other kernels (CGS finalize, norms), 4 the DGKS
 * correction passes (partials kernels; launched every step, exiting at once
 * when not needed), 5 the engine's host work between sweeps (launches = sweeps,
 * seconds = host time: restart logic, launch tables, starts).  bytes are
 * algorithmic HBM bytes (each operand
 * counted once per launch; a correction pass counts only where it ran, from
 * the device's own counters). */
typedef struct {
    long long launches[6];
    double seconds[6];
    double bytes[6];
} hl_profile;
int hl_profile_last(hl_profile *out);

/* Fills *o with the reference defaults (tol = eps, maxiter = 200, generator
 * seed 0, the FP64 basis with the reference's DGKS rule, values not refined)
 * plus this library's: MFMA V.Q, graphs on, the residual check on. */
void hl_options_default(hl_options *o);

/* ds_solver_gpu<dcsrmv_operator_gpu> (HLanc.hpp:1868-1872 driving 1649-1850):
 * computes the nev largest-algebraic eigenpairs.  values (host, >= nev) gets
 * them ascending; vectors (column-major ldv x ncv, ldv >= n) gets all ncv
 * basis columns, the first nev being the unit Ritz vectors.  vectors may be
 * host memory, as the reference's (pageable, or page-locked for a copy at
 * link rate), or device memory on the operator's device, which keeps the
 * result in HBM.  Both are written only on success.  The operator's SpMV runs
 * on device pointers: no PCIe traffic per Lanczos step.  opts/info may be NULL. */
int hl_ds_solve_csr(hl_csr_t A, int nev, int ncv, double *values, double *vectors, int ldv,
                    const hl_options *opts, hl_info *info);

/* One solve of a batch (hl_ds_solve_batch). */
typedef struct {
    hl_csr_t A;          /* square operator on the batch's device */
    int nev, ncv;        /* as hl_ds_solve_csr */
    double *values;      /* host, >= nev: ascending eigenvalues on success */
    double *vectors;     /* column-major ldv x ncv, host or device memory as
                            hl_ds_solve_csr, or NULL: values only (the
                            Ritz-vector GEMM is skipped) */
    int ldv;
    int pad0;
    hl_options opts;     /* any basis_prec */
    double arrival;      /* seconds after the call starts; the solve is admitted
                            no earlier (0: at once); must not decrease along the
                            array */
    hl_info info;        /* out */
    int status;          /* out: HL_OK, HL_NOT_CONVERGED or an error code */
    int pad1;
    double t_admit;      /* out: seconds after the call starts */
    double t_done;       /* out: seconds after the call starts */
} hl_solve_item;

/* Many eigensolves on one device with continuous batching.  No counterpart in
 * the reference, which solves one operator per ds_solver call: every solve in
 * flight advances in lockstep, one segmented SpMV launch and four segmented
 * CGS passes per Lanczos step serving all of them; solves are admitted in
 * array order at restart boundaries, once their arrival time has passed,
 * while fewer than max_active (<= 0: no cap) are in flight and their
 * workspaces fit in hbm_budget bytes (<= 0: no cap; one solve is always
 * admitted when nothing is in flight).  The solves in flight are spread over
 * "batch_engines" engines (hl_tune_set; default 3), each advancing its own in
 * lockstep on its own stream and host thread, so that one engine's host work
 * between sweeps overlaps the others' kernels; the caps hold over all of them.
 * A workspace is allocated on admission and freed when its solve
 * ends.  Each item's results are bitwise those of hl_ds_solve_csr on the same
 * operator and options.  Returns HL_OK once every valid item has run (see each
 * item's status), or the HIP error that stopped the batch. */
int hl_ds_solve_batch(int dev, hl_solve_item *items, int count, int max_active, double hbm_budget);

/* ds_solver_gpu<any callable> (HLanc.hpp:1868-1872): the operator is a host
 * callback called with host pointers, as the reference does (HLanc.hpp:1424). */
int hl_ds_solve_host_op(int dev, int n, hl_host_op_fn op, void *ctx, int nev, int ncv,
                        double *values, double *vectors, int ldv, const hl_options *opts,
                        hl_info *info);

/* ---------------------------------------------------------------- serving
 * A request loop in place of the blocking call: no counterpart in the reference,
 * which solves one operator per ds_solver_gpu call (HLanc.hpp:1868-1872).  A
 * server owns a device and two host threads.  Requests are submitted one by one
 * while others run; the serving thread advances every request in flight in
 * lockstep (one segmented SpMV and four segmented CGS launches per Lanczos step
 * for all of them, each request's basis, tridiagonal and convergence state
 * resident on the device), and between sweeps retires finished requests and
 * admits queued ones in submission order, while fewer than max_active (<= 0: no
 * cap) are in flight and their footprints fit in hbm_budget bytes (<= 0: no cap;
 * an idle server always admits one).  A request's footprint is its workspace
 * (basis pair, vectors, partial sums) plus, for hl_serve_submit_csr, its
 * operator: 12 nnz + 4 (n + 1) bytes and 16 n of staging, uploaded by the
 * server's loader thread once the request is admitted (overlapping the sweeps
 * of the requests already running) and freed when it is done.  Once the
 * loader has built such an operator, the request's footprint is counted
 * again from the operator's device bytes (its SpMV form's arrays included,
 * hl_csr_device_bytes) and the workspace its form needs.  Each request's
 * results are bitwise those of hl_ds_solve_csr on the same operator and
 * options, whatever runs beside it.  All entry points are thread-safe. */
typedef struct hl_server_s *hl_server_t;

#define HL_PENDING 2          /* hl_serve_poll / hl_serve_wait: not done yet */
#define HL_SERVE_SELL_EVEN 1  /* flag: operators the server uploads take the sliced-ELL form
                                 where it holds at most 1.1 nnz entries (stencils) */

typedef struct {
    int status;          /* HL_OK, HL_NOT_CONVERGED or an error code */
    int pad;
    hl_info info;
    double t_submit;     /* seconds after hl_serve_create */
    double t_admit;      /* admitted: counted against the budget */
    double t_start;      /* operator on the device: joined the lockstep sweeps */
    double t_done;       /* outputs written */
    double footprint;    /* bytes counted against the budget */
} hl_serve_result;

typedef struct {
    long long submitted, done;
    int queued;          /* submitted, not admitted */
    int in_flight;       /* admitted, not done */
    double bytes_in_use; /* footprints of the requests in flight */
    double seconds;      /* since hl_serve_create */
} hl_serve_stats_t;

int hl_serve_create(int dev, int max_active, double hbm_budget, int flags, hl_server_t *out);
/* A request on an operator the caller keeps alive until the request is done.
 * values (host, >= nev) and vectors (as hl_ds_solve_csr, or NULL: values only)
 * are written when it is done; *ticket names it. */
int hl_serve_submit(hl_server_t s, hl_csr_t A, int nev, int ncv, const hl_options *opts, double *values,
                    double *vectors, int ldv, long long *ticket);
/* A request on a host CSR matrix (dcsr_matrix_t's layout, square, sym_lower as
 * hl_csr_create): the server uploads it after admission and frees it when the
 * request is done.  The arrays are read, not copied: they must stay valid
 * until the request is done. */
int hl_serve_submit_csr(hl_server_t s, int nrows, int nnz, const int *row_ptr, const int *col,
                        const double *val, int sym_lower, int nev, int ncv, const hl_options *opts,
                        double *values, double *vectors, int ldv, long long *ticket);
/* HL_OK with *out filled once the request is done, HL_PENDING before. */
int hl_serve_poll(hl_server_t s, long long ticket, hl_serve_result *out);
/* Blocks until the request is done (timeout_s < 0: no limit); HL_PENDING on timeout. */
int hl_serve_wait(hl_server_t s, long long ticket, double timeout_s, hl_serve_result *out);
/* Forgets a done request (its result can no longer be polled). */
int hl_serve_release(hl_server_t s, long long ticket);
/* Counters; returns the HIP failure that stopped the server, if any. */
int hl_serve_stats(hl_server_t s, hl_serve_stats_t *out);
/* Takes no more requests, finishes every submitted one, stops the threads and
 * frees the server.  Returns the HIP failure that stopped the server, if any. */
int hl_serve_destroy(hl_server_t s);

#ifdef __cplusplus
}
#endif
#endif
