// hl_internal.hpp -- shared definitions of libhlanc_hip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "hlanc_capi.h"
#include "hlanc_probes.h"

namespace hl {

constexpr int kWave = 64;        // CDNA wavefront
constexpr int kBlock = 256;      // threads per block in every kernel here
constexpr double kDgks = 0.717;  // DGKS threshold (HLanc.hpp:1444,1451)
constexpr double kEps = 1.1102230246251565e-16;    // dlamch("E") = 2^-53
constexpr double kSafmin = 2.2250738585072014e-308; // dlamch("S")

// ----------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
// The error code of the exception being handled (called from a catch block):
// std::bad_alloc is HL_ERR_OOM, anything else HL_ERR_STATE, with the message
// set.  Every extern "C" entry point ends in such a handler, so no C++
// exception crosses the C ABI.
int exception_code() noexcept;
int hip_status(hipError_t e, const char *what);

#define HL_TRY_HIP(call)                                           \
    do {                                                           \
        hipError_t hl_e_ = (call);                                 \
        if (hl_e_ != hipSuccess) return ::hl::hip_status(hl_e_, #call); \
    } while (0)

// Restores the caller's current device on scope exit.
struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// --------------------------------------------------- device-side solve state
// Scalars of one Lanczos factorization, kept in device memory so that a whole
// dsaitr sweep runs without host round trips.  Written only by the last block
// of a reducing kernel ("epilogue") or by the host between sweeps.
struct Ctl {
    double rnorm;    // ||resid||, becomes beta_{j+1}
    double wnorm;    // ||A v_j|| (DGKS reference norm)
    double xscale;   // factor turning resid into v_{j+1}: 1/rnorm, 1, or 0
    double rscale;   // 1, or 0 once DGKS declared an invariant subspace
    int need;        // another DGKS correction is pending
    int iter;        // DGKS failure count of the current step
    long long cgs;   // CGS passes executed (statistics)
    double nrm2;     // last reduced squared norm (standalone passes)
    long long p2, p3;       // DGKS corrections executed (first, second)
    long long m2sum, m3sum; // sum of their basis widths (for byte counts)
    int dgks_always;        // FP32-copy projections: the first correction is pending after every
                            // subtracting pass (dgks_decide_seg_kernel may cancel it)
    int exact;              // this step's corrections read the FP64 basis (FP32-copy projections)
    long long nexact, mxsum; // steps with exact corrections and the sum of their basis widths
    long long nskip;        // corrections skipped as negligible (FP32-copy projections)
};

static_assert(sizeof(Ctl) % sizeof(double) == 0, "Ctl is read back as doubles");
constexpr int kCtlWords = sizeof(Ctl) / sizeof(double);

// Epilogue modes of the reducing kernels.
enum Epi : int {
    EPI_NONE = 0,    // standalone pass: write h_out and ctl->nrm2 only
    EPI_WNORM = 1,   // first CGS half (h = V^T w, ||w||) right after the SpMV
    EPI_PASS1 = 2,   // w -= V h: alpha_j, beta_j, rnorm, DGKS decision
    EPI_PASS2 = 3,   // first DGKS correction
    EPI_PASS3 = 4,   // second DGKS correction (may zero the residual)
    EPI_RNORM = 5,   // rnorm = ||resid|| after start vector or restart
};

// ------------------------------------------------------------ tuning
// Process-wide kernel parameters (hl_tune_set); fixed while solves run.  The
// solver always runs the split, tiled FP64 CGS pass; cgs_flags selects the
// variant of the standalone pass (hl_cgs_pass, hl_bench_cgs) only.
struct Tuning {
    int f32_phase_tol_exp = 6;  // mixed precision: the FP32 phase stops at a 10^-exp Ritz estimate
    int cgs_f32_tiles = 1;      // tiles per trip of the FP32-basis CGS passes (1 or 2; equal speed, 1 fewer VGPRs)
    int dgks_skip_exp = 13;     // FP32-copy projections: skip the correction when ||h1|| <= 10^-exp ||r|| (0: never)
    int cgs_blocks_max = 2048; // grid cap of the CGS passes (<= kMaxReduceBlocks)
    int batch_engines = 3;     // hl_ds_solve_batch: engines (streams and host threads) sharing a batch
    int cgs_flags = 9;         // bit 0: non-temporal basis loads, bit 1: prefetch next tile,
                               // bit 3: split (partials kernel + finalize kernel); bit 2 unused
};
Tuning &tuning();
constexpr int kMaxReduceBlocks = 4096;
// Basis widths: a fused CGS pass holds up to kNarrowCols columns of a tile in
// registers (32 per lane group); wider passes (up to kMaxNcv columns) run as a
// subtracting kernel and a projecting kernel over 128-column chunks.
constexpr int kNarrowCols = HL_MAX_NCV_F32;
constexpr int kMaxNcv = HL_MAX_NCV;

// ------------------------------------------------------------ basis layout
// Storage of a basis V (n rows): column-major with leading dimension ld, or
// "tiled": blocks of 16 consecutive rows, each block holding ld columns of 16
// contiguous doubles, so a 16-row slice of V(:, 0:m) is one contiguous run.
struct VLayout {
    long long ld = 0;
    bool tiled = false;
};
template <bool TILED>
__host__ __device__ inline long long vpos(long long r, long long c, long long ld)
{
    return TILED ? (((r >> 4) * ld + c) << 4) + (r & 15) : c * ld + r;
}
// Columns per tile of a solver-owned tiled basis: ncv rounded up to a multiple
// of 4, so that each of a wave's 4 lane groups reads its ceil(m/4) columns of
// a tile inside that tile.
inline long long tiled_ld(int ncv) { return (static_cast<long long>(ncv) + 3) / 4 * 4; }
// One column j of a basis, as a write target (base == nullptr: none).
struct VColumn {
    double *base = nullptr;
    VLayout L;
    int j = 0;
};

// ------------------------------------------------------------ SpMV plan
// One work item of the adaptive CSR SpMV.  tpr > 0: a "stream" block of whole
// short rows [row0, row1) covering nnz [p0, p1), reduced with tpr lanes per
// row.  tpr == 0: one chunk of a long row row0, nnz [p0, p1); nchunks > 1
// means partial sums meet in slot [slot, slot + nchunks).
struct SpmvBlk {
    int row0, row1, p0, p1;
    int tpr, chunk, nchunks, slot;
};

constexpr int kNnzPerBlock = 2048;  // stream item capacity (8 nnz per thread), default and maximum
constexpr int kRowsPerBlock = 256;  // default rows per stream item
constexpr int kMaxRowsPerBlock = 2048;
constexpr int kLongChunk = 8192;    // nnz per chunk of a long row

// The work items of one CSR (its rows, or the virtual rows of the sliced form)
// and the long-row slots of standalone calls.
struct ItemSet {
    SpmvBlk *blks = nullptr;  // device
    int nblk = 0;
    double *chunk_part = nullptr;
    unsigned *chunk_cnt = nullptr;
    int nslots = 0;
    int nnz_cap = kNnzPerBlock, rows_cap = kRowsPerBlock; // work-item shape in use
};

struct CsrPlan {
    int dev = 0;
    int nrows = 0, ncols = 0;
    long long nnz = 0;        // general (expanded) nnz
    int *row_ptr = nullptr;   // device
    int *col = nullptr;       // device
    double *val = nullptr;    // device FP64 values
    float *val32 = nullptr;   // device FP32 values (optional)
    ItemSet items;            // over the rows
    double *x_stage = nullptr, *y_stage = nullptr;  // for host-pointer calls
    // Column-sliced form of the FP64 matrix (slice_w > 0): the nonzeros regrouped
    // into virtual rows (row, slice) of slice_w columns each, slice-major and
    // rows ascending within a slice, so that the x gathers of one stretch of
    // work items stay within one slice (L2-sized).  Each virtual row's sum is a
    // partial; a second kernel adds a row's partials in slice order.
    int slice_w = 0;
    int npairs = 0;           // virtual rows
    int *s_rp = nullptr;      // npairs + 1
    int *s_col = nullptr;     // nnz, regrouped
    double *s_val = nullptr;  // nnz, regrouped
    int *r_pp = nullptr;      // nrows + 1: each row's partials in r_pairs
    int *r_pairs = nullptr;   // npairs: virtual row of each partial, slice order
    double *s_part = nullptr; // npairs: partials of standalone calls
    ItemSet sitems;           // over the virtual rows
    // Sliced-ELL form of the FP64 matrix (sell_on; exclusive with the column
    // slices): chunks of 64 consecutive rows, each stored column-major (entry k
    // of the chunk's row l at e_ptr[c] + 64 k + l) and as wide as its longest
    // row.  One lane sums one row in column order; padding is never read past
    // a row's own length.
    bool sell_on = false;
    int nchunks = 0;
    long long sell_entries = 0;  // e_ptr[nchunks]
    long long *e_ptr = nullptr;  // nchunks + 1
    int *e_col = nullptr;        // e_ptr[nchunks] entries (padding: column 0); null when banded:
    short *e_d16 = nullptr;      // banded operators (every |col - row| < 2^15): col - row (padding: 0)
    double *e_val = nullptr;     // (padding: 0)
    // Staged-slice form of the FP64 matrix (stg_on; exclusive with the two
    // forms above): column slices of 2^kStagedLw columns, small enough that a
    // block holds the slice's x in LDS.  Each row's entries of a slice, cut
    // into pieces of <= kStagedPiece, are one partial sum.  Pieces are stored
    // slice-major in ELL chunks of 64 (a lane per piece, column-major, 16-bit
    // slice-local columns); within a slice they are grouped by row bin, and by
    // descending length inside a (slice, bin) run.  Row bins hold <= kBinCap
    // pieces; the second kernel copies a bin's runs into LDS and adds each
    // row's partials in slice order through a per-row list of LDS offsets.
    // Within each group of chunks a wave of the first kernel takes at once,
    // lanes are re-ordered by piece length; a lane's code names the slot its
    // partial goes to.
    bool stg_on = false;
    int stg_nsl = 0;                       // slices
    int stg_nchunks = 0;                   // 64-piece chunks (each slice padded to whole chunks)
    int stg_nsup = 0;                      // blocks of the first kernel
    int stg_nbins = 0;                     // blocks of the second kernel
    long long stg_npieces = 0;             // pieces, without the padding
    long long stg_entries = 0;             // s_cptr[nchunks]: stored entries (ELL padding included)
    long long *s_cptr = nullptr;           // nchunks + 1: first entry of each chunk
    unsigned short *s_plen = nullptr;      // 64 nchunks: lane codes, entries of the piece (padding 0) | slot << 7
                                           // (the piece's slot in its wave's group of chunks)
    unsigned short *s_ecol = nullptr;      // s_cptr[nchunks]: slice-local columns
    double *s_eval = nullptr;              // s_cptr[nchunks]: values
    int4 *s_sup = nullptr;                 // nsup: {slice, first chunk, end chunk, 0}
    int4 *s_bmeta = nullptr;               // nbins: {first row, end row, first list offset, end list offset}
    int *s_run = nullptr;                  // nbins x (2 nsl + 1 + kBinCap / kBinStride): each bin's run table
                                           // (LDS offset of each slice's run and the total, first piece minus
                                           // LDS offset, the run of every kBinStride-th LDS index)
    int *s_lpp = nullptr;                  // nrows + 1: each row's pieces in s_lst
    unsigned short *s_lst = nullptr;       // npieces: LDS offset of each piece, slice order
    double *s_spart = nullptr;             // 64 nchunks: partials of standalone calls
};

// Staged-slice form limits (the form refuses matrices beyond them).
constexpr int kStagedLw = 14;            // slices of 2^14 columns (x of a slice: 128 KB of LDS)
constexpr int kStagedThreads = 1024;     // threads per block of the first staged kernel
constexpr int kBinThreads = 512;         // threads per block of the second (bins) kernel
constexpr int kStagedPiece = 64;         // entries per piece (a piece's length fits a byte)
constexpr int kBinCap = 3584;            // pieces per row bin (28 KB of partials and 7 KB of offsets in LDS:
                                         // four bins per CU)
constexpr int kBinStride = 16;           // the run table names the run of every 16th partial
constexpr int kStagedMaxSlices = 4000;   // slices (the bins kernel keeps two ints per slice in LDS)
constexpr int kStagedSupEntries = 65536; // entries per block of the first kernel (about)

int csr_plan_build(CsrPlan &p, int nrows, int ncols, const int *rp, const int *col,
                   const double *val, int sym_lower, int val_prec);
void csr_plan_free(CsrPlan &p);
// Re-cuts the work items of a CSR of `nrows` rows (host row pointers R) into
// `set` (see hl_csr_set_items).
int plan_items(ItemSet &set, const int *R, int nrows, int nnz_cap, int rows_cap);
// Builds (width > 0) or drops (width == 0) the column-sliced form; its items
// take the plan's current item shape.
int plan_slices(CsrPlan &p, int width);
// Builds (on) or drops the sliced-ELL form; building it drops the column slices.
// max_fill > 0 refuses (HL_ERR_ARG) a form of more than max_fill x nnz entries.
int plan_sell(CsrPlan &p, bool on, double max_fill);
// Builds (on) or drops the staged-slice form; building it drops the other two
// forms.
int plan_staged(CsrPlan &p, bool on);
// Device bytes the plan holds: the CSR, its FP32 values, staging, work items
// and the arrays of the form in use.
double csr_device_bytes(const CsrPlan &p);
// The SpMV form chosen from the matrix's structure, without timing (see
// hl_csr_pick_form): *form is HL_FORM_SELL, HL_FORM_STAGED or HL_FORM_ITEMS.
int csr_pick_form(CsrPlan &p, int *form);
constexpr double kPickSellFill = 1.1;        // sliced-ELL where chunks hold <= 1.1 nnz entries
constexpr long long kPickStagedCols = 786432; // staged slices where x spans >= 6 MB (beyond one XCD's L2)

// y = A (x * xscale) with the optional fused write vcol(r) = x[r] * xscale.
// xscale == nullptr means 1.  val_prec selects the FP64 or FP32 value array.
int launch_spmv(const CsrPlan &p, const double *x, const double *xscale, double *y, VColumn vcol,
                int val_prec, hipStream_t s);

// ------------------------------------------------------- segmented launches
// Several solves share one launch per kernel: each solve is a segment of the
// grid, and every block computes exactly what it computes when the solve runs
// alone (same block count, same partial-sum order), so results are bitwise the
// same alone and in a batch.
// Precision kinds of a segment: each segmented launch serves solves of one
// kind (the engine groups a step's solves by kind).  `solo` launches serve
// member 0 of the table alone (an engine with one solve): their blocks read
// the member without waiting for their step entry.
enum SegKind {
    SEG_F64 = 0,     // FP64 basis and values
    SEG_V32 = 1,     // FP32 basis, FP64 values
    SEG_V32_A32 = 2, // FP32 basis and values
    SEG_PROJ32 = 3,  // FP64 basis and values; projection passes read an FP32 copy
    SEG_NKIND = 4
};
// Order of the kinds within a step's entries, so that every launch group is
// one contiguous run: SpMV {V32} {V32_A32} {PROJ32, F64}; CGS passes that read
// the FP64 basis of PROJ32 solves (EPI_PASS1, EPI_PASS3; numbered by cgs0s)
// FP32 {V32, V32_A32}, FP64 {PROJ32, F64}; passes that read its FP32 copy
// (EPI_WNORM, EPI_PASS2; numbered by cgs0) FP32 {V32, V32_A32, PROJ32}, FP64
// {F64}; and the exact first corrections of PROJ32 solves, FP64 {PROJ32}
// (numbered by cgs0s, as PROJ32 leads its cgs0s group).
constexpr int kSegOrder[SEG_NKIND] = {SEG_V32, SEG_V32_A32, SEG_PROJ32, SEG_F64};
// FP32-copy projections: a step whose first pass leaves ||r|| < kProjExact
// ||w|| (strong cancellation) runs its DGKS correction on the FP64 basis,
// since the FP32 copy's rounding (~6e-8 ||w||) is then no longer small
// against r; the second correction always reads the FP64 basis.
constexpr double kProjExact = 1e-2;
struct SegMember {               // per solve, fixed during a sweep
    const SpmvBlk *blks;         // CSR operator (null for host operators): items over rp/col/val
    const int *rp, *col;
    const double *val;
    const float *val32;          // FP32 values (SEG_V32_A32 solves)
    double *part;                // sliced form: partials of the virtual rows (rp/col/val are the
    const int *r_pp, *r_pairs;   // virtual CSR), added per row by the second kernel; else null
    double *chunk_part;
    unsigned *chunk_cnt;
    const long long *e_ptr;      // sliced-ELL form (FP64 values), else null: then blks/rp/col/val
    const int *e_col;
    const short *e_d16;          // banded sliced-ELL: columns relative to the row (else null, e_col set)
    const double *e_val;
    int nchunks;
    // staged-slice form (FP64 values and basis), else s_sup null: its first
    // kernel writes the pieces' partials into part, its bins kernel adds them
    const int4 *s_sup;
    const long long *s_cptr;
    const unsigned short *s_plen;
    const unsigned short *s_ecol, *s_lst;
    const double *s_eval;
    const int4 *s_bmeta;
    const int *s_run, *s_lpp;
    int stg_nsl;
    void *V;                     // tiled basis in use (double, or float for SEG_V32*)
    const void *Vproj;           // basis read by the projection-only CGS passes (V, or the FP32 copy)
    float *V32;                  // FP32 copy written beside V (SEG_PROJ32), else null
    long long vld;
    double *r, *w, *h0, *h1, *hsub, *hdia, *partials;
    Ctl *ctl;
    int n, nb;                   // rows; CGS blocks = cgs_blocks(n)
    int ncv;                     // basis width (read-back size)
};
struct SegStep {                 // per solve and Lanczos step
    int mem;                     // index into the SegMember table
    int j;                       // step index (projects onto j + 1 columns)
    int spmv0;                   // first block in the SpMV launch
    int cgs0;                    // first block in the launches of EPI_WNORM and EPI_PASS2
    int cgs0s;                   // first block in the launches of EPI_PASS1 and EPI_PASS3
    int fin0;                    // first block in the finalize launch
    int red0;                    // first block in the row-sum launch of sliced SpMVs
    int stg0;                    // first block in the first launch of staged-slice SpMVs
    int bin0;                    // first block in their bins launch
};
// Partials kernel of CGS pass `mode` (EPI_WNORM .. EPI_PASS3) for every listed
// step, then the per-column finalize kernel (j + 2 blocks per step).
// v32: the launch reads FP32 bases (SegMember::Vproj); sub: its blocks are
// numbered by SegStep::cgs0s (else cgs0).
int launch_cgs_seg(const SegMember *M, const SegStep *S, int ns, int max_cpg, int cgs_blocks_total, int mode,
                   bool v32, bool sub, bool solo, hipStream_t s);
int launch_cgs_seg_finalize(const SegMember *M, const SegStep *S, int ns, int fin_blocks_total, int mode,
                            bool solo, hipStream_t s);
// Skips the pending first correction of FP32-copy solves whose exact projection
// h1 satisfies ||h1|| <= tau ||r|| (one block per listed step; tau <= 0: none).
int launch_dgks_decide_seg(const SegMember *M, const SegStep *S, int ns, double tau, bool solo, hipStream_t s);

// Launch-cost probes (probes.hip): the co-resident block count of the grid
// barrier kernel, its cooperative launch (buf: 2 blocks x kBlock doubles;
// counter, abort and errors zeroed), and an empty kernel.
int grid_barrier_blocks_max(int dev);
int launch_grid_barrier(int blocks, unsigned *counter, int *abort, double *buf, int barriers, int *errors,
                        hipStream_t s);
int launch_empty(int blocks, hipStream_t s);
int launch_spmv_seg(const SegMember *M, const SegStep *S, int ns, int blocks_total, int kind, bool solo, hipStream_t s);
// Second kernel of the sliced SpMVs among S[0, ns): w = row sums of the
// partials, V(:, j) = r xscale (and its FP32 copy).
int launch_spmv_rows_seg(const SegMember *M, const SegStep *S, int ns, int blocks_total, bool solo, hipStream_t s);
// The two kernels of the staged-slice SpMVs among S[0, ns): the pieces'
// partials (stg_blocks blocks), then the row bins (bin_blocks; nsl_max: the
// most slices of a listed operator) writing w and V(:, j) = r xscale (and its
// FP32 copy).
int launch_spmv_staged_seg(const SegMember *M, const SegStep *S, int ns, int stg_blocks, int bin_blocks, int nsl_max,
                           bool solo, hipStream_t s);
constexpr int kRowsPerSumBlock = 256; // rows per block of the row-sum kernels

// Restart V.Q of one solve within a segmented launch (tiled bases, ld K).
struct SegVQ {
    const void *V;               // input basis (double, or float for FP32 bases)
    void *Out;                   // output basis (the other ping-pong buffer)
    float *Out32;                // FP32 copy of Out (FP64 bases with an FP32 copy), else null
    const double *Q;             // K x N, column-major, ld K
    double *r;                   // residual, updated in place
    Ctl *ctl;                    // rnorm / xscale for the next sweep
    double *partials;            // nb block partials of ||r||^2
    double qlast, beta;          // dsapps residual coefficients
    int n, K, N, res_col;        // res_col < 0: no beta term
    long long ldv;               // columns per tile of V, Out and Out32 (tiled)
    int nb, blk0;                // blocks (vq_seg_blocks(n)); first block in the launch
};
int vq_seg_blocks(int n);
// LDS of the MFMA restart kernel: Q^T of one solve, K rounded up to 4 by N
// rounded up to 16.
constexpr size_t kVqMfmaLds = 48 * 1024;
bool vq_seg_mfma_fits(int K, int N);
// T[0, nt) run on the FMA kernel (blocks numbered 0..blocks_total), then
// T[nt, nt + nt_mfma) on the MFMA kernel (numbered 0..blocks_mfma, at most
// mfma_nct 16-column tiles and mfma_lds bytes of Q^T each), then one
// finalize over all entries.
int launch_vq_seg(const SegVQ *T, int nt, int blocks_total, int max_K, int max_N, int nt_mfma, int blocks_mfma,
                  int mfma_nct, size_t mfma_lds, bool v32, hipStream_t s);
// Residual check of one returned Ritz pair within a segmented launch: G blocks
// write their partial sums of ||y - lam v||^2, v^T (y - lam v) and ||v||^2 to
// out[3 b .. 3 b + 2] (y = A v, both contiguous vectors of n rows); the host
// adds them in block order.
struct SegRes {
    const double *v, *y;
    double *out;
    double lam;
    int n, G, blk0;
};
constexpr int kResBlocksMax = 64; // blocks per checked pair
int res_blocks(int n);
int launch_resid_seg(const SegRes *T, int nt, int blocks_total, hipStream_t s);

struct Reduce;
// Restarts of bases wider than kNarrowCols run V.Q through launch_vq (K-chunked)
// and then these: the dsapps residual update r = qlast rscale r + beta
// Out(:, res_col) with ctl->rnorm := ||r|| (res_col < 0: no beta term), and the
// FP32 copy of the new basis's first N columns (tiled, ld columns per tile).
int launch_vq_residual(int n, const double *Out, long long ld, int res_col, double *r, double qlast, double beta,
                       Reduce &red, Ctl *ctl, hipStream_t s);
int launch_basis_copy32(int n, int N, long long ld, const double *src, float *dst, hipStream_t s);

// out = V(:, 0:K) q (tiled basis, ld K, FP64 or FP32; q on the device).
int launch_basis_combine(int n, int K, const void *V, long long ld, bool v32, const double *q, double *out,
                         hipStream_t s);
// Packs hsub, hdia (ncv each) and the Ctl words of every member at
// out + member * stride (stride >= 2 ncv + kCtlWords).
int launch_seg_readback(const SegMember *M, int nm, int stride, double *out, hipStream_t s);

// ------------------------------------------------------- basis kernels
struct Reduce {
    double *partials = nullptr;  // [nblocks_max * (ncv_max + 1)]
    unsigned *counter = nullptr; // one counter, self-resetting
    int max_blocks = 0;
    int max_cols = 0;
};
int reduce_alloc(Reduce &r, int max_blocks, int max_cols);
int cgs_blocks(int n);

// hl_cgs_pass_noalloc: the split pass of hl_cgs_pass (column-major FP64 V)
// with its block partials in a file-scope device array of kNoAllocScratch
// doubles (enough for kMaxReduceBlocks blocks of 129 partial sums); writes
// h_out / *nrm2_out directly; no allocation of any kind.
constexpr size_t kNoAllocScratch = (size_t)kMaxReduceBlocks * (128 + 1);
constexpr int kMaxDevices = 64;
int launch_cgs_noalloc(int dev, int n, int m, const double *V, long long ldv, const double *w_in, double *w_out,
                       const double *h_in, double *h_out, double *nrm2_out, hipStream_t s);

// One fused CGS pass over an FP64 basis (see hl_cgs_pass).  mode selects the
// epilogue; j is the Lanczos step for the alpha/beta bookkeeping.
int launch_cgs(int n, int m, const double *V, VLayout L, const double *w_in, double *w_out,
               const double *h_in, double *h_out, Reduce &red, Ctl *ctl, double *h_sub,
               double *h_dia, int mode, int j, hipStream_t s);

// Out(:, 0:N) = V(:, 0:K) Q(0:K, 0:N); optionally fused residual update on
// output column res_col: r = qlast * rscale * r + beta * Out(:, res_col) and
// ctl->rnorm := ||r|| (EPI_RNORM).  res_col < 0 disables it.
int launch_vq(int n, int K, int N, const double *V, VLayout Lin, const double *Q, int ldq,
              double *Out, VLayout Lout, int use_mfma, int res_col, double *r, double qlast,
              double beta, Reduce &red, Ctl *ctl, hipStream_t s);

// r = start vector (counter-based splitmix64, seed) or keep r; then ctl->rnorm.
int launch_start_vector(int n, uint64_t seed, double *r, hipStream_t s);
// v = r * ctl->xscale into a contiguous vector v (may be null) and the basis
// column col (host-operator path normalisation).
int launch_scale_copy(int n, const double *r, const Ctl *ctl, double *v, VColumn col, hipStream_t s);
// ||r||^2 -> epilogue (mode EPI_RNORM sets rnorm/xscale/rscale).
int launch_norm(int n, const double *r, Reduce &red, Ctl *ctl, int mode, hipStream_t s);
// HBM read probe: sums n doubles (n even, 16-byte aligned) into out[blocks*4].
int launch_stream_read(const double *a, long long n, double *out, int blocks, hipStream_t s);
// Gather probe: sums x[idx[i]] for i < m into out[blocks*4].
int launch_gather(const double *x, const int *idx, long long m, double *out, int blocks, hipStream_t s);

// ------------------------------------------------------------ profiling
// Profile classes: SpMV; CGS partials kernels of the two passes every step
// runs; V.Q; finalize and other small kernels; CGS partials kernels of the
// DGKS corrections (launched every step, exiting at once when not needed);
// host work of the engine between sweeps (one "launch" per sweep).
enum KClass : int { K_SPMV = 0, K_CGS = 1, K_VQ = 2, K_MISC = 3, K_DGKS = 4, K_HOST = 5, K_NCLASS = 6 };

// HIP-event timing of individual launches on one stream (profiled solves only).
struct Profiler {
    bool on = false;
    hl_profile tot{};
    struct Rec {
        int cls;
        double bytes;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> pool;
    std::vector<Rec> pending;
    hipEvent_t open_ev = nullptr;

    hipEvent_t take();
    int begin(hipStream_t s);                       // record a start event
    int end(hipStream_t s, int cls, double bytes);  // record the stop event
    int flush();                                    // after a stream sync
    void release();
};
hl_profile &thread_profile();

// ------------------------------------------------------------ host dense
// dlartg with MKL's sign convention (called at HLanc.hpp:1514,1536).
void dlartg(double f, double g, double &c, double &s, double &r);
// Symmetric tridiagonal eigen-decomposition by implicit-shift QL with
// Wilkinson shifts; ascending eigenvalues, eigenvectors in z (ldz >= n).
// Plays dsteqr("I") (HLanc.hpp:1691,1805).  Returns 0 or 1 (no convergence /
// non-finite input).
int tridiag_eig(int n, double *d, double *e, double *z, int ldz);
// Same eigenvalues, and only the last row of the eigenvector matrix
// (zlast[i] = Z(n-1, i)): what the restart loop needs (HLanc.hpp:1691-1696),
// in O(n^2); bitwise equal to the corresponding entries of tridiag_eig.
int tridiag_eig_last(int n, double *d, double *e, double *zlast);
// Householder QR of a (m x k, lda), LAPACK dgeqr2 storage (HLanc.hpp:1837),
// then Q = H(0)...H(k-1) formed explicitly as m x m (ldq).
void householder_q(int m, int k, double *a, int lda, double *tau, double *q, int ldq);

}  // namespace hl
