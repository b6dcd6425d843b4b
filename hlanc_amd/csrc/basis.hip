// basis.hip -- kernels over the Krylov basis V for CDNA4 (gfx950).
//
// Replaces mat_v_gpu_t's cuBLAS calls (HLanc.hpp:973-1119) and the host BLAS-1
// of dsaitr (HLanc.hpp:1411-1465):
//   * cgs_block: one fused classical Gram-Schmidt pass
//       w_out = w_in - V h_in ;  h_out = V^T w_out ;  ||w_out||^2
//     A wave covers a 16-row x m tile as 4 lane groups of 16 rows x ceil(m/4)
//     columns, so each load instruction touches four 128-byte column segments.
//     Blocks write their partial sums; a second kernel with one block per
//     column adds them in block order, and the ||w||^2 block runs the
//     Lanczos/DGKS bookkeeping (Ctl), so a whole dsaitr sweep needs no host
//     round trip.  Grid shape depends only on n: results are bitwise
//     repeatable, the same for a solve alone or beside others.  The segmented
//     kernels (cgs_seg_kernel, cgs_finalize_seg_kernel) run one pass of many
//     solves in one launch; cgs_kernel is the standalone form (hl_cgs_pass).
//   * cgs_block_f32 / coldot_block / dgks_decide_seg_kernel: the FP32-copy
//     basis mode (an FP32 copy of the FP64 basis for the projection-only passes):
//     the FP32 pass loads row pairs of adjacent column pairs (full 128-byte
//     lines per 16 lanes), the first pass is the Lanczos three-term projection
//     (one FP64 column), and a correction whose exact coefficients are
//     negligible is skipped.
//   * vq kernels: Out = V Q.  vq_seg_mfma_kernel (v_mfma_f64_16x16x4f64,
//     Out^T = Q^T V^T so stores are column-coalesced) and vq_seg_kernel (FMA)
//     restart every listed solve in one pass with the dsapps residual update
//     r = qlast r + beta Out(:, k) and ||r|| partials fused; vq_mfma_kernel and
//     vq_fma_kernel form the Ritz vectors and serve hl_vq.
#include <cmath>
#include <mutex>
#include <type_traits>

#include "hl_device.hpp"
#include "hl_internal.hpp"

namespace hl {
namespace {
using namespace dev;

template <typename T, bool NT, typename P>
__device__ __forceinline__ T ld_stream(P p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// Block `bid` of an `nb`-block CGS pass.  FL bit 0: non-temporal basis loads;
// bit 1: software prefetch of the next tile; bit 3: split (write this block's
// partial sums and leave the rest to cgs_finalize_kernel); bit 4: lean loads.
// Lean loads serve the solver's own bases (tiled FP64, a tile's leading
// dimension a multiple of 4, every column and padding row finite): a lane
// reads its ceil(m/4) columns of a tile as global loads at constant offsets
// from one address.  Only the last 3 columns of a lane group can lie past m
// (in the last group), so only their loads carry a per-lane mask.  Padding
// rows past n are read and multiplied by a = 0.
template <int MAXC, typename VT, bool TILED, int FL>
__device__ __forceinline__ void cgs_block(int bid, int nb, int n, int m, const VT *__restrict__ V_, long long ldv,
                                          const double *w_in_, double *w_out_, const double *__restrict__ h_in,
                                          double *h_out, double *partials_, unsigned *counter, Ctl *ctl,
                                          double *h_sub, double *h_dia, int mode, int j)
{
    if ((mode == EPI_PASS2 || mode == EPI_PASS3) && !ctl->need) return; // DGKS decided on the device
    const auto V = gptr(V_);
    const auto w_in = gptr(w_in_);
    const auto w_out = gptr(w_out_);
    const auto partials = gptr(partials_);
    __shared__ double hs[kGroups * MAXC];
    __shared__ double red[kBlock / kWave][kGroups * MAXC + 1];
    __shared__ double vals[kGroups * MAXC + 1];
    __shared__ double lds_tmp[kBlock];
    __shared__ double fin[kGroups * MAXC + 1];

    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int g = lane / kTileRows;
    const int rl = lane % kTileRows;
    const int cpg = (m + kGroups - 1) / kGroups;
    const int c0 = g * cpg;
    const bool has_h = h_in != nullptr;
    for (int c = threadIdx.x; c < kGroups * MAXC; c += kBlock) hs[c] = (has_h && c < m) ? h_in[c] : 0.0;
    __syncthreads();

    double acc[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) acc[c] = 0.0;
    double nrm = 0.0;
    const long long ntiles = (n + kTileRows - 1) / kTileRows;
    const long long W = (long long)nb * (kBlock / kWave);
    long long t = (long long)bid * (kBlock / kWave) + wv;

    // software pipeline: the next tile's V and w are in flight while this one is used
    double vn[MAXC], wn;
#define HL_CGS_FETCH(TT)                                                                         \
    {                                                                                            \
        const long long rr = (TT) * kTileRows + rl;                                              \
        const bool okk = (TT) < ntiles && rr < n;                                                \
        wn = okk ? w_in[rr] : 0.0;                                                               \
        _Pragma("unroll") for (int c = 0; c < MAXC; ++c)                                         \
        {                                                                                        \
            const int colj = c0 + c;                                                             \
            vn[c] = (okk && c < cpg && colj < m)                                                 \
                        ? static_cast<double>(ld_stream<VT, (FL & 1) != 0>(&V[vpos<TILED>(rr, colj, ldv)])) \
                        : 0.0;                                                                   \
        }                                                                                        \
    }
    constexpr bool PF = (FL & 2) != 0;
    constexpr bool LEAN = (FL & 16) != 0;
    static_assert(!LEAN || (TILED && !PF && std::is_same<VT, double>::value), "lean loads: tiled FP64, no prefetch");
#define HL_CGS_FETCH_LEAN(TT)                                                                    \
    {                                                                                            \
        const long long rr = (TT) * kTileRows + rl;                                              \
        wn = w_in[rr < n ? rr : n - 1];                                                          \
        const auto pv = V + ((TT) * ldv + c0) * kTileRows + rl;                                  \
        _Pragma("unroll") for (int c = 0; c < MAXC; ++c)                                         \
            vn[c] = c < cpg && (c + 3 < cpg || c0 + c < m)                                       \
                        ? ((FL & 1) ? __builtin_nontemporal_load(pv + c * kTileRows) : pv[c * kTileRows]) \
                        : 0.0;                                                                   \
    }
    if (PF) HL_CGS_FETCH(t);
    for (; t < ntiles; t += W) {
        if constexpr (LEAN)
            HL_CGS_FETCH_LEAN(t)
        else if (!PF)
            HL_CGS_FETCH(t);
        double v[MAXC];
#pragma unroll
        for (int c = 0; c < MAXC; ++c) v[c] = vn[c];
        const double win = wn;
        const long long r = t * kTileRows + rl;
        const bool ok = r < n;
        if (PF) HL_CGS_FETCH(t + W);
        double a = win;
        if (has_h) {
            double p = 0.0;
#pragma unroll
            for (int c = 0; c < MAXC; ++c) p += v[c] * hs[c0 + c];
            p += __shfl_xor(p, 16);
            p += __shfl_xor(p, 32);
            a = win - p;
        }
        if (LEAN && !ok) a = 0.0;
        if (g == 0 && ok) {
            if (w_out) w_out[r] = a;
            nrm += a * a;
        }
#pragma unroll
        for (int c = 0; c < MAXC; ++c) acc[c] += v[c] * a;
    }
#undef HL_CGS_FETCH
#undef HL_CGS_FETCH_LEAN

#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
        double s = acc[c];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        s += __shfl_xor(s, 8);
        acc[c] = s;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off);
    if (rl == 0) {
#pragma unroll
        for (int c = 0; c < MAXC; ++c)
            if (c < cpg && c0 + c < m) red[wv][c0 + c] = acc[c];
    }
    if (lane == 0) red[wv][m] = nrm;
    __syncthreads();
    const int M1 = m + 1;
    for (int c = threadIdx.x; c < M1; c += kBlock) {
        double s = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) s += red[w][c];
        vals[c] = s;
    }
    __syncthreads();
    if constexpr ((FL & 8) != 0) { // split: partials only, cgs_finalize_kernel finishes
        for (int c = threadIdx.x; c < M1; c += kBlock) partials[(long long)c * nb + bid] = vals[c];
        return;
    }
    if (!reduce_blocks(vals, M1, partials_, counter, lds_tmp, fin)) return;
    if (h_out)
        for (int c = threadIdx.x; c < m; c += kBlock) h_out[c] = fin[c];
    if (threadIdx.x == 0) run_epilogue(mode, j, fin[m], h_in, ctl, h_sub, h_dia);
}

typedef float flt2 __attribute__((ext_vector_type(2)));

// The same split pass over an FP32 tiled basis (always tiled and split; FL
// bit 0: non-temporal loads).  A lane holds two consecutive rows of a column
// (one 8-byte load); 16 lanes cover a pair of adjacent columns of the 16-row
// tile (one 128-byte line), so a wave covers the tile as 4 lane groups of
// ceil(m/8) column pairs and each load instruction moves 4 full lines, as
// the FP64 pass does.  Arithmetic is FP64.
template <int MAXC8, int FL, int TB>
__device__ __forceinline__ void cgs_block_f32(int bid, int nb, int n, int m, const float *__restrict__ V_,
                                              long long ldv, const double *w_in_, double *w_out_,
                                              const double *__restrict__ h_in, double *h_out, double *partials_)
{
    const auto V = gptr(V_);
    const auto w_in = gptr(w_in_);
    const auto w_out = gptr(w_out_);
    const auto partials = gptr(partials_);
    constexpr int G = 4;                 // lane groups per wave, 2 columns each per step
    constexpr bool LEAN = (FL & 16) != 0; // solver-owned basis: finite zero padding rows
    __shared__ double hs[2 * G * MAXC8];
    __shared__ double red[kBlock / kWave][2 * G * MAXC8 + 1];
    __shared__ double vals[2 * G * MAXC8 + 1];

    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int g = lane / 16;             // column group
    const int sc = (lane / 8) & 1;       // column of the pair
    const int q = lane % 8;              // row pair: rows 2q, 2q + 1 of the tile
    const int cpp = (m + 2 * G - 1) / (2 * G); // column pairs per group
    const int c0 = g * 2 * cpp + sc;     // this lane's columns: c0 + 2c
    const bool has_h = h_in != nullptr;
    for (int c = threadIdx.x; c < 2 * G * MAXC8; c += kBlock) hs[c] = (has_h && c < m) ? h_in[c] : 0.0;
    __syncthreads();

    double acc[MAXC8];
#pragma unroll
    for (int c = 0; c < MAXC8; ++c) acc[c] = 0.0;
    double nrm = 0.0;
    const long long ntiles = (n + kTileRows - 1) / kTileRows;
    const long long W = (long long)nb * (kBlock / kWave);
    // TB tiles per trip: all their loads are in flight before any is used
    for (long long t = (long long)bid * (kBlock / kWave) + wv; t < ntiles; t += TB * W) {
        flt2 v[TB][MAXC8]; // kept in FP32 registers, widened where used
        double w0[TB], w1[TB];
        bool ok0[TB], ok1[TB];
        long long r0[TB];
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            const long long tt = t + b * W;
            r0[b] = tt * kTileRows + 2 * q;
            // the tile's rows exist in the padded basis; rows >= n are masked to 0
            ok0[b] = tt < ntiles && r0[b] < n;
            ok1[b] = tt < ntiles && r0[b] + 1 < n;
            // one address per tile, the column pairs at constant offsets; only the
            // last 4 pairs of a lane group can lie past m.  Lean: padding rows
            // are read (finite, multiplied by a = 0)
            const auto pv = reinterpret_cast<gmem<const flt2> *>(
                V + ((tt < ntiles ? tt : 0) * ldv + c0) * kTileRows + 2 * q);
#pragma unroll
            for (int c = 0; c < MAXC8; ++c) {
                v[b][c] = flt2{0.0f, 0.0f};
                if ((LEAN ? tt < ntiles : ok0[b]) && c < cpp && (c + 4 < cpp || c0 + 2 * c < m))
                    v[b][c] = ld_stream<flt2, (FL & 1) != 0>(pv + c * kTileRows);
            }
            w0[b] = ok0[b] ? w_in[r0[b]] : 0.0;
            w1[b] = ok1[b] ? w_in[r0[b] + 1] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            if (!LEAN && !ok1[b])
#pragma unroll
                for (int c = 0; c < MAXC8; ++c) v[b][c].y = 0.0f;
            double a0 = w0[b], a1 = w1[b];
            if (has_h) {
                double p0 = 0.0, p1 = 0.0;
#pragma unroll
                for (int c = 0; c < MAXC8; ++c) {
                    p0 += static_cast<double>(v[b][c].x) * hs[c0 + 2 * c];
                    p1 += static_cast<double>(v[b][c].y) * hs[c0 + 2 * c];
                }
                p0 += __shfl_xor(p0, 8);
                p1 += __shfl_xor(p1, 8);
                p0 += __shfl_xor(p0, 16);
                p1 += __shfl_xor(p1, 16);
                p0 += __shfl_xor(p0, 32);
                p1 += __shfl_xor(p1, 32);
                a0 -= p0;
                a1 -= p1;
            }
            if (LEAN) {
                a0 = ok0[b] ? a0 : 0.0;
                a1 = ok1[b] ? a1 : 0.0;
            }
            if (lane < 8) {
                if (ok0[b]) {
                    if (w_out) w_out[r0[b]] = a0;
                    nrm += a0 * a0;
                }
                if (ok1[b]) {
                    if (w_out) w_out[r0[b] + 1] = a1;
                    nrm += a1 * a1;
                }
            }
#pragma unroll
            for (int c = 0; c < MAXC8; ++c)
                acc[c] += static_cast<double>(v[b][c].x) * a0 + static_cast<double>(v[b][c].y) * a1;
        }
    }

#pragma unroll
    for (int c = 0; c < MAXC8; ++c) {
        double sm = acc[c];
        sm += __shfl_xor(sm, 1);
        sm += __shfl_xor(sm, 2);
        sm += __shfl_xor(sm, 4);
        acc[c] = sm;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off);
    if (q == 0) {
#pragma unroll
        for (int c = 0; c < MAXC8; ++c)
            if (c < cpp && c0 + 2 * c < m) red[wv][c0 + 2 * c] = acc[c];
    }
    if (lane == 0) red[wv][m] = nrm;
    __syncthreads();
    const int M1 = m + 1;
    for (int c = threadIdx.x; c < M1; c += kBlock) {
        double sm = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) sm += red[w][c];
        vals[c] = sm;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < M1; c += kBlock) partials[(long long)c * nb + bid] = vals[c];
}

// Block partials of alpha = v^T w and ||w||^2 for one column v of a tiled
// basis (v points at its first tile entry; ld columns per tile row):
// every lane active, row-coalesced.  Written as partials[c * nb + bid],
// c = 0 (alpha) and 1 (norm), the layout the finalize reads for m = 1.
template <typename VT>
__device__ __forceinline__ void coldot_block(int bid, int nb, int n, const VT *__restrict__ v_, long long ldv,
                                             const double *__restrict__ w_, double *partials_)
{
    const auto v = gptr(v_);
    const auto w = gptr(w_);
    const auto partials = gptr(partials_);
    __shared__ double red[2][kBlock / kWave];
    double a = 0.0, s2 = 0.0;
    for (long long r = (long long)bid * kBlock + threadIdx.x; r < n; r += (long long)nb * kBlock) {
        const double wr = w[r];
        a += static_cast<double>(__builtin_nontemporal_load(&v[vpos<true>(r, 0, ldv)])) * wr;
        s2 += wr * wr;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        s2 += __shfl_xor(s2, off);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[0][threadIdx.x / kWave] = a;
        red[1][threadIdx.x / kWave] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) t += red[threadIdx.x][k];
        partials[(long long)threadIdx.x * nb + bid] = t;
    }
}

template <int MAXC, typename VT, bool TILED, int FL>
__global__ __launch_bounds__(kBlock) void cgs_kernel(int n, int m, const VT *__restrict__ V, long long ldv,
                                                     const double *w_in, double *w_out,
                                                     const double *__restrict__ h_in, double *h_out,
                                                     double *partials, unsigned *counter, Ctl *ctl,
                                                     double *h_sub, double *h_dia, int mode, int j)
{
    cgs_block<MAXC, VT, TILED, FL>(blockIdx.x, gridDim.x, n, m, V, ldv, w_in, w_out, h_in, h_out, partials, counter,
                                   ctl, h_sub, h_dia, mode, j);
}

// Second kernel of a split CGS pass: block c adds column c of the column-major
// block partials in block order; block m (the squared norm) runs the epilogue.
// Columns whose block starts after the epilogue cleared ctl->need are not
// written, and no later kernel reads them.
// Column c of the column-major block partials added in block order (a fixed
// tree over the block's lanes); the sum is valid in thread 0.
__device__ __forceinline__ double column_sum(int c, int nb, const double *__restrict__ partials_)
{
    const auto partials = gptr(partials_);
    __shared__ double sred[kBlock];
    const int tid = threadIdx.x;
    double t = 0.0;
    for (int b = tid; b < nb; b += kBlock) t += partials[(long long)c * nb + b];
    sred[tid] = t;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) sred[tid] += sred[tid + w];
        __syncthreads();
    }
    return sred[0];
}

__device__ __forceinline__ void cgs_finalize_block(int c, int nb, int m, const double *__restrict__ partials_,
                                                   const double *h_in, double *h_out, Ctl *ctl, double *h_sub,
                                                   double *h_dia, int mode, int j)
{
    if ((mode == EPI_PASS2 || mode == EPI_PASS3) && !ctl->need) return;
    const double sum = column_sum(c, nb, partials_);
    if (threadIdx.x != 0) return;
    if (c < m) {
        if (h_out) h_out[c] = sum;
    } else {
        run_epilogue(mode, j, sum, h_in, ctl, h_sub, h_dia);
    }
}

// Finalize of a standalone split pass writing straight to the caller's
// outputs: h_out[c] for c < m, *nrm2_out for c == m (either may be null).
__global__ __launch_bounds__(kBlock) void cgs_finalize_out_kernel(int nb, int m, const double *__restrict__ partials,
                                                                  double *h_out, double *nrm2_out)
{
    const int c = blockIdx.x;
    const double sum = column_sum(c, nb, partials);
    if (threadIdx.x != 0) return;
    if (c < m) {
        if (h_out) h_out[c] = sum;
    } else if (nrm2_out) {
        *nrm2_out = sum;
    }
}

__global__ __launch_bounds__(kBlock) void cgs_finalize_kernel(int nb, int m, const double *__restrict__ partials,
                                                              const double *h_in, double *h_out, Ctl *ctl,
                                                              double *h_sub, double *h_dia, int mode, int j)
{
    cgs_finalize_block(blockIdx.x, nb, m, partials, h_in, h_out, ctl, h_sub, h_dia, mode, j);
}

// Largest i < ns with first[i * stride] <= b: the segment owning block b
// (segments are listed in launch order; empty ones are skipped over).
__device__ __forceinline__ int seg_find(const int *first, int stride, int ns, int b)
{
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid * stride] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Operands of CGS pass `mode` of a Lanczos step: the first pass projects w
// (h0 = V^T w), the second forms r = w - V h0 (h1 = V^T r), the DGKS
// corrections update r in place with h1, then h0.
struct PassPtrs {
    const double *w_in;
    double *w_out;
    const double *h_in;
    double *h_out;
};
__device__ __forceinline__ PassPtrs pass_ptrs(const SegMember &mb, int mode)
{
    switch (mode) {
    case EPI_WNORM:
        return {mb.w, nullptr, nullptr, mb.h0};
    case EPI_PASS1:
        return {mb.w, mb.r, mb.h0, mb.h1};
    case EPI_PASS2:
        return {mb.r, mb.r, mb.h1, mb.h0};
    default:
        return {mb.r, mb.r, mb.h0, nullptr};
    }
}

// One CGS pass of every listed solve in a single launch (split form).  The
// pass (MODE) is a template argument so that profiles tell the always-run
// passes (EPI_WNORM, EPI_PASS1) from the DGKS corrections.
// One CGS pass of the listed steps.  FP64 launches read mb.V, FP32 launches
// mb.Vproj (the FP32 basis, or the FP32 copy of an FP64 basis).  `sub`
// selects the block numbering of the launch (SegStep::cgs0s, else cgs0).
// With an FP32 copy, a step's first correction runs on the copy unless the
// step cancelled strongly (ctl->exact): then an FP64 launch of its own runs it.
template <int MAXC, int MODE, typename VT, int TB = 2>
__global__ __launch_bounds__(kBlock) void cgs_seg_kernel(const SegMember *__restrict__ M,
                                                         const SegStep *__restrict__ S, int ns, int sub, int solo)
{
    constexpr bool F32 = std::is_same<VT, float>::value;
    const int i = seg_find(sub ? &S[0].cgs0s : &S[0].cgs0, sizeof(SegStep) / sizeof(int), ns, blockIdx.x);
    const SegStep st = S[i];
    const SegMember mb = M[solo ? 0 : st.mem];
    if (MODE == EPI_PASS2 && mb.V32 && (mb.ctl->exact != 0) == F32) return; // the other launch's step
    const PassPtrs p = pass_ptrs(mb, MODE);
    const int bid = blockIdx.x - (sub ? st.cgs0s : st.cgs0);
    if constexpr (F32) {
        if ((MODE == EPI_PASS2 || MODE == EPI_PASS3) && !mb.ctl->need) return; // DGKS decided on the device
        const float *V = static_cast<const float *>(mb.Vproj);
        if (MODE == EPI_WNORM && mb.V32) { // FP32 copy: only alpha_j = v_j^T w, on the FP64 column
            coldot_block(bid, mb.nb, mb.n, static_cast<const double *>(mb.V) + 16 * st.j, mb.vld, p.w_in,
                         mb.partials);
            return;
        }
        cgs_block_f32<(MAXC + 1) / 2, 17, TB>(bid, mb.nb, mb.n, st.j + 1, V, mb.vld, p.w_in, p.w_out, p.h_in,
                                             p.h_out, mb.partials);
    } else {
        cgs_block<MAXC, double, true, 25>(bid, mb.nb, mb.n, st.j + 1, static_cast<const double *>(mb.V), mb.vld,
                                         p.w_in, p.w_out, p.h_in, p.h_out, mb.partials, nullptr, mb.ctl, mb.hsub,
                                         mb.hdia, MODE, st.j);
    }
}

// Wide passes (m > kNarrowCols, FP64 tiled solver bases): the fused pass
// split in two launches over the same block numbering.  The subtracting kernel
// forms a = w_in - V h_in for the block's tiles (V read in 128-column chunks),
// writes w_out and the block's partial of ||a||^2 (column m of the partials);
// the projecting kernel then reads a back and forms the partials of V^T a one
// 128-column chunk after another.  Tile assignment, lane mapping, load masks and
// reduction order are those of cgs_block, so a pass of m <= 128 columns run
// this way gives the fused pass's sums.
template <int MODE>
__global__ __launch_bounds__(kBlock) void cgs_seg_sub_kernel(const SegMember *__restrict__ M,
                                                             const SegStep *__restrict__ S, int ns, int sub, int solo)
{
    const int i = seg_find(sub ? &S[0].cgs0s : &S[0].cgs0, sizeof(SegStep) / sizeof(int), ns, blockIdx.x);
    const SegStep st = S[i];
    const SegMember mb = M[solo ? 0 : st.mem];
    if (MODE == EPI_PASS2 && mb.V32 && mb.ctl->exact == 0) return; // the FP32 launch's step
    if ((MODE == EPI_PASS2 || MODE == EPI_PASS3) && !mb.ctl->need) return;
    const PassPtrs p = pass_ptrs(mb, MODE);
    const int bid = blockIdx.x - (sub ? st.cgs0s : st.cgs0);
    const int n = mb.n, m = st.j + 1, nb = mb.nb;
    const long long ldv = mb.vld;
    const auto V = gptr(static_cast<const double *>(mb.V));
    const auto w_in = gptr(p.w_in);
    const auto w_out = gptr(p.w_out);
    const auto partials = gptr(mb.partials);
    __shared__ double hs[kMaxNcv + 4 * 32];
    __shared__ double red[kBlock / kWave];
    const bool has_h = p.h_in != nullptr;
    for (int c = threadIdx.x; c < kMaxNcv + 4 * 32; c += kBlock) hs[c] = (has_h && c < m) ? p.h_in[c] : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int g = lane / kTileRows;
    const int rl = lane % kTileRows;
    const long long ntiles = (n + kTileRows - 1) / kTileRows;
    const long long W = (long long)nb * (kBlock / kWave);
    double nrm = 0.0;
    for (long long t = (long long)bid * (kBlock / kWave) + wv; t < ntiles; t += W) {
        const long long r = t * kTileRows + rl;
        const bool ok = r < n;
        const double win = w_in[ok ? r : n - 1];
        double a = win;
        if (has_h) {
            double pr = 0.0;
            for (int k0 = 0; k0 < m; k0 += kNarrowCols) {
                const int mk = m - k0 < kNarrowCols ? m - k0 : kNarrowCols;
                const int cpg = (mk + kGroups - 1) / kGroups;
                const int c0 = k0 + g * cpg;
                const auto pv = V + (t * ldv + c0) * kTileRows + rl;
                double v[32];
#pragma unroll
                for (int c = 0; c < 32; ++c)
                    v[c] = c < cpg && (c + 3 < cpg || c0 + c < k0 + mk) ? __builtin_nontemporal_load(pv + c * kTileRows)
                                                                       : 0.0;
#pragma unroll
                for (int c = 0; c < 32; ++c) pr += v[c] * hs[c0 + c];
            }
            pr += __shfl_xor(pr, 16);
            pr += __shfl_xor(pr, 32);
            a = win - pr;
        }
        if (!ok) a = 0.0;
        if (g == 0 && ok) {
            if (p.w_out) w_out[r] = a;
            nrm += a * a;
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off);
    if (lane == 0) red[wv] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s2 = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) s2 += red[w];
        partials[(long long)m * nb + bid] = s2;
    }
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void cgs_seg_proj_kernel(const SegMember *__restrict__ M,
                                                              const SegStep *__restrict__ S, int ns, int sub, int solo)
{
    const int i = seg_find(sub ? &S[0].cgs0s : &S[0].cgs0, sizeof(SegStep) / sizeof(int), ns, blockIdx.x);
    const SegStep st = S[i];
    const SegMember mb = M[solo ? 0 : st.mem];
    if (MODE == EPI_PASS2 && mb.V32 && mb.ctl->exact == 0) return;
    if ((MODE == EPI_PASS2 || MODE == EPI_PASS3) && !mb.ctl->need) return;
    const PassPtrs p = pass_ptrs(mb, MODE);
    if (!p.h_out) return; // the pass projects nothing
    const int bid = blockIdx.x - (sub ? st.cgs0s : st.cgs0);
    const int n = mb.n, m = st.j + 1, nb = mb.nb;
    const long long ldv = mb.vld;
    const auto V = gptr(static_cast<const double *>(mb.V));
    const auto src = gptr(p.w_out ? p.w_out : p.w_in); // the pass's a, written by the subtracting kernel
    const auto partials = gptr(mb.partials);
    __shared__ double red[kBlock / kWave][kNarrowCols];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int g = lane / kTileRows;
    const int rl = lane % kTileRows;
    const long long ntiles = (n + kTileRows - 1) / kTileRows;
    const long long W = (long long)nb * (kBlock / kWave);
    for (int k0 = 0; k0 < m; k0 += kNarrowCols) {
        const int mk = m - k0 < kNarrowCols ? m - k0 : kNarrowCols;
        const int cpg = (mk + kGroups - 1) / kGroups;
        const int c0 = g * cpg; // within the chunk
        double acc[32];
#pragma unroll
        for (int c = 0; c < 32; ++c) acc[c] = 0.0;
        for (long long t = (long long)bid * (kBlock / kWave) + wv; t < ntiles; t += W) {
            const long long r = t * kTileRows + rl;
            const bool ok = r < n;
            const double a = ok ? src[r] : 0.0;
            const auto pv = V + (t * ldv + k0 + c0) * kTileRows + rl;
            double v[32];
#pragma unroll
            for (int c = 0; c < 32; ++c)
                v[c] = c < cpg && (c + 3 < cpg || c0 + c < mk) ? __builtin_nontemporal_load(pv + c * kTileRows) : 0.0;
#pragma unroll
            for (int c = 0; c < 32; ++c) acc[c] += v[c] * a;
        }
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            double sv = acc[c];
            sv += __shfl_xor(sv, 1);
            sv += __shfl_xor(sv, 2);
            sv += __shfl_xor(sv, 4);
            sv += __shfl_xor(sv, 8);
            acc[c] = sv;
        }
        __syncthreads(); // the previous chunk's readers of red are done
        if (rl == 0) {
#pragma unroll
            for (int c = 0; c < 32; ++c)
                if (c < cpg && c0 + c < mk) red[wv][c0 + c] = acc[c];
        }
        __syncthreads();
        for (int c = threadIdx.x; c < mk; c += kBlock) {
            double sv = 0.0;
            for (int w = 0; w < kBlock / kWave; ++w) sv += red[w][c];
            partials[(long long)(k0 + c) * nb + bid] = sv;
        }
    }
}

__global__ __launch_bounds__(kBlock) void cgs_finalize_seg_kernel(const SegMember *__restrict__ M,
                                                                  const SegStep *__restrict__ S, int ns, int mode,
                                                                  int solo)
{
    const int i = seg_find(&S[0].fin0, sizeof(SegStep) / sizeof(int), ns, blockIdx.x);
    const SegStep st = S[i];
    const SegMember mb = M[solo ? 0 : st.mem];
    const PassPtrs p = pass_ptrs(mb, mode);
    const int c = blockIdx.x - st.fin0;
    if (mode == EPI_WNORM && mb.V32) {
        // FP32 copy: the first projection is the Lanczos three-term one,
        // h0 = (0, .., 0, beta_j, alpha_j) with beta_j = ||r|| of the last
        // step (HLanc.hpp:1431-1438 keeps it as h_sub[j]); the subtracting
        // pass's exact FP64 projection catches whatever else w holds along V
        if (c > 1) return;
        if (c == 1) {
            for (int x = threadIdx.x; x + 1 < st.j; x += kBlock) p.h_out[x] = 0.0;
            if (threadIdx.x == 0 && st.j > 0) p.h_out[st.j - 1] = mb.ctl->rnorm;
        }
        cgs_finalize_block(c, mb.nb, 1, mb.partials, p.h_in, p.h_out + st.j, mb.ctl, mb.hsub, mb.hdia, mode, st.j);
        return;
    }
    cgs_finalize_block(c, mb.nb, st.j + 1, mb.partials, p.h_in, p.h_out, mb.ctl, mb.hsub, mb.hdia, mode, st.j);
}

// After the subtracting pass of FP32-copy solves (their first correction is
// pending after every such pass): the pass's exact FP64 projection h1 = V^T r is
// what the correction would remove.  When ||h1|| <= tau ||r|| the correction
// would move r by at most tau ||r||, so it is skipped (ctl->need = 0).  One
// block per listed step.
__global__ __launch_bounds__(kBlock) void dgks_decide_seg_kernel(const SegMember *__restrict__ M,
                                                                 const SegStep *__restrict__ S, double tau2, int solo)
{
    const SegStep st = S[blockIdx.x];
    const SegMember mb = M[solo ? 0 : st.mem];
    if (!mb.V32 || tau2 <= 0.0) return;
    __shared__ double red[kBlock / kWave];
    double s2 = 0.0;
    for (int c = threadIdx.x; c <= st.j; c += kBlock) s2 += mb.h1[c] * mb.h1[c];
    for (int off = kWave / 2; off > 0; off >>= 1) s2 += __shfl_xor(s2, off);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = s2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += red[w];
        Ctl *ctl = mb.ctl;
        if (ctl->need && t <= tau2 * ctl->rnorm * ctl->rnorm) {
            ctl->need = 0;
            ctl->nskip += 1;
        }
    }
}

typedef double dv2 __attribute__((ext_vector_type(2)));

// Read-only streaming probe: every block sums a contiguous share with 16-byte
// loads; measures the HBM read rate the box reaches.
__global__ __launch_bounds__(kBlock) void stream_read_kernel(const dv2 *__restrict__ a, long long n2,
                                                             double *__restrict__ out)
{
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n2; i += (long long)gridDim.x * kBlock) {
        const dv2 v = __builtin_nontemporal_load(&a[i]);
        s += v.x + v.y;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & (kWave - 1)) == 0) out[blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave] = s;
}

// Gather probe: sums x[idx[i]] over m indices, eight index loads in flight
// per thread and then their eight gathers (the SpMV's access pattern without
// the matrix values).
__global__ __launch_bounds__(kBlock) void gather_kernel(const double *__restrict__ x, const int *__restrict__ idx,
                                                        long long m, double *__restrict__ out)
{
    constexpr int kPer = 8;
    double s = 0.0;
    const long long stride = (long long)gridDim.x * kBlock * kPer;
    for (long long base = (long long)blockIdx.x * kBlock * kPer + threadIdx.x; base < m; base += stride) {
        int ci[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const long long i = base + (long long)k * kBlock;
            ci[k] = i < m ? __builtin_nontemporal_load(&idx[i]) : -1;
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (ci[k] >= 0) s += x[ci[k]];
    }
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & (kWave - 1)) == 0) out[blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave] = s;
}

// Out(:, c0:c0+CB) = V(:, 0:K) Q(0:K, c0:c0+CB); thread per row, FP64 FMA.
template <int CB, typename VT, bool TIN, bool TOUT>
__global__ __launch_bounds__(kBlock) void vq_fma_kernel(int n, int K, int N, const VT *__restrict__ V,
                                                        long long ldv, const double *__restrict__ Q, int ldq,
                                                        VT *__restrict__ Out, long long ldo, int c0,
                                                        int res_local, double *r, double qlast,
                                                        double beta, double *partials, unsigned *counter,
                                                        Ctl *ctl, int accum)
{
    extern __shared__ double qs[]; // K x CB
    __shared__ double red[kBlock / kWave];
    __shared__ double vals[2];
    for (int i = threadIdx.x; i < K * CB; i += kBlock) {
        const int k = i / CB, c = i % CB;
        qs[i] = (c0 + c < N) ? Q[(long long)(c0 + c) * ldq + k] : 0.0;
    }
    __syncthreads();
    const double qrs = (res_local >= 0) ? qlast * ctl->rscale : 0.0;
    double nrm = 0.0;
    for (long long row = (long long)blockIdx.x * kBlock + threadIdx.x; row < n;
         row += (long long)gridDim.x * kBlock) {
        double acc[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c)
            acc[c] = (accum && c0 + c < N) ? static_cast<double>(Out[vpos<TOUT>(row, c0 + c, ldo)]) : 0.0;
        for (int k = 0; k < K; ++k) {
            const double v = static_cast<double>(V[vpos<TIN>(row, k, ldv)]);
#pragma unroll
            for (int c = 0; c < CB; ++c) acc[c] += v * qs[k * CB + c];
        }
        double rk = 0.0;
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            if (c0 + c < N) Out[vpos<TOUT>(row, c0 + c, ldo)] = static_cast<VT>(acc[c]);
            if (c == res_local) rk = acc[c];
        }
        if (res_local >= 0) { // dsapps residual (HLanc.hpp:1605-1614)
            const double t = qrs * r[row];
            const double rr = t + beta * rk;
            r[row] = rr;
            nrm += rr * rr;
        }
    }
    if (res_local < 0) return;
    for (int off = kWave / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) s += red[w];
        vals[0] = s;
    }
    __syncthreads();
    // block partial of ||r||^2; cgs_finalize_kernel (one block, m = 0) adds them
    if (threadIdx.x == 0) partials[blockIdx.x] = vals[0];
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// The restart V.Q of every listed solve in one launch: Out = V Q(:, 0:N) on
// tiled bases (ld K), then the dsapps residual update
// r = Q(ncv-1, nev-1) rscale r + beta Out(:, res_col) (res_col < 0: beta term
// absent) and each block's partial ||r||^2 in partials[block].
template <int CB, typename VT>
__global__ __launch_bounds__(kBlock) void vq_seg_kernel(const SegVQ *__restrict__ T, int nt)
{
    extern __shared__ double qs[]; // K x CB
    __shared__ double red[kBlock / kWave];
    const int i = seg_find(&T[0].blk0, sizeof(SegVQ) / sizeof(int), nt, blockIdx.x);
    const SegVQ t = T[i];
    const VT *__restrict__ V = static_cast<const VT *>(t.V);
    VT *__restrict__ Out = static_cast<VT *>(t.Out);
    const int bid = blockIdx.x - t.blk0;
    const double qrs = t.qlast * t.ctl->rscale;
    double nrm = 0.0;
    for (int c0 = 0; c0 < t.N; c0 += CB) {
        __syncthreads(); // the previous chunk's readers of qs are done
        for (int x = threadIdx.x; x < t.K * CB; x += kBlock) {
            const int k = x / CB, c = x % CB;
            qs[x] = (c0 + c < t.N) ? t.Q[(long long)(c0 + c) * t.K + k] : 0.0;
        }
        __syncthreads();
        const int rl = (t.res_col >= c0 && t.res_col < c0 + CB) ? t.res_col - c0 : -1;
        const bool upd = t.res_col >= 0 ? rl >= 0 : c0 == 0;
        for (long long row = (long long)bid * kBlock + threadIdx.x; row < t.n; row += (long long)t.nb * kBlock) {
            double acc[CB];
#pragma unroll
            for (int c = 0; c < CB; ++c) acc[c] = 0.0;
            for (int k = 0; k < t.K; ++k) {
                const double v = static_cast<double>(V[vpos<true>(row, k, t.ldv)]);
#pragma unroll
                for (int c = 0; c < CB; ++c) acc[c] += v * qs[k * CB + c];
            }
            double rk = 0.0;
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                if (c0 + c < t.N) {
                    Out[vpos<true>(row, c0 + c, t.ldv)] = static_cast<VT>(acc[c]);
                    if (t.Out32) t.Out32[vpos<true>(row, c0 + c, t.ldv)] = static_cast<float>(acc[c]);
                }
                if (c == rl) rk = acc[c];
            }
            if (upd) { // dsapps residual (HLanc.hpp:1605-1614)
                const double rr = qrs * t.r[row] + t.beta * rk;
                t.r[row] = rr;
                nrm += rr * rr;
            }
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) s += red[w];
        t.partials[bid] = s;
    }
}

// The 16-row tiles of this wave for a solve of exactly NCT output column
// tiles: Out = V Q and the fused dsapps residual update; returns the wave's
// partial of ||r||^2.  An exact tile count keeps the accumulators in place
// (a per-tile guard made the compiler copy them between register files).
template <int NCT, typename VT>
__device__ __forceinline__ double vq_mfma_rows(const SegVQ &t, const double *qs, int NC, int Kp, int bid, int wv,
                                               int lane, double qrs)
{
    const auto V = gptr(static_cast<const VT *>(t.V));
    const auto Out = gptr(static_cast<VT *>(t.Out));
    const auto Out32 = gptr(t.Out32);
    const auto r = gptr(t.r);
    const int lr = lane & 15, lk = lane >> 4;
    double nrm = 0.0;
    const long long ntiles = (t.n + 15) / 16;
    for (long long tile = (long long)bid * (kBlock / kWave) + wv; tile < ntiles;
         tile += (long long)t.nb * (kBlock / kWave)) {
        const long long row = tile * 16 + lr;
        const bool ok = row < t.n;
        dbl4 acc[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
        // 8 loads of V in flight per wave before their 8 x NCT MFMAs
        for (int k0 = 0; k0 < Kp; k0 += 32) {
            double b[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + 4 * j + lk;
                b[j] = (ok && k < t.K) ? static_cast<double>(V[vpos<true>(row, k, t.ldv)]) : 0.0;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (k0 + 4 * j >= Kp) break;
                const int k = k0 + 4 * j + lk;
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct)
                    acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(qs[k * NC + ct * 16 + lr], b[j], acc[ct], 0, 0, 0);
            }
        }
        double rk = 0.0;
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int cl = ct * 16 + lk + 4 * e;
                if (ok && cl < t.N) {
                    Out[vpos<true>(row, cl, t.ldv)] = static_cast<VT>(acc[ct][e]);
                    if (Out32) Out32[vpos<true>(row, cl, t.ldv)] = static_cast<float>(acc[ct][e]);
                }
                if (cl == t.res_col) rk = acc[ct][e];
            }
        }
        // the lanes holding column res_col of this tile hand it to lanes 0..15
        const double rkv = t.res_col >= 0 ? __shfl(rk, ((t.res_col % 16) % 4) * 16 + lr) : 0.0;
        if (lk == 0 && ok) { // dsapps residual (HLanc.hpp:1605-1614)
            const double rr = qrs * r[row] + t.beta * rkv;
            r[row] = rr;
            nrm += rr * rr;
        }
    }
    return nrm;
}

// One block's share of the restart V.Q of one listed solve: up to NT 16-column
// output tiles; Q^T sits in LDS.  Used for the solves whose K x N fits
// (vq_seg_mfma_fits).
template <int NT, typename VT>
__global__ __launch_bounds__(kBlock) void vq_seg_mfma_kernel(const SegVQ *__restrict__ T, int nt)
{
    extern __shared__ double qs[]; // Kp x NC, row k
    __shared__ double red[kBlock / kWave];
    const int i = seg_find(&T[0].blk0, sizeof(SegVQ) / sizeof(int), nt, blockIdx.x);
    const SegVQ t = T[i];
    const int bid = blockIdx.x - t.blk0;
    const int nct = (t.N + 15) / 16; // this solve's column tiles, <= NT
    const int NC = nct * 16;
    const int Kp = (t.K + 3) & ~3;
    for (int x = threadIdx.x; x < Kp * NC; x += kBlock) {
        const int k = x / NC, c = x % NC;
        qs[x] = (k < t.K && c < t.N) ? t.Q[(long long)c * t.K + k] : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const double qrs = t.qlast * t.ctl->rscale;
    double nrm = 0.0;
    switch (nct) { // the exact tile count (wave-uniform)
#define HL_VQ_NCT(C)                                                                      \
    case C:                                                                               \
        if constexpr (C <= NT) nrm = vq_mfma_rows<C, VT>(t, qs, NC, Kp, bid, wv, lane, qrs); \
        break;
        HL_VQ_NCT(1)
        HL_VQ_NCT(2)
        HL_VQ_NCT(3)
        HL_VQ_NCT(4)
        HL_VQ_NCT(5)
        HL_VQ_NCT(6)
        HL_VQ_NCT(7)
        HL_VQ_NCT(8)
#undef HL_VQ_NCT
    default:
        break;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off);
    if (lane == 0) red[wv] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) sum += red[w];
        t.partials[bid] = sum;
    }
}

// out = V(:, 0:K) q on a tiled basis (ld columns per tile): one FP64 vector
// from a basis of either precision (the start vector of a mixed-precision
// solve's FP64 phase).
template <typename VT>
__global__ __launch_bounds__(kBlock) void basis_combine_kernel(int n, int K, const VT *__restrict__ V, long long ld,
                                                               const double *__restrict__ q, double *__restrict__ out)
{
    __shared__ double qs[128];
    for (int k = threadIdx.x; k < K; k += kBlock) qs[k] = q[k];
    __syncthreads();
    for (long long row = (long long)blockIdx.x * kBlock + threadIdx.x; row < n; row += (long long)gridDim.x * kBlock) {
        double acc = 0.0;
        for (int k = 0; k < K; ++k) acc += static_cast<double>(V[vpos<true>(row, k, ld)]) * qs[k];
        out[row] = acc;
    }
}

// Residual check after the Ritz-vector step: block b of pair T[i] sums, over
// its rows, d = y - lam v into ||d||^2, v^T d and ||v||^2 (FP64; y = A v from
// the FP64 SpMV) and writes the three partials.  v^T d is formed directly, so
// the Rayleigh quotient lam + v^T d / v^T v and the residual at it,
// ||d||^2 - (v^T d)^2 / v^T v, carry no cancellation of the large terms.
__global__ __launch_bounds__(kBlock) void resid_seg_kernel(const SegRes *__restrict__ T, int nt)
{
    __shared__ double red[3][kBlock / kWave];
    const int i = seg_find(&T[0].blk0, sizeof(SegRes) / sizeof(int), nt, blockIdx.x);
    const SegRes t = T[i];
    const int b = blockIdx.x - t.blk0;
    const auto v = gptr(t.v);
    const auto y = gptr(t.y);
    double a = 0.0, p = 0.0, c = 0.0;
    for (long long r = (long long)b * kBlock + threadIdx.x; r < t.n; r += (long long)t.G * kBlock) {
        const double vr = v[r];
        const double d = y[r] - t.lam * vr;
        a += d * d;
        p += vr * d;
        c += vr * vr;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        p += __shfl_xor(p, off);
        c += __shfl_xor(c, off);
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[0][w] = a;
        red[1][w] = p;
        red[2][w] = c;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) s += red[threadIdx.x][k];
        t.out[3 * b + threadIdx.x] = s;
    }
}

// The dsapps residual update of a restart whose V.Q ran unfused (bases wider
// than the segmented restart kernels take): r = qlast rscale r + beta Out(:, k)
// (res_col < 0: no beta term) on a tiled Out, and each block's partial ||r||^2.
__global__ __launch_bounds__(kBlock) void vq_resid_kernel(int n, const double *__restrict__ Out, long long ld,
                                                          int res_col, double *__restrict__ r, double qlast,
                                                          double beta, const Ctl *ctl, double *partials)
{
    __shared__ double red[kBlock / kWave];
    const double qrs = qlast * ctl->rscale;
    double nrm = 0.0;
    for (long long row = (long long)blockIdx.x * kBlock + threadIdx.x; row < n; row += (long long)gridDim.x * kBlock) {
        const double rk = res_col >= 0 ? Out[vpos<true>(row, res_col, ld)] : 0.0;
        const double rr = qrs * r[row] + beta * rk; // HLanc.hpp:1605-1614
        r[row] = rr;
        nrm += rr * rr;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += red[w];
        partials[blockIdx.x] = t;
    }
}

// dst(:, 0:N) = FP32 copy of src(:, 0:N), both tiled with ld columns per tile.
__global__ __launch_bounds__(kBlock) void basis_copy32_kernel(int n, int N, long long ld, const double *__restrict__ src,
                                                              float *__restrict__ dst)
{
    const long long total = ((long long)n + 15) / 16 * ld * 16; // whole tiles, padding rows included
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < total; i += (long long)gridDim.x * kBlock) {
        const long long c = (i >> 4) % ld;
        if (c < N) dst[i] = static_cast<float>(src[i]);
    }
}

// One block per listed solve: rnorm = sqrt(sum of its V.Q block partials).
__global__ __launch_bounds__(kBlock) void vq_seg_finalize_kernel(const SegVQ *__restrict__ T)
{
    const SegVQ t = T[blockIdx.x];
    cgs_finalize_block(0, t.nb, 0, t.partials, nullptr, nullptr, t.ctl, nullptr, nullptr, EPI_RNORM, 0);
}

// One block per member: hsub, hdia (ncv each) and the Ctl words, packed at
// out + member * stride, so that one copy brings every member's state home.
__global__ __launch_bounds__(kBlock) void seg_readback_kernel(const SegMember *__restrict__ M, int stride,
                                                              double *__restrict__ out)
{
    const SegMember mb = M[blockIdx.x];
    double *o = out + (long long)blockIdx.x * stride;
    for (int x = threadIdx.x; x < mb.ncv; x += kBlock) {
        o[x] = mb.hsub[x];
        o[mb.ncv + x] = mb.hdia[x];
    }
    const double *c = reinterpret_cast<const double *>(mb.ctl);
    for (int x = threadIdx.x; x < kCtlWords; x += kBlock) o[2 * mb.ncv + x] = c[x];
}


// Out^T(c, row) = sum_k Q(k, c) V(row, k) on v_mfma_f64_16x16x4f64.
// A (16 x 4): lane l holds Q^T(ct*16 + l%16, k0 + l/16); B (4 x 16): lane l holds
// V(r0 + l%16, k0 + l/16); D (16 x 16): lane l, element i is D(l/16 + 4 i, l%16),
// i.e. Out(r0 + l%16, ct*16 + l/16 + 4 i) (layout checked on the device by
// tests/test_basis_gpu.py).  NT column tiles of 16 per launch.
template <int NT, bool TIN, bool TOUT>
__global__ __launch_bounds__(kBlock) void vq_mfma_kernel(int n, int K, int N, const double *__restrict__ V,
                                                         long long ldv, const double *__restrict__ Q, int ldq,
                                                         double *__restrict__ Out, long long ldo, int c0,
                                                         int res_local, double *r, double qlast, double beta,
                                                         double *partials, unsigned *counter, Ctl *ctl, int accum)
{
    extern __shared__ double qs[]; // Kp x (NT*16), row k
    __shared__ double red[kBlock / kWave];
    __shared__ double vals[2];
    const int NC = NT * 16;
    const int Kp = (K + 3) & ~3;
    for (int i = threadIdx.x; i < Kp * NC; i += kBlock) {
        const int k = i / NC, c = i % NC;
        qs[i] = (k < K && c0 + c < N) ? Q[(long long)(c0 + c) * ldq + k] : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int lr = lane & 15, lk = lane >> 4;
    const double qrs = (res_local >= 0) ? qlast * ctl->rscale : 0.0;
    double nrm = 0.0;
    const long long ntiles = (n + 15) / 16;
    for (long long t = (long long)blockIdx.x * (kBlock / kWave) + wv; t < ntiles;
         t += (long long)gridDim.x * (kBlock / kWave)) {
        const long long r0 = t * 16;
        const long long row = r0 + lr;
        const bool ok = row < n;
        dbl4 acc[NT];
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
            acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
            if (accum) // K-chunked product: this launch adds to what the previous chunks wrote
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int cl = ct * 16 + lk + 4 * i;
                    if (ok && c0 + cl < N) acc[ct][i] = Out[vpos<TOUT>(row, c0 + cl, ldo)];
                }
        }
        for (int k0 = 0; k0 < Kp; k0 += 32) { // 8 loads of V in flight per wave
            double b[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + 4 * j + lk;
                b[j] = (ok && k < K) ? V[vpos<TIN>(row, k, ldv)] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (k0 + 4 * j >= Kp) break;
                const int k = k0 + 4 * j + lk;
#pragma unroll
                for (int ct = 0; ct < NT; ++ct)
                    acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(qs[k * NC + ct * 16 + lr], b[j], acc[ct], 0, 0, 0);
            }
        }
        double rk = 0.0;
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int cl = ct * 16 + lk + 4 * i;
                if (ok && c0 + cl < N) Out[vpos<TOUT>(row, c0 + cl, ldo)] = acc[ct][i];
                if (cl == res_local) rk = acc[ct][i];
            }
        }
        if (res_local >= 0) { // the lane holding column res_local for this row shares it
            const int owner = ((res_local % 16) % 4) * 16 + lr;
            const double rkv = __shfl(rk, owner);
            if (lk == 0 && ok) {
                const double tt = qrs * r[row];
                const double rr = tt + beta * rkv;
                r[row] = rr;
                nrm += rr * rr;
            }
        }
    }
    if (res_local < 0) return;
    for (int off = kWave / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off);
    if (lane == 0) red[wv] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) s += red[w];
        vals[0] = s;
    }
    __syncthreads();
    // block partial of ||r||^2; cgs_finalize_kernel (one block, m = 0) adds them
    if (threadIdx.x == 0) partials[blockIdx.x] = vals[0];
}

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void start_vector_kernel(int n, uint64_t base, double *r)
{
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        const uint64_t z = mix64(base + 0x9E3779B97F4A7C15ULL * (uint64_t)(i + 1));
        r[i] = 2.0 * (static_cast<double>(z >> 11) * 0x1.0p-53) - 1.0;
    }
}

// v = r * xscale into a contiguous vector and/or a basis column.
template <typename VT>
__global__ __launch_bounds__(kBlock) void scale_copy_kernel(int n, const double *r, const Ctl *ctl, double *v,
                                                            VT *col, long long ld, int j, int tiled)
{
    const double s = ctl->xscale;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        const double x = r[i] * s;
        if (v) v[i] = x;
        if (col) col[tiled ? vpos<true>(i, j, ld) : vpos<false>(i, j, ld)] = static_cast<VT>(x);
    }
}

int grid_for(long long n, int per_block, int cap)
{
    long long g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<int>(g);
}

template <typename VT, bool TILED>
int launch_cgs_t(int n, int m, const VT *V, long long ldv, const double *w_in, double *w_out,
                 const double *h_in, double *h_out, Reduce &red, Ctl *ctl, double *h_sub, double *h_dia,
                 int mode, int j, hipStream_t s)
{
    if (m < 0 || m > 4 * 32 || m > red.max_cols) {
        set_error("CGS pass: m=%d outside [0, %d]", m, red.max_cols < 128 ? red.max_cols : 128);
        return HL_ERR_ARG;
    }
    const int nb = cgs_blocks(n);
    if (nb > red.max_blocks) {
        set_error("CGS pass: reduction workspace too small");
        return HL_ERR_STATE;
    }
    const int cpg = (m + kGroups - 1) / kGroups;
    const int fl = (tuning().cgs_flags & 8) ? 9 : (tuning().cgs_flags & 3);
#define HL_CGS1(MC, F)                                                                                         \
    cgs_kernel<MC, VT, TILED, F><<<nb, kBlock, 0, s>>>(n, m, V, ldv, w_in, w_out, h_in, h_out, red.partials, \
                                                       red.counter, ctl, h_sub, h_dia, mode, j)
#define HL_CGS(MC)              \
    if (fl == 0)                \
        HL_CGS1(MC, 0);         \
    else if (fl == 1)           \
        HL_CGS1(MC, 1);         \
    else if (fl == 2)           \
        HL_CGS1(MC, 2);         \
    else if (fl == 3)           \
        HL_CGS1(MC, 3);         \
    else                        \
        HL_CGS1(MC, 9)
    if (cpg <= 2) {
        HL_CGS(2);
    } else if (cpg <= 4) {
        HL_CGS(4);
    } else if (cpg <= 8) {
        HL_CGS(8);
    } else if (cpg <= 16) {
        HL_CGS(16);
    } else if (cpg <= 24) {
        HL_CGS(24);
    } else {
        HL_CGS(32);
    }
#undef HL_CGS
#undef HL_CGS1
    HL_TRY_HIP(hipGetLastError());
    if (fl == 9) {
        cgs_finalize_kernel<<<m + 1, kBlock, 0, s>>>(nb, m, red.partials, h_in, h_out, ctl, h_sub, h_dia, mode, j);
        HL_TRY_HIP(hipGetLastError());
    }
    return HL_OK;
}

template <typename VT, bool TIN, bool TOUT>
int launch_vq_fma(int n, int K, int N, const VT *V, long long ldv, const double *Q, int ldq, VT *Out,
                  long long ldo, int res_col, double *r, double qlast, double beta, Reduce &red, Ctl *ctl,
                  hipStream_t s, int accum)
{
    constexpr int CB = 32;
    const int nb = grid_for(n, kBlock * 4, 1024);
    if (res_col >= 0 && nb > red.max_blocks) {
        set_error("V.Q: reduction workspace too small");
        return HL_ERR_STATE;
    }
    for (int c0 = 0; c0 < N; c0 += CB) {
        const int rl = (res_col >= c0 && res_col < c0 + CB) ? res_col - c0 : -1;
        const size_t lds = sizeof(double) * (size_t)K * CB;
        vq_fma_kernel<CB, VT, TIN, TOUT><<<nb, kBlock, lds, s>>>(n, K, N, V, ldv, Q, ldq, Out, ldo, c0, rl, r,
                                                                  qlast, beta, red.partials, red.counter, ctl,
                                                                  accum);
        HL_TRY_HIP(hipGetLastError());
    }
    if (res_col >= 0) {
        cgs_finalize_kernel<<<1, kBlock, 0, s>>>(nb, 0, red.partials, nullptr, nullptr, ctl, nullptr, nullptr,
                                                 EPI_RNORM, 0);
        HL_TRY_HIP(hipGetLastError());
    }
    return HL_OK;
}

template <bool TIN, bool TOUT>
int launch_vq_mfma(int n, int K, int N, const double *V, long long ldv, const double *Q, int ldq, double *Out,
                   long long ldo, int res_col, double *r, double qlast, double beta, Reduce &red, Ctl *ctl,
                   hipStream_t s, int accum)
{
    const long long ntiles = (static_cast<long long>(n) + 15) / 16;
    const int nb = grid_for(ntiles, 4 * 4, 1024);
    if (res_col >= 0 && nb > red.max_blocks) {
        set_error("V.Q: reduction workspace too small");
        return HL_ERR_STATE;
    }
    const int Kp = (K + 3) & ~3;
    // column tiles per launch: Q's K x 16*nt slice must fit 48 KB of LDS
    int max_nt = 6144 / (Kp * 16);
    max_nt = max_nt < 1 ? 1 : (max_nt > 4 ? 4 : max_nt);
    const int chunk = 16 * max_nt;
    for (int c0 = 0; c0 < N; c0 += chunk) {
        const int width = N - c0 < chunk ? N - c0 : chunk;
        const int nt = (width + 15) / 16;
        const int rl = (res_col >= c0 && res_col < c0 + chunk) ? res_col - c0 : -1;
        const size_t lds = sizeof(double) * (size_t)Kp * nt * 16;
#define HL_VQM(NTV)                                                                                          \
    vq_mfma_kernel<NTV, TIN, TOUT><<<nb, kBlock, lds, s>>>(n, K, N, V, ldv, Q, ldq, Out, ldo, c0, rl, r, qlast, \
                                                           beta, red.partials, red.counter, ctl, accum)
        if (nt == 1)
            HL_VQM(1);
        else if (nt == 2)
            HL_VQM(2);
        else if (nt == 3)
            HL_VQM(3);
        else
            HL_VQM(4);
#undef HL_VQM
        HL_TRY_HIP(hipGetLastError());
    }
    if (res_col >= 0) {
        cgs_finalize_kernel<<<1, kBlock, 0, s>>>(nb, 0, red.partials, nullptr, nullptr, ctl, nullptr, nullptr,
                                                 EPI_RNORM, 0);
        HL_TRY_HIP(hipGetLastError());
    }
    return HL_OK;
}

}  // namespace

int cgs_blocks(int n)
{
    const long long ntiles = (static_cast<long long>(n) + kTileRows - 1) / kTileRows;
    return grid_for(ntiles, 16, tuning().cgs_blocks_max);
}

Tuning &tuning()
{
    static Tuning t;
    return t;
}

int reduce_alloc(Reduce &r, int max_blocks, int max_cols)
{
    r.max_blocks = max_blocks;
    r.max_cols = max_cols;
    HL_TRY_HIP(hipMalloc(&r.partials, sizeof(double) * (size_t)max_blocks * (max_cols + 1)));
    HL_TRY_HIP(hipMalloc(&r.counter, sizeof(unsigned) * 4));
    HL_TRY_HIP(hipMemset(r.counter, 0, sizeof(unsigned) * 4));
    return HL_OK;
}

int launch_cgs(int n, int m, const double *V, VLayout L, const double *w_in, double *w_out, const double *h_in,
               double *h_out, Reduce &red, Ctl *ctl, double *h_sub, double *h_dia, int mode, int j, hipStream_t s)
{
    if (L.tiled)
        return launch_cgs_t<double, true>(n, m, V, L.ld, w_in, w_out, h_in, h_out, red, ctl, h_sub, h_dia, mode, j, s);
    return launch_cgs_t<double, false>(n, m, V, L.ld, w_in, w_out, h_in, h_out, red, ctl, h_sub, h_dia, mode, j, s);
}

// ------------------------------------------------- allocation-free CGS pass
// The block partials of hl_cgs_pass_noalloc live in this file-scope device
// array: every GPU the library runs on holds its own copy until the process
// ends, and every call on one GPU shares it (calls that overlap on one GPU
// race on it).  Host code never passes the array's name to a kernel -- on the
// host it is only a stand-in for the symbol; the device address comes from
// hipGetSymbolAddress, looked up once per device with that device current.
static __device__ __align__(256) double scratch[kNoAllocScratch];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device dev (current on entry).
int scratch_address(int dev, double **out)
{
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("no-allocation CGS pass: device %d beyond %d", dev, kMaxDevices);
        return HL_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_status(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<double *>(g_scratch_ptr[dev]);
    return HL_OK;
}
} // namespace

int launch_cgs_noalloc(int dev, int n, int m, const double *V, long long ldv, const double *w_in, double *w_out,
                       const double *h_in, double *h_out, double *nrm2_out, hipStream_t s)
{
    if (m < 0 || m > 4 * 32) {
        set_error("no-allocation CGS pass: m=%d outside [0, 128]", m);
        return HL_ERR_ARG;
    }
    const int nb = cgs_blocks(n);
    if ((size_t)nb * (size_t)(m + 1) > kNoAllocScratch) {
        set_error("no-allocation CGS pass: %d blocks x %d columns exceed the %zu-double scratch", nb, m + 1,
                  kNoAllocScratch);
        return HL_ERR_ARG;
    }
    double *part = nullptr;
    int rc = scratch_address(dev, &part);
    if (rc) return rc;
    // the split pass of hl_cgs_pass (cgs_kernel with FL = 9: non-temporal loads,
    // partials only), then its per-column sums into the caller's outputs
    const int cpg = (m + kGroups - 1) / kGroups;
#define HL_CGSN(MC)                                                                                          \
    cgs_kernel<MC, double, false, 9><<<nb, kBlock, 0, s>>>(n, m, V, ldv, w_in, w_out, h_in, nullptr, part,   \
                                                           nullptr, nullptr, nullptr, nullptr, EPI_NONE, 0)
    if (cpg <= 2)
        HL_CGSN(2);
    else if (cpg <= 4)
        HL_CGSN(4);
    else if (cpg <= 8)
        HL_CGSN(8);
    else if (cpg <= 16)
        HL_CGSN(16);
    else if (cpg <= 24)
        HL_CGSN(24);
    else
        HL_CGSN(32);
#undef HL_CGSN
    HL_TRY_HIP(hipGetLastError());
    if (h_out || nrm2_out) {
        cgs_finalize_out_kernel<<<m + 1, kBlock, 0, s>>>(nb, m, part, h_out, nrm2_out);
        HL_TRY_HIP(hipGetLastError());
    }
    return HL_OK;
}

int launch_norm(int n, const double *r, Reduce &red, Ctl *ctl, int mode, hipStream_t s)
{
    return launch_cgs_t<double, false>(n, 0, static_cast<const double *>(nullptr), 0, r, nullptr, nullptr, nullptr,
                                       red, ctl, nullptr, nullptr, mode, 0, s);
}

int launch_vq(int n, int K, int N, const double *V, VLayout Lin, const double *Q, int ldq, double *Out,
              VLayout Lout, int use_mfma, int res_col, double *r, double qlast, double beta, Reduce &red, Ctl *ctl,
              hipStream_t s)
{
    if (K <= 0 || N <= 0 || K > kMaxNcv) {
        set_error("V.Q: K=%d N=%d unsupported (K <= %d)", K, N, kMaxNcv);
        return HL_ERR_ARG;
    }
    // rows of Q per pass: the FMA kernel keeps a K x 32 slice of Q in 48 KB of LDS,
    // the MFMA kernel a K x 16 one; wider K runs in chunks, each pass adding its
    // chunk's product to what the previous passes wrote
    const int kpass = use_mfma ? 384 : 192;
    if (K > kpass && res_col >= 0) {
        set_error("V.Q: the fused residual update needs K <= %d (K = %d)", kpass, K);
        return HL_ERR_ARG;
    }
    for (int k0 = 0; k0 < K; k0 += kpass) {
        const int kc = K - k0 < kpass ? K - k0 : kpass;
        const double *Vk = V + (Lin.tiled ? 16LL * k0 : (long long)k0 * Lin.ld); // column k0 in either layout
        const double *Qk = Q + k0;
        const int acc = k0 > 0;
        int rc;
#define HL_VQ_DISPATCH(TI, TO)                                                                                      \
    rc = use_mfma ? launch_vq_mfma<TI, TO>(n, kc, N, Vk, Lin.ld, Qk, ldq, Out, Lout.ld, res_col, r, qlast, beta, \
                                           red, ctl, s, acc)                                                      \
                  : launch_vq_fma<double, TI, TO>(n, kc, N, Vk, Lin.ld, Qk, ldq, Out, Lout.ld, res_col, r, qlast, \
                                                  beta, red, ctl, s, acc)
        if (Lin.tiled && Lout.tiled)
            HL_VQ_DISPATCH(true, true);
        else if (Lin.tiled)
            HL_VQ_DISPATCH(true, false);
        else if (Lout.tiled)
            HL_VQ_DISPATCH(false, true);
        else
            HL_VQ_DISPATCH(false, false);
#undef HL_VQ_DISPATCH
        if (rc) return rc;
    }
    return HL_OK;
}

int launch_start_vector(int n, uint64_t seed, double *r, hipStream_t s)
{
    const uint64_t base = mix64(seed ^ 0x6A09E667F3BCC909ULL);
    start_vector_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(n, base, r);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_scale_copy(int n, const double *r, const Ctl *ctl, double *v, VColumn col, hipStream_t s)
{
    scale_copy_kernel<double><<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(n, r, ctl, v, col.base, col.L.ld, col.j,
                                                                            col.L.tiled ? 1 : 0);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

namespace {
template <int MODE, typename VT, int TB = 2>
void cgs_seg_dispatch(const SegMember *M, const SegStep *S, int ns, int max_cpg, int blocks, int sub, int solo,
                      hipStream_t s)
{
#define HL_SEG(MC) cgs_seg_kernel<MC, MODE, VT, TB><<<blocks, kBlock, 0, s>>>(M, S, ns, sub, solo)
    if (max_cpg <= 2)
        HL_SEG(2);
    else if (max_cpg <= 4)
        HL_SEG(4);
    else if (max_cpg <= 8)
        HL_SEG(8);
    else if (max_cpg <= 16)
        HL_SEG(16);
    else if (max_cpg <= 24)
        HL_SEG(24);
    else
        HL_SEG(32);
#undef HL_SEG
}
} // namespace

int launch_cgs_seg(const SegMember *M, const SegStep *S, int ns, int max_cpg, int cgs_blocks_total, int mode,
                   bool v32, bool sub, bool solo, hipStream_t s)
{
    if (ns <= 0) return HL_OK;
    if (max_cpg > kNarrowCols / kGroups) { // a step wider than the fused pass holds: two kernels for all
        if (v32 || max_cpg > kMaxNcv / kGroups) {
            set_error("CGS pass: %d columns per lane group unsupported%s", max_cpg, v32 ? " on an FP32 basis" : "");
            return HL_ERR_ARG;
        }
#define HL_WIDE(MODE)                                                                         \
    cgs_seg_sub_kernel<MODE><<<cgs_blocks_total, kBlock, 0, s>>>(M, S, ns, sub ? 1 : 0, solo ? 1 : 0); \
    cgs_seg_proj_kernel<MODE><<<cgs_blocks_total, kBlock, 0, s>>>(M, S, ns, sub ? 1 : 0, solo ? 1 : 0)
        switch (mode) {
        case EPI_WNORM:
            HL_WIDE(EPI_WNORM);
            break;
        case EPI_PASS1:
            HL_WIDE(EPI_PASS1);
            break;
        case EPI_PASS2:
            HL_WIDE(EPI_PASS2);
            break;
        case EPI_PASS3:
            HL_WIDE(EPI_PASS3);
            break;
        default:
            set_error("CGS pass: bad mode %d", mode);
            return HL_ERR_ARG;
        }
#undef HL_WIDE
        HL_TRY_HIP(hipGetLastError());
        return HL_OK;
    }
#define HL_CGS_MODE(MODE)                                                                              \
    (!v32 ? cgs_seg_dispatch<MODE, double>(M, S, ns, max_cpg, cgs_blocks_total, sub, solo, s)           \
     : tuning().cgs_f32_tiles == 1                                                                    \
         ? cgs_seg_dispatch<MODE, float, 1>(M, S, ns, max_cpg, cgs_blocks_total, sub, solo, s)        \
         : cgs_seg_dispatch<MODE, float, 2>(M, S, ns, max_cpg, cgs_blocks_total, sub, solo, s))
    switch (mode) {
    case EPI_WNORM:
        HL_CGS_MODE(EPI_WNORM);
        break;
    case EPI_PASS1:
        HL_CGS_MODE(EPI_PASS1);
        break;
    case EPI_PASS2:
        HL_CGS_MODE(EPI_PASS2);
        break;
    case EPI_PASS3:
        HL_CGS_MODE(EPI_PASS3);
        break;
    default:
        set_error("CGS pass: bad mode %d", mode);
        return HL_ERR_ARG;
    }
#undef HL_CGS_MODE
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_dgks_decide_seg(const SegMember *M, const SegStep *S, int ns, double tau, bool solo, hipStream_t s)
{
    if (ns <= 0 || tau <= 0.0) return HL_OK;
    dgks_decide_seg_kernel<<<ns, kBlock, 0, s>>>(M, S, tau * tau, solo ? 1 : 0);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_cgs_seg_finalize(const SegMember *M, const SegStep *S, int ns, int fin_blocks_total, int mode,
                            bool solo, hipStream_t s)
{
    if (ns <= 0) return HL_OK;
    cgs_finalize_seg_kernel<<<fin_blocks_total, kBlock, 0, s>>>(M, S, ns, mode, solo ? 1 : 0);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int vq_seg_blocks(int n) { return grid_for(n, kBlock * 4, 1024); }

bool vq_seg_mfma_fits(int K, int N)
{
    const int Kp = (K + 3) & ~3, nct = (N + 15) / 16;
    return K > 0 && N > 0 && nct <= 8 && (size_t)Kp * nct * 16 * sizeof(double) <= kVqMfmaLds;
}

namespace {
template <typename VT>
int launch_vq_seg_t(const SegVQ *T, int nt, int blocks_total, int max_K, int max_N, int nt_mfma, int blocks_mfma,
                    int mfma_nct, size_t mfma_lds, hipStream_t s)
{
    if (nt + nt_mfma <= 0) return HL_OK;
    if (max_K > 128) {
        set_error("V.Q: K=%d unsupported (K <= 128)", max_K);
        return HL_ERR_ARG;
    }
    if (nt_mfma > 0) { // T[nt, nt + nt_mfma): their own block numbering from 0
        if (mfma_nct < 1 || mfma_nct > 8 || mfma_lds > kVqMfmaLds) {
            set_error("V.Q: %d column tiles / %zu B of Q unsupported on MFMA", mfma_nct, mfma_lds);
            return HL_ERR_ARG;
        }
        const SegVQ *Tm = T + nt;
#define HL_VQM2(NT) vq_seg_mfma_kernel<NT, VT><<<blocks_mfma, kBlock, mfma_lds, s>>>(Tm, nt_mfma)
        switch (mfma_nct) {
        case 1: HL_VQM2(1); break;
        case 2: HL_VQM2(2); break;
        case 3: HL_VQM2(3); break;
        case 4: HL_VQM2(4); break;
        case 5: HL_VQM2(5); break;
        case 6: HL_VQM2(6); break;
        case 7: HL_VQM2(7); break;
        default: HL_VQM2(8); break;
        }
#undef HL_VQM2
        HL_TRY_HIP(hipGetLastError());
    }
    if (nt <= 0) {
        vq_seg_finalize_kernel<<<nt_mfma, kBlock, 0, s>>>(T);
        HL_TRY_HIP(hipGetLastError());
        return HL_OK;
    }
    // one chunk of output columns when Q's K x CB slice fits 48 KB of LDS: V is
    // then read once (the accumulation order over k does not depend on CB)
#define HL_VQS(CB) vq_seg_kernel<CB, VT><<<blocks_total, kBlock, sizeof(double) * (size_t)max_K * (CB), s>>>(T, nt)
    if (max_N <= 8)
        HL_VQS(8);
    else if (max_N <= 16)
        HL_VQS(16);
    else if (max_N <= 32 || max_K * 48 > 6144)
        HL_VQS(32);
    else if (max_N <= 48 || max_K * 64 > 6144)
        HL_VQS(48);
    else
        HL_VQS(64);
#undef HL_VQS
    HL_TRY_HIP(hipGetLastError());
    vq_seg_finalize_kernel<<<nt + nt_mfma, kBlock, 0, s>>>(T);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}
} // namespace

int launch_vq_seg(const SegVQ *T, int nt, int blocks_total, int max_K, int max_N, int nt_mfma, int blocks_mfma,
                  int mfma_nct, size_t mfma_lds, bool v32, hipStream_t s)
{
    return v32 ? launch_vq_seg_t<float>(T, nt, blocks_total, max_K, max_N, nt_mfma, blocks_mfma, mfma_nct, mfma_lds, s)
               : launch_vq_seg_t<double>(T, nt, blocks_total, max_K, max_N, nt_mfma, blocks_mfma, mfma_nct, mfma_lds,
                                         s);
}

int launch_basis_combine(int n, int K, const void *V, long long ld, bool v32, const double *q, double *out,
                         hipStream_t s)
{
    if (n <= 0) return HL_OK;
    if (K <= 0 || K > 128) {
        set_error("basis combination: K=%d unsupported (1..128)", K);
        return HL_ERR_ARG;
    }
    const int blocks = grid_for(n, kBlock, 2048);
    if (v32)
        basis_combine_kernel<float><<<blocks, kBlock, 0, s>>>(n, K, static_cast<const float *>(V), ld, q, out);
    else
        basis_combine_kernel<double><<<blocks, kBlock, 0, s>>>(n, K, static_cast<const double *>(V), ld, q, out);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_vq_residual(int n, const double *Out, long long ld, int res_col, double *r, double qlast, double beta,
                       Reduce &red, Ctl *ctl, hipStream_t s)
{
    const int nb = vq_seg_blocks(n);
    if (nb > red.max_blocks) {
        set_error("V.Q residual: reduction workspace too small");
        return HL_ERR_STATE;
    }
    vq_resid_kernel<<<nb, kBlock, 0, s>>>(n, Out, ld, res_col, r, qlast, beta, ctl, red.partials);
    HL_TRY_HIP(hipGetLastError());
    cgs_finalize_kernel<<<1, kBlock, 0, s>>>(nb, 0, red.partials, nullptr, nullptr, ctl, nullptr, nullptr, EPI_RNORM,
                                             0);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_basis_copy32(int n, int N, long long ld, const double *src, float *dst, hipStream_t s)
{
    if (n <= 0 || N <= 0) return HL_OK;
    basis_copy32_kernel<<<grid_for((long long)n * ld, kBlock, 4096), kBlock, 0, s>>>(n, N, ld, src, dst);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int res_blocks(int n) { return grid_for(n, kBlock * 16, kResBlocksMax); }

int launch_resid_seg(const SegRes *T, int nt, int blocks_total, hipStream_t s)
{
    if (nt <= 0) return HL_OK;
    resid_seg_kernel<<<blocks_total, kBlock, 0, s>>>(T, nt);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_seg_readback(const SegMember *M, int nm, int stride, double *out, hipStream_t s)
{
    if (nm <= 0) return HL_OK;
    seg_readback_kernel<<<nm, kBlock, 0, s>>>(M, stride, out);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_gather(const double *x, const int *idx, long long m, double *out, int blocks, hipStream_t s)
{
    gather_kernel<<<blocks, kBlock, 0, s>>>(x, idx, m, out);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_stream_read(const double *a, long long n, double *out, int blocks, hipStream_t s)
{
    stream_read_kernel<<<blocks, kBlock, 0, s>>>(reinterpret_cast<const dv2 *>(a), n / 2, out);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

}  // namespace hl
