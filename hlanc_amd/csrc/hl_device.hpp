// hl_device.hpp -- device helpers shared by the kernels of libhlanc_hip.so:
// the Lanczos/DGKS epilogue run by the last block of a reducing kernel, and the
// deterministic two-stage block reduction.
#pragma once

#include "hl_internal.hpp"

namespace hl {
namespace dev {

// A pointer into global memory, typed so: loads and stores through it compile
// to global (not flat) instructions.  Pointers that a kernel reads from a
// struct in memory are generic otherwise.
template <typename T> using gmem = __attribute__((address_space(1))) T;
template <typename T> __device__ __forceinline__ gmem<const T> *gptr(const T *p) { return (gmem<const T> *)p; }
template <typename T> __device__ __forceinline__ gmem<T> *gptr(T *p) { return (gmem<T> *)p; }

constexpr int kGroups = 4;                 // column groups per wave
constexpr int kTileRows = kWave / kGroups; // 16 rows per wave tile

// Field by field (a whole-struct copy from registers is lowered through scratch).
__device__ __forceinline__ void store_ctl(Ctl *d, const Ctl &c)
{
    static_assert(sizeof(Ctl) == 120, "store_ctl lists every field of Ctl");
    d->rnorm = c.rnorm;
    d->wnorm = c.wnorm;
    d->xscale = c.xscale;
    d->rscale = c.rscale;
    d->need = c.need;
    d->iter = c.iter;
    d->cgs = c.cgs;
    d->nrm2 = c.nrm2;
    d->p2 = c.p2;
    d->p3 = c.p3;
    d->m2sum = c.m2sum;
    d->m3sum = c.m3sum;
    d->dgks_always = c.dgks_always;
    d->exact = c.exact;
    d->nexact = c.nexact;
    d->mxsum = c.mxsum;
    d->nskip = c.nskip;
}

// The bookkeeping after pass `mode` of step j reduced its squared norm nrm2:
// hin_j is entry j of the pass's input projection, hsub_j / hdia_j point at the
// step's tridiagonal entries.
__device__ __forceinline__ void run_epilogue_at(int mode, int j, double nrm2, double hin_j, Ctl *ctl,
                                                double *hsub_j, double *hdia_j)
{
    switch (mode) {
    case EPI_NONE:
        ctl->nrm2 = nrm2;
        break;
    case EPI_WNORM:
        ctl->wnorm = sqrt(nrm2);
        break;
    case EPI_PASS1: // HLanc.hpp:1431-1440
        *hdia_j = hin_j;
        *hsub_j = (j == 0) ? 0.0 : ctl->rnorm;
        ctl->rnorm = sqrt(nrm2);
        if (ctl->dgks_always) {
            ctl->need = 1;
            // a strongly cancelling step, or one wider than the FP32 passes take,
            // corrects on the FP64 basis
            ctl->exact = ctl->rnorm < kProjExact * ctl->wnorm || j + 1 > kNarrowCols;
            if (ctl->exact) {
                ctl->nexact += 1;
                ctl->mxsum += j + 1;
            }
        } else {
            ctl->need = ctl->rnorm < kDgks * ctl->wnorm;
        }
        ctl->iter = 0;
        ctl->cgs += 1;
        break;
    case EPI_PASS2: { // HLanc.hpp:1444-1458, first correction
        ctl->p2 += 1;
        ctl->m2sum += j + 1;
        *hdia_j += hin_j;
        const double r1 = sqrt(nrm2);
        if (r1 > kDgks * ctl->rnorm) {
            ctl->rnorm = r1;
            ctl->need = 0;
        } else {
            ctl->rnorm = r1;
            ctl->iter = 1;
            ctl->need = r1 < kDgks * ctl->wnorm;
        }
        ctl->cgs += 1;
        break;
    }
    case EPI_PASS3: { // second correction; failure zeroes resid (HLanc.hpp:1459-1463)
        ctl->p3 += 1;
        ctl->m3sum += j + 1;
        *hdia_j += hin_j;
        const double r1 = sqrt(nrm2);
        if (r1 > kDgks * ctl->rnorm) {
            ctl->rnorm = r1;
        } else {
            ctl->rnorm = 0.0;
            ctl->rscale = 0.0;
        }
        ctl->need = 0;
        ctl->cgs += 1;
        break;
    }
    case EPI_RNORM:
        ctl->rnorm = sqrt(nrm2);
        ctl->rscale = 1.0;
        break;
    default:
        break;
    }
    if (mode >= EPI_PASS1) // STEP 2's scale for v_{j+1} (HLanc.hpp:1412-1419)
        ctl->xscale = (ctl->rnorm > kSafmin ? 1.0 / ctl->rnorm : 1.0) * ctl->rscale;
}

__device__ __forceinline__ void run_epilogue(int mode, int j, double nrm2, const double *h_in, Ctl *ctl,
                                             double *h_sub, double *h_dia)
{
    const bool steps = mode >= EPI_PASS1 && mode <= EPI_PASS3; // the modes that touch h and T
    run_epilogue_at(mode, j, nrm2, steps ? h_in[j] : 0.0, ctl, steps ? h_sub + j : nullptr,
                    steps ? h_dia + j : nullptr);
}

// Block partials of M1 values (already in LDS vals[0..M1)) -> global partials;
// the last block to arrive reduces all of them in block order into fin[0..M1)
// (LDS) and returns true.
__device__ __forceinline__ bool reduce_blocks(const double *vals, int M1, double *partials, unsigned *counter,
                                     double *scratch /* LDS kBlock */, double *fin /* LDS >= M1 */)
{
    __shared__ int is_last;
    const int tid = threadIdx.x;
    for (int c = tid; c < M1; c += kBlock) partials[(long long)blockIdx.x * M1 + c] = vals[c];
    __threadfence(); // release: this block's partials before its arrival is counted
    __syncthreads();
    if (tid == 0) {
        const unsigned prev =
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (prev + 1 == gridDim.x);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    const int nb = gridDim.x;
    const int nsplit = kBlock / M1 > 0 ? kBlock / M1 : 1;
    for (int c0 = 0; c0 < M1; c0 += kBlock) {
        const int span = (M1 - c0) < kBlock ? (M1 - c0) : kBlock;
        const int ns = span == M1 ? nsplit : 1;
        const int c = c0 + tid % span;
        const int part = tid / span;
        double s = 0.0;
        if (part < ns) {
            // plain loads: the acquire fence above made the other blocks' partials visible
            for (int b = part; b < nb; b += ns) s += partials[(long long)b * M1 + c];
        }
        scratch[tid] = s;
        __syncthreads();
        if (tid < span) {
            double t = 0.0;
            for (int q = 0; q < ns; ++q) t += scratch[q * span + tid];
            fin[c0 + tid] = t;
        }
        __syncthreads();
    }
    if (tid == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Folded passes: the block-order sums of columns [0, M1) of column-major block
// partials P (nb <= kBlock blocks) into out[] (LDS), each the same tree of
// additions as column_sum in the finalize kernel (leaf t = 0 + block t; then
// pairs at distance 128, 64, .., 1), so a folded pass reduces bitwise what the
// split pass's finalize reduces.  Wave w takes columns w, w + 4, ..; a round
// issues the loads of 8 columns before the first addition, so the partials
// cost one memory round trip per 32 columns.  Ends with a block barrier; call
// from every thread of the block.
__device__ __forceinline__ void fold_sums(const double *P_, int nb, int M1, double *out)
{
    static_assert(kBlock == 4 * kWave, "the leaf layout assumes 4 waves of 64 lanes");
    constexpr int kCols = 8; // columns per wave and round
    const auto P = gptr(P_);
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    for (int c0 = wv; c0 < M1; c0 += kCols * (kBlock / kWave)) {
        // unconditional loads (absent entries read slot 0 and are masked after),
        // so that none waits for another
        double x[kCols][4];
#pragma unroll
        for (int k = 0; k < kCols; ++k) {
            const int c = c0 + k * (kBlock / kWave);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int b = lane + q * kWave;
                x[k][q] = P[(c < M1 && b < nb) ? (long long)c * nb + b : 0];
            }
        }
#pragma unroll
        for (int k = 0; k < kCols; ++k) {
            const int c = c0 + k * (kBlock / kWave);
            double leaf[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double t = 0.0; // as the finalize's leaf sum starts
                t += (c < M1 && lane + q * kWave < nb) ? x[k][q] : 0.0;
                leaf[q] = t;
            }
            double v = (leaf[0] + leaf[2]) + (leaf[1] + leaf[3]);
            for (int w = kWave / 2; w > 0; w >>= 1) v += __shfl_down(v, w);
            if (lane == 0 && c < M1) out[c] = v;
        }
    }
    __syncthreads();
}

// Folded solves (SegMember::fold): closes the second DGKS correction of step j
// where it ran (its squared norm from the partials at partials + nb (ncv + 1),
// then the epilogue) on a private copy of ctl2; block-wide.  `publish`: this
// block writes the result to ctl and the step's tridiagonal entry.  Returns
// the address of the resulting scale of v_{j+1}: ctl2's own when the
// correction did not run (the usual case: no barrier then), else LDS.
__device__ __forceinline__ const double *fold_close_pass3(const SegMember &mb, int j, bool publish)
{
    __shared__ Ctl lc;
    __shared__ double nrm2[1];
    __shared__ double unused[2];
    if (!mb.ctl2->need) { // uniform: one value for the block
        if (publish && threadIdx.x == 0) *mb.ctl = *mb.ctl2;
        return &mb.ctl2->xscale;
    }
    if (threadIdx.x == 0) lc = *mb.ctl2;
    const double *Pb = mb.partials + (long long)mb.nb * (mb.ncv + 1);
    fold_sums(Pb + (long long)(j + 1) * mb.nb, mb.nb, 1, nrm2); // column j + 1: the norm; a barrier
    if (threadIdx.x == 0)
        run_epilogue_at(EPI_PASS3, j, nrm2[0], mb.h0[j], &lc, publish ? mb.hsub + j : &unused[0],
                        publish ? mb.hdia + j : &unused[1]);
    if (publish && threadIdx.x == 0) *mb.ctl = lc;
    __syncthreads();
    return &lc.xscale;
}

}  // namespace dev
}  // namespace hl
