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

__device__ __forceinline__ void run_epilogue(int mode, int j, double nrm2, const double *h_in, Ctl *ctl,
                                    double *h_sub, double *h_dia)
{
    switch (mode) {
    case EPI_NONE:
        ctl->nrm2 = nrm2;
        break;
    case EPI_WNORM:
        ctl->wnorm = sqrt(nrm2);
        break;
    case EPI_PASS1: // HLanc.hpp:1431-1440
        h_dia[j] = h_in[j];
        h_sub[j] = (j == 0) ? 0.0 : ctl->rnorm;
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
        h_dia[j] += h_in[j];
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
        h_dia[j] += h_in[j];
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

}  // namespace dev
}  // namespace hl
