// probes.hip -- kernels of the launch-cost probes (hl_bench_grid_barrier,
// hl_bench_graph_nodes): a cooperative kernel that runs grid-wide barriers with
// bounded waits, and an empty kernel for graphs of kernel nodes.
#include "hl_internal.hpp"

namespace hl {
namespace {

constexpr long long kSpinLimit = 1ll << 22; // polls per barrier before the wait is abandoned

// Waits until *count >= target (acquire); false when the poll budget ran out or
// another block abandoned its wait (then *abort is set for every block).
__device__ bool barrier_wait(unsigned *count, int *abort, unsigned target)
{
    for (long long polls = 0;; ++polls) {
        if (__hip_atomic_load(count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
        if (polls > kSpinLimit || __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// `barriers` rounds: every block writes its 256 doubles of buffer (k & 1),
// crosses the barrier (thread 0: release add, acquire polls) and checks its
// neighbour's; stale reads are counted in *errors.  counter, abort and errors
// start at zero.
__global__ __launch_bounds__(kBlock) void grid_barrier_kernel(unsigned *counter, int *abort, double *buf,
                                                              int barriers, int *errors)
{
    __shared__ int ok;
    const unsigned nb = gridDim.x;
    int err = 0;
    for (int k = 0; k < barriers; ++k) {
        double *b = buf + (long long)(k & 1) * nb * kBlock;
        b[(long long)blockIdx.x * kBlock + threadIdx.x] = static_cast<double>(k + blockIdx.x);
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            ok = barrier_wait(counter, abort, (unsigned)(k + 1) * nb);
        }
        __syncthreads();
        if (!ok) break;
        const unsigned nbr = (blockIdx.x + 1 + k) % nb;
        err += b[(long long)nbr * kBlock + threadIdx.x] != static_cast<double>(k + nbr);
    }
    if (err) atomicAdd(errors, err);
}

__global__ __launch_bounds__(kBlock) void empty_kernel(int *never)
{
    if (never && threadIdx.x == kBlock) *never = 1; // never true: keeps the argument live
}

} // namespace

int grid_barrier_blocks_max(int dev)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grid_barrier_kernel, kBlock, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}

int launch_grid_barrier(int blocks, unsigned *counter, int *abort, double *buf, int barriers, int *errors,
                        hipStream_t s)
{
    void *args[] = {&counter, &abort, &buf, &barriers, &errors};
    HL_TRY_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(grid_barrier_kernel), dim3(blocks),
                                          dim3(kBlock), args, 0, s));
    return HL_OK;
}

int launch_empty(int blocks, hipStream_t s)
{
    empty_kernel<<<blocks, kBlock, 0, s>>>(nullptr);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

} // namespace hl
