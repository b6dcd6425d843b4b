// spmv.hip -- adaptive CSR SpMV for CDNA4 (gfx950).
//
// Replaces cusparseDcsrmv inside dcsrmv_operator_gpu::operator() (HLanc.hpp:553-567).
// The matrix is cut on the host into work items by an nnz walk over row_ptr
// (plan_items; the shape is a plan parameter, default 256 rows / 2048 nnz):
//   * stream items: whole short rows up to the item's row and nnz caps.  The
//     block streams val/col coalesced (non-temporal, all of a thread's loads
//     in flight before its x gathers), multiplies by the gathered x into LDS,
//     then reduces each row from LDS with tpr lanes per row (tpr = 1 is "a
//     lane per short row", tpr = 64 "a wavefront per medium row");
//   * chunk items: a row longer than the nnz cap gets blocks of <= 8192 nnz;
//     the partial sums of a row are added in chunk order by whichever block
//     finishes last (a deterministic split of heavy rows).
// spmv_seg_kernel runs the SpMV of many solves in one launch (one segment of
// the grid per solve).  Every row sum is formed in a fixed order, so results
// are bitwise repeatable.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "hl_device.hpp"
#include "hl_internal.hpp"

namespace hl {
namespace {

__device__ inline double block_sum_to_thread0(double v, double *wred)
{
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & (kWave - 1)) == 0) wred[threadIdx.x / kWave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kBlock / kWave; ++w) t += wred[w];
    }
    return t;
}

__device__ inline long long vcol_pos(long long r, long long ld, int j, bool tiled)
{
    return tiled ? vpos<true>(r, j, ld) : vpos<false>(r, j, ld);
}

// One work item of the SpMV (one block).  AT: matrix values, XT: x, VT: the
// basis column written alongside (vout).
template <typename AT, typename XT, typename VT>
__device__ __forceinline__ void spmv_item(const SpmvBlk b, const int *__restrict__ rp_, const int *__restrict__ col_,
                                          const AT *__restrict__ val_, const XT *__restrict__ x_,
                                          const double *__restrict__ xscale, double *__restrict__ y_,
                                          VT *__restrict__ vout_, float *__restrict__ vout32_, long long vld,
                                          int vj, bool vtiled, double *__restrict__ chunk_part,
                                          unsigned *__restrict__ chunk_cnt)
{
    // global (not flat) memory instructions for the streams and the gathers
    const auto rp = dev::gptr(rp_);
    const auto col = dev::gptr(col_);
    const auto val = dev::gptr(val_);
    const auto x = dev::gptr(x_);
    const auto y = dev::gptr(y_);
    const auto vout = dev::gptr(vout_);
    const auto vout32 = dev::gptr(vout32_);
    __shared__ double prod[kNnzPerBlock];
    __shared__ double wred[kBlock / kWave];
    const double s = xscale ? *xscale : 1.0;
    const int tid = threadIdx.x;

    if (b.tpr > 0) {
        const int nnz = b.p1 - b.p0;
        const int tpr = b.tpr;
        const int nrows = b.row1 - b.row0;
        const int lr = tid / tpr, ll = tid & (tpr - 1);
        const int rstep = kBlock / tpr;
        // loads that do not wait on the gathers go out first: the bounds of
        // this thread's first row of the reduction and, for the basis column,
        // x of row row0 + tid
        int e0 = 0, e1 = 0;
        if (lr < nrows) {
            e0 = rp[b.row0 + lr];
            e1 = rp[b.row0 + lr + 1];
        }
        const double xr = (vout && tid < nrows) ? static_cast<double>(x[b.row0 + tid]) : 0.0;
        // all of this thread's (<= 8) column indices and values in flight at
        // once, then all of its x gathers: the matrix is streamed once
        // (non-temporal), so L2 keeps x
        constexpr int kPer = kNnzPerBlock / kBlock;
        int ci[kPer];
        AT av[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = tid + k * kBlock;
            if (i < nnz) {
                ci[k] = __builtin_nontemporal_load(&col[b.p0 + i]);
                av[k] = __builtin_nontemporal_load(&val[b.p0 + i]);
            }
        }
        XT xv[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (tid + k * kBlock < nnz) xv[k] = x[ci[k]];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = tid + k * kBlock;
            if (i < nnz) prod[i] = static_cast<double>(av[k]) * (static_cast<double>(xv[k]) * s);
        }
        if (vout) {
            for (int r = b.row0 + tid; r < b.row1; r += kBlock) {
                const double xs = (r == b.row0 + tid ? xr : static_cast<double>(x[r])) * s;
                vout[vcol_pos(r, vld, vj, vtiled)] = static_cast<VT>(xs);
                if (vout32) vout32[vcol_pos(r, vld, vj, vtiled)] = static_cast<float>(xs);
            }
        }
        __syncthreads();
        for (int base = 0; base < nrows; base += rstep) {
            const int rr = base + lr;
            double acc = 0.0;
            if (rr < nrows) {
                const int r = b.row0 + rr;
                const int a0 = (base == 0 ? e0 : rp[r]) - b.p0, a1 = (base == 0 ? e1 : rp[r + 1]) - b.p0;
                for (int k = a0 + ll; k < a1; k += tpr) acc += prod[k];
            }
            for (int off = tpr >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
            if (rr < nrows && ll == 0) y[b.row0 + rr] = acc;
        }
        return;
    }

    // chunk of a long row (loads of later terms may run ahead; the sum keeps its order)
    double acc = 0.0;
#pragma unroll 4
    for (int p = b.p0 + tid; p < b.p1; p += kBlock)
        acc += static_cast<double>(__builtin_nontemporal_load(&val[p])) *
               (static_cast<double>(x[__builtin_nontemporal_load(&col[p])]) * s);
    const double tot = block_sum_to_thread0(acc, wred);
    if (tid != 0) return;
    if (vout && b.chunk == 0) {
        const double xs = static_cast<double>(x[b.row0]) * s;
        vout[vcol_pos(b.row0, vld, vj, vtiled)] = static_cast<VT>(xs);
        if (vout32) vout32[vcol_pos(b.row0, vld, vj, vtiled)] = static_cast<float>(xs);
    }
    if (b.nchunks == 1) {
        y[b.row0] = tot;
        return;
    }
    __hip_atomic_store(&chunk_part[b.slot + b.chunk], tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned prev = __hip_atomic_fetch_add(&chunk_cnt[b.slot], 1u, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    if (prev + 1 == static_cast<unsigned>(b.nchunks)) {
        double t = 0.0;
        for (int c = 0; c < b.nchunks; ++c)
            t += __hip_atomic_load(&chunk_part[b.slot + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        y[b.row0] = t;
        __hip_atomic_store(&chunk_cnt[b.slot], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Chunk c of the sliced-ELL form, one wave: lane l sums row 64 c + l in its
// own column order (U entries of the row in flight at once), y(r) = that sum,
// and writes the basis column v(r) = x(r) s beside.
// CT int: absolute columns; CT short: columns relative to the row (banded
// operators: every |col - row| < 2^15), 2 bytes an entry instead of 4.
template <typename VT, typename CT = int>
__device__ __forceinline__ void sell_chunk(int c, int n, const long long *__restrict__ e_ptr_,
                                           const CT *__restrict__ e_col_, const double *__restrict__ e_val_,
                                           const int *__restrict__ rp_, const double *__restrict__ x_,
                                           const double *__restrict__ xscale, double *__restrict__ y_,
                                           VT *__restrict__ vout_, float *__restrict__ vout32_, long long vld,
                                           int vj, bool vtiled)
{
    const auto e_col = dev::gptr(e_col_);
    const auto e_val = dev::gptr(e_val_);
    const auto rp = dev::gptr(rp_);
    const auto x = dev::gptr(x_);
    const auto y = dev::gptr(y_);
    const auto vout = dev::gptr(vout_);
    const auto vout32 = dev::gptr(vout32_);
    const int lane = threadIdx.x & (kWave - 1);
    const int r = c * kWave + lane;
    const long long b0 = e_ptr_[c] + lane;
    const int w = static_cast<int>((e_ptr_[c + 1] - e_ptr_[c]) / kWave);
    const double s = xscale ? *xscale : 1.0;
    int len = 0;
    double xr = 0.0;
    if (r < n) {
        len = rp[r + 1] - rp[r];
        if (vout) xr = x[r];
    }
    constexpr int U = 8;
    double acc = 0.0;
    for (int k0 = 0; k0 < w; k0 += U) {
        int ci[U];
        double av[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < w) { // the whole chunk column: coalesced, padding included
                ci[u] = __builtin_nontemporal_load(&e_col[b0 + (long long)(k0 + u) * kWave]);
                av[u] = __builtin_nontemporal_load(&e_val[b0 + (long long)(k0 + u) * kWave]);
            }
        double xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < len) xv[u] = x[std::is_same<CT, short>::value ? r + ci[u] : ci[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < len) acc += av[u] * (xv[u] * s);
    }
    if (r < n) {
        y[r] = acc;
        if (vout) {
            const double xs = xr * s;
            vout[vcol_pos(r, vld, vj, vtiled)] = static_cast<VT>(xs);
            if (vout32) vout32[vcol_pos(r, vld, vj, vtiled)] = static_cast<float>(xs);
        }
    }
}

__global__ __launch_bounds__(kBlock) void spmv_sell_kernel(int nchunks, int n, const long long *__restrict__ e_ptr,
                                                           const int *__restrict__ e_col,
                                                           const short *__restrict__ e_d16,
                                                           const double *__restrict__ e_val,
                                                           const int *__restrict__ rp, const double *__restrict__ x,
                                                           const double *__restrict__ xscale, double *__restrict__ y,
                                                           double *__restrict__ vout, long long vld, int vj,
                                                           bool vtiled)
{
    const int c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (c >= nchunks) return;
    if (e_d16)
        sell_chunk<double, short>(c, n, e_ptr, e_d16, e_val, rp, x, xscale, y, vout, nullptr, vld, vj, vtiled);
    else
        sell_chunk<double>(c, n, e_ptr, e_col, e_val, rp, x, xscale, y, vout, nullptr, vld, vj, vtiled);
}

template <typename AT, typename XT>
__global__ __launch_bounds__(kBlock) void spmv_adaptive(
    const SpmvBlk *__restrict__ blks, const int *__restrict__ rp, const int *__restrict__ col,
    const AT *__restrict__ val, const XT *__restrict__ x, const double *__restrict__ xscale,
    double *__restrict__ y, XT *__restrict__ vout, long long vld, int vj, bool vtiled,
    double *__restrict__ chunk_part, unsigned *__restrict__ chunk_cnt)
{
    spmv_item<AT, XT, XT>(blks[blockIdx.x], rp, col, val, x, xscale, y, vout, nullptr, vld, vj, vtiled, chunk_part,
                          chunk_cnt);
}

__device__ __forceinline__ int seg_find_spmv(const SegStep *S, int ns, int b)
{
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S[mid].spmv0 <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <typename AT> __device__ __forceinline__ const AT *seg_values(const SegMember &mb);
template <> __device__ __forceinline__ const double *seg_values<double>(const SegMember &mb) { return mb.val; }
template <> __device__ __forceinline__ const float *seg_values<float>(const SegMember &mb) { return mb.val32; }

// The SpMV step of every listed solve in one launch: y = A (r s), V(:, j) = r s.
// AT: the listed solves' matrix values, VT: their basis.
template <typename AT, typename VT>
__global__ __launch_bounds__(kBlock) void spmv_seg_kernel(const SegMember *__restrict__ M,
                                                          const SegStep *__restrict__ S, int ns, int solo)
{
    const int i = seg_find_spmv(S, ns, blockIdx.x);
    const SegStep st = S[i];
    const SegMember mb = M[solo ? 0 : st.mem];
    if constexpr (std::is_same<AT, double>::value) {
        if (mb.e_ptr) { // sliced-ELL form: a wave per 64-row chunk
            const int c = (blockIdx.x - st.spmv0) * (kBlock / kWave) + threadIdx.x / kWave;
            if (c < mb.nchunks) {
                if (mb.e_d16)
                    sell_chunk<VT, short>(c, mb.n, mb.e_ptr, mb.e_d16, mb.e_val, mb.rp, mb.r, &mb.ctl->xscale, mb.w,
                                          static_cast<VT *>(mb.V), mb.V32, mb.vld, st.j, true);
                else
                    sell_chunk<VT>(c, mb.n, mb.e_ptr, mb.e_col, mb.e_val, mb.rp, mb.r, &mb.ctl->xscale, mb.w,
                                   static_cast<VT *>(mb.V), mb.V32, mb.vld, st.j, true);
            }
            return;
        }
    }
    const bool sliced = mb.part != nullptr; // partials now; rows and the basis column in the row-sum kernel
    spmv_item<AT, double, VT>(mb.blks[blockIdx.x - st.spmv0], mb.rp, mb.col, seg_values<AT>(mb), mb.r,
                              &mb.ctl->xscale, sliced ? mb.part : mb.w, sliced ? nullptr : static_cast<VT *>(mb.V),
                              sliced ? nullptr : mb.V32, mb.vld, st.j, true, mb.chunk_part, mb.chunk_cnt);
}

// Rows [r0, r1) of a sliced SpMV: y(r) = its partials added in slice order,
// and the basis column v(r) = x(r) s (and its FP32 copy) written beside.
__device__ __forceinline__ void spmv_rows(int r0, int r1, const int *__restrict__ r_pp_,
                                          const int *__restrict__ r_pairs_, const double *__restrict__ part_,
                                          const double *__restrict__ x_, double s, double *__restrict__ y_,
                                          double *__restrict__ vout_, float *__restrict__ vout32_, long long vld,
                                          int vj, bool vtiled)
{
    const auto r_pp = dev::gptr(r_pp_);
    const auto r_pairs = dev::gptr(r_pairs_);
    const auto part = dev::gptr(part_);
    const auto x = dev::gptr(x_);
    const auto y = dev::gptr(y_);
    const auto vout = dev::gptr(vout_);
    const auto vout32 = dev::gptr(vout32_);
    for (int r = r0 + threadIdx.x; r < r1; r += kBlock) {
        const double xr = vout ? x[r] : 0.0; // independent of the sum: issued first
        const int k0 = r_pp[r], k1 = r_pp[r + 1];
        double acc = 0.0;
        for (int k = k0; k < k1; k += 4) { // four partials in flight, added in slice order
            int id[4];
            double pv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) id[u] = k + u < k1 ? r_pairs[k + u] : -1;
#pragma unroll
            for (int u = 0; u < 4; ++u) pv[u] = id[u] >= 0 ? part[id[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (k + u < k1) acc += pv[u];
        }
        y[r] = acc;
        if (vout) {
            const double xs = xr * s;
            vout[vcol_pos(r, vld, vj, vtiled)] = xs;
            if (vout32) vout32[vcol_pos(r, vld, vj, vtiled)] = static_cast<float>(xs);
        }
    }
}

__global__ __launch_bounds__(kBlock) void spmv_rows_seg_kernel(const SegMember *__restrict__ M,
                                                               const SegStep *__restrict__ S, int ns, int solo)
{
    int lo = 0, hi = ns - 1; // the step owning this block (by red0)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S[mid].red0 <= (int)blockIdx.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const SegStep st = S[lo];
    const SegMember mb = M[solo ? 0 : st.mem];
    const int r0 = (blockIdx.x - st.red0) * kRowsPerSumBlock;
    spmv_rows(r0, min(r0 + kRowsPerSumBlock, mb.n), mb.r_pp, mb.r_pairs, mb.part, mb.r, mb.ctl->xscale, mb.w,
              static_cast<double *>(mb.V), mb.V32, mb.vld, st.j, true);
}

__global__ __launch_bounds__(kBlock) void spmv_rows_kernel(int n, const int *__restrict__ r_pp,
                                                           const int *__restrict__ r_pairs,
                                                           const double *__restrict__ part, const double *x,
                                                           const double *xscale, double *y, double *vout,
                                                           long long vld, int vj, bool vtiled)
{
    const int r0 = blockIdx.x * kRowsPerSumBlock;
    spmv_rows(r0, min(r0 + kRowsPerSumBlock, n), r_pp, r_pairs, part, x, xscale ? *xscale : 1.0, y, vout, nullptr,
              vld, vj, vtiled);
}

// ------------------------------------------------------------ staged slices
// First kernel of the staged-slice form, one block per run of chunks within a
// slice (su = {slice, first chunk, end chunk}): x of the slice into LDS, then
// each wave takes kStagedGroup chunks at a time (their loads in flight
// together), a lane per piece summing its entries in order (lanes stop at
// their own length, so padding is neither loaded nor added).  Within a wave's
// group, lanes are ordered by descending piece length (chunks only as wide as
// their own longest piece) and each lane writes its partial to its piece's
// slot in the group (plen: length | slot << 7), so partials keep the run
// order the bins kernel reads.  A chunk is column-major: entry k of lane l at
// cptr[c] + 64 k + l, so a wave's k-th loads are one contiguous stretch.
constexpr int kStagedGroup = 4;
constexpr int kStagedUnroll = 4; // entries of a piece in flight per trip
__device__ __forceinline__ void staged_block(int4 su, long long ncols, const long long *__restrict__ cptr_,
                                             const unsigned short *__restrict__ plen_,
                                             const unsigned short *__restrict__ ecol_,
                                             const double *__restrict__ eval_, const double *__restrict__ x_,
                                             double s, double *__restrict__ part_, double *xs)
{
    const auto cptr = dev::gptr(cptr_);
    const auto plen = dev::gptr(plen_);
    const auto ecol = dev::gptr(ecol_);
    const auto eval = dev::gptr(eval_);
    const auto x = dev::gptr(x_);
    const auto part = dev::gptr(part_);
    constexpr int W = 1 << kStagedLw;
    constexpr int G = kStagedGroup;
    constexpr int U = kStagedUnroll;
    constexpr int NW = kStagedThreads / kWave;
    // block-uniform values (su may come through a generic pointer), so chunk
    // indices are known uniform and the chunk metadata loads are scalar
    su.x = __builtin_amdgcn_readfirstlane(su.x);
    su.y = __builtin_amdgcn_readfirstlane(su.y);
    su.z = __builtin_amdgcn_readfirstlane(su.z);
    const long long c0 = static_cast<long long>(su.x) << kStagedLw;
    const int nx = static_cast<int>(min(static_cast<long long>(W), ncols - c0));
    if (nx == W && (reinterpret_cast<uintptr_t>(x_ + c0) & 15) == 0) { // a whole, aligned slice: 16-byte loads
        const double2 *x2 = reinterpret_cast<const double2 *>(x_ + c0);
        for (int i = threadIdx.x; i < W / 2; i += kStagedThreads) {
            const double2 t = x2[i];
            xs[2 * i] = t.x;
            xs[2 * i + 1] = t.y;
        }
    } else {
        for (int i = threadIdx.x; i < W; i += kStagedThreads) xs[i] = i < nx ? x[c0 + i] : 0.0;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    // chunk bounds and lane codes (piece length | slot << 7) of a wave's group
    // of chunks; the next group's are loaded while this one's entries are in flight
    long long nb[G], ne[G];
    int nl[G];
    auto meta = [&](int cb) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int c = cb + g * NW;
            nb[g] = ne[g] = 0;
            nl[g] = 0;
            if (c < su.z) {
                nb[g] = cptr[c];
                ne[g] = cptr[c + 1];
                nl[g] = plen[static_cast<long long>(c) * kWave + lane];
            }
        }
    };
    meta(su.y + wave);
    __syncthreads();
    for (int cb = su.y + wave; cb < su.z; cb += G * NW) {
        long long base[G];
        int len[G], slot[G];
        int wmax = 0;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            base[g] = nb[g];
            len[g] = nl[g] & 127;
            slot[g] = nl[g] >> 7;
            wmax = max(wmax, static_cast<int>((ne[g] - nb[g]) / kWave));
        }
        meta(cb + G * NW);
        double acc[G];
#pragma unroll
        for (int g = 0; g < G; ++g) acc[g] = 0.0;
        for (int k0 = 0; k0 < wmax; k0 += U) {
            int ci[G][U];
            double av[G][U];
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (k0 + u < len[g]) {
                        const long long e = base[g] + static_cast<long long>(k0 + u) * kWave + lane;
                        ci[g][u] = __builtin_nontemporal_load(&ecol[e]);
                        av[g][u] = __builtin_nontemporal_load(&eval[e]);
                    }
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (k0 + u < len[g]) acc[g] += av[g][u] * (xs[ci[g][u]] * s);
        }
        // each partial to its piece's slot in the group (the group's stores
        // together write every slot of its chunks); non-temporal stores, which
        // took less time over a whole solve than plain ones
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int c = cb + g * NW;
            if (c < su.z)
                __builtin_nontemporal_store(acc[g], &part[static_cast<long long>(cb + (slot[g] >> 6) * NW) * kWave +
                                                          (slot[g] & (kWave - 1))]);
        }
    }
}

// Second kernel of the staged-slice form, one block (kBinThreads) per row bin
// b of at most kBinCap pieces: the bin's run of every slice copied into LDS
// back to back (ps), and beside it the bin's stretch of the per-row offset
// lists (lsh); then y(r) = the row's partials added in slice order, and the
// basis column v(r) = x(r) s beside.  The bin's run table (built with the
// plan: 2 nsl + 1 ints of LDS offsets and deltas, then the coarse table) is
// read first.
template <typename VT>
__device__ __forceinline__ void bins_block(int b, int nsl, const int4 *__restrict__ bmeta_, const int *__restrict__ run_,
                                           const int *__restrict__ lpp_, const unsigned short *__restrict__ lst_,
                                           const double *__restrict__ part_, const double *__restrict__ x_, double s,
                                           double *__restrict__ y_, VT *__restrict__ vout_,
                                           float *__restrict__ vout32_, long long vld, int vj, bool vtiled,
                                           double *ps, unsigned short *lsh, unsigned short *coarse, int *roff,
                                           int *rdel)
{
    const auto run = dev::gptr(run_);
    const auto lpp = dev::gptr(lpp_);
    const auto lst = dev::gptr(lst_);
    const auto part = dev::gptr(part_);
    const auto x = dev::gptr(x_);
    const auto y = dev::gptr(y_);
    const auto vout = dev::gptr(vout_);
    const auto vout32 = dev::gptr(vout32_);
    constexpr int T = kBinThreads;
    const int tid = threadIdx.x;
    const int4 bm = bmeta_[b]; // rows [r0, r1), offsets lst[l0, l0 + nl)
    const int r0 = bm.x, r1 = bm.y, l0 = bm.z, nl = bm.w - bm.z;
    // this thread's first row: its bounds and x(r), independent of the copy
    const int rf = r0 + tid;
    int f0 = 0, f1 = 0;
    double xf = 0.0;
    if (rf < r1) {
        f0 = lpp[rf] - l0;
        f1 = lpp[rf + 1] - l0;
        if (vout) xf = x[rf];
    }
    // the bin's run table into LDS: roff[sl] (LDS offset of run sl, roff[nsl]
    // = the bin's total) and rdel[sl] (first piece of run sl - roff[sl]),
    // then the coarse table: the run of every kBinStride-th LDS index
    const auto tb = run + static_cast<long long>(b) * (2 * nsl + 1 + kBinCap / kBinStride);
    for (int i = tid; i < 2 * nsl + 1; i += T) roff[i] = tb[i]; // rdel = roff + nsl + 1
    for (int q = tid; q < kBinCap / kBinStride; q += T) coarse[q] = static_cast<unsigned short>(tb[2 * nsl + 1 + q]);
    __syncthreads();
    // the copy: LDS index i comes from run sl, the last run starting at or
    // before i (empty runs never are): from the coarse table's run of the
    // index's group, a short walk; all of a thread's loads in flight, the
    // bin's offsets with them.  The partials are read with ordinary (cached)
    // loads: a run averages under two 128-byte lines, and its first and last
    // lines hold the neighbouring bins' runs of the same slice, which other
    // blocks read at about the same time
    constexpr int kCopy = kBinCap / T;
    const int total = roff[nsl];
    double v[kCopy];
    int li[kCopy];
#pragma unroll
    for (int u = 0; u < kCopy; ++u) {
        const int i = tid + u * T;
        if (i < total) {
            int sl = coarse[i / kBinStride];
            while (roff[sl + 1] <= i) ++sl;
            v[u] = part[rdel[sl] + i];
        }
        if (i < nl) li[u] = lst[l0 + i];
    }
#pragma unroll
    for (int u = 0; u < kCopy; ++u) {
        const int i = tid + u * T;
        if (i < total) ps[i] = v[u];
        if (i < nl) lsh[i] = static_cast<unsigned short>(li[u]);
    }
    __syncthreads();
    for (int r = rf; r < r1; r += T) {
        int k0 = f0, k1 = f1;
        double xr = xf;
        if (r != rf) {
            k0 = lpp[r] - l0;
            k1 = lpp[r + 1] - l0;
            xr = vout ? x[r] : 0.0;
        }
        double acc = 0.0;
        for (int k = k0; k < k1; ++k) acc += ps[lsh[k]];
        y[r] = acc;
        if (vout) {
            const double xs = xr * s;
            vout[vcol_pos(r, vld, vj, vtiled)] = static_cast<VT>(xs);
            if (vout32) vout32[vcol_pos(r, vld, vj, vtiled)] = static_cast<float>(xs);
        }
    }
}

// The work unit of block i of an nb-block launch, dealt to the XCDs in
// groups of K consecutive units.  Blocks are dispatched to the 8 XCDs in turn
// (block i to XCD i % 8), so consecutive units would run on different XCDs;
// here each XCD takes K consecutive units at a time, and what neighbouring
// units share is read again from that XCD's L2: a slice's x for the first
// staged kernel's blocks (groups of 4), the lines that a bin's runs share
// with the neighbouring bins' runs of the same slice for the bins kernel
// (groups of 8).  Groups still go out in unit order, so every XCD gets its
// share of the heavy units (slice 0, the hub-row bins) at the front.  A
// bijection on [0, nb): the last nb mod (8 K) units keep their own block.
template <int K>
__device__ __forceinline__ int xcd_deal(int i, int nb)
{
    const int full = nb / (8 * K) * (8 * K);
    if (i >= full) return i;
    const int j = i >> 3;
    return ((j / K) * 8 + (i & 7)) * K + j % K;
}
constexpr int kStagedGroupXcd = 4, kBinGroupXcd = 8;

// Dynamic LDS of the bins kernels: the run offsets and deltas (2 nsl + 1 ints).
inline size_t bins_dyn_lds(int nsl) { return sizeof(int) * (2 * static_cast<size_t>(nsl) + 1); }

__global__ __launch_bounds__(kStagedThreads) void spmv_staged_kernel(const int4 *__restrict__ sup, long long ncols,
                                                                     const long long *__restrict__ cptr,
                                                                     const unsigned short *__restrict__ plen,
                                                                     const unsigned short *__restrict__ ecol,
                                                                     const double *__restrict__ eval,
                                                                     const double *__restrict__ x,
                                                                     const double *__restrict__ xscale,
                                                                     double *__restrict__ part)
{
    __shared__ double xs[1 << kStagedLw];
    staged_block(sup[xcd_deal<kStagedGroupXcd>(blockIdx.x, gridDim.x)], ncols, cptr, plen, ecol, eval, x,
                 xscale ? *xscale : 1.0, part, xs);
}

__global__ __launch_bounds__(kBinThreads) void spmv_bins_kernel(int nsl, const int4 *__restrict__ bmeta,
                                                                const int *__restrict__ run,
                                                                const int *__restrict__ lpp,
                                                                const unsigned short *__restrict__ lst,
                                                                const double *__restrict__ part,
                                                                const double *__restrict__ x,
                                                                const double *__restrict__ xscale,
                                                                double *__restrict__ y, double *__restrict__ vout,
                                                                long long vld, int vj, bool vtiled)
{
    __shared__ double ps[kBinCap];
    __shared__ unsigned short lsh[kBinCap];
    __shared__ unsigned short coarse[kBinCap / kBinStride];
    extern __shared__ int bins_dyn[];
    bins_block<double>(xcd_deal<kBinGroupXcd>(blockIdx.x, gridDim.x), nsl, bmeta, run, lpp, lst, part, x,
                       xscale ? *xscale : 1.0, y, vout, nullptr, vld, vj, vtiled, ps, lsh, coarse, bins_dyn,
                       bins_dyn + nsl + 1);
}

// The staged-slice SpMVs of every listed solve: blocks numbered by stg0
// (first kernel) and bin0 (bins kernel), as spmv_seg_kernel numbers by spmv0.
__device__ __forceinline__ int seg_find_stg(const SegStep *S, int ns, int b, bool bins)
{
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((bins ? S[mid].bin0 : S[mid].stg0) <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(kStagedThreads) void spmv_staged_seg_kernel(const SegMember *__restrict__ M,
                                                                         const SegStep *__restrict__ S, int ns,
                                                                         int solo)
{
    __shared__ double xs[1 << kStagedLw];
    const int gb = xcd_deal<kStagedGroupXcd>(blockIdx.x, gridDim.x); // a bijection on the whole grid, then the member
    const SegStep st = S[seg_find_stg(S, ns, gb, false)];
    const SegMember mb = M[solo ? 0 : st.mem];
    staged_block(mb.s_sup[gb - st.stg0], mb.n, mb.s_cptr, mb.s_plen, mb.s_ecol, mb.s_eval, mb.r,
                 mb.ctl->xscale, mb.part, xs);
}

__global__ __launch_bounds__(kBinThreads) void spmv_bins_seg_kernel(const SegMember *__restrict__ M,
                                                                    const SegStep *__restrict__ S, int ns, int solo)
{
    __shared__ double ps[kBinCap];
    __shared__ unsigned short lsh[kBinCap];
    __shared__ unsigned short coarse[kBinCap / kBinStride];
    extern __shared__ int bins_dyn[];
    const int gb = xcd_deal<kBinGroupXcd>(blockIdx.x, gridDim.x); // a bijection on the whole grid, then the member
    const SegStep st = S[seg_find_stg(S, ns, gb, true)];
    const SegMember mb = M[solo ? 0 : st.mem];
    bins_block<double>(gb - st.bin0, mb.stg_nsl, mb.s_bmeta, mb.s_run, mb.s_lpp, mb.s_lst, mb.part, mb.r,
                       mb.ctl->xscale, mb.w, static_cast<double *>(mb.V), mb.V32, mb.vld, st.j, true, ps, lsh, coarse,
                       bins_dyn, bins_dyn + mb.stg_nsl + 1);
}

template <typename T>
int dev_upload(T **dst, const T *src, size_t count)
{
    *dst = nullptr;
    if (count == 0) count = 1;
    HL_TRY_HIP(hipMalloc(reinterpret_cast<void **>(dst), count * sizeof(T)));
    if (src) HL_TRY_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return HL_OK;
}

}  // namespace

int csr_plan_build(CsrPlan &p, int nrows, int ncols, const int *rp, const int *col,
                   const double *val, int sym_lower, int val_prec)
{
    if (nrows < 0 || ncols < 0 || (nrows > 0 && !rp)) {
        set_error("hl_csr_create: bad shape");
        return HL_ERR_ARG;
    }
    if (sym_lower && nrows != ncols) {
        set_error("hl_csr_create: symmetric storage needs a square matrix");
        return HL_ERR_ARG;
    }
    const long long nnz_in = nrows > 0 ? rp[nrows] : 0;
    for (long long i = 0; i < nnz_in; ++i) {
        if (col[i] < 0 || col[i] >= ncols) {
            set_error("hl_csr_create: column index %d out of range at %lld", col[i], i);
            return HL_ERR_ARG;
        }
    }
    // Symmetric-lower storage expands to general storage with the per-row entry
    // order that dcsr_matrix_t::save_bin produces (HLanc.hpp:339-364).
    std::vector<int> grp, gcol;
    std::vector<double> gval;
    const int *R = rp, *C = col;
    const double *A = val;
    if (sym_lower) {
        std::vector<long long> cnt(static_cast<size_t>(nrows) + 1, 0);
        for (int r = 0; r < nrows; ++r)
            for (int q = rp[r]; q < rp[r + 1]; ++q) {
                cnt[r + 1]++;
                if (col[q] != r) cnt[col[q] + 1]++;
            }
        for (int r = 0; r < nrows; ++r) cnt[r + 1] += cnt[r];
        if (cnt[nrows] > 0x7fffffffLL) {
            set_error("hl_csr_create: expanded nnz exceeds int32");
            return HL_ERR_ARG;
        }
        grp.resize(static_cast<size_t>(nrows) + 1);
        for (int r = 0; r <= nrows; ++r) grp[r] = static_cast<int>(cnt[r]);
        gcol.resize(static_cast<size_t>(cnt[nrows]));
        gval.resize(static_cast<size_t>(cnt[nrows]));
        std::vector<long long> cur(cnt.begin(), cnt.end() - 1);
        for (int r = 0; r < nrows; ++r)
            for (int q = rp[r]; q < rp[r + 1]; ++q) {
                const int c = col[q];
                gcol[cur[r]] = c;
                gval[cur[r]++] = val[q];
                if (c != r) {
                    gcol[cur[c]] = r;
                    gval[cur[c]++] = val[q];
                }
            }
        R = grp.data();
        C = gcol.data();
        A = gval.data();
    }
    const long long nnz = nrows > 0 ? R[nrows] : 0;

    p.nrows = nrows;
    p.ncols = ncols;
    p.nnz = nnz;
    int rc;
    if ((rc = dev_upload(&p.row_ptr, R, static_cast<size_t>(nrows) + 1))) return rc;
    if ((rc = dev_upload(&p.col, C, static_cast<size_t>(nnz)))) return rc;
    if ((rc = dev_upload(&p.val, A, static_cast<size_t>(nnz)))) return rc;
    if (val_prec == HL_PREC_F32) {
        std::vector<float> v32(static_cast<size_t>(nnz));
        for (long long i = 0; i < nnz; ++i) v32[i] = static_cast<float>(A[i]);
        if ((rc = dev_upload(&p.val32, v32.data(), static_cast<size_t>(nnz)))) return rc;
    }
    if ((rc = dev_upload<double>(&p.x_stage, nullptr, static_cast<size_t>(ncols)))) return rc;
    if ((rc = dev_upload<double>(&p.y_stage, nullptr, static_cast<size_t>(nrows)))) return rc;
    return plan_items(p.items, R, nrows, kNnzPerBlock, kRowsPerBlock);
}

// Cuts the rows into work items by an nnz walk over row_ptr (host copy R):
// stream items of <= rows_cap whole rows holding <= nnz_cap nnz, and chunk items
// of <= kLongChunk nnz for every row longer than nnz_cap.  Replaces the plan's
// items and long-row slots.
int plan_items(ItemSet &p, const int *R, int nrows, int nnz_cap, int rows_cap)
{
    if (nnz_cap < kBlock || nnz_cap > kNnzPerBlock || rows_cap < 1 || rows_cap > kMaxRowsPerBlock) {
        set_error("SpMV items: nnz cap %d (256..%d) or rows cap %d (1..%d) out of range", nnz_cap, kNnzPerBlock,
                  rows_cap, kMaxRowsPerBlock);
        return HL_ERR_ARG;
    }
    std::vector<SpmvBlk> blks;
    int slots = 0;
    for (int r = 0; r < nrows;) {
        const int len = R[r + 1] - R[r];
        if (len > nnz_cap) {
            const int nch = (len + kLongChunk - 1) / kLongChunk;
            for (int c = 0; c < nch; ++c) {
                SpmvBlk b{};
                b.row0 = r;
                b.row1 = r + 1;
                b.p0 = R[r] + c * kLongChunk;
                b.p1 = std::min(R[r + 1], b.p0 + kLongChunk);
                b.tpr = 0;
                b.chunk = c;
                b.nchunks = nch;
                b.slot = nch > 1 ? slots : 0;
                blks.push_back(b);
            }
            if (nch > 1) slots += nch;
            ++r;
            continue;
        }
        const int r0 = r;
        int nz = 0;
        while (r < nrows && r - r0 < rows_cap) {
            const int l = R[r + 1] - R[r];
            if (l > nnz_cap || nz + l > nnz_cap) break;
            nz += l;
            ++r;
        }
        const int nr = r - r0;
        int tpr = 1;
        while (tpr < kWave && tpr * 2 * nr <= kBlock) tpr *= 2;
        SpmvBlk b{};
        b.row0 = r0;
        b.row1 = r;
        b.p0 = R[r0];
        b.p1 = R[r];
        b.tpr = tpr;
        blks.push_back(b);
    }
    void *old[] = {p.blks, p.chunk_part, p.chunk_cnt};
    for (void *q : old)
        if (q) (void)hipFree(q);
    p.blks = nullptr;
    p.chunk_part = nullptr;
    p.chunk_cnt = nullptr;
    p.nblk = static_cast<int>(blks.size());
    p.nslots = slots;
    p.nnz_cap = nnz_cap;
    p.rows_cap = rows_cap;
    int rc;
    if ((rc = dev_upload(&p.blks, blks.data(), blks.size()))) return rc;
    if ((rc = dev_upload<double>(&p.chunk_part, nullptr, static_cast<size_t>(slots)))) return rc;
    if ((rc = dev_upload<unsigned>(&p.chunk_cnt, nullptr, static_cast<size_t>(slots)))) return rc;
    HL_TRY_HIP(hipMemset(p.chunk_cnt, 0, sizeof(unsigned) * std::max(slots, 1)));
    HL_TRY_HIP(hipDeviceSynchronize());
    return HL_OK;
}

namespace {
void items_free(ItemSet &t)
{
    void *ptrs[] = {t.blks, t.chunk_part, t.chunk_cnt};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    t = ItemSet{};
}

void sell_free(CsrPlan &p)
{
    void *ptrs[] = {p.e_ptr, p.e_col, p.e_d16, p.e_val};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p.e_ptr = nullptr;
    p.e_col = nullptr;
    p.e_d16 = nullptr;
    p.e_val = nullptr;
    p.nchunks = 0;
    p.sell_entries = 0;
    p.sell_on = false;
}

void slices_free(CsrPlan &p)
{
    void *ptrs[] = {p.s_rp, p.s_col, p.s_val, p.r_pp, p.r_pairs, p.s_part};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p.s_rp = p.s_col = p.r_pp = p.r_pairs = nullptr;
    p.s_val = p.s_part = nullptr;
    p.slice_w = 0;
    p.npairs = 0;
    items_free(p.sitems);
}

void staged_free(CsrPlan &p)
{
    void *ptrs[] = {p.s_cptr, p.s_plen, p.s_ecol, p.s_eval, p.s_sup, p.s_bmeta, p.s_run, p.s_lpp, p.s_lst, p.s_spart};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p.s_cptr = nullptr;
    p.s_plen = nullptr;
    p.s_ecol = p.s_lst = nullptr;
    p.s_eval = p.s_spart = nullptr;
    p.s_sup = nullptr;
    p.s_bmeta = nullptr;
    p.s_run = p.s_lpp = nullptr;
    p.stg_on = false;
    p.stg_nsl = p.stg_nchunks = p.stg_nsup = p.stg_nbins = 0;
    p.stg_npieces = 0;
    p.stg_entries = 0;
}
} // namespace

// Virtual rows (row, slice): slice-major, rows ascending within a slice, each
// virtual row's entries in the row's own order.  r_pp/r_pairs list every row's
// virtual rows in slice order (the order its partials are added in).
int plan_slices(CsrPlan &p, int width)
{
    if (width < 0) {
        set_error("SpMV slices: width %d < 0", width);
        return HL_ERR_ARG;
    }
    slices_free(p);
    if (width == 0 || p.nrows == 0) return HL_OK;
    sell_free(p);
    staged_free(p);
    const int n = p.nrows;
    const long long nnz = p.nnz;
    std::vector<int> R(static_cast<size_t>(n) + 1), C(static_cast<size_t>(nnz));
    std::vector<double> A(static_cast<size_t>(nnz));
    HL_TRY_HIP(hipMemcpy(R.data(), p.row_ptr, sizeof(int) * R.size(), hipMemcpyDeviceToHost));
    if (nnz > 0) {
        HL_TRY_HIP(hipMemcpy(C.data(), p.col, sizeof(int) * C.size(), hipMemcpyDeviceToHost));
        HL_TRY_HIP(hipMemcpy(A.data(), p.val, sizeof(double) * A.size(), hipMemcpyDeviceToHost));
    }
    const int ns = (p.ncols + width - 1) / width;
    // pass 1: virtual rows per slice and entries per virtual row
    std::vector<long long> srows(static_cast<size_t>(ns) + 1, 0);
    std::vector<int> last(static_cast<size_t>(ns), -1);
    for (int r = 0; r < n; ++r)
        for (int q = R[r]; q < R[r + 1]; ++q) {
            const int sl = C[q] / width;
            if (last[sl] != r) {
                last[sl] = r;
                ++srows[sl + 1];
            }
        }
    for (int sl = 0; sl < ns; ++sl) srows[sl + 1] += srows[sl];
    const long long np = srows[ns];
    if (np > 0x7fffffffLL) {
        set_error("SpMV slices: %lld virtual rows exceed int32", np);
        return HL_ERR_ARG;
    }
    // pass 2: each row's virtual rows (slice order) and their sizes
    std::vector<int> vrow_len(static_cast<size_t>(np), 0), rpp(static_cast<size_t>(n) + 1, 0),
        rpairs(static_cast<size_t>(np));
    std::vector<long long> fill(srows.begin(), srows.end() - 1);
    std::vector<int> slot(static_cast<size_t>(ns), -1);
    std::fill(last.begin(), last.end(), -1);
    std::vector<int> touched;
    for (int r = 0; r < n; ++r) {
        touched.clear();
        for (int q = R[r]; q < R[r + 1]; ++q) {
            const int sl = C[q] / width;
            if (last[sl] != r) {
                last[sl] = r;
                slot[sl] = static_cast<int>(fill[sl]++);
                touched.push_back(sl);
            }
            ++vrow_len[slot[sl]];
        }
        std::sort(touched.begin(), touched.end());
        rpp[r + 1] = rpp[r] + static_cast<int>(touched.size());
        for (size_t t = 0; t < touched.size(); ++t) rpairs[rpp[r] + t] = slot[touched[t]];
    }
    std::vector<int> srp(static_cast<size_t>(np) + 1, 0);
    for (long long v = 0; v < np; ++v) srp[v + 1] = srp[v] + vrow_len[v];
    // pass 3: the entries, each row's entries of a slice in their own order
    std::vector<int> sc(static_cast<size_t>(nnz));
    std::vector<double> sv(static_cast<size_t>(nnz));
    std::vector<int> put(srp.begin(), srp.end() - 1);
    std::fill(last.begin(), last.end(), -1);
    std::fill(fill.begin(), fill.end(), 0);
    for (int sl = 0; sl < ns; ++sl) fill[sl] = srows[sl];
    for (int r = 0; r < n; ++r)
        for (int q = R[r]; q < R[r + 1]; ++q) {
            const int sl = C[q] / width;
            if (last[sl] != r) {
                last[sl] = r;
                slot[sl] = static_cast<int>(fill[sl]++);
            }
            const int at = put[slot[sl]]++;
            sc[at] = C[q];
            sv[at] = A[q];
        }
    int rc;
    p.slice_w = width;
    p.npairs = static_cast<int>(np);
    if ((rc = dev_upload(&p.s_rp, srp.data(), srp.size())) || (rc = dev_upload(&p.s_col, sc.data(), sc.size())) ||
        (rc = dev_upload(&p.s_val, sv.data(), sv.size())) || (rc = dev_upload(&p.r_pp, rpp.data(), rpp.size())) ||
        (rc = dev_upload(&p.r_pairs, rpairs.data(), rpairs.size())) ||
        (rc = dev_upload<double>(&p.s_part, nullptr, static_cast<size_t>(np))) ||
        (rc = plan_items(p.sitems, srp.data(), p.npairs, p.items.nnz_cap, p.items.rows_cap))) {
        slices_free(p);
        return rc;
    }
    return HL_OK;
}

// Chunks of kWave rows, entry k of a chunk's row l at e_ptr[c] + kWave k + l;
// a chunk is as wide as its longest row, padding is column 0 and value 0.
int plan_sell(CsrPlan &p, bool on, double max_fill)
{
    sell_free(p);
    if (!on || p.nrows == 0) return HL_OK;
    const int n = p.nrows;
    const long long nnz = p.nnz;
    std::vector<int> R(static_cast<size_t>(n) + 1);
    HL_TRY_HIP(hipMemcpy(R.data(), p.row_ptr, sizeof(int) * R.size(), hipMemcpyDeviceToHost));
    const int nch = (n + kWave - 1) / kWave;
    std::vector<long long> ptr(static_cast<size_t>(nch) + 1, 0);
    int wmax = 0;
    for (int c = 0; c < nch; ++c) {
        int w = 0;
        for (int r = c * kWave; r < std::min(n, (c + 1) * kWave); ++r) w = std::max(w, R[r + 1] - R[r]);
        ptr[c + 1] = ptr[c] + static_cast<long long>(kWave) * w;
        wmax = std::max(wmax, w);
    }
    const long long ne = ptr[nch];
    if (wmax > kNnzPerBlock) { // one wave would walk the whole row: the items split it into chunks
        set_error("SpMV sliced-ELL: a row of %d nonzeros (more than %d)", wmax, kNnzPerBlock);
        return HL_ERR_ARG;
    }
    if (max_fill > 0.0 && static_cast<double>(ne) > max_fill * static_cast<double>(std::max(nnz, 1LL))) {
        set_error("SpMV sliced-ELL: %lld entries for %lld nonzeros exceed %.2fx", ne, nnz, max_fill);
        return HL_ERR_ARG;
    }
    std::vector<int> C(static_cast<size_t>(nnz));
    std::vector<double> A(static_cast<size_t>(nnz));
    if (nnz > 0) {
        HL_TRY_HIP(hipMemcpy(C.data(), p.col, sizeof(int) * C.size(), hipMemcpyDeviceToHost));
        HL_TRY_HIP(hipMemcpy(A.data(), p.val, sizeof(double) * A.size(), hipMemcpyDeviceToHost));
    }
    // banded (every |col - row| < 2^15): columns as 16-bit offsets from the row
    bool banded = true;
    for (int r = 0; r < n && banded; ++r)
        for (int q = R[r]; q < R[r + 1]; ++q)
            if (C[q] - r < -32768 || C[q] - r > 32767) {
                banded = false;
                break;
            }
    std::vector<int> ec(banded ? 0 : static_cast<size_t>(ne), 0);
    std::vector<short> ed(banded ? static_cast<size_t>(ne) : 0, 0);
    std::vector<double> ev(static_cast<size_t>(ne), 0.0);
    for (int r = 0; r < n; ++r) {
        const int c = r / kWave, l = r % kWave;
        for (int q = R[r]; q < R[r + 1]; ++q) {
            const long long at = ptr[c] + static_cast<long long>(kWave) * (q - R[r]) + l;
            if (banded)
                ed[at] = static_cast<short>(C[q] - r);
            else
                ec[at] = C[q];
            ev[at] = A[q];
        }
    }
    int rc;
    if ((rc = dev_upload(&p.e_ptr, ptr.data(), ptr.size())) ||
        (rc = banded ? dev_upload(&p.e_d16, ed.data(), ed.size()) : dev_upload(&p.e_col, ec.data(), ec.size())) ||
        (rc = dev_upload(&p.e_val, ev.data(), ev.size()))) {
        sell_free(p);
        return rc;
    }
    slices_free(p);
    staged_free(p);
    p.nchunks = nch;
    p.sell_entries = ne;
    p.sell_on = true;
    return HL_OK;
}

// The staged-slice form (see CsrPlan).  Host work: pieces generated row by row
// (slices ascending, each slice's entries in the row's own order, cut every
// kStagedPiece entries; this is also each row's summation order), row bins cut
// greedily at kBinCap pieces, pieces bucketed by slice (row order kept) and
// sorted by descending length within each (slice, bin) run (the partials'
// order), each slice padded to whole 64-piece chunks; the first kernel's
// blocks cut from that order; then, within each wave's group of chunks, the
// lanes re-ordered by length with each lane's slot recorded; each bin's run
// table for the bins kernel.
int plan_staged(CsrPlan &p, bool on)
{
    staged_free(p);
    if (!on || p.nrows == 0) return HL_OK;
    constexpr int lw = kStagedLw, bincap = kBinCap;
    const int n = p.nrows;
    const long long nnz = p.nnz;
    const long long nsl_ll = (static_cast<long long>(p.ncols) + (1LL << lw) - 1) >> lw;
    if (nsl_ll > kStagedMaxSlices) {
        set_error("SpMV staged slices: %lld slices of 2^%d columns (at most %d)", nsl_ll, lw, kStagedMaxSlices);
        return HL_ERR_ARG;
    }
    const int nsl = static_cast<int>(std::max(nsl_ll, 1LL));
    std::vector<int> R(static_cast<size_t>(n) + 1), C(static_cast<size_t>(nnz));
    std::vector<double> A(static_cast<size_t>(nnz));
    HL_TRY_HIP(hipMemcpy(R.data(), p.row_ptr, sizeof(int) * R.size(), hipMemcpyDeviceToHost));
    if (nnz > 0) {
        HL_TRY_HIP(hipMemcpy(C.data(), p.col, sizeof(int) * C.size(), hipMemcpyDeviceToHost));
        HL_TRY_HIP(hipMemcpy(A.data(), p.val, sizeof(double) * A.size(), hipMemcpyDeviceToHost));
    }
    // pass 1: pieces per row and per slice
    std::vector<int> last(static_cast<size_t>(nsl), -1), cnt(static_cast<size_t>(nsl), 0), touched;
    std::vector<long long> lpp(static_cast<size_t>(n) + 1, 0), per_slice(static_cast<size_t>(nsl), 0);
    for (int r = 0; r < n; ++r) {
        touched.clear();
        for (int q = R[r]; q < R[r + 1]; ++q) {
            const int sl = C[q] >> lw;
            if (last[sl] != r) {
                last[sl] = r;
                cnt[sl] = 0;
                touched.push_back(sl);
            }
            ++cnt[sl];
        }
        long long k = 0;
        for (int sl : touched) {
            const int pc = (cnt[sl] + kStagedPiece - 1) / kStagedPiece;
            k += pc;
            per_slice[sl] += pc;
        }
        if (k > bincap) {
            set_error("SpMV staged slices: row %d makes %lld pieces (at most %d)", r, k, bincap);
            return HL_ERR_ARG;
        }
        lpp[r + 1] = lpp[r] + k;
    }
    const long long np = lpp[n];
    // row bins of <= bincap pieces
    std::vector<int> brow{0};
    for (int r = 0; r < n; ++r)
        if (lpp[r + 1] - lpp[brow.back()] > bincap) brow.push_back(r);
    brow.push_back(n);
    const int nbins = static_cast<int>(brow.size()) - 1;
    // padded piece index ranges of the slices
    std::vector<long long> sbase(static_cast<size_t>(nsl) + 1, 0);
    for (int sl = 0; sl < nsl; ++sl) sbase[sl + 1] = sbase[sl] + (per_slice[sl] + kWave - 1) / kWave * kWave;
    const long long npad = sbase[nsl];
    if (npad > 0x7fffffffLL) {
        set_error("SpMV staged slices: %lld pieces exceed int32", npad);
        return HL_ERR_ARG;
    }
    // pass 2: pieces in generation order g (row, slice, cut), entries regrouped
    // so that each piece's entries are contiguous
    std::vector<unsigned short> g_sl(static_cast<size_t>(np)), g_len(static_cast<size_t>(np));
    std::vector<long long> g_e(static_cast<size_t>(np));
    std::vector<int> gc(static_cast<size_t>(nnz));
    std::vector<double> ga(static_cast<size_t>(nnz));
    std::vector<int> start(static_cast<size_t>(nsl), 0);
    std::fill(last.begin(), last.end(), -1);
    for (int r = 0; r < n; ++r) {
        touched.clear();
        for (int q = R[r]; q < R[r + 1]; ++q) {
            const int sl = C[q] >> lw;
            if (last[sl] != r) {
                last[sl] = r;
                cnt[sl] = 0;
                touched.push_back(sl);
            }
            ++cnt[sl];
        }
        std::sort(touched.begin(), touched.end());
        int at = R[r];
        long long g = lpp[r];
        for (int sl : touched) {
            start[sl] = at;
            for (int k = 0; k < cnt[sl]; k += kStagedPiece) {
                g_sl[g] = static_cast<unsigned short>(sl);
                g_len[g] = static_cast<unsigned short>(std::min(kStagedPiece, cnt[sl] - k));
                g_e[g] = at + k;
                ++g;
            }
            at += cnt[sl];
        }
        for (int q = R[r]; q < R[r + 1]; ++q) {
            const int sl = C[q] >> lw;
            gc[start[sl]] = C[q] - (sl << lw);
            ga[start[sl]++] = A[q];
        }
    }
    // bucket by slice (generation order, hence row order, kept), then order each
    // (slice, bin) run by descending length (ties: generation order)
    std::vector<long long> fill(sbase.begin(), sbase.end() - 1);
    std::vector<int> at_pos(static_cast<size_t>(npad), -1); // padded index -> generation index
    for (long long g = 0; g < np; ++g) at_pos[fill[g_sl[g]]++] = static_cast<int>(g);
    // run table: first padded index of (bin, slice); row b = nbins holds the ends
    std::vector<int> run(static_cast<size_t>(nbins + 1) * nsl);
    for (int sl = 0; sl < nsl; ++sl) {
        long long i = sbase[sl];
        const long long e = sbase[sl] + per_slice[sl];
        for (int b = 0; b < nbins; ++b) {
            run[static_cast<size_t>(b) * nsl + sl] = static_cast<int>(i);
            const long long i0 = i;
            // pieces of rows below brow[b + 1]: their generation index is below lpp[brow[b + 1]]
            while (i < e && at_pos[i] < lpp[brow[b + 1]]) ++i;
            std::stable_sort(at_pos.begin() + i0, at_pos.begin() + i,
                             [&](int x, int y) { return g_len[x] > g_len[y]; });
        }
        run[static_cast<size_t>(nbins) * nsl + sl] = static_cast<int>(e);
    }
    // LDS offset of every piece within its bin, listed in generation order
    std::vector<unsigned short> lst(static_cast<size_t>(np));
    for (int b = 0; b < nbins; ++b) {
        int o = 0;
        for (int sl = 0; sl < nsl; ++sl) {
            const int i0 = run[static_cast<size_t>(b) * nsl + sl], i1 = run[static_cast<size_t>(b + 1) * nsl + sl];
            for (int i = i0; i < i1; ++i) lst[at_pos[i]] = static_cast<unsigned short>(o + (i - i0));
            o += i1 - i0;
        }
        if (o > bincap) {
            set_error("SpMV staged slices: bin %d holds %d pieces", b, o);
            return HL_ERR_ARG;
        }
    }
    // each bin's run table for the bins kernel: roff[0..nsl] (LDS offset of
    // each slice's run, then the bin's total), rdel[0..nsl) (first piece of the
    // run minus its LDS offset), and the run of every kBinStride-th LDS index
    constexpr int kCoarse = kBinCap / kBinStride;
    const size_t tstride = 2 * static_cast<size_t>(nsl) + 1 + kCoarse;
    std::vector<int> rtab(static_cast<size_t>(nbins) * tstride, 0);
    for (int b = 0; b < nbins; ++b) {
        int *ro = rtab.data() + static_cast<size_t>(b) * tstride, *rd = ro + nsl + 1, *co = rd + nsl;
        int o = 0;
        for (int sl = 0; sl < nsl; ++sl) {
            const int i0 = run[static_cast<size_t>(b) * nsl + sl], i1 = run[static_cast<size_t>(b + 1) * nsl + sl];
            ro[sl] = o;
            rd[sl] = i0 - o;
            o += i1 - i0;
        }
        ro[nsl] = o;
        for (int q = 0, sl = 0; q < kCoarse && q * kBinStride < o; ++q) {
            while (ro[sl + 1] <= q * kBinStride) ++sl; // the last run starting at or before the index
            co[q] = sl;
        }
    }
    // blocks of the first kernel: runs of chunks within a slice, about
    // kStagedSupEntries entries each (chunk widths of the order above)
    const long long nch = npad / kWave;
    auto width0 = [&](long long c) {
        int w = 0;
        for (int l = 0; l < kWave; ++l) {
            const int g = at_pos[c * kWave + l];
            if (g >= 0) w = std::max(w, static_cast<int>(g_len[g]));
        }
        return w;
    };
    std::vector<int4> sup;
    for (int sl = 0; sl < nsl; ++sl) {
        const long long c0 = sbase[sl] / kWave, c1 = sbase[sl + 1] / kWave;
        long long c = c0;
        while (c < c1) {
            const long long a = c;
            long long ent = 0;
            while (c < c1) {
                const long long e = static_cast<long long>(kWave) * width0(c);
                if (c > a && ent + e > kStagedSupEntries) break;
                ent += e;
                ++c;
            }
            sup.push_back(make_int4(sl, static_cast<int>(a), static_cast<int>(c), 0));
        }
    }
    // within each wave's group of chunks (the kStagedGroup chunks a wave of
    // the first kernel takes at once: cb, cb + NW, ..), lanes are re-ordered by
    // descending piece length, so that a chunk is only as wide as its own
    // longest piece; each lane keeps the slot of its piece in the group (the
    // partial is written there, so partials stay in run order)
    constexpr int NW = kStagedThreads / kWave;
    static_assert(kStagedPiece < 128 && kStagedGroup * kWave <= 512, "lane code: 7-bit length, 9-bit slot");
    std::vector<int> lane_g(static_cast<size_t>(npad), -1);      // piece of each lane
    std::vector<unsigned short> plen(static_cast<size_t>(npad), 0); // length | slot << 7
    {
        std::vector<std::pair<int, int>> grp; // (length, slot in the group)
        std::vector<long long> pos;           // the group's lane positions, slot order
        for (const int4 &su : sup)
            for (long long base = su.y; base < su.z; base += static_cast<long long>(kStagedGroup) * NW)
                for (int w = 0; w < NW && base + w < su.z; ++w) {
                    grp.clear();
                    pos.clear();
                    for (int g = 0; g < kStagedGroup; ++g) {
                        const long long c = base + w + static_cast<long long>(g) * NW;
                        if (c >= su.z) break;
                        for (int l = 0; l < kWave; ++l) {
                            const long long q = c * kWave + l;
                            const int pg = at_pos[q];
                            grp.emplace_back(pg >= 0 ? g_len[pg] : 0, static_cast<int>(pos.size()));
                            pos.push_back(q);
                        }
                    }
                    std::stable_sort(grp.begin(), grp.end(),
                                     [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                                         return x.first > y.first;
                                     });
                    for (size_t i = 0; i < grp.size(); ++i) {
                        const int slot = grp[i].second;
                        lane_g[pos[i]] = at_pos[pos[slot]];
                        plen[pos[i]] = static_cast<unsigned short>(grp[i].first | (slot << 7));
                    }
                }
    }
    // chunks (column-major, as wide as their longest piece) and entries
    std::vector<long long> cptr(static_cast<size_t>(nch) + 1, 0);
    for (long long c = 0; c < nch; ++c) {
        int w = 0;
        for (int l = 0; l < kWave; ++l) w = std::max(w, plen[c * kWave + l] & 127);
        cptr[c + 1] = cptr[c] + static_cast<long long>(kWave) * w;
    }
    std::vector<unsigned short> ec(static_cast<size_t>(cptr[nch]), 0);
    std::vector<double> ev(static_cast<size_t>(cptr[nch]), 0.0);
    for (long long c = 0; c < nch; ++c)
        for (int l = 0; l < kWave; ++l) {
            const int g = lane_g[c * kWave + l];
            if (g < 0) continue;
            for (int k = 0; k < g_len[g]; ++k) {
                const long long e = cptr[c] + static_cast<long long>(k) * kWave + l;
                ec[e] = static_cast<unsigned short>(gc[g_e[g] + k]);
                ev[e] = ga[g_e[g] + k];
            }
        }
    std::vector<int> lpp32(static_cast<size_t>(n) + 1);
    for (int r = 0; r <= n; ++r) lpp32[r] = static_cast<int>(lpp[r]);
    std::vector<int4> bmeta(static_cast<size_t>(nbins)); // {first row, end row, first offset, end offset}
    for (int b = 0; b < nbins; ++b)
        bmeta[b] = make_int4(brow[b], brow[b + 1], lpp32[brow[b]], lpp32[brow[b + 1]]);
    int rc;
    if ((rc = dev_upload(&p.s_cptr, cptr.data(), cptr.size())) ||
        (rc = dev_upload(&p.s_plen, plen.data(), plen.size())) ||
        (rc = dev_upload(&p.s_ecol, ec.data(), ec.size())) || (rc = dev_upload(&p.s_eval, ev.data(), ev.size())) ||
        (rc = dev_upload(&p.s_sup, sup.data(), sup.size())) ||
        (rc = dev_upload(&p.s_bmeta, bmeta.data(), bmeta.size())) ||
        (rc = dev_upload(&p.s_run, rtab.data(), rtab.size())) ||
        (rc = dev_upload(&p.s_lpp, lpp32.data(), lpp32.size())) ||
        (rc = dev_upload(&p.s_lst, lst.data(), lst.size())) ||
        (rc = dev_upload<double>(&p.s_spart, nullptr, static_cast<size_t>(npad)))) {
        staged_free(p);
        return rc;
    }
    sell_free(p);
    slices_free(p);
    p.stg_on = true;
    p.stg_nsl = nsl;
    p.stg_nchunks = static_cast<int>(nch);
    p.stg_nsup = static_cast<int>(sup.size());
    p.stg_nbins = nbins;
    p.stg_npieces = np;
    p.stg_entries = cptr[nch];
    return HL_OK;
}

double csr_device_bytes(const CsrPlan &p)
{
    const double n = p.nrows, nnz = static_cast<double>(p.nnz);
    auto items = [](const ItemSet &t) {
        return static_cast<double>(sizeof(SpmvBlk)) * std::max(t.nblk, 1) + 12.0 * std::max(t.nslots, 1);
    };
    // row_ptr, col and values, FP32 values, x and y staging, work items
    double b = 4.0 * (n + 1) + 12.0 * nnz + (p.val32 ? 4.0 * nnz : 0.0) + 8.0 * (p.ncols + n) + items(p.items);
    if (p.slice_w > 0) // virtual-row CSR, row lists, partials
        b += 4.0 * (p.npairs + 1.0) + 12.0 * nnz + 4.0 * (n + 1) + 12.0 * p.npairs + items(p.sitems);
    if (p.sell_on) // chunk pointers, columns (2 or 4 bytes), values
        b += 8.0 * (p.nchunks + 1.0) + (p.e_d16 ? 10.0 : 12.0) * static_cast<double>(p.sell_entries);
    if (p.stg_on) { // chunks, lane codes and partials, entries, blocks, bins, run table, lists
        const double pad = static_cast<double>(kWave) * p.stg_nchunks;
        b += 8.0 * (p.stg_nchunks + 1.0) + 10.0 * pad + 10.0 * static_cast<double>(p.stg_entries) +
             16.0 * (p.stg_nsup + p.stg_nbins) + 4.0 * p.stg_nbins * (2.0 * p.stg_nsl + 1 + kBinCap / kBinStride) +
             4.0 * (n + 1) +
             2.0 * static_cast<double>(p.stg_npieces);
    }
    return b;
}

// Structural choice of the SpMV form: the sliced-ELL form where its chunks
// hold at most kPickSellFill x nnz entries (even rows: stencils, meshes), else
// the staged-slice form where x spans at least kPickStagedCols columns (too
// large for one XCD's L2, so gathers from global memory would miss), else the
// stream items.  A form that refuses the matrix passes to the next choice.
// Measured on one MI355X (standalone SpMV): Erdos-Renyi 1e6 / 2e6 rows staged
// 196 / 215 us against items 253 / 353; Chung-Lu 1e6 / 2e6 staged 97 / 185
// against 121 / 265; at 5e5 rows items win (66 against 81, 56 against 58).
int csr_pick_form(CsrPlan &p, int *form)
{
    int rc;
    if ((rc = plan_slices(p, 0)) || (rc = plan_staged(p, false)) || (rc = plan_sell(p, false, 0.0))) return rc;
    rc = plan_sell(p, true, kPickSellFill);
    if (rc == HL_OK) {
        *form = HL_FORM_SELL;
        return HL_OK;
    }
    if (rc != HL_ERR_ARG) return rc;
    if (static_cast<long long>(p.ncols) >= kPickStagedCols) {
        rc = plan_staged(p, true);
        if (rc == HL_OK) {
            *form = HL_FORM_STAGED;
            return HL_OK;
        }
        if (rc != HL_ERR_ARG) return rc;
    }
    *form = HL_FORM_ITEMS;
    return HL_OK;
}

void csr_plan_free(CsrPlan &p)
{
    sell_free(p);
    slices_free(p);
    staged_free(p);
    items_free(p.items);
    void *ptrs[] = {p.row_ptr, p.col, p.val, p.val32, p.x_stage, p.y_stage};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p = CsrPlan{};
}

int launch_spmv(const CsrPlan &p, const double *x, const double *xscale, double *y, VColumn vcol,
                int val_prec, hipStream_t s)
{
    if (val_prec == HL_PREC_F32) {
        if (!p.val32) {
            set_error("SpMV: FP32 values were not requested at creation");
            return HL_ERR_STATE;
        }
        if (p.items.nblk == 0) return HL_OK;
        spmv_adaptive<float, double><<<p.items.nblk, kBlock, 0, s>>>(
            p.items.blks, p.row_ptr, p.col, p.val32, x, xscale, y, vcol.base, vcol.L.ld, vcol.j, vcol.L.tiled,
            p.items.chunk_part, p.items.chunk_cnt);
    } else if (p.sell_on) { // a wave per 64-row chunk
        if (p.nchunks > 0)
            spmv_sell_kernel<<<(p.nchunks + kBlock / kWave - 1) / (kBlock / kWave), kBlock, 0, s>>>(
                p.nchunks, p.nrows, p.e_ptr, p.e_col, p.e_d16, p.e_val, p.row_ptr, x, xscale, y, vcol.base, vcol.L.ld,
                vcol.j,
                vcol.L.tiled);
    } else if (p.stg_on) { // partials of the pieces (x slices in LDS), then the row bins
        if (p.stg_nsup > 0)
            spmv_staged_kernel<<<p.stg_nsup, kStagedThreads, 0, s>>>(p.s_sup, p.ncols, p.s_cptr, p.s_plen, p.s_ecol,
                                                                     p.s_eval, x, xscale, p.s_spart);
        HL_TRY_HIP(hipGetLastError());
        if (p.stg_nbins > 0)
            spmv_bins_kernel<<<p.stg_nbins, kBinThreads, bins_dyn_lds(p.stg_nsl), s>>>(
                p.stg_nsl, p.s_bmeta, p.s_run, p.s_lpp, p.s_lst, p.s_spart, x, xscale, y, vcol.base, vcol.L.ld, vcol.j,
                vcol.L.tiled);
    } else if (p.slice_w > 0) { // partials of the virtual rows, then the row sums
        if (p.sitems.nblk > 0)
            spmv_adaptive<double, double><<<p.sitems.nblk, kBlock, 0, s>>>(
                p.sitems.blks, p.s_rp, p.s_col, p.s_val, x, xscale, p.s_part, nullptr, 0, 0, false,
                p.sitems.chunk_part, p.sitems.chunk_cnt);
        HL_TRY_HIP(hipGetLastError());
        if (p.nrows > 0)
            spmv_rows_kernel<<<(p.nrows + kRowsPerSumBlock - 1) / kRowsPerSumBlock, kBlock, 0, s>>>(
                p.nrows, p.r_pp, p.r_pairs, p.s_part, x, xscale, y, vcol.base, vcol.L.ld, vcol.j, vcol.L.tiled);
    } else {
        if (p.items.nblk == 0) return HL_OK;
        spmv_adaptive<double, double><<<p.items.nblk, kBlock, 0, s>>>(
            p.items.blks, p.row_ptr, p.col, p.val, x, xscale, y, vcol.base, vcol.L.ld, vcol.j, vcol.L.tiled,
            p.items.chunk_part, p.items.chunk_cnt);
    }
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

}  // namespace hl

namespace hl {

int launch_spmv_staged_seg(const SegMember *M, const SegStep *S, int ns, int stg_blocks, int bin_blocks, int nsl_max,
                           bool solo, hipStream_t s)
{
    if (ns <= 0) return HL_OK;
    if (stg_blocks > 0) spmv_staged_seg_kernel<<<stg_blocks, kStagedThreads, 0, s>>>(M, S, ns, solo ? 1 : 0);
    HL_TRY_HIP(hipGetLastError());
    if (bin_blocks > 0)
        spmv_bins_seg_kernel<<<bin_blocks, kBinThreads, bins_dyn_lds(nsl_max), s>>>(M, S, ns, solo ? 1 : 0);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_spmv_rows_seg(const SegMember *M, const SegStep *S, int ns, int blocks_total, bool solo, hipStream_t s)
{
    if (ns <= 0 || blocks_total <= 0) return HL_OK;
    spmv_rows_seg_kernel<<<blocks_total, kBlock, 0, s>>>(M, S, ns, solo ? 1 : 0);
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

int launch_spmv_seg(const SegMember *M, const SegStep *S, int ns, int blocks_total, int kind, bool solo, hipStream_t s)
{
    if (ns <= 0 || blocks_total <= 0) return HL_OK;
    switch (kind) {
    case SEG_F64:
    case SEG_PROJ32: // (writes the FP32 copy of the basis column too)
        spmv_seg_kernel<double, double><<<blocks_total, kBlock, 0, s>>>(M, S, ns, solo ? 1 : 0);
        break;
    case SEG_V32:
        spmv_seg_kernel<double, float><<<blocks_total, kBlock, 0, s>>>(M, S, ns, solo ? 1 : 0);
        break;
    case SEG_V32_A32:
        spmv_seg_kernel<float, float><<<blocks_total, kBlock, 0, s>>>(M, S, ns, solo ? 1 : 0);
        break;
    default:
        set_error("SpMV: bad segment kind %d", kind);
        return HL_ERR_ARG;
    }
    HL_TRY_HIP(hipGetLastError());
    return HL_OK;
}

}  // namespace hl
