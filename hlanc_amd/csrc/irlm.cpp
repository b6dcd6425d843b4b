// irlm.cpp -- host half of the implicitly restarted Lanczos method.
//
// The restart logic of ds_solver (HLanc.hpp:1376-1401, 1492-1589) on the
// ncv x ncv tridiagonal problem, the per-solve device workspace and the HIP
// event profiler.  The code that sequences sweeps, restarts and the Ritz
// vector step over these pieces is engine.cpp.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "hl_internal.hpp"
#include "irlm.hpp"

namespace hl {

struct Pair {
    double d1, d2;
};

// dsgets for "largest algebraic" (HLanc.hpp:1386-1401): ascending by value, then
// the first np (the unwanted ones, used as shifts) by bound descending.
void sort_wanted(int nev, int np, double *ritz, double *bounds)
{
    const int ncv = nev + np;
    std::vector<Pair> p(ncv);
    for (int i = 0; i < ncv; ++i) p[i] = {ritz[i], bounds[i]};
    std::sort(p.begin(), p.end(), [](const Pair &a, const Pair &b) { return a.d1 < b.d1; });
    std::sort(p.begin(), p.begin() + np, [](const Pair &a, const Pair &b) { return a.d2 > b.d2; });
    for (int i = 0; i < ncv; ++i) {
        ritz[i] = p[i].d1;
        bounds[i] = p[i].d2;
    }
}

// dsconv (HLanc.hpp:1376-1384).
int count_converged(int n, const double *ritz, const double *bounds, double tol)
{
    const double eps23 = std::pow(kEps, 2.0 / 3.0);
    int c = 0;
    for (int i = 0; i < n; ++i) c += bounds[i] <= tol * std::max(eps23, std::fabs(ritz[i]));
    return c;
}

static inline void rotate_cols(double *q, int ldq, int c0, int rows, double c, double s)
{
    double *__restrict__ qa = q + (size_t)c0 * ldq;
    double *__restrict__ qb = qa + ldq; // columns c0 and c0 + 1: disjoint
    for (int r = 0; r < rows; ++r) {
        const double a1 = c * qa[r] + s * qb[r];
        qb[r] = -s * qa[r] + c * qb[r];
        qa[r] = a1;
    }
}

// The shift sweeps of dsapps (HLanc.hpp:1492-1589): np implicit QR steps with
// exact shifts on T = (dia, sub), by bulge chasing with deflation; Q (ncv x ncv,
// column-major) accumulates the rotations.
void restart_chase(int nev, int np, const double *shift, double *sub, double *dia, double *q)
{
    const int ncv = nev + np, ldq = ncv;
    std::fill(q, q + (size_t)ncv * ncv, 0.0);
    for (int i = 0; i < ncv; ++i) q[(size_t)i * ldq + i] = 1.0;
    int itop = 1;
    double c = 0.0, s = 0.0, r = 0.0;
    for (int jj = 1; jj <= np; ++jj) {
        int istart = itop;
        for (;;) {
            int iend = ncv;
            for (int i = istart; i <= ncv - 1; ++i) {
                if (sub[i] <= kEps * (std::fabs(dia[i - 1]) + std::fabs(dia[i]))) {
                    sub[i] = 0.0;
                    iend = i;
                    break;
                }
            }
            if (istart < iend) {
                dlartg(dia[istart - 1] - shift[jj - 1], sub[istart], c, s, r);
                {
                    const double d0 = dia[istart - 1], d1 = dia[istart], e1 = sub[istart];
                    const double a1 = c * d0 + s * e1;
                    const double a2 = c * e1 + s * d1;
                    const double a4 = c * d1 - s * e1;
                    const double a3 = c * e1 - s * d0;
                    dia[istart - 1] = c * a1 + s * a2;
                    dia[istart] = c * a4 - s * a3;
                    sub[istart] = c * a3 + s * a4;
                }
                rotate_cols(q, ldq, istart - 1, std::min(istart + jj, ncv), c, s);
                for (int i = istart + 1; i <= iend - 1; ++i) {
                    const double f = sub[i - 1];
                    const double g = s * sub[i];
                    sub[i] = c * sub[i];
                    dlartg(f, g, c, s, r);
                    if (r < 0.0) {
                        r = -r;
                        c = -c;
                        s = -s;
                    }
                    sub[i - 1] = r;
                    const double d0 = dia[i - 1], d1 = dia[i], e1 = sub[i];
                    const double a1 = c * d0 + s * e1;
                    const double a2 = c * e1 + s * d1;
                    const double a3 = c * e1 - s * d0;
                    const double a4 = c * d1 - s * e1;
                    dia[i - 1] = c * a1 + s * a2;
                    dia[i] = c * a4 - s * a3;
                    sub[i] = c * a3 + s * a4;
                    rotate_cols(q, ldq, i - 1, std::min(i + jj, ncv), c, s);
                }
            }
            istart = iend + 1;
            if (sub[iend - 1] < 0.0) {
                sub[iend - 1] = -sub[iend - 1];
                for (int i = 0; i < ncv; ++i) q[(size_t)(iend - 1) * ldq + i] = -q[(size_t)(iend - 1) * ldq + i];
            }
            if (iend < ncv) continue;
            for (int i = itop; i <= ncv - 1; ++i) {
                if (sub[i] > 0.0) break;
                ++itop;
            }
            break;
        }
    }
    for (int i = itop; i <= ncv - 1; ++i)
        if (sub[i] <= kEps * (std::fabs(dia[i - 1]) + std::fabs(dia[i]))) sub[i] = 0.0;
}

// ------------------------------------------------------------------ Profiler

hipEvent_t Profiler::take()
{
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int Profiler::begin(hipStream_t s)
{
    if (!on) return HL_OK;
    open_ev = take();
    if (!open_ev) return HL_ERR_HIP_BASE;
    HL_TRY_HIP(hipEventRecord(open_ev, s));
    return HL_OK;
}

int Profiler::end(hipStream_t s, int cls, double bytes)
{
    if (!on) return HL_OK;
    hipEvent_t b = take();
    if (!b) return HL_ERR_HIP_BASE;
    HL_TRY_HIP(hipEventRecord(b, s));
    pending.push_back({cls, bytes, open_ev, b});
    open_ev = nullptr;
    return HL_OK;
}

int Profiler::flush()
{
    for (const Rec &r : pending) {
        float ms = 0.0f;
        HL_TRY_HIP(hipEventSynchronize(r.b));
        HL_TRY_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        tot.launches[r.cls] += 1;
        tot.seconds[r.cls] += 1e-3 * ms;
        tot.bytes[r.cls] += r.bytes;
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    pending.clear();
    return HL_OK;
}

void Profiler::release()
{
    for (const Rec &r : pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    pending.clear();
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
    if (open_ev) (void)hipEventDestroy(open_ev);
    open_ev = nullptr;
}

hl_profile &thread_profile()
{
    static thread_local hl_profile p{};
    return p;
}

// ------------------------------------------------------------------ Workspace

Workspace::~Workspace() { release(); }

std::pair<void *, void *> Workspace::detach()
{
    std::pair<void *, void *> p{blob, hblob};
    blob = hblob = nullptr;
    release();
    return p;
}

void Workspace::release()
{
    if (blob) (void)hipFree(blob);
    if (hblob) (void)hipHostFree(hblob);
    blob = hblob = nullptr;
    V[0] = V[1] = nullptr;
    V32[0] = V32[1] = nullptr;
    part = nullptr;
    r = w = hsub = hdia = hbuf = xv = chunk_part = nullptr;
    chunk_cnt = nullptr;
    ctl = nullptr;
    ctl_h = nullptr;
    xh = yh = vres_h = nullptr;
    red = Reduce{};
}

namespace {
long long padded_rows(int n) { return ((long long)n + 63) / 64 * 64; } // 512-byte aligned columns

// Offsets of the workspace's pieces in its device allocation (256-byte aligned).
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes)
    {
        const size_t at = off;
        off += (bytes + 255) / 256 * 256;
        return at;
    }
};

struct WsLayout {
    size_t V0, V1, S0, S1, r, w, hsub, hdia, hbuf, ctl, xv, spart, cpart, ccnt, part, cnt, total;
    WsLayout(int n, int ncv, bool host_op, int nslots, int npairs, bool copy32)
    {
        const size_t ldv = static_cast<size_t>(padded_rows(n));
        const size_t ldt = static_cast<size_t>(tiled_ld(ncv));
        Carve c;
        V0 = c.take(8 * ldv * ldt);
        V1 = c.take(8 * ldv * ldt);
        S0 = copy32 ? c.take(4 * ldv * ldt) : 0;
        S1 = copy32 ? c.take(4 * ldv * ldt) : 0;
        r = c.take(8 * ldv);
        w = c.take(8 * ldv);
        hsub = c.take(8 * (size_t)ncv);
        hdia = c.take(8 * (size_t)ncv);
        hbuf = c.take(16 * (size_t)ncv);
        ctl = c.take(sizeof(Ctl));
        xv = host_op ? c.take(8 * ldv) : 0;
        spart = npairs > 0 ? c.take(8 * (size_t)npairs) : 0;
        cpart = nslots > 0 ? c.take(8 * (size_t)nslots) : 0;
        ccnt = nslots > 0 ? c.take(4 * (size_t)nslots) : 0;
        part = c.take(8 * (size_t)kMaxReduceBlocks * (ncv + 1));
        cnt = c.take(16);
        total = c.off;
    }
};
} // namespace

double Workspace::bytes_needed(int n, int ncv, bool host_op, int nslots, int npairs, bool copy32)
{
    return static_cast<double>(WsLayout(n, ncv, host_op, nslots, npairs, copy32).total);
}

int Workspace::init(int dev_, int n_, int ncv_, bool host_op, int nslots, int npairs, bool copy32, int nres,
                    hipStream_t s)
{
    dev = dev_;
    n = n_;
    ncv = ncv_;
    ldv = padded_rows(n);
    cur = 0;
    const WsLayout L(n, ncv, host_op, nslots, npairs, copy32);
    HL_TRY_HIP(hipMalloc(&blob, L.total));
    char *b = static_cast<char *>(blob);
    V[0] = reinterpret_cast<double *>(b + L.V0);
    V[1] = reinterpret_cast<double *>(b + L.V1);
    V32[0] = copy32 ? reinterpret_cast<float *>(b + L.S0) : nullptr;
    V32[1] = copy32 ? reinterpret_cast<float *>(b + L.S1) : nullptr;
    r = reinterpret_cast<double *>(b + L.r);
    w = reinterpret_cast<double *>(b + L.w);
    hsub = reinterpret_cast<double *>(b + L.hsub);
    hdia = reinterpret_cast<double *>(b + L.hdia);
    hbuf = reinterpret_cast<double *>(b + L.hbuf);
    ctl = reinterpret_cast<Ctl *>(b + L.ctl);
    xv = host_op ? reinterpret_cast<double *>(b + L.xv) : nullptr;
    part = npairs > 0 ? reinterpret_cast<double *>(b + L.spart) : nullptr;
    chunk_part = nslots > 0 ? reinterpret_cast<double *>(b + L.cpart) : nullptr;
    chunk_cnt = nslots > 0 ? reinterpret_cast<unsigned *>(b + L.ccnt) : nullptr;
    red.partials = reinterpret_cast<double *>(b + L.part);
    red.counter = reinterpret_cast<unsigned *>(b + L.cnt);
    red.max_blocks = kMaxReduceBlocks;
    red.max_cols = ncv;
    // the CGS passes read whole tiles: unwritten columns and padding rows must be finite
    HL_TRY_HIP(hipMemsetAsync(V[0], 0, L.V1 - L.V0, s));
    HL_TRY_HIP(hipMemsetAsync(V[1], 0, 8 * static_cast<size_t>(ldv) * static_cast<size_t>(tiled_ld(ncv)), s));
    if (copy32) {
        HL_TRY_HIP(hipMemsetAsync(V32[0], 0, 4 * static_cast<size_t>(ldv) * static_cast<size_t>(tiled_ld(ncv)), s));
        HL_TRY_HIP(hipMemsetAsync(V32[1], 0, 4 * static_cast<size_t>(ldv) * static_cast<size_t>(tiled_ld(ncv)), s));
    }
    if (chunk_cnt) HL_TRY_HIP(hipMemsetAsync(chunk_cnt, 0, sizeof(unsigned) * (size_t)nslots, s));
    HL_TRY_HIP(hipMemsetAsync(red.counter, 0, 16, s));
    const size_t hres = 3 * (size_t)kResBlocksMax * (size_t)(nres > 0 ? nres : 0);
    const size_t hbytes = sizeof(Ctl) + 8 * hres + (host_op ? 16 * (size_t)n : 0);
    HL_TRY_HIP(hipHostMalloc(&hblob, hbytes, hipHostMallocDefault));
    ctl_h = static_cast<Ctl *>(hblob);
    double *hd = reinterpret_cast<double *>(static_cast<char *>(hblob) + sizeof(Ctl));
    vres_h = nres > 0 ? hd : nullptr;
    if (host_op) {
        xh = hd + hres;
        yh = xh + n;
    }
    return HL_OK;
}

} // namespace hl
