// irlm.cpp -- the implicitly restarted Lanczos driver (host C++ over HIP kernels).
//
// Follows ds_solver / dsaitr / dsapps / dsgets / dsconv / dsesrt of the reference
// (HLanc.hpp:1353-1850) step for step, with the dense work moved on the device:
//   * the basis V, the residual and the alpha/beta recurrence stay in HBM; a
//     dsaitr sweep of np Lanczos steps is a fixed sequence of kernel launches
//     (SpMV, then CGS passes whose DGKS decisions are taken on the device) and
//     the host reads back 2 ncv doubles once per restart;
//   * the ncv x ncv work (tridiagonal eigen, Givens bulge chase, selection,
//     Householder Q) stays on the host in FP64;
//   * V Q of dsapps and of the Ritz-vector step runs as one GEMM into a
//     ping-pong copy of V (no in-place column shuffling).
// Return values and the default tol / restart cap are the reference's.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "hl_internal.hpp"
#include "irlm.hpp"

namespace hl {

namespace {

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

inline void rotate_cols(double *q, int ldq, int c0, int rows, double c, double s)
{
    double *qa = q + (size_t)c0 * ldq, *qb = qa + ldq;
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

} // namespace

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

void Workspace::release()
{
    void *ptrs[] = {V[0], V[1], Vf[0], Vf[1], r, w, hsub, hdia, hbuf, Qd, ctl, vd_tmp, xv, ipart, fin};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (ctl_h) (void)hipHostFree(ctl_h);
    if (hh) (void)hipHostFree(hh);
    if (xh) (void)hipHostFree(xh);
    if (yh) (void)hipHostFree(yh);
    reduce_free(red);
    if (stream) (void)hipStreamDestroy(stream);
    V[0] = V[1] = nullptr;
    Vf[0] = Vf[1] = nullptr;
    r = w = hsub = hdia = hbuf = Qd = vd_tmp = xv = ipart = fin = nullptr;
    ctl = nullptr;
    ctl_h = nullptr;
    hh = xh = yh = nullptr;
    stream = nullptr;
}

int Workspace::init(int dev_, int n_, int ncv_, int prec_, bool host_op, int items)
{
    dev = dev_;
    n = n_;
    ncv = ncv_;
    prec = prec_;
    ldv = ((long long)n + 63) / 64 * 64; // 512-byte aligned columns
    tiled = tuning().basis_tiled != 0;
    HL_TRY_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    const size_t vbytes = sizeof(double) * (size_t)ldv * ncv;
    if (prec == HL_PREC_F32) {
        HL_TRY_HIP(hipMalloc(&Vf[0], vbytes / 2));
        HL_TRY_HIP(hipMalloc(&Vf[1], vbytes / 2));
        HL_TRY_HIP(hipMalloc(&vd_tmp, vbytes)); // FP64 staging of the final basis
    } else {
        HL_TRY_HIP(hipMalloc(&V[0], vbytes));
        HL_TRY_HIP(hipMalloc(&V[1], vbytes));
    }
    HL_TRY_HIP(hipMalloc(&r, sizeof(double) * (size_t)ldv));
    HL_TRY_HIP(hipMalloc(&w, sizeof(double) * (size_t)ldv));
    HL_TRY_HIP(hipMalloc(&hsub, sizeof(double) * ncv));
    HL_TRY_HIP(hipMalloc(&hdia, sizeof(double) * ncv));
    HL_TRY_HIP(hipMalloc(&hbuf, sizeof(double) * 2 * ncv));
    HL_TRY_HIP(hipMalloc(&Qd, sizeof(double) * (size_t)ncv * ncv));
    HL_TRY_HIP(hipMalloc(&ctl, sizeof(Ctl)));
    if (items > 0) { // partial sums of the fused SpMV + projection kernel
        HL_TRY_HIP(hipMalloc(&ipart, sizeof(double) * (size_t)items * (ncv + 1)));
        HL_TRY_HIP(hipMalloc(&fin, sizeof(double) * (ncv + 1)));
    }
    HL_TRY_HIP(hipHostMalloc(&ctl_h, sizeof(Ctl), hipHostMallocDefault));
    HL_TRY_HIP(hipHostMalloc(&hh, sizeof(double) * 2 * ncv, hipHostMallocDefault));
    if (host_op) {
        HL_TRY_HIP(hipMalloc(&xv, sizeof(double) * (size_t)ldv));
        HL_TRY_HIP(hipHostMalloc(&xh, sizeof(double) * (size_t)n, hipHostMallocDefault));
        HL_TRY_HIP(hipHostMalloc(&yh, sizeof(double) * (size_t)n, hipHostMallocDefault));
    }
    const int nb = kMaxReduceBlocks;
    int rc = reduce_alloc(red, nb, ncv);
    if (rc) return rc;
    HL_TRY_HIP(hipMemsetAsync(hsub, 0, sizeof(double) * ncv, stream));
    HL_TRY_HIP(hipMemsetAsync(hdia, 0, sizeof(double) * ncv, stream));
    return HL_OK;
}

// ------------------------------------------------------------------ Solver

int Solver::reset_ctl()
{
    std::memset(ws.ctl_h, 0, sizeof(Ctl));
    ws.ctl_h->rscale = 1.0;
    ws.ctl_h->xscale = 1.0;
    HL_TRY_HIP(hipMemcpyAsync(ws.ctl, ws.ctl_h, sizeof(Ctl), hipMemcpyHostToDevice, ws.stream));
    return HL_OK;
}

// STEP 3 of dsaitr: w = A v_j with v_j = resid * xscale written into V(:, j).
int Solver::apply_op(int j)
{
    Workspace &W = ws;
    ++spmv;
    if (A) {
        int rc = prof.begin(W.stream);
        if (!rc) rc = launch_spmv(*A, W.r, &W.ctl->xscale, W.w, VColumn{W.V[W.cur], W.vl(), j}, HL_PREC_F64,
                                  W.stream);
        if (!rc) {
            const double b = 12.0 * A->nnz + 4.0 * (A->nrows + 1) + 8.0 * A->ncols + 16.0 * A->nrows;
            rc = prof.end(W.stream, K_SPMV, b);
        }
        return rc;
    }
    // host operator: v_j on the device, then x/y through pinned host buffers
    int rc = launch_scale_copy(W.n, W.r, W.ctl, W.xv, VColumn{W.V[W.cur], W.vl(), j}, W.stream);
    if (rc) return rc;
    HL_TRY_HIP(hipMemcpyAsync(W.xh, W.xv, sizeof(double) * W.n, hipMemcpyDeviceToHost, W.stream));
    HL_TRY_HIP(hipStreamSynchronize(W.stream));
    host_op(host_ctx, W.xh, W.yh);
    HL_TRY_HIP(hipMemcpyAsync(W.w, W.yh, sizeof(double) * W.n, hipMemcpyHostToDevice, W.stream));
    return HL_OK;
}

// One Lanczos step j: SpMV, then CGS + up to two DGKS corrections
// (HLanc.hpp:1410-1477); every decision is taken on the device.
int Solver::step(int j)
{
    Workspace &W = ws;
    const double *V = W.V[W.cur];
    double *h0 = W.hbuf, *h1 = W.hbuf + W.ncv;
    const int m = j + 1;
    int rc;
    if (A && fuse && W.tiled) { // SpMV and the first CGS half in one persistent kernel
        ++spmv;
        if ((rc = prof.begin(W.stream))) return rc;
        if ((rc = launch_spmv_project(*A, W.r, &W.ctl->xscale, W.w, W.V[W.cur], W.vl(), j, h0, W.ipart, W.fin,
                                      W.red, W.ctl, W.stream)))
            return rc;
        const double b = 12.0 * A->nnz + 4.0 * (A->nrows + 1) + 8.0 * A->ncols + 16.0 * A->nrows + 8.0 * W.n * j;
        if ((rc = prof.end(W.stream, K_SPMV_PROJ, b))) return rc;
    } else {
        if ((rc = apply_op(j))) return rc;
        if ((rc = cgs_step(m, V, W.w, nullptr, nullptr, h0, EPI_WNORM, j))) return rc;
    }
    if ((rc = cgs_step(m, V, W.w, W.r, h0, h1, EPI_PASS1, j))) return rc;
    if ((rc = cgs_step(m, V, W.r, W.r, h1, h0, EPI_PASS2, j))) return rc;
    return cgs_step(m, V, W.r, W.r, h0, nullptr, EPI_PASS3, j);
}

// One CGS launch.  Profiled solves decide on the host whether a DGKS
// correction runs (so that only executed passes are timed); otherwise the
// kernel itself exits early when the device-side decision says so.
int Solver::cgs_step(int m, const double *V, const double *w_in, double *w_out, const double *h_in, double *h_out,
                     int mode, int j)
{
    Workspace &W = ws;
    if (prof.on && (mode == EPI_PASS2 || mode == EPI_PASS3)) {
        HL_TRY_HIP(hipMemcpyAsync(W.ctl_h, W.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, W.stream));
        HL_TRY_HIP(hipStreamSynchronize(W.stream));
        if (!W.ctl_h->need) return HL_OK;
    }
    int rc = prof.begin(W.stream);
    if (rc) return rc;
    if ((rc = launch_cgs(W.n, m, V, W.vl(), w_in, w_out, h_in, h_out, W.red, W.ctl, W.hsub, W.hdia, mode, j,
                         W.stream)))
        return rc;
    const double bytes = 8.0 * W.n * (m + 1 + (w_out ? 1 : 0));
    return prof.end(W.stream, K_CGS, bytes);
}

// dsaitr(k, np) and the read-back of the new alpha/beta entries and rnorm.
int Solver::extend(int k, int np)
{
    Workspace &W = ws;
    for (int j = k; j < k + np; ++j) {
        int rc = step(j);
        if (rc) return rc;
    }
    const int cnt = W.ncv - k;
    HL_TRY_HIP(hipMemcpyAsync(W.hh, W.hsub + k, sizeof(double) * cnt, hipMemcpyDeviceToHost, W.stream));
    HL_TRY_HIP(hipMemcpyAsync(W.hh + W.ncv, W.hdia + k, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                              W.stream));
    HL_TRY_HIP(hipMemcpyAsync(W.ctl_h, W.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, W.stream));
    HL_TRY_HIP(hipStreamSynchronize(W.stream));
    if (prof.on) {
        int rc = prof.flush();
        if (rc) return rc;
    }
    for (int i = 0; i < cnt; ++i) {
        sub[k + i] = W.hh[i];
        dia[k + i] = W.hh[W.ncv + i];
    }
    rnorm = W.ctl_h->rnorm;
    cgs = W.ctl_h->cgs;
    return HL_OK;
}

// Device half of dsapps (HLanc.hpp:1591-1614): V := V Q(:, 0:nev) into the other
// ping-pong buffer and resid := Q(ncv-1, nev-1) resid + beta V Q(:, nev).
int Solver::apply_restart(int nev, const double *q)
{
    Workspace &W = ws;
    const int ncv = W.ncv;
    const bool has_beta = sub[nev] > 0.0;
    const int N = has_beta ? nev + 1 : nev;
    HL_TRY_HIP(hipMemcpyAsync(W.Qd, q, sizeof(double) * (size_t)ncv * N, hipMemcpyHostToDevice, W.stream));
    const double qlast = q[(size_t)(nev - 1) * ncv + ncv - 1];
    const int nxt = 1 - W.cur;
    int rc = prof.begin(W.stream);
    if (rc) return rc;
    if (has_beta) {
        rc = launch_vq(W.n, ncv, N, W.V[W.cur], W.vl(), W.Qd, ncv, W.V[nxt], W.vl(), use_mfma, nev, W.r, qlast,
                       sub[nev], W.red, W.ctl, W.stream);
        if (!rc) rc = prof.end(W.stream, K_VQ, 8.0 * W.n * (ncv + N + 2));
    } else {
        rc = launch_vq(W.n, ncv, N, W.V[W.cur], W.vl(), W.Qd, ncv, W.V[nxt], W.vl(), use_mfma, -1, nullptr, 0.0,
                       0.0, W.red, W.ctl, W.stream);
        if (!rc) rc = prof.end(W.stream, K_VQ, 8.0 * W.n * (ncv + N));
        if (!rc) rc = launch_scale_rscale(W.n, W.r, qlast, W.ctl, W.stream);
        if (!rc) rc = launch_norm(W.n, W.r, W.red, W.ctl, EPI_RNORM, W.stream);
    }
    if (rc) return rc;
    // keep the host copy of Q alive until the async copy is done
    HL_TRY_HIP(hipStreamSynchronize(W.stream));
    W.cur = nxt;
    return HL_OK;
}

int Solver::run(int nev0, double *values, double *vectors, int ldv_host, const hl_options &o, hl_info &info)
{
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    Workspace &W = ws;
    const int n = W.n, ncv = W.ncv;
    const double tol = o.tol > 0.0 ? o.tol : kEps;
    const int maxiter = o.maxiter > 0 ? o.maxiter : 200;
    const double eps23 = std::pow(kEps, 2.0 / 3.0);
    use_mfma = o.use_mfma;
    fuse = tuning().fuse_head != 0 && (o.flags & HL_FLAG_NO_FUSE) == 0;
    prof.on = (o.flags & HL_FLAG_PROFILE) != 0;
    prof.tot = hl_profile{};
    sub.assign(ncv, 0.0);
    dia.assign(ncv, 0.0);

    int rc = reset_ctl();
    if (rc) return rc;
    if (o.v0) {
        HL_TRY_HIP(hipMemcpyAsync(W.r, o.v0, sizeof(double) * n, hipMemcpyHostToDevice, W.stream));
    } else {
        if ((rc = launch_start_vector(n, o.seed, W.r, W.stream))) return rc; // dgetv0 (HLanc.hpp:1368-1374)
    }
    if ((rc = launch_norm(n, W.r, W.red, W.ctl, EPI_RNORM, W.stream))) return rc;

    int nev = nev0, np = ncv - nev0, nconv = 0;
    const int np0 = np;
    std::vector<double> ritz(ncv), bounds(ncv), q((size_t)ncv * ncv), e(ncv), work(ncv);
    std::vector<double> ritz_saved(ncv), bounds_saved(ncv);

    if ((rc = extend(0, nev0))) return rc;
    int iter = 0;
    int ret = HL_NOT_CONVERGED;
    for (iter = 0; iter < maxiter; ++iter) {
        restarts = iter;
        if ((rc = extend(nev, np))) return rc;
        // Ritz values and bounds (HLanc.hpp:1687-1698)
        std::copy(dia.begin(), dia.end(), ritz.begin());
        for (int i = 0; i + 1 < ncv; ++i) e[i] = sub[i + 1];
        if (tridiag_eig(ncv, ritz.data(), e.data(), q.data(), ncv)) goto finish;
        for (int i = 0; i < ncv; ++i) bounds[i] = rnorm * std::fabs(q[(size_t)i * ncv + ncv - 1]);
        ritz_saved = ritz;
        bounds_saved = bounds;

        nev = nev0;
        np = np0;
        sort_wanted(nev, np, ritz.data(), bounds.data());
        std::copy(bounds.begin() + np, bounds.end(), work.begin() + np);
        nconv = count_converged(nev, ritz.data() + np, work.data() + np, tol);
        for (int j = 0; j < np; ++j) { // HLanc.hpp:1708-1713 (np re-read each trip)
            if (bounds[j] == 0.0) {
                --np;
                ++nev;
            }
        }
        if (nconv >= nev0 || np == 0) break;
        {
            const int nevbef = nev; // HLanc.hpp:1719-1730
            nev += std::min(nconv, np / 2);
            if (nev == 1 && ncv >= 6)
                nev = ncv / 2;
            else if (nev == 1 && ncv > 2)
                nev = 2;
            np = ncv - nev;
            if (nevbef < nev) sort_wanted(nev, np, ritz.data(), bounds.data());
        }
        restart_chase(nev, np, ritz.data(), sub.data(), dia.data(), q.data());
        if ((rc = apply_restart(nev, q.data()))) return rc;
    }
    restarts = iter;
    if (nconv < nev0) goto finish; // HLanc.hpp:1762

    { // Ritz vectors (HLanc.hpp:1765-1843)
        nev = nev0;
        np = np0;
        std::vector<double> idx(ncv);
        std::vector<char> select(ncv, 0);
        for (int j = 0; j < ncv; ++j) idx[j] = j + 1;
        sort_wanted(nev, np, ritz_saved.data(), idx.data());
        int reord = 0, numcnv = 0;
        for (int j = 0; j < ncv; ++j) {
            const double temp1 = std::max(eps23, std::fabs(ritz_saved[ncv - j - 1]));
            const int jj = static_cast<int>(idx[ncv - j - 1]);
            if (numcnv < nconv && bounds_saved[jj - 1] <= tol * temp1) {
                select[jj - 1] = 1;
                ++numcnv;
                if (jj > nev) reord = 1;
            }
        }
        if (numcnv != nconv) goto finish;
        std::vector<double> ihd(dia), ihb(ncv, 0.0), iq((size_t)ncv * ncv), tau(ncv), qf((size_t)ncv * ncv);
        for (int i = 0; i + 1 < ncv; ++i) ihb[i] = sub[i + 1];
        if (tridiag_eig(ncv, ihd.data(), ihb.data(), iq.data(), ncv)) goto finish;
        if (reord) { // selected columns to the front (HLanc.hpp:1808-1831)
            int left = 0, right = ncv - 1;
            for (;;) {
                while (select[left]) ++left;
                while (!select[right]) --right;
                if (left > right) break;
                std::swap(ihd[left], ihd[right]);
                for (int i = 0; i < ncv; ++i) std::swap(iq[(size_t)left * ncv + i], iq[(size_t)right * ncv + i]);
                ++left;
                --right;
            }
        }
        // dsesrt (HLanc.hpp:1626-1647): ascending values, Q columns alongside
        std::vector<int> order(nconv);
        for (int i = 0; i < nconv; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ihd[a] < ihd[b]; });
        std::vector<double> vals(nconv), qs((size_t)ncv * nconv);
        for (int i = 0; i < nconv; ++i) {
            vals[i] = ihd[order[i]];
            std::copy(iq.begin() + (size_t)order[i] * ncv, iq.begin() + (size_t)(order[i] + 1) * ncv,
                      qs.begin() + (size_t)i * ncv);
        }
        // dgeqr2 + dorm2r("R","N") (HLanc.hpp:1837-1840) as one V Q_full GEMM
        householder_q(ncv, nconv, qs.data(), ncv, tau.data(), qf.data(), ncv);
        HL_TRY_HIP(hipMemcpyAsync(W.Qd, qf.data(), sizeof(double) * (size_t)ncv * ncv, hipMemcpyHostToDevice,
                                  W.stream));
        const int nxt = 1 - W.cur;
        if ((rc = prof.begin(W.stream))) return rc;
        if ((rc = launch_vq(n, ncv, ncv, W.V[W.cur], W.vl(), W.Qd, ncv, W.V[nxt], VLayout{W.ldv, false}, use_mfma,
                            -1, nullptr, 0.0, 0.0, W.red, W.ctl, W.stream)))
            return rc;
        if ((rc = prof.end(W.stream, K_VQ, 16.0 * n * ncv))) return rc;
        // sync_host (HLanc.hpp:1063-1071): all ncv columns to the caller
        HL_TRY_HIP(hipMemcpy2DAsync(vectors, sizeof(double) * (size_t)ldv_host, W.V[nxt],
                                    sizeof(double) * (size_t)W.ldv, sizeof(double) * (size_t)n, ncv,
                                    hipMemcpyDeviceToHost, W.stream));
        HL_TRY_HIP(hipStreamSynchronize(W.stream));
        std::copy(vals.begin(), vals.end(), values);
        ret = HL_OK;
    }

finish:
    if (prof.on) {
        (void)hipStreamSynchronize(W.stream);
        (void)prof.flush();
        thread_profile() = prof.tot;
    }
    prof.release();
    info.ret = ret;
    info.restarts = restarts;
    info.nconv = nconv;
    info.spmv = spmv;
    info.cgs_passes = cgs;
    info.rnorm = rnorm;
    info.seconds = std::chrono::duration<double>(clock::now() - t0).count();
    info.max_rel_resid = -1.0;
    info.refine_shift = 0.0;
    return ret;
}

} // namespace hl
