// engine.cpp -- eigensolves driven sweep by sweep, alone or many at once.
//
// A Member carries one solve of ds_solver (HLanc.hpp:1649-1850): its device
// workspace and the host half of the algorithm (restart logic, tridiagonal
// eigen, selection, Householder Q) as a state machine.  The Engine advances all
// active members in lockstep.  A Lanczos step of every member with a step left
// in its current sweep is one segmented SpMV launch plus four segmented CGS
// passes (each a partials kernel and a per-column finalize kernel):
//   * the basis V, the residual and the alpha/beta recurrence stay in HBM; the
//     DGKS decisions are taken on the device (a correction pass that is not
//     needed exits at once) and the host reads back 2 ncv doubles per sweep;
//   * the ncv x ncv work stays on the host in FP64;
//   * V Q of dsapps and of the Ritz-vector step is one GEMM into a ping-pong
//     copy of V (no in-place column shuffling).
// A solve run alone is an engine with one member; every block of a segmented
// launch computes what it computes in a one-member launch, so a member's
// results are bitwise the same alone and in any batch.  Return values and the
// default tol / restart cap are the reference's.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "hl_device.hpp"
#include "hl_internal.hpp"
#include "irlm.hpp"

namespace hl {

namespace {
double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
} // namespace

// ------------------------------------------------------------------ Member

double Member::cgs_bytes_pass(int m) const { return 8.0 * n * (m + 2); }

int Member::start(hipStream_t s)
{
    Workspace &W = ws;
    t0 = now_s();
    tol = opts.tol > 0.0 ? opts.tol : kEps;
    maxiter = opts.maxiter > 0 ? opts.maxiter : 200;
    sub.assign(ncv, 0.0);
    dia.assign(ncv, 0.0);
    ritz.assign(ncv, 0.0);
    bounds.assign(ncv, 0.0);
    q.assign((size_t)ncv * ncv, 0.0);
    e.assign(ncv, 0.0);
    work.assign(ncv, 0.0);
    ritz_saved.clear();
    bounds_saved.clear();
    spmv = cgs = p2 = p3 = m2sum = m3sum = 0;
    nconv = 0;
    rnorm = 0.0;
    ret = HL_NOT_CONVERGED;
    std::memset(W.ctl_h, 0, sizeof(Ctl));
    W.ctl_h->rscale = 1.0;
    W.ctl_h->xscale = 1.0;
    HL_TRY_HIP(hipMemcpyAsync(W.ctl, W.ctl_h, sizeof(Ctl), hipMemcpyHostToDevice, s));
    HL_TRY_HIP(hipMemsetAsync(W.hsub, 0, sizeof(double) * ncv, s));
    HL_TRY_HIP(hipMemsetAsync(W.hdia, 0, sizeof(double) * ncv, s));
    int rc;
    if (opts.v0) {
        HL_TRY_HIP(hipMemcpyAsync(W.r, opts.v0, sizeof(double) * n, hipMemcpyHostToDevice, s));
    } else if ((rc = launch_start_vector(n, opts.seed, W.r, s))) { // dgetv0 (HLanc.hpp:1368-1374)
        return rc;
    }
    if ((rc = launch_norm(n, W.r, W.red, W.ctl, EPI_RNORM, s))) return rc;
    nev = nev0;
    np = ncv - nev0;
    iter = 0;
    stage = 0;
    sweep_k = 0;
    sweep_np = nev0; // dsaitr(0, nev0) (HLanc.hpp:1680)
    phase = ACTIVE;
    return HL_OK;
}

// STEP 3 of dsaitr for a host operator: v_j = resid * xscale on the device,
// then y = A v_j through pinned host buffers (HLanc.hpp:1412-1424).
int Member::host_apply(int j, hipStream_t s)
{
    Workspace &W = ws;
    int rc = launch_scale_copy(n, W.r, W.ctl, W.xv, VColumn{W.V[W.cur], W.vl(), j}, s);
    if (rc) return rc;
    HL_TRY_HIP(hipMemcpyAsync(W.xh, W.xv, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HL_TRY_HIP(hipStreamSynchronize(s));
    host_op(host_ctx, W.xh, W.yh);
    HL_TRY_HIP(hipMemcpyAsync(W.w, W.yh, sizeof(double) * n, hipMemcpyHostToDevice, s));
    return HL_OK;
}

// Device half of dsapps (HLanc.hpp:1591-1614): V := V Q(:, 0:nev) into the other
// ping-pong buffer and resid := Q(ncv-1, nev-1) resid + beta V Q(:, nev).  The
// host q stays untouched until the engine's next stream synchronisation.
int Member::restart(hipStream_t s)
{
    Workspace &W = ws;
    const bool has_beta = sub[nev] > 0.0;
    const int N = has_beta ? nev + 1 : nev;
    HL_TRY_HIP(hipMemcpyAsync(W.Qd, q.data(), sizeof(double) * (size_t)ncv * N, hipMemcpyHostToDevice, s));
    const double qlast = q[(size_t)(nev - 1) * ncv + ncv - 1];
    const int nxt = 1 - W.cur;
    int rc = prof->begin(s);
    if (rc) return rc;
    if (has_beta) {
        rc = launch_vq(n, ncv, N, W.V[W.cur], W.vl(), W.Qd, ncv, W.V[nxt], W.vl(), opts.use_mfma, nev, W.r, qlast,
                       sub[nev], W.red, W.ctl, s);
        if (!rc) rc = prof->end(s, K_VQ, 8.0 * n * (ncv + N + 2));
    } else {
        rc = launch_vq(n, ncv, N, W.V[W.cur], W.vl(), W.Qd, ncv, W.V[nxt], W.vl(), opts.use_mfma, -1, nullptr, 0.0,
                       0.0, W.red, W.ctl, s);
        if (!rc) rc = prof->end(s, K_VQ, 8.0 * n * (ncv + N));
        if (!rc) rc = launch_scale_rscale(n, W.r, qlast, W.ctl, s);
        if (!rc) rc = launch_norm(n, W.r, W.red, W.ctl, EPI_RNORM, s);
    }
    if (rc) return rc;
    W.cur = nxt;
    return HL_OK;
}

// The rest of ds_solver after a sweep (HLanc.hpp:1687-1736): Ritz values and
// bounds, convergence, nev adaptation, exact-shift restart; or the end.
int Member::after_sweep(hipStream_t s)
{
    const int k = sweep_k, cnt = ncv - k;
    for (int i = 0; i < cnt; ++i) {
        sub[k + i] = ws.hh[i];
        dia[k + i] = ws.hh[ncv + i];
    }
    rnorm = ws.ctl_h->rnorm;
    cgs = ws.ctl_h->cgs;
    p2 = ws.ctl_h->p2;
    p3 = ws.ctl_h->p3;
    m2sum = ws.ctl_h->m2sum;
    m3sum = ws.ctl_h->m3sum;
    spmv += sweep_np;
    if (stage == 0) { // the initial dsaitr(0, nev0) is followed directly by the main loop
        stage = 1;
        restarts = iter;
        sweep_k = nev;
        sweep_np = np;
        return HL_OK;
    }
    std::copy(dia.begin(), dia.end(), ritz.begin());
    for (int i = 0; i + 1 < ncv; ++i) e[i] = sub[i + 1];
    if (tridiag_eig(ncv, ritz.data(), e.data(), q.data(), ncv)) // dsteqr failure (HLanc.hpp:1692)
        return finish(s, false);
    for (int i = 0; i < ncv; ++i) bounds[i] = rnorm * std::fabs(q[(size_t)i * ncv + ncv - 1]);
    ritz_saved = ritz;
    bounds_saved = bounds;
    const int np0 = ncv - nev0;
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
    if (nconv >= nev0 || np == 0) return finish(s, true);
    const int nevbef = nev; // HLanc.hpp:1719-1730
    nev += std::min(nconv, np / 2);
    if (nev == 1 && ncv >= 6)
        nev = ncv / 2;
    else if (nev == 1 && ncv > 2)
        nev = 2;
    np = ncv - nev;
    if (nevbef < nev) sort_wanted(nev, np, ritz.data(), bounds.data());
    restart_chase(nev, np, ritz.data(), sub.data(), dia.data(), q.data());
    int rc = restart(s);
    if (rc) return rc;
    ++iter;
    restarts = iter;
    if (iter >= maxiter) return finish(s, true);
    sweep_k = nev;
    sweep_np = np;
    return HL_OK;
}

// Ritz vectors and outputs (HLanc.hpp:1762-1843).  vectors_ok is false after a
// tridiagonal-solver failure, which ends the solve without them.
int Member::finish(hipStream_t s, bool vectors_ok)
{
    Workspace &W = ws;
    const double eps23 = std::pow(kEps, 2.0 / 3.0);
    phase = DONE;
    ret = HL_NOT_CONVERGED;
    if (vectors_ok && nconv >= nev0) {
        const int np0 = ncv - nev0;
        std::vector<double> idx(ncv);
        std::vector<char> select(ncv, 0);
        for (int j = 0; j < ncv; ++j) idx[j] = j + 1;
        std::vector<double> rz = ritz_saved;
        sort_wanted(nev0, np0, rz.data(), idx.data());
        int reord = 0, numcnv = 0;
        for (int j = 0; j < ncv; ++j) {
            const double temp1 = std::max(eps23, std::fabs(rz[ncv - j - 1]));
            const int jj = static_cast<int>(idx[ncv - j - 1]);
            if (numcnv < nconv && bounds_saved[jj - 1] <= tol * temp1) {
                select[jj - 1] = 1;
                ++numcnv;
                if (jj > nev0) reord = 1;
            }
        }
        std::vector<double> ihd(dia), ihb(ncv, 0.0), iq((size_t)ncv * ncv), tau(ncv);
        for (int i = 0; i + 1 < ncv; ++i) ihb[i] = sub[i + 1];
        if (numcnv == nconv && !tridiag_eig(ncv, ihd.data(), ihb.data(), iq.data(), ncv)) {
            if (reord) { // selected columns to the front (HLanc.hpp:1808-1831)
                int left = 0, right = ncv - 1;
                for (;;) {
                    while (select[left]) ++left;
                    while (!select[right]) --right;
                    if (left > right) break;
                    std::swap(ihd[left], ihd[right]);
                    for (int i = 0; i < ncv; ++i)
                        std::swap(iq[(size_t)left * ncv + i], iq[(size_t)right * ncv + i]);
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
            qf.assign((size_t)ncv * ncv, 0.0);
            householder_q(ncv, nconv, qs.data(), ncv, tau.data(), qf.data(), ncv);
            if (vectors) {
                HL_TRY_HIP(hipMemcpyAsync(W.Qd, qf.data(), sizeof(double) * (size_t)ncv * ncv,
                                          hipMemcpyHostToDevice, s));
                const int nxt = 1 - W.cur;
                int rc = prof->begin(s);
                if (!rc)
                    rc = launch_vq(n, ncv, ncv, W.V[W.cur], W.vl(), W.Qd, ncv, W.V[nxt], VLayout{W.ldv, false},
                                   opts.use_mfma, -1, nullptr, 0.0, 0.0, W.red, W.ctl, s);
                if (!rc) rc = prof->end(s, K_VQ, 16.0 * n * ncv);
                if (rc) return rc;
                // sync_host (HLanc.hpp:1063-1071): all ncv columns to the caller
                HL_TRY_HIP(hipMemcpy2DAsync(vectors, sizeof(double) * (size_t)ldv_host, W.V[nxt],
                                            sizeof(double) * (size_t)W.ldv, sizeof(double) * (size_t)n, ncv,
                                            hipMemcpyDeviceToHost, s));
                HL_TRY_HIP(hipStreamSynchronize(s));
            }
            std::copy(vals.begin(), vals.end(), values);
            ret = HL_OK;
        }
    }
    info.ret = ret;
    info.restarts = restarts;
    info.nconv = nconv;
    info.spmv = spmv;
    info.cgs_passes = cgs;
    info.rnorm = rnorm;
    info.seconds = now_s() - t0;
    info.max_rel_resid = -1.0;
    info.refine_shift = 0.0;
    return HL_OK;
}

// ------------------------------------------------------------------ Engine

int Engine::init(int dev_)
{
    dev = dev_;
    HL_TRY_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    return HL_OK;
}

void Engine::release()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (dM) (void)hipFree(dM);
    if (dS) (void)hipFree(dS);
    if (hM) (void)hipHostFree(hM);
    if (hS) (void)hipHostFree(hS);
    prof.release();
    if (stream) (void)hipStreamDestroy(stream);
    dM = nullptr;
    dS = nullptr;
    hM = nullptr;
    hS = nullptr;
    stream = nullptr;
    capM = capS = 0;
}

int Engine::grow(size_t nm, size_t ns)
{
    if (nm > capM) {
        HL_TRY_HIP(hipStreamSynchronize(stream));
        if (dM) (void)hipFree(dM);
        if (hM) (void)hipHostFree(hM);
        dM = nullptr;
        hM = nullptr;
        capM = std::max(nm, 2 * capM);
        HL_TRY_HIP(hipMalloc(&dM, sizeof(SegMember) * capM));
        HL_TRY_HIP(hipHostMalloc(&hM, sizeof(SegMember) * capM, hipHostMallocDefault));
    }
    if (ns > capS) {
        HL_TRY_HIP(hipStreamSynchronize(stream));
        if (dS) (void)hipFree(dS);
        if (hS) (void)hipHostFree(hS);
        dS = nullptr;
        hS = nullptr;
        capS = std::max(ns, 2 * capS);
        HL_TRY_HIP(hipMalloc(&dS, sizeof(SegStep) * capS));
        HL_TRY_HIP(hipHostMalloc(&hS, sizeof(SegStep) * capS, hipHostMallocDefault));
    }
    return HL_OK;
}

int Engine::sweep(std::vector<Member *> &ms)
{
    int rc;
    for (Member *m : ms)
        if (m->phase == Member::NEW && (rc = m->start(stream))) return rc;
    std::vector<Member *> act;
    for (Member *m : ms)
        if (m->phase == Member::ACTIVE) act.push_back(m);
    if (act.empty()) return HL_OK;

    // member table and all steps of the sweep, uploaded once
    int maxsteps = 0;
    for (Member *m : act) maxsteps = std::max(maxsteps, m->sweep_np);
    const size_t nm = act.size();
    if ((rc = grow(nm, nm * (size_t)std::max(maxsteps, 1)))) return rc;
    HL_TRY_HIP(hipStreamSynchronize(stream)); // the pinned tables are rewritten every sweep
    for (size_t i = 0; i < nm; ++i) {
        Member &m = *act[i];
        Workspace &W = m.ws;
        SegMember &M = hM[i];
        M.blks = m.A ? m.A->blks : nullptr;
        M.rp = m.A ? m.A->row_ptr : nullptr;
        M.col = m.A ? m.A->col : nullptr;
        M.val = m.A ? m.A->val : nullptr;
        M.chunk_part = W.chunk_part;
        M.chunk_cnt = W.chunk_cnt;
        M.V = W.V[W.cur];
        M.vld = W.vl().ld;
        M.r = W.r;
        M.w = W.w;
        M.h0 = W.hbuf;
        M.h1 = W.hbuf + W.ncv;
        M.hsub = W.hsub;
        M.hdia = W.hdia;
        M.partials = W.red.partials;
        M.ctl = W.ctl;
        M.n = m.n;
        M.nb = cgs_blocks(m.n);
    }
    std::vector<int> off(maxsteps + 1, 0), nspmv(maxsteps, 0), ncgs(maxsteps, 0), nfin(maxsteps, 0),
        maxcpg(maxsteps, 0);
    size_t ns = 0;
    for (int st = 0; st < maxsteps; ++st) {
        off[st] = static_cast<int>(ns);
        for (size_t i = 0; i < nm; ++i) {
            Member &m = *act[i];
            if (st >= m.sweep_np) continue;
            SegStep &S = hS[ns++];
            S.mem = static_cast<int>(i);
            S.j = m.sweep_k + st;
            S.spmv0 = nspmv[st];
            S.cgs0 = ncgs[st];
            S.fin0 = nfin[st];
            if (m.A) nspmv[st] += m.A->nblk;
            ncgs[st] += hM[i].nb;
            nfin[st] += S.j + 2; // j + 1 projections and the squared norm
            maxcpg[st] = std::max(maxcpg[st], (S.j + 1 + dev::kGroups - 1) / dev::kGroups);
        }
    }
    off[maxsteps] = static_cast<int>(ns);
    HL_TRY_HIP(hipMemcpyAsync(dM, hM, sizeof(SegMember) * nm, hipMemcpyHostToDevice, stream));
    HL_TRY_HIP(hipMemcpyAsync(dS, hS, sizeof(SegStep) * ns, hipMemcpyHostToDevice, stream));

    const int modes[4] = {EPI_WNORM, EPI_PASS1, EPI_PASS2, EPI_PASS3};
    for (int st = 0; st < maxsteps; ++st) {
        const SegStep *S = dS + off[st];
        const int cnt = off[st + 1] - off[st];
        double bytes = 0.0;
        for (int q = off[st]; q < off[st + 1]; ++q) {
            Member &m = *act[hS[q].mem];
            if (!m.A && (rc = m.host_apply(hS[q].j, stream))) return rc;
            if (m.A) bytes += 12.0 * m.A->nnz + 4.0 * (m.A->nrows + 1) + 8.0 * m.A->ncols + 16.0 * m.A->nrows;
        }
        if (nspmv[st] > 0) {
            if ((rc = prof.begin(stream))) return rc;
            if ((rc = launch_spmv_seg(dM, S, cnt, nspmv[st], stream))) return rc;
            if ((rc = prof.end(stream, K_SPMV, bytes))) return rc;
        }
        for (int p = 0; p < 4; ++p) {
            bytes = 0.0;
            if (p < 2) // executed DGKS corrections are counted from the device counters
                for (int q = off[st]; q < off[st + 1]; ++q) {
                    const Member &m = *act[hS[q].mem];
                    bytes += m.cgs_bytes_pass(hS[q].j + 1) - (p == 0 ? 8.0 * m.n : 0.0);
                }
            if ((rc = prof.begin(stream))) return rc;
            if ((rc = launch_cgs_seg(dM, S, cnt, maxcpg[st], ncgs[st], nfin[st], modes[p], stream))) return rc;
            if ((rc = prof.end(stream, K_CGS, bytes))) return rc;
        }
    }
    // the new alpha/beta entries and the residual norm of every member
    for (Member *m : act) {
        Workspace &W = m->ws;
        const int k = m->sweep_k, cnt = m->ncv - k;
        HL_TRY_HIP(hipMemcpyAsync(W.hh, W.hsub + k, sizeof(double) * cnt, hipMemcpyDeviceToHost, stream));
        HL_TRY_HIP(hipMemcpyAsync(W.hh + m->ncv, W.hdia + k, sizeof(double) * cnt, hipMemcpyDeviceToHost, stream));
        HL_TRY_HIP(hipMemcpyAsync(W.ctl_h, W.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
    }
    HL_TRY_HIP(hipStreamSynchronize(stream));
    if (prof.on) {
        if ((rc = prof.flush())) return rc;
        for (Member *m : act) {
            const Ctl &c = *m->ws.ctl_h;
            prof.tot.bytes[K_CGS] += 8.0 * m->n *
                                     ((double)(c.m2sum - m->m2sum) + 2.0 * (double)(c.p2 - m->p2) +
                                      (double)(c.m3sum - m->m3sum) + 2.0 * (double)(c.p3 - m->p3));
        }
    }
    for (Member *m : act)
        if ((rc = m->after_sweep(stream))) return rc;
    return HL_OK;
}

int Engine::run(std::vector<Member *> &queue, int max_active, double hbm_budget)
{
    t_start = now_s();
    std::vector<Member *> active;
    size_t next = 0;
    double in_use = 0.0;
    int rc = HL_OK;
    auto need = [](const Member *m) {
        return Workspace::bytes_needed(m->n, m->ncv, m->opts.basis_prec, m->A == nullptr,
                                       m->A ? m->A->nslots : 0);
    };
    while (next < queue.size() || !active.empty()) {
        // admission at the sweep boundary, in queue order
        double t = now_s() - t_start;
        if (active.empty() && next < queue.size() && queue[next]->arrival > t) {
            std::this_thread::sleep_for(std::chrono::duration<double>(queue[next]->arrival - t));
            t = now_s() - t_start;
        }
        while (next < queue.size() && queue[next]->arrival <= t &&
               (max_active <= 0 || (int)active.size() < max_active) &&
               (hbm_budget <= 0.0 || active.empty() || in_use + need(queue[next]) <= hbm_budget)) {
            Member *m = queue[next++];
            m->prof = &prof;
            m->t_admit = t;
            m->phase = Member::NEW;
            if ((m->err = m->ws.init(dev, m->n, m->ncv, m->opts.basis_prec, m->A == nullptr,
                                     m->A ? m->A->nslots : 0))) {
                m->ws.release();
                m->phase = Member::DONE;
                m->t_done = now_s() - t_start;
                continue;
            }
            in_use += need(m);
            active.push_back(m);
        }
        if (active.empty()) continue;
        if ((rc = sweep(active))) break;
        // retire finished members and free their workspaces
        size_t keep = 0;
        for (Member *m : active) {
            if (m->phase == Member::DONE) {
                m->t_done = now_s() - t_start;
                m->ws.release();
                in_use -= need(m);
            } else {
                active[keep++] = m;
            }
        }
        active.resize(keep);
    }
    if (!rc && prof.on) { // the V.Q launches after the last sweep
        HL_TRY_HIP(hipStreamSynchronize(stream));
        rc = prof.flush();
    }
    if (rc) { // a HIP failure ends every member in flight
        (void)hipStreamSynchronize(stream);
        for (Member *m : active) {
            m->err = rc;
            m->ws.release();
            m->phase = Member::DONE;
        }
    }
    return rc;
}

} // namespace hl
