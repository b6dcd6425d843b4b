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
//     needed exits at once);
//   * a step's solves are grouped by precision kind (FP64, the default; FP64
//     with an FP32 copy for the projection passes; FP32 basis), each kind's
//     launches reading its own basis type;
//   * after a sweep one kernel packs every member's alpha/beta and Ctl and one
//     copy brings them home; the members' ncv x ncv work runs on host threads
//     in FP64;
//   * V Q of dsapps for every restarting member is one segmented GEMM launch
//     into the other ping-pong basis (no in-place column shuffling), and the
//     Ritz-vector step is one GEMM per finishing member.
// A solve run alone is an engine with one member; every block of a segmented
// launch computes what it computes in a one-member launch, so a member's
// results are bitwise the same alone and in any batch.  Return values and the
// default tol / restart cap are the reference's.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
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

double Member::cgs_bytes_pass(int m, bool sub) const
{
    if (!sub && kind == SEG_PROJ32) return 8.0 * n + 16.0 * n; // three-term first pass: FP64 column j
    const double b = sub ? (f32_phase() ? 4.0 : 8.0) : vbytes();
    return b * n * m + 16.0 * n;
}

int Member::reset(hipStream_t s)
{
    Workspace &W = ws;
    tol = opts.tol > 0.0 ? opts.tol : kEps;
    if (f32_phase()) tol = std::max(tol, std::pow(10.0, -tuning().f32_phase_tol_exp));
    // the restart cap covers both phases of a two-phase solve
    maxiter = std::max(1, (opts.maxiter > 0 ? opts.maxiter : 200) - restarts_f32);
    sub.assign(ncv, 0.0);
    dia.assign(ncv, 0.0);
    ritz.assign(ncv, 0.0);
    bounds.assign(ncv, 0.0);
    zlast.assign(ncv, 0.0);
    q.assign((size_t)ncv * ncv, 0.0);
    e.assign(ncv, 0.0);
    work.assign(ncv, 0.0);
    ritz_saved.clear();
    bounds_saved.clear();
    spmv = cgs = p2 = p3 = m2sum = m3sum = nexact = mxsum = 0;
    dgks_bytes = 0.0;
    nconv = 0;
    rnorm = 0.0;
    ret = HL_NOT_CONVERGED;
    act = ACT_NONE;
    std::memset(W.ctl_h, 0, sizeof(Ctl));
    W.ctl_h->rscale = 1.0;
    W.ctl_h->xscale = 1.0;
    // FP32-copy projections: the first correction is decided from its exact
    // coefficients (dgks_decide_seg_kernel), not by the 0.717 cancellation test
    W.ctl_h->dgks_always = kind == SEG_PROJ32;
    HL_TRY_HIP(hipMemcpyAsync(W.ctl, W.ctl_h, sizeof(Ctl), hipMemcpyHostToDevice, s));
    HL_TRY_HIP(hipMemsetAsync(W.hsub, 0, sizeof(double) * ncv, s));
    HL_TRY_HIP(hipMemsetAsync(W.hdia, 0, sizeof(double) * ncv, s));
    return HL_OK;
}

int Member::begin(hipStream_t s)
{
    Workspace &W = ws;
    int rc;
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

int Member::start(hipStream_t s)
{
    Workspace &W = ws;
    t0 = now_s();
    mixed = opts.basis_prec == HL_PREC_F32 && A != nullptr;
    kind = mixed ? (A->val32 ? SEG_V32_A32 : SEG_V32) : copy32() ? SEG_PROJ32 : SEG_F64;
    spmv_f32 = 0;
    restarts_f32 = 0;
    vals_f32.clear();
    graph_sweeps = 0;
    int rc;
    if ((rc = reset(s))) return rc;
    if (opts.v0) {
        HL_TRY_HIP(hipMemcpyAsync(W.r, opts.v0, sizeof(double) * n, hipMemcpyHostToDevice, s));
    } else if ((rc = launch_start_vector(n, opts.seed, W.r, s))) { // dgetv0 (HLanc.hpp:1368-1374)
        return rc;
    }
    return begin(s);
}

// End of the FP32 phase: the engine has formed ws.r = V q (rephase_host's q);
// the solve starts over in FP64 from it.
int Member::spmv_blocks() const
{
    if (!A) return 0;
    if (sell()) return (A->nchunks + kBlock / kWave - 1) / (kBlock / kWave); // a wave per chunk
    if (staged()) return 0;                                                    // kernels of their own
    return sliced() ? A->sitems.nblk : A->items.nblk;
}

int Member::rephase(hipStream_t s)
{
    spmv_f32 = spmv;
    restarts_f32 = restarts;
    kind = SEG_F64;
    // the FP64 passes read whole tiles: clear what the FP32 phase left in the basis
    const size_t vbytes = 8 * static_cast<size_t>(ws.ldv) * static_cast<size_t>(ws.vl().ld);
    HL_TRY_HIP(hipMemsetAsync(ws.V[0], 0, vbytes, s));
    HL_TRY_HIP(hipMemsetAsync(ws.V[1], 0, vbytes, s));
    int rc;
    if ((rc = reset(s))) return rc;
    return begin(s);
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

// The read-back of a sweep: entries [sweep_k, ncv) of the tridiagonal (the rest
// is the host's own, as the restart left it) and the device counters.
void Member::absorb(const double *rb)
{
    for (int i = sweep_k; i < ncv; ++i) {
        sub[i] = rb[i];
        dia[i] = rb[ncv + i];
    }
    Ctl c;
    std::memcpy(&c, rb + 2 * ncv, sizeof(Ctl));
    rnorm = c.rnorm;
    cgs = c.cgs;
    // basis columns read by the corrections (plus w read and written): first
    // corrections on the projected basis (FP64 for exact steps), second ones
    // on the FP64 basis for FP32-copy projections
    const double c2 = (double)(c.m2sum - m2sum) + 2.0 * (double)(c.p2 - p2);
    const double c3 = (double)(c.m3sum - m3sum) + 2.0 * (double)(c.p3 - p3);
    const double cx = (double)(c.mxsum - mxsum) + 2.0 * (double)(c.nexact - nexact);
    if (kind == SEG_PROJ32)
        dgks_bytes = n * (4.0 * c2 + 4.0 * cx + 8.0 * c3);
    else
        dgks_bytes = vbytes() * n * (c2 + c3);
    p2 = c.p2;
    p3 = c.p3;
    m2sum = c.m2sum;
    m3sum = c.m3sum;
    nexact = c.nexact;
    mxsum = c.mxsum;
    spmv += sweep_np;
}

// The rest of ds_solver after a sweep (HLanc.hpp:1687-1736): Ritz values and
// bounds, convergence, nev adaptation, exact-shift chase; or the end.
void Member::decide(double *qslot)
{
    act = ACT_NONE;
    if (stage == 0) { // the initial dsaitr(0, nev0) is followed directly by the main loop
        stage = 1;
        restarts = iter;
        sweep_k = nev;
        sweep_np = np;
        act = ACT_SWEEP;
        return;
    }
    std::copy(dia.begin(), dia.end(), ritz.begin());
    for (int i = 0; i + 1 < ncv; ++i) e[i] = sub[i + 1];
    // dsteqr("I") (HLanc.hpp:1691); only Q's last row is used here (1694-1696)
    if (tridiag_eig_last(ncv, ritz.data(), e.data(), zlast.data())) {
        finish_host(qslot, false); // dsteqr failure (HLanc.hpp:1692)
        return;
    }
    for (int i = 0; i < ncv; ++i) bounds[i] = rnorm * std::fabs(zlast[i]);
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
    if (nconv >= nev0 || np == 0) {
        if (f32_phase())
            rephase_host(qslot);
        else
            finish_host(qslot, true);
        return;
    }
    const int nevbef = nev; // HLanc.hpp:1719-1730
    nev += std::min(nconv, np / 2);
    if (nev == 1 && ncv >= 6)
        nev = ncv / 2;
    else if (nev == 1 && ncv > 2)
        nev = 2;
    np = ncv - nev;
    if (nevbef < nev) sort_wanted(nev, np, ritz.data(), bounds.data());
    if (f32_phase() && iter + 1 >= maxiter) { // the FP32 phase's last restart: hand over what
        ++iter;                                 // T and V hold now (before a chase rotates T)
        restarts = iter;
        rephase_host(qslot);
        return;
    }
    restart_chase(nev, np, ritz.data(), sub.data(), dia.data(), q.data());
    // device half of dsapps (HLanc.hpp:1591-1614): V := V Q(:, 0:nev) and
    // resid := Q(ncv-1, nev-1) resid + beta V Q(:, nev)
    const bool has_beta = sub[nev] > 0.0;
    vq_N = has_beta ? nev + 1 : nev;
    vq_res = has_beta ? nev : -1;
    vq_beta = has_beta ? sub[nev] : 0.0;
    vq_qlast = q[(size_t)(nev - 1) * ncv + ncv - 1];
    std::copy(q.begin(), q.begin() + (size_t)ncv * vq_N, qslot);
    ++iter;
    restarts = iter;
    if (iter >= maxiter) { // the restart would feed no further sweep: the solve ends unconverged
        finish_host(qslot, true);
        return;
    }
    sweep_k = nev;
    sweep_np = np;
    act = ACT_RESTART;
}

// Ritz values and the Ritz-vector Q (HLanc.hpp:1762-1840) into qslot (ncv x
// ncv), or the failure return.  vectors_ok is false after a tridiagonal-solver
// failure, which ends the solve without them.
void Member::finish_host(double *qslot, bool vectors_ok)
{
    act = ACT_FINISH;
    with_vectors = need_ritz = verify = false;
    ret = HL_NOT_CONVERGED;
    if (!vectors_ok || nconv < nev0) return; // HLanc.hpp:1762
    const double eps23 = std::pow(kEps, 2.0 / 3.0);
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
    if (numcnv != nconv) return; // HLanc.hpp:1791
    std::vector<double> ihd(dia), ihb(ncv, 0.0), iq((size_t)ncv * ncv), tau(ncv);
    for (int i = 0; i + 1 < ncv; ++i) ihb[i] = sub[i + 1];
    if (tridiag_eig(ncv, ihd.data(), ihb.data(), iq.data(), ncv)) return; // HLanc.hpp:1805-1806
    anorm = 0.0; // ||A||_est for the residual check: the extreme Ritz values approach ||A||_2 from below
    for (int i = 0; i < ncv; ++i) anorm = std::max(anorm, std::fabs(ihd[i]));
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
    vals.assign(nconv, 0.0);
    std::vector<double> qs((size_t)ncv * nconv);
    for (int i = 0; i < nconv; ++i) {
        vals[i] = ihd[order[i]];
        std::copy(iq.begin() + (size_t)order[i] * ncv, iq.begin() + (size_t)(order[i] + 1) * ncv,
                  qs.begin() + (size_t)i * ncv);
    }
    // dgeqr2 + dorm2r("R","N") (HLanc.hpp:1837-1840) as one V Q_full GEMM
    householder_q(ncv, nconv, qs.data(), ncv, tau.data(), qslot, ncv);
    with_vectors = vectors != nullptr;
    verify = verifies() && anorm > 0.0;
    need_ritz = with_vectors || verify;
    ret = HL_OK;
}

// The FP64 phase's start vector: the sum of the FP32 phase's nev0 wanted
// (largest algebraic) Ritz vectors, V z_i for the eigenvectors z_i of its
// current T; qslot receives q = sum z_i (ncv entries).
void Member::rephase_host(double *qslot)
{
    act = ACT_REPHASE;
    std::vector<double> d(dia), b(ncv, 0.0), z((size_t)ncv * ncv);
    for (int i = 0; i + 1 < ncv; ++i) b[i] = sub[i + 1];
    std::fill(qslot, qslot + ncv, 0.0);
    vals_f32.clear();
    if (tridiag_eig(ncv, d.data(), b.data(), z.data(), ncv)) { // keep the last Lanczos vector's direction
        qslot[ncv - 1] = 1.0;
        return;
    }
    for (int c = ncv - nev0; c < ncv; ++c) { // ascending eigenvalues: the last nev0 are wanted
        vals_f32.push_back(d[c]);
        for (int i = 0; i < ncv; ++i) qslot[i] += z[(size_t)c * ncv + i];
    }
}

// The residual check's sums (block partials added in block order) for pair i:
// ||A v - theta v||^2, v^T (A v - theta v) and ||v||^2.
void Member::check_sums(int i, double &a, double &p, double &c) const
{
    const int G = res_blocks(n);
    const double *o = ws.vres_h + (size_t)3 * G * i;
    a = p = c = 0.0;
    for (int b = 0; b < G; ++b) {
        a += o[3 * b];
        p += o[3 * b + 1];
        c += o[3 * b + 2];
    }
}

void Member::complete()
{
    info.max_rel_resid = -1.0;
    info.anorm_est = 0.0;
    info.rq_shift = 0.0;
    info.verified = 0;
    if (ret == HL_OK && verify) {
        // true residuals of the returned pairs; with HL_REFINE_RQ the value moves
        // to the Rayleigh quotient theta + v^T d / v^T v (d = A v - theta v), where
        // ||A v - rq v||^2 = ||d||^2 - (v^T d)^2 / v^T v
        const bool rq = opts.refine == HL_REFINE_RQ;
        double worst = 0.0, shift = 0.0;
        for (int i = 0; i < nconv; ++i) {
            double a, p, c;
            check_sums(i, a, p, c);
            const double delta = c > 0.0 ? p / c : 0.0;
            const double r2 = rq ? std::max(a - delta * p, 0.0) : a;
            const double rel = std::sqrt(r2) / (anorm * std::sqrt(c > 0.0 ? c : 1.0));
            shift = std::max(shift, std::fabs(delta) / std::max(std::fabs(vals[i]), 1e-300));
            if (rq) vals[i] += delta;
            worst = std::max(worst, rel);
            if (opts.residuals) opts.residuals[i] = rel;
        }
        info.max_rel_resid = worst;
        info.anorm_est = anorm;
        info.rq_shift = shift;
        info.verified = nconv;
    }
    if (ret == HL_OK) std::copy(vals.begin(), vals.end(), values);
    info.ret = ret;
    info.restarts = restarts_f32 + restarts;
    info.nconv = nconv;
    info.spmv = spmv_f32 + spmv;
    info.cgs_passes = cgs;
    info.rnorm = rnorm;
    info.seconds = now_s() - t0;
    info.refine_shift = 0.0;
    if (ret == HL_OK && mixed) // how far the FP32 phase's values were from the result
        for (size_t i = 0; i < vals_f32.size() && i < vals.size(); ++i)
            info.refine_shift = std::max(info.refine_shift, std::fabs(vals_f32[i] - vals[i]) /
                                                                std::max(std::fabs(vals[i]), 1e-300));
    info.spmv_f32 = spmv_f32;
    info.restarts_f32 = restarts_f32;
    info.proj64_steps = nexact;
    info.graph_sweeps = graph_sweeps;
    phase = DONE;
}

// ------------------------------------------------------------------ Engine

// Host threads for the members' restart logic: a parallel loop over members,
// the calling thread taking part.
struct Engine::Pool {
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable go, done;
    std::function<void(int)> job;
    int count = 0, busy = 0;
    long gen = 0;
    bool stop = false;
    std::atomic<int> next{0};

    explicit Pool(int nt)
    {
        for (int i = 0; i < nt; ++i) th.emplace_back([this] { work(); });
    }
    ~Pool()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        go.notify_all();
        for (std::thread &t : th) t.join();
    }
    void work()
    {
        long seen = 0;
        for (;;) {
            std::function<void(int)> f;
            int cnt;
            {
                std::unique_lock<std::mutex> lk(mu);
                go.wait(lk, [&] { return stop || gen != seen; });
                if (stop) return;
                seen = gen;
                f = job;
                cnt = count;
                ++busy;
            }
            for (int i; (i = next.fetch_add(1)) < cnt;) f(i);
            std::lock_guard<std::mutex> lk(mu);
            if (--busy == 0) done.notify_all();
        }
    }
    void run(int n, const std::function<void(int)> &f)
    {
        if (n < 16 || th.empty()) {
            for (int i = 0; i < n; ++i) f(i);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            job = f;
            count = n;
            next.store(0);
            ++gen;
        }
        go.notify_all();
        for (int i; (i = next.fetch_add(1)) < n;) f(i);
        std::unique_lock<std::mutex> lk(mu);
        done.wait(lk, [&] { return busy == 0; });
    }
};

int Engine::init(int dev_, int host_threads)
{
    dev = dev_;
    auto lk = hip_lock(); // stream and allocations: not while another engine captures
    HL_TRY_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    Ctl one{};
    one.xscale = one.rscale = 1.0;
    HL_TRY_HIP(hipMalloc(&dOne, sizeof(Ctl)));
    HL_TRY_HIP(hipMemcpy(dOne, &one, sizeof(Ctl), hipMemcpyHostToDevice));
    const unsigned hw = std::thread::hardware_concurrency();
    pool = new Pool(host_threads >= 0 ? host_threads : static_cast<int>(std::min(15u, hw > 1 ? hw - 1 : 0u)));
    return HL_OK;
}

void Engine::drop_graphs()
{
    for (GraphEntry &g : gcache) (void)hipGraphExecDestroy(g.exec);
    gcache.clear();
    gseen.clear();
}

void Engine::release()
{
    if (stream) (void)sync_stream();
    bury(); // takes the lock itself
    auto lk = hip_lock();
    for (Pending &p : pend) (void)hipEventDestroy(p.ev);
    pend.clear();
    drop_graphs();
    void *d[] = {dM, dS, dV, dRB, dQ, dMv, dSv, dRv, dOne};
    for (void *p : d)
        if (p) (void)hipFree(p);
    void *h[] = {hMs[0], hMs[1], hSs[0], hSs[1], hVs[0], hVs[1], hRB, hQs[0], hQs[1], hMv, hSv, hRv};
    for (void *p : h)
        if (p) (void)hipHostFree(p);
    dMv = hMv = nullptr;
    dSv = hSv = nullptr;
    dRv = hRv = nullptr;
    dOne = nullptr;
    capMv = capSv = capRv = 0;
    dM = hM = hMs[0] = hMs[1] = nullptr;
    dS = hS = hSs[0] = hSs[1] = nullptr;
    dV = hV = hVs[0] = hVs[1] = nullptr;
    dRB = hRB = dQ = hQ = hQs[0] = hQs[1] = nullptr;
    capM = capS = capV = capRB = capQ = 0;
    prof.release();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    delete pool;
    pool = nullptr;
}

namespace {
// A device table and its `nh` pinned staging copies, grown to `need` entries.
template <typename T>
int grow_set(T *&d, T **h, int nh, size_t &cap, size_t need)
{
    if (need <= cap) return HL_OK;
    if (d) (void)hipFree(d);
    d = nullptr;
    for (int i = 0; i < nh; ++i) {
        if (h[i]) (void)hipHostFree(h[i]);
        h[i] = nullptr;
    }
    cap = std::max(need, 2 * cap);
    HL_TRY_HIP(hipMalloc(&d, sizeof(T) * cap));
    for (int i = 0; i < nh; ++i) HL_TRY_HIP(hipHostMalloc(&h[i], sizeof(T) * cap, hipHostMallocDefault));
    return HL_OK;
}
} // namespace

int Engine::grow(size_t nm, size_t ns, size_t nv, size_t rb, size_t qd)
{
    std::unique_lock<std::mutex> lk;
    if (nm > capM || ns > capS || nv > capV || rb > capRB || qd > capQ) {
        lk = hip_lock();
        HL_TRY_HIP(sync_stream()); // uploads from the old buffers may be in flight
        if (nm > capM || ns > capS) drop_graphs(); // captured launches name the old tables
    }
    int rc;
    if ((rc = grow_set(dM, hMs, 2, capM, nm))) return rc;
    if ((rc = grow_set(dS, hSs, 2, capS, ns))) return rc;
    if ((rc = grow_set(dV, hVs, 2, capV, nv))) return rc;
    if ((rc = grow_set(dRB, &hRB, 1, capRB, rb))) return rc;
    if ((rc = grow_set(dQ, hQs, 2, capQ, qd))) return rc;
    // the other staging set: its last uploads were ordered before the previous
    // sweep's read-back, which the host has waited for
    par ^= 1;
    hM = hMs[par];
    hS = hSs[par];
    hV = hVs[par];
    hQ = hQs[par];
    return HL_OK;
}

void Engine::bury()
{
    if (graveyard.empty()) return;
    auto lk = hip_lock();
    for (auto &p : graveyard) {
        if (p.first) (void)hipFree(p.first);
        if (p.second) (void)hipHostFree(p.second);
    }
    graveyard.clear();
}

int Engine::poll_finished(bool wait)
{
    size_t keep = 0;
    for (size_t k = 0; k < pend.size(); ++k) {
        Pending &p = pend[k];
        hipError_t e;
        {
            auto wl = wait_lock();
            e = wait ? hipEventSynchronize(p.ev) : hipEventQuery(p.ev);
        }
        if (e == hipErrorNotReady) {
            pend[keep++] = p;
            continue;
        }
        {
            auto lk = hip_lock();
            (void)hipEventDestroy(p.ev);
        }
        if (e != hipSuccess) {
            for (size_t r = k + 1; r < pend.size(); ++r) pend[keep++] = pend[r];
            pend.resize(keep);
            return hip_status(e, "Ritz-vector copy");
        }
        for (Member *m : p.ms) m->complete();
    }
    pend.resize(keep);
    return HL_OK;
}

namespace {
// The operator side of a member's SpMV entry: its current form (items over the
// rows or the virtual rows, sliced-ELL chunks) and its long-row slots.
void spmv_fields(SegMember &M, const Member &m)
{
    const Workspace &W = m.ws;
    const bool sl = m.sliced(), sg = m.staged();
    M.blks = !m.A ? nullptr : sl ? m.A->sitems.blks : m.A->items.blks;
    M.rp = !m.A ? nullptr : sl ? m.A->s_rp : m.A->row_ptr;
    M.col = !m.A ? nullptr : sl ? m.A->s_col : m.A->col;
    M.val = !m.A ? nullptr : sl ? m.A->s_val : m.A->val;
    M.val32 = m.A ? m.A->val32 : nullptr;
    M.part = sl || sg ? W.part : nullptr;
    M.r_pp = sl ? m.A->r_pp : nullptr;
    M.r_pairs = sl ? m.A->r_pairs : nullptr;
    M.chunk_part = W.chunk_part;
    M.chunk_cnt = W.chunk_cnt;
    const bool el = m.sell();
    M.e_ptr = el ? m.A->e_ptr : nullptr;
    M.e_col = el ? m.A->e_col : nullptr;
    M.e_d16 = el ? m.A->e_d16 : nullptr;
    M.e_val = el ? m.A->e_val : nullptr;
    M.nchunks = el ? m.A->nchunks : 0;
    M.s_sup = sg ? m.A->s_sup : nullptr;
    M.s_cptr = sg ? m.A->s_cptr : nullptr;
    M.s_plen = sg ? m.A->s_plen : nullptr;
    M.s_ecol = sg ? m.A->s_ecol : nullptr;
    M.s_lst = sg ? m.A->s_lst : nullptr;
    M.s_eval = sg ? m.A->s_eval : nullptr;
    M.s_bmeta = sg ? m.A->s_bmeta : nullptr;
    M.s_run = sg ? m.A->s_run : nullptr;
    M.s_lpp = sg ? m.A->s_lpp : nullptr;
    M.stg_nsl = sg ? m.A->stg_nsl : 0;
}
} // namespace

int Engine::sweep(std::vector<Member *> &ms)
{
    int rc;
    const double th0 = now_s();
    for (Member *m : ms)
        if (m->phase == Member::NEW) { // may copy a caller's pageable v0: not while another engine captures
            auto wl = wait_lock();
            if ((rc = m->start(stream))) return rc;
        }
    std::vector<Member *> act;
    for (Member *m : ms)
        if (m->phase == Member::ACTIVE) act.push_back(m);
    if (act.empty()) return HL_OK;

    // member table and all steps of the sweep, uploaded once
    int maxsteps = 0, ncvmax = 0;
    for (Member *m : act) {
        maxsteps = std::max(maxsteps, m->sweep_np);
        ncvmax = std::max(ncvmax, m->ncv);
    }
    const size_t nm = act.size();
    const int stride = 2 * ncvmax + kCtlWords;
    if ((rc = grow(nm, nm * (size_t)std::max(maxsteps, 1), nm, nm * (size_t)stride, nm * (size_t)ncvmax * ncvmax)))
        return rc;
    for (size_t i = 0; i < nm; ++i) {
        Member &m = *act[i];
        Workspace &W = m.ws;
        SegMember &M = hM[i];
        spmv_fields(M, m);
        M.V = W.V[W.cur];
        M.V32 = m.kind == SEG_PROJ32 ? W.V32[W.cur] : nullptr;
        M.Vproj = M.V32 ? static_cast<const void *>(M.V32) : M.V;
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
        M.ncv = m.ncv;
    }
    // every step index's entries grouped by precision kind (kSegOrder): one SpMV
    // launch per value/basis type, one CGS partials launch per basis type read,
    // one finalize launch over all
    auto spmv_group = [](int k) { return k == SEG_V32 ? 0 : k == SEG_V32_A32 ? 1 : 2; };
    auto proj_group = [](int k) { return k == SEG_F64 ? 1 : 0; };            // 0: FP32 reads
    auto sub_group = [](int k) { return k == SEG_F64 || k == SEG_PROJ32 ? 1 : 0; };
    static const int kSpmvKind[3] = {SEG_V32, SEG_V32_A32, SEG_F64};
    std::vector<std::array<int, 4>> goff_spmv(maxsteps), goff_proj(maxsteps), goff_sub(maxsteps);
    std::vector<std::array<int, 3>> nspmv(maxsteps);
    std::vector<std::array<int, 2>> nproj(maxsteps), nsub(maxsteps), cpg_proj(maxsteps), cpg_sub(maxsteps);
    std::vector<int> off(maxsteps + 1, 0), nfin(maxsteps, 0), n32(maxsteps, 0), nb32(maxsteps, 0), cpg32(maxsteps, 0),
        nred(maxsteps, 0), nstg(maxsteps, 0), nbin(maxsteps, 0), nslx(maxsteps, 0);
    size_t ns = 0;
    for (int st = 0; st < maxsteps; ++st) {
        off[st] = static_cast<int>(ns);
        nspmv[st] = {0, 0, 0};
        nproj[st] = nsub[st] = cpg_proj[st] = cpg_sub[st] = {0, 0};
        int cur_sp = -1, cur_pr = -1, cur_sb = -1;
        for (int ko = 0; ko < SEG_NKIND; ++ko) {
            const int k = kSegOrder[ko];
            for (size_t i = 0; i < nm; ++i) {
                Member &m = *act[i];
                if (st >= m.sweep_np || m.kind != k) continue;
                // group boundaries: every group is a contiguous run of entries
                const int gs = spmv_group(k), gp = proj_group(k), gb = sub_group(k);
                for (; cur_sp < gs; ++cur_sp) goff_spmv[st][cur_sp + 1] = static_cast<int>(ns);
                for (; cur_pr < gp; ++cur_pr) goff_proj[st][cur_pr + 1] = static_cast<int>(ns);
                for (; cur_sb < gb; ++cur_sb) goff_sub[st][cur_sb + 1] = static_cast<int>(ns);
                SegStep &S = hS[ns++];
                S.mem = static_cast<int>(i);
                S.j = m.sweep_k + st;
                S.spmv0 = nspmv[st][gs];
                S.cgs0 = nproj[st][gp];
                S.cgs0s = nsub[st][gb];
                S.fin0 = nfin[st];
                S.red0 = nred[st];
                S.stg0 = nstg[st];
                S.bin0 = nbin[st];
                nspmv[st][gs] += m.spmv_blocks();
                if (m.sliced()) nred[st] += (m.n + kRowsPerSumBlock - 1) / kRowsPerSumBlock;
                if (m.staged()) {
                    nstg[st] += m.A->stg_nsup;
                    nbin[st] += m.A->stg_nbins;
                    nslx[st] = std::max(nslx[st], m.A->stg_nsl);
                }
                nproj[st][gp] += hM[i].nb;
                nsub[st][gb] += hM[i].nb;
                nfin[st] += S.j + 2; // j + 1 projections and the squared norm
                const int cpg = (S.j + 1 + dev::kGroups - 1) / dev::kGroups; // columns per lane group
                // PROJ32 steps in the FP32 launches: the first pass is a column dot and a
                // step wider than kNarrowCols corrects on the FP64 basis, so neither sets the width
                const bool f32_wide = gp == 0 && k == SEG_PROJ32 && S.j + 1 > kNarrowCols;
                cpg_proj[st][gp] = std::max(cpg_proj[st][gp], f32_wide ? 1 : cpg);
                cpg_sub[st][gb] = std::max(cpg_sub[st][gb], cpg);
                if (k == SEG_PROJ32) { // leads its cgs0s group: the exact corrections launch over these alone
                    ++n32[st];
                    nb32[st] += hM[i].nb;
                    cpg32[st] = std::max(cpg32[st], cpg);
                }
            }
        }
        for (; cur_sp < 3; ++cur_sp) goff_spmv[st][cur_sp + 1] = static_cast<int>(ns);
        for (; cur_pr < 2; ++cur_pr) goff_proj[st][cur_pr + 1] = static_cast<int>(ns);
        for (; cur_sb < 2; ++cur_sb) goff_sub[st][cur_sb + 1] = static_cast<int>(ns);
    }
    off[maxsteps] = static_cast<int>(ns);
    HL_TRY_HIP(hipMemcpyAsync(dM, hM, sizeof(SegMember) * nm, hipMemcpyHostToDevice, stream));
    HL_TRY_HIP(hipMemcpyAsync(dS, hS, sizeof(SegStep) * ns, hipMemcpyHostToDevice, stream));
    double host_s = now_s() - th0; // host time of this sweep (profiled runs report it)

    const int modes[4] = {EPI_WNORM, EPI_PASS1, EPI_PASS2, EPI_PASS3};
    bool any_host = false, any_proj32 = false;
    for (Member *m : act) {
        any_host |= (m->A == nullptr);
        any_proj32 |= m->kind == SEG_PROJ32;
    }
    const bool solo = nm == 1; // every entry's member is entry 0 of the table
    auto launch_steps = [&]() -> int {
        int r;
        for (int st = 0; st < maxsteps; ++st) {
            const SegStep *S = dS + off[st];
            const int cnt = off[st + 1] - off[st];
            double bytes = 0.0;
            for (int q = off[st]; q < off[st + 1]; ++q) {
                Member &m = *act[hS[q].mem];
                if (!m.A) { // waits on the stream inside: not while another engine captures
                    auto wl = wait_lock();
                    if ((r = m.host_apply(hS[q].j, stream))) return r;
                }
                if (m.A) // values + indices, row pointers, x, y, the basis column (and its FP32 copy)
                    bytes += (m.kind == SEG_V32_A32 ? 8.0 : 12.0) * m.A->nnz + 4.0 * (m.A->nrows + 1) +
                             8.0 * m.A->ncols + 8.0 * m.A->nrows +
                             (m.kind == SEG_PROJ32 ? 12.0 : m.vbytes()) * m.A->nrows;
            }
            if (nspmv[st][0] + nspmv[st][1] + nspmv[st][2] + nstg[st] > 0) {
                if ((r = prof.begin(stream))) return r;
                for (int g = 0; g < 3; ++g)
                    if (nspmv[st][g] > 0 &&
                        (r = launch_spmv_seg(dM, dS + goff_spmv[st][g], goff_spmv[st][g + 1] - goff_spmv[st][g],
                                             nspmv[st][g], kSpmvKind[g], solo, stream)))
                        return r;
                if (nred[st] > 0 && // row sums of the sliced SpMVs (FP64 bases: the last group)
                    (r = launch_spmv_rows_seg(dM, dS + goff_spmv[st][2], goff_spmv[st][3] - goff_spmv[st][2],
                                              nred[st], solo, stream)))
                    return r;
                if (nstg[st] > 0 && // staged-slice SpMVs (FP64 bases: the last group)
                    (r = launch_spmv_staged_seg(dM, dS + goff_spmv[st][2], goff_spmv[st][3] - goff_spmv[st][2],
                                                nstg[st], nbin[st], nslx[st], solo, stream)))
                    return r;
                if ((r = prof.end(stream, K_SPMV, bytes))) return r;
            }
            for (int p = 0; p < 4; ++p) {
                // EPI_PASS1 and EPI_PASS3 read the FP64 basis of PROJ32 solves
                const bool sub = modes[p] == EPI_PASS1 || modes[p] == EPI_PASS3;
                bytes = 0.0;
                if (p < 2) // executed DGKS corrections are counted from the device counters
                    for (int q = off[st]; q < off[st + 1]; ++q) {
                        const Member &m = *act[hS[q].mem];
                        bytes += m.cgs_bytes_pass(hS[q].j + 1, sub) - (p == 0 ? 8.0 * m.n : 0.0);
                    }
                const auto &go = sub ? goff_sub[st] : goff_proj[st];
                const auto &nb = sub ? nsub[st] : nproj[st];
                const auto &cp = sub ? cpg_sub[st] : cpg_proj[st];
                if ((r = prof.begin(stream))) return r;
                for (int g = 0; g < 2; ++g) // group 0 reads FP32 bases, group 1 FP64
                    if (nb[g] > 0 && (r = launch_cgs_seg(dM, dS + go[g], go[g + 1] - go[g], cp[g], nb[g], modes[p],
                                                         g == 0, sub, solo, stream)))
                        return r;
                if (modes[p] == EPI_PASS2 && n32[st] > 0 && // strongly cancelling PROJ32 steps, on FP64
                    (r = launch_cgs_seg(dM, dS + goff_sub[st][1], n32[st], cpg32[st], nb32[st], modes[p], false, true,
                                        solo, stream)))
                    return r;
                if ((r = prof.end(stream, p < 2 ? K_CGS : K_DGKS, bytes))) return r;
                if ((r = prof.begin(stream))) return r;
                if ((r = launch_cgs_seg_finalize(dM, S, cnt, nfin[st], modes[p], solo, stream))) return r;
                if (modes[p] == EPI_PASS1 && any_proj32 && tuning().dgks_skip_exp > 0 &&
                    (r = launch_dgks_decide_seg(dM, S, cnt, std::pow(10.0, -tuning().dgks_skip_exp), solo, stream)))
                    return r;
                if ((r = prof.end(stream, K_MISC, 0.0))) return r;
            }
        }
        return HL_OK;
    };
    if (!graphs || prof.on || any_host) {
        if ((rc = launch_steps())) return rc;
    } else {
        // the sweep's shape: table addresses and every step's segment and grid sizes
        std::vector<long long> key;
        key.reserve(2 + 28 * (size_t)maxsteps);
        key.push_back(reinterpret_cast<long long>(dM));
        key.push_back(reinterpret_cast<long long>(dS));
        key.push_back(solo ? 1 : 0);
        for (int st = 0; st < maxsteps; ++st) {
            key.push_back(off[st]);
            for (int g = 0; g < 4; ++g) {
                key.push_back(goff_spmv[st][g]);
                if (g < 3) key.push_back(goff_proj[st][g]);
                if (g < 3) key.push_back(goff_sub[st][g]);
            }
            for (int g = 0; g < 3; ++g) key.push_back(nspmv[st][g]);
            for (int g = 0; g < 2; ++g) {
                key.push_back(nproj[st][g]);
                key.push_back(nsub[st][g]);
                key.push_back(cpg_proj[st][g]);
                key.push_back(cpg_sub[st][g]);
            }
            key.push_back(nfin[st]);
            key.push_back(nred[st]);
            key.push_back(nstg[st]);
            key.push_back(nbin[st]);
            key.push_back(nslx[st]);
            key.push_back(n32[st]);
            key.push_back(nb32[st]);
            key.push_back(cpg32[st]);
        }
        ++gclock;
        GraphEntry *hit = nullptr;
        for (GraphEntry &g : gcache)
            if (g.key == key) hit = &g;
        if (!hit && std::find(gseen.begin(), gseen.end(), key) != gseen.end()) {
            // second sighting of this shape: capture it once, replay from now on
            hipGraph_t graph = nullptr;
            std::unique_lock<std::mutex> cap_lk;
            if (capture_mu) cap_lk = std::unique_lock<std::mutex>(*capture_mu);
            std::unique_lock<std::shared_mutex> cap_wl; // no other engine waits meanwhile
            if (sync_rw) cap_wl = std::unique_lock<std::shared_mutex>(*sync_rw);
            HL_TRY_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            rc = launch_steps();
            hipError_t ce = hipStreamEndCapture(stream, &graph);
            // the lock stays held through instantiation and eviction (both allocate
            // or free device memory) and ends with this block
            const int inval = HL_ERR_HIP_BASE + static_cast<int>(hipErrorStreamCaptureInvalidated);
            if (rc == inval || ce == hipErrorStreamCaptureInvalidated) {
                // another host thread's HIP call overlapped the capture: nothing of it
                // ran, so launch the sweep directly, and capture no more in this engine
                if (graph) (void)hipGraphDestroy(graph);
                (void)hipGetLastError();
                graphs = false;
                if ((rc = launch_steps())) return rc;
                goto readback;
            }
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ce != hipSuccess) return hip_status(ce, "hipStreamEndCapture");
            hipGraphExec_t exec = nullptr;
            ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ce != hipSuccess) return hip_status(ce, "hipGraphInstantiate");
            constexpr size_t kMaxGraphs = 16;
            if (gcache.size() >= kMaxGraphs) { // evict the least recently used
                auto lru = std::min_element(gcache.begin(), gcache.end(),
                                            [](const GraphEntry &a, const GraphEntry &b) { return a.last < b.last; });
                (void)hipGraphExecDestroy(lru->exec);
                gcache.erase(lru);
            }
            gcache.push_back(GraphEntry{key, exec, gclock});
            hit = &gcache.back();
            ++graph_captures;
        }
        if (hit) {
            hit->last = gclock;
            HL_TRY_HIP(hipGraphLaunch(hit->exec, stream));
            ++graph_launches;
            for (Member *m : act) ++m->graph_sweeps;
        } else {
            if (gseen.size() >= 64) gseen.erase(gseen.begin());
            gseen.push_back(key);
            if ((rc = launch_steps())) return rc;
        }
    }
readback:
    // every member's alpha/beta entries and counters in one copy
    if ((rc = launch_seg_readback(dM, static_cast<int>(nm), stride, dRB, stream))) return rc;
    HL_TRY_HIP(hipMemcpyAsync(hRB, dRB, sizeof(double) * nm * stride, hipMemcpyDeviceToHost, stream));
    HL_TRY_HIP(sync_stream());
    const double th1 = now_s();
    bury(); // the device is idle: frees here cost no stall
    if (prof.on && (rc = prof.flush())) return rc;
    for (size_t i = 0; i < nm; ++i) {
        act[i]->absorb(hRB + i * stride);
        if (prof.on) prof.tot.bytes[K_DGKS] += act[i]->dgks_bytes;
    }
    rc = after_sweep(act);
    if (prof.on) {
        host_s += now_s() - th1;
        prof.tot.launches[K_HOST] += 1;
        prof.tot.seconds[K_HOST] += host_s;
    }
    return rc;
}

// Restart logic of every member on host threads, then one segmented V.Q launch
// for the restarting members and a V.Q GEMM plus copy-out per finishing one.
int Engine::after_sweep(std::vector<Member *> &act)
{
    const size_t nm = act.size();
    int ncvmax = 0;
    for (Member *m : act) ncvmax = std::max(ncvmax, m->ncv);
    const size_t qstride = (size_t)ncvmax * ncvmax;
    pool->run(static_cast<int>(nm), [&](int i) { act[i]->decide(hQ + (size_t)i * qstride); });

    // Restart entries in four runs: FP64 bases on the FMA kernel, FP64 on MFMA,
    // FP32 on FMA, FP32 on MFMA.  Each member's run follows from its own
    // options, phase and shape alone, so a solve gives the same bits in any
    // batch; each basis type is one launch_vq_seg call.
    int nfinish = 0, nrephase = 0;
    std::vector<int> cat[4], wide; // wide: bases beyond the segmented restart kernels, one launch_vq each
    for (size_t i = 0; i < nm; ++i) {
        const Member &m = *act[i];
        nfinish += m.act == Member::ACT_FINISH;
        nrephase += m.act == Member::ACT_REPHASE;
        if (m.act != Member::ACT_RESTART) continue;
        if (m.ncv > kNarrowCols) {
            wide.push_back(static_cast<int>(i));
            continue;
        }
        const bool mf = m.opts.use_mfma && vq_seg_mfma_fits(m.ncv, m.vq_N);
        const bool v32 = m.kind == SEG_V32 || m.kind == SEG_V32_A32;
        cat[(v32 ? 2 : 0) + (mf ? 1 : 0)].push_back(static_cast<int>(i));
    }
    int nres = 0, nblk[4] = {0, 0, 0, 0}, maxK[4] = {0, 0, 0, 0}, maxN[4] = {0, 0, 0, 0}, mnct[4] = {0, 0, 0, 0};
    size_t mlds[4] = {0, 0, 0, 0};
    double bytes = 0.0;
    for (int c = 0; c < 4; ++c)
        for (int i : cat[c]) {
            Member &m = *act[i];
            Workspace &W = m.ws;
            SegVQ &t = hV[nres++];
            t.V = W.V[W.cur];
            t.Out = W.V[1 - W.cur];
            t.Out32 = m.kind == SEG_PROJ32 ? W.V32[1 - W.cur] : nullptr;
            t.Q = dQ + i * qstride;
            t.r = W.r;
            t.ctl = W.ctl;
            t.partials = W.red.partials;
            t.qlast = m.vq_qlast;
            t.beta = m.vq_beta;
            t.n = m.n;
            t.K = m.ncv;
            t.ldv = W.vl().ld;
            t.N = m.vq_N;
            t.res_col = m.vq_res;
            t.nb = vq_seg_blocks(m.n);
            t.blk0 = nblk[c];
            nblk[c] += t.nb;
            maxK[c] = std::max(maxK[c], t.K);
            maxN[c] = std::max(maxN[c], t.N);
            const int nct = (t.N + 15) / 16;
            mnct[c] = std::max(mnct[c], nct);
            mlds[c] = std::max(mlds[c], sizeof(double) * (size_t)((t.K + 3) & ~3) * nct * 16);
            const double vb = c >= 2 ? 4.0 : 8.0;
            bytes += vb * m.n * (t.K + t.N) + 16.0 * m.n + (t.Out32 ? 4.0 * m.n * t.N : 0.0);
        }
    int rc;
    if (nres > 0 || nfinish > 0 || nrephase > 0 || !wide.empty())
        HL_TRY_HIP(hipMemcpyAsync(dQ, hQ, sizeof(double) * nm * qstride, hipMemcpyHostToDevice, stream));
    if (nres > 0) {
        HL_TRY_HIP(hipMemcpyAsync(dV, hV, sizeof(SegVQ) * nres, hipMemcpyHostToDevice, stream));
        if ((rc = prof.begin(stream))) return rc;
        const int n01 = static_cast<int>(cat[0].size() + cat[1].size());
        if (n01 > 0 && (rc = launch_vq_seg(dV, static_cast<int>(cat[0].size()), nblk[0], maxK[0], maxN[0],
                                           static_cast<int>(cat[1].size()), nblk[1], mnct[1], mlds[1], false, stream)))
            return rc;
        if (nres > n01 && (rc = launch_vq_seg(dV + n01, static_cast<int>(cat[2].size()), nblk[2], maxK[2], maxN[2],
                                              static_cast<int>(cat[3].size()), nblk[3], mnct[3], mlds[3], true,
                                              stream)))
            return rc;
        if ((rc = prof.end(stream, K_VQ, bytes))) return rc;
    }
    for (int i : wide) { // the K-chunked V.Q, then the dsapps residual and the FP32 copy on their own
        Member &m = *act[i];
        Workspace &W = m.ws;
        if ((rc = prof.begin(stream))) return rc;
        if ((rc = launch_vq(m.n, m.ncv, m.vq_N, W.V[W.cur], W.vl(), dQ + i * qstride, m.ncv, W.V[1 - W.cur], W.vl(),
                            m.opts.use_mfma, -1, nullptr, 0.0, 0.0, W.red, W.ctl, stream)))
            return rc;
        if ((rc = launch_vq_residual(m.n, W.V[1 - W.cur], W.vl().ld, m.vq_res, W.r, m.vq_qlast, m.vq_beta, W.red,
                                     W.ctl, stream)))
            return rc;
        if (m.kind == SEG_PROJ32 &&
            (rc = launch_basis_copy32(m.n, m.vq_N, W.vl().ld, W.V[1 - W.cur], W.V32[1 - W.cur], stream)))
            return rc;
        if ((rc = prof.end(stream, K_VQ, 8.0 * m.n * (m.ncv + m.vq_N) + 16.0 * m.n))) return rc;
    }
    for (Member *m : act)
        if (m->act == Member::ACT_RESTART) m->ws.cur = 1 - m->ws.cur;
    for (size_t i = 0; i < nm && nrephase > 0; ++i) { // FP32 phase done: FP64 start vector V q, then restart
        Member &m = *act[i];
        if (m.act != Member::ACT_REPHASE) continue;
        Workspace &W = m.ws;
        if ((rc = launch_basis_combine(m.n, m.ncv, W.V[W.cur], W.vl().ld, true, dQ + i * qstride, W.r, stream)))
            return rc;
        if ((rc = m.rephase(stream))) return rc;
    }
    if (nfinish == 0) return HL_OK;
    Pending fin{nullptr, {}};
    std::vector<Member *> ver;
    for (size_t i = 0; i < nm; ++i) {
        Member &m = *act[i];
        if (m.act != Member::ACT_FINISH) continue;
        if (!m.need_ritz) { // nothing on the device: the outputs are ready
            m.complete();
            continue;
        }
        m.phase = Member::FINISHING;
        fin.ms.push_back(&m);
        Workspace &W = m.ws;
        const int nxt = 1 - W.cur;
        if ((rc = prof.begin(stream))) return rc;
        if ((rc = launch_vq(m.n, m.ncv, m.ncv, W.V[W.cur], W.vl(), dQ + i * qstride, m.ncv, W.V[nxt],
                            VLayout{W.ldv, false}, m.opts.use_mfma, -1, nullptr, 0.0, 0.0, W.red, W.ctl, stream)))
            return rc;
        if ((rc = prof.end(stream, K_VQ, 16.0 * m.n * m.ncv))) return rc;
        // sync_host (HLanc.hpp:1063-1071): all ncv columns to the caller's
        // buffer, host (pageable or page-locked) or device memory
        if (m.with_vectors) { // pageable host memory makes the copy wait: not while another engine captures
            auto wl = wait_lock();
            HL_TRY_HIP(hipMemcpy2DAsync(m.vectors, sizeof(double) * (size_t)m.ldv_host, W.V[nxt],
                                        sizeof(double) * (size_t)W.ldv, sizeof(double) * (size_t)m.n, m.ncv,
                                        hipMemcpyDefault, stream));
        }
        if (m.verify) ver.push_back(&m);
    }
    if (!ver.empty() && (rc = verify_pairs(ver))) return rc;
    if (fin.ms.empty()) return HL_OK;
    // completed when this event has passed: the next sweep is enqueued meanwhile
    {
        auto lk = hip_lock();
        HL_TRY_HIP(hipEventCreateWithFlags(&fin.ev, hipEventDisableTiming));
    }
    HL_TRY_HIP(hipEventRecord(fin.ev, stream));
    pend.push_back(std::move(fin));
    return HL_OK;
}

// The residual check of the finishing members' returned pairs, after their
// Ritz-vector GEMM (vectors in column-major V[1 - cur], ld ldv): y_i = A v_i into
// column i of the old basis V[cur] (no longer read), then one launch of the
// residual sums.  Device operators: one segmented SpMV launch per pair index
// over every finishing member (each member's own partial and long-row slots,
// so its pairs run one launch after another); host operators: the callback on
// host copies.  The partial sums come home with the vectors.
int Engine::verify_pairs(std::vector<Member *> &ver)
{
    int rc;
    size_t npairs = 0;
    int maxnev = 0;
    for (Member *m : ver) {
        npairs += m->nev0;
        maxnev = std::max(maxnev, m->nev0);
    }
    if (npairs > capMv) {
        auto lk = hip_lock();
        if ((rc = grow_set(dMv, &hMv, 1, capMv, npairs))) return rc;
        if ((rc = grow_set(dSv, &hSv, 1, capSv, npairs))) return rc;
        if ((rc = grow_set(dRv, &hRv, 1, capRv, npairs))) return rc;
    }
    if ((rc = prof.begin(stream))) return rc;
    // host operators: y_i = A v_i through the pinned staging
    for (Member *m : ver) {
        if (m->A) continue;
        Workspace &W = m->ws;
        const double *Vr = W.V[1 - W.cur];
        double *Y = W.V[W.cur];
        for (int i = 0; i < m->nev0; ++i) {
            HL_TRY_HIP(hipMemcpyAsync(W.xh, Vr + (size_t)i * W.ldv, sizeof(double) * m->n, hipMemcpyDeviceToHost,
                                      stream));
            HL_TRY_HIP(sync_stream());
            m->host_op(m->host_ctx, W.xh, W.yh);
            HL_TRY_HIP(hipMemcpyAsync(Y + (size_t)i * W.ldv, W.yh, sizeof(double) * m->n, hipMemcpyHostToDevice,
                                      stream));
        }
        HL_TRY_HIP(sync_stream()); // yh is reused by the next member
    }
    // device operators: entries ordered by pair index, so launch i is one run
    std::vector<int> off(maxnev + 1, 0), nblk(maxnev, 0), nred(maxnev, 0), nstg(maxnev, 0), nbin(maxnev, 0),
        nslx(maxnev, 0);
    size_t ne = 0;
    for (int i = 0; i < maxnev; ++i) {
        off[i] = static_cast<int>(ne);
        for (Member *m : ver) {
            if (!m->A || i >= m->nev0) continue;
            Workspace &W = m->ws;
            SegMember &M = hMv[ne];
            std::memset(&M, 0, sizeof(M));
            spmv_fields(M, *m);
            M.r = W.V[1 - W.cur] + (size_t)i * W.ldv;
            M.w = W.V[W.cur] + (size_t)i * W.ldv;
            M.ctl = dOne;
            M.n = m->n;
            M.ncv = m->ncv;
            SegStep &S = hSv[ne];
            std::memset(&S, 0, sizeof(S));
            S.mem = static_cast<int>(ne);
            S.spmv0 = nblk[i];
            S.red0 = nred[i];
            S.stg0 = nstg[i];
            S.bin0 = nbin[i];
            nblk[i] += m->spmv_blocks();
            if (m->sliced()) nred[i] += (m->n + kRowsPerSumBlock - 1) / kRowsPerSumBlock;
            if (m->staged()) {
                nstg[i] += m->A->stg_nsup;
                nbin[i] += m->A->stg_nbins;
                nslx[i] = std::max(nslx[i], m->A->stg_nsl);
            }
            ++ne;
        }
    }
    off[maxnev] = static_cast<int>(ne);
    if (ne > 0) {
        HL_TRY_HIP(hipMemcpyAsync(dMv, hMv, sizeof(SegMember) * ne, hipMemcpyHostToDevice, stream));
        HL_TRY_HIP(hipMemcpyAsync(dSv, hSv, sizeof(SegStep) * ne, hipMemcpyHostToDevice, stream));
        for (int i = 0; i < maxnev; ++i) {
            const int cnt = off[i + 1] - off[i];
            if (cnt <= 0) continue;
            const bool solo = ne == 1; // one pair: the table's only entry
            if (nblk[i] > 0 && (rc = launch_spmv_seg(dMv, dSv + off[i], cnt, nblk[i], SEG_F64, solo, stream)))
                return rc;
            if (nred[i] > 0 && (rc = launch_spmv_rows_seg(dMv, dSv + off[i], cnt, nred[i], solo, stream))) return rc;
            if (nstg[i] > 0 &&
                (rc = launch_spmv_staged_seg(dMv, dSv + off[i], cnt, nstg[i], nbin[i], nslx[i], solo, stream)))
                return rc;
        }
    }
    // the residual sums of every pair in one launch
    int blocks = 0;
    size_t nr = 0;
    for (Member *m : ver) {
        Workspace &W = m->ws;
        const int G = res_blocks(m->n);
        for (int i = 0; i < m->nev0; ++i) {
            SegRes &t = hRv[nr++];
            t.v = W.V[1 - W.cur] + (size_t)i * W.ldv;
            t.y = W.V[W.cur] + (size_t)i * W.ldv;
            t.out = W.red.partials + (size_t)3 * G * i;
            t.lam = m->vals[i];
            t.n = m->n;
            t.G = G;
            t.blk0 = blocks;
            blocks += G;
        }
    }
    HL_TRY_HIP(hipMemcpyAsync(dRv, hRv, sizeof(SegRes) * nr, hipMemcpyHostToDevice, stream));
    if ((rc = launch_resid_seg(dRv, static_cast<int>(nr), blocks, stream))) return rc;
    for (Member *m : ver)
        HL_TRY_HIP(hipMemcpyAsync(m->ws.vres_h, m->ws.red.partials,
                                  sizeof(double) * 3 * (size_t)res_blocks(m->n) * m->nev0, hipMemcpyDeviceToHost,
                                  stream));
    return prof.end(stream, K_MISC, 0.0);
}

int Engine::run(std::vector<Member *> &queue, int max_active, double hbm_budget)
{
    Admission adm;
    adm.queue = &queue;
    adm.max_active = max_active;
    adm.budget = hbm_budget;
    adm.t_start = now_s();
    return run_shared(adm);
}

int Engine::run_shared(Admission &adm)
{
    t_start = adm.t_start;
    std::vector<Member *> &queue = *adm.queue;
    std::vector<Member *> active;
    int rc = HL_OK;
    auto need = [](const Member *m) {
        return Workspace::bytes_needed(m->n, m->ncv, m->A == nullptr, m->nslots(), m->npairs(), m->copy32());
    };
    // members done since the last call leave the active set (and the shared
    // counts); their workspaces are freed at the next point where the device is idle
    auto retire = [&]() {
        size_t keep = 0;
        int gone = 0;
        double freed = 0.0;
        for (Member *m : active) {
            if (m->phase == Member::DONE) {
                m->t_done = now_s() - t_start;
                graveyard.push_back(m->ws.detach());
                freed += need(m);
                ++gone;
            } else {
                active[keep++] = m;
            }
        }
        active.resize(keep);
        if (gone) {
            std::lock_guard<std::mutex> lk(adm.mu);
            adm.active -= gone;
            adm.in_use -= freed;
            adm.cv.notify_all();
        }
    };
    auto sweeping = [&]() {
        for (Member *m : active)
            if (m->phase != Member::FINISHING) return true;
        return false;
    };
    for (;;) {
        if ((rc = poll_finished(false))) break; // Ritz vectors that have arrived
        retire();
        // admission at the sweep boundary, in queue order
        std::vector<Member *> admitted;
        bool queue_done = false;
        {
            std::unique_lock<std::mutex> lk(adm.mu);
            for (;;) {
                const double t = now_s() - t_start;
                auto fits = [&](const Member *m) {
                    return (adm.max_active <= 0 || adm.active < adm.max_active) &&
                           (adm.budget <= 0.0 || adm.active == 0 || adm.in_use + need(m) <= adm.budget);
                };
                // this engine's share: its part of what is in flight or has arrived
                size_t arrived = adm.next;
                while (arrived < queue.size() && queue[arrived]->arrival <= t) ++arrived;
                const size_t avail = static_cast<size_t>(adm.active) + (arrived - adm.next);
                const size_t share = (avail + adm.engines - 1) / adm.engines;
                while (adm.next < arrived && fits(queue[adm.next]) &&
                       (adm.engines == 1 || active.size() + admitted.size() < share)) {
                    Member *m = queue[adm.next++];
                    m->t_admit = t;
                    adm.active += 1;
                    adm.in_use += need(m);
                    admitted.push_back(m);
                }
                queue_done = adm.next >= queue.size();
                if (!active.empty() || !admitted.empty() || !pend.empty() || queue_done) break;
                // nothing to do here: wait for the next arrival, or for another
                // engine to finish a member (nothing in flight means it fits)
                const Member *q = queue[adm.next];
                if (q->arrival > t && fits(q))
                    adm.cv.wait_for(lk, std::chrono::duration<double>(q->arrival - t));
                else
                    adm.cv.wait(lk);
            }
        }
        if (queue_done && active.empty() && admitted.empty() && pend.empty()) break;
        for (Member *m : admitted) {
            m->phase = Member::NEW;
            auto hk = hip_lock(); // allocations: not while another engine captures a sweep
            if ((m->err = m->ws.init(dev, m->n, m->ncv, m->A == nullptr, m->nslots(), m->npairs(), m->copy32(),
                                     m->verifies() ? m->nev0 : 0, stream))) {
                m->ws.release();
                m->phase = Member::DONE; // retired below, out of the shared counts
            }
            active.push_back(m);
        }
        retire();
        if (sweeping()) {
            if ((rc = sweep(active))) break;
            retire();
        } else if (!pend.empty()) { // only finishing members are left: wait for their vectors
            if ((rc = poll_finished(true))) break;
            retire();
        }
    }
    if (!rc) rc = poll_finished(true);
    if (!rc) retire();
    if (!rc && prof.on) { // the V.Q launches after the last sweep
        HL_TRY_HIP(sync_stream());
        rc = prof.flush();
    }
    if (rc) { // a HIP failure ends every member in flight on this engine
        (void)sync_stream();
        {
            auto lk = hip_lock();
            for (Pending &p : pend) (void)hipEventDestroy(p.ev);
        }
        pend.clear();
        double freed = 0.0;
        for (Member *m : active) {
            m->err = rc;
            freed += need(m);
            auto lk = hip_lock();
            m->ws.release();
            m->phase = Member::DONE;
        }
        std::lock_guard<std::mutex> lk(adm.mu);
        adm.active -= static_cast<int>(active.size());
        adm.in_use -= freed;
        adm.cv.notify_all();
        active.clear();
    }
    HL_TRY_HIP(sync_stream());
    bury();
    return rc;
}

} // namespace hl
