// serve.cpp -- the serving loop: eigensolve requests submitted one by one
// while others run, each keeping its basis, tridiagonal and convergence state
// on the device for its whole life.
//
// The reference has only the blocking ds_solver_gpu call (HLanc.hpp:1868-1872).
// Here a server owns one device and two host threads:
//   * the serving thread runs an Engine: every in-flight request advances by
//     one dsaitr sweep per lockstep round (one segmented SpMV launch and four
//     segmented CGS passes per Lanczos step for all of them); between rounds it
//     retires finished requests and admits queued ones, in submission order,
//     while fewer than max_active are in flight and the admitted footprints fit
//     the HBM budget.  A request's footprint is its workspace (the basis pair,
//     vectors, partial sums) plus, when the server uploads the operator, the
//     operator's device bytes (12 nnz + 4 (n + 1) + 16 n of staging);
//   * the loader thread uploads the operators of admitted requests submitted as
//     host CSR arrays, so that uploads overlap the sweeps of the requests
//     already running; a request joins the lockstep rounds once its operator
//     is on the device.
// Each request computes bitwise what hl_ds_solve_csr computes on the same
// operator (same plan) and options, whatever runs beside it.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hl_internal.hpp"
#include "irlm.hpp"

namespace hl {
namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    long long ticket = 0;
    Member m;
    // host CSR to upload (the caller keeps the arrays valid until the request is done)
    int nrows = 0, nnz = 0, sym = 0;
    const int *rp = nullptr, *col = nullptr;
    const double *val = nullptr;
    hl_csr_t A = nullptr;     // the operator: the caller's, or uploaded by the loader
    bool owned = false;       // uploaded by the server (freed when the request is done)
    int load_rc = HL_OK;      // the upload's status
    double footprint = 0.0;   // bytes counted against the budget while admitted
    enum State { QUEUED, LOADING, READY, RUNNING, DONE } state = QUEUED;
    int status = HL_ERR_STATE;
    double t_submit = 0.0, t_admit = 0.0, t_start = 0.0, t_done = 0.0;
};

} // namespace
} // namespace hl

struct hl_server_s {
    int dev = 0;
    int max_active = 0;
    double budget = 0.0;
    int flags = 0;
    double t0 = 0.0;
    std::mutex mu;
    std::mutex hip_mu;                 // the loader's HIP calls vs. the serving thread's graph captures
    std::condition_variable cv_work;   // the serving thread: new requests, uploads done, stop
    std::condition_variable cv_load;   // the loader: uploads to do, stop
    std::condition_variable cv_done;   // waiters: a request finished
    std::deque<hl::Request *> queued;  // submitted, not admitted (submission order)
    std::deque<hl::Request *> to_load; // admitted, operator to upload
    std::deque<hl::Request *> ready;   // admitted, operator on the device: join the next round
    std::unordered_map<long long, std::unique_ptr<hl::Request>> reqs;
    long long next_ticket = 1;
    long long n_done = 0;
    double in_use = 0.0;               // footprints of admitted, unfinished requests
    int in_flight = 0;                 // admitted, unfinished requests
    bool stop = false;                 // no more submissions; drain and end the serving thread
    bool load_stop = false;            // the serving thread has ended: end the loader
    int fatal = HL_OK;                 // a HIP failure: every later request fails with it
    std::thread serve_th, load_th;
};

namespace hl {
namespace {

double operator_bytes(int n, long long nnz_general)
{
    return 12.0 * static_cast<double>(nnz_general) + 4.0 * (n + 1.0) + 16.0 * n;
}

// A request's workspace bytes, once its operator is known.
double workspace_bytes(const Member &m)
{
    return Workspace::bytes_needed(m.n, m.ncv, false, m.nslots(), m.npairs(), m.copy32());
}

// The loader thread: uploads admitted requests' operators in admission order.
void load_loop(hl_server_s *s)
{
    DeviceGuard g(s->dev);
    for (;;) {
        Request *q = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_load.wait(lk, [&] { return s->load_stop || !s->to_load.empty(); });
            if (s->to_load.empty()) return; // the serving thread has ended and nothing is left
            q = s->to_load.front();
            s->to_load.pop_front();
        }
        hl_csr_t h = nullptr;
        int rc;
        {
            std::lock_guard<std::mutex> hk(s->hip_mu); // not while a sweep is being captured
            rc = hl_csr_create(s->dev, q->nrows, q->nrows, q->nnz, q->rp, q->col, q->val, q->sym, HL_PREC_F64, &h);
            if (rc == HL_OK && (s->flags & HL_SERVE_SELL_EVEN)) {
                // the sliced-ELL form where rows are even (fill <= 1.1); a refusal keeps the items
                long long entries = 0;
                if (hl_csr_set_sell(h, 1) == HL_OK && hl_csr_get_sell(h, nullptr, &entries) == HL_OK &&
                    static_cast<double>(entries) > 1.1 * static_cast<double>(csr_plan(h)->nnz))
                    rc = hl_csr_set_sell(h, 0);
            }
        }
        std::lock_guard<std::mutex> lk(s->mu);
        if (h) { // the footprint again: the operator's device bytes and the workspace its form needs
            q->m.A = csr_plan(h);
            const double fp = workspace_bytes(q->m) + csr_device_bytes(*csr_plan(h));
            s->in_use += fp - q->footprint;
            q->footprint = fp;
        }
        q->A = h;
        q->owned = h != nullptr;
        q->load_rc = rc;
        q->state = Request::READY;
        s->ready.push_back(q);
        s->cv_work.notify_all();
    }
}

void serve_loop(hl_server_s *s)
{
    DeviceGuard g(s->dev);
    Engine eng;
    int rc = eng.init(s->dev);
    eng.capture_mu = &s->hip_mu;
    std::vector<Member *> active;
    std::unordered_map<Member *, Request *> owner;
    std::vector<hl_csr_t> release; // server-owned operators of finished requests (freed when idle)
    auto finish = [&](Request *q, int status) { // under s->mu; q was admitted
        q->status = status;
        q->t_done = now_s() - s->t0;
        q->state = Request::DONE;
        s->in_use -= q->footprint;
        s->in_flight -= 1;
        s->n_done += 1;
        if (q->owned) release.push_back(q->A);
        s->cv_done.notify_all();
    };
    // members done since the last call: record results, free workspaces at the next idle point
    auto retire = [&]() {
        size_t keep = 0;
        std::lock_guard<std::mutex> lk(s->mu);
        for (Member *m : active) {
            if (m->phase == Member::DONE) {
                Request *q = owner[m];
                eng.graveyard.push_back(m->ws.detach());
                owner.erase(m);
                finish(q, m->err ? m->err : m->ret);
            } else {
                active[keep++] = m;
            }
        }
        active.resize(keep);
    };
    auto free_released = [&]() { // the device is idle after a sweep's synchronisation
        for (hl_csr_t h : release) csr_release(h);
        release.clear();
    };
    auto fail_active = [&](int err) { // after a HIP failure
        (void)hipStreamSynchronize(eng.stream);
        for (Engine::Pending &p : eng.pend) (void)hipEventDestroy(p.ev);
        eng.pend.clear();
        std::lock_guard<std::mutex> lk(s->mu);
        s->fatal = err;
        for (Member *m : active) {
            m->ws.release();
            finish(owner[m], err);
        }
        active.clear();
        owner.clear();
    };
    if (rc) fail_active(rc);
    for (;;) {
        if (!rc) {
            if ((rc = eng.poll_finished(false))) {
                fail_active(rc);
            } else {
                retire();
            }
        }
        std::vector<Request *> joining;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // admission in submission order: concurrency cap and HBM budget (the
            // first request of an idle server is always admitted)
            auto admissible = [&]() {
                if (s->queued.empty()) return false;
                const Request *q = s->queued.front();
                return (s->max_active <= 0 || s->in_flight < s->max_active) &&
                       (s->budget <= 0.0 || s->in_flight == 0 || s->in_use + q->footprint <= s->budget);
            };
            while (admissible()) {
                Request *q = s->queued.front();
                s->queued.pop_front();
                q->t_admit = now_s() - s->t0;
                s->in_use += q->footprint;
                s->in_flight += 1;
                if (q->A || s->fatal) {
                    q->state = Request::READY;
                    s->ready.push_back(q);
                } else {
                    q->state = Request::LOADING;
                    s->to_load.push_back(q);
                    s->cv_load.notify_one();
                }
            }
            while (!s->ready.empty()) { // operators on the device: these join this round
                Request *q = s->ready.front();
                s->ready.pop_front();
                if (s->fatal || q->load_rc) {
                    finish(q, s->fatal ? s->fatal : q->load_rc);
                    continue;
                }
                q->state = Request::RUNNING;
                joining.push_back(q);
            }
            if (joining.empty() && active.empty() && eng.pend.empty()) {
                auto idle_done = [&] { return s->stop && s->queued.empty() && s->in_flight == 0; };
                if (idle_done()) break;
                s->cv_work.wait(lk, [&] { return idle_done() || admissible() || !s->ready.empty(); });
                continue;
            }
        }
        for (Request *q : joining) { // the operator is known now: the member's own workspace
            Member &m = q->m;
            m.A = csr_plan(q->A);
            q->t_start = now_s() - s->t0;
            m.phase = Member::NEW;
            if ((m.err = m.ws.init(s->dev, m.n, m.ncv, false, m.nslots(), m.npairs(), m.copy32(),
                                   m.verifies() ? m.nev0 : 0, eng.stream))) {
                m.ws.release();
                m.phase = Member::DONE;
            }
            owner[&m] = q;
            active.push_back(&m);
        }
        bool sweeping = false;
        for (Member *m : active) sweeping |= m->phase == Member::NEW || m->phase == Member::ACTIVE;
        if (sweeping) {
            if ((rc = eng.sweep(active))) {
                fail_active(rc);
                continue;
            }
            free_released();
            retire();
        } else if (!eng.pend.empty()) {
            if ((rc = eng.poll_finished(true))) {
                fail_active(rc);
                continue;
            }
            retire();
        } else {
            retire(); // members that failed to start
        }
    }
    (void)hipStreamSynchronize(eng.stream);
    eng.bury();
    free_released();
}

int submit(hl_server_s *s, Request *q, hl_csr_t A, int nev, int ncv, const hl_options *opts, double *values,
           double *vectors, int ldv, long long *ticket)
{
    hl_options o;
    hl_options_default(&o);
    if (opts) o = *opts;
    Member &m = q->m;
    m.dev = s->dev;
    m.n = A ? csr_plan(A)->nrows : q->nrows;
    m.nev0 = nev;
    m.ncv = ncv;
    m.opts = o;
    m.values = values;
    m.vectors = vectors;
    m.ldv_host = ldv;
    if (!solve_args_ok(m.n, nev, ncv, ldv, values, false, vectors, o, true, "hl_serve_submit")) return HL_ERR_ARG;
    q->A = A;
    m.A = A ? csr_plan(A) : nullptr; // a caller's operator: its plan sizes the workspace
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        set_error("hl_serve_submit: the server is shutting down");
        return HL_ERR_STATE;
    }
    if (s->fatal) {
        set_error("hl_serve_submit: the server stopped after a HIP failure (%d)", s->fatal);
        return s->fatal;
    }
    // the footprint: the workspace (long-row slots and sliced partials of a
    // server-built operator follow its default plan: none counted beyond the
    // rows) plus the operator itself when the server uploads it
    const double ws =
        A ? workspace_bytes(m)
          : Workspace::bytes_needed(m.n, ncv, false, 0, 0, o.basis_prec == HL_PREC_F64_PROJ32);
    long long nnz_general = q->nnz;
    if (!A && q->sym) { // lower storage expands to both triangles
        long long diag = 0;
        for (int r = 0; r < q->nrows; ++r)
            for (int p = q->rp[r]; p < q->rp[r + 1]; ++p) diag += q->col[p] == r;
        nnz_general = 2LL * q->nnz - diag;
    }
    q->footprint = ws + (A ? 0.0 : operator_bytes(m.n, nnz_general));
    q->ticket = s->next_ticket++;
    q->t_submit = now_s() - s->t0;
    q->state = Request::QUEUED;
    *ticket = q->ticket;
    s->queued.push_back(q);
    s->reqs.emplace(q->ticket, std::unique_ptr<Request>(q));
    s->cv_work.notify_all();
    return HL_OK;
}

void fill_result(const Request &q, hl_serve_result *out)
{
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    out->status = q.status;
    out->info = q.m.info;
    out->info.ret = q.status;
    out->t_submit = q.t_submit;
    out->t_admit = q.t_admit;
    out->t_start = q.t_start;
    out->t_done = q.t_done;
    out->footprint = q.footprint;
}

} // namespace
} // namespace hl

using namespace hl;

extern "C" {

int hl_serve_create(int dev, int max_active, double hbm_budget, int flags, hl_server_t *out)
try {
    if (!out) return HL_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        set_error("hl_serve_create: no HIP device visible");
        return HL_ERR_NO_DEVICE;
    }
    if (dev < 0 || dev >= count || (flags & ~HL_SERVE_SELL_EVEN)) {
        set_error("hl_serve_create: bad device %d or flags 0x%x", dev, flags);
        return HL_ERR_ARG;
    }
    hl_server_s *s = new (std::nothrow) hl_server_s;
    if (!s) return HL_ERR_OOM;
    s->dev = dev;
    s->max_active = max_active;
    s->budget = hbm_budget;
    s->flags = flags;
    s->t0 = now_s();
    s->serve_th = std::thread(serve_loop, s);
    s->load_th = std::thread(load_loop, s);
    *out = s;
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_serve_submit(hl_server_t s, hl_csr_t A, int nev, int ncv, const hl_options *opts, double *values,
                    double *vectors, int ldv, long long *ticket)
try {
    if (!s || !A || !ticket) return HL_ERR_ARG;
    const CsrPlan *p = csr_plan(A);
    if (p->dev != s->dev || p->nrows != p->ncols) {
        set_error("hl_serve_submit: the operator must be square and on the server's device %d", s->dev);
        return HL_ERR_ARG;
    }
    Request *q = new (std::nothrow) Request;
    if (!q) return HL_ERR_OOM;
    const int rc = hl::submit(s, q, A, nev, ncv, opts, values, vectors, ldv, ticket);
    if (rc) delete q;
    return rc;
} catch (...) {
    return exception_code();
}

int hl_serve_submit_csr(hl_server_t s, int nrows, int nnz, const int *row_ptr, const int *col, const double *val,
                        int sym_lower, int nev, int ncv, const hl_options *opts, double *values, double *vectors,
                        int ldv, long long *ticket)
try {
    if (!s || !ticket || nrows <= 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!col || !val))) return HL_ERR_ARG;
    if (row_ptr[0] != 0 || row_ptr[nrows] != nnz) {
        set_error("hl_serve_submit_csr: inconsistent CSR (row_ptr[0] must be 0 and row_ptr[nrows] == nnz)");
        return HL_ERR_ARG;
    }
    Request *q = new (std::nothrow) Request;
    if (!q) return HL_ERR_OOM;
    q->nrows = nrows;
    q->nnz = nnz;
    q->sym = sym_lower ? 1 : 0;
    q->rp = row_ptr;
    q->col = col;
    q->val = val;
    const int rc = hl::submit(s, q, nullptr, nev, ncv, opts, values, vectors, ldv, ticket);
    if (rc) delete q;
    return rc;
} catch (...) {
    return exception_code();
}

int hl_serve_poll(hl_server_t s, long long ticket, hl_serve_result *out)
try {
    if (!s) return HL_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) {
        set_error("hl_serve_poll: unknown ticket %lld", ticket);
        return HL_ERR_ARG;
    }
    if (it->second->state != Request::DONE) return HL_PENDING;
    fill_result(*it->second, out);
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_serve_wait(hl_server_t s, long long ticket, double timeout_s, hl_serve_result *out)
try {
    if (!s) return HL_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) {
        set_error("hl_serve_wait: unknown ticket %lld", ticket);
        return HL_ERR_ARG;
    }
    Request *q = it->second.get();
    auto done = [&] { return q->state == Request::DONE; };
    if (timeout_s < 0.0)
        s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
        return HL_PENDING;
    fill_result(*q, out);
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_serve_release(hl_server_t s, long long ticket)
try {
    if (!s) return HL_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end() || it->second->state != Request::DONE) {
        set_error("hl_serve_release: ticket %lld unknown or not done", ticket);
        return HL_ERR_ARG;
    }
    s->reqs.erase(it);
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_serve_stats(hl_server_t s, hl_serve_stats_t *out)
try {
    if (!s || !out) return HL_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->next_ticket - 1;
    out->done = s->n_done;
    out->queued = static_cast<int>(s->queued.size());
    out->in_flight = s->in_flight;
    out->bytes_in_use = s->in_use;
    out->seconds = now_s() - s->t0;
    return s->fatal;
} catch (...) {
    return exception_code();
}

int hl_serve_destroy(hl_server_t s)
try {
    if (!s) return HL_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
        s->cv_load.notify_all();
    }
    if (s->serve_th.joinable()) s->serve_th.join(); // every request is done: nothing left to upload
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->load_stop = true;
        s->cv_load.notify_all();
    }
    if (s->load_th.joinable()) s->load_th.join();
    const int rc = s->fatal;
    delete s;
    return rc;
} catch (...) {
    return exception_code();
}

} // extern "C"
