// irlm.hpp -- solve state and the multi-solve engine of libhlanc_hip.so.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <shared_mutex>
#include <utility>
#include <vector>

#include "hl_internal.hpp"

namespace hl {

// Device workspace of one solve: one device allocation and one pinned host
// allocation, carved up (freed on every return path).
struct Workspace {
    int dev = 0, n = 0, ncv = 0;
    long long ldv = 0;                  // column-major leading dimension, multiple of 64
    double *V[2] = {nullptr, nullptr};  // FP64 tiled basis, ping-pong for V Q
    float *V32[2] = {nullptr, nullptr}; // FP32 copies of V (projection passes), if requested
    int cur = 0;
    double *r = nullptr, *w = nullptr;  // residual and A v_j
    double *hsub = nullptr, *hdia = nullptr, *hbuf = nullptr;
    double *xv = nullptr;               // contiguous v_j for host operators
    double *part = nullptr;             // partials of a sliced SpMV (npairs)
    double *chunk_part = nullptr;       // this solve's long-row partial sums (nslots)
    unsigned *chunk_cnt = nullptr;      // and arrival counters, so that solves of
                                        // one operator can share an SpMV launch
    Ctl *ctl = nullptr;
    Ctl *ctl_h = nullptr;               // pinned
    double *xh = nullptr, *yh = nullptr; // pinned, host operators only
    double *vres_h = nullptr;           // pinned: residual-check partials (3 kResBlocksMax per pair)
    Reduce red;                         // block partial sums (carved, not owned)
    void *blob = nullptr, *hblob = nullptr;

    // The ping-pong bases are tiled: 16-row blocks of ncv columns.
    VLayout vl() const { return VLayout{tiled_ld(ncv), true}; }
    ~Workspace();
    int init(int dev, int n, int ncv, bool host_op, int nslots, int npairs, bool copy32, int nres,
             hipStream_t s); // zeroing on s; nres: pairs the residual check covers
    // Hands the two allocations to the caller (to free once the device is idle)
    // and clears the workspace.
    std::pair<void *, void *> detach();
    static double bytes_needed(int n, int ncv, bool host_op, int nslots, int npairs,
                               bool copy32); // device bytes of init
    void release();
};

// Host restart logic shared by every solve (irlm.cpp).
void sort_wanted(int nev, int np, double *ritz, double *bounds);                  // dsgets
int count_converged(int n, const double *ritz, const double *bounds, double tol);  // dsconv
void restart_chase(int nev, int np, const double *shift, double *sub, double *dia, double *q); // dsapps sweeps

// One eigensolve: the host half of ds_solver (HLanc.hpp:1649-1850) as a state
// machine that the engine drives sweep by sweep.  A sweep is one dsaitr call,
// Lanczos steps [sweep_k, sweep_k + sweep_np).  After a sweep the engine hands
// the member its read-back alpha/beta entries and Ctl (absorb), the member
// decides on the host what comes next (decide: restart logic, restart Q or
// Ritz-vector Q), and the engine enqueues the device work of all members.
struct Member {
    // inputs
    const CsrPlan *A = nullptr;          // device operator, or
    hl_host_op_fn host_op = nullptr;     // host callback
    void *host_ctx = nullptr;
    int dev = 0, n = 0, nev0 = 0, ncv = 0;
    hl_options opts{};
    double *values = nullptr, *vectors = nullptr; // caller's host outputs (vectors may be null)
    int ldv_host = 0;
    double arrival = 0.0;                // seconds after the engine's start; admitted no earlier
    double t_admit = 0.0, t_done = 0.0;  // seconds after the engine's start
    // device state
    Workspace ws;
    // host state
    enum Phase { NEW, ACTIVE, FINISHING, DONE } phase = NEW; // FINISHING: Ritz vectors in flight
    enum Act { ACT_NONE, ACT_SWEEP, ACT_RESTART, ACT_FINISH, ACT_REPHASE } act = ACT_NONE;
    // Mixed precision (opts.basis_prec == HL_PREC_F32): a first phase keeps V
    // in FP32 (and reads FP32 values when the operator has them) and stops at
    // kF32PhaseTol; the solve then restarts in FP64 from the sum of that
    // phase's wanted Ritz vectors (ACT_REPHASE) and ends as an FP64 solve.
    // The FP32 phase's relative Ritz-estimate tolerance is 10^-f32_phase_tol_exp
    // (a tuning key, 1e-6 by default: above the ~6e-8 relative rounding of an
    // FP32 basis, so that the phase ends).
    int kind = SEG_F64;                  // SegKind of the current phase
    bool mixed = false;
    long long spmv_f32 = 0;              // FP32-phase counters
    int restarts_f32 = 0;
    std::vector<double> vals_f32;        // wanted Ritz values at the end of the FP32 phase
    int stage = 0;                       // 0: initial sweep pending, 1: main loop
    int sweep_k = 0, sweep_np = 0;
    int nev = 0, np = 0, nconv = 0, iter = 0, restarts = 0, maxiter = 200;
    double tol = 0.0, rnorm = 0.0, t0 = 0.0;
    long long spmv = 0, cgs = 0, p2 = 0, p3 = 0, m2sum = 0, m3sum = 0, nexact = 0, mxsum = 0;
    double dgks_bytes = 0.0;             // algorithmic bytes of the DGKS corrections of the last sweep
    int graph_sweeps = 0;                // sweeps replayed from a captured graph
    int ret = HL_NOT_CONVERGED;
    int err = HL_OK;                     // a HIP or allocation failure of this member
    hl_info info{};
    // pending device work chosen by decide()
    int vq_N = 0, vq_res = -1;           // restart: Q columns and the beta column
    double vq_qlast = 0.0, vq_beta = 0.0;
    bool with_vectors = false;           // finish: Ritz vectors to the caller
    bool need_ritz = false;              // finish: the Ritz-vector GEMM runs (vectors or the check)
    bool verify = false;                 // finish: the residual check runs on the returned pairs
    double anorm = 0.0;                  // largest |eigenvalue| of the final tridiagonal
    std::vector<double> sub, dia, ritz, bounds, zlast, q, e, work, ritz_saved, bounds_saved, vals;

    int start(hipStream_t s);            // start vector, norm, first sweep [0, nev0)
    int rephase(hipStream_t s);          // FP64 phase from the start vector already in ws.r
    bool f32_phase() const { return kind == SEG_V32 || kind == SEG_V32_A32; } // FP32 phase of a two-phase solve
    double vbytes() const { return kind == SEG_F64 ? 8.0 : 4.0; } // bytes per entry of the projected basis
    bool copy32() const { return A != nullptr && opts.basis_prec == HL_PREC_F64_PROJ32; } // FP32 basis copy
    // the operator's column-sliced form serves the SpMV (FP64 values and basis)
    bool sliced() const { return A != nullptr && A->slice_w > 0 && (kind == SEG_F64 || kind == SEG_PROJ32); }
    // the operator's staged-slice form serves the SpMV (FP64 values and basis)
    bool staged() const { return A != nullptr && A->stg_on && (kind == SEG_F64 || kind == SEG_PROJ32); }
    bool sell() const { return A != nullptr && A->sell_on && kind != SEG_V32_A32; } // FP64 values only
    bool verifies() const { return (opts.flags & HL_FLAG_NO_VERIFY) == 0; } // residual check requested
    int spmv_blocks() const; // blocks of this solve's SpMV launch segment
    int nslots() const { return A ? std::max(A->items.nslots, A->sitems.nslots) : 0; }
    // partials of a sliced or staged SpMV (the forms exclude each other)
    int npairs() const { return A ? (A->stg_on ? kWave * A->stg_nchunks : A->npairs) : 0; }
    void absorb(const double *rb);       // hsub, hdia (ncv each) and the Ctl words
    // Host-only (thread-safe across members): restart logic; writes the restart
    // Q (ncv x vq_N) or the Ritz-vector Q (ncv x ncv) to qslot.
    void decide(double *qslot);
    int host_apply(int j, hipStream_t s); // SpMV of a host operator for step j
    double cgs_bytes_pass(int m, bool sub) const; // algorithmic bytes of one CGS pass (sub: EPI_PASS1)
    void complete();                     // outputs and info once the engine synchronised
    void check_sums(int i, double &a, double &p, double &c) const; // residual check of pair i

  private:
    int reset(hipStream_t s);            // host state, Ctl and T for a new phase
    int begin(hipStream_t s);            // norm of ws.r, first sweep
    void finish_host(double *qslot, bool vectors_ok);
    void rephase_host(double *qslot);
};


// Runs members to completion, all active members advancing in lockstep: one
// segmented SpMV launch and four segmented CGS passes per Lanczos step serve
// every member that has a step left in its current sweep; one segmented V.Q
// launch restarts every member that restarts after the sweep.
//
// Admission state of one batch, shared by the engines that serve it (each on
// its own stream and host thread): members are taken from the queue in order,
// once their arrival time has passed, while fewer than max_active (<= 0: no
// cap) are in flight over all engines and their workspaces fit in budget bytes
// (<= 0: no cap; when nothing is in flight one member is always admitted).
struct Admission {
    std::mutex mu;
    std::condition_variable cv;     // a member finished: an idle engine may admit now
    std::vector<Member *> *queue = nullptr;
    size_t next = 0;                // first member not admitted yet
    int active = 0;                 // admitted and not done, all engines
    double in_use = 0.0;            // their workspace bytes
    int max_active = 0;
    double budget = 0.0;
    double t_start = 0.0;           // steady-clock seconds when the batch started
    int engines = 1;                // engines sharing the queue: each admits up to its share of the
                                    // members in flight or arrived, so that all of them get work
};

struct Engine {
    int dev = 0;
    hipStream_t stream = nullptr;
    Profiler prof;
    // launch tables (device) and their pinned staging, two sets used in turn
    // so that a sweep's uploads may still be in flight while the next sweep's
    // tables are written
    SegMember *dM = nullptr, *hM = nullptr, *hMs[2] = {nullptr, nullptr};
    SegStep *dS = nullptr, *hS = nullptr, *hSs[2] = {nullptr, nullptr};
    SegVQ *dV = nullptr, *hV = nullptr, *hVs[2] = {nullptr, nullptr};
    size_t capM = 0, capS = 0, capV = 0;
    // read-back (device + pinned) and Q arenas (device + two pinned)
    double *dRB = nullptr, *hRB = nullptr, *dQ = nullptr, *hQ = nullptr, *hQs[2] = {nullptr, nullptr};
    size_t capRB = 0, capQ = 0;
    // residual check of finishing members: one SpMV entry and one SegRes per
    // returned pair (device tables and their pinned staging), and a Ctl whose
    // xscale is 1 (the SpMV kernels scale x by ctl->xscale)
    SegMember *dMv = nullptr, *hMv = nullptr;
    SegStep *dSv = nullptr, *hSv = nullptr;
    SegRes *dRv = nullptr, *hRv = nullptr;
    size_t capMv = 0, capSv = 0, capRv = 0;
    Ctl *dOne = nullptr;
    int par = 0;
    // finishing members whose Ritz vectors are in flight, completed once
    // their event has passed; workspaces freed at the next idle point
    struct Pending {
        hipEvent_t ev;
        std::vector<Member *> ms;
    };
    std::vector<Pending> pend;
    std::vector<std::pair<void *, void *>> graveyard;
    struct Pool;
    Pool *pool = nullptr;                // host threads for decide()
    // Sweep launch sequences replayed as HIP graphs: a sweep's launches depend
    // only on the table addresses and per-step grid shapes, so a shape seen
    // twice is captured once and replayed after that (no profiling, no host
    // operators, not with HL_FLAG_NO_GRAPH).
    bool graphs = true;
    struct GraphEntry {
        std::vector<long long> key;
        hipGraphExec_t exec;
        long long last;
    };
    std::vector<GraphEntry> gcache;
    std::vector<std::vector<long long>> gseen; // shapes met once
    long long gclock = 0;
    long long graph_launches = 0, graph_captures = 0;

    // Held while a sweep is captured into a graph and around the engine's
    // allocations and frees, when other host threads of the same process issue
    // HIP calls on the device meanwhile (the serving loop's loader, the other
    // engines of a batch): HIP invalidates a capture that such calls overlap.
    std::mutex *capture_mu = nullptr;
    std::unique_lock<std::mutex> hip_lock() const
    {
        return capture_mu ? std::unique_lock<std::mutex>(*capture_mu) : std::unique_lock<std::mutex>();
    }
    // Shared between the engines of a batch: each engine's waits on its own
    // stream or events hold it shared, a capture holds it exclusively, so that
    // no wait of one engine overlaps another engine's capture (HIP refuses
    // such a wait with hipErrorStreamCaptureUnsupported).
    std::shared_mutex *sync_rw = nullptr;
    std::shared_lock<std::shared_mutex> wait_lock() const
    {
        return sync_rw ? std::shared_lock<std::shared_mutex>(*sync_rw) : std::shared_lock<std::shared_mutex>();
    }
    hipError_t sync_stream() const
    {
        auto lk = wait_lock();
        return hipStreamSynchronize(stream);
    }
    double t_start = 0.0;                // steady-clock seconds at run()
    // host_threads: workers of the decide() pool (< 0: one per host core, up to 15)
    int init(int dev, int host_threads = -1);
    void release();
    ~Engine() { release(); }
    // Start NEW members (start vector, norm) and run one sweep of every ACTIVE
    // member; members that finish become DONE.
    int sweep(std::vector<Member *> &ms);
    // Continuous batching: members are admitted in order, once their arrival
    // time has passed, while fewer than max_active (<= 0: no cap) are in flight
    // and their workspaces fit in hbm_budget bytes (<= 0: no cap; one member is
    // always admitted).  Admission happens between sweeps; a member's workspace
    // is allocated on admission and freed when it is done.
    int run(std::vector<Member *> &queue, int max_active, double hbm_budget);
    // The same loop over admission state shared with other engines.
    int run_shared(Admission &adm);

    int poll_finished(bool wait);         // complete members whose vectors have arrived
    void bury();                          // free detached workspaces (device idle)

  private:
    int grow(size_t nm, size_t ns, size_t nv, size_t rb, size_t qd);
    int after_sweep(std::vector<Member *> &act);
    void drop_graphs();
    int verify_pairs(std::vector<Member *> &ver); // residual check, enqueued after the Ritz vectors
};

// Shared with the serving loop (capi.cpp): argument rules of a solve, the plan
// behind an operator handle, and the release of a handle without the device
// synchronisation of hl_csr_destroy (the caller knows the device is done with it).
bool solve_args_ok(int n, int nev, int ncv, int ldv, const double *values, bool need_vectors,
                   const double *vectors, const hl_options &o, bool device_op, const char *who);
const CsrPlan *csr_plan(hl_csr_t h);
void csr_release(hl_csr_t h);

}  // namespace hl
