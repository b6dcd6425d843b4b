// irlm.hpp -- solve state and the multi-solve engine of libhlanc_hip.so.
#pragma once

#include <vector>

#include "hl_internal.hpp"

namespace hl {

// Device workspace of one solve (freed on every return path).
struct Workspace {
    int dev = 0, n = 0, ncv = 0, prec = HL_PREC_F64;
    long long ldv = 0;               // column-major leading dimension, multiple of 64
    bool tiled = true;               // the solver's basis is always tiled
    double *V[2] = {nullptr, nullptr};  // FP64 basis, ping-pong for V Q
    float *Vf[2] = {nullptr, nullptr};  // FP32 basis (mixed precision)
    int cur = 0;
    double *r = nullptr, *w = nullptr;  // residual and A v_j
    double *hsub = nullptr, *hdia = nullptr, *hbuf = nullptr, *Qd = nullptr, *vd_tmp = nullptr;
    double *xv = nullptr;               // contiguous v_j for host operators
    double *chunk_part = nullptr;       // this solve's long-row partial sums (nslots)
    unsigned *chunk_cnt = nullptr;      // and arrival counters, so that solves of
                                        // one operator can share an SpMV launch
    Ctl *ctl = nullptr;
    Ctl *ctl_h = nullptr;               // pinned mirrors
    double *hh = nullptr, *xh = nullptr, *yh = nullptr;
    Reduce red;

    // The ping-pong bases are tiled (16-row blocks of ncv columns) by default.
    VLayout vl() const { return tiled ? VLayout{ncv, true} : VLayout{ldv, false}; }
    ~Workspace();
    int init(int dev, int n, int ncv, int prec, bool host_op, int nslots);
    static double bytes_needed(int n, int ncv, int prec, bool host_op, int nslots); // device bytes of init
    void release();
};

// Host restart logic shared by every solve (irlm.cpp).
void sort_wanted(int nev, int np, double *ritz, double *bounds);                  // dsgets
int count_converged(int n, const double *ritz, const double *bounds, double tol);  // dsconv
void restart_chase(int nev, int np, const double *shift, double *sub, double *dia, double *q); // dsapps sweeps

// One eigensolve: the host half of ds_solver (HLanc.hpp:1649-1850) as a state
// machine that the engine drives sweep by sweep.  A sweep is one dsaitr call,
// Lanczos steps [sweep_k, sweep_k + sweep_np); between sweeps the member reads
// back its new alpha/beta entries and runs the restart logic on the host.
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
    Profiler *prof = nullptr;            // the engine's (set on admission)
    // device state
    Workspace ws;
    // host state
    enum Phase { NEW, ACTIVE, DONE } phase = NEW;
    int stage = 0;                       // 0: initial sweep pending, 1: main loop
    int sweep_k = 0, sweep_np = 0;
    int nev = 0, np = 0, nconv = 0, iter = 0, restarts = 0, maxiter = 200;
    double tol = 0.0, rnorm = 0.0, t0 = 0.0;
    long long spmv = 0, cgs = 0, p2 = 0, p3 = 0, m2sum = 0, m3sum = 0;
    int ret = HL_NOT_CONVERGED;
    int err = HL_OK;                     // a HIP or allocation failure of this member
    hl_info info{};
    std::vector<double> sub, dia, ritz, bounds, q, e, work, ritz_saved, bounds_saved, qf;

    int start(hipStream_t s);            // start vector, norm, first sweep [0, nev0)
    int after_sweep(hipStream_t s);      // restart logic; next sweep, or finish
    int host_apply(int j, hipStream_t s); // SpMV of a host operator for step j
    double cgs_bytes_pass(int m) const;  // algorithmic bytes of one CGS pass

  private:
    int restart(hipStream_t s);
    int finish(hipStream_t s, bool vectors_ok);
};

// Runs members to completion, all active members advancing in lockstep: one
// segmented SpMV launch and four segmented CGS passes per Lanczos step serve
// every member that has a step left in its current sweep.
struct Engine {
    int dev = 0;
    hipStream_t stream = nullptr;
    Profiler prof;
    // device tables
    SegMember *dM = nullptr;
    SegStep *dS = nullptr;
    size_t capM = 0, capS = 0;
    // pinned staging
    SegMember *hM = nullptr;
    SegStep *hS = nullptr;
    size_t hcapM = 0, hcapS = 0;

    double t_start = 0.0;                // steady-clock seconds at run()
    int init(int dev);
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

  private:
    int grow(size_t nm, size_t ns);
};

}  // namespace hl
