// irlm.hpp -- solver state of one implicitly restarted Lanczos solve.
#pragma once

#include <vector>

#include "hl_internal.hpp"

namespace hl {

// Device workspace of one solve (freed on every return path).
struct Workspace {
    int dev = 0, n = 0, ncv = 0, prec = HL_PREC_F64;
    long long ldv = 0;               // device leading dimension, multiple of 64
    hipStream_t stream = nullptr;
    double *V[2] = {nullptr, nullptr};  // FP64 basis, ping-pong for V Q
    float *Vf[2] = {nullptr, nullptr};  // FP32 basis (mixed precision)
    int cur = 0;
    double *r = nullptr, *w = nullptr;  // residual and A v_j
    double *hsub = nullptr, *hdia = nullptr, *hbuf = nullptr, *Qd = nullptr, *vd_tmp = nullptr;
    double *xv = nullptr;               // contiguous v_j for host operators
    double *ipart = nullptr, *fin = nullptr; // fused SpMV + projection partial sums
    Ctl *ctl = nullptr;
    Ctl *ctl_h = nullptr;               // pinned mirrors
    double *hh = nullptr, *xh = nullptr, *yh = nullptr;
    Reduce red;

    // The ping-pong bases are tiled (16-row blocks of ncv columns).
    bool tiled = true;                  // fixed at init from tuning().basis_tiled
    VLayout vl() const { return tiled ? VLayout{ncv, true} : VLayout{ldv, false}; }
    ~Workspace();
    int init(int dev, int n, int ncv, int prec, bool host_op, int items);
    void release();
};

struct Solver {
    Workspace ws;
    const CsrPlan *A = nullptr;          // device operator, or
    hl_host_op_fn host_op = nullptr;     // host callback
    void *host_ctx = nullptr;
    int use_mfma = 1;
    bool fuse = true;                    // fused SpMV + first CGS half (CSR operator)
    std::vector<double> sub, dia;        // host copy of T: sub[j] = beta_j, dia[j] = alpha_j
    double rnorm = 0.0;
    long long spmv = 0, cgs = 0;
    int restarts = 0;
    Profiler prof;                       // HL_FLAG_PROFILE

    int run(int nev, double *values, double *vectors, int ldv_host, const hl_options &o, hl_info &info);

  private:
    int reset_ctl();
    int apply_op(int j);
    int step(int j);
    int extend(int k, int np);
    int apply_restart(int nev, const double *q);
    int cgs_step(int m, const double *V, const double *w_in, double *w_out, const double *h_in, double *h_out,
                 int mode, int j);
};

}  // namespace hl
