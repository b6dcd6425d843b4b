// capi.cpp -- the extern "C" entry points declared in include/hlanc_capi.h.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <exception>
#include <new>
#include <vector>

#include "hl_internal.hpp"
#include "irlm.hpp"

struct hl_csr_s {
    hl::CsrPlan plan;
    std::mutex mu; // serializes host-staged applications (shared x/y staging)
};

namespace hl {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int exception_code() noexcept
{
    try {
        throw;
    } catch (const std::bad_alloc &) {
        set_error("out of host memory");
        return HL_ERR_OOM;
    } catch (const std::exception &e) {
        set_error("internal error: %s", e.what());
        return HL_ERR_STATE;
    } catch (...) {
        set_error("internal error: unknown exception");
        return HL_ERR_STATE;
    }
}

int hip_status(hipError_t e, const char *what)
{
    set_error("%s: %s", what, hipGetErrorString(e));
    return HL_ERR_HIP_BASE + static_cast<int>(e);
}

namespace {

int check_device(int dev)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        set_error("no HIP device visible");
        return HL_ERR_NO_DEVICE;
    }
    if (dev < 0 || dev >= count) {
        set_error("device %d out of range (%d visible)", dev, count);
        return HL_ERR_ARG;
    }
    return HL_OK;
}

// true when p is device memory of device dev (HLanc.hpp:539,545 probe).  Plain
// malloc memory may report an error or an unregistered type: both mean host.
bool is_device_ptr(const void *p, int dev)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice && a.device == dev;
}

} // namespace

// Argument rules of ds_solver (HLanc.hpp:1649-1660) plus this library's
// limits (ncv <= kMaxNcv; the two-phase FP32 mode ncv <= kNarrowCols).
bool solve_args_ok(int n, int nev, int ncv, int ldv, const double *values, bool need_vectors,
                   const double *vectors, const hl_options &o, bool device_op, const char *who)
{
    if (n <= 0 || nev <= 0 || ncv <= nev || ncv > n || !values || (need_vectors && !vectors) ||
        (vectors && ldv < n)) {
        set_error("%s: need 0 < nev < ncv <= n and ldv >= n (n=%d nev=%d ncv=%d ldv=%d)", who, n, nev, ncv, ldv);
        return false;
    }
    if (ncv > kMaxNcv || (o.basis_prec == HL_PREC_F32 && ncv > kNarrowCols)) {
        set_error("%s: ncv=%d exceeds this library's basis width (%d; %d for basis_prec HL_PREC_F32)", who, ncv,
                  kMaxNcv, kNarrowCols);
        return false;
    }
    if (o.basis_prec != HL_PREC_F64 && o.basis_prec != HL_PREC_F32 && o.basis_prec != HL_PREC_F64_PROJ32) {
        set_error("%s: basis_prec must be HL_PREC_F64, HL_PREC_F64_PROJ32 or HL_PREC_F32", who);
        return false;
    }
    if (o.refine != HL_REFINE_NONE && o.refine != HL_REFINE_RQ) {
        set_error("%s: refine must be HL_REFINE_NONE or HL_REFINE_RQ", who);
        return false;
    }
    if (o.refine == HL_REFINE_RQ && (o.flags & HL_FLAG_NO_VERIFY)) {
        set_error("%s: HL_REFINE_RQ needs the residual check (HL_FLAG_NO_VERIFY is set)", who);
        return false;
    }
    if (o.basis_prec == HL_PREC_F32 && !device_op) {
        set_error("%s: the mixed-precision solve (basis_prec = HL_PREC_F32) needs a device CSR operator", who);
        return false;
    }
    return true;
}

const CsrPlan *csr_plan(hl_csr_t h) { return &h->plan; }

void csr_release(hl_csr_t h)
{
    if (!h) return;
    DeviceGuard g(h->plan.dev);
    csr_plan_free(h->plan);
    delete h;
}

namespace {

// One solve is an engine with one member.
int solve_common(hl::Member &m, int dev, int n, int nev, int ncv, double *values, double *vectors, int ldv,
                 const hl_options *opts, hl_info *info)
{
    hl_options o;
    hl_options_default(&o);
    if (opts) o = *opts;
    hl_info inf;
    std::memset(&inf, 0, sizeof(inf));
    if (info) *info = inf;
    if (!solve_args_ok(n, nev, ncv, ldv, values, true, vectors, o, m.A != nullptr, "ds_solver")) return HL_ERR_ARG;
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    m.dev = dev;
    m.n = n;
    m.nev0 = nev;
    m.ncv = ncv;
    m.opts = o;
    m.values = values;
    m.vectors = vectors;
    m.ldv_host = ldv;
    hl::Engine eng;
    if ((rc = eng.init(dev))) return rc;
    eng.prof.on = (o.flags & HL_FLAG_PROFILE) != 0;
    eng.graphs = (o.flags & HL_FLAG_NO_GRAPH) == 0;
    std::vector<hl::Member *> q{&m};
    rc = eng.run(q, 1, 0.0);
    if (!rc) rc = m.err;
    if (eng.prof.on && !rc) thread_profile() = eng.prof.tot;
    if (rc) return rc;
    if (info) *info = m.info;
    return m.ret;
}

} // namespace
} // namespace hl

using namespace hl;

extern "C" {

const char *hl_last_error(void) { return g_err; }

const char *hl_version(void) { return "1.0"; }

int hl_host_alloc(size_t bytes, void **ptr)
try {
    if (!ptr) return HL_ERR_ARG;
    *ptr = nullptr;
    HL_TRY_HIP(hipHostMalloc(ptr, bytes > 0 ? bytes : 1, hipHostMallocDefault));
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_host_free(void *ptr)
try {
    if (ptr) HL_TRY_HIP(hipHostFree(ptr));
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_device_count(int *count)
try {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    if (count) *count = c;
    return HL_OK;
} catch (...) {
    return exception_code();
}

void hl_options_default(hl_options *o)
{
    std::memset(o, 0, sizeof(*o));
    o->tol = 0.0;
    o->maxiter = 200;
    o->v0 = nullptr;
    o->seed = 0;
    o->basis_prec = HL_PREC_F64; // the reference's DGKS rule; HL_PREC_F64_PROJ32 is the faster opt-in
    o->use_mfma = 1; // measured faster than the FMA kernel on every benchmarked workload
    o->refine = HL_REFINE_NONE;
    o->residuals = nullptr;
}

int hl_csr_create(int dev, int nrows, int ncols, int nnz, const int *row_ptr, const int *col,
                  const double *val, int sym_lower, int val_prec, hl_csr_t *out)
try {
    if (!out) return HL_ERR_ARG;
    *out = nullptr;
    if (nrows < 0 || ncols < 0 || (nrows > 0 && (!row_ptr || row_ptr[0] != 0 || row_ptr[nrows] != nnz)) ||
        (nnz > 0 && (!col || !val))) {
        set_error("hl_csr_create: inconsistent CSR (row_ptr[0] must be 0 and row_ptr[nrows] == nnz)");
        return HL_ERR_ARG;
    }
    for (int r = 0; r < nrows; ++r)
        if (row_ptr[r + 1] < row_ptr[r]) {
            set_error("hl_csr_create: row_ptr decreases at row %d", r);
            return HL_ERR_ARG;
        }
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    hl_csr_s *h = new (std::nothrow) hl_csr_s;
    if (!h) return HL_ERR_OOM;
    h->plan.dev = dev;
    rc = csr_plan_build(h->plan, nrows, ncols, row_ptr, col, val, sym_lower, val_prec);
    if (rc) {
        csr_plan_free(h->plan);
        delete h;
        return rc;
    }
    *out = h;
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_csr_destroy(hl_csr_t h)
try {
    if (!h) return HL_OK;
    DeviceGuard g(h->plan.dev);
    (void)hipDeviceSynchronize();
    csr_plan_free(h->plan);
    delete h;
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_csr_info(hl_csr_t h, int *dev, int *nrows, int *ncols, long long *nnz_general)
try {
    if (!h) return HL_ERR_ARG;
    if (dev) *dev = h->plan.dev;
    if (nrows) *nrows = h->plan.nrows;
    if (ncols) *ncols = h->plan.ncols;
    if (nnz_general) *nnz_general = h->plan.nnz;
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_csr_apply(hl_csr_t h, const double *x, double *y)
try {
    if (!h || (h->plan.ncols > 0 && !x) || (h->plan.nrows > 0 && !y)) return HL_ERR_ARG;
    std::lock_guard<std::mutex> lk(h->mu);
    CsrPlan &p = h->plan;
    DeviceGuard g(p.dev);
    const bool xd = is_device_ptr(x, p.dev), yd = is_device_ptr(y, p.dev);
    const double *xx = x;
    if (!xd) {
        HL_TRY_HIP(hipMemcpy(p.x_stage, x, sizeof(double) * p.ncols, hipMemcpyDefault));
        xx = p.x_stage;
    }
    double *yy = yd ? y : p.y_stage;
    int rc = launch_spmv(p, xx, nullptr, yy, VColumn{}, HL_PREC_F64, nullptr);
    if (rc) return rc;
    if (!yd) HL_TRY_HIP(hipMemcpy(y, yy, sizeof(double) * p.nrows, hipMemcpyDefault));
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_csr_spmv(hl_csr_t h, const double *x_dev, double *y_dev, void *stream)
try {
    if (!h) return HL_ERR_ARG;
    DeviceGuard g(h->plan.dev);
    return launch_spmv(h->plan, x_dev, nullptr, y_dev, VColumn{}, HL_PREC_F64, static_cast<hipStream_t>(stream));
} catch (...) {
    return exception_code();
}

int hl_csr_spmv_f32val(hl_csr_t h, const double *x_dev, double *y_dev, void *stream)
try {
    if (!h) return HL_ERR_ARG;
    DeviceGuard g(h->plan.dev);
    return launch_spmv(h->plan, x_dev, nullptr, y_dev, VColumn{}, HL_PREC_F32, static_cast<hipStream_t>(stream));
} catch (...) {
    return exception_code();
}

double hl_csr_spmv_bytes(hl_csr_t h, int val_prec)
{
    if (!h) return 0.0;
    const double sa = val_prec == HL_PREC_F32 ? 4.0 : 8.0;
    return (sa + 4.0) * static_cast<double>(h->plan.nnz) + 4.0 * (h->plan.nrows + 1.0) + 8.0 * h->plan.ncols +
           8.0 * h->plan.nrows;
}

namespace {
int csr_items(hl_csr_t h, int nnz_cap, int rows_cap)
{
    CsrPlan &p = h->plan;
    std::vector<int> R(static_cast<size_t>(p.nrows) + 1);
    HL_TRY_HIP(hipMemcpy(R.data(), p.row_ptr, sizeof(int) * R.size(), hipMemcpyDeviceToHost));
    int rc = plan_items(p.items, R.data(), p.nrows, nnz_cap, rows_cap);
    if (rc || p.slice_w == 0) return rc;
    R.resize(static_cast<size_t>(p.npairs) + 1); // the sliced form's items take the same shape
    HL_TRY_HIP(hipMemcpy(R.data(), p.s_rp, sizeof(int) * R.size(), hipMemcpyDeviceToHost));
    return plan_items(p.sitems, R.data(), p.npairs, nnz_cap, rows_cap);
}
} // namespace

int hl_csr_set_items(hl_csr_t h, int nnz_cap, int rows_cap)
try {
    if (!h) return HL_ERR_ARG;
    DeviceGuard g(h->plan.dev);
    std::lock_guard<std::mutex> lk(h->mu);
    return csr_items(h, nnz_cap, rows_cap);
} catch (...) {
    return exception_code();
}

int hl_csr_set_slices(hl_csr_t h, int width)
try {
    if (!h) return HL_ERR_ARG;
    DeviceGuard g(h->plan.dev);
    std::lock_guard<std::mutex> lk(h->mu);
    return plan_slices(h->plan, width);
} catch (...) {
    return exception_code();
}

int hl_csr_get_slices(hl_csr_t h, int *width, int *npairs)
try {
    if (!h) return HL_ERR_ARG;
    if (width) *width = h->plan.slice_w;
    if (npairs) *npairs = h->plan.npairs;
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_csr_set_sell(hl_csr_t h, int on)
try {
    if (!h) return HL_ERR_ARG;
    DeviceGuard g(h->plan.dev);
    std::lock_guard<std::mutex> lk(h->mu);
    return plan_sell(h->plan, on != 0, 2.0);
} catch (...) {
    return exception_code();
}

int hl_csr_get_sell(hl_csr_t h, int *on, long long *entries)
try {
    if (!h) return HL_ERR_ARG;
    long long ne = 0;
    if (h->plan.sell_on)
        HL_TRY_HIP(hipMemcpy(&ne, h->plan.e_ptr + h->plan.nchunks, sizeof(long long), hipMemcpyDeviceToHost));
    if (on) *on = h->plan.sell_on ? 1 : 0;
    if (entries) *entries = ne;
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_csr_get_sell_col_bytes(hl_csr_t h, int *bytes)
try {
    if (!h || !bytes) return HL_ERR_ARG;
    *bytes = !h->plan.sell_on ? 0 : h->plan.e_d16 ? 2 : 4;
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_csr_pick_form(hl_csr_t h, int *form)
try {
    if (!h || !form) return HL_ERR_ARG;
    DeviceGuard g(h->plan.dev);
    std::lock_guard<std::mutex> lk(h->mu);
    return csr_pick_form(h->plan, form);
} catch (...) {
    return exception_code();
}

int hl_csr_device_bytes(hl_csr_t h, double *bytes)
try {
    if (!h || !bytes) return HL_ERR_ARG;
    std::lock_guard<std::mutex> lk(h->mu);
    *bytes = csr_device_bytes(h->plan);
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_csr_set_staged(hl_csr_t h, int on)
try {
    if (!h) return HL_ERR_ARG;
    DeviceGuard g(h->plan.dev);
    std::lock_guard<std::mutex> lk(h->mu);
    return plan_staged(h->plan, on != 0);
} catch (...) {
    return exception_code();
}

int hl_csr_get_staged(hl_csr_t h, int *on, long long *pieces, long long *entries, int *bins)
try {
    if (!h) return HL_ERR_ARG;
    const CsrPlan &p = h->plan;
    long long ne = 0;
    if (p.stg_on)
        HL_TRY_HIP(hipMemcpy(&ne, p.s_cptr + p.stg_nchunks, sizeof(long long), hipMemcpyDeviceToHost));
    if (on) *on = p.stg_on ? 1 : 0;
    if (pieces) *pieces = p.stg_npieces;
    if (entries) *entries = ne;
    if (bins) *bins = p.stg_nbins;
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_csr_autotune(hl_csr_t h, int reps, int *nnz_cap, int *rows_cap, double *seconds)
try {
    if (!h || reps <= 0) return HL_ERR_ARG;
    DeviceGuard g(h->plan.dev);
    std::lock_guard<std::mutex> lk(h->mu);
    CsrPlan &p = h->plan;
    double *x = nullptr, *y = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = HL_OK, best_n = kNnzPerBlock, best_r = kRowsPerBlock;
    double best_t = 1e30;
    auto cleanup = [&]() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (x) (void)hipFree(x);
        if (y) (void)hipFree(y);
    };
    hipError_t e;
    if ((e = hipMalloc(&x, sizeof(double) * std::max(p.ncols, 1))) != hipSuccess ||
        (e = hipMalloc(&y, sizeof(double) * std::max(p.nrows, 1))) != hipSuccess ||
        (e = hipMemset(x, 0, sizeof(double) * std::max(p.ncols, 1))) != hipSuccess ||
        (e = hipEventCreate(&a)) != hipSuccess || (e = hipEventCreate(&b)) != hipSuccess) {
        cleanup();
        return hip_status(e, "hl_csr_autotune");
    }
    // seconds per launch of the operator's current form (3 untimed launches first)
    auto time_form = [&](double &t) -> int {
        int r = HL_OK;
        for (int i = 0; i < 3 && !r; ++i) r = launch_spmv(p, x, nullptr, y, VColumn{}, HL_PREC_F64, nullptr);
        if (r) return r;
        float ms = 0.0f;
        if ((e = hipEventRecord(a, nullptr)) != hipSuccess) return hip_status(e, "hl_csr_autotune");
        for (int i = 0; i < reps && !r; ++i) r = launch_spmv(p, x, nullptr, y, VColumn{}, HL_PREC_F64, nullptr);
        if (r) return r;
        if ((e = hipEventRecord(b, nullptr)) != hipSuccess || (e = hipEventSynchronize(b)) != hipSuccess ||
            (e = hipEventElapsedTime(&ms, a, b)) != hipSuccess)
            return hip_status(e, "hl_csr_autotune");
        t = 1e-3 * ms / reps;
        return HL_OK;
    };
    const int ncap[] = {512, 1024, 2048};
    const int rcap[] = {64, 128, 256, 512, 1024, 2048};
    // column slices only where x (8 ncols bytes) exceeds an XCD's 4 MB L2 twice
    // over; widest first, and a narrower width must win by 2 %: it makes more
    // virtual rows, whose partials a solve's row-sum pass also writes into the basis
    std::vector<int> widths{0};
    if (p.ncols >= (1 << 20))
        for (int w : {1 << 19, 1 << 18, 1 << 17}) widths.push_back(w);
    int best_w = 0;
    bool best_sell = false;
    rc = plan_sell(p, false, 0.0); // items first
    for (size_t wi = 0; wi < widths.size() && !rc; ++wi) {
        const int wdt = widths[wi];
        if ((rc = plan_slices(p, wdt))) break;
        for (int nc : ncap)
            for (int rcp : rcap) {
                double t = 0.0;
                if (rc || (rc = csr_items(h, nc, rcp)) || (rc = time_form(t))) break;
                if (t < (wdt != best_w && best_w != 0 ? 0.98 : 1.0) * best_t) {
                    best_t = t;
                    best_n = nc;
                    best_r = rcp;
                    best_w = wdt;
                }
            }
    }
    // then the sliced-ELL form, where its padding stays within 1.5 nnz
    if (!rc) {
        const int src = plan_sell(p, true, 1.5);
        double t = 0.0;
        if (src == HL_OK) {
            if (!(rc = time_form(t)) && t < best_t) {
                best_t = t;
                best_sell = true;
            }
        } else if (src != HL_ERR_ARG) {
            rc = src;
        }
    }
    // then the staged-slice form, which must win by 2 % (it takes more memory)
    bool best_stg = false;
    if (!rc && p.nnz >= (1 << 16)) {
        const int src = plan_staged(p, true);
        double t = 0.0;
        if (src == HL_OK) {
            if (!(rc = time_form(t)) && t < 0.98 * best_t) {
                best_t = t;
                best_stg = true;
            }
        } else if (src != HL_ERR_ARG) {
            rc = src;
        }
    }
    cleanup();
    if (rc) return rc;
    if (best_stg) {
        // the staged form is in place
    } else if (best_sell) {
        if ((rc = plan_sell(p, true, 0.0))) return rc;
    } else if ((rc = plan_staged(p, false)) || (rc = plan_sell(p, false, 0.0)) || (rc = plan_slices(p, best_w))) {
        return rc;
    }
    if ((rc = csr_items(h, best_n, best_r))) return rc; // (FP32-value SpMVs use the items)
    if (nnz_cap) *nnz_cap = best_n;
    if (rows_cap) *rows_cap = best_r;
    if (seconds) *seconds = best_t;
    return HL_OK;
} catch (...) {
    return exception_code();
}

namespace {
// Reduction workspace of the standalone basis entry points, one per host
// thread and device, kept for the life of the process so that the calls are
// asynchronous and cheap to repeat.
struct PassWs {
    Reduce red;
    Ctl *ctl = nullptr;
    double *hout = nullptr;
};
int pass_ws(int dev, PassWs *&out)
{
    static thread_local PassWs ws[16];
    if (dev < 0 || dev >= 16) return HL_ERR_ARG;
    PassWs &w = ws[dev];
    if (!w.ctl) {
        int rc = reduce_alloc(w.red, kMaxReduceBlocks, 128);
        if (rc) return rc;
        HL_TRY_HIP(hipMalloc(&w.ctl, sizeof(Ctl)));
        HL_TRY_HIP(hipMemset(w.ctl, 0, sizeof(Ctl)));
        HL_TRY_HIP(hipMalloc(&w.hout, sizeof(double) * 128));
    }
    out = &w;
    return HL_OK;
}
} // namespace

int hl_cgs_pass(int dev, int n, int m, const double *V, long long ldv, const double *w_in, double *w_out,
                const double *h_in, double *h_out, double *nrm2_out, void *stream)
try {
    if (n < 0 || m < 0 || m > 128 || (m > 0 && (!V || ldv < n)) || !w_in) {
        set_error("hl_cgs_pass: bad arguments");
        return HL_ERR_ARG;
    }
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    hipStream_t s = static_cast<hipStream_t>(stream);
    PassWs *ws = nullptr;
    if ((rc = pass_ws(dev, ws))) return rc;
    rc = launch_cgs(n, m, V, VLayout{ldv, false}, w_in, w_out, h_in, ws->hout, ws->red, ws->ctl, nullptr, nullptr,
                    EPI_NONE, 0, s);
    if (rc) return rc;
    if (h_out && m > 0)
        HL_TRY_HIP(hipMemcpyAsync(h_out, ws->hout, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    if (nrm2_out) HL_TRY_HIP(hipMemcpyAsync(nrm2_out, &ws->ctl->nrm2, sizeof(double), hipMemcpyDeviceToDevice, s));
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_cgs_pass_noalloc(int dev, int n, int m, const double *V, long long ldv, const double *w_in, double *w_out,
                        const double *h_in, double *h_out, double *nrm2_out, void *stream)
try {
    if (n < 0 || m < 0 || (m > 0 && (!V || ldv < n)) || !w_in) {
        set_error("hl_cgs_pass_noalloc: bad arguments");
        return HL_ERR_ARG;
    }
    if (m > 128) {
        set_error("hl_cgs_pass_noalloc: m=%d exceeds the 128 columns its scratch holds", m);
        return HL_ERR_ARG;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || dev < 0 || dev >= count) {
        set_error("hl_cgs_pass_noalloc: device %d not visible", dev);
        return count <= 0 ? HL_ERR_NO_DEVICE : HL_ERR_ARG;
    }
    DeviceGuard g(dev);
    return launch_cgs_noalloc(dev, n, m, V, ldv, w_in, w_out, h_in, h_out, nrm2_out, static_cast<hipStream_t>(stream));
} catch (...) {
    return exception_code();
}

int hl_bench_cgs(int dev, int n, int m, int tiled, int reps, double *seconds)
try {
    if (n <= 0 || m <= 0 || m > 128 || reps <= 0 || !seconds) return HL_ERR_ARG;
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    PassWs *ws = nullptr;
    if ((rc = pass_ws(dev, ws))) return rc;
    const long long ld = ((long long)n + 63) / 64 * 64;
    double *V = nullptr, *w = nullptr, *w2 = nullptr, *h = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    HL_TRY_HIP(hipMalloc(&V, sizeof(double) * ld * m));
    HL_TRY_HIP(hipMalloc(&w, sizeof(double) * ld));
    HL_TRY_HIP(hipMalloc(&w2, sizeof(double) * ld));
    HL_TRY_HIP(hipMalloc(&h, sizeof(double) * 128));
    HL_TRY_HIP(hipMemset(V, 0, sizeof(double) * ld * m));
    HL_TRY_HIP(hipMemset(w, 0, sizeof(double) * ld));
    HL_TRY_HIP(hipMemset(h, 0, sizeof(double) * 128));
    HL_TRY_HIP(hipEventCreate(&a));
    HL_TRY_HIP(hipEventCreate(&b));
    const VLayout L = tiled ? VLayout{m, true} : VLayout{ld, false};
    for (int i = 0; i < 3 && !rc; ++i)
        rc = launch_cgs(n, m, V, L, w, w2, h, ws->hout, ws->red, ws->ctl, nullptr, nullptr, EPI_NONE, 0, nullptr);
    if (!rc) rc = hipEventRecord(a, nullptr) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    for (int i = 0; i < reps && !rc; ++i)
        rc = launch_cgs(n, m, V, L, w, w2, h, ws->hout, ws->red, ws->ctl, nullptr, nullptr, EPI_NONE, 0, nullptr);
    float ms = 0.0f;
    if (!rc) rc = hipEventRecord(b, nullptr) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    if (!rc) rc = hipEventSynchronize(b) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    if (!rc) rc = hipEventElapsedTime(&ms, a, b) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    *seconds = 1e-3 * ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(V);
    (void)hipFree(w);
    (void)hipFree(w2);
    (void)hipFree(h);
    return rc;
} catch (...) {
    return exception_code();
}

int hl_bench_stream(int dev, long long bytes, int reps, double *seconds)
try {
    if (bytes < 1024 || reps <= 0 || !seconds) return HL_ERR_ARG;
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    const long long n = bytes / 16 * 2;
    const int blocks = 2048;
    double *a = nullptr, *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HL_TRY_HIP(hipMalloc(&a, sizeof(double) * n));
    HL_TRY_HIP(hipMalloc(&out, sizeof(double) * blocks * (kBlock / kWave)));
    HL_TRY_HIP(hipMemset(a, 0, sizeof(double) * n));
    HL_TRY_HIP(hipEventCreate(&e0));
    HL_TRY_HIP(hipEventCreate(&e1));
    rc = launch_stream_read(a, n, out, blocks, nullptr);
    if (!rc) rc = hipEventRecord(e0, nullptr) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    for (int i = 0; i < reps && !rc; ++i) rc = launch_stream_read(a, n, out, blocks, nullptr);
    float ms = 0.0f;
    if (!rc) rc = hipEventRecord(e1, nullptr) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    if (!rc) rc = hipEventElapsedTime(&ms, e0, e1) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    *seconds = 1e-3 * ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(out);
    return rc;
} catch (...) {
    return exception_code();
}

int hl_bench_grid_barrier(int dev, int blocks, int barriers, double *seconds)
try {
    if (blocks < 1 || barriers < 1 || barriers > (1 << 20) || !seconds) return HL_ERR_ARG;
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    int coop = 0;
    HL_TRY_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    const int cap = grid_barrier_blocks_max(dev);
    if (!coop || blocks > cap) {
        set_error("hl_bench_grid_barrier: %d blocks cannot all be co-resident (at most %d, cooperative launch %d)",
                  blocks, cap, coop);
        return HL_ERR_ARG;
    }
    if ((long long)barriers * blocks >= 0xffffffffLL) return HL_ERR_ARG; // counter range
    unsigned *counter = nullptr;
    double *buf = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HL_TRY_HIP(hipMalloc(&counter, 256));
    HL_TRY_HIP(hipMalloc(&buf, sizeof(double) * 2 * (size_t)blocks * kBlock));
    HL_TRY_HIP(hipEventCreate(&e0));
    HL_TRY_HIP(hipEventCreate(&e1));
    int *abort_flag = reinterpret_cast<int *>(counter + 16), *errors = reinterpret_cast<int *>(counter + 32);
    double best = 1e30;
    int host[2] = {0, 0};
    for (int rep = 0; rep < 3 && !rc; ++rep) { // the best of three launches
        if (hipMemset(counter, 0, 256) != hipSuccess) rc = HL_ERR_HIP_BASE;
        if (!rc) rc = hipEventRecord(e0, nullptr) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
        if (!rc) rc = launch_grid_barrier(blocks, counter, abort_flag, buf, barriers, errors, nullptr);
        if (!rc) rc = hipEventRecord(e1, nullptr) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
        if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
        float ms = 0.0f;
        if (!rc) rc = hipEventElapsedTime(&ms, e0, e1) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
        if (!rc) rc = hipMemcpy(host, abort_flag, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
        if (!rc) rc = hipMemcpy(host + 1, errors, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
        if (!rc && (host[0] || host[1])) {
            set_error("hl_bench_grid_barrier: %s", host[0] ? "a barrier wait expired" : "stale data after a barrier");
            rc = HL_ERR_STATE;
        }
        if (!rc && ms < best) best = ms;
    }
    if (!rc) *seconds = 1e-3 * best / barriers;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(counter);
    (void)hipFree(buf);
    return rc;
} catch (...) {
    return exception_code();
}

int hl_bench_graph_nodes(int dev, int nodes, int blocks, int reps, double *seconds)
try {
    if (nodes < 1 || nodes > 4096 || blocks < 1 || blocks > (1 << 20) || reps < 1 || !seconds) return HL_ERR_ARG;
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    hipStream_t s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HL_TRY_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HL_TRY_HIP(hipEventCreate(&e0));
    HL_TRY_HIP(hipEventCreate(&e1));
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = HL_ERR_HIP_BASE;
    for (int i = 0; i < nodes && !rc; ++i) rc = launch_empty(blocks, s);
    if (hipStreamEndCapture(s, &graph) != hipSuccess && !rc) rc = HL_ERR_HIP_BASE;
    if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) rc = HL_ERR_HIP_BASE;
    if (!rc && hipGraphLaunch(exec, s) != hipSuccess) rc = HL_ERR_HIP_BASE; // warm-up
    if (!rc) rc = hipEventRecord(e0, s) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    for (int r = 0; r < reps && !rc; ++r)
        if (hipGraphLaunch(exec, s) != hipSuccess) rc = HL_ERR_HIP_BASE;
    if (!rc) rc = hipEventRecord(e1, s) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    float ms = 0.0f;
    if (!rc) rc = hipEventElapsedTime(&ms, e0, e1) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    if (!rc) *seconds = 1e-3 * ms / ((double)reps * nodes);
    if (rc) set_error("hl_bench_graph_nodes: a HIP call failed");
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    return rc;
} catch (...) {
    return exception_code();
}

int hl_bench_gather(int dev, long long nx, long long m, int random, int reps, double *seconds)
try {
    if (nx <= 0 || nx > 0x7fffffffLL || m < 1024 || reps <= 0 || !seconds) return HL_ERR_ARG;
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    std::vector<int> idx(static_cast<size_t>(m));
    uint64_t z = 0x9E3779B97F4A7C15ULL;
    for (long long i = 0; i < m; ++i) {
        if (random) { // splitmix64 stream, reduced to [0, nx)
            z += 0x9E3779B97F4A7C15ULL;
            uint64_t v = z;
            v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
            v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
            v ^= v >> 31;
            idx[i] = static_cast<int>(v % static_cast<uint64_t>(nx));
        } else {
            idx[i] = static_cast<int>(i % nx);
        }
    }
    const int blocks = 2048;
    double *x = nullptr, *out = nullptr;
    int *di = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HL_TRY_HIP(hipMalloc(&x, sizeof(double) * nx));
    HL_TRY_HIP(hipMalloc(&di, sizeof(int) * m));
    HL_TRY_HIP(hipMalloc(&out, sizeof(double) * blocks * (kBlock / kWave)));
    HL_TRY_HIP(hipMemset(x, 0, sizeof(double) * nx));
    HL_TRY_HIP(hipMemcpy(di, idx.data(), sizeof(int) * m, hipMemcpyHostToDevice));
    HL_TRY_HIP(hipEventCreate(&e0));
    HL_TRY_HIP(hipEventCreate(&e1));
    rc = launch_gather(x, di, m, out, blocks, nullptr);
    if (!rc) rc = hipEventRecord(e0, nullptr) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    for (int i = 0; i < reps && !rc; ++i) rc = launch_gather(x, di, m, out, blocks, nullptr);
    float ms = 0.0f;
    if (!rc) rc = hipEventRecord(e1, nullptr) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    if (!rc) rc = hipEventElapsedTime(&ms, e0, e1) == hipSuccess ? 0 : HL_ERR_HIP_BASE;
    *seconds = 1e-3 * ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(x);
    (void)hipFree(di);
    (void)hipFree(out);
    return rc;
} catch (...) {
    return exception_code();
}

int hl_tune_set(const char *key, int value)
try {
    if (!key) return HL_ERR_ARG;
    Tuning &t = tuning();
    if (!std::strcmp(key, "cgs_blocks_max") && value >= 1 && value <= kMaxReduceBlocks)
        t.cgs_blocks_max = value;
    else if (!std::strcmp(key, "cgs_flags") && value >= 0 && value <= 15 && !(value & 4))
        t.cgs_flags = value;
    else if (!std::strcmp(key, "f32_phase_tol_exp") && value >= 3 && value <= 9)
        t.f32_phase_tol_exp = value;
    else if (!std::strcmp(key, "cgs_f32_tiles") && (value == 1 || value == 2))
        t.cgs_f32_tiles = value;
    else if (!std::strcmp(key, "dgks_skip_exp") && value >= 0 && value <= 16)
        t.dgks_skip_exp = value;
    else if (!std::strcmp(key, "batch_engines") && value >= 1 && value <= 8)
        t.batch_engines = value;
    else {
        set_error("hl_tune_set: unknown key or value out of range: %s=%d", key, value);
        return HL_ERR_ARG;
    }
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_tune_get(const char *key, int *value)
try {
    if (!key || !value) return HL_ERR_ARG;
    const Tuning &t = tuning();
    if (!std::strcmp(key, "cgs_blocks_max"))
        *value = t.cgs_blocks_max;
    else if (!std::strcmp(key, "cgs_flags"))
        *value = t.cgs_flags;
    else if (!std::strcmp(key, "f32_phase_tol_exp"))
        *value = t.f32_phase_tol_exp;
    else if (!std::strcmp(key, "cgs_f32_tiles"))
        *value = t.cgs_f32_tiles;
    else if (!std::strcmp(key, "dgks_skip_exp"))
        *value = t.dgks_skip_exp;
    else if (!std::strcmp(key, "batch_engines"))
        *value = t.batch_engines;
    else {
        set_error("hl_tune_get: unknown key %s", key);
        return HL_ERR_ARG;
    }
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_vq(int dev, int n, int K, int N, const double *V, long long ldv, const double *Q, int ldq, double *Out,
          long long ldo, int use_mfma, void *stream)
try {
    if (n < 0 || K <= 0 || N <= 0 || K > kMaxNcv || !V || !Q || !Out || ldv < n || ldo < n || ldq < K) {
        set_error("hl_vq: bad arguments");
        return HL_ERR_ARG;
    }
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    Reduce red;
    return launch_vq(n, K, N, V, VLayout{ldv, false}, Q, ldq, Out, VLayout{ldo, false}, use_mfma, -1, nullptr, 0.0,
                     0.0, red, nullptr, static_cast<hipStream_t>(stream));
} catch (...) {
    return exception_code();
}

int hl_host_tridiag_eig(int n, double *d, double *e, double *z, int ldz)
try {
    if (n < 0 || (n > 0 && (!d || (n > 1 && !e) || !z || ldz < n))) return HL_ERR_ARG;
    return tridiag_eig(n, d, e, z, ldz);
} catch (...) {
    return exception_code();
}

int hl_host_tridiag_eig_last(int n, double *d, double *e, double *zlast)
try {
    if (n < 0 || (n > 0 && (!d || (n > 1 && !e) || !zlast))) return HL_ERR_ARG;
    return tridiag_eig_last(n, d, e, zlast);
} catch (...) {
    return exception_code();
}

void hl_host_dlartg(double f, double g, double *c, double *s, double *r)
{
    double cc = 0.0, ss = 0.0, rr = 0.0;
    dlartg(f, g, cc, ss, rr);
    if (c) *c = cc;
    if (s) *s = ss;
    if (r) *r = rr;
}

int hl_host_householder_q(int m, int k, double *a, int lda, double *tau, double *q, int ldq)
try {
    if (m < 0 || k < 0 || k > m || (m > 0 && (!a || !tau || !q || lda < m || ldq < m))) return HL_ERR_ARG;
    householder_q(m, k, a, lda, tau, q, ldq);
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_host_restart_chase(int nev, int np, const double *shift, double *sub, double *dia, double *q)
try {
    if (nev <= 0 || np < 0 || (np > 0 && !shift) || !sub || !dia || !q) return HL_ERR_ARG;
    restart_chase(nev, np, shift, sub, dia, q);
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_profile_last(hl_profile *out)
try {
    if (!out) return HL_ERR_ARG;
    *out = thread_profile();
    return HL_OK;
} catch (...) {
    return exception_code();
}

int hl_ds_solve_csr(hl_csr_t A, int nev, int ncv, double *values, double *vectors, int ldv,
                    const hl_options *opts, hl_info *info)
try {
    if (!A) return HL_ERR_ARG;
    if (A->plan.nrows != A->plan.ncols) {
        set_error("ds_solver: the operator must be square");
        return HL_ERR_ARG;
    }
    hl::Member m;
    m.A = &A->plan;
    return solve_common(m, A->plan.dev, A->plan.nrows, nev, ncv, values, vectors, ldv, opts, info);
} catch (...) {
    return exception_code();
}

int hl_ds_solve_host_op(int dev, int n, hl_host_op_fn op, void *ctx, int nev, int ncv, double *values,
                        double *vectors, int ldv, const hl_options *opts, hl_info *info)
try {
    if (!op) return HL_ERR_ARG;
    hl::Member m;
    m.host_op = op;
    m.host_ctx = ctx;
    return solve_common(m, dev, n, nev, ncv, values, vectors, ldv, opts, info);
} catch (...) {
    return exception_code();
}

int hl_ds_solve_batch(int dev, hl_solve_item *items, int count, int max_active, double hbm_budget)
try {
    if (count < 0 || (count > 0 && !items)) return HL_ERR_ARG;
    int rc = check_device(dev);
    if (rc) return rc;
    DeviceGuard g(dev);
    std::vector<hl::Member> ms(static_cast<size_t>(count));
    std::vector<hl::Member *> queue;
    bool profile = false, graphs = true;
    double last_arrival = 0.0;
    for (int i = 0; i < count; ++i) {
        hl_solve_item &it = items[i];
        std::memset(&it.info, 0, sizeof(it.info));
        it.status = HL_ERR_ARG;
        it.t_admit = it.t_done = -1.0;
        const hl::CsrPlan *p = it.A ? &it.A->plan : nullptr;
        if (!p || p->dev != dev || p->nrows != p->ncols) {
            set_error("hl_ds_solve_batch: item %d needs a square operator on device %d", i, dev);
            continue;
        }
        if (!solve_args_ok(p->nrows, it.nev, it.ncv, it.ldv, it.values, false, it.vectors, it.opts, true,
                           "hl_ds_solve_batch") ||
            it.arrival < last_arrival) {
            if (it.arrival < last_arrival) set_error("hl_ds_solve_batch: arrivals must not decrease (item %d)", i);
            continue;
        }
        last_arrival = it.arrival;
        hl::Member &m = ms[i];
        m.A = p;
        m.dev = dev;
        m.n = p->nrows;
        m.nev0 = it.nev;
        m.ncv = it.ncv;
        m.opts = it.opts;
        m.values = it.values;
        m.vectors = it.vectors;
        m.ldv_host = it.ldv;
        m.arrival = it.arrival;
        profile |= (it.opts.flags & HL_FLAG_PROFILE) != 0;
        graphs &= (it.opts.flags & HL_FLAG_NO_GRAPH) == 0;
        queue.push_back(&m);
    }
    // several engines, each on its own stream and host thread, share the
    // queue: while one runs its restart logic on the host, the others' sweeps
    // keep the device busy (a profiled batch runs on one engine)
    const int engines =
        profile ? 1 : std::max(1, std::min(tuning().batch_engines, static_cast<int>(queue.size())));
    if (engines == 1) {
        hl::Engine eng;
        if ((rc = eng.init(dev))) return rc;
        eng.prof.on = profile;
        eng.graphs = graphs;
        rc = eng.run(queue, max_active, hbm_budget);
        if (profile && !rc) thread_profile() = eng.prof.tot;
    } else {
        hl::Admission adm;
        adm.queue = &queue;
        adm.max_active = max_active;
        adm.budget = hbm_budget;
        adm.t_start = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
        adm.engines = engines;
        std::mutex hip_mu;            // captures vs. the other engines' allocations
        std::shared_mutex sync_rw;    // captures vs. the other engines' waits
        std::vector<int> rcs(static_cast<size_t>(engines), HL_OK);
        auto serve = [&](int k) {
            DeviceGuard gk(dev);
            hl::Engine eng;
            eng.capture_mu = &hip_mu;
            eng.sync_rw = &sync_rw;
            eng.graphs = graphs;
            int r = eng.init(dev); // a full host pool each: their restart phases rarely coincide
            if (!r) r = eng.run_shared(adm);
            rcs[static_cast<size_t>(k)] = r;
        };
        std::vector<std::thread> th;
        for (int k = 1; k < engines; ++k) th.emplace_back(serve, k);
        serve(0);
        for (std::thread &t : th) t.join();
        for (int r : rcs)
            if (r && !rc) rc = r;
    }
    for (int i = 0; i < count; ++i) {
        const hl::Member &m = ms[i];
        if (!m.A) continue;
        hl_solve_item &it = items[i];
        it.status = m.err ? m.err : (m.phase == hl::Member::DONE ? m.ret : HL_ERR_ARG);
        it.info = m.info;
        it.info.ret = it.status;
        it.t_admit = m.t_admit;
        it.t_done = m.t_done;
    }
    return rc;
} catch (...) {
    return exception_code();
}

} // extern "C"

