// HLanc.hpp -- drop-in C++ interface of the MI355X (gfx950) build of HLanc.
//
// Keeps the public names, signatures and behaviour of the reference header
// (HLanc.hpp:15-67 utilities, 74-392 dcsr_matrix_t, 476-572 dcsrmv_operator_gpu,
// 750-893 dcsrmv_operator_mgpu, 1868-1872 ds_solver_gpu) on top of the C ABI in
// hlanc_capi.h (library libhlanc_hip.so).  Host code only: include it from any
// C++11 compiler and link -lhlanc_hip.
//
//  * ds_solver_gpu<Op>(n, nev, ncv, values, vectors, ldv, op, dev): returns 0 on
//    success and 1 on the reference's numerical failures; values/vectors are
//    written only on success.  When Op is this header's dcsrmv_operator_gpu or
//    dcsrmv_operator_mgpu on device dev, the whole solve stays on the GPU;
//    any other callable is called with host pointers, as the reference does.
//  * An extended overload takes hl_options (tol, maxiter, start vector, ...)
//    and fills hl_info (restarts, SpMV count, timing).
//  * Hard errors (HIP failures, bad arguments, unreadable files) print a
//    message and exit(), like the reference's SAFE_CALL / error_exit.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "hlanc_capi.h"

#define HLANC_VERSION_MAJOR 1
#define HLANC_VERSION_MINOR 0

#ifndef CEIL_DIV
#define CEIL_DIV(a, b) (((a) + (b) - 1) / (b))
#endif
#ifndef MIN
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#endif

double const c_23 = 0.66666666666666663;
double const c_d0 = 0.0;
double const c_d1 = 1.0;
double const c_dn1 = -1.0;

namespace HLanc {

// ----------------------------------------------------------------- utilities
// check_call / SAFE_CALL (HLanc.hpp:38-50): non-zero status prints and exits.
// hl_* status codes print the library's last error message.
template <typename status_t>
inline void check_call(status_t error, char const *file, size_t line, char const *call)
{
    if (0 == error) return;
    std::fprintf(stderr, "%s failed in \"%s:%lu\": %d (%s)\n", call, file, (unsigned long)line, (int)error,
                 hl_last_error());
    std::exit((int)error);
}
#define SAFE_CALL(call) HLanc::check_call((call), __FILE__, __LINE__, #call)

inline void error_exit(char const *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::exit(1);
}

inline void warn(char const *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
}

// ------------------------------------------------------------- dcsr_matrix_t
class dcsr_matrix_t {
  public:
    // Reads "pp", "mm" or "bin" (HLanc.hpp:103-113); errors print and exit.
    dcsr_matrix_t(char const *format, char const *filename, bool is_symmetric = false)
    {
        hl_matrix m;
        const int rc = hl_matrix_read(format, filename, is_symmetric ? 1 : 0, &m);
        if (rc) error_exit("%s\n", hl_last_error());
        csr_row_ptr.assign(m.row_ptr, m.row_ptr + m.nrows + 1);
        csr_column_index.assign(m.col, m.col + m.nnz);
        csr_value.assign(m.val, m.val + m.nnz);
        num_columns = m.ncols;
        this->is_symmetric = m.sym != 0;
        hl_matrix_free(&m);
    }
    // Extension: adopt in-memory CSR arrays (zero-based, int32).
    dcsr_matrix_t(int num_rows, int num_columns_, int const *row_ptr, int const *col, double const *val,
                  bool is_symmetric_ = false)
        : csr_row_ptr(row_ptr, row_ptr + num_rows + 1), csr_column_index(col, col + row_ptr[num_rows]),
          csr_value(val, val + row_ptr[num_rows]), num_columns(num_columns_), is_symmetric(is_symmetric_)
    {
    }
    ~dcsr_matrix_t() {}

    void save_pp(char const *filename) const { save("pp", filename); }
    void save_mm(char const *filename) const { save("mm", filename); }
    void save_bin(char const *filename) const { save("bin", filename); }

    int get_num_rows() const { return (int)csr_row_ptr.size() - 1; }
    int get_nnz() const { return (int)csr_value.size(); }
    int get_num_columns() const { return num_columns; }
    int const *get_row_ptr() const { return csr_row_ptr.data(); }
    int const *get_column_index() const { return csr_column_index.data(); }
    double const *get_value() const { return csr_value.data(); }
    bool symmetric() const { return is_symmetric; }

  protected:
    void save(char const *fmt, char const *filename) const
    {
        hl_matrix m;
        m.nrows = get_num_rows();
        m.ncols = num_columns;
        m.nnz = get_nnz();
        m.sym = is_symmetric ? 1 : 0;
        m.row_ptr = const_cast<int *>(csr_row_ptr.data());
        m.col = const_cast<int *>(csr_column_index.data());
        m.val = const_cast<double *>(csr_value.data());
        if (hl_matrix_write(fmt, filename, &m)) error_exit("%s\n", hl_last_error());
    }

    std::vector<int> csr_row_ptr;
    std::vector<int> csr_column_index;
    std::vector<double> csr_value;
    int num_columns;
    bool is_symmetric;
};

// ------------------------------------------------------- dcsrmv_operator_naive
// Host reference SpMV kept for API compatibility (deprecated/debug in the
// reference, HLanc.hpp:399-434; README "Deprecated").
class dcsrmv_operator_naive {
  public:
    dcsrmv_operator_naive(dcsr_matrix_t const &m) : m(m) {}
    ~dcsrmv_operator_naive() {}
    void operator()(double const *x, double *y) const
    {
        const int n = m.get_num_rows();
        int const *rp = m.get_row_ptr();
        int const *ci = m.get_column_index();
        double const *v = m.get_value();
        if (m.symmetric()) {
            for (int i = 0; i < n; ++i) y[i] = 0.0;
            for (int r = 0; r < n; ++r)
                for (int p = rp[r]; p < rp[r + 1]; ++p) {
                    y[r] += x[ci[p]] * v[p];
                    if (r != ci[p]) y[ci[p]] += x[r] * v[p];
                }
        } else {
            for (int r = 0; r < n; ++r) {
                double s = 0.0;
                for (int p = rp[r]; p < rp[r + 1]; ++p) s += x[ci[p]] * v[p];
                y[r] = s;
            }
        }
    }

  protected:
    dcsr_matrix_t const &m;
};

// --------------------------------------------------------- dcsrmv_operator_gpu
// y = A x on one GPU (HLanc.hpp:476-572).  x and y may be host memory or device
// memory of `dev`; host y means the result is complete on return.
class dcsrmv_operator_gpu {
  public:
    dcsrmv_operator_gpu(dcsr_matrix_t const &m, int dev = 0) : dev(dev), h(nullptr)
    {
        SAFE_CALL(hl_csr_create(dev, m.get_num_rows(), m.get_num_columns(), m.get_nnz(), m.get_row_ptr(),
                                m.get_column_index(), m.get_value(), m.symmetric() ? 1 : 0, HL_PREC_F64, &h));
    }
    ~dcsrmv_operator_gpu()
    {
        if (h) hl_csr_destroy(h);
    }
    dcsrmv_operator_gpu(dcsrmv_operator_gpu const &) = delete;
    dcsrmv_operator_gpu &operator=(dcsrmv_operator_gpu const &) = delete;

    void operator()(double const *x, double *y) const { SAFE_CALL(hl_csr_apply(h, x, y)); }

    hl_csr_t handle() const { return h; }
    int device() const { return dev; }

  protected:
    int dev;
    hl_csr_t h;
};

// -------------------------------------------------------- dcsrmv_operator_mgpu
// Multi-GPU SpMV of the reference (HLanc.hpp:750-893) on a one-GPU build: the
// first device selected by `mask` holds the whole matrix.  Symmetric storage is
// rejected as in the reference (HLanc.hpp:798-800).
class dcsrmv_operator_mgpu : public dcsrmv_operator_gpu {
  public:
    dcsrmv_operator_mgpu(dcsr_matrix_t const &m, unsigned mask = (unsigned)-1)
        : dcsrmv_operator_gpu(check(m), first_device(mask))
    {
    }

  private:
    static dcsr_matrix_t const &check(dcsr_matrix_t const &m)
    {
        if (m.symmetric()) error_exit("dcsrmv_operator_mgpu can only be constructed with a general matrix.\n");
        return m;
    }
    static int first_device(unsigned mask)
    {
        int count = 0;
        hl_device_count(&count);
        for (int d = 0; d < count && d < 32; ++d)
            if (mask & (1u << d)) return d;
        error_exit("dcsrmv_operator_mgpu: no device selected by mask 0x%x\n", mask);
        return -1;
    }
};

// -------------------------------------------------------------- ds_solver_gpu
namespace detail {

template <typename Op>
struct host_op_ref {
    Op const &op;
};

template <typename Op>
void host_op_trampoline(void *ctx, double const *x, double *y)
{
    static_cast<host_op_ref<Op> *>(ctx)->op(x, y);
}

template <typename Op>
struct is_device_operator
    : std::integral_constant<bool, std::is_base_of<dcsrmv_operator_gpu, typename std::decay<Op>::type>::value> {
};

template <typename Op>
inline int solve(std::true_type, int n, int nev, int ncv, double *values, double *vectors, int ldv, Op const &op,
                 int dev, hl_options const *o, hl_info *info)
{
    if (op.device() == dev) return hl_ds_solve_csr(op.handle(), nev, ncv, values, vectors, ldv, o, info);
    // operator on another device: call it through host pointers as the reference would
    host_op_ref<Op> ref{op};
    return hl_ds_solve_host_op(dev, n, &host_op_trampoline<Op>, &ref, nev, ncv, values, vectors, ldv, o, info);
}

template <typename Op>
inline int solve(std::false_type, int n, int nev, int ncv, double *values, double *vectors, int ldv, Op const &op,
                 int dev, hl_options const *o, hl_info *info)
{
    host_op_ref<Op> ref{op};
    return hl_ds_solve_host_op(dev, n, &host_op_trampoline<Op>, &ref, nev, ncv, values, vectors, ldv, o, info);
}

inline int finish(int rc)
{
    if (rc == HL_OK || rc == HL_NOT_CONVERGED) return rc;
    std::fprintf(stderr, "ds_solver_gpu failed: %d (%s)\n", rc, hl_last_error());
    std::exit(rc);
    return rc;
}

} // namespace detail

// Extended form: tolerance, restart cap, start vector, precision; statistics.
// A basis wider than this build takes (HL_MAX_NCV, HL_MAX_NCV_F32) is not a hard
// error: the call warns and returns 1, as for a solve that did not converge.
template <typename dspmv_operator_t>
int ds_solver_gpu(int n, int nev, int ncv, double *values, double *vectors, int ldv,
                  dspmv_operator_t const &dspmv_operator, int dev, hl_options const &opts, hl_info *info = nullptr)
{
    const int width = opts.basis_prec == HL_PREC_F32 ? HL_MAX_NCV_F32 : HL_MAX_NCV;
    if (ncv > width && ncv <= n && nev < ncv) {
        warn("ds_solver_gpu: ncv = %d exceeds this build's basis width (%d); returning 1\n", ncv, width);
        return 1;
    }
    return detail::finish(detail::solve(detail::is_device_operator<dspmv_operator_t>(), n, nev, ncv, values, vectors,
                                        ldv, dspmv_operator, dev, &opts, info));
}

// ---------------------------------------------------------- batched solves
// Extension without a counterpart in the reference: many solves of operators on
// one device in one call, continuously batched (hl_ds_solve_batch).  Each
// solve's values and vectors equal what ds_solver_gpu returns for the same
// operator and options, bit for bit.
struct batch_solve {
    dcsrmv_operator_gpu const *op = nullptr;
    int nev = 0, ncv = 0;
    double *values = nullptr;  // host, >= nev
    double *vectors = nullptr; // host, column-major ldv x ncv, or nullptr: values only
    int ldv = 0;
    hl_options opts;           // the reference defaults unless changed
    double arrival = 0.0;      // seconds after the call; admitted no earlier
    hl_info info;              // out
    int status = HL_ERR_ARG;   // out: HL_OK, HL_NOT_CONVERGED or an error code
    batch_solve() { hl_options_default(&opts); std::memset(&info, 0, sizeof(info)); }
};

// Returns HL_OK once every solve has run (see each status); hard errors exit
// like SAFE_CALL.  max_active <= 0 and hbm_budget <= 0 mean no cap.
inline int ds_solver_gpu_batch(std::vector<batch_solve> &solves, int dev = 0, int max_active = 0,
                               double hbm_budget = 0.0)
{
    std::vector<hl_solve_item> items(solves.size());
    for (size_t i = 0; i < solves.size(); ++i) {
        hl_solve_item &it = items[i];
        std::memset(&it, 0, sizeof(it));
        const batch_solve &b = solves[i];
        it.A = b.op ? b.op->handle() : nullptr;
        it.nev = b.nev;
        it.ncv = b.ncv;
        it.values = b.values;
        it.vectors = b.vectors;
        it.ldv = b.ldv;
        it.opts = b.opts;
        it.arrival = b.arrival;
    }
    const int rc = hl_ds_solve_batch(dev, items.data(), static_cast<int>(items.size()), max_active, hbm_budget);
    if (rc) {
        std::fprintf(stderr, "ds_solver_gpu_batch failed: %d (%s)\n", rc, hl_last_error());
        std::exit(rc);
    }
    for (size_t i = 0; i < solves.size(); ++i) {
        solves[i].info = items[i].info;
        solves[i].status = items[i].status;
    }
    return HL_OK;
}

// HLanc.hpp:1868-1872: tol = machine epsilon, 200 restarts, largest algebraic.
template <typename dspmv_operator_t>
int ds_solver_gpu(int n, int nev, int ncv, double *values, double *vectors, int ldv,
                  dspmv_operator_t const &dspmv_operator, int dev = 0)
{
    hl_options o;
    hl_options_default(&o);
    return ds_solver_gpu(n, nev, ncv, values, vectors, ldv, dspmv_operator, dev, o, nullptr);
}

} // namespace HLanc
