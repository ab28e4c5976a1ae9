// Vendor-library control: C = A * B with rocSPARSE's SpGEMM, timed at the
// boundary of the reference's cuSPARSE control (src/cusparse_core_benchmark.h:4-30):
// inputs and descriptors are device resident before a sample starts; a
// sample runs the buffer-size stage, the nnz stage (row pointers of C),
// allocates C's columns and values, runs the compute stage and synchronises.
// The scratch buffer is allocated in the untimed warm-up sample and reused
// (the reference's steady-state protocol).  B = A^T for aat = 1 is built
// with rocSPARSE's csr2csc before timing.
//
// This is a separate library (librtt_control.so): the SpGEMM library does not
// depend on rocSPARSE.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rtt_spgemm.h"

namespace {

char g_err[512];

int fail(const char *what, int code)
{
    std::snprintf(g_err, sizeof(g_err), "%s failed (%d)", what, code);
    return 1;
}

#define CK_HIP(call)                                                   \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess)                                          \
            return fail(#call, static_cast<int>(e_));                  \
    } while (0)
#define CK_SP(call)                                                    \
    do {                                                               \
        rocsparse_status e_ = (call);                                  \
        if (e_ != rocsparse_status_success)                            \
            return fail(#call, static_cast<int>(e_));                  \
    } while (0)

// Device CSR (int32 indices, FP64 values) that frees itself.
struct DevCsr {
    int rows = 0, cols = 0;
    int64_t nnz = 0;
    int *rp = nullptr, *ci = nullptr;
    double *val = nullptr;
    ~DevCsr()
    {
        (void)hipFree(rp);
        (void)hipFree(ci);
        (void)hipFree(val);
    }
};

int upload(const rtt_smatrix &h, DevCsr *d)
{
    d->rows = h.m;
    d->cols = h.n;
    d->nnz = h.nnz;
    CK_HIP(hipMalloc(&d->rp, (static_cast<size_t>(h.m) + 1) * sizeof(int)));
    CK_HIP(hipMalloc(&d->ci, std::max<size_t>(1, h.nnz) * sizeof(int)));
    CK_HIP(hipMalloc(&d->val, std::max<size_t>(1, h.nnz) * sizeof(double)));
    CK_HIP(hipMemcpy(d->rp, h.rowpointer, (static_cast<size_t>(h.m) + 1) * sizeof(int),
                     hipMemcpyHostToDevice));
    if (h.nnz > 0) {
        CK_HIP(hipMemcpy(d->ci, h.columnindex, static_cast<size_t>(h.nnz) * sizeof(int),
                         hipMemcpyHostToDevice));
        CK_HIP(hipMemcpy(d->val, h.value, static_cast<size_t>(h.nnz) * sizeof(double),
                         hipMemcpyHostToDevice));
    }
    return 0;
}

// A device allocation freed when it goes out of scope.
template <typename T>
struct DevPtr {
    T *p = nullptr;
    DevPtr() = default;
    DevPtr(const DevPtr &) = delete;
    DevPtr &operator=(const DevPtr &) = delete;
    ~DevPtr() { (void)hipFree(p); }
};

struct Handle {
    rocsparse_handle h = nullptr;
    ~Handle()
    {
        if (h)
            rocsparse_destroy_handle(h);
    }
};

struct Descr {
    rocsparse_spmat_descr d = nullptr;
    ~Descr()
    {
        if (d)
            rocsparse_destroy_spmat_descr(d);
    }
};

int make_descr(const DevCsr &m, Descr *out)
{
    CK_SP(rocsparse_create_csr_descr(&out->d, m.rows, m.cols, m.nnz, m.rp, m.ci, m.val,
                                     rocsparse_indextype_i32, rocsparse_indextype_i32,
                                     rocsparse_index_base_zero, rocsparse_datatype_f64_r));
    return 0;
}

// One product in the order rocsparse_spgemm documents for one C descriptor
// (rocsparse/internal/generic/rocsparse_spgemm.h, "requires three stages"):
// buffer-size stage, nnz stage (C's row pointers), allocation of C's columns
// and values, compute stage, synchronisation.  The scratch buffer is owned by
// the caller and grown here when the buffer-size stage asks for more (only
// the untimed warm-up sample does, since the inputs do not change).
int spgemm_once(rocsparse_handle h, const Descr &a, const Descr &b, const Descr &d, int rows,
                int cols, void **buffer, size_t *capacity, double *ms, int64_t *nnz_c,
                double *checksum)
{
    const double alpha = 1.0, beta = 0.0;
    DevPtr<int> crp, cci;
    DevPtr<double> cval;
    CK_HIP(hipMalloc(&crp.p, (static_cast<size_t>(rows) + 1) * sizeof(int)));
    Descr c;
    CK_SP(rocsparse_create_csr_descr(&c.d, rows, cols, 0, crp.p, nullptr, nullptr,
                                     rocsparse_indextype_i32, rocsparse_indextype_i32,
                                     rocsparse_index_base_zero, rocsparse_datatype_f64_r));
    CK_HIP(hipDeviceSynchronize());
    const auto t0 = std::chrono::steady_clock::now();
    size_t bytes = 0;
    CK_SP(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, a.d,
                           b.d, &beta, d.d, c.d, rocsparse_datatype_f64_r,
                           rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_buffer_size,
                           &bytes, nullptr));
    if (bytes > *capacity || *buffer == nullptr) {
        (void)hipFree(*buffer);
        *buffer = nullptr;
        *capacity = 0;
        CK_HIP(hipMalloc(buffer, std::max<size_t>(bytes, 1)));
        *capacity = std::max<size_t>(bytes, 1);
    }
    CK_SP(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, a.d,
                           b.d, &beta, d.d, c.d, rocsparse_datatype_f64_r,
                           rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_nnz, &bytes,
                           *buffer));
    int64_t r = 0, cc = 0, nz = 0;
    CK_SP(rocsparse_spmat_get_size(c.d, &r, &cc, &nz));
    CK_HIP(hipMalloc(&cci.p, std::max<int64_t>(1, nz) * sizeof(int)));
    CK_HIP(hipMalloc(&cval.p, std::max<int64_t>(1, nz) * sizeof(double)));
    CK_SP(rocsparse_csr_set_pointers(c.d, crp.p, cci.p, cval.p));
    CK_SP(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, a.d,
                           b.d, &beta, d.d, c.d, rocsparse_datatype_f64_r,
                           rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_compute, &bytes,
                           *buffer));
    CK_HIP(hipDeviceSynchronize());
    const auto t1 = std::chrono::steady_clock::now();
    *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    *nnz_c = nz;
    if (checksum) {
        std::vector<double> v(static_cast<size_t>(nz));
        if (nz > 0)
            CK_HIP(hipMemcpy(v.data(), cval.p, static_cast<size_t>(nz) * sizeof(double),
                             hipMemcpyDeviceToHost));
        double s = 0.0;
        for (double x : v)
            s += x;
        *checksum = s;
    }
    return 0;
}

} // namespace

extern "C" {

const char *rtt_control_last_error(void) { return g_err; }

// C = A * A (aat = 0) or A * A^T (aat = 1) with rocsparse_spgemm.  median_ms
// receives the median of `iterations` timed samples after `warmup` untimed
// ones; nnz_c the nonzeros of C; checksum (may be NULL) the sum of C's
// values of the last sample.  Returns 0 on success.
int rtt_control_rocsparse(const rtt_smatrix *a_host, int aat, int warmup, int iterations,
                          double *median_ms, long long *nnz_c, double *checksum)
{
    g_err[0] = '\0';
    if (!a_host || !median_ms || !nnz_c || iterations < 1 || warmup < 0)
        return fail("arguments", -1);
    DevCsr a, b;
    if (upload(*a_host, &a))
        return 1;
    Handle h;
    CK_SP(rocsparse_create_handle(&h.h));
    // The library resolved at load time must be the one these declarations
    // come from: a process that already holds another rocsparse (same soname,
    // other release) would bind this code to it.
    int version = 0;
    CK_SP(rocsparse_get_version(h.h, &version));
    const int built = ROCSPARSE_VERSION_MAJOR * 100000 + ROCSPARSE_VERSION_MINOR * 100 +
                      ROCSPARSE_VERSION_PATCH;
    if (version != built) {
        std::snprintf(g_err, sizeof(g_err), "rocsparse %d loaded, built against %d", version,
                      built);
        return 1;
    }
    if (aat) {
        // B = A^T: the CSC of A is the CSR of A^T.
        b.rows = a.cols;
        b.cols = a.rows;
        b.nnz = a.nnz;
        CK_HIP(hipMalloc(&b.rp, (static_cast<size_t>(b.rows) + 1) * sizeof(int)));
        CK_HIP(hipMalloc(&b.ci, std::max<int64_t>(1, b.nnz) * sizeof(int)));
        CK_HIP(hipMalloc(&b.val, std::max<int64_t>(1, b.nnz) * sizeof(double)));
        size_t tb = 0;
        CK_SP(rocsparse_csr2csc_buffer_size(h.h, a.rows, a.cols, static_cast<int>(a.nnz), a.rp,
                                            a.ci, rocsparse_action_numeric, &tb));
        DevPtr<void> tmp;
        CK_HIP(hipMalloc(&tmp.p, std::max<size_t>(tb, 1)));
        CK_SP(rocsparse_dcsr2csc(h.h, a.rows, a.cols, static_cast<int>(a.nnz), a.val, a.rp, a.ci,
                                 b.val, b.ci, b.rp, rocsparse_action_numeric,
                                 rocsparse_index_base_zero, tmp.p));
    } else if (upload(*a_host, &b)) {
        return 1;
    }
    // D (beta = 0) is an empty matrix of C's shape.
    DevCsr d;
    d.rows = a.rows;
    d.cols = b.cols;
    d.nnz = 0;
    CK_HIP(hipMalloc(&d.rp, (static_cast<size_t>(d.rows) + 1) * sizeof(int)));
    CK_HIP(hipMemset(d.rp, 0, (static_cast<size_t>(d.rows) + 1) * sizeof(int)));
    CK_HIP(hipMalloc(&d.ci, sizeof(int)));
    CK_HIP(hipMalloc(&d.val, sizeof(double)));
    Descr da, db, dd;
    if (make_descr(a, &da) || make_descr(b, &db) || make_descr(d, &dd))
        return 1;
    DevPtr<void> buffer;
    size_t capacity = 0;
    std::vector<double> times;
    int64_t nz = 0;
    for (int it = 0; it < warmup + iterations; ++it) {
        double ms = 0.0;
        const bool last = it == warmup + iterations - 1;
        if (spgemm_once(h.h, da, db, dd, a.rows, b.cols, &buffer.p, &capacity, &ms, &nz,
                        last ? checksum : nullptr))
            return 1;
        if (it >= warmup)
            times.push_back(ms);
    }
    std::sort(times.begin(), times.end());
    *median_ms = times[times.size() / 2];
    *nnz_c = static_cast<long long>(nz);
    return 0;
}

} // extern "C"
