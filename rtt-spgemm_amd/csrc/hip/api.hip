// extern "C" entry points of the device path: A preparation, B rebuild,
// nnzCub, the SpGEMM core and the host export of C.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <unordered_set>
#include <vector>

#include "internal.hpp"

// Scratch of rtt_tile_values_noalloc: a fixed device array instead of an
// allocation (each GPU holds its own copy until the process ends).  Host code
// uses only the per-device address hipGetSymbolAddress returned for it.
constexpr size_t kNoAllocScratchBytes = 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_dev[kMaxDevices] = {};   // address of `scratch` on each device
} // namespace

struct rtt_prepared_a {
    rtt::DevCsr csr;          // input CSR, rows column-sorted (B source)
    rtt::DevCsr csr_tiled;    // reordered copy when a permutation is given
    rtt::DevCsr csr_t;        // transposed input (reordered AAT rebuilds)
    rtt::DevTiles tiles;      // A: 8x4 row-major hybrid tiles
    int rows = 0, cols = 0;
    int active_rows = 0, active_inner = 0;
    int dense_threshold = 24;
    bool identity = true;     // both permutations identity
    bool row_identity = true; // row permutation identity
    std::vector<int> row_new_to_old, inner_new_to_old;
    rtt::DevBuf d_inner_new_to_old;
    rtt::BuildWork build;
    rtt::SpgemmWork core;
    rtt::Events events;
    rtt::DevC last_c;
    bool has_c = false;
    uint64_t id = 0;          // unique per preparation (never reused in a process)
    // The last product: B handle and build generation, and whether every row
    // ran on the tile-pair kernel (C's whole structure was written first).
    const rtt_dynamic_b *last_b = nullptr;
    uint64_t last_b_gen = 0;
    bool last_all_tile = false;
};

// A B is built for one prepared A: its per-A-tile pair extents are indexed by
// that A's tiles and, for B = A^T at tile level, its payload is that A's.
struct rtt_dynamic_b {
    rtt::DevTiles tiles;
    rtt::BuildWork build;
    int aat = 0;
    uint64_t a_id = 0;        // rtt_prepared_a::id of the A it was built for
    uint64_t gen = 0;         // structure builds of this handle (values updates keep it)
};

struct rtt_device_csr {
    rtt::DevCsr csr;
    rtt::BuildWork build;
};

namespace {

using rtt::fail;

// Thread per row of C: its entry count and value sum, from C's tile
// structure (a checking utility; walks the row's tiles in order).
__global__ void c_row_stats_kernel(int64_t rows, const int64_t *__restrict__ tile_ptr,
                                   const uint64_t *__restrict__ mask,
                                   const int64_t *__restrict__ tile_nnz,
                                   const double *__restrict__ val, int64_t *__restrict__ nnz_out,
                                   double *__restrict__ sum_out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows)
        return;
    const int64_t tr = i / 8;
    const int r = static_cast<int>(i % 8);
    int64_t n = 0;
    double sum = 0.0;
    for (int64_t t = tile_ptr[tr]; t < tile_ptr[tr + 1]; ++t) {
        const uint64_t m = mask[t];
        const uint64_t below = r ? (m & ((1ull << (8 * r)) - 1ull)) : 0ull;
        const int before = __popcll(below);
        const int here = __popc(static_cast<uint32_t>((m >> (8 * r)) & 0xFFull));
        const int64_t base = tile_nnz[t] + before;
        for (int x = 0; x < here; ++x)
            sum += val[base + x];
        n += here;
    }
    nnz_out[i] = n;
    sum_out[i] = sum;
}

// Identities of the prepared A handles alive in this process, so that a B
// bound to a destroyed A is refused instead of dereferenced.
std::atomic<uint64_t> g_next_a_id{1};
std::mutex g_live_mu;
std::unordered_set<uint64_t> g_live_a;

uint64_t register_a()
{
    const uint64_t id = g_next_a_id.fetch_add(1);
    std::lock_guard<std::mutex> lock(g_live_mu);
    g_live_a.insert(id);
    return id;
}

void unregister_a(uint64_t id)
{
    std::lock_guard<std::mutex> lock(g_live_mu);
    g_live_a.erase(id);
}

bool a_alive(uint64_t id)
{
    std::lock_guard<std::mutex> lock(g_live_mu);
    return g_live_a.count(id) != 0;
}

// B must have been built (rtt_rebuild_b / rtt_rebuild_b_general) for this A.
int check_binding(const rtt_prepared_a *a, const rtt_dynamic_b *b, const char *who)
{
    if (b->a_id == 0)
        return fail(RTT_ERR_INVALID, "%s: B has not been built", who);
    if (b->a_id != a->id)
        return fail(RTT_ERR_INVALID,
                    "%s: B was built for another prepared A (a B is bound to the A passed to "
                    "its rebuild)", who);
    return RTT_OK;
}

int export_c(const rtt::DevC &c, rtt_smatrix *out)
{
    if (c.num_tiles > INT_MAX || c.nnz > INT_MAX || c.tilem > INT_MAX)
        return fail(RTT_ERR_OVERFLOW,
                    "C has %lld tiles / %lld nonzeros: beyond the int32 SMatrix ABI "
                    "(use rtt_last_c_view for the 64-bit device result)",
                    (long long)c.num_tiles, (long long)c.nnz);
    hipStream_t s = rtt::ctx().stream;
    const size_t T = static_cast<size_t>(c.num_tiles);
    const size_t Z = static_cast<size_t>(c.nnz);
    const size_t R = static_cast<size_t>(c.tilem);
    rtt_smatrix r;
    std::memset(&r, 0, sizeof(r));
    r.m = c.rows;
    r.n = c.cols;
    r.nnz = static_cast<int>(Z);
    r.tilem = static_cast<int>(c.tilem);
    r.tilen = static_cast<int>(c.tilen);
    r.numtile = static_cast<int>(T);
    r.tile_ptr = static_cast<int *>(std::malloc((R + 1) * sizeof(int)));
    r.tile_columnidx = static_cast<int *>(std::malloc(std::max<size_t>(T, 1) * sizeof(int)));
    r.tile_nnz = static_cast<int *>(std::malloc((T + 1) * sizeof(int)));
    r.tile_csr_Ptr = static_cast<unsigned char *>(std::malloc(std::max<size_t>(T * 8, 1)));
    r.tile_csr_Col = static_cast<unsigned char *>(std::malloc(std::max<size_t>(Z, 1)));
    r.tile_csr_Value = static_cast<double *>(std::malloc(std::max<size_t>(Z, 1) * sizeof(double)));
    std::vector<int64_t> tp(R + 1), tn(T + 1);
    if (!r.tile_ptr || !r.tile_columnidx || !r.tile_nnz || !r.tile_csr_Ptr || !r.tile_csr_Col ||
        !r.tile_csr_Value) {
        rtt_destroy_smatrix(&r);
        return fail(RTT_ERR_OOM, "host allocation of the C export failed");
    }
    auto copy = [&](void *dst, const void *src, size_t bytes) -> int {
        if (bytes == 0)
            return RTT_OK;
        RTT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        return RTT_OK;
    };
    int rc = RTT_OK;
    if ((rc = copy(tp.data(), c.tile_ptr.ptr, (R + 1) * sizeof(int64_t))) != RTT_OK ||
        (rc = copy(tn.data(), c.tile_nnz.ptr, (T + 1) * sizeof(int64_t))) != RTT_OK ||
        (rc = copy(r.tile_columnidx, c.tile_col.ptr, T * sizeof(int))) != RTT_OK ||
        (rc = copy(r.tile_csr_Ptr, c.tile_csr_ptr.ptr, T * 8)) != RTT_OK ||
        (rc = copy(r.tile_csr_Col, c.tile_csr_col.ptr, Z)) != RTT_OK ||
        (rc = copy(r.tile_csr_Value, c.tile_csr_val.ptr, Z * sizeof(double))) != RTT_OK) {
        rtt_destroy_smatrix(&r);
        return rc;
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
        rtt_destroy_smatrix(&r);
        return fail(RTT_ERR_DEVICE, "C export synchronisation failed");
    }
    for (size_t i = 0; i <= R; ++i)
        r.tile_ptr[i] = static_cast<int>(tp[i]);
    for (size_t i = 0; i <= T; ++i)
        r.tile_nnz[i] = static_cast<int>(tn[i]);
    rtt_destroy_smatrix(out);
    *out = r;
    return RTT_OK;
}

bool is_bijection(const int *p, int n)
{
    std::vector<char> seen(static_cast<size_t>(n), 0);
    for (int i = 0; i < n; ++i) {
        if (p[i] < 0 || p[i] >= n || seen[p[i]])
            return false;
        seen[p[i]] = 1;
    }
    return true;
}

bool is_identity(const int *p, int n)
{
    for (int i = 0; i < n; ++i)
        if (p[i] != i)
            return false;
    return true;
}

} // namespace

extern "C" {

int rtt_init(int device)
{
    rtt::clear_error();
    rtt::Context &c = rtt::ctx();
    if (c.stream && c.device == device)
        return RTT_OK;
    if (c.stream)
        return fail(RTT_ERR_INVALID, "rtt_init: already bound to device %d", c.device);
    int count = 0;
    RTT_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(RTT_ERR_INVALID, "rtt_init: device %d not present (%d devices)", device,
                    count);
    RTT_HIP(hipSetDevice(device));
    if (device < kMaxDevices && !g_scratch_dev[device]) {
        void *p = nullptr;
        RTT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = p;
    }
    RTT_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    RTT_HIP(hipDeviceGetAttribute(&c.num_cus, hipDeviceAttributeMultiprocessorCount, device));
    c.device = device;
    return RTT_OK;
}

int rtt_sync(void)
{
    RTT_TRY(rtt::require_ctx());
    RTT_HIP(hipStreamSynchronize(rtt::ctx().stream));
    return RTT_OK;
}

int rtt_stream(void **stream)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!stream)
        return fail(RTT_ERR_INVALID, "rtt_stream: NULL argument");
    *stream = rtt::ctx().stream;
    return RTT_OK;
}

int rtt_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    RTT_TRY(rtt::require_ctx());
    if (bytes == 0)
        return RTT_OK;
    RTT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, rtt::ctx().stream));
    RTT_HIP(hipStreamSynchronize(rtt::ctx().stream));
    return RTT_OK;
}

int rtt_download_tiles(const rtt_device_tiles *v, int *tile_row_ptr, int *tile_col_idx,
                       int *value_offsets, uint32_t *masks, double *values, int *tile_col_ptr,
                       int *tile_row_idx, int *csc_tile_ids)
{
    rtt::clear_error();
    if (!v)
        return fail(RTT_ERR_INVALID, "rtt_download_tiles: NULL view");
    const size_t T = static_cast<size_t>(v->num_tiles);
    struct Item {
        void *dst;
        const void *src;
        size_t bytes;
    } items[] = {
        {tile_row_ptr, v->tile_row_ptr, (static_cast<size_t>(v->tile_row_count) + 1) * sizeof(int)},
        {tile_col_idx, v->tile_col_idx, T * sizeof(int)},
        {value_offsets, v->value_offsets, (T + 1) * sizeof(int)},
        {masks, v->masks, T * sizeof(uint32_t)},
        {values, v->values, static_cast<size_t>(v->payload_size) * sizeof(double)},
        {tile_col_ptr, v->tile_col_ptr, (static_cast<size_t>(v->tile_col_count) + 1) * sizeof(int)},
        {tile_row_idx, v->tile_row_idx, T * sizeof(int)},
        {csc_tile_ids, v->csc_tile_ids, T * sizeof(int)},
    };
    for (const Item &it : items) {
        if (!it.dst)
            continue;
        if (!it.src && it.bytes)
            return fail(RTT_ERR_INVALID, "rtt_download_tiles: requested array is absent");
        RTT_TRY(rtt_memcpy_d2h(it.dst, it.src, it.bytes));
    }
    return RTT_OK;
}

int rtt_prepare_external_a(const rtt_smatrix *host, const int *row_new_to_old,
                           const int *inner_new_to_old, int dense_threshold,
                           rtt_prepared_a **out, rtt_offline_a_stats *stats)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!host || !out || dense_threshold < 1 || dense_threshold > rtt::kInputElems)
        return fail(RTT_ERR_INVALID, "rtt_prepare_*_a: bad argument");
    if (host->m < 0 || host->n < 0 || host->nnz < 0 || !host->rowpointer ||
        host->rowpointer[0] != 0 || host->rowpointer[host->m] != host->nnz)
        return fail(RTT_ERR_INVALID, "rtt_prepare_*_a: malformed CSR");
    const bool external = row_new_to_old || inner_new_to_old;
    if (external && (!row_new_to_old || !inner_new_to_old))
        return fail(RTT_ERR_INVALID, "external orders need both axes");
    if (external && (!is_bijection(row_new_to_old, host->m) ||
                     !is_bijection(inner_new_to_old, host->n)))
        return fail(RTT_ERR_INVALID, "external order is not a bijection");
    rtt_offline_a_stats st;
    std::memset(&st, 0, sizeof(st));
    auto t_total = rtt::Clock::now();
    std::unique_ptr<rtt_prepared_a> a(new (std::nothrow) rtt_prepared_a());
    if (!a)
        return fail(RTT_ERR_OOM, "host allocation failed");
    a->rows = host->m;
    a->cols = host->n;
    a->dense_threshold = dense_threshold;
    std::vector<int> row_old_to_new, inner_old_to_new;
    if (external) {
        a->row_new_to_old.assign(row_new_to_old, row_new_to_old + host->m);
        a->inner_new_to_old.assign(inner_new_to_old, inner_new_to_old + host->n);
        row_old_to_new.resize(host->m);
        inner_old_to_new.resize(host->n);
        for (int i = 0; i < host->m; ++i)
            row_old_to_new[row_new_to_old[i]] = i;
        for (int i = 0; i < host->n; ++i)
            inner_old_to_new[inner_new_to_old[i]] = i;
        a->row_identity = is_identity(row_new_to_old, host->m);
        a->identity = a->row_identity && is_identity(inner_new_to_old, host->n);
    }
    // Active prefixes: the last non-empty (new) row and the largest used
    // (new) inner index, as gpu_prepare_identity_a / the external plan do
    // (src/gpu_dmma_tiles.h:2458-2469, src/dmma_reorder.h:2257-2284).
    for (int r = 0; r < host->m; ++r) {
        const int b = host->rowpointer[r], e = host->rowpointer[r + 1];
        if (b < 0 || e < b || e > host->nnz)
            return fail(RTT_ERR_INVALID, "row %d has invalid bounds", r);
        if (b != e)
            a->active_rows = std::max(a->active_rows, (external ? row_old_to_new[r] : r) + 1);
        for (int x = b; x < e; ++x) {
            const int c = host->columnindex[x];
            if (c < 0 || c >= host->n)
                return fail(RTT_ERR_INVALID, "column %d out of range", c);
            a->active_inner = std::max(a->active_inner, (external ? inner_old_to_new[c] : c) + 1);
        }
    }
    double sort_ms = 0.0;
    RTT_TRY(rtt::upload_csr(*host, &a->csr, &a->build, &st.h2d_ms, &st.validation_ms, &sort_ms,
                            &st.rows_were_sorted));
    st.canonicalize_ms = sort_ms;
    const rtt::DevCsr *tiled = &a->csr;
    if (external && !a->identity) {
        auto t_re = rtt::Clock::now();
        // Materialise the reordered CSR once (offline): rows gathered by
        // row_new_to_old, columns mapped by inner_old_to_new.
        rtt_smatrix p;
        std::memset(&p, 0, sizeof(p));
        p.m = host->m;
        p.n = host->n;
        p.nnz = host->nnz;
        std::vector<int> rp(host->m + 1), ci(host->nnz);
        std::vector<double> vv(host->nnz);
        int e = 0;
        for (int nr = 0; nr < host->m; ++nr) {
            rp[nr] = e;
            const int orow = row_new_to_old[nr];
            for (int x = host->rowpointer[orow]; x < host->rowpointer[orow + 1]; ++x) {
                ci[e] = inner_old_to_new[host->columnindex[x]];
                vv[e] = host->value[x];
                ++e;
            }
        }
        rp[host->m] = e;
        p.rowpointer = rp.data();
        p.columnindex = ci.data();
        p.value = vv.data();
        double h2d = 0, val = 0, srt = 0;
        int sorted = 0;
        RTT_TRY(rtt::upload_csr(p, &a->csr_tiled, &a->build, &h2d, &val, &srt, &sorted));
        st.reorder_ms = rtt::ms_since(t_re);
        tiled = &a->csr_tiled;
        RTT_TRY(a->d_inner_new_to_old.ensure(static_cast<size_t>(host->n) * sizeof(int)));
        if (host->n > 0)
            RTT_HIP(hipMemcpy(a->d_inner_new_to_old.ptr, inner_new_to_old,
                              static_cast<size_t>(host->n) * sizeof(int),
                              hipMemcpyHostToDevice));
    }
    auto t_tiles = rtt::Clock::now();
    RTT_TRY(rtt::build_tiles(*tiled, nullptr, a->active_rows, a->active_inner, host->m, host->n,
                             false, dense_threshold, false, &a->tiles, &a->build));
    RTT_TRY(rtt::fill_tile_row_of(&a->tiles));
    RTT_TRY(rtt::build_colmajor_values(&a->tiles));
    RTT_HIP(hipStreamSynchronize(rtt::ctx().stream));
    auto t_tr = rtt::Clock::now();
    RTT_TRY(rtt::build_transpose_source(&a->tiles, &a->build));
    RTT_HIP(hipStreamSynchronize(rtt::ctx().stream));
    st.transpose_src_ms = rtt::ms_since(t_tr);
    st.tile_build_ms = rtt::ms_since(t_tiles);
    st.total_ms = rtt::ms_since(t_total);
    if (stats)
        *stats = st;
    a->id = register_a();
    *out = a.release();
    return RTT_OK;
}

int rtt_prepare_identity_a(const rtt_smatrix *host, int dense_threshold, rtt_prepared_a **out,
                           rtt_offline_a_stats *stats)
{
    return rtt_prepare_external_a(host, nullptr, nullptr, dense_threshold, out, stats);
}

int rtt_prepared_a_view(const rtt_prepared_a *a, rtt_device_tiles *view)
{
    if (!a || !view)
        return fail(RTT_ERR_INVALID, "rtt_prepared_a_view: NULL argument");
    a->tiles.view(view);
    return RTT_OK;
}

int rtt_prepared_a_info(const rtt_prepared_a *a, int *active_rows, int *active_inner,
                        int *identity)
{
    if (!a)
        return fail(RTT_ERR_INVALID, "rtt_prepared_a_info: NULL handle");
    if (active_rows)
        *active_rows = a->active_rows;
    if (active_inner)
        *active_inner = a->active_inner;
    if (identity)
        *identity = a->identity ? 1 : 0;
    return RTT_OK;
}

const int *rtt_prepared_a_row_new_to_old(const rtt_prepared_a *a)
{
    if (!a || a->identity || a->row_new_to_old.empty())
        return nullptr;
    return a->row_new_to_old.data();
}

void rtt_destroy_prepared_a(rtt_prepared_a *a)
{
    if (a && rtt::ctx().stream)
        (void)hipStreamSynchronize(rtt::ctx().stream);
    if (a)
        unregister_a(a->id);
    delete a;
}

int rtt_rebuild_b(const rtt_prepared_a *a_const, int aat, rtt_dynamic_b **inout,
                  rtt_b_update_stats *stats)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a_const || !inout || (aat != 0 && aat != 1))
        return fail(RTT_ERR_INVALID, "rtt_rebuild_b: bad argument");
    rtt_prepared_a *a = const_cast<rtt_prepared_a *>(a_const);
    if (!aat && a->rows != a->cols)
        return fail(RTT_ERR_INVALID, "C = A*A needs a square A (%d x %d)", a->rows, a->cols);
    if (!*inout) {
        *inout = new (std::nothrow) rtt_dynamic_b();
        if (!*inout)
            return fail(RTT_ERR_OOM, "host allocation failed");
    }
    rtt_dynamic_b *b = *inout;
    b->aat = aat;
    b->a_id = 0;   // unbound until the rebuild completes
    rtt_b_update_stats st;
    std::memset(&st, 0, sizeof(st));
    auto t0 = rtt::Clock::now();
    const int *inner_map = a->identity ? nullptr : a->d_inner_new_to_old.as<int>();
    if (!aat) {
        // B = A with rows gathered through the inner permutation, columns kept.
        st.path = 0;
        RTT_TRY(rtt::build_tiles(a->csr, inner_map, a->active_inner, a->cols, a->cols, a->cols,
                                 true, a->dense_threshold, false, &b->tiles, &b->build));
    } else if (a->row_identity) {
        // B = A^T straight from A's tiles.
        st.path = 1;
        RTT_TRY(rtt::transpose_tiles(a->tiles, a->rows, &b->tiles, &b->build));
    } else {
        st.path = 2;
        auto tp = rtt::Clock::now();
        RTT_TRY(rtt::transpose_csr(a->csr, &a->csr_t, &b->build));
        RTT_HIP(hipStreamSynchronize(rtt::ctx().stream));
        st.csr_row_permute_ms = rtt::ms_since(tp);
        RTT_TRY(rtt::build_tiles(a->csr_t, inner_map, a->active_inner, a->rows, a->cols, a->rows,
                                 true, a->dense_threshold, false, &b->tiles, &b->build));
    }
    if (!b->tiles.transpose_of)   // the transpose writes its records itself
        RTT_TRY(rtt::order_by_class(&b->tiles, &b->build));
    RTT_TRY(rtt::build_pair_extents(a->tiles, &b->tiles));
    RTT_HIP(hipStreamSynchronize(rtt::ctx().stream));
    b->a_id = a->id;
    ++b->gen;
    st.total_ms = rtt::ms_since(t0);
    st.csr2tile_ms = st.total_ms - st.csr_row_permute_ms;
    if (stats)
        *stats = st;
    return RTT_OK;
}

int rtt_upload_csr(const rtt_smatrix *host, rtt_device_csr **out, double *h2d_ms,
                   double *validation_ms)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!host || !out)
        return fail(RTT_ERR_INVALID, "rtt_upload_csr: NULL argument");
    std::unique_ptr<rtt_device_csr> d(new (std::nothrow) rtt_device_csr());
    if (!d)
        return fail(RTT_ERR_OOM, "host allocation failed");
    double h2d = 0, val = 0, sort = 0;
    int sorted = 1;
    RTT_TRY(rtt::upload_csr(*host, &d->csr, &d->build, &h2d, &val, &sort, &sorted));
    if (h2d_ms)
        *h2d_ms = h2d;
    if (validation_ms)
        *validation_ms = val + sort;
    *out = d.release();
    return RTT_OK;
}

int rtt_device_csr_set_values(rtt_device_csr *csr, const double *host_values)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!csr)
        return fail(RTT_ERR_INVALID, "rtt_device_csr_set_values: NULL handle");
    return rtt::set_csr_values(&csr->csr, host_values, &csr->build);
}

void rtt_destroy_device_csr(rtt_device_csr *csr)
{
    if (csr && rtt::ctx().stream)
        (void)hipStreamSynchronize(rtt::ctx().stream);
    delete csr;
}

int rtt_rebuild_b_general(const rtt_prepared_a *a_const, const rtt_device_csr *src,
                          rtt_dynamic_b **inout, rtt_b_update_stats *stats)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a_const || !src || !inout)
        return fail(RTT_ERR_INVALID, "rtt_rebuild_b_general: NULL argument");
    if (src->csr.rows != a_const->cols)
        return fail(RTT_ERR_INVALID, "A is %d x %d but B has %d rows", a_const->rows,
                    a_const->cols, src->csr.rows);
    if (!*inout) {
        *inout = new (std::nothrow) rtt_dynamic_b();
        if (!*inout)
            return fail(RTT_ERR_OOM, "host allocation failed");
    }
    rtt_dynamic_b *b = *inout;
    b->aat = 0;
    b->a_id = 0;
    rtt_b_update_stats st;
    std::memset(&st, 0, sizeof(st));
    st.path = 3;
    auto t0 = rtt::Clock::now();
    const int *inner_map = a_const->identity ? nullptr : a_const->d_inner_new_to_old.as<int>();
    // B's rows through A's inner permutation, all of B's columns, and the
    // source-to-payload map for later values-only updates.
    RTT_TRY(rtt::build_tiles(src->csr, inner_map, a_const->active_inner, src->csr.cols,
                             a_const->cols, src->csr.cols, true, a_const->dense_threshold, false,
                             &b->tiles, &b->build, true));
    RTT_TRY(rtt::order_by_class(&b->tiles, &b->build));
    RTT_TRY(rtt::build_pair_extents(a_const->tiles, &b->tiles));
    RTT_HIP(hipStreamSynchronize(rtt::ctx().stream));
    b->a_id = a_const->id;
    ++b->gen;
    st.total_ms = rtt::ms_since(t0);
    st.csr2tile_ms = st.total_ms;
    if (stats)
        *stats = st;
    return RTT_OK;
}

int rtt_compute_nnz_cub_ab(const rtt_prepared_a *a_const, const rtt_device_csr *b,
                           unsigned long long *nnz_cub, double *elapsed_ms)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a_const || !b || !nnz_cub)
        return fail(RTT_ERR_INVALID, "rtt_compute_nnz_cub_ab: bad argument");
    rtt_prepared_a *a = const_cast<rtt_prepared_a *>(a_const);
    auto t0 = rtt::Clock::now();
    RTT_TRY(rtt::nnz_cub_ab(a->csr, b->csr, nnz_cub, &a->build));
    if (elapsed_ms)
        *elapsed_ms = rtt::ms_since(t0);
    return RTT_OK;
}

int rtt_update_b_values(const rtt_prepared_a *a_const, const rtt_device_csr *src, int policy,
                        rtt_dynamic_b *b, rtt_b_update_stats *stats,
                        rtt_b_values_decision *decision)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a_const || !src || !b)
        return fail(RTT_ERR_INVALID, "rtt_update_b_values: NULL argument");
    if (b->a_id != 0)
        RTT_TRY(check_binding(a_const, b, "rtt_update_b_values"));
    const rtt::DevTiles &t = b->tiles;
    const bool valid = t.map_source_id != 0 && t.map_source_id == src->csr.structure_id &&
                       t.map_nnz == src->csr.nnz && !t.transpose_of &&
                       t.tile_row_count == a_const->tiles.tile_col_count;
    // Every active source entry has a slot by construction, duplicates share
    // one (so the map is injective exactly when there are none), the build
    // writes every payload slot (dense padding as zeros the update never
    // touches), and the numeric kernels only read B.
    rtt_b_values_decision d;
    RTT_TRY(rtt_choose_b_values_clear(policy, 1, valid ? 1 : 0, t.map_dups ? 1 : 0, 1,
                                      t.map_dups ? 0 : 1, 1, 1, 1, &d));
    if (decision)
        *decision = d;
    if (d.reason == RTT_B_VALUES_FALLBACK_INVALID_POLICY)
        return fail(RTT_ERR_INVALID, "unknown B values clear policy %d", policy);
    if (!valid)
        return rtt_rebuild_b_general(a_const, src, &b, stats);
    rtt_b_update_stats st;
    std::memset(&st, 0, sizeof(st));
    st.path = 4;
    auto t0 = rtt::Clock::now();
    RTT_TRY(rtt::update_b_values(src->csr, d.clear_payload != 0, &b->tiles));
    RTT_HIP(hipStreamSynchronize(rtt::ctx().stream));
    st.total_ms = rtt::ms_since(t0);
    if (stats)
        *stats = st;
    return RTT_OK;
}

int rtt_dynamic_b_view(const rtt_dynamic_b *b_const, rtt_device_tiles *view)
{
    rtt::clear_error();
    if (!b_const || !view)
        return fail(RTT_ERR_INVALID, "rtt_dynamic_b_view: NULL argument");
    RTT_TRY(rtt::require_ctx());
    // The timed rebuild produces what the SpGEMM core reads (tile row
    // pointer + packed per-tile records); the remaining DmmaDeviceTiles
    // arrays and the tile CSC are completed here, on first request.
    rtt_dynamic_b *b = const_cast<rtt_dynamic_b *>(b_const);
    if (b->tiles.transpose_of && !a_alive(b->a_id))
        return fail(RTT_ERR_INVALID,
                    "rtt_dynamic_b_view: the A this B = A^T was built from has been destroyed");
    RTT_TRY(rtt::materialize_tiles(&b->tiles, &b->build));
    b->tiles.view(view);
    return RTT_OK;
}

void rtt_destroy_dynamic_b(rtt_dynamic_b *b)
{
    if (b && rtt::ctx().stream)
        (void)hipStreamSynchronize(rtt::ctx().stream);
    delete b;
}

int rtt_compute_nnz_cub(const rtt_prepared_a *a_const, int aat, unsigned long long *nnz_cub,
                        double *elapsed_ms)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a_const || !nnz_cub)
        return fail(RTT_ERR_INVALID, "rtt_compute_nnz_cub: bad argument");
    rtt_prepared_a *a = const_cast<rtt_prepared_a *>(a_const);
    auto t0 = rtt::Clock::now();
    RTT_TRY(rtt::nnz_cub(a->csr, aat != 0, nnz_cub, &a->build));
    if (elapsed_ms)
        *elapsed_ms = rtt::ms_since(t0);
    return RTT_OK;
}

void rtt_schedule_defaults(rtt_schedule *s)
{
    if (!s)
        return;
    s->lb_mode = RTT_LB_ADAPTIVE_TAIL;
    s->materialize_output = 1;
    s->split_threshold = 16384.0;
    s->split_chunk_target = 8192.0;
    s->split_max_chunks = 8;
    s->max_heavy_fraction = 0.01;
    s->tail_record_capacity = 1 << 20;
    s->force_tail_split = 0;
    s->numeric_path = RTT_NUMERIC_AUTO;
    s->tile_fill_threshold = 12.0;
    s->precision = RTT_PRECISION_FP64;
}

int rtt_tilespgemm(const rtt_prepared_a *a_const, const rtt_dynamic_b *b, rtt_smatrix *output,
                   rtt_stats *stats, const rtt_schedule *sched)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a_const || !b || !stats)
        return fail(RTT_ERR_INVALID, "rtt_tilespgemm: NULL argument");
    RTT_TRY(check_binding(a_const, b, "rtt_tilespgemm"));
    rtt_prepared_a *a = const_cast<rtt_prepared_a *>(a_const);
    rtt_schedule defaults;
    rtt_schedule_defaults(&defaults);
    const rtt_schedule &sc = sched ? *sched : defaults;
    if (sc.lb_mode < RTT_LB_STATIC || sc.lb_mode > RTT_LB_ADAPTIVE_TAIL)
        return fail(RTT_ERR_INVALID, "unknown lb_mode %d", sc.lb_mode);
    // The tail planner's own validity conditions (dmma_plan_tile_tail_queue,
    // src/dmma_tile_tail_queue.h:190-292; max_chunks 2..32).
    if (!(sc.split_threshold > 0.0) || !(sc.split_chunk_target > 0.0) ||
        sc.split_max_chunks < 2 || sc.split_max_chunks > 32 || !(sc.max_heavy_fraction >= 0.0) ||
        !(sc.max_heavy_fraction <= 1.0) || sc.tail_record_capacity <= 0)
        return fail(RTT_ERR_INVALID,
                    "bad tail schedule: threshold %g, chunk target %g, max chunks %d, heavy "
                    "fraction %g, record capacity %d",
                    sc.split_threshold, sc.split_chunk_target, sc.split_max_chunks,
                    sc.max_heavy_fraction, sc.tail_record_capacity);
    if (sc.numeric_path < RTT_NUMERIC_AUTO || sc.numeric_path > RTT_NUMERIC_TILE ||
        !(sc.tile_fill_threshold >= 0.0))
        return fail(RTT_ERR_INVALID, "unknown numeric path %d (fill threshold %g)",
                    sc.numeric_path, sc.tile_fill_threshold);
    if (sc.precision != RTT_PRECISION_FP64 && sc.precision != RTT_PRECISION_SPLIT_BF16)
        return fail(RTT_ERR_INVALID, "unknown precision %d", sc.precision);
    if (sc.precision == RTT_PRECISION_SPLIT_BF16 && sc.numeric_path == RTT_NUMERIC_VECTOR)
        return fail(RTT_ERR_INVALID, "split precision runs on the TILE path only");
    // The same shape contract as dmma_tilespgemm (src/dmma_spgemm.h:6495-6510).
    if (a->tiles.tile_col_count != b->tiles.tile_row_count || a->tiles.tile_rows != rtt::kTileM ||
        a->tiles.tile_cols != rtt::kTileK || b->tiles.tile_rows != rtt::kTileK ||
        b->tiles.tile_cols != rtt::kTileN)
        return fail(RTT_ERR_INVALID, "A and B tile shapes do not conform");
    if (sc.materialize_output && !output)
        return fail(RTT_ERR_INVALID, "materialize_output needs an output matrix");
    std::memset(stats, 0, sizeof(*stats));
    auto t0 = rtt::Clock::now();
    rtt::CoreTimes times;
    a->has_c = false;
    // Numeric routing per tile row: TILE takes every row to the tile-pair
    // kernel, VECTOR none, AUTO the rows whose A tiles hold at least
    // tile_fill_threshold stored values on average.  Split precision runs
    // every row on the tile-pair kernel.
    int path = sc.numeric_path;
    if (sc.precision == RTT_PRECISION_SPLIT_BF16)
        path = RTT_NUMERIC_TILE;
    float fill = path == RTT_NUMERIC_TILE ? 0.f : static_cast<float>(sc.tile_fill_threshold);
    int64_t tile_rows = 0, nonempty = 0;
    if (path == RTT_NUMERIC_AUTO) {
        RTT_TRY(rtt::count_tile_rows(a->tiles, fill, &a->core.tile_rows, &tile_rows, &nonempty));
        if (tile_rows == 0)
            path = RTT_NUMERIC_VECTOR;
        else if (tile_rows == nonempty)
            path = RTT_NUMERIC_TILE;
    }
    rtt::TileOpts topt;
    topt.fill = path == RTT_NUMERIC_TILE ? 0.f : fill;
    topt.precision = sc.precision;
    // The adaptive-tail schedule splits heavy C tiles of the tile-pair path
    // (src/main.cu:132-166: the tail queue belongs to --lb-mode adaptive-tail).
    topt.tail = sc.lb_mode == RTT_LB_ADAPTIVE_TAIL;
    topt.threshold = static_cast<float>(sc.split_threshold);
    topt.target = static_cast<float>(sc.split_chunk_target);
    topt.max_chunks = sc.split_max_chunks;
    topt.max_heavy_fraction = sc.max_heavy_fraction;
    topt.force = sc.force_tail_split;
    topt.record_capacity = sc.tail_record_capacity;
    if (path != RTT_NUMERIC_TILE)
        RTT_TRY(rtt::ensure_record_values(&const_cast<rtt_dynamic_b *>(b)->tiles));
    if (path != RTT_NUMERIC_VECTOR) {
        rtt_dynamic_b *bm = const_cast<rtt_dynamic_b *>(b);
        auto tc = rtt::Clock::now();
        RTT_TRY(rtt::build_b_columns(&bm->tiles, &bm->build));
        RTT_HIP(hipStreamSynchronize(rtt::ctx().stream));
        stats->b_columns_ms = rtt::ms_since(tc);
        topt.bc = rtt::b_columns(b->tiles);
    }
    RTT_TRY(rtt::spgemm_core(a->tiles, b->tiles, &a->last_c, &a->core, &a->events, &times,
                            sc.lb_mode, path == RTT_NUMERIC_VECTOR ? nullptr : &topt));
    stats->numeric_path = path;
    stats->precision = path == RTT_NUMERIC_VECTOR ? RTT_PRECISION_FP64 : sc.precision;
    a->last_b = b;
    a->last_b_gen = b->gen;
    a->last_all_tile = path == RTT_NUMERIC_TILE;
    stats->mfma_pairs = times.mfma_pairs;
    stats->tail_reason = times.tail_reason;
    stats->tail_records = times.tail_records;
    stats->heavy_chunks = times.tail_chunks;
    a->has_c = true;
    stats->symbolic_ms = times.symbolic_ms;
    stats->allocation_ms = times.allocation_ms;
    stats->numeric_ms = times.numeric_ms;
    stats->total_ms = rtt::ms_since(t0);
    stats->symbolic_count_kernel_ms = times.sym_count_kernel_ms;
    stats->symbolic_fill_kernel_ms = times.sym_fill_kernel_ms;
    stats->numeric_kernel_ms = times.numeric_kernel_ms;
    stats->pair_count = times.pairs;
    stats->output_tiles = static_cast<unsigned long long>(a->last_c.num_tiles);
    stats->output_nnz = static_cast<unsigned long long>(a->last_c.nnz);
    stats->heavy_rows = times.heavy_rows;
    if (sc.materialize_output) {
        auto te = rtt::Clock::now();
        RTT_TRY(export_c(a->last_c, output));
        stats->output_copy_ms = rtt::ms_since(te);
        stats->output_materialized = 1;
    }
    return RTT_OK;
}

int rtt_batch_extract(const rtt_prepared_a *a, const int64_t *row_offset, const int64_t *col_offset,
                      int m, int n, rtt_smatrix *out)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a || !row_offset || !col_offset || !out || m < 0 || n < 0)
        return fail(RTT_ERR_INVALID, "rtt_batch_extract: bad argument");
    if (!a->has_c)
        return fail(RTT_ERR_INVALID, "rtt_batch_extract: no C has been produced");
    const int64_t r0 = *row_offset, c0 = *col_offset;
    if (r0 % rtt::kTileM || c0 % rtt::kTileN || r0 < 0 || c0 < 0 || r0 + m > a->rows ||
        c0 + n > a->last_c.cols)
        return fail(RTT_ERR_INVALID,
                    "rtt_batch_extract: block at (%lld, %lld) of %d x %d is not 8-aligned "
                    "inside C",
                    (long long)r0, (long long)c0, m, n);
    const rtt::DevC &c = a->last_c;
    const int64_t tr0 = r0 / rtt::kTileM;
    const int64_t tr_want = (m + rtt::kTileM - 1) / rtt::kTileM;
    const int64_t tr1 = std::min<int64_t>(c.tilem, tr0 + tr_want);
    const int64_t nrows = std::max<int64_t>(0, tr1 - tr0);
    std::vector<int64_t> tp(static_cast<size_t>(nrows) + 1, 0);
    if (nrows > 0)
        RTT_TRY(rtt_memcpy_d2h(tp.data(), c.tile_ptr.as<int64_t>() + tr0,
                               (static_cast<size_t>(nrows) + 1) * sizeof(int64_t)));
    const int64_t t0 = nrows > 0 ? tp[0] : 0, t1 = nrows > 0 ? tp[nrows] : 0;
    const int64_t T = t1 - t0;
    std::vector<int64_t> tn(static_cast<size_t>(T) + 1, 0);
    if (T > 0)
        RTT_TRY(rtt_memcpy_d2h(tn.data(), c.tile_nnz.as<int64_t>() + t0,
                               (static_cast<size_t>(T) + 1) * sizeof(int64_t)));
    const int64_t z0 = T > 0 ? tn[0] : 0, Z = T > 0 ? tn[T] - z0 : 0;
    if (T > INT_MAX || Z > INT_MAX)
        return fail(RTT_ERR_OVERFLOW, "problem block has %lld tiles / %lld nonzeros",
                    (long long)T, (long long)Z);
    rtt_smatrix r;
    std::memset(&r, 0, sizeof(r));
    r.m = m;
    r.n = n;
    r.tilem = static_cast<int>(tr_want);
    r.tilen = (n + rtt::kTileN - 1) / rtt::kTileN;
    r.numtile = static_cast<int>(T);
    r.nnz = static_cast<int>(Z);
    r.tile_ptr = static_cast<int *>(std::malloc((static_cast<size_t>(tr_want) + 1) * sizeof(int)));
    r.tile_columnidx = static_cast<int *>(std::malloc(std::max<size_t>(T, 1) * sizeof(int)));
    r.tile_nnz = static_cast<int *>(std::malloc((static_cast<size_t>(T) + 1) * sizeof(int)));
    r.tile_csr_Ptr = static_cast<unsigned char *>(std::malloc(std::max<size_t>(8 * T, 1)));
    r.tile_csr_Col = static_cast<unsigned char *>(std::malloc(std::max<size_t>(Z, 1)));
    r.tile_csr_Value = static_cast<double *>(std::malloc(std::max<size_t>(Z, 1) * sizeof(double)));
    if (!r.tile_ptr || !r.tile_columnidx || !r.tile_nnz || !r.tile_csr_Ptr || !r.tile_csr_Col ||
        !r.tile_csr_Value) {
        rtt_destroy_smatrix(&r);
        return fail(RTT_ERR_OOM, "host allocation of the extracted block failed");
    }
    int rc = RTT_OK;
    if (T > 0 && ((rc = rtt_memcpy_d2h(r.tile_columnidx, c.tile_col.as<int>() + t0,
                                       static_cast<size_t>(T) * sizeof(int))) != RTT_OK ||
                  (rc = rtt_memcpy_d2h(r.tile_csr_Ptr, c.tile_csr_ptr.as<uint8_t>() + 8 * t0,
                                       static_cast<size_t>(8 * T))) != RTT_OK)) {
        rtt_destroy_smatrix(&r);
        return rc;
    }
    if (Z > 0 && ((rc = rtt_memcpy_d2h(r.tile_csr_Col, c.tile_csr_col.as<uint8_t>() + z0,
                                       static_cast<size_t>(Z))) != RTT_OK ||
                  (rc = rtt_memcpy_d2h(r.tile_csr_Value, c.tile_csr_val.as<double>() + z0,
                                       static_cast<size_t>(Z) * sizeof(double))) != RTT_OK)) {
        rtt_destroy_smatrix(&r);
        return rc;
    }
    const int shift = static_cast<int>(c0 / rtt::kTileN);
    for (int64_t i = 0; i <= tr_want; ++i)
        r.tile_ptr[i] = static_cast<int>((i <= nrows ? tp[i] : t1) - t0);
    for (int64_t t = 0; t < T; ++t)
        r.tile_columnidx[t] -= shift;
    for (int64_t t = 0; t <= T; ++t)
        r.tile_nnz[t] = static_cast<int>(tn[t] - z0);
    rtt_destroy_smatrix(out);
    *out = r;
    return RTT_OK;
}

int rtt_last_c_row_stats(const rtt_prepared_a *a, int64_t *row_nnz, double *row_sum)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a || !row_nnz || !row_sum)
        return fail(RTT_ERR_INVALID, "rtt_last_c_row_stats: NULL argument");
    if (!a->has_c)
        return fail(RTT_ERR_INVALID, "rtt_last_c_row_stats: no C has been produced");
    const rtt::DevC &c = a->last_c;
    const int64_t rows = c.tilem * rtt::kTileM;
    std::fill(row_nnz, row_nnz + a->rows, int64_t(0));
    std::fill(row_sum, row_sum + a->rows, 0.0);
    if (rows == 0)
        return RTT_OK;
    rtt::DevBuf dn, ds;
    RTT_TRY(dn.ensure(static_cast<size_t>(rows) * sizeof(int64_t)));
    RTT_TRY(ds.ensure(static_cast<size_t>(rows) * sizeof(double)));
    hipStream_t s = rtt::ctx().stream;
    hipLaunchKernelGGL(c_row_stats_kernel, dim3(rtt::blocks_for(rows, 256)), dim3(256), 0, s, rows,
                       c.tile_ptr.as<int64_t>(), c.tile_mask.as<uint64_t>(),
                       c.tile_nnz.as<int64_t>(), c.tile_csr_val.as<double>(), dn.as<int64_t>(),
                       ds.as<double>());
    RTT_HIP(hipGetLastError());
    const int64_t keep = std::min<int64_t>(rows, a->rows);
    RTT_TRY(rtt_memcpy_d2h(row_nnz, dn.ptr, static_cast<size_t>(keep) * sizeof(int64_t)));
    RTT_TRY(rtt_memcpy_d2h(row_sum, ds.ptr, static_cast<size_t>(keep) * sizeof(double)));
    return RTT_OK;
}

int rtt_autotune(const rtt_prepared_a *a, const rtt_dynamic_b *b, int repeats, rtt_schedule *sched,
                 rtt_tune_report *report)
{
    rtt::clear_error();
    if (!a || !b || !sched || repeats < 1 || repeats > 64)
        return fail(RTT_ERR_INVALID, "rtt_autotune: bad argument");
    // Candidates: the numeric routing threshold (stored values per A tile
    // above which a tile row takes the FP64 matrix-core tile-pair kernel
    // instead of the LDS sparse accumulator; 0 = every row, inf = none) and,
    // for the slab rows of the sparse kernels, the queue assignment.
    struct Cand {
        int path;
        double fill;
        int lb;
    };
    const Cand cands[] = {
        {RTT_NUMERIC_VECTOR, 0.0, RTT_LB_ADAPTIVE_TAIL}, {RTT_NUMERIC_VECTOR, 0.0, RTT_LB_DYNAMIC},
        {RTT_NUMERIC_AUTO, 24.0, RTT_LB_ADAPTIVE_TAIL},  {RTT_NUMERIC_AUTO, 16.0, RTT_LB_ADAPTIVE_TAIL},
        {RTT_NUMERIC_AUTO, 12.0, RTT_LB_ADAPTIVE_TAIL},  {RTT_NUMERIC_AUTO, 8.0, RTT_LB_ADAPTIVE_TAIL},
        {RTT_NUMERIC_TILE, 0.0, RTT_LB_ADAPTIVE_TAIL},
    };
    constexpr int kCands = static_cast<int>(sizeof(cands) / sizeof(cands[0]));
    static_assert(kCands <= RTT_TUNE_MAX_CANDIDATES, "report size");
    rtt_schedule base = *sched;
    base.materialize_output = 0;
    base.precision = RTT_PRECISION_FP64;
    rtt_tune_report rep;
    std::memset(&rep, 0, sizeof(rep));
    rep.candidates = kCands;
    int best = 0;
    for (int i = 0; i < kCands; ++i) {
        rtt_schedule sc = base;
        sc.numeric_path = cands[i].path;
        sc.tile_fill_threshold = cands[i].fill;
        sc.lb_mode = cands[i].lb;
        std::vector<double> t;
        rtt_stats st;
        for (int r = 0; r <= repeats; ++r) {   // the first run warms up
            auto t0 = rtt::Clock::now();
            RTT_TRY(rtt_tilespgemm(a, b, nullptr, &st, &sc));
            if (r > 0)
                t.push_back(rtt::ms_since(t0));
        }
        std::sort(t.begin(), t.end());
        rep.ms[i] = t[t.size() / 2];
        rep.numeric_path[i] = cands[i].path;
        rep.fill[i] = cands[i].fill;
        rep.lb_mode[i] = cands[i].lb;
        if (rep.ms[i] < rep.ms[best])
            best = i;
    }
    rep.chosen = best;
    sched->numeric_path = cands[best].path;
    sched->tile_fill_threshold = cands[best].path == RTT_NUMERIC_AUTO ? cands[best].fill
                                                                      : sched->tile_fill_threshold;
    sched->lb_mode = cands[best].lb;
    if (report)
        *report = rep;
    return RTT_OK;
}

int rtt_tile_values_noalloc(const rtt_prepared_a *a, const rtt_dynamic_b *b, int precision,
                            double *c_values)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a || !b || !c_values)
        return fail(RTT_ERR_INVALID, "rtt_tile_values_noalloc: NULL argument");
    RTT_TRY(check_binding(a, b, "rtt_tile_values_noalloc"));
    if (!a->has_c || !a->last_all_tile || a->last_b != b || a->last_b_gen != b->gen)
        return fail(RTT_ERR_INVALID,
                    "rtt_tile_values_noalloc: needs C's structure from an rtt_tilespgemm on this "
                    "(A, B) whose rows all took the TILE path, with B not rebuilt since");
    if (!b->tiles.transpose_of && !b->tiles.cols_ready)
        return fail(RTT_ERR_INVALID, "rtt_tile_values_noalloc: B's tile columns are not built");
    const int dev = rtt::ctx().device;
    if (dev < 0 || dev >= kMaxDevices || !g_scratch_dev[dev])
        return fail(RTT_ERR_INVALID, "rtt_tile_values_noalloc: no scratch address for device %d",
                    dev);
    return rtt::tile_values_noalloc(a->tiles, rtt::b_columns(b->tiles), a->last_c, precision,
                                    g_scratch_dev[dev], kNoAllocScratchBytes, c_values,
                                    rtt::ctx().stream);
}

int rtt_last_tail_records(const rtt_prepared_a *a, int max_records, long long *task, int *scans,
                          int *matches, int *chunks, int *count)
{
    rtt::clear_error();
    RTT_TRY(rtt::require_ctx());
    if (!a || !count || max_records < 0 || (max_records > 0 && (!task || !scans || !matches ||
                                                                 !chunks)))
        return fail(RTT_ERR_INVALID, "rtt_last_tail_records: bad argument");
    *count = 0;
    rtt_prepared_a *am = const_cast<rtt_prepared_a *>(a);
    if (!am->core.tail_state.ptr)
        return RTT_OK;
    return rtt::tail_records(&am->core, max_records, task, scans, matches, chunks, count);
}

int rtt_last_c_view(const rtt_prepared_a *a, rtt_device_c *v)
{
    if (!a || !v)
        return fail(RTT_ERR_INVALID, "rtt_last_c_view: NULL argument");
    if (!a->has_c)
        return fail(RTT_ERR_INVALID, "rtt_last_c_view: no C has been produced");
    const rtt::DevC &c = a->last_c;
    v->tilem = c.tilem;
    v->tilen = c.tilen;
    v->num_tiles = c.num_tiles;
    v->nnz = c.nnz;
    v->tile_ptr = c.tile_ptr.as<int64_t>();
    v->tile_col = c.tile_col.as<int>();
    v->tile_mask = c.tile_mask.as<uint64_t>();
    v->tile_nnz = c.tile_nnz.as<int64_t>();
    v->tile_csr_ptr = c.tile_csr_ptr.as<uint8_t>();
    v->tile_csr_col = c.tile_csr_col.as<uint8_t>();
    v->tile_csr_val = c.tile_csr_val.as<double>();
    return RTT_OK;
}

} // extern "C"
