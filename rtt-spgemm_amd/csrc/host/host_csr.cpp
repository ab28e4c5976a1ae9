// Host-side pieces of the C ABI: error text, Matrix Market I/O, order files,
// seeded generators and the tile -> CSR conversion.  Pure C++/OpenMP; no
// function in this file touches the GPU.
#include <algorithm>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

#include "../rtt_common.hpp"

namespace rtt {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

const char *last_error_text() { return g_last_error.c_str(); }

} // namespace rtt

using rtt::fail;

namespace {

template <typename T>
T *host_alloc(size_t count)
{
    if (count == 0)
        count = 1;
    return static_cast<T *>(std::malloc(count * sizeof(T)));
}

// Allocates the CSR arrays of *out (zeroed struct first).
int alloc_csr(int m, int n, int64_t nnz, rtt_smatrix *out)
{
    if (nnz > INT_MAX)
        return fail(RTT_ERR_OVERFLOW, "nnz %lld exceeds the int32 CSR ABI",
                    (long long)nnz);
    std::memset(out, 0, sizeof(*out));
    out->m = m;
    out->n = n;
    out->nnz = static_cast<int>(nnz);
    out->rowpointer = host_alloc<int>(static_cast<size_t>(m) + 1);
    out->columnindex = host_alloc<int>(static_cast<size_t>(nnz));
    out->value = host_alloc<double>(static_cast<size_t>(nnz));
    if (!out->rowpointer || !out->columnindex || !out->value) {
        rtt_destroy_smatrix(out);
        return fail(RTT_ERR_OOM, "host CSR allocation failed");
    }
    return RTT_OK;
}

} // namespace

int rtt::alloc_host_csr(int m, int n, int64_t nnz, rtt_smatrix *out)
{
    return alloc_csr(m, n, nnz, out);
}

namespace {

// Matrix Market banner and size line (the NIST mmio conventions used by
// src/mmio.h:398-600).
struct MtxHeader {
    long long m = 0, n = 0, nnz = 0;
    bool pattern = false, real = false, complex_ = false, integer = false;
    bool symmetric = false;
};

int read_mtx_header(FILE *f, MtxHeader *h)
{
    char line[1025];
    if (!std::fgets(line, sizeof(line), f))
        return fail(RTT_ERR_INVALID, "empty Matrix Market file");
    char banner[64], obj[64], fmt[64], field[64], sym[64];
    if (std::sscanf(line, "%63s %63s %63s %63s %63s", banner, obj, fmt, field,
                    sym) != 5)
        return fail(RTT_ERR_INVALID, "could not process Matrix Market banner");
    for (char *p : {obj, fmt, field, sym})
        for (; *p; ++p)
            *p = static_cast<char>(std::tolower(static_cast<unsigned char>(*p)));
    if (std::strncmp(banner, "%%MatrixMarket", 14) != 0 ||
        std::strcmp(obj, "matrix") != 0 || std::strcmp(fmt, "coordinate") != 0)
        return fail(RTT_ERR_INVALID, "not a coordinate Matrix Market matrix");
    h->real = !std::strcmp(field, "real");
    h->complex_ = !std::strcmp(field, "complex");
    h->pattern = !std::strcmp(field, "pattern");
    h->integer = !std::strcmp(field, "integer");
    if (!(h->real || h->complex_ || h->pattern || h->integer))
        return fail(RTT_ERR_INVALID, "unsupported Matrix Market field '%s'", field);
    h->symmetric = !std::strcmp(sym, "symmetric") || !std::strcmp(sym, "hermitian");
    do {
        if (!std::fgets(line, sizeof(line), f))
            return fail(RTT_ERR_INVALID, "premature end of Matrix Market file");
    } while (line[0] == '%');
    while (std::sscanf(line, "%lld %lld %lld", &h->m, &h->n, &h->nnz) != 3) {
        if (!std::fgets(line, sizeof(line), f))
            return fail(RTT_ERR_INVALID, "missing Matrix Market size line");
    }
    if (h->m < 0 || h->n < 0 || h->nnz < 0 || h->m > INT_MAX || h->n > INT_MAX ||
        h->nnz > INT_MAX)
        return fail(RTT_ERR_OVERFLOW, "Matrix Market sizes exceed int32");
    return RTT_OK;
}

} // namespace

extern "C" {

const char *rtt_last_error(void) { return rtt::last_error_text(); }

const char *rtt_version(void) { return "rtt-spgemm-mi355x 0.1 (gfx950)"; }

void rtt_destroy_smatrix(rtt_smatrix *m)
{
    if (!m)
        return;
    std::free(m->value);
    std::free(m->columnindex);
    std::free(m->rowpointer);
    std::free(m->tile_ptr);
    std::free(m->tile_columnidx);
    std::free(m->tile_rowidx);
    std::free(m->tile_nnz);
    std::free(m->tile_csr_Value);
    std::free(m->tile_csr_Col);
    std::free(m->tile_csr_Ptr);
    std::free(m->mask);
    std::free(m->csc_tile_ptr);
    std::free(m->csc_tile_rowidx);
    std::memset(m, 0, sizeof(*m));
}

// mmio_allinone (src/mmio_highlevel.h:593-759): counting placement in file
// order; a symmetric off-diagonal entry (i,j) goes to row i, then its mirror
// to row j; duplicates and unsorted rows are kept.
int rtt_read_mtx(const char *path, int reference_values, rtt_smatrix *out)
{
    rtt::clear_error();
    if (!path || !out)
        return fail(RTT_ERR_INVALID, "rtt_read_mtx: NULL argument");
    FILE *f = std::fopen(path, "r");
    if (!f)
        return fail(RTT_ERR_INVALID, "unable to open %s", path);
    MtxHeader h;
    int rc = read_mtx_header(f, &h);
    if (rc != RTT_OK) {
        std::fclose(f);
        return rc;
    }
    std::vector<int> ri(h.nnz), ci(h.nnz);
    std::vector<double> vv(h.nnz);
    for (long long i = 0; i < h.nnz; ++i) {
        int a = 0, b = 0, iv = 0;
        double x = 0.0, im = 0.0;
        int got;
        if (h.real)
            got = std::fscanf(f, "%d %d %lg\n", &a, &b, &x);
        else if (h.complex_)
            got = std::fscanf(f, "%d %d %lg %lg\n", &a, &b, &x, &im);
        else if (h.integer) {
            got = std::fscanf(f, "%d %d %d\n", &a, &b, &iv);
            x = iv;
        } else {
            got = std::fscanf(f, "%d %d\n", &a, &b);
            x = 1.0;
        }
        if (got < 2 || a < 1 || b < 1 || a > h.m || b > h.n) {
            std::fclose(f);
            return fail(RTT_ERR_INVALID, "%s: bad entry %lld", path, i + 1);
        }
        ri[i] = a - 1;
        ci[i] = b - 1;
        vv[i] = x;
    }
    std::fclose(f);
    std::vector<long long> cnt(h.m + 1, 0);
    for (long long i = 0; i < h.nnz; ++i) {
        cnt[ri[i]]++;
        if (h.symmetric && ri[i] != ci[i])
            cnt[ci[i]]++;
    }
    long long total = 0;
    for (long long r = 0; r < h.m; ++r)
        total += cnt[r];
    rc = alloc_csr(static_cast<int>(h.m), static_cast<int>(h.n), total, out);
    if (rc != RTT_OK)
        return rc;
    out->isSymmetric = h.symmetric ? 1 : 0;
    out->rowpointer[0] = 0;
    for (long long r = 0; r < h.m; ++r)
        out->rowpointer[r + 1] = out->rowpointer[r] + static_cast<int>(cnt[r]);
    std::vector<int> fill(h.m, 0);
    for (long long i = 0; i < h.nnz; ++i) {
        const int r = ri[i], c = ci[i];
        int o = out->rowpointer[r] + fill[r]++;
        out->columnindex[o] = c;
        out->value[o] = vv[i];
        if (h.symmetric && r != c) {
            o = out->rowpointer[c] + fill[c]++;
            out->columnindex[o] = r;
            out->value[o] = vv[i];
        }
    }
    if (reference_values)
        for (int i = 0; i < out->nnz; ++i)
            out->value[i] = static_cast<double>(i % 10);
    return RTT_OK;
}

int rtt_write_mtx(const char *path, const rtt_smatrix *csr)
{
    rtt::clear_error();
    if (!path || !csr || !csr->rowpointer)
        return fail(RTT_ERR_INVALID, "rtt_write_mtx: NULL argument");
    FILE *f = std::fopen(path, "w");
    if (!f)
        return fail(RTT_ERR_INVALID, "unable to create %s", path);
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%d %d %d\n", csr->m, csr->n, csr->nnz);
    for (int r = 0; r < csr->m; ++r)
        for (int e = csr->rowpointer[r]; e < csr->rowpointer[r + 1]; ++e)
            std::fprintf(f, "%d %d %.17g\n", r + 1, csr->columnindex[e] + 1,
                         csr->value[e]);
    const bool ok = std::fclose(f) == 0;
    return ok ? RTT_OK : fail(RTT_ERR_INVALID, "write to %s failed", path);
}

// Order files: exactly `expected` 0-based decimal tokens, token new_id gives
// old_id; short files, extra tokens, duplicates and out-of-range ids are
// rejected (src/dmma_reorder.h:2048-2120).
int rtt_read_order_file(const char *path, int expected, int *out)
{
    rtt::clear_error();
    if (!path || !out || expected < 0)
        return fail(RTT_ERR_INVALID, "rtt_read_order_file: bad argument");
    FILE *f = std::fopen(path, "r");
    if (!f)
        return fail(RTT_ERR_INVALID, "unable to open order file %s", path);
    std::vector<char> seen(static_cast<size_t>(expected), 0);
    int count = 0;
    char tok[64];
    int rc = RTT_OK;
    while (std::fscanf(f, "%63s", tok) == 1) {
        char *end = nullptr;
        const long long v = std::strtoll(tok, &end, 10);
        if (*end != '\0' || tok[0] == '-' || tok[0] == '+') {
            rc = fail(RTT_ERR_INVALID, "%s: non-decimal token '%s'", path, tok);
            break;
        }
        if (count >= expected) {
            rc = fail(RTT_ERR_INVALID, "%s: more than %d tokens", path, expected);
            break;
        }
        if (v < 0 || v >= expected || seen[v]) {
            rc = fail(RTT_ERR_INVALID, "%s: token %lld out of range or repeated",
                      path, v);
            break;
        }
        seen[v] = 1;
        out[count++] = static_cast<int>(v);
    }
    std::fclose(f);
    if (rc == RTT_OK && count != expected)
        rc = fail(RTT_ERR_INVALID, "%s: %d tokens, expected %d", path, count,
                  expected);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Generators                                                                */
/* ------------------------------------------------------------------------ */

int rtt_gen_laplace2d(int g, double pert, uint64_t seed, rtt_smatrix *out)
{
    rtt::clear_error();
    if (g <= 0 || !out)
        return fail(RTT_ERR_INVALID, "rtt_gen_laplace2d: bad argument");
    const long long m = static_cast<long long>(g) * g;
    if (m > INT_MAX)
        return fail(RTT_ERR_OVERFLOW, "grid too large");
    const long long nnz = 5 * m - 4LL * g;
    int rc = alloc_csr(static_cast<int>(m), static_cast<int>(m), nnz, out);
    if (rc != RTT_OK)
        return rc;
    long long e = 0;
    for (int y = 0; y < g; ++y)
        for (int x = 0; x < g; ++x) {
            const int row = x + g * y;
            out->rowpointer[row] = static_cast<int>(e);
            const int cols[5] = {row - g, row - 1, row, row + 1, row + g};
            const bool ok[5] = {y > 0, x > 0, true, x < g - 1, y < g - 1};
            for (int t = 0; t < 5; ++t) {
                if (!ok[t])
                    continue;
                out->columnindex[e] = cols[t];
                out->value[e] = t == 2 ? 4.0
                                       : -1.0 + pert * (2.0 * rtt::uniform_at(seed, e) - 1.0);
                ++e;
            }
        }
    out->rowpointer[m] = static_cast<int>(e);
    return RTT_OK;
}

int rtt_gen_stencil27(int g, double lo, double hi, uint64_t seed,
                      rtt_smatrix *out)
{
    rtt::clear_error();
    if (g <= 0 || !out)
        return fail(RTT_ERR_INVALID, "rtt_gen_stencil27: bad argument");
    const long long g2 = static_cast<long long>(g) * g;
    const long long m = g2 * g;
    const long long side = 3LL * g - 2;
    const long long nnz = g == 1 ? 1 : side * side * side;
    if (m > INT_MAX)
        return fail(RTT_ERR_OVERFLOW, "grid too large");
    int rc = alloc_csr(static_cast<int>(m), static_cast<int>(m), nnz, out);
    if (rc != RTT_OK)
        return rc;
    auto span = [g](int v) { return (v > 0 ? 1 : 0) + 1 + (v < g - 1 ? 1 : 0); };
    // Row lengths are a product of per-axis spans; prefix-sum them.
    std::vector<long long> len(static_cast<size_t>(m));
#pragma omp parallel for
    for (long long row = 0; row < m; ++row) {
        const int x = static_cast<int>(row % g), y = static_cast<int>((row / g) % g),
                  z = static_cast<int>(row / g2);
        len[row] = static_cast<long long>(span(x)) * span(y) * span(z);
    }
    long long run = 0;
    for (long long row = 0; row < m; ++row) {
        out->rowpointer[row] = static_cast<int>(run);
        run += len[row];
    }
    out->rowpointer[m] = static_cast<int>(run);
#pragma omp parallel for schedule(static)
    for (long long row = 0; row < m; ++row) {
        const int x = static_cast<int>(row % g), y = static_cast<int>((row / g) % g),
                  z = static_cast<int>(row / g2);
        long long e = out->rowpointer[row];
        for (int dz = -1; dz <= 1; ++dz) {
            if (z + dz < 0 || z + dz >= g)
                continue;
            for (int dy = -1; dy <= 1; ++dy) {
                if (y + dy < 0 || y + dy >= g)
                    continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    if (x + dx < 0 || x + dx >= g)
                        continue;
                    out->columnindex[e] =
                        static_cast<int>(row + dx + static_cast<long long>(g) * dy + g2 * dz);
                    out->value[e] = lo + (hi - lo) * rtt::uniform_at(seed, e);
                    ++e;
                }
            }
        }
    }
    return RTT_OK;
}

int rtt_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                 uint64_t seed, rtt_smatrix *out)
{
    rtt::clear_error();
    if (scale <= 0 || scale > 30 || edge_factor <= 0 || !out || a < 0 || b < 0 ||
        c < 0 || a + b + c > 1.0)
        return fail(RTT_ERR_INVALID, "rtt_gen_rmat: bad argument");
    const long long n = 1LL << scale;
    const long long draws = n * edge_factor;
    std::vector<uint64_t> keys(static_cast<size_t>(draws));
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < draws; ++i) {
        uint64_t row = 0, col = 0;
        for (int level = 0; level < scale; ++level) {
            const double u = rtt::uniform_at(seed, static_cast<uint64_t>(i) * scale + level);
            const uint64_t rb = u >= ab ? 1 : 0;               // c or d quadrant
            const uint64_t cb = (u >= a && u < ab) || u >= abc ? 1 : 0;  // b or d
            row = (row << 1) | rb;
            col = (col << 1) | cb;
        }
        keys[i] = (row << 32) | col;
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const long long nnz = static_cast<long long>(keys.size());
    int rc = alloc_csr(static_cast<int>(n), static_cast<int>(n), nnz, out);
    if (rc != RTT_OK)
        return rc;
    std::vector<int> cnt(static_cast<size_t>(n) + 1, 0);
    for (long long e = 0; e < nnz; ++e)
        cnt[keys[e] >> 32]++;
    out->rowpointer[0] = 0;
    for (long long r = 0; r < n; ++r)
        out->rowpointer[r + 1] = out->rowpointer[r] + cnt[r];
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < nnz; ++e) {
        out->columnindex[e] = static_cast<int>(keys[e] & 0xffffffffu);
        out->value[e] = 2.0 * rtt::uniform_at(seed ^ 0x5bd1e995ull, e) - 1.0;
    }
    return RTT_OK;
}

int rtt_gen_random(int m, int n, int per_row, uint64_t seed, rtt_smatrix *out)
{
    rtt::clear_error();
    if (m < 0 || n < 0 || per_row < 0 || !out || (n == 0 && per_row > 0))
        return fail(RTT_ERR_INVALID, "rtt_gen_random: bad argument");
    const int k = std::min(per_row, n);
    const long long nnz = static_cast<long long>(m) * k;
    int rc = alloc_csr(m, n, nnz, out);
    if (rc != RTT_OK)
        return rc;
    for (int r = 0; r <= m; ++r)
        out->rowpointer[r] = r * k;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < m; ++r) {
        rtt::SplitMix64 gen(seed ^ (0x9E3779B97F4A7C15ull * (static_cast<uint64_t>(r) + 1)));
        std::vector<int> cols;
        cols.reserve(k);
        while (static_cast<int>(cols.size()) < k) {
            const int c = static_cast<int>(gen.next() % static_cast<uint64_t>(n));
            if (std::find(cols.begin(), cols.end(), c) == cols.end())
                cols.push_back(c);
        }
        std::sort(cols.begin(), cols.end());
        for (int t = 0; t < k; ++t) {
            out->columnindex[static_cast<long long>(r) * k + t] = cols[t];
            out->value[static_cast<long long>(r) * k + t] = 2.0 * gen.uniform() - 1.0;
        }
    }
    return RTT_OK;
}

/* ------------------------------------------------------------------------ */
/* tile2csr / tile2csr_restore_rows (src/tile2csr.h:232-523)                 */
/* ------------------------------------------------------------------------ */

} // extern "C"

namespace {

constexpr int kOutRows = 8;
constexpr int kOutCols = 8;

inline int out_row_count(int tile_row, int m)
{
    const long long rem = static_cast<long long>(m) - static_cast<long long>(tile_row) * kOutRows;
    if (rem <= 0)
        return 0;
    return rem < kOutRows ? static_cast<int>(rem) : kOutRows;
}

// Clamps a tile's row bounds into the tile (src/tile2csr.h:48-73).
inline void row_bounds(const unsigned char *ptr, size_t off, int lr, int rc,
                       int tnnz, int *s, int *e)
{
    int a = ptr[off + lr];
    int b = lr + 1 == rc ? tnnz : ptr[off + lr + 1];
    a = std::min(std::max(a, 0), tnnz);
    b = std::min(std::max(b, a), tnnz);
    *s = a;
    *e = b;
}

inline bool valid_col(int tile_col, unsigned char lc, int n, int *col)
{
    if (lc >= kOutCols || tile_col < 0)
        return false;
    const long long c = static_cast<long long>(tile_col) * kOutCols + lc;
    if (c >= n)
        return false;
    *col = static_cast<int>(c);
    return true;
}

void sort_rows(rtt_smatrix *m)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int r = 0; r < m->m; ++r) {
        const int b = m->rowpointer[r], e = m->rowpointer[r + 1];
        std::vector<std::pair<int, double>> row;
        row.reserve(e - b);
        for (int x = b; x < e; ++x)
            row.push_back({m->columnindex[x], m->value[x]});
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<int, double> &p, const std::pair<int, double> &q) {
                             return p.first < q.first;
                         });
        for (int x = b; x < e; ++x) {
            m->columnindex[x] = row[x - b].first;
            m->value[x] = row[x - b].second;
        }
    }
}

int tile2csr_impl(rtt_smatrix *m, const int *row_new_to_old, int perm_size,
                  int canonical)
{
    rtt::clear_error();
    if (!m || m->m < 0 || m->n < 0 || m->tilem < 0 || m->numtile < 0 ||
        !m->tile_ptr || !m->tile_nnz ||
        (m->numtile > 0 && (!m->tile_columnidx || !m->tile_csr_Ptr)) ||
        (m->nnz > 0 && (!m->tile_csr_Col || !m->tile_csr_Value)))
        return fail(RTT_ERR_INVALID, "tile2csr: malformed tile matrix");
    const bool restore = row_new_to_old != nullptr;
    if (restore || perm_size != 0) {
        if (perm_size < 0 || perm_size > m->m || (perm_size > 0 && !row_new_to_old))
            return fail(RTT_ERR_INVALID, "tile2csr_restore_rows: bad permutation size");
        std::vector<unsigned char> seen(static_cast<size_t>(m->m), 0);
        for (int i = 0; i < perm_size; ++i) {
            const int o = row_new_to_old[i];
            if (o < 0 || o >= m->m || seen[o])
                return fail(RTT_ERR_INVALID,
                            "tile2csr_restore_rows: permutation is not injective");
            seen[o] = 1;
        }
    }
    std::vector<long long> counts(static_cast<size_t>(m->m) + 1, 0);
    int tail_nonempty = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : tail_nonempty)
    for (int tr = 0; tr < m->tilem; ++tr) {
        const int rc = out_row_count(tr, m->m);
        long long local[kOutRows] = {};
        for (int t = m->tile_ptr[tr]; t < m->tile_ptr[tr + 1]; ++t) {
            const int tnnz = m->tile_nnz[t + 1] - m->tile_nnz[t];
            const int base = m->tile_nnz[t];
            for (int lr = 0; lr < rc; ++lr) {
                int s, e, col;
                row_bounds(m->tile_csr_Ptr, static_cast<size_t>(t) * kOutRows, lr, rc,
                           tnnz, &s, &e);
                for (int x = s; x < e; ++x)
                    if (valid_col(m->tile_columnidx[t], m->tile_csr_Col[base + x], m->n, &col))
                        ++local[lr];
            }
        }
        for (int lr = 0; lr < rc; ++lr) {
            const int row = tr * kOutRows + lr;
            if (!restore)
                counts[row] = local[lr];
            else if (row >= perm_size)
                tail_nonempty |= local[lr] != 0 ? 1 : 0;
            else
                counts[row_new_to_old[row]] = local[lr];
        }
    }
    if (tail_nonempty)
        return fail(RTT_ERR_INVALID,
                    "tile2csr_restore_rows: non-empty output row outside the active prefix");
    long long run = 0;
    for (int r = 0; r <= m->m; ++r) {
        const long long c = counts[r];
        counts[r] = run;
        run += c;
    }
    if (run > INT_MAX)
        return fail(RTT_ERR_OVERFLOW, "tile2csr: nnz(C) %lld exceeds int32", run);
    int *rp = static_cast<int *>(std::malloc((static_cast<size_t>(m->m) + 1) * sizeof(int)));
    int *ci = static_cast<int *>(std::malloc(std::max<size_t>(run, 1) * sizeof(int)));
    double *vv = static_cast<double *>(std::malloc(std::max<size_t>(run, 1) * sizeof(double)));
    if (!rp || !ci || !vv) {
        std::free(rp);
        std::free(ci);
        std::free(vv);
        return fail(RTT_ERR_OOM, "tile2csr: host allocation failed");
    }
    for (int r = 0; r <= m->m; ++r)
        rp[r] = static_cast<int>(counts[r]);
#pragma omp parallel for schedule(dynamic, 64)
    for (int tr = 0; tr < m->tilem; ++tr) {
        const int rc = out_row_count(tr, m->m);
        long long local[kOutRows] = {};
        for (int t = m->tile_ptr[tr]; t < m->tile_ptr[tr + 1]; ++t) {
            const int tnnz = m->tile_nnz[t + 1] - m->tile_nnz[t];
            const int base = m->tile_nnz[t];
            for (int lr = 0; lr < rc; ++lr) {
                const int row = tr * kOutRows + lr;
                if (restore && row >= perm_size)
                    continue;
                const int orow = restore ? row_new_to_old[row] : row;
                int s, e, col;
                row_bounds(m->tile_csr_Ptr, static_cast<size_t>(t) * kOutRows, lr, rc,
                           tnnz, &s, &e);
                for (int x = s; x < e; ++x) {
                    if (!valid_col(m->tile_columnidx[t], m->tile_csr_Col[base + x], m->n, &col))
                        continue;
                    const long long o = rp[orow] + local[lr]++;
                    ci[o] = col;
                    vv[o] = m->tile_csr_Value[base + x];
                }
            }
        }
    }
    std::free(m->rowpointer);
    std::free(m->columnindex);
    std::free(m->value);
    m->rowpointer = rp;
    m->columnindex = ci;
    m->value = vv;
    m->nnz = static_cast<int>(run);
    if (!canonical)
        sort_rows(m);
    return RTT_OK;
}

} // namespace

extern "C" {

int rtt_tile2csr(rtt_smatrix *m, int canonical)
{
    return tile2csr_impl(m, nullptr, 0, canonical);
}

int rtt_tile2csr_restore_rows(rtt_smatrix *m, const int *row_new_to_old,
                              int permutation_size, int canonical)
{
    if (!row_new_to_old && permutation_size > 0)
        return fail(RTT_ERR_INVALID, "tile2csr_restore_rows: NULL permutation");
    if (!row_new_to_old) {
        // An empty prefix: every output row must be empty.
        static const int dummy = 0;
        return tile2csr_impl(m, &dummy, 0, canonical);
    }
    return tile2csr_impl(m, row_new_to_old, permutation_size, canonical);
}

} // extern "C"
