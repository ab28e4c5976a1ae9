/*
 * CPU oracle for the tiled FP64 SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY: the shipped library never links or loads this
 * file.  It is a fresh restatement of the reference's CPU semantics; each
 * function cites the reference file:line it follows.  It is deliberately
 * plain (scalar loops, OpenMP only over independent rows) so that it is easy
 * to audit against the cited sources.
 */
#include "rtt_oracle.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

/* ------------------------------------------------------------------------ */
/* Matrix Market banner / size parsing (src/mmio.h:398-600, the NIST mmio)   */
/* ------------------------------------------------------------------------ */

struct MmHeader {
    int m = 0, n = 0;
    int64_t nnz = 0;
    bool pattern = false, real = false, complex_ = false, integer = false;
    bool symmetric = false;
};

/* Returns 0 and leaves `f` positioned after the size line, or a negative code
 * mirroring mmio_allinone's early returns (-2 banner, -4 size). */
int read_header(FILE *f, MmHeader *h)
{
    char line[1025];
    if (!std::fgets(line, sizeof(line), f))
        return -2;
    char banner[64], mtx[64], crd[64], dtype[64], storage[64];
    if (std::sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dtype,
                    storage) != 5)
        return -2;
    for (char *p : {mtx, crd, dtype, storage})
        for (; *p; ++p)
            *p = static_cast<char>(std::tolower(static_cast<unsigned char>(*p)));
    if (std::strncmp(banner, "%%MatrixMarket", 14) != 0 ||
        std::strcmp(mtx, "matrix") != 0 || std::strcmp(crd, "coordinate") != 0)
        return -2;
    h->real = std::strcmp(dtype, "real") == 0;
    h->complex_ = std::strcmp(dtype, "complex") == 0;
    h->pattern = std::strcmp(dtype, "pattern") == 0;
    h->integer = std::strcmp(dtype, "integer") == 0;
    if (!(h->real || h->complex_ || h->pattern || h->integer))
        return -2;
    /* mmio_allinone expands symmetric and hermitian storage only
     * (src/mmio_highlevel.h:631-635); skew-symmetric is read as stored. */
    h->symmetric = std::strcmp(storage, "symmetric") == 0 ||
                   std::strcmp(storage, "hermitian") == 0;
    do {
        if (!std::fgets(line, sizeof(line), f))
            return -4;
    } while (line[0] == '%');
    long long m = 0, n = 0, nz = 0;
    while (std::sscanf(line, "%lld %lld %lld", &m, &n, &nz) != 3) {
        if (!std::fgets(line, sizeof(line), f))
            return -4;
    }
    h->m = static_cast<int>(m);
    h->n = static_cast<int>(n);
    h->nnz = nz;
    return 0;
}

struct Triplets {
    std::vector<int> r, c;
    std::vector<double> v;
};

int read_triplets(const char *path, MmHeader *h, Triplets *t)
{
    FILE *f = std::fopen(path, "r");
    if (!f)
        return -1;
    int rc = read_header(f, h);
    if (rc != 0) {
        std::fclose(f);
        return rc;
    }
    t->r.resize(h->nnz);
    t->c.resize(h->nnz);
    t->v.resize(h->nnz);
    /* src/mmio_highlevel.h:646-681: one entry per line; pattern entries are
     * 1.0, integer entries are converted, complex keeps the real part. */
    for (int64_t i = 0; i < h->nnz; ++i) {
        int ii = 0, jj = 0;
        double fv = 0.0, fim = 0.0;
        int iv = 0;
        if (h->real)
            (void)std::fscanf(f, "%d %d %lg\n", &ii, &jj, &fv);
        else if (h->complex_)
            (void)std::fscanf(f, "%d %d %lg %lg\n", &ii, &jj, &fv, &fim);
        else if (h->integer) {
            (void)std::fscanf(f, "%d %d %d\n", &ii, &jj, &iv);
            fv = iv;
        } else {
            (void)std::fscanf(f, "%d %d\n", &ii, &jj);
            fv = 1.0;
        }
        t->r[i] = ii - 1;
        t->c[i] = jj - 1;
        t->v[i] = fv;
    }
    std::fclose(f);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Double-double helpers (Dekker / Knuth error-free transformations)         */
/* ------------------------------------------------------------------------ */

inline void two_sum(double a, double b, double *s, double *e)
{
    const double x = a + b;
    const double z = x - a;
    *e = (a - (x - z)) + (b - z);
    *s = x;
}

inline void dd_add_prod(double *hi, double *lo, double a, double b)
{
    const double p = a * b;
    const double pe = std::fma(a, b, -p);
    double s, e;
    two_sum(*hi, p, &s, &e);
    e += *lo + pe;
    two_sum(s, e, hi, lo);
}

/* ------------------------------------------------------------------------ */
/* Hybrid tile enumeration (src/dmma_tiles.h:163-234)                        */
/* ------------------------------------------------------------------------ */

struct RawEntry {
    int tile_col;
    int physical;
    double value;
};

/* Calls consume(tile_col, mask, physical_values[32], unique) for every tile
 * of one tile row in increasing tile-column order.  Duplicates are summed in
 * CSR order (the reference sorts with an unstable std::sort, so its order of
 * duplicate summation is unspecified; a stable order is one valid choice). */
template <typename F>
void for_each_tile(int m, const int *rp, const int *ci, const double *v,
                   int tile_rows, int tile_cols, bool col_major, int tile_row,
                   F &&consume)
{
    const int first = tile_row * tile_rows;
    const int last = std::min(m, first + tile_rows);
    std::vector<RawEntry> entries;
    for (int row = first; row < last; ++row) {
        const int local_row = row - first;
        for (int e = rp[row]; e < rp[row + 1]; ++e) {
            const int col = ci[e];
            const int local_col = col % tile_cols;
            const int phys = col_major ? local_col * tile_rows + local_row
                                       : local_row * tile_cols + local_col;
            entries.push_back({col / tile_cols, phys, v ? v[e] : 0.0});
        }
    }
    std::stable_sort(entries.begin(), entries.end(),
                     [](const RawEntry &a, const RawEntry &b) {
                         if (a.tile_col != b.tile_col)
                             return a.tile_col < b.tile_col;
                         return a.physical < b.physical;
                     });
    size_t cur = 0;
    while (cur < entries.size()) {
        const int tc = entries[cur].tile_col;
        double pv[32] = {};
        uint32_t mask = 0;
        int unique = 0;
        while (cur < entries.size() && entries[cur].tile_col == tc) {
            const int phys = entries[cur].physical;
            double sum = 0.0;
            do {
                sum += entries[cur].value;
                ++cur;
            } while (cur < entries.size() && entries[cur].tile_col == tc &&
                     entries[cur].physical == phys);
            pv[phys] = sum;
            mask |= 1u << phys;
            ++unique;
        }
        consume(tc, mask, pv, unique);
    }
}

/* validate_csr (src/dmma_tiles.h:126-161) */
bool valid_csr(int m, int n, const int *rp, const int *ci)
{
    if (m < 0 || n < 0)
        return false;
    if (m == 0)
        return true;
    if (!rp || rp[0] != 0)
        return false;
    for (int r = 0; r < m; ++r)
        if (rp[r + 1] < rp[r])
            return false;
    for (int e = 0; e < rp[m]; ++e)
        if (ci[e] < 0 || ci[e] >= n)
            return false;
    return true;
}

/* tile2csr_row_bounds (src/tile2csr.h:48-73) */
inline void row_bounds(const uint8_t *ptr, size_t off, int local_row,
                       int row_count, int tilennz, int *s, int *e)
{
    int a = ptr[off + local_row];
    int b = local_row + 1 == row_count ? tilennz : ptr[off + local_row + 1];
    if (a < 0)
        a = 0;
    if (a > tilennz)
        a = tilennz;
    if (b < a)
        b = a;
    if (b > tilennz)
        b = tilennz;
    *s = a;
    *e = b;
}

inline int tile_row_count(int tile_row, int m)
{
    const long long rem = (long long)m - (long long)tile_row * 8;
    if (rem <= 0)
        return 0;
    return rem < 8 ? (int)rem : 8;
}

} // namespace

extern "C" {

/* ---- Matrix Market ------------------------------------------------------ */

int orc_mmio_info(const char *path, int *m, int *n, int64_t *nnz_file,
                  int *is_symmetric)
{
    FILE *f = std::fopen(path, "r");
    if (!f)
        return -1;
    MmHeader h;
    const int rc = read_header(f, &h);
    std::fclose(f);
    if (rc != 0)
        return rc;
    *m = h.m;
    *n = h.n;
    *nnz_file = h.nnz;
    *is_symmetric = h.symmetric ? 1 : 0;
    return 0;
}

int64_t orc_mmio_expanded_nnz(const char *path)
{
    MmHeader h;
    Triplets t;
    if (read_triplets(path, &h, &t) != 0)
        return -1;
    int64_t nnz = h.nnz;
    if (h.symmetric)
        for (int64_t i = 0; i < h.nnz; ++i)
            if (t.r[i] != t.c[i])
                ++nnz;
    return nnz;
}

/* mmio_allinone (src/mmio_highlevel.h:683-741): counting placement in file
 * order; a symmetric off-diagonal entry (i,j) is placed in row i and then
 * mirrored into row j at the same file position. */
int orc_mmio_read(const char *path, int *row_ptr, int *col_idx,
                  double *values)
{
    MmHeader h;
    Triplets t;
    const int rc = read_triplets(path, &h, &t);
    if (rc != 0)
        return rc;
    std::vector<int64_t> cnt(h.m + 1, 0);
    for (int64_t i = 0; i < h.nnz; ++i) {
        cnt[t.r[i]]++;
        if (h.symmetric && t.r[i] != t.c[i])
            cnt[t.c[i]]++;
    }
    std::vector<int64_t> start(h.m + 1, 0);
    for (int r = 0; r < h.m; ++r)
        start[r + 1] = start[r] + cnt[r];
    for (int r = 0; r <= h.m; ++r)
        row_ptr[r] = static_cast<int>(start[r]);
    std::vector<int64_t> fill(h.m, 0);
    for (int64_t i = 0; i < h.nnz; ++i) {
        const int r = t.r[i], c = t.c[i];
        int64_t o = start[r] + fill[r]++;
        col_idx[o] = c;
        values[o] = t.v[i];
        if (h.symmetric && r != c) {
            o = start[c] + fill[c]++;
            col_idx[o] = r;
            values[o] = t.v[i];
        }
    }
    return 0;
}

/* ---- nnzCub ------------------------------------------------------------- */

/* AA: sum over entries e of rowlen(col e); AAT: sum of coldeg(col e)
 * (src/gpu_dmma_tiles.h:744-763, 2052-2101).  Duplicates count separately. */
uint64_t orc_nnz_cub(int m, int n, const int *row_ptr, const int *col_idx,
                     int aat)
{
    const int nnz = m > 0 ? row_ptr[m] : 0;
    std::vector<uint64_t> deg;
    if (aat) {
        deg.assign(n, 0);
        for (int e = 0; e < nnz; ++e)
            deg[col_idx[e]]++;
    }
    uint64_t total = 0;
    for (int e = 0; e < nnz; ++e) {
        const int c = col_idx[e];
        total += aat ? deg[c] : (uint64_t)(row_ptr[c + 1] - row_ptr[c]);
    }
    return total;
}

/* ---- Transpose (src/utils.h:158-197) ------------------------------------ */

void orc_transpose(int m, int n, const int *row_ptr, const int *col_idx,
                   const double *values, int *t_row_ptr, int *t_col_idx,
                   double *t_values)
{
    const int nnz = m > 0 ? row_ptr[m] : 0;
    std::vector<int64_t> ptr(n + 1, 0);
    for (int e = 0; e < nnz; ++e)
        ptr[col_idx[e] + 1]++;
    for (int c = 0; c < n; ++c)
        ptr[c + 1] += ptr[c];
    for (int c = 0; c <= n; ++c)
        t_row_ptr[c] = static_cast<int>(ptr[c]);
    std::vector<int64_t> next(ptr.begin(), ptr.end() - 1);
    for (int r = 0; r < m; ++r)
        for (int e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
            const int64_t o = next[col_idx[e]]++;
            t_col_idx[o] = r;
            if (values && t_values)
                t_values[o] = values[e];
        }
}

/* ---- Row SPA SpGEMM ------------------------------------------------------ */

/* spgemm_spa structure pass (src/spgemm_serialref_spa_new.h:29-62): one bit
 * flag per output column, popcount per row, exclusive scan. */
int64_t orc_spgemm_count(int m_a, int n_a, const int *rp_a, const int *ci_a,
                         int m_b, int n_b, const int *rp_b, const int *ci_b,
                         int64_t *rp_c)
{
    if (n_a != m_b)
        return -1;
    const int words = n_b / 32 + 1;
#pragma omp parallel
    {
        // Column bitmap of the row, and the words the row set: only those
        // are counted and cleared, so a row costs its products, not n_b.
        std::vector<uint32_t> flag(words, 0u);
        std::vector<int> touched;
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < m_a; ++i) {
            for (int e = rp_a[i]; e < rp_a[i + 1]; ++e) {
                const int k = ci_a[e];
                for (int l = rp_b[k]; l < rp_b[k + 1]; ++l) {
                    const int key = ci_b[l];
                    if (!flag[key / 32])
                        touched.push_back(key / 32);
                    flag[key / 32] |= 1u << (key % 32);
                }
            }
            int64_t cnt = 0;
            for (const int w : touched) {
                cnt += __builtin_popcount(flag[w]);
                flag[w] = 0u;
            }
            touched.clear();
            rp_c[i + 1] = cnt;
        }
    }
    rp_c[0] = 0;
    for (int i = 0; i < m_a; ++i)
        rp_c[i + 1] += rp_c[i];
    return rp_c[m_a];
}

/* Columns: spgemm_spa fill pass (src/spgemm_serialref_spa_new.h:63-103),
 * ascending bit order.  Values: spgemm_serialref's dense row
 * (src/external/cusparse/spgemm_serialref_spa.h:116-139, 198-221): for each A
 * entry in row order, for each B entry in row order, acc[col] += a*b. */
int orc_spgemm_fill(int m_a, int n_a, const int *rp_a, const int *ci_a,
                    const double *v_a, int m_b, int n_b, const int *rp_b,
                    const int *ci_b, const double *v_b, const int64_t *rp_c,
                    int *ci_c, double *v_c, double *v_c_hi, double *v_c_lo,
                    double *abs_c)
{
    if (n_a != m_b)
        return -1;
    const int words = n_b / 32 + 1;
    int status = 0;
#pragma omp parallel
    {
        std::vector<uint32_t> flag(words, 0u);
        std::vector<int> touched;   // words the row set, emitted in ascending order
        std::vector<double> acc(v_c ? n_b : 0);
        std::vector<double> hi(v_c_hi ? n_b : 0), lo(v_c_hi ? n_b : 0);
        std::vector<double> ab(abs_c ? n_b : 0);
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < m_a; ++i) {
            for (int e = rp_a[i]; e < rp_a[i + 1]; ++e) {
                const int k = ci_a[e];
                const double a = v_a ? v_a[e] : 0.0;
                for (int l = rp_b[k]; l < rp_b[k + 1]; ++l) {
                    const int j = ci_b[l];
                    const uint32_t bit = 1u << (j % 32);
                    const double b = v_b ? v_b[l] : 0.0;
                    if (!(flag[j / 32] & bit)) {
                        if (!flag[j / 32])
                            touched.push_back(j / 32);
                        flag[j / 32] |= bit;
                        if (v_c)
                            acc[j] = 0.0;
                        if (v_c_hi) {
                            hi[j] = 0.0;
                            lo[j] = 0.0;
                        }
                        if (abs_c)
                            ab[j] = 0.0;
                    }
                    if (v_c)
                        acc[j] += a * b;
                    if (v_c_hi)
                        dd_add_prod(&hi[j], &lo[j], a, b);
                    if (abs_c)
                        ab[j] += std::fabs(a) * std::fabs(b);
                }
            }
            int64_t o = rp_c[i];
            std::sort(touched.begin(), touched.end());
            for (const int w : touched) {
                uint32_t f = flag[w];
                flag[w] = 0u;
                while (f) {
                    const int bitpos = __builtin_ctz(f);
                    f &= f - 1;
                    const int j = w * 32 + bitpos;
                    if (o >= rp_c[i + 1]) {
#pragma omp atomic write
                        status = -2;
                        break;
                    }
                    ci_c[o] = j;
                    if (v_c)
                        v_c[o] = acc[j];
                    if (v_c_hi) {
                        v_c_hi[o] = hi[j];
                        v_c_lo[o] = lo[j];
                    }
                    if (abs_c)
                        abs_c[o] = ab[j];
                    ++o;
                }
            }
            touched.clear();
        }
    }
    return status;
}

/* ---- Hybrid tile format -------------------------------------------------- */

/* csr_to_hybrid_tiles sizing pass (src/dmma_tiles.h:243-325). */
int orc_hybrid_tiles_count(int m, int n, const int *row_ptr,
                           const int *col_idx, int tile_rows, int tile_cols,
                           int col_major, int dense_threshold, int64_t *counts)
{
    if (!valid_csr(m, n, row_ptr, col_idx) || tile_rows <= 0 ||
        tile_cols <= 0 || tile_rows * tile_cols != 32 || dense_threshold < 1 ||
        dense_threshold > 32)
        return -1;
    const int trc = (m + tile_rows - 1) / tile_rows;
    int64_t tiles = 0, payload = 0, dense = 0, sparse = 0;
    for (int tr = 0; tr < trc; ++tr)
        for_each_tile(m, row_ptr, col_idx, nullptr, tile_rows, tile_cols,
                      col_major != 0, tr,
                      [&](int, uint32_t, const double *, int unique) {
                          ++tiles;
                          if (unique >= dense_threshold) {
                              payload += 32;
                              ++dense;
                          } else {
                              payload += unique;
                              ++sparse;
                          }
                      });
    counts[0] = tiles;
    counts[1] = payload;
    counts[2] = dense;
    counts[3] = sparse;
    return 0;
}

/* csr_to_hybrid_tiles fill pass and tile CSC (src/dmma_tiles.h:341-424):
 * dense payload keeps all 32 physical slots (absent slots are 0); sparse
 * payload is packed in increasing mask-bit order; the CSC lists tiles of
 * each tile column in increasing tile-row order. */
int orc_hybrid_tiles_fill(int m, int n, const int *row_ptr,
                          const int *col_idx, const double *values,
                          int tile_rows, int tile_cols, int col_major,
                          int dense_threshold, int build_csc,
                          int *tile_row_ptr, int *tile_col_idx,
                          int *value_offsets, uint32_t *masks,
                          double *payload, int *tile_col_ptr,
                          int *tile_row_idx, int *csc_tile_ids)
{
    if (!valid_csr(m, n, row_ptr, col_idx) || tile_rows * tile_cols != 32)
        return -1;
    const int trc = (m + tile_rows - 1) / tile_rows;
    const int tcc = (n + tile_cols - 1) / tile_cols;
    int64_t t = 0, p = 0;
    tile_row_ptr[0] = 0;
    value_offsets[0] = 0;
    for (int tr = 0; tr < trc; ++tr) {
        for_each_tile(m, row_ptr, col_idx, values, tile_rows, tile_cols,
                      col_major != 0, tr,
                      [&](int tc, uint32_t mask, const double *pv, int unique) {
                          tile_col_idx[t] = tc;
                          masks[t] = mask;
                          if (unique >= dense_threshold) {
                              for (int q = 0; q < 32; ++q)
                                  payload[p++] = pv[q];
                          } else {
                              for (int q = 0; q < 32; ++q)
                                  if (mask & (1u << q))
                                      payload[p++] = pv[q];
                          }
                          ++t;
                          value_offsets[t] = static_cast<int>(p);
                      });
        tile_row_ptr[tr + 1] = static_cast<int>(t);
    }
    if (build_csc) {
        std::vector<int64_t> ptr(tcc + 1, 0);
        for (int64_t i = 0; i < t; ++i)
            ptr[tile_col_idx[i] + 1]++;
        for (int c = 0; c < tcc; ++c)
            ptr[c + 1] += ptr[c];
        for (int c = 0; c <= tcc; ++c)
            tile_col_ptr[c] = static_cast<int>(ptr[c]);
        std::vector<int64_t> next(ptr.begin(), ptr.end() - 1);
        for (int tr = 0; tr < trc; ++tr)
            for (int i = tile_row_ptr[tr]; i < tile_row_ptr[tr + 1]; ++i) {
                const int64_t d = next[tile_col_idx[i]]++;
                tile_row_idx[d] = tr;
                csc_tile_ids[d] = i;
            }
    }
    return 0;
}

/* ---- Output 8x8 tiles of C --------------------------------------------- */

int64_t orc_c_tiles_count(int m, int n, const int64_t *rp_c, const int *ci_c,
                          int64_t *tile_ptr)
{
    (void)n;
    const int tilem = (m + 7) / 8;
    tile_ptr[0] = 0;
    for (int tr = 0; tr < tilem; ++tr) {
        std::vector<int> cols;
        for (int r = tr * 8; r < std::min(m, tr * 8 + 8); ++r)
            for (int64_t e = rp_c[r]; e < rp_c[r + 1]; ++e)
                cols.push_back(ci_c[e] / 8);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        tile_ptr[tr + 1] = tile_ptr[tr] + (int64_t)cols.size();
    }
    return tile_ptr[tilem];
}

/* Output compression of the numeric writer (src/dmma_spgemm.h:4261-4281):
 * positions r*8+c present in the tile mask, in increasing position order;
 * tile_csr_ptr[8*t + r] = number of entries of rows < r in the tile. */
int orc_c_tiles_fill(int m, int n, const int64_t *rp_c, const int *ci_c,
                     const double *v_c, const int64_t *tile_ptr,
                     int *tile_col, int64_t *tile_nnz, uint8_t *tile_csr_ptr,
                     uint8_t *tile_csr_col, double *tile_csr_val,
                     uint64_t *tile_mask)
{
    (void)n;
    const int tilem = (m + 7) / 8;
    int64_t vo = 0;
    tile_nnz[0] = 0;
    for (int tr = 0; tr < tilem; ++tr) {
        const int64_t t0 = tile_ptr[tr], t1 = tile_ptr[tr + 1];
        std::vector<int> cols;
        for (int r = tr * 8; r < std::min(m, tr * 8 + 8); ++r)
            for (int64_t e = rp_c[r]; e < rp_c[r + 1]; ++e)
                cols.push_back(ci_c[e] / 8);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        if ((int64_t)cols.size() != t1 - t0)
            return -1;
        std::vector<uint64_t> mask(cols.size(), 0);
        std::vector<double> val(cols.size() * 64, 0.0);
        for (int r = tr * 8; r < std::min(m, tr * 8 + 8); ++r)
            for (int64_t e = rp_c[r]; e < rp_c[r + 1]; ++e) {
                const int J = ci_c[e] / 8;
                const size_t q = std::lower_bound(cols.begin(), cols.end(), J) -
                                 cols.begin();
                const int pos = (r - tr * 8) * 8 + ci_c[e] % 8;
                mask[q] |= 1ull << pos;
                val[q * 64 + pos] = v_c ? v_c[e] : 0.0;
            }
        for (size_t q = 0; q < cols.size(); ++q) {
            const int64_t t = t0 + (int64_t)q;
            tile_col[t] = cols[q];
            if (tile_mask)
                tile_mask[t] = mask[q];
            for (int r = 0; r < 8; ++r) {
                const uint64_t below = r == 0 ? 0 : (mask[q] & ((1ull << (8 * r)) - 1));
                tile_csr_ptr[t * 8 + r] = (uint8_t)__builtin_popcountll(below);
            }
            for (int pos = 0; pos < 64; ++pos)
                if (mask[q] & (1ull << pos)) {
                    tile_csr_col[vo] = (uint8_t)(pos % 8);
                    tile_csr_val[vo] = val[q * 64 + pos];
                    ++vo;
                }
            tile_nnz[t + 1] = vo;
        }
    }
    return 0;
}

/* ---- tile2csr ------------------------------------------------------------ */

/* Counting pass of tile2csr (src/tile2csr.h:243-276) and of
 * tile2csr_restore_rows (src/tile2csr.h:354-442).  Column validity follows
 * tile2csr_valid_global_col (src/tile2csr.h:75-90). */
int64_t orc_tile2csr_count(int m, int n, int tilem, const int *tile_ptr,
                           const int *tile_columnidx, const int *tile_nnz,
                           const uint8_t *tile_csr_ptr,
                           const uint8_t *tile_csr_col,
                           const int *row_new_to_old, int permutation_size,
                           int64_t *row_ptr)
{
    if (m < 0 || n < 0 || tilem < 0 || !tile_ptr || !tile_nnz)
        return -1;
    const bool restore = row_new_to_old != nullptr;
    if (restore) {
        if (permutation_size < 0 || permutation_size > m)
            return -1;
        std::vector<char> seen(m, 0);
        for (int i = 0; i < permutation_size; ++i) {
            const int o = row_new_to_old[i];
            if (o < 0 || o >= m || seen[o])
                return -1;
            seen[o] = 1;
        }
    }
    std::fill(row_ptr, row_ptr + m + 1, 0);
    bool tail_nonempty = false;
    for (int tr = 0; tr < tilem; ++tr) {
        const int rc = tile_row_count(tr, m);
        int64_t local[8] = {};
        for (int t = tile_ptr[tr]; t < tile_ptr[tr + 1]; ++t) {
            const int tnnz = tile_nnz[t + 1] - tile_nnz[t];
            const int base = tile_nnz[t];
            for (int lr = 0; lr < rc; ++lr) {
                int s, e;
                row_bounds(tile_csr_ptr, (size_t)t * 8, lr, rc, tnnz, &s, &e);
                for (int x = s; x < e; ++x) {
                    const int lc = tile_csr_col[base + x];
                    const long long col = (long long)tile_columnidx[t] * 8 + lc;
                    if (lc < 8 && tile_columnidx[t] >= 0 && col < n)
                        ++local[lr];
                }
            }
        }
        for (int lr = 0; lr < rc; ++lr) {
            const int row = tr * 8 + lr;
            if (!restore) {
                row_ptr[row] = local[lr];
            } else if (row >= permutation_size) {
                if (local[lr] != 0)
                    tail_nonempty = true;
            } else {
                row_ptr[row_new_to_old[row]] = local[lr];
            }
        }
    }
    if (tail_nonempty)
        return -1;
    int64_t run = 0;
    for (int r = 0; r <= m; ++r) {
        const int64_t c = row_ptr[r];
        row_ptr[r] = run;
        run += c;
    }
    return row_ptr[m];
}

/* Fill pass (src/tile2csr.h:301-327, 464-521) plus the optional per-row
 * column sort of tile2csr_sort_rows (src/tile2csr.h:205-223) when the caller
 * does not assert canonical order.  The reference's heap sort is not stable;
 * rows with distinct columns (all valid tile outputs) sort identically. */
int orc_tile2csr_fill(int m, int n, int tilem, const int *tile_ptr,
                      const int *tile_columnidx, const int *tile_nnz,
                      const uint8_t *tile_csr_ptr, const uint8_t *tile_csr_col,
                      const double *tile_csr_val, const int *row_new_to_old,
                      int permutation_size, int canonical,
                      const int64_t *row_ptr, int *col_idx, double *values)
{
    const bool restore = row_new_to_old != nullptr;
    for (int tr = 0; tr < tilem; ++tr) {
        const int rc = tile_row_count(tr, m);
        int64_t local[8] = {};
        for (int t = tile_ptr[tr]; t < tile_ptr[tr + 1]; ++t) {
            const int tnnz = tile_nnz[t + 1] - tile_nnz[t];
            const int base = tile_nnz[t];
            for (int lr = 0; lr < rc; ++lr) {
                const int row = tr * 8 + lr;
                if (restore && row >= permutation_size)
                    continue;
                const int orow = restore ? row_new_to_old[row] : row;
                int s, e;
                row_bounds(tile_csr_ptr, (size_t)t * 8, lr, rc, tnnz, &s, &e);
                for (int x = s; x < e; ++x) {
                    const int lc = tile_csr_col[base + x];
                    const long long col = (long long)tile_columnidx[t] * 8 + lc;
                    if (!(lc < 8 && tile_columnidx[t] >= 0 && col < n))
                        continue;
                    const int64_t o = row_ptr[orow] + local[lr]++;
                    col_idx[o] = (int)col;
                    values[o] = tile_csr_val[base + x];
                }
            }
        }
    }
    if (!canonical) {
        for (int r = 0; r < m; ++r) {
            const int64_t b = row_ptr[r], e = row_ptr[r + 1];
            std::vector<std::pair<int, double>> row;
            for (int64_t x = b; x < e; ++x)
                row.push_back({col_idx[x], values[x]});
            std::stable_sort(row.begin(), row.end(),
                             [](const std::pair<int, double> &a,
                                const std::pair<int, double> &c) {
                                 return a.first < c.first;
                             });
            for (int64_t x = b; x < e; ++x) {
                col_idx[x] = row[x - b].first;
                values[x] = row[x - b].second;
            }
        }
    }
    return 0;
}

/* ---- Tail queue ---------------------------------------------------------- */

/* dmma_tile_tail_queue_chunk_count (src/dmma_tile_tail_queue.h:154-176):
 * FP32 work = scan*scans + match*matches; strictly above the threshold and
 * more than one scan step; chunks = clamp(ceil(work/target), 2, max, scans). */
int orc_tail_chunk_count(int scans, int matches, double admission_threshold,
                         double chunk_target, int max_chunks,
                         double scan_weight, double match_weight)
{
    const float work = (float)scan_weight * (float)scans +
                       (float)match_weight * (float)matches;
    if (!(work > (float)admission_threshold) || scans <= 1)
        return 0;
    const float requested = std::ceil(work / (float)chunk_target);
    int chunks = requested >= (float)max_chunks ? max_chunks : (int)requested;
    if (chunks < 2)
        chunks = 2;
    if (chunks > scans)
        chunks = scans;
    return chunks >= 2 ? chunks : 0;
}

/* dmma_plan_tile_tail_queue (src/dmma_tile_tail_queue.h:190-292) with the
 * default cost model (scan 1, match 8).  Reason codes follow
 * DmmaTileTailQueueGateReason (src/dmma_tile_tail_queue.h:59-70). */
int orc_plan_tail_queue(int n, const int *scans, const int *matches,
                        const int *output_nnz, int enabled,
                        double admission_threshold, double chunk_target,
                        int max_chunks, double max_heavy_fraction,
                        int record_capacity, uint64_t workspace_limit,
                        int workspace_reduction, int64_t *out, int *heavy_ids,
                        int *heavy_chunks)
{
    out[0] = 0;
    out[1] = 0;
    out[2] = 0;
    out[3] = 0;
    if (!enabled)
        return 0;
    const bool valid = std::isfinite(admission_threshold) &&
                       admission_threshold > 0.0 && std::isfinite(chunk_target) &&
                       chunk_target > 0.0 && max_chunks >= 2 && max_chunks <= 32 &&
                       std::isfinite(max_heavy_fraction) &&
                       max_heavy_fraction >= 0.0 && max_heavy_fraction <= 1.0 &&
                       record_capacity > 0;
    if (!valid) {
        out[0] = 2;
        return 0;
    }
    int64_t heavy = 0, chunks = 0;
    for (int i = 0; i < n; ++i) {
        if (scans[i] < 0 || matches[i] < 0 || matches[i] > scans[i] ||
            output_nnz[i] < 0 || output_nnz[i] > 64) {
            out[0] = 3;
            return 0;
        }
        const int c = orc_tail_chunk_count(scans[i], matches[i],
                                           admission_threshold, chunk_target,
                                           max_chunks, 1.0, 8.0);
        if (c == 0)
            continue;
        if (heavy >= record_capacity) {
            out[0] = 5;
            return 0;
        }
        heavy_ids[heavy] = i;
        heavy_chunks[heavy] = c;
        ++heavy;
        chunks += c;
    }
    out[1] = heavy;
    out[2] = chunks;
    out[3] = chunks * 64 * (int64_t)sizeof(double);
    const double frac = n == 0 ? 0.0 : (double)heavy / (double)n;
    if (heavy == 0)
        out[0] = 4;
    else if (frac > max_heavy_fraction)
        out[0] = 6;
    else if (workspace_reduction && (uint64_t)out[3] > workspace_limit)
        out[0] = 7;
    else
        out[0] = 1;
    return 0;
}

} // extern "C"

/* ------------------------------------------------------------------------ */
/* Legacy 16x16 tile SpGEMM on the CPU (src/spgemm-cpu.h:3-455)             */
/* ------------------------------------------------------------------------ */

namespace {

constexpr int kLegacyTile = 16;   /* BLOCK_SIZE, src/common.h:31-32 */

/* A matrix in 16x16 tiles, each tile a small CSR: tile rows of tiles (tile
 * CSR), per tile its first entry, per tile and local row the entry offset
 * (u8), per entry its local column (u8) and value.  With by_column the tiles
 * are also listed per tile column (tile CSC: ids in increasing tile row). */
struct Tiles16 {
    int tm = 0, tn = 0;
    std::vector<int> row_ptr, col_idx;        /* tile CSR */
    std::vector<int> col_ptr, csc_row, csc_id; /* tile CSC */
    std::vector<int> first;                    /* first entry of each tile (+1) */
    std::vector<unsigned char> lrow;           /* 16 per tile: local row offsets */
    std::vector<unsigned char> lcol;
    std::vector<double> val;
};

/* csr2tile for one operand (src/csr2tile.h semantics: entries of a tile in
 * row order, columns ascending; duplicates summed in CSR order). */
Tiles16 make_tiles16(int m, int n, const int *rp, const int *ci, const double *v, bool by_column)
{
    Tiles16 t;
    t.tm = (m + kLegacyTile - 1) / kLegacyTile;
    t.tn = (n + kLegacyTile - 1) / kLegacyTile;
    t.row_ptr.assign(t.tm + 1, 0);
    t.first.push_back(0);
    for (int tr = 0; tr < t.tm; ++tr) {
        const int r0 = tr * kLegacyTile, r1 = std::min(m, r0 + kLegacyTile);
        /* (tile column, local row, local column, value) of the tile row */
        struct E { int tc, r, c; double v; };
        std::vector<E> es;
        for (int r = r0; r < r1; ++r)
            for (int e = rp[r]; e < rp[r + 1]; ++e)
                es.push_back({ci[e] / kLegacyTile, r - r0, ci[e] % kLegacyTile, v ? v[e] : 0.0});
        std::stable_sort(es.begin(), es.end(), [](const E &a, const E &b) {
            if (a.tc != b.tc) return a.tc < b.tc;
            if (a.r != b.r) return a.r < b.r;
            return a.c < b.c;
        });
        size_t i = 0;
        while (i < es.size()) {
            const int tc = es[i].tc;
            t.col_idx.push_back(tc);
            unsigned char cnt[kLegacyTile] = {};
            const size_t base = t.lcol.size();
            while (i < es.size() && es[i].tc == tc) {
                const int r = es[i].r, c = es[i].c;
                double sum = 0.0;
                do {
                    sum += es[i].v;
                    ++i;
                } while (i < es.size() && es[i].tc == tc && es[i].r == r && es[i].c == c);
                t.lcol.push_back(static_cast<unsigned char>(c));
                t.val.push_back(sum);
                ++cnt[r];
            }
            unsigned char off = 0;
            for (int r = 0; r < kLegacyTile; ++r) {
                t.lrow.push_back(off);
                off = static_cast<unsigned char>(off + cnt[r]);
            }
            t.first.push_back(static_cast<int>(t.lcol.size()));
            (void)base;
        }
        t.row_ptr[tr + 1] = static_cast<int>(t.col_idx.size());
    }
    if (by_column) {
        const int nt = static_cast<int>(t.col_idx.size());
        t.col_ptr.assign(t.tn + 1, 0);
        for (int id = 0; id < nt; ++id)
            ++t.col_ptr[t.col_idx[id] + 1];
        for (int c = 0; c < t.tn; ++c)
            t.col_ptr[c + 1] += t.col_ptr[c];
        t.csc_row.resize(nt);
        t.csc_id.resize(nt);
        std::vector<int> fill(t.col_ptr.begin(), t.col_ptr.end() - 1);
        for (int tr = 0; tr < t.tm; ++tr)
            for (int id = t.row_ptr[tr]; id < t.row_ptr[tr + 1]; ++id) {
                const int at = fill[t.col_idx[id]]++;
                t.csc_row[at] = tr;
                t.csc_id[at] = id;
            }
    }
    return t;
}

/* Entries [lo, hi) of local row r of tile id. */
inline void tile_row_range(const Tiles16 &t, int id, int r, int *lo, int *hi)
{
    const int base = t.first[id];
    *lo = base + t.lrow[static_cast<size_t>(id) * kLegacyTile + r];
    *hi = r + 1 < kLegacyTile ? base + t.lrow[static_cast<size_t>(id) * kLegacyTile + r + 1]
                              : t.first[id + 1];
}

} // namespace

extern "C" {

/* The legacy tile SpGEMM (spgemm_cpu, src/spgemm-cpu.h:369-455): C's tiles
 * per tile row are the B tile columns whose tile K list meets A's tile row
 * (step1/step2, a merge of the two sorted K lists per candidate column,
 * src/spgemm-cpu.h:3-92); each C tile's entries come from a 16x16 flag grid
 * (step3, :186-263) and its values from a 16x16 dense accumulator (step4,
 * :266-364), the matched pairs taken in increasing K and each pair's products
 * in A-row then B-row order.  C is returned as CSR (rows ascending, columns
 * ascending: tiles of a tile row in column order).  With ci_c == NULL only
 * rp_c is filled; the return value is nnz(C), or -1 on a shape mismatch. */
int64_t orc_legacy_tile_spgemm(int m_a, int n_a, const int *rp_a, const int *ci_a,
                               const double *v_a, int m_b, int n_b, const int *rp_b,
                               const int *ci_b, const double *v_b, int64_t *rp_c, int *ci_c,
                               double *v_c)
{
    if (n_a != m_b)
        return -1;
    const Tiles16 A = make_tiles16(m_a, n_a, rp_a, ci_a, v_a, false);
    const Tiles16 B = make_tiles16(m_b, n_b, rp_b, ci_b, v_b, true);
    const bool fill = ci_c != nullptr;
    /* Per C row: its entries (column, value), assembled tile by tile. */
    std::vector<std::vector<std::pair<int, double>>> rows(static_cast<size_t>(m_a));
#pragma omp parallel
    {
        double acc[kLegacyTile * kLegacyTile];
        unsigned char flag[kLegacyTile * kLegacyTile];
#pragma omp for schedule(dynamic, 4)
        for (int ti = 0; ti < A.tm; ++ti) {
            const int ra0 = A.row_ptr[ti], ra1 = A.row_ptr[ti + 1];
            const int rowlen = std::min(kLegacyTile, m_a - ti * kLegacyTile);
            for (int tj = 0; tj < B.tn; ++tj) {
                int pa = ra0, pb = B.col_ptr[tj];
                const int pb1 = B.col_ptr[tj + 1];
                bool any = false;
                while (pa < ra1 && pb < pb1) {
                    const int ka = A.col_idx[pa], kb = B.csc_row[pb];
                    if (ka < kb) {
                        ++pa;
                    } else if (ka > kb) {
                        ++pb;
                    } else {
                        if (!any) {
                            std::fill(flag, flag + kLegacyTile * kLegacyTile, 0);
                            std::fill(acc, acc + kLegacyTile * kLegacyTile, 0.0);
                        }
                        any = true;
                        const int ida = pa, idb = B.csc_id[pb];
                        for (int r = 0; r < rowlen; ++r) {
                            int a0, a1;
                            tile_row_range(A, ida, r, &a0, &a1);
                            for (int i = a0; i < a1; ++i) {
                                const int ka_local = A.lcol[i];
                                int b0, b1;
                                tile_row_range(B, idb, ka_local, &b0, &b1);
                                for (int j = b0; j < b1; ++j) {
                                    const int c = B.lcol[j];
                                    acc[r * kLegacyTile + c] += A.val[i] * B.val[j];
                                    flag[r * kLegacyTile + c] = 1;
                                }
                            }
                        }
                        ++pa;
                        ++pb;
                    }
                }
                if (!any)
                    continue;
                for (int r = 0; r < rowlen; ++r)
                    for (int c = 0; c < kLegacyTile; ++c)
                        if (flag[r * kLegacyTile + c])
                            rows[static_cast<size_t>(ti) * kLegacyTile + r].push_back(
                                {tj * kLegacyTile + c, acc[r * kLegacyTile + c]});
            }
        }
    }
    rp_c[0] = 0;
    for (int i = 0; i < m_a; ++i)
        rp_c[i + 1] = rp_c[i] + static_cast<int64_t>(rows[i].size());
    if (fill)
        for (int i = 0; i < m_a; ++i) {
            int64_t o = rp_c[i];
            for (const auto &e : rows[i]) {
                ci_c[o] = e.first;
                if (v_c)
                    v_c[o] = e.second;
                ++o;
            }
        }
    return rp_c[m_a];
}

} // extern "C"
