// Seeded synthetic operands of the block-FEM, banded-batch and AMG
// workloads, the host transpose and the block-diagonal batch layout.  Every
// random draw is a splitmix64 value at a fixed position of a fixed stream,
// so all hosts and thread counts produce identical matrices.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#include <omp.h>

#include "rtt_common.hpp"

using rtt::fail;

namespace {

// Spreads the low 10 bits of v to every third bit (Morton interleave).
uint32_t spread3(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Stream of one block's 16 entries: independent of every other block.
uint64_t block_stream(uint64_t seed, int i, int j)
{
    return rtt::mix_at(seed ^ 0xB10C4B10C4ull, (static_cast<uint64_t>(i) << 32) |
                                                   static_cast<uint32_t>(j));
}

// Exact k nearest neighbours (squared distance, ties by index) of every point,
// self excluded, on a uniform cell grid searched in growing cubes of cells.
void knn_grid(const std::vector<double> &x, int n, int k, std::vector<int> *nb)
{
    const int G = std::max(1, static_cast<int>(std::cbrt(n / 4.0)));
    auto cell_of = [G](double v) { return std::min(G - 1, static_cast<int>(v * G)); };
    std::vector<int> cell(n), start(static_cast<size_t>(G) * G * G + 1, 0), items(n);
    for (int i = 0; i < n; ++i) {
        cell[i] = (cell_of(x[3 * i]) * G + cell_of(x[3 * i + 1])) * G + cell_of(x[3 * i + 2]);
        ++start[cell[i] + 1];
    }
    for (size_t c = 0; c + 1 < start.size(); ++c)
        start[c + 1] += start[c];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < n; ++i)
            items[fill[cell[i]]++] = i;
    }
    nb->assign(static_cast<size_t>(n) * k, -1);
    const int kk = std::min(k, n - 1);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < n; ++i) {
        const double px = x[3 * i], py = x[3 * i + 1], pz = x[3 * i + 2];
        const int cx = cell_of(px), cy = cell_of(py), cz = cell_of(pz);
        std::vector<std::pair<double, int>> best;   // sorted ascending, at most kk
        for (int r = 0;; ++r) {
            // Cells on the surface of the cube of radius r around (cx, cy, cz).
            for (int dx = -r; dx <= r; ++dx) {
                const int gx = cx + dx;
                if (gx < 0 || gx >= G)
                    continue;
                for (int dy = -r; dy <= r; ++dy) {
                    const int gy = cy + dy;
                    if (gy < 0 || gy >= G)
                        continue;
                    for (int dz = -r; dz <= r; ++dz) {
                        const int gz = cz + dz;
                        if (gz < 0 || gz >= G)
                            continue;
                        if (std::max({std::abs(dx), std::abs(dy), std::abs(dz)}) != r)
                            continue;
                        const int c = (gx * G + gy) * G + gz;
                        for (int s = start[c]; s < start[c + 1]; ++s) {
                            const int j = items[s];
                            if (j == i)
                                continue;
                            const double ddx = x[3 * j] - px, ddy = x[3 * j + 1] - py,
                                         ddz = x[3 * j + 2] - pz;
                            const std::pair<double, int> cand(ddx * ddx + ddy * ddy + ddz * ddz, j);
                            if (static_cast<int>(best.size()) == kk && !(cand < best.back()))
                                continue;
                            best.insert(std::upper_bound(best.begin(), best.end(), cand), cand);
                            if (static_cast<int>(best.size()) > kk)
                                best.pop_back();
                        }
                    }
                }
            }
            // Distance from the point to the nearest face of the searched cube
            // that has cells beyond it: nothing unsearched is closer.
            double margin = 1e300;
            const int lo[3] = {cx - r, cy - r, cz - r}, hi[3] = {cx + r, cy + r, cz + r};
            const double p[3] = {px, py, pz};
            bool all_covered = true;
            for (int d = 0; d < 3; ++d) {
                if (lo[d] > 0) {
                    margin = std::min(margin, p[d] - static_cast<double>(lo[d]) / G);
                    all_covered = false;
                }
                if (hi[d] < G - 1) {
                    margin = std::min(margin, static_cast<double>(hi[d] + 1) / G - p[d]);
                    all_covered = false;
                }
            }
            if (all_covered)
                break;
            if (static_cast<int>(best.size()) == kk && best.back().first <= margin * margin)
                break;
        }
        for (int t = 0; t < static_cast<int>(best.size()); ++t)
            (*nb)[static_cast<size_t>(i) * k + t] = best[t].second;
    }
}

double normal_at(uint64_t seed, uint64_t i)
{
    // Box-Muller on two independent uniforms.
    const double u1 = 1.0 - rtt::uniform_at(seed, 2 * i);
    const double u2 = rtt::uniform_at(seed, 2 * i + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

} // namespace

extern "C" {

int rtt_gen_fem_blocks(int nodes, int neighbours, int spatial_order, uint64_t seed,
                       rtt_smatrix *out)
{
    rtt::clear_error();
    if (nodes <= 0 || neighbours < 0 || !out || nodes > INT_MAX / 4)
        return fail(RTT_ERR_INVALID, "rtt_gen_fem_blocks: bad argument");
    const int n = nodes;
    std::vector<double> x(static_cast<size_t>(3) * n);
    for (int i = 0; i < 3 * n; ++i)
        x[i] = rtt::uniform_at(seed, static_cast<uint64_t>(i));
    if (spatial_order) {
        // Number the nodes along a Morton curve of a 1024^3 lattice (ties by
        // draw order), as a mesher numbering nodes with locality would.
        std::vector<std::pair<uint32_t, int>> key(n);
        for (int i = 0; i < n; ++i) {
            auto q = [](double v) { return std::min(1023u, static_cast<uint32_t>(v * 1024.0)); };
            key[i] = {spread3(q(x[3 * i])) | (spread3(q(x[3 * i + 1])) << 1) |
                          (spread3(q(x[3 * i + 2])) << 2),
                      i};
        }
        std::sort(key.begin(), key.end());
        std::vector<double> y(x.size());
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d)
                y[3 * i + d] = x[3 * key[i].second + d];
        x.swap(y);
    }
    std::vector<int> nb;
    knn_grid(x, n, neighbours, &nb);
    // Symmetrised block pattern with the diagonal: i ~ j when either is among
    // the other's nearest neighbours.
    std::vector<int> deg(n + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int t = 0; t < neighbours; ++t) {
            const int j = nb[static_cast<size_t>(i) * neighbours + t];
            if (j >= 0) {
                ++deg[i + 1];
                ++deg[j + 1];
            }
        }
    for (int i = 0; i < n; ++i)
        deg[i + 1] += deg[i] + 0;
    std::vector<int> adj(static_cast<size_t>(deg[n]) + n);
    std::vector<int64_t> fill(n);
    for (int i = 0; i < n; ++i)
        fill[i] = deg[i] + i;   // room for the diagonal of every row before i
    std::vector<int64_t> rstart(n + 1);
    for (int i = 0; i <= n; ++i)
        rstart[i] = deg[i] + i;
    for (int i = 0; i < n; ++i)
        adj[fill[i]++] = i;
    for (int i = 0; i < n; ++i)
        for (int t = 0; t < neighbours; ++t) {
            const int j = nb[static_cast<size_t>(i) * neighbours + t];
            if (j >= 0) {
                adj[fill[i]++] = j;
                adj[fill[j]++] = i;
            }
        }
    std::vector<int> bdeg(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        int *b = adj.data() + rstart[i], *e = adj.data() + fill[i];
        std::sort(b, e);
        bdeg[i] = static_cast<int>(std::unique(b, e) - b);
    }
    std::vector<int64_t> rowblk(n + 1, 0);
    for (int i = 0; i < n; ++i)
        rowblk[i + 1] = rowblk[i] + bdeg[i];
    const int64_t nnz = rowblk[n] * 16;
    if (nnz > INT_MAX)
        return fail(RTT_ERR_OVERFLOW, "rtt_gen_fem_blocks: nnz %lld beyond int32",
                    (long long)nnz);
    int rc = rtt::alloc_host_csr(4 * n, 4 * n, nnz, out);
    if (rc != RTT_OK)
        return rc;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const int d = bdeg[i];
        const int *cols = adj.data() + rstart[i];
        for (int r = 0; r < 4; ++r) {
            const int64_t row = 4LL * i + r;
            const int64_t base = rowblk[i] * 16 + static_cast<int64_t>(r) * 4 * d;
            out->rowpointer[row] = static_cast<int>(base);
            for (int b = 0; b < d; ++b) {
                const uint64_t st = block_stream(seed, i, cols[b]);
                for (int c = 0; c < 4; ++c) {
                    const int64_t e = base + 4LL * b + c;
                    const uint64_t h = rtt::mix_at(st, static_cast<uint64_t>(r * 4 + c));
                    const double u = (h >> 11) * (1.0 / 9007199254740992.0);
                    const double mag = std::pow(10.0, -3.0 + 6.0 * u);
                    out->columnindex[e] = 4 * cols[b] + c;
                    out->value[e] = (h & 1u) ? -mag : mag;
                }
            }
        }
    }
    out->rowpointer[4LL * n] = static_cast<int>(nnz);
    return RTT_OK;
}

int rtt_gen_banded(int n, int bandwidth, int per_row, uint64_t seed, rtt_smatrix *out)
{
    rtt::clear_error();
    if (n < 0 || bandwidth < 0 || per_row < 0 || !out)
        return fail(RTT_ERR_INVALID, "rtt_gen_banded: bad argument");
    std::vector<std::vector<int>> cols(n);
    std::vector<int64_t> len(n + 1, 0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const int lo = std::max(0, i - bandwidth), hi = std::min(n - 1, i + bandwidth);
        const int width = hi - lo + 1;
        const int k = std::min(per_row, width);
        rtt::SplitMix64 g(seed ^ (0xBA4DEDull * (static_cast<uint64_t>(i) + 1)));
        std::vector<int> &c = cols[i];
        // k distinct draws from the band (rejection), plus the diagonal.
        while (static_cast<int>(c.size()) < k) {
            const int v = lo + static_cast<int>(g.next() % static_cast<uint64_t>(width));
            if (std::find(c.begin(), c.end(), v) == c.end())
                c.push_back(v);
        }
        if (std::find(c.begin(), c.end(), i) == c.end())
            c.push_back(i);
        std::sort(c.begin(), c.end());
        len[i + 1] = static_cast<int64_t>(c.size());
    }
    for (int i = 0; i < n; ++i)
        len[i + 1] += len[i];
    int rc = rtt::alloc_host_csr(n, n, len[n], out);
    if (rc != RTT_OK)
        return rc;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        out->rowpointer[i] = static_cast<int>(len[i]);
        for (size_t t = 0; t < cols[i].size(); ++t) {
            const int64_t e = len[i] + static_cast<int64_t>(t);
            out->columnindex[e] = cols[i][t];
            out->value[e] = 2.0 * rtt::uniform_at(seed ^ 0x5EEDBA4Dull, static_cast<uint64_t>(e)) - 1.0;
        }
    }
    out->rowpointer[n] = static_cast<int>(len[n]);
    return RTT_OK;
}

int rtt_gen_banded_batch_shapes(int problems, double median_rows, double sigma, int min_rows,
                                int max_rows, int min_bw, int max_bw, int min_per_row,
                                int max_per_row, uint64_t seed, int *rows, int *bandwidth,
                                int *per_row)
{
    rtt::clear_error();
    if (problems < 0 || !(median_rows > 0) || !(sigma >= 0) || min_rows < 1 ||
        max_rows < min_rows || min_bw < 0 || max_bw < min_bw || min_per_row < 0 ||
        max_per_row < min_per_row || (problems > 0 && (!rows || !bandwidth || !per_row)))
        return fail(RTT_ERR_INVALID, "rtt_gen_banded_batch_shapes: bad argument");
    for (int p = 0; p < problems; ++p) {
        const double r = std::round(std::exp(std::log(median_rows) + sigma * normal_at(seed, p)));
        rows[p] = static_cast<int>(std::min<double>(max_rows, std::max<double>(min_rows, r)));
        const uint64_t h = rtt::mix_at(seed ^ 0xBA5Eull, static_cast<uint64_t>(p));
        bandwidth[p] = min_bw + static_cast<int>((h >> 32) % static_cast<uint64_t>(max_bw - min_bw + 1));
        per_row[p] = min_per_row + static_cast<int>((h & 0xFFFFFFFFu) %
                                                    static_cast<uint64_t>(max_per_row - min_per_row + 1));
    }
    return RTT_OK;
}

int rtt_gen_poisson7_jumps(int g, double decades, uint64_t seed, rtt_smatrix *out)
{
    rtt::clear_error();
    if (g <= 0 || !(decades >= 0) || !out)
        return fail(RTT_ERR_INVALID, "rtt_gen_poisson7_jumps: bad argument");
    const int64_t n = static_cast<int64_t>(g) * g * g;
    if (n > INT_MAX / 8)
        return fail(RTT_ERR_OVERFLOW, "rtt_gen_poisson7_jumps: grid too large");
    std::vector<double> kc(n);
    for (int64_t i = 0; i < n; ++i)
        kc[i] = std::pow(10.0, decades * (2.0 * rtt::uniform_at(seed, static_cast<uint64_t>(i)) - 1.0));
    auto span = [g](int v) { return (v > 0 ? 1 : 0) + (v < g - 1 ? 1 : 0); };
    std::vector<int64_t> len(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int x = static_cast<int>(i % g), y = static_cast<int>((i / g) % g),
                  z = static_cast<int>(i / (static_cast<int64_t>(g) * g));
        len[i + 1] = len[i] + 1 + span(x) + span(y) + span(z);
    }
    int rc = rtt::alloc_host_csr(static_cast<int>(n), static_cast<int>(n), len[n], out);
    if (rc != RTT_OK)
        return rc;
    const int64_t gg = static_cast<int64_t>(g) * g;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const int c[3] = {static_cast<int>(i % g), static_cast<int>((i / g) % g),
                          static_cast<int>(i / gg)};
        const int64_t stride[3] = {1, g, gg};
        // Faces in column order: -z, -y, -x, diagonal, +x, +y, +z.  The face
        // weight is the mean of the two cells' coefficients; a boundary face
        // adds the cell's own (Dirichlet) to the diagonal.
        double diag = 0.0;
        int64_t e = len[i];
        out->rowpointer[i] = static_cast<int>(e);
        int64_t diag_slot = -1;
        for (int pass = 0; pass < 7; ++pass) {
            if (pass == 3) {
                diag_slot = e;
                out->columnindex[e++] = static_cast<int>(i);
                continue;
            }
            const int axis = pass < 3 ? 2 - pass : pass - 4;
            const int sgn = pass < 3 ? -1 : 1;
            const int v = c[axis] + sgn;
            if (v < 0 || v >= g) {
                diag += kc[i];
                continue;
            }
            const int64_t j = i + sgn * stride[axis];
            const double w = 0.5 * (kc[i] + kc[j]);
            out->columnindex[e] = static_cast<int>(j);
            out->value[e++] = -w;
            diag += w;
        }
        out->value[diag_slot] = diag;
    }
    out->rowpointer[n] = static_cast<int>(len[n]);
    return RTT_OK;
}

int rtt_gen_aggregation_prolongator(const rtt_smatrix *a, int g, double omega, rtt_smatrix *out)
{
    rtt::clear_error();
    if (!a || !out || g <= 0 || static_cast<int64_t>(g) * g * g != a->m || a->m != a->n ||
        !a->rowpointer || (a->nnz > 0 && (!a->columnindex || !a->value)))
        return fail(RTT_ERR_INVALID, "rtt_gen_aggregation_prolongator: bad argument");
    const int h = (g + 1) / 2;
    const int n = a->m;
    const int64_t gg = static_cast<int64_t>(g) * g;
    auto agg = [&](int64_t i) {
        const int x = static_cast<int>(i % g), y = static_cast<int>((i / g) % g),
                  z = static_cast<int>(i / gg);
        return static_cast<int>((static_cast<int64_t>(z / 2) * h + y / 2) * h + x / 2);
    };
    // P = (I - omega D^-1 A) T with T the 2x2x2 aggregation: row i holds, for
    // every aggregate its stencil touches, T(i,.) - omega/a_ii * sum of A(i,j)
    // over the j of that aggregate (summed in A's column order).
    std::vector<std::vector<std::pair<int, double>>> rows(n);
    std::vector<int64_t> len(static_cast<size_t>(n) + 1, 0);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int i = 0; i < n; ++i) {
        double aii = 0.0;
        for (int e = a->rowpointer[i]; e < a->rowpointer[i + 1]; ++e)
            if (a->columnindex[e] == i)
                aii += a->value[e];
        if (aii == 0.0) {
            bad = 1;
            continue;
        }
        std::vector<std::pair<int, double>> &r = rows[i];
        auto add = [&r](int col, double v) {
            for (auto &p : r)
                if (p.first == col) {
                    p.second += v;
                    return;
                }
            r.emplace_back(col, v);
        };
        add(agg(i), 1.0);
        for (int e = a->rowpointer[i]; e < a->rowpointer[i + 1]; ++e)
            add(agg(a->columnindex[e]), -omega * a->value[e] / aii);
        std::sort(r.begin(), r.end());
        len[static_cast<size_t>(i) + 1] = static_cast<int64_t>(r.size());
    }
    if (bad)
        return fail(RTT_ERR_INVALID, "rtt_gen_aggregation_prolongator: zero diagonal");
    for (int i = 0; i < n; ++i)
        len[i + 1] += len[i];
    int rc = rtt::alloc_host_csr(n, h * h * h, len[n], out);
    if (rc != RTT_OK)
        return rc;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        out->rowpointer[i] = static_cast<int>(len[i]);
        for (size_t t = 0; t < rows[i].size(); ++t) {
            out->columnindex[len[i] + t] = rows[i][t].first;
            out->value[len[i] + t] = rows[i][t].second;
        }
    }
    out->rowpointer[n] = static_cast<int>(len[n]);
    return RTT_OK;
}

int rtt_csr_transpose(const rtt_smatrix *in, rtt_smatrix *out)
{
    rtt::clear_error();
    if (!in || !out || in->m < 0 || in->n < 0 || !in->rowpointer ||
        (in->nnz > 0 && (!in->columnindex || !in->value)))
        return fail(RTT_ERR_INVALID, "rtt_csr_transpose: bad argument");
    rtt_smatrix t;
    int rc = rtt::alloc_host_csr(in->n, in->m, in->nnz, &t);
    if (rc != RTT_OK)
        return rc;
    std::vector<int64_t> cnt(static_cast<size_t>(in->n) + 1, 0);
    for (int e = 0; e < in->nnz; ++e) {
        const int c = in->columnindex[e];
        if (c < 0 || c >= in->n) {
            rtt_destroy_smatrix(&t);
            return fail(RTT_ERR_INVALID, "rtt_csr_transpose: column %d out of range", c);
        }
        ++cnt[static_cast<size_t>(c) + 1];
    }
    for (int c = 0; c < in->n; ++c)
        cnt[c + 1] += cnt[c];
    for (int c = 0; c <= in->n; ++c)
        t.rowpointer[c] = static_cast<int>(cnt[c]);
    // Rows in increasing order within each column: a stable counting sort.
    for (int r = 0; r < in->m; ++r)
        for (int e = in->rowpointer[r]; e < in->rowpointer[r + 1]; ++e) {
            const int64_t d = cnt[in->columnindex[e]]++;
            t.columnindex[d] = r;
            t.value[d] = in->value[e];
        }
    rtt_destroy_smatrix(out);
    *out = t;
    return RTT_OK;
}

int rtt_block_diag(int count, const rtt_smatrix *const *mats, int align, rtt_smatrix *out,
                   int64_t *row_offsets, int64_t *col_offsets)
{
    rtt::clear_error();
    if (count < 0 || !out || align < 1 || (count > 0 && !mats))
        return fail(RTT_ERR_INVALID, "rtt_block_diag: bad argument");
    int64_t m = 0, n = 0, nnz = 0;
    auto up = [align](int64_t v) { return (v + align - 1) / align * align; };
    std::vector<int64_t> ro(static_cast<size_t>(count) + 1), co(static_cast<size_t>(count) + 1);
    for (int p = 0; p < count; ++p) {
        const rtt_smatrix *q = mats[p];
        if (!q || q->m < 0 || q->n < 0 || !q->rowpointer || q->rowpointer[0] != 0 ||
            q->rowpointer[q->m] != q->nnz || (q->nnz > 0 && (!q->columnindex || !q->value)))
            return fail(RTT_ERR_INVALID, "rtt_block_diag: matrix %d is malformed", p);
        ro[p] = m;
        co[p] = n;
        m += up(q->m);
        n += up(q->n);
        nnz += q->nnz;
    }
    ro[count] = m;
    co[count] = n;
    if (m > INT_MAX || n > INT_MAX || nnz > INT_MAX)
        return fail(RTT_ERR_OVERFLOW, "rtt_block_diag: %lld x %lld with %lld entries beyond int32",
                    (long long)m, (long long)n, (long long)nnz);
    rtt_smatrix r;
    int rc = rtt::alloc_host_csr(static_cast<int>(m), static_cast<int>(n), nnz, &r);
    if (rc != RTT_OK)
        return rc;
    std::vector<int64_t> ebase(static_cast<size_t>(count) + 1, 0);
    for (int p = 0; p < count; ++p)
        ebase[p + 1] = ebase[p] + mats[p]->nnz;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < count; ++p) {
        const rtt_smatrix *q = mats[p];
        const int64_t rend = p + 1 < count ? ro[p + 1] : m;
        for (int64_t i = 0; i < rend - ro[p]; ++i)
            r.rowpointer[ro[p] + i] =
                static_cast<int>(ebase[p] + (i < q->m ? q->rowpointer[i] : q->nnz));
        for (int e = 0; e < q->nnz; ++e) {
            r.columnindex[ebase[p] + e] = static_cast<int>(co[p] + q->columnindex[e]);
            r.value[ebase[p] + e] = q->value[e];
        }
    }
    r.rowpointer[m] = static_cast<int>(nnz);
    if (row_offsets)
        std::memcpy(row_offsets, ro.data(), (static_cast<size_t>(count) + 1) * sizeof(int64_t));
    if (col_offsets)
        std::memcpy(col_offsets, co.data(), (static_cast<size_t>(count) + 1) * sizeof(int64_t));
    rtt_destroy_smatrix(out);
    *out = r;
    return RTT_OK;
}

} // extern "C"
