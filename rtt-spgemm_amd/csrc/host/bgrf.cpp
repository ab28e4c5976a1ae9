// BGRF-v1 offline reorder of A: one row permutation and one inner (K)
// permutation that lower the number of occupied 8x4 A tiles.  A is prepared
// once, outside the timed region, so this runs on the host (OpenMP over
// windows and tile rows); its output feeds rtt_prepare_external_a.
//
// Stages, as the reference's design document states them
// (docs/BGRF_V1_PPOPP_ALGORITHM_AND_DATA_FORMAT_CN.md sections 3-6) and
// build_dmma_reorder_plan implements them (src/dmma_reorder.h:2651-3085):
//  I   coarse: a Cuthill-McKee traversal of the row-K bipartite graph with
//      neighbours ordered by (degree bucket, id) and seeds by (degree, id),
//      at most 2048 frontier rounds over the whole graph, then reversed and
//      projected onto both axes; accepted only if the exact A tile count
//      drops (both axes together);
//  II  rows: inside each current window of 32 rows, greedy packing into
//      groups of 8 by panel fingerprints; a window's proposal is committed
//      when its exact tile count drops, or stays equal while the summed
//      per-group panel span (the fanout proxy) drops;
//  III K: the same inside windows of 16 K with groups of 4, over row
//      panels of the final row order.
// Degrees count CSR entries (duplicates included); adjacency, panels and
// tile counts use each row's distinct columns.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "../rtt_common.hpp"

namespace {

constexpr int kBuckets = 8;
constexpr int kLevelBudget = 2048;
constexpr int kRowWindow = 32, kRowGroup = 8;
constexpr int kInnerWindow = 16, kInnerGroup = 4;
constexpr int kTileM = 8, kTileK = 4;

int log2_floor(unsigned v)
{
    return 31 - __builtin_clz(v);
}

// Coarse adjacency bucket: 0 for degree <= 1, else floor(log2 d) capped at 7.
int degree_bucket(int d)
{
    return d <= 1 ? 0 : std::min(log2_floor(static_cast<unsigned>(d)), kBuckets - 1);
}

// Fine-stage bucket: -1 for an empty axis element, else floor(log2 d) / 2.
int soft_bucket(int d)
{
    return d <= 0 ? -1 : log2_floor(static_cast<unsigned>(d)) / 2;
}

// SplitMix64 output function (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014).
uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Two 64-bit words, one hashed bit per panel in each: a set sketch whose
// popcount approximates the number of distinct panels.
struct Fingerprint {
    uint64_t w[2] = {0, 0};
    void add(int panel)
    {
        for (int k = 0; k < 2; ++k)
            w[k] |= 1ull << (splitmix64(2ull * static_cast<uint32_t>(panel) + k) & 63);
    }
    int bits() const { return __builtin_popcountll(w[0]) + __builtin_popcountll(w[1]); }
};

struct Profile {
    int id = 0, degree = 0, bucket = -1, occupancy = 0;
    int pmin = -1, pmax = -1, span = 0;
    Fingerprint fp;
};

// Distinct, sorted columns per row (the structure the tiles see).
struct Pattern {
    int rows = 0, cols = 0;
    std::vector<int64_t> rp;
    std::vector<int> ci;
    std::vector<int> row_degree, col_degree;   // CSR entry counts
};

Pattern make_pattern(const rtt_smatrix &a)
{
    Pattern p;
    p.rows = a.m;
    p.cols = a.n;
    p.rp.assign(static_cast<size_t>(a.m) + 1, 0);
    p.row_degree.assign(a.m, 0);
    p.col_degree.assign(a.n, 0);
    std::vector<std::vector<int>> per_row(a.m);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < a.m; ++i) {
        std::vector<int> &r = per_row[i];
        r.assign(a.columnindex + a.rowpointer[i], a.columnindex + a.rowpointer[i + 1]);
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        p.row_degree[i] = a.rowpointer[i + 1] - a.rowpointer[i];
    }
    for (int64_t e = 0; e < a.nnz; ++e)
        ++p.col_degree[a.columnindex[e]];
    for (int i = 0; i < a.m; ++i)
        p.rp[i + 1] = p.rp[i] + static_cast<int64_t>(per_row[i].size());
    p.ci.resize(static_cast<size_t>(p.rp[a.m]));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < a.m; ++i)
        std::copy(per_row[i].begin(), per_row[i].end(), p.ci.begin() + p.rp[i]);
    return p;
}

// Occupied 8x4 tiles under (row old->new, inner old->new).
int64_t count_tiles(const Pattern &p, const std::vector<int> &row_new_to_old,
                    const std::vector<int> &inner_old_to_new)
{
    const int tile_rows = (p.rows + kTileM - 1) / kTileM;
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
    for (int t = 0; t < tile_rows; ++t) {
        std::vector<int> ks;
        for (int r = t * kTileM; r < std::min(p.rows, (t + 1) * kTileM); ++r) {
            const int old = row_new_to_old[r];
            for (int64_t e = p.rp[old]; e < p.rp[old + 1]; ++e)
                ks.push_back(inner_old_to_new[p.ci[e]] / kTileK);
        }
        std::sort(ks.begin(), ks.end());
        total += std::unique(ks.begin(), ks.end()) - ks.begin();
    }
    return total;
}

std::vector<int> invert(const std::vector<int> &perm)
{
    std::vector<int> inv(perm.size());
    for (size_t i = 0; i < perm.size(); ++i)
        inv[perm[i]] = static_cast<int>(i);
    return inv;
}

// ---- stage I: bounded bucketed Cuthill-McKee, reversed --------------------
// Vertices 0..m-1 are rows, m..m+k-1 are K indices.  Returns the number of
// frontier rounds processed.
int coarse_proposal(const Pattern &p, std::vector<int> *row_new_to_old,
                    std::vector<int> *inner_new_to_old, int *components)
{
    const int m = p.rows, k = p.cols;
    const int64_t nv = static_cast<int64_t>(m) + k;
    std::vector<int> degree(static_cast<size_t>(nv));
    for (int i = 0; i < m; ++i)
        degree[i] = p.row_degree[i];
    for (int j = 0; j < k; ++j)
        degree[m + j] = p.col_degree[j];
    // Adjacency: row -> K (distinct columns), K -> rows; each list ordered by
    // (degree bucket of the neighbour, neighbour id).
    std::vector<int64_t> adj_ptr(static_cast<size_t>(nv) + 1, 0);
    std::vector<int64_t> col_count(k + 1, 0);
    for (int64_t e = 0; e < p.rp[m]; ++e)
        ++col_count[p.ci[e] + 1];
    for (int j = 0; j < k; ++j)
        col_count[j + 1] += col_count[j];
    for (int i = 0; i < m; ++i)
        adj_ptr[i + 1] = p.rp[i + 1];
    for (int j = 0; j < k; ++j)
        adj_ptr[m + j + 1] = p.rp[m] + col_count[j + 1];
    std::vector<int> adj(static_cast<size_t>(adj_ptr[nv]));
    for (int64_t e = 0; e < p.rp[m]; ++e)
        adj[e] = m + p.ci[e];
    {
        std::vector<int64_t> fill(col_count.begin(), col_count.end() - 1);
        for (int i = 0; i < m; ++i)
            for (int64_t e = p.rp[i]; e < p.rp[i + 1]; ++e)
                adj[p.rp[m] + fill[p.ci[e]]++] = i;   // rows ascending per column
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < nv; ++v)
        std::stable_sort(adj.begin() + adj_ptr[v], adj.begin() + adj_ptr[v + 1],
                         [&](int x, int y) { return degree_bucket(degree[x]) < degree_bucket(degree[y]); });
    // Seeds: vertices with entries, by (degree, id).
    std::vector<int> seeds;
    seeds.reserve(static_cast<size_t>(nv));
    for (int64_t v = 0; v < nv; ++v)
        if (degree[v] > 0)
            seeds.push_back(static_cast<int>(v));
    std::stable_sort(seeds.begin(), seeds.end(),
                     [&](int x, int y) { return degree[x] < degree[y]; });
    std::vector<char> visited(static_cast<size_t>(nv), 0);
    std::vector<int> order;
    order.reserve(seeds.size());
    size_t seed_pos = 0;
    int levels = 0;
    std::vector<int> frontier, next;
    while (levels < kLevelBudget) {
        while (seed_pos < seeds.size() && visited[seeds[seed_pos]])
            ++seed_pos;
        if (seed_pos == seeds.size())
            break;
        const int s = seeds[seed_pos];
        ++*components;
        visited[s] = 1;
        order.push_back(s);
        frontier.assign(1, s);
        while (!frontier.empty() && levels < kLevelBudget) {
            // Parents in frontier order claim their unvisited neighbours in
            // adjacency order: the earliest parent owns a shared neighbour.
            next.clear();
            for (int u : frontier)
                for (int64_t e = adj_ptr[u]; e < adj_ptr[u + 1]; ++e) {
                    const int v = adj[e];
                    if (!visited[v]) {
                        visited[v] = 1;
                        next.push_back(v);
                    }
                }
            order.insert(order.end(), next.begin(), next.end());
            frontier.swap(next);
            ++levels;
        }
    }
    // Vertices the budget left unvisited follow in reverse seed order; then
    // the whole sequence is reversed.
    for (size_t i = seeds.size(); i-- > 0;)
        if (!visited[seeds[i]])
            order.push_back(seeds[i]);
    std::reverse(order.begin(), order.end());
    row_new_to_old->clear();
    inner_new_to_old->clear();
    for (int v : order) {
        if (v < m)
            row_new_to_old->push_back(v);
        else
            inner_new_to_old->push_back(v - m);
    }
    // Empty rows and K indices after the active ones, by id.
    for (int i = 0; i < m; ++i)
        if (degree[i] == 0)
            row_new_to_old->push_back(i);
    for (int j = 0; j < k; ++j)
        if (degree[m + j] == 0)
            inner_new_to_old->push_back(j);
    return levels;
}

// ---- stages II and III: window packing and exact window commit ----------
// Anchor order: nonempty first, larger bucket, larger occupancy, smaller
// panel minimum, smaller span, larger degree, smaller fingerprint, smaller id.
bool anchor_before(const Profile &a, const Profile &b)
{
    const bool na = a.degree > 0, nb = b.degree > 0;
    if (na != nb)
        return na;
    if (a.bucket != b.bucket)
        return a.bucket > b.bucket;
    if (a.occupancy != b.occupancy)
        return a.occupancy > b.occupancy;
    if (a.pmin != b.pmin)
        return a.pmin < b.pmin;
    if (a.span != b.span)
        return a.span < b.span;
    if (a.degree != b.degree)
        return a.degree > b.degree;
    if (a.fp.w[0] != b.fp.w[0])
        return a.fp.w[0] < b.fp.w[0];
    if (a.fp.w[1] != b.fp.w[1])
        return a.fp.w[1] < b.fp.w[1];
    return a.id < b.id;
}

struct Group {
    Fingerprint fp;
    int pmin = -1, pmax = -1;
    void add(const Profile &x)
    {
        fp.w[0] |= x.fp.w[0];
        fp.w[1] |= x.fp.w[1];
        if (x.degree > 0) {
            pmin = pmin < 0 ? x.pmin : std::min(pmin, x.pmin);
            pmax = std::max(pmax, x.pmax);
        }
    }
    int span() const { return pmax < 0 ? 0 : pmax - pmin + 1; }
};

// Candidate score against anchor a and the group so far, compared
// lexicographically, smaller first: bucket distance, new fingerprint bits,
// larger overlap/union (compared by cross-multiplication), larger overlap,
// span growth, degree distance, id.
struct Score {
    int bucket_dist, new_bits, overlap, uni, span_growth, degree_dist, id;
    bool operator<(const Score &o) const
    {
        if (bucket_dist != o.bucket_dist)
            return bucket_dist < o.bucket_dist;
        if (new_bits != o.new_bits)
            return new_bits < o.new_bits;
        const int64_t lhs = static_cast<int64_t>(overlap) * o.uni;
        const int64_t rhs = static_cast<int64_t>(o.overlap) * uni;
        if (lhs != rhs)
            return lhs > rhs;
        if (overlap != o.overlap)
            return overlap > o.overlap;
        if (span_growth != o.span_growth)
            return span_growth < o.span_growth;
        if (degree_dist != o.degree_dist)
            return degree_dist < o.degree_dist;
        return id < o.id;
    }
};

Score score(const Profile &a, const Group &g, const Profile &x)
{
    Score s;
    s.bucket_dist = std::abs(a.bucket - x.bucket);
    s.new_bits = __builtin_popcountll(x.fp.w[0] & ~g.fp.w[0]) +
                 __builtin_popcountll(x.fp.w[1] & ~g.fp.w[1]);
    s.overlap = __builtin_popcountll(x.fp.w[0] & g.fp.w[0]) +
                __builtin_popcountll(x.fp.w[1] & g.fp.w[1]);
    s.uni = __builtin_popcountll(x.fp.w[0] | g.fp.w[0]) +
            __builtin_popcountll(x.fp.w[1] | g.fp.w[1]);
    Group h = g;
    h.add(x);
    s.span_growth = h.span() - g.span();
    s.degree_dist = std::abs(a.degree - x.degree);
    s.id = x.id;
    return s;
}

// Greedy packing of one window (positions [w0, w1) of `cur`, new -> old):
// writes the proposed new -> old order of those positions into `out`.
void pack_window(const std::vector<Profile> &prof, const int *cur, int n, int group,
                 int *out)
{
    std::vector<int> unused(cur, cur + n);
    int pos = 0;
    while (!unused.empty()) {
        size_t best = 0;
        for (size_t i = 1; i < unused.size(); ++i)
            if (anchor_before(prof[unused[i]], prof[unused[best]]))
                best = i;
        const Profile &a = prof[unused[best]];
        Group g;
        g.add(a);
        out[pos++] = a.id;
        unused.erase(unused.begin() + static_cast<long>(best));
        for (int size = 1; size < group && !unused.empty(); ++size) {
            size_t pick = 0;
            Score ps = score(a, g, prof[unused[0]]);
            for (size_t i = 1; i < unused.size(); ++i) {
                const Score si = score(a, g, prof[unused[i]]);
                if (si < ps) {
                    ps = si;
                    pick = i;
                }
            }
            g.add(prof[unused[pick]]);
            out[pos++] = unused[pick];
            unused.erase(unused.begin() + static_cast<long>(pick));
        }
    }
}

// Exact tiles and fanout of one window's order: distinct (group, opposite
// tile) pairs, and the summed per-group span of opposite tiles.
void window_cost(const int *order, int n, int group, const std::vector<std::vector<int>> &tiles_of,
                 int64_t *tiles, int64_t *fanout)
{
    int64_t t = 0, f = 0;
    std::vector<int> ks;
    for (int g0 = 0; g0 < n; g0 += group) {
        ks.clear();
        for (int i = g0; i < std::min(n, g0 + group); ++i) {
            const std::vector<int> &v = tiles_of[order[i]];
            ks.insert(ks.end(), v.begin(), v.end());
        }
        if (ks.empty())
            continue;
        std::sort(ks.begin(), ks.end());
        t += std::unique(ks.begin(), ks.end()) - ks.begin();
        f += ks.back() - ks.front() + 1;
    }
    *tiles = t;
    *fanout = f;
}

// One fine stage over the axis whose new -> old order is `order` (size n),
// with `panels_of[x]` = the opposite-axis panels (distinct, sorted) element
// x touches and `degree[x]` its CSR entry count.  Returns windows accepted.
int fine_stage(std::vector<int> *order, const std::vector<std::vector<int>> &panels_of,
               const std::vector<int> &degree, int window, int group, int *windows,
               int64_t *fan_before, int64_t *fan_after)
{
    const int n = static_cast<int>(order->size());
    std::vector<Profile> prof(n);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int x = 0; x < n; ++x) {
        Profile &p = prof[x];
        p.id = x;
        p.degree = degree[x];
        p.bucket = soft_bucket(p.degree);
        const std::vector<int> &pv = panels_of[x];
        for (int q : pv)
            p.fp.add(q);
        p.occupancy = p.fp.bits();
        if (p.degree > 0 && !pv.empty()) {
            p.pmin = pv.front();
            p.pmax = pv.back();
            p.span = p.pmax - p.pmin + 1;
        }
    }
    const int nw = (n + window - 1) / window;
    int accepted = 0;
    int64_t fb = 0, fa = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : accepted, fb, fa)
    for (int w = 0; w < nw; ++w) {
        const int w0 = w * window, len = std::min(n, w0 + window) - w0;
        int *cur = order->data() + w0;
        std::vector<int> prop(len);
        pack_window(prof, cur, len, group, prop.data());
        int64_t t0, f0, t1, f1;
        window_cost(cur, len, group, panels_of, &t0, &f0);
        window_cost(prop.data(), len, group, panels_of, &t1, &f1);
        const bool take = t1 < t0 || (t1 == t0 && f1 < f0);
        if (take) {
            std::copy(prop.begin(), prop.end(), cur);
            ++accepted;
        }
        fb += f0;
        fa += take ? f1 : f0;
    }
    *windows = nw;
    *fan_before = fb;
    *fan_after = fa;
    return accepted;
}

// Distinct opposite tiles each row touches (K panels under inner old->new).
std::vector<std::vector<int>> row_panels(const Pattern &p, const std::vector<int> &inner_old_to_new)
{
    std::vector<std::vector<int>> out(p.rows);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < p.rows; ++i) {
        std::vector<int> &v = out[i];
        for (int64_t e = p.rp[i]; e < p.rp[i + 1]; ++e)
            v.push_back(inner_old_to_new[p.ci[e]] / kTileK);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    return out;
}

// Distinct row panels each K index touches (row tiles under row old->new).
std::vector<std::vector<int>> inner_panels(const Pattern &p, const std::vector<int> &row_old_to_new)
{
    std::vector<std::vector<int>> out(p.cols);
    for (int i = 0; i < p.rows; ++i) {
        const int panel = row_old_to_new[i] / kTileM;
        for (int64_t e = p.rp[i]; e < p.rp[i + 1]; ++e)
            out[p.ci[e]].push_back(panel);
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int j = 0; j < p.cols; ++j) {
        std::vector<int> &v = out[j];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    return out;
}

} // namespace

extern "C" {

int rtt_bgrf_plan(const rtt_smatrix *a, int *row_new_to_old, int *inner_new_to_old,
                  rtt_bgrf_stats *stats)
{
    rtt::clear_error();
    if (!a || !row_new_to_old || !inner_new_to_old || a->m < 0 || a->n < 0 || a->nnz < 0 ||
        (a->m > 0 && !a->rowpointer) || (a->nnz > 0 && !a->columnindex))
        return rtt::fail(RTT_ERR_INVALID, "rtt_bgrf_plan: bad argument");
    for (int i = 0; i < a->m; ++i)
        if (a->rowpointer[i + 1] < a->rowpointer[i])
            return rtt::fail(RTT_ERR_INVALID, "rtt_bgrf_plan: row pointer falls at row %d", i);
    if (a->m > 0 && (a->rowpointer[0] != 0 || a->rowpointer[a->m] != a->nnz))
        return rtt::fail(RTT_ERR_INVALID, "rtt_bgrf_plan: row pointer ends do not match nnz");
    for (int64_t e = 0; e < a->nnz; ++e)
        if (a->columnindex[e] < 0 || a->columnindex[e] >= a->n)
            return rtt::fail(RTT_ERR_INVALID, "rtt_bgrf_plan: column %d out of range at entry %lld",
                             a->columnindex[e], static_cast<long long>(e));
    const auto t0 = std::chrono::steady_clock::now();
    rtt_bgrf_stats st;
    std::memset(&st, 0, sizeof(st));
    const Pattern p = make_pattern(*a);
    std::vector<int> rows(a->m), inner(a->n);
    std::iota(rows.begin(), rows.end(), 0);
    std::iota(inner.begin(), inner.end(), 0);
    st.tiles_identity = count_tiles(p, rows, inner);

    // I: coarse proposal, kept only when it lowers the exact tile count.
    std::vector<int> crow, cinner;
    st.coarse_levels = coarse_proposal(p, &crow, &cinner, &st.coarse_components);
    st.tiles_coarse = count_tiles(p, crow, invert(cinner));
    if (st.tiles_coarse < st.tiles_identity) {
        rows.swap(crow);
        inner.swap(cinner);
        st.coarse_accepted = 1;
    }
    st.tiles_after_coarse = st.coarse_accepted ? st.tiles_coarse : st.tiles_identity;
    const auto t_fine = std::chrono::steady_clock::now();
    st.coarse_ms = std::chrono::duration<double, std::milli>(t_fine - t0).count();

    // II: rows in windows of 32, groups of 8, over K panels of the inner order.
    {
        const auto panels = row_panels(p, invert(inner));
        int64_t fb = 0, fa = 0;
        st.row_windows_accepted = fine_stage(&rows, panels, p.row_degree, kRowWindow, kRowGroup,
                                             &st.row_windows, &fb, &fa);
        st.row_fanout_before = fb;
        st.row_fanout_after = fa;
    }
    st.tiles_after_rows = count_tiles(p, rows, invert(inner));

    // III: K in windows of 16, groups of 4, over row panels of the row order.
    {
        const auto panels = inner_panels(p, invert(rows));
        int64_t fb = 0, fa = 0;
        st.inner_windows_accepted = fine_stage(&inner, panels, p.col_degree, kInnerWindow,
                                               kInnerGroup, &st.inner_windows, &fb, &fa);
        st.inner_fanout_before = fb;
        st.inner_fanout_after = fa;
    }
    st.tiles_final = count_tiles(p, rows, invert(inner));
    std::copy(rows.begin(), rows.end(), row_new_to_old);
    std::copy(inner.begin(), inner.end(), inner_new_to_old);
    const auto t_end = std::chrono::steady_clock::now();
    st.fine_ms = std::chrono::duration<double, std::milli>(t_end - t_fine).count();
    st.plan_ms = std::chrono::duration<double, std::milli>(t_end - t0).count();
    if (stats)
        *stats = st;
    return RTT_OK;
}

int rtt_prepare_bgrf_a(const rtt_smatrix *host, int dense_threshold, rtt_prepared_a **out,
                       rtt_offline_a_stats *stats, rtt_bgrf_stats *bgrf)
{
    if (!host || !out)
        return rtt::fail(RTT_ERR_INVALID, "rtt_prepare_bgrf_a: NULL argument");
    std::vector<int> rows(host->m > 0 ? host->m : 1), inner(host->n > 0 ? host->n : 1);
    const int rc = rtt_bgrf_plan(host, rows.data(), inner.data(), bgrf);
    if (rc != RTT_OK)
        return rc;
    return rtt_prepare_external_a(host, rows.data(), inner.data(), dense_threshold, out, stats);
}

} // extern "C"
