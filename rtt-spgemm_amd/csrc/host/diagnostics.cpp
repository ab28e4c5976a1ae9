// On-disk diagnostics of a reorder plan, in the reference driver's formats:
//  <prefix>_permutations.csv, <prefix>_unified_stats.csv and
//  <prefix>_A_reordered.mtx (dump_dmma_reorder_plan, src/dmma_reorder.h:3086-3209,
//  formats in docs/BGRF_V1_PPOPP_ALGORITHM_AND_DATA_FORMAT_CN.md section 9.12);
//  <prefix>_heatmap.csv and <prefix>_heatmap_meta.csv (dump_reorder_heatmap,
//  src/main.cu:1114-1246).
// Host only; the plan is given as new -> old maps.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../rtt_common.hpp"

#define RTT_TRY_HOST(expr)             \
    do {                               \
        const int rc_ = (expr);        \
        if (rc_ != RTT_OK)             \
            return rc_;                \
    } while (0)

namespace {

struct PlanFacts {
    std::vector<int> row_o2n, inner_o2n;
    int active_rows = 0, active_inner = 0;
    unsigned long long moved_rows = 0, moved_inner = 0, row_disp = 0, inner_disp = 0;
    long long tiles = 0, sparse = 0, dense = 0;
    unsigned long long payload = 0;
    int active_row_tiles = 0, active_k_tiles = 0;
    bool external = false;   // maps given without a BGRF plan
};

// Inverse maps, movement, active prefixes and the 8x4 tile statistics of A
// laid out by the plan (dense iff at least 24 distinct entries).
int plan_facts(const rtt_smatrix &a, const int *row_n2o, const int *inner_n2o, PlanFacts *f)
{
    f->row_o2n.assign(a.m, -1);
    f->inner_o2n.assign(a.n, -1);
    f->external = row_n2o || inner_n2o;
    for (int i = 0; i < a.m; ++i) {
        const int o = row_n2o ? row_n2o[i] : i;
        if (o < 0 || o >= a.m || f->row_o2n[o] >= 0)
            return rtt::fail(RTT_ERR_INVALID, "row map is not a bijection at new id %d", i);
        f->row_o2n[o] = i;
    }
    for (int j = 0; j < a.n; ++j) {
        const int o = inner_n2o ? inner_n2o[j] : j;
        if (o < 0 || o >= a.n || f->inner_o2n[o] >= 0)
            return rtt::fail(RTT_ERR_INVALID, "inner map is not a bijection at new id %d", j);
        f->inner_o2n[o] = j;
    }
    if (a.m > 0 && (a.rowpointer[0] != 0 || a.rowpointer[a.m] != a.nnz))
        return rtt::fail(RTT_ERR_INVALID, "row pointer ends do not match nnz");
    for (int o = 0; o < a.m; ++o)
        if (a.rowpointer[o + 1] < a.rowpointer[o])
            return rtt::fail(RTT_ERR_INVALID, "row pointer falls at row %d", o);
    for (int o = 0; o < a.m; ++o) {
        const int n = f->row_o2n[o];
        f->moved_rows += n != o;
        f->row_disp += static_cast<unsigned long long>(std::abs(n - o));
        if (a.rowpointer[o + 1] > a.rowpointer[o])
            f->active_rows = std::max(f->active_rows, n + 1);
    }
    for (int o = 0; o < a.n; ++o) {
        const int n = f->inner_o2n[o];
        f->moved_inner += n != o;
        f->inner_disp += static_cast<unsigned long long>(std::abs(n - o));
    }
    for (int64_t e = 0; e < a.nnz; ++e) {
        if (a.columnindex[e] < 0 || a.columnindex[e] >= a.n)
            return rtt::fail(RTT_ERR_INVALID, "column out of range at entry %lld",
                             static_cast<long long>(e));
        f->active_inner = std::max(f->active_inner, f->inner_o2n[a.columnindex[e]] + 1);
    }
    f->active_row_tiles = (f->active_rows + 7) / 8;
    f->active_k_tiles = (f->active_inner + 3) / 4;
    std::vector<long long> keys;
    keys.reserve(static_cast<size_t>(a.nnz));
    const long long kt = (a.n + 3) / 4;
    for (int o = 0; o < a.m; ++o)
        for (int e = a.rowpointer[o]; e < a.rowpointer[o + 1]; ++e) {
            const int r = f->row_o2n[o], c = f->inner_o2n[a.columnindex[e]];
            // (tile, slot) with the slot in the low 5 bits: distinct entries.
            keys.push_back(((r / 8) * kt + c / 4) * 32 + (r % 8) * 4 + c % 4);
        }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    for (size_t i = 0; i < keys.size();) {
        size_t j = i;
        while (j < keys.size() && keys[j] / 32 == keys[i] / 32)
            ++j;
        const long long cnt = static_cast<long long>(j - i);
        ++f->tiles;
        if (cnt >= 24) {
            ++f->dense;
            f->payload += 32;
        } else {
            ++f->sparse;
            f->payload += static_cast<unsigned long long>(cnt);
        }
        i = j;
    }
    return RTT_OK;
}

int heatmap_bin(int index, int extent, int bins)
{
    if (extent <= 0)
        return 0;
    return std::min(bins - 1, static_cast<int>(static_cast<uint64_t>(index) *
                                               static_cast<uint64_t>(bins) / extent));
}

bool close_ok(FILE *f)
{
    return std::fclose(f) == 0;
}

void write_plan_keys(FILE *f, const rtt_bgrf_stats *st, const PlanFacts &p)
{
    const bool bgrf = st != nullptr;
    std::fprintf(f, "algorithm,%s\nsweeps,%d\nrow_window,%d\ninner_window,%d\n",
                 bgrf ? "bgrf-v1" : (p.external ? "external" : "identity_baseline"),
                 bgrf ? 1 : 0, 32, 16);
    std::fprintf(f, "active_rows,%d\nactive_inner,%d\n", p.active_rows, p.active_inner);
    std::fprintf(f,
                 "moved_rows,%llu\nmoved_inner,%llu\nrow_displacement_sum,%llu\n"
                 "inner_displacement_sum,%llu\n",
                 p.moved_rows, p.moved_inner, p.row_disp, p.inner_disp);
    if (bgrf) {
        const long long row_red = st->tiles_after_coarse - st->tiles_after_rows;
        const long long inner_red = st->tiles_after_rows - st->tiles_final;
        const long long coarse_red = st->tiles_identity - st->tiles_after_coarse;
        std::fprintf(f,
                     "accepted_row_windows,%d\naccepted_inner_windows,%d\n"
                     "row_tile_reduction,%lld\ninner_tile_reduction,%lld\n"
                     "row_fanout_before,%lld\nrow_fanout_after,%lld\n"
                     "inner_fanout_before,%lld\ninner_fanout_after,%lld\n"
                     "coarse_components,%d\ncoarse_levels,%d\ncoarse_level_budget,%d\n"
                     "coarse_candidate_accepted,%d\ncoarse_tile_reduction,%lld\n"
                     "coarse_ms,%.6f\nfine_ms,%.6f\n",
                     st->row_windows_accepted, st->inner_windows_accepted, row_red, inner_red,
                     st->row_fanout_before, st->row_fanout_after, st->inner_fanout_before,
                     st->inner_fanout_after, st->coarse_components, st->coarse_levels, 2048,
                     st->coarse_accepted, coarse_red, st->coarse_ms, st->fine_ms);
    }
    std::fprintf(f,
                 "num_tiles,%lld\nactive_row_tiles,%d\nactive_k_tiles,%d\nsparse_tiles,%lld\n"
                 "dense_tiles,%lld\npayload,%llu\n",
                 p.tiles, p.active_row_tiles, p.active_k_tiles, p.sparse, p.dense, p.payload);
}

bool valid_matrix(const rtt_smatrix *a)
{
    return a && a->m >= 0 && a->n >= 0 && a->nnz >= 0 && (a->m == 0 || a->rowpointer) &&
           (a->nnz == 0 || a->columnindex);
}

} // namespace

extern "C" {

int rtt_dump_reorder(const char *prefix, const rtt_smatrix *a, const int *row_new_to_old,
                     const int *inner_new_to_old, const rtt_bgrf_stats *stats)
{
    rtt::clear_error();
    if (!prefix || !valid_matrix(a))
        return rtt::fail(RTT_ERR_INVALID, "rtt_dump_reorder: bad argument");
    PlanFacts p;
    RTT_TRY_HOST(plan_facts(*a, row_new_to_old, inner_new_to_old, &p));
    const std::string stem(prefix);
    FILE *f = std::fopen((stem + "_permutations.csv").c_str(), "w");
    if (!f)
        return rtt::fail(RTT_ERR_INVALID, "cannot write %s_permutations.csv", prefix);
    std::fprintf(f, "axis,old_id,new_id,new_to_old_at_new_id\n");
    for (int o = 0; o < a->m; ++o) {
        const int n = p.row_o2n[o];
        std::fprintf(f, "row,%d,%d,%d\n", o, n, row_new_to_old ? row_new_to_old[n] : n);
    }
    for (int o = 0; o < a->n; ++o) {
        const int n = p.inner_o2n[o];
        std::fprintf(f, "inner,%d,%d,%d\n", o, n, inner_new_to_old ? inner_new_to_old[n] : n);
    }
    if (!close_ok(f))
        return rtt::fail(RTT_ERR_INVALID, "write of %s_permutations.csv failed", prefix);
    f = std::fopen((stem + "_unified_stats.csv").c_str(), "w");
    if (!f)
        return rtt::fail(RTT_ERR_INVALID, "cannot write %s_unified_stats.csv", prefix);
    std::fprintf(f, "key,value\n");
    write_plan_keys(f, stats, p);
    if (!close_ok(f))
        return rtt::fail(RTT_ERR_INVALID, "write of %s_unified_stats.csv failed", prefix);
    // Structure only, 1-based, every CSR entry (duplicates kept), written in
    // old-row order.
    f = std::fopen((stem + "_A_reordered.mtx").c_str(), "w");
    if (!f)
        return rtt::fail(RTT_ERR_INVALID, "cannot write %s_A_reordered.mtx", prefix);
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate pattern general\n");
    std::fprintf(f, "%% Unified 8x4 row/inner reordered A\n");
    std::fprintf(f, "%d %d %d\n", a->m, a->n, a->nnz);
    for (int o = 0; o < a->m; ++o)
        for (int e = a->rowpointer[o]; e < a->rowpointer[o + 1]; ++e)
            std::fprintf(f, "%d %d\n", p.row_o2n[o] + 1, p.inner_o2n[a->columnindex[e]] + 1);
    if (!close_ok(f))
        return rtt::fail(RTT_ERR_INVALID, "write of %s_A_reordered.mtx failed", prefix);
    return RTT_OK;
}

int rtt_dump_reorder_heatmap(const char *prefix, const rtt_smatrix *a, const int *row_new_to_old,
                             const int *inner_new_to_old, int bins,
                             const rtt_bgrf_stats *stats)
{
    rtt::clear_error();
    if (!prefix || !valid_matrix(a) || bins < 1)
        return rtt::fail(RTT_ERR_INVALID, "rtt_dump_reorder_heatmap: bad argument");
    PlanFacts p;
    RTT_TRY_HOST(plan_facts(*a, row_new_to_old, inner_new_to_old, &p));
    const size_t cells = static_cast<size_t>(bins) * static_cast<size_t>(bins);
    std::vector<unsigned long long> before(cells, 0), after(cells, 0);
    // Every CSR entry counts (duplicates included), in both layouts.
    for (int o = 0; o < a->m; ++o) {
        const int br0 = heatmap_bin(o, a->m, bins), br1 = heatmap_bin(p.row_o2n[o], a->m, bins);
        for (int e = a->rowpointer[o]; e < a->rowpointer[o + 1]; ++e) {
            const int c = a->columnindex[e];
            ++before[static_cast<size_t>(br0) * bins + heatmap_bin(c, a->n, bins)];
            ++after[static_cast<size_t>(br1) * bins + heatmap_bin(p.inner_o2n[c], a->n, bins)];
        }
    }
    const std::string stem(prefix);
    FILE *f = std::fopen((stem + "_heatmap.csv").c_str(), "w");
    if (!f)
        return rtt::fail(RTT_ERR_INVALID, "cannot write %s_heatmap.csv", prefix);
    std::fprintf(f, "bin_row,bin_col,original_nnz,reordered_nnz\n");
    for (int r = 0; r < bins; ++r)
        for (int c = 0; c < bins; ++c) {
            const size_t cell = static_cast<size_t>(r) * bins + c;
            std::fprintf(f, "%d,%d,%llu,%llu\n", r, c, before[cell], after[cell]);
        }
    if (!close_ok(f))
        return rtt::fail(RTT_ERR_INVALID, "write of %s_heatmap.csv failed", prefix);
    f = std::fopen((stem + "_heatmap_meta.csv").c_str(), "w");
    if (!f)
        return rtt::fail(RTT_ERR_INVALID, "cannot write %s_heatmap_meta.csv", prefix);
    std::fprintf(f, "key,value\nrows,%d\ncols,%d\nnnz,%d\nbins,%d\n", a->m, a->n, a->nnz, bins);
    write_plan_keys(f, stats, p);
    if (!close_ok(f))
        return rtt::fail(RTT_ERR_INVALID, "write of %s_heatmap_meta.csv failed", prefix);
    return RTT_OK;
}

} // extern "C"
