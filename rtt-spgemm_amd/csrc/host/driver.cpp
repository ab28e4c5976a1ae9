// rtt_spgemm: drop-in driver for the C = A*A / C = A*A^T / C = A*B benchmark
// path.
//
// Mirrors the reference driver's flags (src/main.cu:168-309), its timed
// boundary (B rebuild + symbolic + allocation + numeric, median over the
// iterations, A preparation untimed: src/main.cu:2171-2318) and its
// RTT_RESULT summary line (src/main.cu:2804-2854).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rtt_spgemm.h"

namespace {

struct Options {
    int device = 0;
    int aat = 0;
    int dense_threshold = 24;
    int warmup = 0;
    int iterations = 5;
    int lb_mode = RTT_LB_ADAPTIVE_TAIL;
    const char *lb_name = "adaptive-tail";
    bool no_reorder = false;
    bool reference_values = false;
    bool prepare_only = false;
    const char *row_order = nullptr;
    const char *inner_order = nullptr;
    const char *reorder_name = nullptr;
    const char *a_file = nullptr;
    const char *b_file = nullptr;      // --b: general A*B
    int b_update = 0;                  // 0 none, 1 values, 2 permutation, 3 structure
    int b_clear_policy = RTT_B_VALUES_SAFE_AUTO;
    const char *dump_prefix = nullptr;    // --dump-reorder-prefix
    const char *heatmap_prefix = nullptr; // --dump-reorder-heatmap
    int heatmap_bins = 256;
};

const char *const kBUpdateNames[] = {"none", "values", "permutation", "structure"};

void usage(const char *prog)
{
    std::printf(
        "Usage: %s [options] A.mtx\n"
        "  -d GPU                         HIP device (default: 0)\n"
        "  -aat 0|1                       C = A*A (0) or C = A*A^T (1)\n"
        "  --b B.mtx                      compute general A*B; otherwise A*A\n"
        "  --b-update none|values|permutation|structure (default: none)\n"
        "                                 values with --b: timed values-only\n"
        "                                 update of B; otherwise a full rebuild\n"
        "  --b-values-clear always-clear|noclear|safe-auto (default: safe-auto)\n"
        "  --lb-mode static|dynamic|adaptive-tail (default: adaptive-tail)\n"
        "  --warmup-iterations N          warmups (default: 0)\n"
        "  --iterations N                 repeats (default: 5)\n"
        "  --no-reorder                   identity layout (default: the BGRF-v1 plan)\n"
        "  --row-order FILE --inner-order FILE --reorder-name NAME\n"
        "  --dump-reorder-prefix P | --dump-reorder-heatmap P [--heatmap-bins N (16..1024)]\n"
        "  --reference-values             replace values by i %% 10 as the\n"
        "                                 reference driver does\n"
        "  --prepare-only\n",
        prog);
}

bool parse(int argc, char **argv, Options *o)
{
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        auto next = [&](void) -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
        if (!std::strcmp(a, "-d")) {
            const char *v = next();
            if (!v)
                return false;
            o->device = std::atoi(v);
        } else if (!std::strcmp(a, "-aat")) {
            const char *v = next();
            if (!v)
                return false;
            o->aat = std::atoi(v);
        } else if (!std::strcmp(a, "--iterations")) {
            const char *v = next();
            if (!v)
                return false;
            o->iterations = std::atoi(v);
        } else if (!std::strcmp(a, "--warmup-iterations")) {
            const char *v = next();
            if (!v)
                return false;
            o->warmup = std::atoi(v);
        } else if (!std::strcmp(a, "--lb-mode")) {
            const char *v = next();
            if (!v)
                return false;
            if (!std::strcmp(v, "static"))
                o->lb_mode = RTT_LB_STATIC, o->lb_name = "static";
            else if (!std::strcmp(v, "dynamic"))
                o->lb_mode = RTT_LB_DYNAMIC, o->lb_name = "dynamic";
            else if (!std::strcmp(v, "adaptive-tail") || !std::strcmp(v, "tail"))
                o->lb_mode = RTT_LB_ADAPTIVE_TAIL, o->lb_name = "adaptive-tail";
            else
                return false;
        } else if (!std::strcmp(a, "--no-reorder")) {
            o->no_reorder = true;
        } else if (!std::strcmp(a, "--row-order")) {
            if (!(o->row_order = next()))
                return false;
        } else if (!std::strcmp(a, "--inner-order")) {
            if (!(o->inner_order = next()))
                return false;
        } else if (!std::strcmp(a, "--reorder-name")) {
            if (!(o->reorder_name = next()))
                return false;
        } else if (!std::strcmp(a, "--reference-values")) {
            o->reference_values = true;
        } else if (!std::strcmp(a, "--prepare-only")) {
            o->prepare_only = true;
        } else if (!std::strcmp(a, "--dump-reorder-prefix")) {
            if (!(o->dump_prefix = next()))
                return false;
        } else if (!std::strcmp(a, "--dump-reorder-heatmap")) {
            if (!(o->heatmap_prefix = next()))
                return false;
        } else if (!std::strcmp(a, "--heatmap-bins")) {
            const char *v = next();
            if (!v)
                return false;
            o->heatmap_bins = std::atoi(v);
        } else if (!std::strcmp(a, "--b")) {
            if (!(o->b_file = next()))
                return false;
        } else if (!std::strcmp(a, "--b-update")) {
            const char *v = next();
            int mode = -1;
            for (int k = 0; v && k < 4; ++k)
                if (!std::strcmp(v, kBUpdateNames[k]))
                    mode = k;
            if (mode < 0)
                return false;
            o->b_update = mode;
        } else if (!std::strcmp(a, "--b-values-clear")) {
            const char *v = next();
            int pol = -1;
            for (int k = 0; v && k < 3; ++k)
                if (!std::strcmp(v, rtt_b_values_clear_policy_name(k)))
                    pol = k;
            if (pol < 0)
                return false;
            o->b_clear_policy = pol;
        } else if (a[0] != '-' && !o->a_file) {
            o->a_file = a;
        } else {
            std::fprintf(stderr, "unsupported option '%s'\n", a);
            return false;
        }
    }
    const bool any_ext = o->row_order || o->inner_order || o->reorder_name;
    const bool all_ext = o->row_order && o->inner_order && o->reorder_name;
    return o->a_file && (o->aat == 0 || o->aat == 1) && o->warmup >= 0 && o->iterations > 0 &&
           (!any_ext || all_ext) && !(o->no_reorder && any_ext) && o->heatmap_bins >= 16 &&
           o->heatmap_bins <= 1024;
}

double median_of(const std::vector<double> &v, size_t *selected)
{
    std::vector<size_t> order(v.size());
    for (size_t i = 0; i < order.size(); ++i)
        order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return v[x] < v[y]; });
    *selected = order[order.size() / 2];
    return v[*selected];
}

int check(int rc, const char *what)
{
    if (rc != RTT_OK)
        std::fprintf(stderr, "%s failed: %s\n", what, rtt_last_error());
    return rc;
}

} // namespace

int main(int argc, char **argv)
{
    Options o;
    if (!parse(argc, argv, &o)) {
        usage(argv[0]);
        return EXIT_FAILURE;
    }
    rtt_smatrix A;
    std::memset(&A, 0, sizeof(A));
    if (check(rtt_read_mtx(o.a_file, o.reference_values ? 1 : 0, &A), "read"))
        return EXIT_FAILURE;
    const bool general_ab = o.b_file != nullptr;
    rtt_smatrix B;
    std::memset(&B, 0, sizeof(B));
    if (general_ab) {
        if (check(rtt_read_mtx(o.b_file, o.reference_values ? 1 : 0, &B), "read B"))
            return EXIT_FAILURE;
        if (A.n != B.m) {
            std::fprintf(stderr, "A*B needs A.n == B.m (got %d x %d times %d x %d).\n", A.m,
                         A.n, B.m, B.n);
            return EXIT_FAILURE;
        }
        if (o.aat)
            std::printf("Note: --b selects general A*B; -aat is ignored.\n");
        o.aat = 0;
    } else if (!o.aat && A.m != A.n) {
        std::fprintf(stderr, "C = A*A needs a square matrix (got %d x %d).\n", A.m, A.n);
        return EXIT_FAILURE;
    }
    if (check(rtt_init(o.device), "rtt_init"))
        return EXIT_FAILURE;
    std::printf("A: %d x %d, nnz %d, aat %d, lb-mode %s\n", A.m, A.n, A.nnz, o.aat, o.lb_name);
    rtt_device_csr *bsrc = nullptr;
    if (general_ab) {
        double h2d = 0, val = 0;
        if (check(rtt_upload_csr(&B, &bsrc, &h2d, &val), "B upload"))
            return EXIT_FAILURE;
        std::printf("B: %d x %d, nnz %d; B CSR H2D = %.3f ms; validation = %.3f ms\n", B.m, B.n,
                    B.nnz, h2d, val);
    } else {
        std::printf("B source reuses A device CSR; additional CSR H2D = 0 ms\n");
    }

    rtt_prepared_a *pa = nullptr;
    rtt_offline_a_stats ast;
    std::vector<int> row_order, inner_order;
    const char *layout = "identity";
    rtt_bgrf_stats bgrf;
    bool have_bgrf = false;
    int rc;
    if (o.row_order) {
        row_order.resize(A.m);
        inner_order.resize(A.n);
        if (check(rtt_read_order_file(o.row_order, A.m, row_order.data()), "row order") ||
            check(rtt_read_order_file(o.inner_order, A.n, inner_order.data()), "inner order"))
            return EXIT_FAILURE;
        rc = rtt_prepare_external_a(&A, row_order.data(), inner_order.data(), o.dense_threshold,
                                    &pa, &ast);
        layout = "external";
    } else if (o.no_reorder) {
        // Identity layout with the empty suffix trimmed (--no-reorder).
        rc = rtt_prepare_identity_a(&A, o.dense_threshold, &pa, &ast);
    } else {
        // The reference's default: the BGRF-v1 plan (src/main.cu:1863-1874).
        row_order.resize(A.m > 0 ? A.m : 1);
        inner_order.resize(A.n > 0 ? A.n : 1);
        rc = rtt_bgrf_plan(&A, row_order.data(), inner_order.data(), &bgrf);
        if (rc == RTT_OK)
            rc = rtt_prepare_external_a(&A, row_order.data(), inner_order.data(),
                                        o.dense_threshold, &pa, &ast);
        layout = "bgrf-v1";
        have_bgrf = rc == RTT_OK;
        if (have_bgrf)
            std::printf("BGRF plan %.3f ms: A tiles identity %lld, coarse %lld (%s), rows %lld "
                        "(%d/%d windows), final %lld (%d/%d K windows)\n",
                        bgrf.plan_ms, bgrf.tiles_identity, bgrf.tiles_coarse,
                        bgrf.coarse_accepted ? "accepted" : "rejected", bgrf.tiles_after_rows,
                        bgrf.row_windows_accepted, bgrf.row_windows, bgrf.tiles_final,
                        bgrf.inner_windows_accepted, bgrf.inner_windows);
    }
    if (check(rc, "prepare A"))
        return EXIT_FAILURE;
    {
        // Reorder diagnostics (src/main.cu:1930-1960): the layout in use.
        const int *rmap = o.no_reorder ? nullptr : row_order.data();
        const int *imap = o.no_reorder ? nullptr : inner_order.data();
        const rtt_bgrf_stats *bs = have_bgrf ? &bgrf : nullptr;
        if (o.dump_prefix &&
            check(rtt_dump_reorder(o.dump_prefix, &A, rmap, imap, bs), "dump reorder"))
            return EXIT_FAILURE;
        if (o.heatmap_prefix &&
            check(rtt_dump_reorder_heatmap(o.heatmap_prefix, &A, rmap, imap, o.heatmap_bins, bs),
                  "dump heatmap"))
            return EXIT_FAILURE;
        if (o.dump_prefix)
            std::printf("REORDER_DUMP prefix=%s\n", o.dump_prefix);
        if (o.heatmap_prefix)
            std::printf("REORDER_HEATMAP prefix=%s bins=%d\n", o.heatmap_prefix, o.heatmap_bins);
    }
    std::printf("A prepared (%s): upload %.3f ms, validate %.3f ms, row sort %.3f ms, "
                "tiles %.3f ms\n",
                layout, ast.h2d_ms, ast.validation_ms, ast.canonicalize_ms, ast.tile_build_ms);
    if (o.prepare_only) {
        rtt_destroy_prepared_a(pa);
        rtt_destroy_smatrix(&A);
        return EXIT_SUCCESS;
    }
    unsigned long long nnz_cub = 0;
    double cub_ms = 0;
    if (check(general_ab ? rtt_compute_nnz_cub_ab(pa, bsrc, &nnz_cub, &cub_ms)
                         : rtt_compute_nnz_cub(pa, o.aat, &nnz_cub, &cub_ms),
              "nnzCub"))
        return EXIT_FAILURE;
    std::printf("SpGEMM nnzCub = %llu; GPU nnzCub time = %.3f ms\n", nnz_cub, cub_ms);

    rtt_schedule sched;
    rtt_schedule_defaults(&sched);
    sched.lb_mode = o.lb_mode;
    rtt_dynamic_b *pb = nullptr;
    rtt_smatrix C;
    std::memset(&C, 0, sizeof(C));
    // The initial (untimed) B build, as the reference does before its
    // samples (src/main.cu:1996-2005).
    rtt_b_update_stats bst0;
    if (check(general_ab ? rtt_rebuild_b_general(pa, bsrc, &pb, &bst0)
                         : rtt_rebuild_b(pa, o.aat, &pb, &bst0),
              "initial B rebuild"))
        return EXIT_FAILURE;
    std::printf("Initial B rebuild %.3f ms (path %d)\n", bst0.total_ms, bst0.path);
    const bool values_only = general_ab && o.b_update == 1;
    std::vector<double> b_ms, b_perm_ms, b_tile_ms, sym_ms, alloc_ms, num_ms, total_ms;
    for (int it = 0; it < o.warmup + o.iterations; ++it) {
        const bool timed = it >= o.warmup;
        const bool last = it + 1 == o.warmup + o.iterations;
        sched.materialize_output = last ? 1 : 0;
        rtt_b_update_stats bst;
        rtt_stats st;
        rtt_b_values_decision dec;
        if (check(rtt_sync(), "sync"))
            return EXIT_FAILURE;
        if (values_only)
            rc = rtt_update_b_values(pa, bsrc, o.b_clear_policy, pb, &bst, &dec);
        else if (general_ab)
            rc = rtt_rebuild_b_general(pa, bsrc, &pb, &bst);
        else
            rc = rtt_rebuild_b(pa, o.aat, &pb, &bst);
        if (check(rc, "B update") || check(rtt_tilespgemm(pa, pb, &C, &st, &sched), "SpGEMM"))
            return EXIT_FAILURE;
        if (values_only && it == 0)
            std::printf("B_UPDATE mode=values path=%d clear=%d reason=%s\n", bst.path,
                        dec.clear_payload, rtt_b_values_clear_reason_name(dec.reason));
        if (!timed)
            continue;
        b_ms.push_back(bst.total_ms);
        b_perm_ms.push_back(bst.csr_row_permute_ms);
        b_tile_ms.push_back(bst.csr2tile_ms);
        sym_ms.push_back(st.symbolic_ms);
        alloc_ms.push_back(st.allocation_ms);
        num_ms.push_back(st.numeric_ms);
        total_ms.push_back(bst.total_ms + st.symbolic_ms + st.allocation_ms + st.numeric_ms);
        std::printf("iteration %d: b_update %.3f ms symbolic %.3f ms malloc %.3f ms "
                    "numeric %.3f ms tiles %llu nnzC %llu\n",
                    it - o.warmup + 1, bst.total_ms, st.symbolic_ms, st.allocation_ms,
                    st.numeric_ms, st.output_tiles, st.output_nnz);
    }
    size_t sel = 0;
    const double runtime = median_of(total_ms, &sel);
    const double gflops = runtime > 0 ? 2.0 * static_cast<double>(nnz_cub) / (runtime * 1e6) : 0.0;
    std::printf("RTT-SpGEMM (%s), repeats = %d, statistic = median, b_update_mode = %s, "
                "product = %s\n",
                o.lb_name, o.iterations, kBUpdateNames[o.b_update],
                general_ab ? "AB" : (o.aat ? "AAT" : "AA"));
    std::printf("B online preparation runtime is %.3f ms\n", b_ms[sel]);
    std::printf("  CSR row permutation is %.3f ms; CSR2Tile is %.3f ms\n", b_perm_ms[sel],
                b_tile_ms[sel]);
    std::printf("Symbolic runtime is %.3f ms\n", sym_ms[sel]);
    std::printf("Malloc runtime is %.3f ms\n", alloc_ms[sel]);
    std::printf("Numeric runtime is %.3f ms\n", num_ms[sel]);
    std::printf("HIP   RTT-SpGEMM runtime is %.3f ms, gflops = %.2f\n", runtime, gflops);
    std::printf("RTT_RESULT matrix=%s mode=%s repeats=%d statistic=median "
                "selected_iteration=%zu b_permutation_ms=%.6f "
                "b_csr_row_permute_ms=%.6f b_csr2tile_ms=%.6f "
                "symbolic_ms=%.6f malloc_ms=%.6f "
                "numeric_ms=%.6f runtime_ms=%.6f gflops=%.6f "
                "input_h2d_included=0 output_d2h_included=0 "
                "validation_included=0 capacity_workspace_reuse=1\n",
                o.a_file, o.lb_name, o.iterations, sel + 1, b_ms[sel], b_perm_ms[sel],
                b_tile_ms[sel], sym_ms[sel], alloc_ms[sel], num_ms[sel], runtime, gflops);
    // Convert the exported tiles to CSR (untimed, as in the reference).
    const int *perm = rtt_prepared_a_row_new_to_old(pa);
    int active_rows = 0;
    rtt_prepared_a_info(pa, &active_rows, nullptr, nullptr);
    rc = perm ? rtt_tile2csr_restore_rows(&C, perm, active_rows, 1) : rtt_tile2csr(&C, 1);
    if (check(rc, "tile2csr"))
        return EXIT_FAILURE;
    std::printf("C: %d x %d, nnz %d, tiles %d\n", C.m, C.n, C.nnz, C.numtile);
    rtt_destroy_smatrix(&C);
    rtt_destroy_dynamic_b(pb);
    rtt_destroy_device_csr(bsrc);
    rtt_destroy_prepared_a(pa);
    rtt_destroy_smatrix(&A);
    if (general_ab)
        rtt_destroy_smatrix(&B);
    std::printf("---------------------------------------------------------------\n");
    return EXIT_SUCCESS;
}
