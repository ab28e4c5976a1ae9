/*
 * rtt_spgemm.h -- C ABI of the MI355X-native tiled FP64 SpGEMM engine.
 *
 * The entry points replace the header-only C++ interface that the
 * reference's driver (src/main.cu) calls for the C = A*A (-aat 0) and
 * C = A*A^T (-aat 1) path.  Each declaration names the reference function it
 * replaces.  All types are plain C: device buffers are opaque handles or raw
 * pointers, host buffers are malloc'ed by the library and released with the
 * matching rtt_destroy_* call.  No HIP or torch type appears here.
 *
 * Threading: calls on one rtt_prepared_a handle must not overlap (the handle
 * owns the reusable C output workspace).  Distinct handles may be used from
 * different host threads.
 */
#ifndef RTT_SPGEMM_H_
#define RTT_SPGEMM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes (the reference returns bool + stderr text; see               */
/* src/dmma_spgemm.h:1192-1199).  rtt_last_error() gives the text.           */
/* ------------------------------------------------------------------------ */
enum rtt_status {
    RTT_OK = 0,
    RTT_ERR_INVALID = 1,     /* malformed input or arguments               */
    RTT_ERR_OOM = 2,         /* host or device allocation failed            */
    RTT_ERR_OVERFLOW = 3,    /* a count exceeds the int32 legacy ABI        */
    RTT_ERR_DEVICE = 4,      /* HIP runtime error or device error flag      */
    RTT_ERR_UNSUPPORTED = 5  /* option not implemented by this build        */
};

/* Thread-local text of the last failure ("" when none). */
const char *rtt_last_error(void);

/* Version string of the library build. */
const char *rtt_version(void);

/* ------------------------------------------------------------------------ */
/* Host matrices: the legacy SMatrix layout, field for field                 */
/* (src/common.h:48-73).  A maintainer can cast an SMatrix* to rtt_smatrix*. */
/* ------------------------------------------------------------------------ */
typedef struct rtt_smatrix {
    int m;
    int n;
    int nnz;
    int isSymmetric;
    double *value;
    int *columnindex;
    int *rowpointer;
    int tilem;
    int tilen;
    int *tile_ptr;
    int *tile_columnidx;
    int *tile_rowidx;
    int *tile_nnz;
    int numtile;
    double *tile_csr_Value;
    unsigned char *tile_csr_Col;
    unsigned char *tile_csr_Ptr;
    unsigned short *mask;
    int *csc_tile_ptr;
    int *csc_tile_rowidx;
} rtt_smatrix;

/* Frees every non-NULL array of *m with free() and zeroes the struct
 * (destroy_output_matrix / free_host_csr, src/main.cu:1059-1089). */
void rtt_destroy_smatrix(rtt_smatrix *m);

/* Matrix Market reader with mmio_allinone semantics
 * (src/mmio_highlevel.h:593-759): 0-based CSR, symmetric/hermitian storage
 * mirrored, rows kept in file order, duplicates kept.  When
 * reference_values != 0 every value is replaced by (i % 10) as the
 * reference driver's load_matrix does (src/main.cu:1091-1112). */
int rtt_read_mtx(const char *path, int reference_values, rtt_smatrix *out);

/* Writes a CSR matrix as "matrix coordinate real general", %.17g values. */
int rtt_write_mtx(const char *path, const rtt_smatrix *csr);

/* ------------------------------------------------------------------------ */
/* Device                                                                    */
/* ------------------------------------------------------------------------ */
/* Selects the HIP device (cudaSetDevice, src/main.cu:1526) and creates the
 * library stream.  Must precede every call that touches the GPU. */
int rtt_init(int device);
/* Blocks until all work issued by the library has completed. */
int rtt_sync(void);

/* Device tile view: DmmaDeviceTiles (src/dmma_device_tiles.h:8-28).
 * Pointers are device addresses.  A: 8x4 tiles, row-major payload
 * (physical = r*4+k).  B: 4x8 tiles, column-major payload
 * (physical = c*4+k) plus a tile CSC.  A tile is dense iff its payload
 * span value_offsets[t+1]-value_offsets[t] is 32; otherwise its values are
 * packed in increasing mask-bit order. */
typedef struct rtt_device_tiles {
    int rows;
    int cols;
    int tile_rows;
    int tile_cols;
    int tile_row_count;
    int tile_col_count;
    int num_tiles;
    int payload_size;
    int dense_tiles;
    int sparse_tiles;
    const int *tile_row_ptr;
    const int *tile_col_idx;
    const int *value_offsets;
    const uint32_t *masks;
    const double *values;
    const int *tile_col_ptr;
    const int *tile_row_idx;
    const int *csc_tile_ids;
} rtt_device_tiles;

/* Copies a device tile view to caller-provided host arrays (any may be NULL;
 * sizes follow the view's counts). */
int rtt_download_tiles(const rtt_device_tiles *view, int *tile_row_ptr,
                       int *tile_col_idx, int *value_offsets,
                       uint32_t *masks, double *values, int *tile_col_ptr,
                       int *tile_row_idx, int *csc_tile_ids);

/* ------------------------------------------------------------------------ */
/* Offline A preparation                                                     */
/* ------------------------------------------------------------------------ */
typedef struct rtt_prepared_a rtt_prepared_a;   /* DmmaPreparedA */

typedef struct rtt_offline_a_stats {            /* DmmaOfflineAStats */
    double h2d_ms;
    double validation_ms;
    double canonicalize_ms;   /* per-row column sort when rows arrive unsorted */
    double reorder_ms;
    double tile_build_ms;
    double total_ms;
    int rows_were_sorted;     /* 1 if the input rows were already column-sorted */
    double transpose_src_ms;  /* part of tile_build_ms: the per-tile records that the
                               * A*A^T B update (identity rows) copies from; the reference
                               * builds the corresponding transpose inside its timed B
                               * rebuild */
} rtt_offline_a_stats;

/* gpu_prepare_identity_a (src/gpu_dmma_tiles.h:2443-2502): upload and
 * validate the CSR, keep the identity layout with only the empty row / inner
 * suffix trimmed, and build A's 8x4 tiles. */
int rtt_prepare_identity_a(const rtt_smatrix *host, int dense_threshold,
                           rtt_prepared_a **out, rtt_offline_a_stats *stats);

/* gpu_prepare_external_a (src/gpu_dmma_tiles.h:2389-2441) with the order
 * files already parsed: row_new_to_old has host->m entries and
 * inner_new_to_old host->n entries, both bijections (new_id -> old_id, the
 * token semantics of src/dmma_reorder.h:2048-2184). */
int rtt_prepare_external_a(const rtt_smatrix *host,
                           const int *row_new_to_old,
                           const int *inner_new_to_old, int dense_threshold,
                           rtt_prepared_a **out, rtt_offline_a_stats *stats);

/* BGRF-v1 offline reorder (build_dmma_reorder_plan, src/dmma_reorder.h:2651;
 * algorithm in docs/BGRF_V1_PPOPP_ALGORITHM_AND_DATA_FORMAT_CN.md sections
 * 3-6), the reference driver's default layout (src/main.cu:1863-1874).
 * Host-side, deterministic.  Statistics of one plan: */
typedef struct rtt_bgrf_stats {               /* DmmaReorderPlan counters */
    long long tiles_identity;      /* occupied 8x4 A tiles, identity layout */
    long long tiles_coarse;        /* ... under the coarse proposal */
    long long tiles_after_coarse;  /* ... after the coarse accept/reject */
    long long tiles_after_rows;    /* ... after the 32/8 row windows */
    long long tiles_final;         /* ... after the 16/4 K windows */
    long long row_fanout_before, row_fanout_after;
    long long inner_fanout_before, inner_fanout_after;
    int coarse_accepted;
    int coarse_levels;             /* frontier rounds of the traversal */
    int coarse_components;         /* traversals started from a seed */
    int row_windows, row_windows_accepted;
    int inner_windows, inner_windows_accepted;
    double coarse_ms, fine_ms, plan_ms;
} rtt_bgrf_stats;

/* Computes the row and inner permutations (new -> old, host->m and host->n
 * entries) for rtt_prepare_external_a.  stats may be NULL. */
int rtt_bgrf_plan(const rtt_smatrix *host, int *row_new_to_old, int *inner_new_to_old,
                  rtt_bgrf_stats *stats);
/* rtt_bgrf_plan followed by rtt_prepare_external_a (gpu_prepare_reordered_a,
 * src/gpu_dmma_tiles.h:2334). */
int rtt_prepare_bgrf_a(const rtt_smatrix *host, int dense_threshold, rtt_prepared_a **out,
                       rtt_offline_a_stats *stats, rtt_bgrf_stats *bgrf);

/* On-disk diagnostics of a layout (new -> old maps; NULL means identity and
 * stats NULL the identity baseline), in the reference's formats
 * (docs/BGRF_V1_PPOPP_ALGORITHM_AND_DATA_FORMAT_CN.md section 9.12):
 * dump_dmma_reorder_plan (src/dmma_reorder.h:3086-3209) writes
 * <prefix>_permutations.csv, <prefix>_unified_stats.csv and
 * <prefix>_A_reordered.mtx; dump_reorder_heatmap (src/main.cu:1114-1246)
 * writes <prefix>_heatmap.csv (bins x bins entry counts before and after)
 * and <prefix>_heatmap_meta.csv. */
int rtt_dump_reorder(const char *prefix, const rtt_smatrix *host, const int *row_new_to_old,
                     const int *inner_new_to_old, const rtt_bgrf_stats *stats);
int rtt_dump_reorder_heatmap(const char *prefix, const rtt_smatrix *host,
                             const int *row_new_to_old, const int *inner_new_to_old, int bins,
                             const rtt_bgrf_stats *stats);

/* Parses one order file (exactly `expected` 0-based decimal tokens, new_id
 * -> old_id, no duplicates or out-of-range ids: src/dmma_reorder.h:2048-2120)
 * into out[expected]. */
int rtt_read_order_file(const char *path, int expected, int *out);

/* Views and metadata of a prepared A. */
int rtt_prepared_a_view(const rtt_prepared_a *a, rtt_device_tiles *view);
int rtt_prepared_a_info(const rtt_prepared_a *a, int *active_rows,
                        int *active_inner, int *identity);
/* Host copy of the row permutation (new -> old) for tile2csr_restore_rows;
 * NULL for the identity layout. */
const int *rtt_prepared_a_row_new_to_old(const rtt_prepared_a *a);
void rtt_destroy_prepared_a(rtt_prepared_a *a);

/* ------------------------------------------------------------------------ */
/* Per-iteration B rebuild (timed in the reference: src/main.cu:2188-2191)   */
/* ------------------------------------------------------------------------ */
typedef struct rtt_dynamic_b rtt_dynamic_b;     /* DmmaDynamicB */

typedef struct rtt_b_update_stats {             /* DmmaBUpdateStats */
    double total_ms;
    double csr_row_permute_ms;
    double csr2tile_ms;
    double csc_ms;
    int path;            /* 0 ordered-CSR merge, 1 tile transpose, 2 CSR transpose,
                            3 independent B CSR (general A*B), 4 values-only update */
} rtt_b_update_stats;

/* gpu_rebuild_dynamic_b_via_permuted_csr (aat = 0, src/gpu_dmma_tiles.h:2275)
 * and gpu_rebuild_dynamic_b_transpose (aat = 1, :2285): B = A (rows mapped by
 * the inner permutation) or B = A^T, as 4x8 column-major tiles with tile
 * CSC.  *inout may hold a previous B, which is rebuilt in place. */
int rtt_rebuild_b(const rtt_prepared_a *a, int aat, rtt_dynamic_b **inout,
                  rtt_b_update_stats *stats);
int rtt_dynamic_b_view(const rtt_dynamic_b *b, rtt_device_tiles *view);
void rtt_destroy_dynamic_b(rtt_dynamic_b *b);

/* gpu_compute_nnz_cub_derived (src/gpu_dmma_tiles.h:2304-2320): number of
 * scalar products, AA: sum_e rowlen(col e); AAT: sum_e coldeg(col e). */
int rtt_compute_nnz_cub(const rtt_prepared_a *a, int aat,
                        unsigned long long *nnz_cub, double *elapsed_ms);

/* ------------------------------------------------------------------------ */
/* General A * B (--b B.mtx, src/main.cu:1554-1566,1960-1972) and the        */
/* values-only B update (docs/BGRF_V1_PPOPP_ALGORITHM_AND_DATA_FORMAT_CN.md  */
/* section 9.10; policy src/dmma_b_values_clear_policy.h)                    */
/* ------------------------------------------------------------------------ */
typedef struct rtt_device_csr rtt_device_csr;   /* DmmaOwnedDeviceCsr
                                                   (src/gpu_dmma_tiles.h:32) */

/* gpu_upload_csr (src/gpu_dmma_tiles.h:2259): H2D copy plus structural
 * validation of an independent B.  Rows with unsorted columns are sorted
 * (stably) on the device; the library keeps where each input entry went, so
 * rtt_device_csr_set_values takes values in the caller's entry order. */
int rtt_upload_csr(const rtt_smatrix *host, rtt_device_csr **out, double *h2d_ms,
                   double *validation_ms);
/* New values for an uploaded CSR whose structure is unchanged: nnz doubles
 * in the entry order of the host CSR passed to rtt_upload_csr. */
int rtt_device_csr_set_values(rtt_device_csr *csr, const double *host_values);
void rtt_destroy_device_csr(rtt_device_csr *csr);

/* gpu_rebuild_dynamic_b_via_permuted_csr with an independent B CSR
 * (src/gpu_dmma_tiles.h:2275, called at src/main.cu:1417-1420): B's rows
 * are taken through A's inner permutation, its columns are kept.  Needs
 * b->rows == A's column count.  The build also records, per B source entry,
 * where its value lands in the tile payload (the source-to-payload map). */
int rtt_rebuild_b_general(const rtt_prepared_a *a, const rtt_device_csr *b,
                          rtt_dynamic_b **inout, rtt_b_update_stats *stats);

/* gpu_compute_nnz_cub_ab (src/gpu_dmma_tiles.h:2295): sum over A's entries
 * of the length of the B row its column selects. */
int rtt_compute_nnz_cub_ab(const rtt_prepared_a *a, const rtt_device_csr *b,
                           unsigned long long *nnz_cub, double *elapsed_ms);

enum rtt_b_values_clear_policy {        /* DmmaBValuesClearPolicy */
    RTT_B_VALUES_ALWAYS_CLEAR = 0,
    RTT_B_VALUES_NOCLEAR = 1,
    RTT_B_VALUES_SAFE_AUTO = 2
};

enum rtt_b_values_clear_reason {        /* DmmaBValuesClearReason */
    RTT_B_VALUES_CLEAR_POLICY_ALWAYS = 0,
    RTT_B_VALUES_NOCLEAR_ELIGIBLE = 1,
    RTT_B_VALUES_AUTO_ELIGIBLE = 2,
    RTT_B_VALUES_FALLBACK_STRUCTURE_UPDATE = 3,
    RTT_B_VALUES_FALLBACK_INVALID_STATE = 4,
    RTT_B_VALUES_FALLBACK_DUPLICATES = 5,
    RTT_B_VALUES_FALLBACK_NONUNIQUE_MAPPING = 6,
    RTT_B_VALUES_FALLBACK_UNINITIALIZED_PAYLOAD = 7,
    RTT_B_VALUES_FALLBACK_NUMERIC_WRITE_CONTRACT = 8,
    RTT_B_VALUES_FALLBACK_INVALID_POLICY = 9,
    RTT_B_VALUES_FALLBACK_INCOMPLETE_MAPPING = 10,
    RTT_B_VALUES_FALLBACK_DENSE_HOLES = 11
};

typedef struct rtt_b_values_decision {  /* DmmaBValuesClearDecision */
    int clear_payload;
    int skip_clear;
    int fallback;
    int reason;                         /* enum rtt_b_values_clear_reason */
} rtt_b_values_decision;

/* dmma_choose_b_values_clear (src/dmma_b_values_clear_policy.h:91-160):
 * host-only gate deciding whether a values-only update may skip clearing
 * the payload.  Returns RTT_ERR_INVALID only for a NULL out. */
int rtt_choose_b_values_clear(int policy, int values_only_update, int dynamic_b_valid,
                              int has_active_duplicates, int active_source_mapping_complete,
                              int active_source_to_payload_injective,
                              int payload_fully_initialized, int every_payload_slot_overwritten,
                              int numeric_b_read_only, rtt_b_values_decision *out);
/* dmma_b_values_clear_policy_name / _reason_name (:43-88). */
const char *rtt_b_values_clear_policy_name(int policy);
const char *rtt_b_values_clear_reason_name(int reason);

/* Values-only B update (section 9.10 of the design document above): the
 * structure of *b was built by rtt_rebuild_b_general from this same
 * rtt_device_csr, and only its values changed (rtt_device_csr_set_values).
 * Clears the payload when the policy gate says so, then writes every mapped
 * slot with the in-order sum of its source entries (duplicates included), so
 * the payload equals what a rebuild computes, bit for bit.  When *b was not
 * built from this CSR's structure, falls back to rtt_rebuild_b_general
 * (decision->reason = RTT_B_VALUES_FALLBACK_INVALID_STATE, stats->path 3).
 * decision may be NULL. */
int rtt_update_b_values(const rtt_prepared_a *a, const rtt_device_csr *b_src, int policy,
                        rtt_dynamic_b *b, rtt_b_update_stats *stats,
                        rtt_b_values_decision *decision);

/* ------------------------------------------------------------------------ */
/* SpGEMM core: symbolic -> allocate -> numeric                              */
/* ------------------------------------------------------------------------ */
enum rtt_lb_mode {               /* --lb-mode (src/main.cu:132-166) */
    RTT_LB_STATIC = 0,
    RTT_LB_DYNAMIC = 1,
    RTT_LB_ADAPTIVE_TAIL = 2
};

/* Numeric kernels of a call, chosen per A tile row.  VECTOR: one workgroup
 * per A tile row adds every scalar product into an LDS (or, for very wide
 * rows, HBM) copy of the row's output with FP64 atomics.  TILE: the row's C
 * structure is written first, then every 8x8 C tile sums its (A tile,
 * B tile) pairs in increasing K on the FP64 matrix cores
 * (v_mfma_f64_4x4x4f64, one instruction per pair), so its values are
 * bit-reproducible.  AUTO sends a tile row to TILE when its A tiles hold at
 * least tile_fill_threshold stored values on average and to VECTOR
 * otherwise, in the same call (the reference's DMMA path multiplies every
 * matched tile pair on tensor cores, src/dmma_spgemm.h:4176-4246; sparse
 * tiles pay there for slots they do not use).  rtt_stats.numeric_path
 * reports TILE or VECTOR when every nonempty row took that path, AUTO for a
 * mix. */
enum rtt_numeric_path {
    RTT_NUMERIC_AUTO = 0,
    RTT_NUMERIC_VECTOR = 1,
    RTT_NUMERIC_TILE = 2
};

/* Arithmetic of the TILE path.  FP64: one FP64 MFMA per tile pair.
 * SPLIT_BF16: the Ozaki scheme on BF16 matrix cores -- every A entry is cut,
 * relative to the largest exponent of its tile row over a batch of at most
 * eight pairs of one C tile, into eight exact 8-bit slices (B entries
 * likewise per tile column); all 64 slice products of a pair run on
 * v_mfma_f32_16x16x32_bf16 with exact FP32 sums (eight instructions per
 * pair), and the slice sums are recombined and accumulated in FP64 on the
 * GPU.  Each entry keeps 64 bits below its row (column) maximum, so
 * |c - c*| <= 2^-62 sum_k (max_row|a| |b_kj| + |a_ik| max_col|b|) plus FP64
 * rounding.  Bit-reproducible like FP64.  Requires the TILE path (AUTO
 * takes it when this is set). */
enum rtt_precision {
    RTT_PRECISION_FP64 = 0,
    RTT_PRECISION_SPLIT_BF16 = 1
};

typedef struct rtt_schedule {    /* DmmaNumericScheduleConfig subset */
    int lb_mode;                 /* enum rtt_lb_mode; default adaptive-tail */
    int materialize_output;      /* export C tiles to the host SMatrix */
    double split_threshold;      /* heavy admission cost (16384) */
    double split_chunk_target;   /* cost per chunk (8192) */
    int split_max_chunks;        /* 8; valid 2..32 */
    double max_heavy_fraction;   /* 0.01; valid 0..1 */
    int tail_record_capacity;    /* 1 << 20 */
    int force_tail_split;        /* split the admitted tiles whatever their share
                                    (skips the heavy-fraction gate) */
    int numeric_path;            /* enum rtt_numeric_path; default AUTO */
    double tile_fill_threshold;  /* AUTO: stored values per A tile for TILE (12) */
    int precision;               /* enum rtt_precision; default FP64 */
} rtt_schedule;

/* Fills *s with the reference defaults (src/main.cu:61-130). */
void rtt_schedule_defaults(rtt_schedule *s);

typedef struct rtt_stats {       /* DmmaSpGemmStats subset */
    double symbolic_ms;          /* candidate + exact structure */
    double allocation_ms;        /* size readback + C allocation */
    double numeric_ms;           /* scheduler + numeric */
    double total_ms;
    double output_copy_ms;       /* D2H export (outside the Core) */
    /* device-event durations of the main kernels, last call */
    double symbolic_count_kernel_ms;
    double symbolic_fill_kernel_ms;
    double numeric_kernel_ms;
    unsigned long long pair_count;     /* (A tile, B tile) products visited */
    unsigned long long output_tiles;   /* 8x8 C tiles */
    unsigned long long output_nnz;     /* nnz(C) */
    unsigned long long heavy_rows;     /* VECTOR: tile rows beyond the 1024-value
                                          LDS budget (larger LDS kernels or HBM
                                          slab) */
    unsigned long long heavy_chunks;   /* TILE, adaptive-tail: chunks the split heavy
                                          C tiles were cut into (0 when not split) */
    int output_materialized;
    int numeric_path;                  /* VECTOR, TILE, or AUTO when rows took both */
    double b_columns_ms;               /* TILE: B's tile columns built for the call */
    unsigned long long mfma_pairs;     /* TILE: tile pairs multiplied: one FP64 MFMA
                                          (4 blocks x 4x4x4, 512 flops) each, or for
                                          SPLIT_BF16 eight BF16 16x16x32 MFMAs */
    int precision;                     /* enum rtt_precision that ran */
    int tail_reason;                   /* TILE, adaptive-tail: tail gate decision --
                                          1 split, 4 no heavy tile, 5 record capacity
                                          exceeded, 6 heavy fraction exceeded, 7
                                          workspace limit exceeded; 0 no admission */
    unsigned long long tail_records;   /* C tiles admitted as heavy */
} rtt_stats;

/* dmma_tilespgemm (src/dmma_spgemm.h:6490): C = A * B with A 8x4 and B 4x8
 * tiles; C is produced as 8x8 tiles.  When sched->materialize_output is set,
 * *output receives the host tile-CSR (m, n, tilem, tilen, numtile, nnz,
 * tile_ptr, tile_columnidx, tile_nnz, tile_csr_Ptr, tile_csr_Col,
 * tile_csr_Value), in canonical order, exactly as dmma_copy_output_to_host
 * (src/dmma_spgemm.h:6399-6482) lays it out.  Returns RTT_ERR_OVERFLOW when
 * the tile count or nnz(C) exceeds INT_MAX and export was requested. */
int rtt_tilespgemm(const rtt_prepared_a *a, const rtt_dynamic_b *b,
                   rtt_smatrix *output, rtt_stats *stats,
                   const rtt_schedule *sched);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Recomputes the values of the last C produced on `a`
 * (by rtt_tilespgemm with B = b, every row on the TILE path, b not rebuilt
 * since; a values-only rtt_update_b_values in between is what it is for)
 * into c_values, nnz(C) doubles of device memory the caller owns, in C's
 * value order, with `precision` (enum rtt_precision).  Allocates nothing:
 * its scratch (8 row-queue heads and a pair counter, 40 bytes) lives in a
 * fixed 256-byte device array per GPU, cleared by a kernel on every call.
 * Calls that overlap on one GPU race on that array.  Heavy tiles run whole
 * (no tail split).  Queues the kernels on the library stream and returns;
 * rtt_sync() waits for them. */
int rtt_tile_values_noalloc(const rtt_prepared_a *a, const rtt_dynamic_b *b, int precision,
                            double *c_values);

/* Tail queue of the TILE path under adaptive-tail (the reference's
 * admission, gate, chunk descriptors and chunk workspace reduction,
 * src/dmma_spgemm.h:2628-2776, 4902-5254, 6003-6049, 6981-7209): a C tile
 * whose K-merge costs scans + 8 matches > split_threshold (FP32) is admitted
 * with chunks = ceil(cost / split_chunk_target) clamped to
 * [2, split_max_chunks] and to its scans; when the gate passes, each chunk
 * (a contiguous range of the tile's B column) is multiplied by its own wave
 * into a 64-value workspace slot and the chunks are summed in order.  This
 * copies the admitted records of the last call on `a` (slot order): tile
 * index, merge scans, matches, chunks; *count receives how many. */
int rtt_last_tail_records(const rtt_prepared_a *a, int max_records, long long *task,
                          int *scans, int *matches, int *chunks, int *count);

/* Device view of the last C produced on handle `a` (valid until the next
 * rtt_tilespgemm on `a`): 64-bit offsets, 64-bit tile masks. */
typedef struct rtt_device_c {
    int64_t tilem;
    int64_t tilen;
    int64_t num_tiles;
    int64_t nnz;
    const int64_t *tile_ptr;     /* tilem + 1 */
    const int *tile_col;         /* num_tiles */
    const uint64_t *tile_mask;   /* num_tiles */
    const int64_t *tile_nnz;     /* num_tiles + 1 */
    const uint8_t *tile_csr_ptr; /* 8 * num_tiles */
    const uint8_t *tile_csr_col; /* nnz */
    const double *tile_csr_val;  /* nnz */
} rtt_device_c;
int rtt_last_c_view(const rtt_prepared_a *a, rtt_device_c *view);

/* No counterpart in the reference (its calls run on CUDA's default stream):
 * the HIP stream (a hipStream_t) every call of this library queues its work
 * on, after rtt_init.  A caller orders its own work after the library's
 * with it, or captures an allocation-free call (rtt_tile_values_noalloc) on
 * it into a HIP graph with hipStreamBeginCapture / hipStreamEndCapture and
 * replays the graph. */
int rtt_stream(void **stream);

/* Copies device memory to host (test and tool helper). */
int rtt_memcpy_d2h(void *dst, const void *src, size_t bytes);

/* Per row of the last C on `a` (rows in A's tiled order, a's row count
 * entries each): its number of entries and the sum of its values, computed on
 * the device from C's tile structure (checking helper for products too large
 * to download). */
int rtt_last_c_row_stats(const rtt_prepared_a *a, int64_t *row_nnz, double *row_sum);

/* ------------------------------------------------------------------------ */
/* Autotuning (no reference counterpart: the reference fixes its schedule   */
/* through driver flags, src/main.cu:61-166)                                 */
/* ------------------------------------------------------------------------ */
#define RTT_TUNE_MAX_CANDIDATES 16
typedef struct rtt_tune_report {
    int candidates;
    int chosen;                                  /* index of the fastest */
    double ms[RTT_TUNE_MAX_CANDIDATES];          /* median call time */
    int numeric_path[RTT_TUNE_MAX_CANDIDATES];
    double fill[RTT_TUNE_MAX_CANDIDATES];
    int lb_mode[RTT_TUNE_MAX_CANDIDATES];
} rtt_tune_report;

/* Times rtt_tilespgemm(a, b) (symbolic + allocate + numeric, no export)
 * `repeats` times after one warm-up for each candidate schedule -- the
 * per-row numeric routing threshold (VECTOR, AUTO at 24/16/12/8 stored values
 * per A tile, TILE) and the slab rows' queue assignment -- and writes the
 * fastest into *sched (other fields kept).  The tile shapes are fixed by the
 * format (8x4 A, 4x8 B, 8x8 C).  report may be NULL. */
int rtt_autotune(const rtt_prepared_a *a, const rtt_dynamic_b *b, int repeats,
                 rtt_schedule *sched, rtt_tune_report *report);

/* One problem of a block-diagonal batch (rtt_block_diag) from the last C on
 * `a`: the m x n block whose first row and column are *row_offset and
 * *col_offset (multiples of 8), as a host tile-CSR SMatrix with the block's
 * own tile columns (no reference counterpart: the reference runs one
 * product per call). */
int rtt_batch_extract(const rtt_prepared_a *a, const int64_t *row_offset,
                      const int64_t *col_offset, int m, int n, rtt_smatrix *out);

/* ------------------------------------------------------------------------ */
/* Output conversion (host, OpenMP)                                          */
/* ------------------------------------------------------------------------ */
/* tile2csr (src/tile2csr.h:232-329): fills rowpointer/columnindex/value and
 * nnz of *m from its 8x8 tile-CSR fields. canonical != 0 skips the row sort. */
int rtt_tile2csr(rtt_smatrix *m, int canonical);
/* tile2csr_restore_rows (src/tile2csr.h:337-523): the same with rows mapped
 * from the reordered space through row_new_to_old[0..permutation_size). */
int rtt_tile2csr_restore_rows(rtt_smatrix *m, const int *row_new_to_old,
                              int permutation_size, int canonical);

/* ------------------------------------------------------------------------ */
/* Seeded synthetic inputs (splitmix64; identical on every host)             */
/* ------------------------------------------------------------------------ */
/* 2D 5-point Laplacian on g x g (row = x + g*y): diagonal 4.0,
 * off-diagonals -1.0 + U[-pert, pert]. */
int rtt_gen_laplace2d(int g, double pert, uint64_t seed, rtt_smatrix *out);
/* 3D 27-point stencil on g^3 (row = x + g*y + g*g*z), values U[lo, hi]. */
int rtt_gen_stencil27(int g, double lo, double hi, uint64_t seed,
                      rtt_smatrix *out);
/* R-MAT with 2^scale rows, edge_factor * 2^scale draws, probabilities
 * (a, b, c, 1-a-b-c), duplicates removed, values U[-1, 1], rows sorted. */
int rtt_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                 uint64_t seed, rtt_smatrix *out);
/* Random m x n matrix with `per_row` distinct sorted columns per row drawn
 * uniformly, values U[-1, 1] (test helper). */
int rtt_gen_random(int m, int n, int per_row, uint64_t seed, rtt_smatrix *out);

/* Block-structured FEM-like matrix: `nodes` points uniform in the unit cube,
 * each joined to its `neighbours` nearest points (exact, ties by index),
 * symmetrised, plus itself; every joined pair (i, j) is a dense 4x4 block
 * (rows 4i.., columns 4j..) whose entries have magnitudes 10^U[-3,3] and
 * random signs.  spatial_order != 0 numbers the nodes along a Morton curve
 * (a mesher's locality numbering); 0 keeps the draw order. */
int rtt_gen_fem_blocks(int nodes, int neighbours, int spatial_order, uint64_t seed,
                       rtt_smatrix *out);
/* Banded n x n matrix: row i holds min(per_row, band width) distinct columns
 * drawn uniformly from [i - bandwidth, i + bandwidth], plus the diagonal;
 * values U[-1, 1]. */
int rtt_gen_banded(int n, int bandwidth, int per_row, uint64_t seed, rtt_smatrix *out);
/* Shapes of a batch of banded problems: rows round(exp(N(ln median, sigma)))
 * clipped to [min_rows, max_rows], bandwidth U{min_bw..max_bw}, entries per
 * row U{min_per_row..max_per_row}. */
int rtt_gen_banded_batch_shapes(int problems, double median_rows, double sigma, int min_rows,
                                int max_rows, int min_bw, int max_bw, int min_per_row,
                                int max_per_row, uint64_t seed, int *rows, int *bandwidth,
                                int *per_row);
/* 3D 7-point Poisson operator on g^3 cells (row = x + g*y + g*g*z) with cell
 * coefficients 10^U[-decades, decades]: a face between cells i and j weighs
 * (k_i + k_j) / 2, a boundary face adds k_i to the diagonal (Dirichlet). */
int rtt_gen_poisson7_jumps(int g, double decades, uint64_t seed, rtt_smatrix *out);
/* Smoothed-aggregation prolongator P = (I - omega D^-1 A) T of a g^3 grid
 * operator A, with T the 2x2x2 cell aggregation onto ceil(g/2)^3 aggregates. */
int rtt_gen_aggregation_prolongator(const rtt_smatrix *a, int g, double omega,
                                    rtt_smatrix *out);
/* Host transpose of a CSR matrix (rows increasing within each column). */
int rtt_csr_transpose(const rtt_smatrix *in, rtt_smatrix *out);

/* ------------------------------------------------------------------------ */
/* Batches                                                                   */
/* ------------------------------------------------------------------------ */
/* Block-diagonal concatenation of `count` CSR matrices, every block starting
 * at row and column offsets that are multiples of `align` (8 keeps every 8x4
 * A, 4x8 B and 8x8 C tile inside one problem).  row_offsets / col_offsets
 * (count + 1 entries each, may be NULL) receive the block starts and the
 * padded totals.  The product of two such concatenations with matching
 * offsets is the concatenation of the problems' products, so one SpGEMM call
 * runs the whole batch: one launch per stage over every problem's tiles. */
int rtt_block_diag(int count, const rtt_smatrix *const *mats, int align, rtt_smatrix *out,
                   int64_t *row_offsets, int64_t *col_offsets);

#ifdef __cplusplus
}
#endif

#endif /* RTT_SPGEMM_H_ */
