/*
 * CPU oracle for the tiled FP64 SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Every function restates one piece of the reference's CPU semantics and
 * cites the reference file:line it follows.  All array outputs are allocated
 * by the caller (two-phase count/fill APIs) so the functions map directly onto
 * numpy buffers through ctypes.
 *
 * Parity pinning of each function is described in oracle/README.md.
 */
#ifndef RTT_ORACLE_H_
#define RTT_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Matrix Market reader (src/mmio_highlevel.h:593-759) -------------- */
/* Phase 1: banner + size line.  Returns 0 or a negative reader code. */
int orc_mmio_info(const char *path, int *m, int *n, int64_t *nnz_file,
                  int *is_symmetric);
/* Phase 2: full read into caller buffers sized from the expanded nnz, which
 * the caller obtains from orc_mmio_expanded_nnz. */
int64_t orc_mmio_expanded_nnz(const char *path);
int orc_mmio_read(const char *path, int *row_ptr, int *col_idx,
                  double *values);

/* ---- nnzCub (src/gpu_dmma_tiles.h:744-775, 2052-2101) ------------------ */
uint64_t orc_nnz_cub(int m, int n, const int *row_ptr, const int *col_idx,
                     int aat);

/* ---- Host transpose (src/utils.h:158-197) ------------------------------ */
void orc_transpose(int m, int n, const int *row_ptr, const int *col_idx,
                   const double *values, int *t_row_ptr, int *t_col_idx,
                   double *t_values);

/* ---- Row SPA SpGEMM, C = A * B ------------------------------------------
 * Structure: bit-flag SPA (src/spgemm_serialref_spa_new.h:7-105), output
 * columns ascending.  Values: dense-row SPA accumulated in A-row order then
 * B-row order (src/external/cusparse/spgemm_serialref_spa.h:116-228).
 * Optional outputs (may be NULL): compensated double-double sum (hi, lo) and
 * sum of |a|*|b| for forward-error scaling.                                 */
int64_t orc_spgemm_count(int m_a, int n_a, const int *rp_a, const int *ci_a,
                         int m_b, int n_b, const int *rp_b, const int *ci_b,
                         int64_t *rp_c);
int orc_spgemm_fill(int m_a, int n_a, const int *rp_a, const int *ci_a,
                    const double *v_a, int m_b, int n_b, const int *rp_b,
                    const int *ci_b, const double *v_b, const int64_t *rp_c,
                    int *ci_c, double *v_c, double *v_c_hi, double *v_c_lo,
                    double *abs_c);

/* ---- Legacy 16x16 tile SpGEMM (src/spgemm-cpu.h:3-455) -------------------
 * C as CSR; ci_c == NULL fills rp_c only.  Returns nnz(C) or -1.          */
int64_t orc_legacy_tile_spgemm(int m_a, int n_a, const int *rp_a, const int *ci_a,
                               const double *v_a, int m_b, int n_b, const int *rp_b,
                               const int *ci_b, const double *v_b, int64_t *rp_c, int *ci_c,
                               double *v_c);

/* ---- Hybrid tile format (src/dmma_tiles.h:163-440) ----------------------
 * counts[0]=num_tiles counts[1]=payload counts[2]=dense counts[3]=sparse
 * Returns 0 on success, -1 on invalid input (validate_csr semantics).     */
int orc_hybrid_tiles_count(int m, int n, const int *row_ptr,
                           const int *col_idx, int tile_rows, int tile_cols,
                           int col_major, int dense_threshold,
                           int64_t *counts);
int orc_hybrid_tiles_fill(int m, int n, const int *row_ptr,
                          const int *col_idx, const double *values,
                          int tile_rows, int tile_cols, int col_major,
                          int dense_threshold, int build_csc,
                          int *tile_row_ptr, int *tile_col_idx,
                          int *value_offsets, uint32_t *masks,
                          double *payload, int *tile_col_ptr,
                          int *tile_row_idx, int *csc_tile_ids);

/* ---- 8x8 output tiles of a CSR matrix C ---------------------------------
 * The tile-CSR the numeric writer exports (src/dmma_spgemm.h:4261-4281,
 * 6399-6482): tile_ptr, tile_columnidx, tile_nnz (value offsets), 8 local
 * row offsets per tile, local columns and values.  Rows must be sorted.    */
int64_t orc_c_tiles_count(int m, int n, const int64_t *rp_c,
                          const int *ci_c, int64_t *tile_ptr);
int orc_c_tiles_fill(int m, int n, const int64_t *rp_c, const int *ci_c,
                     const double *v_c, const int64_t *tile_ptr,
                     int *tile_col, int64_t *tile_nnz,
                     uint8_t *tile_csr_ptr, uint8_t *tile_csr_col,
                     double *tile_csr_val, uint64_t *tile_mask);

/* ---- tile2csr / tile2csr_restore_rows (src/tile2csr.h:232-523) ----------
 * Pass row_new_to_old = NULL for plain tile2csr.  Phase 1 fills row_ptr
 * (m+1) and returns nnz, or -1 when the reference would return false.     */
int64_t orc_tile2csr_count(int m, int n, int tilem, const int *tile_ptr,
                           const int *tile_columnidx, const int *tile_nnz,
                           const uint8_t *tile_csr_ptr,
                           const uint8_t *tile_csr_col,
                           const int *row_new_to_old, int permutation_size,
                           int64_t *row_ptr);
int orc_tile2csr_fill(int m, int n, int tilem, const int *tile_ptr,
                      const int *tile_columnidx, const int *tile_nnz,
                      const uint8_t *tile_csr_ptr, const uint8_t *tile_csr_col,
                      const double *tile_csr_val, const int *row_new_to_old,
                      int permutation_size, int canonical,
                      const int64_t *row_ptr, int *col_idx, double *values);

/* ---- Tail-queue admission (src/dmma_tile_tail_queue.h:151-292) --------- */
int orc_tail_chunk_count(int scans, int matches, double admission_threshold,
                         double chunk_target, int max_chunks,
                         double scan_weight, double match_weight);
/* Plans over n tasks given as (scans, matches, output_nnz) triples.
 * out[0]=reason out[1]=heavy_tasks out[2]=chunks out[3]=workspace bytes;
 * heavy_ids/heavy_chunks (capacity n) receive the admitted tasks.          */
int orc_plan_tail_queue(int n, const int *scans, const int *matches,
                        const int *output_nnz, int enabled,
                        double admission_threshold, double chunk_target,
                        int max_chunks, double max_heavy_fraction,
                        int record_capacity, uint64_t workspace_limit,
                        int workspace_reduction, int64_t *out,
                        int *heavy_ids, int *heavy_chunks);

#ifdef __cplusplus
}
#endif

#endif
