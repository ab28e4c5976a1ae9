// Internal device-side data structures and launcher declarations.  Only the
// library's own .hip files include this header.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../rtt_common.hpp"

namespace rtt {

constexpr int kTileM = 8;   // A tile rows / C tile rows
constexpr int kTileK = 4;   // A tile cols / B tile rows
constexpr int kTileN = 8;   // B tile cols / C tile cols
constexpr int kInputElems = 32;
// B tiles with at least this many entries, or with entries in more than one
// row k, are "heavy"; the others are "light" and lie in one row k.  Every B
// row keeps its heavy tiles first, then its light tiles grouped by k, so
// that pairs of like cost can be taken together and an A tile with entries
// in one column k meets only the light tiles of row k.  A light tile's at
// most two values fill the first value part of its record.
constexpr int kHeavyMin = 3;
constexpr int kClassSlots = 8;   // class slots per B row: heavy, light k = 0..3, unused

// Per-B-tile record read by the numeric kernel: the packed meta and the
// first kRecVals payload values, kRecParts 16-byte parts.  B keeps its
// records as part planes (part i of tile t at planes[i * n + t], n tiles), so
// that the lanes of a wave, which take consecutive tiles, read each part as
// one contiguous stretch.  A's offline transpose source keeps whole records
// (one tile's parts contiguous).
constexpr int kRecVals = 8;
struct alignas(16) BRecord {
    int4 meta;
    double v[kRecVals];
};
constexpr int kRecParts = 1 + kRecVals / 2;
static_assert(kRecVals % 2 == 0 && sizeof(BRecord) == 16 * kRecParts, "record layout");

// Checks a HIP call; on failure records the error and returns RTT_ERR_DEVICE.
#define RTT_HIP(call)                                                            \
    do {                                                                         \
        hipError_t rtt_err_ = (call);                                            \
        if (rtt_err_ != hipSuccess)                                              \
            return ::rtt::fail(RTT_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,   \
                               hipGetErrorString(rtt_err_), __FILE__, __LINE__); \
    } while (0)

#define RTT_TRY(expr)                  \
    do {                               \
        int rtt_rc_ = (expr);          \
        if (rtt_rc_ != RTT_OK)         \
            return rtt_rc_;            \
    } while (0)

// Grow-only device buffer.  ensure() reallocates only when the requested
// capacity exceeds what is held, so per-iteration rebuilds reuse memory.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <typename T>
    T *as() const { return static_cast<T *>(ptr); }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

// Grow-only pinned host buffer (hipHostMalloc) for small readbacks.
struct PinnedBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    template <typename T>
    T *as() const { return static_cast<T *>(ptr); }
    ~PinnedBuf();
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
};

// Device CSR with rows sorted by column (duplicates adjacent).
struct DevCsr {
    int rows = 0, cols = 0;
    int64_t nnz = 0;
    DevBuf row_ptr, col_idx, values;
    // When the upload sorted rows: input_of[x] = input position of sorted entry x.
    bool reordered = false;
    DevBuf input_of;
    uint64_t structure_id = 0;   // unique per upload; kept by value updates
};

// Hybrid tile matrix on the device (the DmmaDeviceTiles fields) plus the
// packed per-tile record the SpGEMM kernels read: meta[t] = {tile column,
// mask, per-k byte sets (byte k = stride-4 bits k, k+4, ..., k+28 of the
// mask), payload offset | dense bit 31}.
struct DevTiles {
    int rows = 0, cols = 0, tile_rows = 0, tile_cols = 0;
    int tile_row_count = 0, tile_col_count = 0;
    int64_t num_tiles = 0, payload = 0, dense = 0, sparse = 0;
    bool has_csc = false;       // tile_col_ptr / tile_row_idx / csc_tile_ids valid
    bool cols_ready = false;    // col_ptr / col_ent valid (B built from a CSR)
    bool rec_values_ready = true;   // B from a CSR: values_cm and record planes 1.. valid
    bool arrays_ready = true;   // tile_col_idx / masks / value_offsets / values valid
    DevBuf tile_row_ptr, tile_col_idx, value_offsets, masks, values;
    DevBuf tile_col_ptr, tile_row_idx, csc_tile_ids;
    DevBuf tile_row_of;   // tile -> tile row (A: used by the B transpose)
    DevBuf meta;          // int4 per tile
    DevBuf records;       // B only: kRecParts planes of int4, num_tiles entries each
    DevBuf a_brange;      // B only: per A tile, {first B tile, heavy, count} of B row K
    DevBuf a_blight;      // B only: per A tile, {first, count} of the light tiles it can meet
    DevBuf pair_total;    // B only: u64, sum over A tiles of their B row's tile count
    DevBuf row_groups;    // B only: start of each (row, class) run, kClassSlots per row
    DevBuf col_ptr;       // B only: tile-column pointer of col_ent (tile_col_count + 1)
    DevBuf col_ent;       // B only: per tile column, its tiles in tile-row order as
                          // {tile row K, mask, per-k byte sets, offset | dense bit}
    DevBuf values_cm;     // each tile's values with the stride-4 bit groups in order:
                          // A (8x4): column-major; B (4x8): k-major (row k's entries
                          // contiguous).  Packed from the tile's offset.  For B built as
                          // A^T this is A's array (payload_kmajor()).

    DevBuf transpose_src; // A only: per tile, the BRecord of the B = A^T tile it becomes
    int max_col_tiles = -1;   // A only: most tiles in one tile column (-1: not counted)
    DevBuf row_payload_off;   // A only: payload offset of each tile row (tile_row_count + 1)
    // B built as A^T at tile level: meta offsets index A's payload, and
    // sorted_ids[b] is the A tile behind B tile b.
    const DevTiles *transpose_of = nullptr;
    DevBuf sorted_ids;
    // B built from a CSR with a source map: src_slot[x] = {values slot,
    // values_cm slot} of source entry x, {-1, -1} outside the tiled extent.
    // Duplicate entries share a slot and are adjacent.
    DevBuf src_slot;
    uint64_t map_source_id = 0;  // DevCsr::structure_id the map belongs to (0: none)
    int64_t map_nnz = 0;
    bool map_dups = false;       // some slot has more than one source entry
    const double *payload_values() const
    {
        return transpose_of ? transpose_of->values.as<double>() : values.as<double>();
    }
    const double *payload_kmajor() const
    {
        return transpose_of ? transpose_of->values_cm.as<double>() : values_cm.as<double>();
    }
    void view(rtt_device_tiles *v) const;
};

// Last C produced on a prepared-A handle (64-bit offsets).
struct DevC {
    int64_t tilem = 0, tilen = 0, num_tiles = 0, nnz = 0;
    int rows = 0, cols = 0;
    DevBuf tile_ptr, tile_col, tile_mask, tile_nnz, tile_csr_ptr, tile_csr_col,
        tile_csr_val;
};

// Scratch of the SpGEMM core, reused across calls.
struct SpgemmWork {
    DevBuf row_tiles, row_nnz, row_level;   // per A tile row
    DevBuf scan_tmp, counters, l3_rows, num_rows, lds_rows, tiny_rows, mid_rows, slab;
    PinnedBuf host_counters;
    DevBuf stage_j, stage_m;                // staged sorted tile lists
    DevBuf c_row_nnz_off;                   // per tile row, int64
    DevBuf queue;                           // work-queue heads of the slab kernels
    DevBuf row_cost, row_cost2, rows_sorted;   // slab rows, heaviest first
    DevBuf trace_start, trace_end, trace_cu;   // numeric timeline samples
    // Slab rows' (column, mask) lists between the symbolic passes: entries
    // allocated, and the most the previous calls needed (grown at the next
    // call's start, never while a fill pass may still read the buffer).
    DevBuf slab_stage_j, slab_stage_m, slab_row_off;
    uint64_t slab_stage_cap = 0, slab_stage_need = 0;
    DevBuf tile_queue;                      // row-queue heads of the tile-pair kernels
    DevBuf tile_pairs;                      // u64: pairs the tile-pair kernel multiplied
    DevBuf tile_rows;                       // u64 x 2: rows counted for routing
    // Tail queue of the tile-pair path: records, state, chunk prefix, heavy
    // tile bitset and the chunks' partial tiles.
    DevBuf tail_records, tail_state, tail_chunk_base, tail_heavy, tail_workspace;
};

// Scratch of the tile builders, reused across calls.
struct BuildWork {
    DevBuf row_tiles, row_payload, row_dense, scan_tmp, counters;
    DevBuf keys_in, keys_out, ids_in, ids_out, spans;
    DevBuf meta2, records2;                     // heavy-first reordering
    DevBuf tcsr_row_ptr, tcsr_col, tcsr_val;   // transposed CSR (reordered AAT)
    DevBuf flag;                                // one int (duplicate detection)
    DevBuf staging;                             // uploaded values before the gather
    DevBuf totals;                              // tile build: tiles, payload, dense
};

struct Events {
    hipEvent_t ev[16] = {};
    bool ready = false;
    int create();
    ~Events();
};

// Library-wide device context (one device per process).
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    int num_cus = 0;
};
Context &ctx();
int require_ctx();

// ---- CSR / tile builders (build.hip) ------------------------------------
// Uploads and validates a host CSR; sorts rows by column when needed.
int upload_csr(const rtt_smatrix &host, DevCsr *out, BuildWork *work,
               double *h2d_ms, double *validate_ms, double *sort_ms,
               int *rows_were_sorted);
// Builds hybrid tiles from a row-sorted CSR.  row_map (device, may be null)
// maps tiled row -> source row for rows < active_rows.
int build_tiles(const DevCsr &src, const int *row_map, int active_rows,
                int active_cols, int logical_rows, int logical_cols,
                bool col_major, int dense_threshold, bool build_csc,
                DevTiles *out, BuildWork *work, bool want_source_map = false);
// Values-only update of a B built with a source map from `src`: optional
// clear, in-order sums of each slot's source entries into values and
// values_cm, then the value parts of the records.
int update_b_values(const DevCsr &src, bool clear, DevTiles *b);
// Uploaded values in the caller's entry order -> src->values (sorted order).
int set_csr_values(DevCsr *csr, const double *host_values, BuildWork *work);
int nnz_cub_ab(const DevCsr &a, const DevCsr &b, unsigned long long *out, BuildWork *work);
// B = A^T at tile level (A tile (J,K) row-major == B tile (K,J) col-major).
// Builds B's tile row pointer and meta only; materialize_tiles() fills the
// remaining DmmaDeviceTiles arrays when a caller asks for them.
int transpose_tiles(const DevTiles &a, int b_cols, DevTiles *b, BuildWork *work);
// Completes the standard arrays (and the tile CSC) of a B built lazily.
int materialize_tiles(DevTiles *t, BuildWork *work);
// Transposed CSR of `src` (stable: rows ascending within each column).
int transpose_csr(const DevCsr &src, DevCsr *out, BuildWork *work);
int nnz_cub(const DevCsr &csr, bool aat, unsigned long long *out, BuildWork *work);
int fill_tile_row_of(DevTiles *t);
// Fills t->values_cm from t->meta and t->values.
int build_colmajor_values(DevTiles *t);
// Fills t->transpose_src (A tiles, offline; needs tile_row_of) and
// t->row_payload_off.
int build_transpose_source(DevTiles *t, BuildWork *w);
// b->a_brange[t] = {B.tile_row_ptr[K], heavy tiles, count} and b->a_blight[t]
// for every A tile t in column K.
int build_pair_extents(const DevTiles &a, DevTiles *b);
// Builds b->col_ptr / b->col_ent (tile columns, tile rows ascending) of a B
// built from a CSR; no-op for B = A^T at tile level, whose columns are A's
// tile rows (b_columns below).
int build_b_columns(DevTiles *b, BuildWork *w);
// Reorders b->meta of a B built from CSR into class order within every row
// (stable), writes it as record plane 0 and fills b->row_groups.  The record
// value planes and the k-major payload are left to ensure_record_values.
int order_by_class(DevTiles *b, BuildWork *w);
// Builds values_cm and the record value planes of a B built from a CSR when
// they are missing (the LDS-accumulator numeric kernels read them; the
// symbolic and tile-pair kernels do not).
int ensure_record_values(DevTiles *b);

// ---- tile-pair numeric path (tiletask.hip) ---------------------------------
// B's tile columns as the tile-pair kernels read them: ptr[J]..ptr[J+1] index
// ent, entries {tile row K, mask, per-k bytes, payload offset | dense bit} in
// increasing K, payload `vals` (each tile column-major).
struct TileColumns {
    const int *ptr = nullptr;
    const int4 *ent = nullptr;
    const double *vals = nullptr;
};
TileColumns b_columns(const DevTiles &b);
// Values of the C tiles of every tile row whose A tiles hold at least `fill`
// stored values on average (0: every row), from C's finished structure:
// precision 0 on FP64 matrix cores, 1 by BF16 slices (Ozaki scheme) with
// FP64 recombination.  Deterministic: each tile sums its pairs in
// increasing K.
struct TileOpts;
int tile_numeric(const DevTiles &a, const TileColumns &bc, const DevC &c, const TileOpts &opt,
                 SpgemmWork *w, hipStream_t s);
// C's values for C's existing structure, every row on the tile-pair kernel,
// written to c_values; the row-queue heads and pair counter live in the
// caller's `scratch` (device memory of scratch_bytes).  Launches kernels
// only: no allocation, no host synchronisation; no tail split.
int tile_values_noalloc(const DevTiles &a, const TileColumns &bc, const DevC &c, int precision,
                        void *scratch, size_t scratch_bytes, double *c_values, hipStream_t s);
// Gate decision, admitted records and chunks of the last tail admission.
int tail_stats(SpgemmWork *w, int *reason, unsigned long long *records,
               unsigned long long *chunks);
// The admitted records (slot order): tile, merge scans and matches, chunks.
int tail_records(SpgemmWork *w, int max_records, long long *task, int *scans, int *matches,
                 int *chunks, int *count);
// Tile rows of A the tile-pair kernel would take at `fill`, and the tile
// rows holding any tile (one readback).
int count_tile_rows(const DevTiles &a, float fill, DevBuf *scratch, int64_t *tile_rows,
                    int64_t *nonempty_rows);
// Per-row numeric routing of one call: rows whose A tiles hold at least
// `fill` stored values on average take the tile-pair kernel in `precision`.
// With `tail`, C tiles whose K-merge costs more than `threshold` (FP32
// scans + 8 matches) are split into chunks (the adaptive-tail queue).
struct TileOpts {
    float fill = 0.f;
    int precision = 0;
    TileColumns bc;
    bool tail = false;
    float threshold = 16384.f, target = 8192.f;
    int max_chunks = 8;
    double max_heavy_fraction = 0.01;
    int force = 0;
    int record_capacity = 1 << 20;
};

// ---- SpGEMM core (spgemm.hip) ---------------------------------------------
struct CoreTimes {
    double symbolic_ms = 0, allocation_ms = 0, numeric_ms = 0;
    double sym_count_kernel_ms = 0, sym_fill_kernel_ms = 0, numeric_kernel_ms = 0;
    unsigned long long pairs = 0, heavy_rows = 0;
    unsigned long long mfma_pairs = 0;   // tile path: (A tile, B tile) products on MFMA
    int tail_reason = 0;                 // tile path: tail gate (0: no admission ran)
    unsigned long long tail_records = 0, tail_chunks = 0;
};
// lb_mode (enum rtt_lb_mode) assigns the rows of the slab kernels, which
// hold one workgroup per row for many rows: static hands row i to
// workgroup i mod grid; dynamic takes rows from a work queue in list order;
// adaptive-tail takes them from the queue heaviest first (most pairs).
// tile == nullptr runs the LDS-accumulator kernels for every row; otherwise
// the rows tile->fill selects have their structure written first and their
// values computed by the tile-pair kernel, the others as before.
int spgemm_core(const DevTiles &a, const DevTiles &b, DevC *c, SpgemmWork *work,
                Events *events, CoreTimes *times, int lb_mode = RTT_LB_ADAPTIVE_TAIL,
                const TileOpts *tile = nullptr);

// ---- scans (scan.hip) -------------------------------------------------------
// Exclusive sum of n int64 values into out[0..n] (out[n] = total).
int exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, DevBuf *tmp,
                       hipStream_t s);

inline unsigned int blocks_for(int64_t n, int threads)
{
    return static_cast<unsigned int>((n + threads - 1) / threads);
}

} // namespace rtt
