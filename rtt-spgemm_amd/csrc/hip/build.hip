// CSR upload/validation and the CSR -> hybrid tile builders.
//
// Tiles are formed without a global sort: rows are kept column-sorted (they
// are sorted once at upload when they arrive unsorted), so the tiles of one
// tile row are the ordered merge of its TR row streams.  A group of TR lanes
// owns one tile row; each lane walks one scalar row.  The group repeatedly
// takes the minimum tile column among its lanes' heads, ORs the lanes' bits
// into the tile mask and writes the tile.  This replaces the reference's
// 64-bit key radix sort + reduce_by_key (src/gpu_dmma_tiles.h:1117-1395) and
// its 4-lane ordered-warp B path (src/gpu_dmma_tiles.h:313-560).
#include <atomic>
#include <hipcub/hipcub.hpp>

#include <climits>

#include "device.hpp"
#include "internal.hpp"

namespace rtt {

namespace {

constexpr int kThreads = 256;

__global__ void set_i64_kernel(int64_t *p, int64_t n, int64_t v)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ void fill_int_kernel(int *p, int64_t n, int v)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ void iota_kernel(int *p, int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = static_cast<int>(i);
}

__global__ void fill_f64_kernel(double *p, int64_t n, double v)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// out[x] = in[idx[x]]
__global__ void gather_f64_kernel(int64_t n, const int *__restrict__ idx,
                                  const double *__restrict__ in, double *__restrict__ out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = in[idx[i]];
}

// Position of physical slot `phys` in the tile's values_cm order: stride-4
// group k = phys & 3 first, then by phys >> 2 within the group.
__device__ __forceinline__ int kmajor_slot(uint32_t mask, int phys)
{
    const int k = phys & 3;
    const uint32_t lower_groups = k == 0 ? 0u : (k == 1 ? 0x11111111u
                                                : (k == 2 ? 0x33333333u : 0x77777777u));
    const uint32_t group = mask & (0x11111111u << k);
    return __popc(mask & lower_groups) + __popc(group & ((1u << phys) - 1u));
}

// flags: 1 row_ptr not monotone / wrong ends, 2 column out of range,
//        4 some row not column-sorted (duplicates allowed).
__global__ void validate_csr_kernel(int rows, int cols, int64_t nnz,
                                    const int *__restrict__ rp,
                                    const int *__restrict__ ci, int *flags)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0 && (rp[0] != 0 || rp[rows] != nnz))
        atomicOr(flags, 1);
    if (i < rows) {
        const int b = rp[i], e = rp[i + 1];
        // Bounds first: a row outside [0, nnz] is flagged, never walked.
        if (b < 0 || e < b || static_cast<int64_t>(e) > nnz) {
            atomicOr(flags, 1);
        } else {
            for (int x = b + 1; x < e; ++x)
                if (ci[x] < ci[x - 1]) {
                    atomicOr(flags, 4);
                    break;
                }
        }
    }
    if (i < nnz) {
        const int c = ci[i];
        if (c < 0 || c >= cols)
            atomicOr(flags, 2);
    }
}

// Columns of its 64 rows a wave stages in LDS before the walk (when the rows
// are contiguous in the CSR and fit), so that the lanes' entry-by-entry walk
// reads LDS instead of issuing one dependent global load per entry.  The
// stage is sized per launch (dynamic LDS, at most kBuildStage columns per
// wave) from the operand's average row length: a fixed 3072-column stage
// held the kernel to 3 waves per SIMD on every operand.
constexpr int kBuildStage = 3072;

int build_stage_cap(const DevCsr &src, bool row_map)
{
    if (row_map || src.rows <= 0)
        return 0;   // a row map disables the staging
    const double per_wave = 1.5 * kWave * static_cast<double>(src.nnz) / src.rows;
    const int cap = (static_cast<int>(per_wave) + 255) / 256 * 256;
    return std::min(kBuildStage, std::max(256, cap));
}

// One group of TR lanes per tile row.  COUNT pass: per tile row the number of
// tiles, payload slots and dense tiles.  FILL pass: tile columns, masks,
// value offsets and payload.  Duplicates (adjacent equal columns in a sorted
// row) are summed in row order.
template <int TR, int TC, bool COL_MAJOR, bool FILL>
__global__ __launch_bounds__(kThreads) void tile_build_kernel(
    int tile_row_count, int active_rows, int active_cols,
    const int *__restrict__ rp, const int *__restrict__ ci,
    const double *__restrict__ val, const int *__restrict__ row_map,
    int dense_threshold, int64_t *__restrict__ row_tiles,
    int64_t *__restrict__ row_payload, int64_t *__restrict__ row_dense,
    const int64_t *__restrict__ tile_off, const int64_t *__restrict__ payload_off,
    int *__restrict__ tile_col_idx, uint32_t *__restrict__ masks,
    int *__restrict__ value_offsets, double *__restrict__ values, int4 *__restrict__ meta,
    int2 *__restrict__ src_slot, int *__restrict__ dup_flag, int stage_cap)
{
    static_assert(TR * TC == 32, "tiles hold 32 elements");
    static_assert(kWave % TR == 0, "a wave holds whole tile rows");
    extern __shared__ int s_build_stage[];   // stage_cap columns per wave
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    // Row gid (= tile row gid / TR, row gid % TR of it).  The wave's rows are
    // 64 consecutive rows; without a row map their entries are one stretch
    // of the CSR, whose first stage_cap columns are staged in LDS (all of
    // them when they fit).
    const int64_t wave_row0 = gid - lane_id();
    const int64_t lim = min<int64_t>(static_cast<int64_t>(tile_row_count) * TR, active_rows);
    int e0 = 0, ns = 0;
    if (!row_map && wave_row0 < lim) {
        e0 = rp[wave_row0];
        ns = min(rp[min<int64_t>(wave_row0 + kWave, lim)] - e0, stage_cap);
    }
    int *stage = s_build_stage + (threadIdx.x / kWave) * stage_cap;
    for (int x = lane_id(); x < ns; x += kWave)
        stage[x] = ci[e0 + x];
    wave_sync();
    auto col = [&](int x) { return x - e0 < ns ? stage[x - e0] : ci[x]; };
    const int64_t tr = gid / TR;
    const int r = static_cast<int>(gid % TR);
    if (tr >= tile_row_count)
        return;   // whole groups leave together
    const int64_t row = tr * TR + r;
    int p = 0, end = 0;
    if (row < active_rows) {
        const int src = row_map ? row_map[row] : static_cast<int>(row);
        p = rp[src];
        end = rp[src + 1];
    }
    // Columns at or beyond active_cols are outside the tiled extent.
    while (end > p && col(end - 1) >= active_cols)
        --end;

    int64_t ntiles = 0, payload = 0, ndense = 0;
    int64_t t_out = FILL ? tile_off[tr] : 0;
    int64_t p_out = FILL ? payload_off[tr] : 0;
    for (;;) {
        int head = p < end ? col(p) / TC : INT_MAX;
#pragma unroll
        for (int o = TR / 2; o > 0; o >>= 1)
            head = min(head, __shfl_xor(head, o, TR));
        if (head == INT_MAX)
            break;
        uint32_t bits = 0;
        int q = p;
        if (p + TC - e0 <= ns) {
            // The first TC columns from p, staged in LDS, are read together;
            // without duplicates a row has at most TC entries in a tile.
            // (Unstaged, the extra reads would be global loads of columns
            // past a sparse tile's few entries.)
            int cb[TC];
#pragma unroll
            for (int u = 0; u < TC; ++u)
                cb[u] = p + u < end ? col(p + u) : -1;
            bool in_tile = true;
#pragma unroll
            for (int u = 0; u < TC; ++u) {
                in_tile = in_tile && cb[u] >= 0 && cb[u] / TC == head;
                if (in_tile) {
                    const int lc = cb[u] % TC;
                    bits |= 1u << (COL_MAJOR ? lc * TR + r : r * TC + lc);
                    ++q;
                }
            }
        }
        while (q < end) {   // the rest (all of it when unstaged)
            const int c = col(q);
            if (c / TC != head)
                break;
            const int lc = c % TC;
            bits |= 1u << (COL_MAJOR ? lc * TR + r : r * TC + lc);
            ++q;
        }
        uint32_t mask = bits;
#pragma unroll
        for (int o = TR / 2; o > 0; o >>= 1)
            mask |= __shfl_xor(mask, o, TR);
        const int cnt = __popc(mask);
        const bool dense = cnt >= dense_threshold;
        if constexpr (FILL) {
            if (r == 0) {
                tile_col_idx[t_out] = head;
                masks[t_out] = mask;
                value_offsets[t_out] = static_cast<int>(p_out);
                uint32_t bytes = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    bytes |= stride4_bits(mask, k) << (8 * k);
                meta[t_out] = make_int4(head, static_cast<int>(mask), static_cast<int>(bytes),
                                        static_cast<int>(p_out) | (dense ? INT_MIN : 0));
            }
            if (dense) {
                // Absent slots of this lane's row are explicit zeros.
#pragma unroll
                for (int lc = 0; lc < TC; ++lc) {
                    const int phys = COL_MAJOR ? lc * TR + r : r * TC + lc;
                    if (!(bits & (1u << phys)))
                        values[p_out + phys] = 0.0;
                }
            }
            // Sum of source entries x0 .. x-1 (duplicates of column c, in
            // entry order) into its payload slot.
            auto store = [&](int c, double s, int x0, int x) {
                const int lc = c % TC;
                const int phys = COL_MAJOR ? lc * TR + r : r * TC + lc;
                const int slot = static_cast<int>(p_out) + payload_slot(mask, dense ? 32 : cnt, phys);
                values[slot] = s;
                if (src_slot) {
                    const int2 sl = make_int2(slot, static_cast<int>(p_out) + kmajor_slot(mask, phys));
                    for (int y = x0; y < x; ++y)
                        src_slot[y] = sl;
                    if (x - x0 > 1)
                        *dup_flag = 1;
                }
            };
            if (q - p == 1) {
                store(col(p), val[p], p, q);   // the common case of sparse operands
            } else if (!row_map && q - p <= TC) {
                // At most TC entries (always, without duplicates): their
                // values are loaded together, not one dependent load each.
                // (Measured slower for rows gathered through a row map.)
                double vb[TC];
#pragma unroll
                for (int u = 0; u < TC; ++u)
                    vb[u] = p + u < q ? val[p + u] : 0.0;
                int cprev = -1, x0 = p;
                double s = 0.0;
#pragma unroll
                for (int u = 0; u < TC; ++u) {
                    if (p + u < q) {
                        const int c = col(p + u);
                        if (c != cprev) {
                            if (cprev >= 0)
                                store(cprev, s, x0, p + u);
                            cprev = c;
                            x0 = p + u;
                            s = vb[u];
                        } else {
                            s += vb[u];
                        }
                    }
                }
                if (cprev >= 0)
                    store(cprev, s, x0, q);
            } else {
                int x = p;
                while (x < q) {
                    const int c = col(x);
                    const int x0 = x;
                    double s = val[x];
                    ++x;
                    while (x < q && col(x) == c) {
                        s += val[x];
                        ++x;
                    }
                    store(c, s, x0, x);
                }
            }
            ++t_out;
            p_out += dense ? 32 : cnt;
        }
        ++ntiles;
        payload += dense ? 32 : cnt;
        ndense += dense ? 1 : 0;
        p = q;
    }
    if constexpr (!FILL) {
        if (r == 0) {
            row_tiles[tr] = ntiles;
            row_payload[tr] = payload;
            row_dense[tr] = ndense;
        }
    }
}

__global__ void gather3_kernel(const int64_t *a, const int64_t *b, const int64_t *c, int64_t *out)
{
    out[0] = *a;
    out[1] = *b;
    out[2] = *c;
}

__global__ void narrow_i64_kernel(const int64_t *in, int *out, int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = static_cast<int>(in[i]);
}

__global__ void tile_row_of_kernel(int rows, const int *__restrict__ trp,
                                   int *__restrict__ out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < rows)
        for (int t = trp[i]; t < trp[i + 1]; ++t)
            out[t] = static_cast<int>(i);
}

__global__ void histogram_kernel(const int *__restrict__ keys, int64_t n,
                                 int64_t *__restrict__ counts)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        atomicAdd(reinterpret_cast<unsigned long long *>(counts + keys[i]), 1ull);
}

// CSC of a tile matrix from its stably sorted tile ids.
__global__ void csc_fill_kernel(int64_t n, const int *__restrict__ sorted_ids,
                                const int *__restrict__ tile_row_of,
                                int *__restrict__ tile_row_idx,
                                int *__restrict__ csc_tile_ids)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const int t = sorted_ids[i];
        csc_tile_ids[i] = t;
        tile_row_idx[i] = tile_row_of[t];
    }
}

// B tile b (sorted by K) comes from A tile ids[b]; gather its metadata.
__global__ void transpose_meta_kernel(int64_t n, const int *__restrict__ ids,
                                      const int *__restrict__ a_row_of,
                                      const uint32_t *__restrict__ a_masks,
                                      const int *__restrict__ a_off,
                                      int *__restrict__ b_col,
                                      uint32_t *__restrict__ b_masks,
                                      int64_t *__restrict__ spans,
                                      int *__restrict__ b_csc_ids)
{
    const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b < n) {
        const int a = ids[b];
        b_col[b] = a_row_of[a];
        b_masks[b] = a_masks[a];
        spans[b] = a_off[a + 1] - a_off[a];
        b_csc_ids[a] = static_cast<int>(b);   // A order == B's CSC order
    }
}

// Payload copy, one wave per 2 tiles (32 lanes per tile).
__global__ void transpose_payload_kernel(int64_t n, const int *__restrict__ ids,
                                         const int *__restrict__ a_off,
                                         const double *__restrict__ a_vals,
                                         const int64_t *__restrict__ b_off64,
                                         int *__restrict__ b_off,
                                         double *__restrict__ b_vals)
{
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t b = gid / 32;
    const int s = static_cast<int>(gid % 32);
    if (b >= n)
        return;
    const int a = ids[b];
    const int src = a_off[a];
    const int span = a_off[a + 1] - src;
    const int64_t dst = b_off64[b];
    if (s == 0)
        b_off[b] = static_cast<int>(dst);
    if (s < span)
        b_vals[dst + s] = a_vals[src + s];
}

// Row pointer of a sorted key array: ptr[k] = first index with key >= k.
__global__ void row_ptr_from_sorted_kernel(const int *__restrict__ keys, int64_t n, int rows,
                                           int *__restrict__ ptr)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > n)
        return;
    const int prev = i == 0 ? -1 : keys[i - 1];
    const int cur = i == n ? rows : keys[i];
    for (int k = prev + 1; k <= cur; ++k)
        ptr[k] = static_cast<int>(i);
}

// Part `part` of a tile's record (meta, then the first kRecVals values of
// the tile's k-major payload: row k's entries contiguous, columns ascending).
// Parts of a record that hold something: the meta, then one part per two of
// the tile's first kRecVals values.  The kernels that read records use a
// value only when the tile has it, so the parts past these are never used;
// the writers that hold the meta anyway skip them.
__device__ __forceinline__ int record_parts_used(const int4 &m)
{
    return 1 + (min(kRecVals, __popc(static_cast<uint32_t>(m.y))) + 1) / 2;
}

__device__ __forceinline__ int4 record_part(int part, const int4 &m, const double *kmajor)
{
    if (part == 0)
        return m;
    const uint32_t off = static_cast<uint32_t>(m.w) & 0x7FFFFFFFu;
    const int nv = min(kRecVals, __popc(static_cast<uint32_t>(m.y)));
    const int j = 2 * (part - 1);
    double2 v;
    v.x = j < nv ? kmajor[off + j] : 0.0;
    v.y = j + 1 < nv ? kmajor[off + j + 1] : 0.0;
    return *reinterpret_cast<const int4 *>(&v);
}

// Value parts (1 .. kRecParts-1) of B's record planes, one tile per thread;
// part 0 (meta) is left as it is.  A thread reads its tile's first values
// once (one contiguous stretch, consecutive tiles' stretches adjacent) and
// writes each plane's entry (consecutive threads, consecutive entries).
__global__ void record_values_kernel(int64_t n, const int4 *__restrict__ meta,
                                     const double *__restrict__ kmajor,
                                     int4 *__restrict__ planes)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int4 m = meta[t];
    const int used = record_parts_used(m);
    // Every value the record holds is loaded before any plane is written.
    const uint32_t off = static_cast<uint32_t>(m.w) & 0x7FFFFFFFu;
    const int nv = min(kRecVals, __popc(static_cast<uint32_t>(m.y)));
    double v[kRecVals];
#pragma unroll
    for (int j = 0; j < kRecVals; ++j)
        v[j] = j < nv ? kmajor[off + j] : 0.0;
#pragma unroll
    for (int part = 1; part < kRecParts; ++part)
        if (part < used)
            planes[part * n + t] = *reinterpret_cast<const int4 *>(
                &reinterpret_cast<const double2 *>(v)[part - 1]);
}

// Values-only update: the first source entry of each slot's run sums the
// run in entry order (the order the tile build sums duplicates in) and
// stores the sum into both payload layouts.  Runs are adjacent because
// duplicates of a coordinate are adjacent in a row-sorted CSR.
__global__ void scatter_runs_kernel(int64_t nnz, const int2 *__restrict__ slot,
                                    const double *__restrict__ v, double *__restrict__ values,
                                    double *__restrict__ kmajor)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nnz)
        return;
    const int2 s = slot[e];
    if (s.x < 0 || (e > 0 && slot[e - 1].x == s.x))
        return;
    double sum = v[e];
    for (int64_t f = e + 1; f < nnz && slot[f].x == s.x; ++f)
        sum += v[f];
    values[s.x] = sum;
    if (kmajor)
        kmajor[s.y] = sum;
}

// Offline, per A tile: the record a B tile built as A^T takes over, its meta
// with the tile column replaced by the tile row.
// The B tile's k-major order is the A tile's column-major order.
__global__ void transpose_source_kernel(int64_t n, const int *__restrict__ a_row_of,
                                        const int4 *__restrict__ a_meta,
                                        const double *__restrict__ a_vals_cm,
                                        BRecord *__restrict__ out)
{
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t t = gid / kRecParts;
    if (t >= n)
        return;
    int4 m = a_meta[t];
    m.x = a_row_of[t];
    const int part = static_cast<int>(gid % kRecParts);
    reinterpret_cast<int4 *>(out + t)[part] = record_part(part, m, a_vals_cm);
}

// B = A^T: B tile b is A tile ids[b], so its record is a copy of that A
// tile's transpose record, one tile per thread: a thread reads its record
// (80 contiguous bytes) and consecutive threads write consecutive entries
// of every plane (3 % faster than a thread per 16-byte part).
__global__ void transposed_records_kernel(int64_t n, const int *__restrict__ ids,
                                          const BRecord *__restrict__ src,
                                          int4 *__restrict__ planes)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int4 *r = reinterpret_cast<const int4 *>(src + ids[t]);
    int4 v[kRecParts];
#pragma unroll
    for (int part = 0; part < kRecParts; ++part)
        v[part] = r[part];
#pragma unroll
    for (int part = 0; part < kRecParts; ++part)
        planes[part * n + t] = v[part];
}

// Class of a B tile within its row: 0 for heavy tiles (kHeavyMin entries or
// more, or entries in more than one row k), else 1 + k for a light tile whose
// entries all lie in row k.  bytes holds the per-k column bytes of the tile
// (for B = A^T, the per-k row bytes of the A tile).
__device__ __forceinline__ int b_tile_class(uint32_t mask, uint32_t bytes)
{
    uint32_t rows = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        rows |= ((bytes >> (8 * k)) & 0xFFu) ? (1u << k) : 0u;
    if (__popc(mask) >= kHeavyMin || __popc(rows) != 1)
        return 0;
    return 1 + __builtin_ctz(rows);
}

// Counting transpose of A's tiles into B = A^T rows, for operands whose tile
// columns hold at most kShortCol tiles: a count per B row, a scan, a scatter
// in arbitrary order within each row, then one workgroup per kOrderRows rows
// ranks every entry within its row by (class, A tile id).  The result is the
// stable sort by (row, class) that the radix path produces, entry for entry.
constexpr int kShortCol = 64;     // most tiles in an A tile column it takes
constexpr int kOrderRows = 64;    // B rows per ordering workgroup
constexpr int kOrderIdBits = 28;  // A tile ids below 2^28; the class above them

// B class of the A^T tile made from each A tile, and its arrival rank in its
// B row (the tile's column of A).
__global__ void column_rank_kernel(int64_t n, const int *__restrict__ col,
                                   const uint32_t *__restrict__ masks, int *__restrict__ counts,
                                   int *__restrict__ rank_cls)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const uint32_t mask = masks[t];
    uint32_t bytes = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        bytes |= stride4_bits(mask, k) << (8 * k);
    const int r = atomicAdd(&counts[col[t]], 1);
    rank_cls[t] = (r << 3) | b_tile_class(mask, bytes);
}

__global__ void column_scatter_kernel(int64_t n, const int *__restrict__ col,
                                      const int *__restrict__ rank_cls,
                                      const int *__restrict__ rp, uint32_t *__restrict__ ent)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int rc = rank_cls[t];
    ent[rp[col[t]] + (rc >> 3)] = (static_cast<uint32_t>(rc & 7) << kOrderIdBits) |
                                  static_cast<uint32_t>(t);
}

// Rows r0 .. r0 + kOrderRows - 1 of B: each entry's place is its row's start
// plus the number of the row's entries below it by (class, id); each
// (row, class) run starts after the row's entries of lower classes.
__global__ void __launch_bounds__(kThreads)
    order_rows_kernel(int R, int64_t n, const int *__restrict__ rp,
                      const uint32_t *__restrict__ ent, int *__restrict__ sorted_ids,
                      int *__restrict__ groups)
{
    __shared__ int s_rp[kOrderRows + 1];
    __shared__ uint32_t s_ent[kOrderRows * kShortCol];
    const int r0 = blockIdx.x * kOrderRows;
    const int nr = min(kOrderRows, R - r0);
    for (int i = threadIdx.x; i <= nr; i += blockDim.x)
        s_rp[i] = rp[r0 + i];
    __syncthreads();
    const int base = s_rp[0];
    const int total = s_rp[nr] - base;
    // At most nr * kShortCol when every column holds at most kShortCol
    // tiles, which the host checks before taking this path; the bound keeps
    // the stage in its LDS whatever the input.
    if (total > kOrderRows * kShortCol)
        return;
    for (int e = threadIdx.x; e < total; e += blockDim.x)
        s_ent[e] = ent[base + e];
    __syncthreads();
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int lo = 0, hi = nr - 1;   // last row starting at or before e
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_rp[mid] - base <= e)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int b = s_rp[lo] - base, end = s_rp[lo + 1] - base;
        const uint32_t v = s_ent[e];
        int rank = 0;
        for (int x = b; x < end; ++x)
            rank += s_ent[x] < v ? 1 : 0;
        sorted_ids[base + b + rank] = static_cast<int>(v & ((1u << kOrderIdBits) - 1u));
    }
    for (int i = threadIdx.x; i < nr * kClassSlots; i += blockDim.x) {
        const int r = i / kClassSlots;
        const uint32_t c = static_cast<uint32_t>(i % kClassSlots);
        const int b = s_rp[r] - base, end = s_rp[r + 1] - base;
        int below = 0;
        for (int x = b; x < end; ++x)
            below += (s_ent[x] >> kOrderIdBits) < c ? 1 : 0;
        groups[static_cast<int64_t>(r0 + r) * kClassSlots + c] = base + b + below;
    }
    if (r0 + nr == R && threadIdx.x == 0)
        groups[static_cast<int64_t>(R) * kClassSlots] = static_cast<int>(n);
}

// Tiles per tile column; column_max_kernel then puts the largest count in
// counts[cols].
__global__ void column_count_kernel(int64_t n, const int *__restrict__ col, int *__restrict__ counts)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n)
        atomicAdd(&counts[col[t]], 1);
}

// One atomic per workgroup: an atomic per wave on the one address
// serialises (0.1 ms on 2^19 columns).
__global__ void __launch_bounds__(kThreads) column_max_kernel(int cols, int *__restrict__ counts)
{
    __shared__ int s_max[kThreads / kWave];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    int m = j < cols ? counts[j] : 0;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
        m = max(m, __shfl_xor(m, o, kWave));
    if (lane_id() == 0)
        s_max[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / kWave; ++w)
            m = max(m, s_max[w]);
        if (m > 0)
            atomicMax(&counts[cols], m);
    }
}

// Per A tile t in column K: {first B tile of row K, heavy tiles, all tiles}
// and the light run its pairs need: for an A tile whose entries all lie in
// one column k, only the light tiles of row k can meet it.
__global__ __launch_bounds__(kThreads) void pair_extent_kernel(
    int64_t n, const int4 *__restrict__ a_meta, const int *__restrict__ groups,
    int4 *__restrict__ out, int2 *__restrict__ light, unsigned long long *__restrict__ pairs)
{
    // Grid-stride, so that the pair total takes one atomic per workgroup.
    __shared__ unsigned long long s_sum[kThreads / kWave];
    unsigned long long sum = 0;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int4 m = a_meta[t];
        const int *gp = groups + static_cast<int64_t>(kClassSlots) * m.x;
        const int b0 = gp[0];
        const int total = gp[kClassSlots] - b0;
        out[t] = make_int4(b0, gp[1] - b0, total, 0);
        sum += static_cast<unsigned long long>(total);
        const uint32_t bytes = static_cast<uint32_t>(m.z);
        uint32_t cols = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            cols |= ((bytes >> (8 * k)) & 0xFFu) ? (1u << k) : 0u;
        if (__popc(cols) == 1) {
            const int k0 = __builtin_ctz(cols);
            light[t] = make_int2(gp[1 + k0], gp[2 + k0] - gp[1 + k0]);
        } else {
            light[t] = make_int2(gp[1], gp[5] - gp[1]);
        }
    }
    sum = wave_sum(sum);
    if (lane_id() == 0)
        s_sum[threadIdx.x / kWave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long all = 0;
        for (int w = 0; w < kThreads / kWave; ++w)
            all += s_sum[w];
        if (all)
            atomicAdd(pairs, all);
    }
}

// Stable partition of each B row's tiles by class (tiles arrive in column
// order): the order a stable sort by (row, class) gives.  Rows of up to
// kPartRow tiles take one thread, which packs its per-class counts into the
// bytes of a 64-bit word; longer rows are listed for a wave each.
constexpr int kPartRow = 64;
constexpr int kPartBatch = 8;    // metas a thread loads at once
constexpr int kClassesUsed = 5;   // heavy, light k = 0..3

__device__ __forceinline__ int meta_class(const int4 &m)
{
    return b_tile_class(static_cast<uint32_t>(m.y), static_cast<uint32_t>(m.z));
}

__global__ void class_partition_kernel(int R, const int *__restrict__ trp,
                                       const int4 *__restrict__ meta, int4 *__restrict__ out,
                                       int4 *__restrict__ plane0, int *__restrict__ groups,
                                       int *__restrict__ long_rows, int *__restrict__ n_long)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > R)
        return;
    const int s = trp[r];
    if (r == R) {
        groups[static_cast<int64_t>(kClassSlots) * R] = s;
        return;
    }
    const int e = trp[r + 1];
    if (e - s > kPartRow) {
        long_rows[atomicAdd(n_long, 1)] = r;
        return;
    }
    // The row's metas come kPartBatch at a time, loads issued together; a
    // row of at most kPartBatch tiles keeps them for the placing pass.
    int4 m[kPartBatch];
#pragma unroll
    for (int u = 0; u < kPartBatch; ++u)
        m[u] = make_int4(0, 0, 0, 0);
    unsigned long long cnt = 0;   // byte c: tiles of class c
    for (int x0 = s; x0 < e; x0 += kPartBatch) {
#pragma unroll
        for (int u = 0; u < kPartBatch; ++u)
            if (x0 + u < e)
                m[u] = meta[x0 + u];
#pragma unroll
        for (int u = 0; u < kPartBatch; ++u)
            if (x0 + u < e)
                cnt += 1ull << (8 * meta_class(m[u]));
    }
    // Byte c: tiles of classes below c (at most kPartRow, so no carries).
    unsigned long long pos = (cnt << 8) * 0x0101010101010101ull;
    int *g = groups + static_cast<int64_t>(kClassSlots) * r;
#pragma unroll
    for (int c = 0; c < kClassSlots; ++c)
        g[c] = s + static_cast<int>((pos >> (8 * c)) & 0xFFu);
    const bool held = e - s <= kPartBatch;
    for (int x0 = s; x0 < e; x0 += kPartBatch) {
        if (!held) {
#pragma unroll
            for (int u = 0; u < kPartBatch; ++u)
                if (x0 + u < e)
                    m[u] = meta[x0 + u];
        }
#pragma unroll
        for (int u = 0; u < kPartBatch; ++u) {
            if (x0 + u < e) {
                const int c = meta_class(m[u]);
                const int p = s + static_cast<int>((pos >> (8 * c)) & 0xFFu);
                pos += 1ull << (8 * c);
                out[p] = m[u];
                plane0[p] = m[u];
            }
        }
    }
}

// The listed long rows, a wave each: class counts by ballot, then every
// tile placed after the row's tiles of lower classes and the earlier tiles
// of its own.
__global__ void __launch_bounds__(kThreads)
    class_partition_long_kernel(const int *__restrict__ trp, const int4 *__restrict__ meta,
                                int4 *__restrict__ out, int4 *__restrict__ plane0,
                                int *__restrict__ groups, const int *__restrict__ long_rows,
                                const int *__restrict__ n_long)
{
    const int lane = lane_id();
    const int waves = gridDim.x * (blockDim.x / kWave);
    const int count = *n_long;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < count; i += waves) {
        const int r = long_rows[i];
        const int s = trp[r], e = trp[r + 1];
        int run[kClassesUsed];
#pragma unroll
        for (int c = 0; c < kClassesUsed; ++c)
            run[c] = 0;
        for (int x0 = s; x0 < e; x0 += kWave) {
            const int x = x0 + lane;
            const int c = x < e ? meta_class(meta[x]) : -1;
#pragma unroll
            for (int k = 0; k < kClassesUsed; ++k)
                run[k] += __popcll(__ballot(c == k));
        }
        int start = s;
#pragma unroll
        for (int k = 0; k < kClassesUsed; ++k) {
            const int nk = run[k];
            run[k] = start;
            start += nk;
        }
        if (lane < kClassSlots) {
            int v = e;
#pragma unroll
            for (int k = 0; k < kClassesUsed; ++k)
                if (lane == k)
                    v = run[k];
            groups[static_cast<int64_t>(kClassSlots) * r + lane] = v;
        }
        for (int x0 = s; x0 < e; x0 += kWave) {
            const int x = x0 + lane;
            int4 m = make_int4(0, 0, 0, 0);
            int c = -1;
            if (x < e) {
                m = meta[x];
                c = meta_class(m);
            }
            int p = 0;
#pragma unroll
            for (int k = 0; k < kClassesUsed; ++k) {
                const uint64_t b = __ballot(c == k);
                if (c == k)
                    p = run[k] + __popcll(b & below);
                run[k] += __popcll(b);
            }
            if (x < e) {
                out[p] = m;
                plane0[p] = m;
            }
        }
    }
}

// Sort key of a B tile: its row times kClassSlots plus its class.
__global__ void class_keys_kernel(int64_t n, const int *__restrict__ row,
                                  const int4 *__restrict__ meta, int *__restrict__ keys)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) {
        const int4 m = meta[t];
        keys[t] = kClassSlots * row[t] +
                  b_tile_class(static_cast<uint32_t>(m.y), static_cast<uint32_t>(m.z));
    }
}

// B's row pointer from the row pointer over (row, class) keys.
__global__ void group_row_ptr_kernel(int rows, const int *__restrict__ groups,
                                     int *__restrict__ trp)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= rows)
        trp[r] = groups[static_cast<int64_t>(kClassSlots) * r];
}

// Gathers a B's meta into class order (the record value planes are built
// afterwards, from the ordered meta, when a kernel needs them).
__global__ void gather_meta_kernel(int64_t n, const int *__restrict__ perm,
                                   const int4 *__restrict__ meta_in, int4 *__restrict__ meta_out,
                                   int4 *__restrict__ plane0)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int4 m = meta_in[perm[t]];
    meta_out[t] = m;
    plane0[t] = m;
}

__global__ void entry_rows_kernel(int rows, const int *__restrict__ rp,
                                  int *__restrict__ out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < rows)
        for (int e = rp[i]; e < rp[i + 1]; ++e)
            out[e] = static_cast<int>(i);
}

__global__ void gather_transposed_kernel(int64_t nnz, const int *__restrict__ order,
                                         const int *__restrict__ entry_row,
                                         const double *__restrict__ val,
                                         int *__restrict__ t_col,
                                         double *__restrict__ t_val)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nnz) {
        const int e = order[i];
        t_col[i] = entry_row[e];
        t_val[i] = val[e];
    }
}

__global__ void nnz_cub_kernel(int64_t nnz, const int *__restrict__ ci,
                               const int *__restrict__ rp,
                               const int64_t *__restrict__ coldeg,
                               unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long part[kThreads / kWave];
    unsigned long long v = 0;
    // Grid-stride, so that only a few hundred workgroups add into *out.
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nnz;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int c = ci[i];
        v += coldeg ? static_cast<unsigned long long>(coldeg[c])
                    : static_cast<unsigned long long>(rp[c + 1] - rp[c]);
    }
    v = wave_sum(v);
    if (lane_id() == 0)
        part[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < kThreads / kWave; ++w)
            s += part[w];
        atomicAdd(out, s);
    }
}

int bits_for(int64_t max_value)
{
    int bits = 1;
    while (bits < 62 && (int64_t(1) << bits) <= max_value)
        ++bits;
    return bits;
}

// Stable sort of int keys with int payload (hipcub LSD radix sort).
int sort_pairs(const int *k_in, int *k_out, const int *v_in, int *v_out, int64_t n,
               int bits, DevBuf *tmp, hipStream_t s)
{
    size_t bytes = 0;
    RTT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k_in, k_out, v_in, v_out,
                                               static_cast<int>(n), 0, bits, s));
    RTT_TRY(tmp->ensure(bytes));
    RTT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp->ptr, bytes, k_in, k_out, v_in, v_out,
                                               static_cast<int>(n), 0, bits, s));
    return RTT_OK;
}

int read_i64(const int64_t *d, int64_t *h, hipStream_t s)
{
    RTT_HIP(hipMemcpyAsync(h, d, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    RTT_HIP(hipStreamSynchronize(s));
    return RTT_OK;
}

// Builds B's tile CSC (tile_col_ptr, tile_row_idx, csc_tile_ids) from its
// tile CSR: tiles of each tile column in increasing tile-row order
// (src/gpu_dmma_tiles.h:1003-1088 semantics).
int build_csc(DevTiles *t, BuildWork *w)
{
    hipStream_t s = ctx().stream;
    const int64_t n = t->num_tiles;
    const int ncols = t->tile_col_count;
    RTT_TRY(t->tile_col_ptr.ensure((ncols + 1) * sizeof(int)));
    RTT_TRY(t->tile_row_idx.ensure(n * sizeof(int)));
    RTT_TRY(t->csc_tile_ids.ensure(n * sizeof(int)));
    RTT_TRY(fill_tile_row_of(t));
    RTT_TRY(w->spans.ensure((ncols + 1) * sizeof(int64_t)));
    RTT_TRY(w->counters.ensure((ncols + 1) * sizeof(int64_t)));
    int64_t *cnt = w->spans.as<int64_t>();
    hipLaunchKernelGGL(set_i64_kernel, dim3(blocks_for(ncols + 1, kThreads)), dim3(kThreads),
                       0, s, cnt, static_cast<int64_t>(ncols) + 1, int64_t(0));
    if (n > 0) {
        hipLaunchKernelGGL(histogram_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0,
                           s, t->tile_col_idx.as<int>(), n, cnt);
        RTT_TRY(w->ids_in.ensure(n * sizeof(int)));
        RTT_TRY(w->ids_out.ensure(n * sizeof(int)));
        RTT_TRY(w->keys_out.ensure(n * sizeof(int)));
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                           w->ids_in.as<int>(), n);
        RTT_TRY(sort_pairs(t->tile_col_idx.as<int>(), w->keys_out.as<int>(),
                           w->ids_in.as<int>(), w->ids_out.as<int>(), n, bits_for(ncols),
                           &w->scan_tmp, s));
        hipLaunchKernelGGL(csc_fill_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0,
                           s, n, w->ids_out.as<int>(), t->tile_row_of.as<int>(),
                           t->tile_row_idx.as<int>(), t->csc_tile_ids.as<int>());
    }
    int64_t *ptr64 = w->counters.as<int64_t>();
    RTT_TRY(exclusive_scan_i64(cnt, ptr64, ncols, &w->scan_tmp, s));
    hipLaunchKernelGGL(narrow_i64_kernel, dim3(blocks_for(ncols + 1, kThreads)),
                       dim3(kThreads), 0, s, ptr64, t->tile_col_ptr.as<int>(),
                       static_cast<int64_t>(ncols) + 1);
    RTT_HIP(hipGetLastError());
    t->has_csc = true;
    return RTT_OK;
}

template <int TR, int TC, bool CM>
int run_tile_build(const DevCsr &src, const int *row_map, int active_rows,
                   int active_cols, int dense_threshold, DevTiles *out,
                   BuildWork *w, int2 *src_slot, int *dup_flag)
{
    hipStream_t s = ctx().stream;
    const int R = out->tile_row_count;
    RTT_TRY(w->row_tiles.ensure((R + 1) * sizeof(int64_t)));
    RTT_TRY(w->row_payload.ensure((R + 1) * sizeof(int64_t)));
    RTT_TRY(w->row_dense.ensure((R + 1) * sizeof(int64_t)));
    // counters: [0, R] tile prefix, [R+1, 2R+1] payload prefix, then dense sum
    RTT_TRY(w->counters.ensure((2 * (R + 1) + 1) * sizeof(int64_t)));
    int64_t *tile_pre = w->counters.as<int64_t>();
    int64_t *pay_pre = tile_pre + (R + 1);
    int64_t *dense_pre = w->spans.ensure((R + 1) * sizeof(int64_t)) == RTT_OK
                             ? w->spans.as<int64_t>()
                             : nullptr;
    if (!dense_pre)
        return RTT_ERR_OOM;
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, s, w->row_tiles.as<int64_t>() + R,
                       int64_t(1), int64_t(0));
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, s,
                       w->row_payload.as<int64_t>() + R, int64_t(1), int64_t(0));
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, s, w->row_dense.as<int64_t>() + R,
                       int64_t(1), int64_t(0));
    const int64_t threads_total = static_cast<int64_t>(R) * TR;
    const int stage_cap = build_stage_cap(src, row_map != nullptr);
    const size_t stage_bytes = static_cast<size_t>(kThreads / kWave) * stage_cap * sizeof(int);
    if (R > 0)
        hipLaunchKernelGGL((tile_build_kernel<TR, TC, CM, false>),
                           dim3(blocks_for(threads_total, kThreads)), dim3(kThreads), stage_bytes,
                           s, R, active_rows, active_cols, src.row_ptr.as<int>(),
                           src.col_idx.as<int>(), src.values.as<double>(), row_map,
                           dense_threshold, w->row_tiles.as<int64_t>(),
                           w->row_payload.as<int64_t>(), w->row_dense.as<int64_t>(),
                           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr, nullptr, stage_cap);
    RTT_HIP(hipGetLastError());
    RTT_TRY(exclusive_scan_i64(w->row_tiles.as<int64_t>(), tile_pre, R, &w->scan_tmp, s));
    RTT_TRY(exclusive_scan_i64(w->row_payload.as<int64_t>(), pay_pre, R, &w->scan_tmp, s));
    RTT_TRY(exclusive_scan_i64(w->row_dense.as<int64_t>(), dense_pre, R, &w->scan_tmp, s));
    // The three totals in one readback.
    int64_t totals[3] = {0, 0, 0};
    RTT_TRY(w->totals.ensure(3 * sizeof(int64_t)));
    hipLaunchKernelGGL(gather3_kernel, dim3(1), dim3(1), 0, s, tile_pre + R, pay_pre + R,
                       dense_pre + R, w->totals.as<int64_t>());
    RTT_HIP(hipMemcpyAsync(totals, w->totals.ptr, sizeof(totals), hipMemcpyDeviceToHost, s));
    RTT_HIP(hipStreamSynchronize(s));
    if (totals[0] > INT_MAX || totals[1] > INT_MAX)
        return fail(RTT_ERR_OVERFLOW,
                    "tile count %lld or payload %lld exceeds the int32 tile ABI",
                    (long long)totals[0], (long long)totals[1]);
    out->num_tiles = totals[0];
    out->payload = totals[1];
    out->dense = totals[2];
    out->sparse = totals[0] - totals[2];
    RTT_TRY(out->tile_row_ptr.ensure((R + 1) * sizeof(int)));
    RTT_TRY(out->tile_col_idx.ensure(out->num_tiles * sizeof(int)));
    RTT_TRY(out->masks.ensure(out->num_tiles * sizeof(uint32_t)));
    RTT_TRY(out->value_offsets.ensure((out->num_tiles + 1) * sizeof(int)));
    RTT_TRY(out->values.ensure(out->payload * sizeof(double)));
    RTT_TRY(out->meta.ensure(out->num_tiles * sizeof(int4)));
    hipLaunchKernelGGL(narrow_i64_kernel, dim3(blocks_for(R + 1, kThreads)), dim3(kThreads), 0,
                       s, tile_pre, out->tile_row_ptr.as<int>(), static_cast<int64_t>(R) + 1);
    // value_offsets[num_tiles] = payload total.
    hipLaunchKernelGGL(narrow_i64_kernel, dim3(1), dim3(1), 0, s, pay_pre + R,
                       out->value_offsets.as<int>() + out->num_tiles, int64_t(1));
    if (R > 0 && out->num_tiles > 0)
        hipLaunchKernelGGL((tile_build_kernel<TR, TC, CM, true>),
                           dim3(blocks_for(threads_total, kThreads)), dim3(kThreads), stage_bytes,
                           s, R, active_rows, active_cols, src.row_ptr.as<int>(),
                           src.col_idx.as<int>(), src.values.as<double>(), row_map,
                           dense_threshold, nullptr, nullptr, nullptr, tile_pre, pay_pre,
                           out->tile_col_idx.as<int>(), out->masks.as<uint32_t>(),
                           out->value_offsets.as<int>(), out->values.as<double>(),
                           out->meta.as<int4>(), src_slot, dup_flag, stage_cap);
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

} // namespace

void DevTiles::view(rtt_device_tiles *v) const
{
    v->rows = rows;
    v->cols = cols;
    v->tile_rows = tile_rows;
    v->tile_cols = tile_cols;
    v->tile_row_count = tile_row_count;
    v->tile_col_count = tile_col_count;
    v->num_tiles = static_cast<int>(num_tiles);
    v->payload_size = static_cast<int>(payload);
    v->dense_tiles = static_cast<int>(dense);
    v->sparse_tiles = static_cast<int>(sparse);
    v->tile_row_ptr = tile_row_ptr.as<int>();
    v->tile_col_idx = tile_col_idx.as<int>();
    v->value_offsets = value_offsets.as<int>();
    v->masks = masks.as<uint32_t>();
    v->values = values.as<double>();
    v->tile_col_ptr = has_csc ? tile_col_ptr.as<int>() : nullptr;
    v->tile_row_idx = has_csc ? tile_row_idx.as<int>() : nullptr;
    v->csc_tile_ids = has_csc ? csc_tile_ids.as<int>() : nullptr;
}

int ensure_record_values(DevTiles *b)
{
    if (b->transpose_of || b->rec_values_ready)
        return RTT_OK;
    RTT_TRY(build_colmajor_values(b));
    RTT_TRY(b->records.ensure(b->num_tiles * sizeof(BRecord)));
    if (b->num_tiles > 0)
        hipLaunchKernelGGL(record_values_kernel, dim3(blocks_for(b->num_tiles, kThreads)),
                           dim3(kThreads), 0, ctx().stream, b->num_tiles, b->meta.as<int4>(),
                           b->values_cm.as<double>(), b->records.as<int4>());
    RTT_HIP(hipGetLastError());
    b->rec_values_ready = true;
    return RTT_OK;
}

int build_pair_extents(const DevTiles &a, DevTiles *b)
{
    RTT_TRY(b->a_brange.ensure(a.num_tiles * sizeof(int4)));
    RTT_TRY(b->a_blight.ensure(a.num_tiles * sizeof(int2)));
    RTT_TRY(b->pair_total.ensure(sizeof(unsigned long long)));
    hipStream_t s = ctx().stream;
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, s, b->pair_total.as<int64_t>(),
                       int64_t(1), int64_t(0));
    const int cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
    if (a.num_tiles > 0)
        hipLaunchKernelGGL(pair_extent_kernel,
                           dim3(static_cast<unsigned>(std::min<int64_t>(4 * cus, blocks_for(a.num_tiles, kThreads)))),
                           dim3(kThreads), 0, s, a.num_tiles, a.meta.as<int4>(),
                           b->row_groups.as<int>(), b->a_brange.as<int4>(),
                           b->a_blight.as<int2>(), b->pair_total.as<unsigned long long>());
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

int order_by_class(DevTiles *b, BuildWork *w)
{
    hipStream_t s = ctx().stream;
    const int64_t n = b->num_tiles;
    const int R = b->tile_row_count;
    const int64_t slots = static_cast<int64_t>(kClassSlots) * R;
    RTT_TRY(b->row_groups.ensure((slots + 1) * sizeof(int)));
    RTT_TRY(w->keys_in.ensure((static_cast<size_t>(R) + 1) * sizeof(int)));   // long rows
    RTT_TRY(w->keys_out.ensure(sizeof(int)));                                 // their count
    RTT_TRY(w->meta2.ensure((n > 0 ? n : 1) * sizeof(int4)));
    RTT_TRY(b->records.ensure((n > 0 ? n : 1) * sizeof(BRecord)));
    int *n_long = w->keys_out.as<int>();
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(1), 0, s, n_long, int64_t(1), 0);
    // Meta in class order, also as record plane 0; the value planes follow
    // from it in ensure_record_values.
    hipLaunchKernelGGL(class_partition_kernel, dim3(blocks_for(static_cast<int64_t>(R) + 1, kThreads)),
                       dim3(kThreads), 0, s, R, b->tile_row_ptr.as<int>(), b->meta.as<int4>(),
                       w->meta2.as<int4>(), b->records.as<int4>(), b->row_groups.as<int>(),
                       w->keys_in.as<int>(), n_long);
    if (n > kPartRow) {
        const int cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
        hipLaunchKernelGGL(class_partition_long_kernel, dim3(static_cast<unsigned>(2 * cus)),
                           dim3(kThreads), 0, s, b->tile_row_ptr.as<int>(), b->meta.as<int4>(),
                           w->meta2.as<int4>(), b->records.as<int4>(), b->row_groups.as<int>(),
                           w->keys_in.as<int>(), n_long);
    }
    if (n > 0)
        RTT_HIP(hipMemcpyAsync(b->meta.ptr, w->meta2.ptr, n * sizeof(int4),
                               hipMemcpyDeviceToDevice, s));
    b->rec_values_ready = false;
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

// Tile columns of B for the tile-pair numeric path: the tiles of each tile
// column J in increasing tile-row order, each as {tile row K, mask, per-k
// byte sets, payload offset | dense bit}.  A stable sort of the (class
// ordered) tiles by column keeps rows ascending within a column.
__global__ void column_keys_kernel(int64_t n, const int4 *__restrict__ meta, int *__restrict__ keys)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n)
        keys[t] = meta[t].x;
}

__global__ void column_entries_kernel(int64_t n, const int *__restrict__ ids,
                                      const int *__restrict__ row_of,
                                      const int4 *__restrict__ meta, int4 *__restrict__ out)
{
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int t = ids[p];
    const int4 m = meta[t];
    out[p] = make_int4(row_of[t], m.y, m.z, m.w);
}

int build_b_columns(DevTiles *b, BuildWork *w)
{
    if (b->transpose_of || b->cols_ready)
        return RTT_OK;   // B = A^T reads A's tile rows as its columns
    hipStream_t s = ctx().stream;
    const int64_t n = b->num_tiles;
    const int BC = b->tile_col_count;
    RTT_TRY(b->col_ptr.ensure((static_cast<size_t>(BC) + 1) * sizeof(int)));
    RTT_TRY(b->col_ent.ensure((n > 0 ? n : 1) * sizeof(int4)));
    if (n > 0) {
        RTT_TRY(fill_tile_row_of(b));
        RTT_TRY(w->keys_in.ensure(n * sizeof(int)));
        RTT_TRY(w->keys_out.ensure(n * sizeof(int)));
        RTT_TRY(w->ids_in.ensure(n * sizeof(int)));
        RTT_TRY(w->ids_out.ensure(n * sizeof(int)));
        hipLaunchKernelGGL(column_keys_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                           n, b->meta.as<int4>(), w->keys_in.as<int>());
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                           w->ids_in.as<int>(), n);
        RTT_TRY(sort_pairs(w->keys_in.as<int>(), w->keys_out.as<int>(), w->ids_in.as<int>(),
                           w->ids_out.as<int>(), n, bits_for(BC), &w->scan_tmp, s));
        hipLaunchKernelGGL(column_entries_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0,
                           s, n, w->ids_out.as<int>(), b->tile_row_of.as<int>(),
                           b->meta.as<int4>(), b->col_ent.as<int4>());
    }
    hipLaunchKernelGGL(row_ptr_from_sorted_kernel, dim3(blocks_for(n + 1, kThreads)),
                       dim3(kThreads), 0, s, w->keys_out.as<int>(), n, BC, b->col_ptr.as<int>());
    RTT_HIP(hipGetLastError());
    b->cols_ready = true;
    return RTT_OK;
}

// An 8x4 tile's payload re-packed column-major: entry (r, k) goes to off +
// (rows of columns < k) + (rows < r in column k).  LANES = 8: lane r takes
// tile row r (physical slots r * 4 .. r * 4 + 3), so the lanes of a
// well-filled tile read its payload as one stretch.  LANES = 1: one lane
// walks the mask of a tile of a few entries, leaving no lanes idle.
template <int LANES>
__global__ void colmajor_values_kernel(int64_t n, const int4 *__restrict__ meta,
                                       const double *__restrict__ vals, double *__restrict__ out)
{
    static_assert(LANES == 1 || LANES == 8, "a lane per tile or per tile row");
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t t = gid / LANES;
    if (t >= n)
        return;
    const int4 m = meta[t];
    const uint32_t mask = static_cast<uint32_t>(m.y);
    const uint32_t rb = static_cast<uint32_t>(m.z);
    const uint32_t w = static_cast<uint32_t>(m.w);
    const bool dense = (w & 0x80000000u) != 0;
    const uint32_t off = w & 0x7FFFFFFFu;
    if constexpr (LANES == 1) {
        int start[4];
        int acc = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            start[k] = acc;
            acc += __popc((rb >> (8 * k)) & 0xFFu);
        }
        uint32_t mm = mask;
        int j = 0;
        while (mm) {
            const int phys = __builtin_ctz(mm);
            mm &= mm - 1;
            const int r = phys >> 2, k = phys & 3;
            const uint32_t col_rows = (rb >> (8 * k)) & 0xFFu;
            out[off + start[k] + __popc(col_rows & ((1u << r) - 1u))] =
                vals[off + (dense ? phys : j)];
            ++j;
        }
    } else {
        const int r = static_cast<int>(gid & 7);
        uint32_t bits = (mask >> (4 * r)) & 0xFu;
        int rank = __popc(mask & ((1u << (4 * r)) - 1u));   // entries of rows < r (r < 8)
        while (bits) {
            const int k = __builtin_ctz(bits);
            bits &= bits - 1;
            const int before = __popc(rb & ((1u << (8 * k)) - 1u));   // rows of columns < k
            const uint32_t col_rows = (rb >> (8 * k)) & 0xFFu;
            out[off + before + __popc(col_rows & ((1u << r) - 1u))] =
                vals[off + (dense ? 4 * r + k : rank)];
            ++rank;
        }
    }
}

__global__ void row_payload_off_kernel(int rows, const int *__restrict__ trp,
                                       const int *__restrict__ off, int *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= rows)
        out[r] = off[trp[r]];
}

int build_transpose_source(DevTiles *t, BuildWork *w)
{
    // The longest tile column decides how the B = A^T update orders B's rows.
    const int C = t->tile_col_count;
    RTT_TRY(w->counters.ensure((static_cast<size_t>(C) + 1) * sizeof(int)));
    int *counts = w->counters.as<int>();
    hipLaunchKernelGGL(fill_int_kernel, dim3(blocks_for(C + 1, kThreads)), dim3(kThreads), 0,
                       ctx().stream, counts, static_cast<int64_t>(C) + 1, 0);
    if (t->num_tiles > 0) {
        hipLaunchKernelGGL(column_count_kernel, dim3(blocks_for(t->num_tiles, kThreads)),
                           dim3(kThreads), 0, ctx().stream, t->num_tiles,
                           t->tile_col_idx.as<int>(), counts);
        hipLaunchKernelGGL(column_max_kernel, dim3(blocks_for(C, kThreads)), dim3(kThreads), 0,
                           ctx().stream, C, counts);
    }
    int longest = 0;
    RTT_HIP(hipMemcpyAsync(&longest, counts + C, sizeof(int), hipMemcpyDeviceToHost,
                           ctx().stream));
    RTT_HIP(hipStreamSynchronize(ctx().stream));
    t->max_col_tiles = longest;
    RTT_TRY(t->row_payload_off.ensure((static_cast<size_t>(t->tile_row_count) + 1) * sizeof(int)));
    hipLaunchKernelGGL(row_payload_off_kernel, dim3(blocks_for(t->tile_row_count + 1, kThreads)),
                       dim3(kThreads), 0, ctx().stream, t->tile_row_count,
                       t->tile_row_ptr.as<int>(), t->value_offsets.as<int>(),
                       t->row_payload_off.as<int>());
    RTT_TRY(t->transpose_src.ensure(static_cast<size_t>(t->num_tiles > 0 ? t->num_tiles : 1) *
                                    sizeof(BRecord)));
    if (t->num_tiles > 0)
        hipLaunchKernelGGL(transpose_source_kernel, dim3(blocks_for(t->num_tiles * kRecParts, kThreads)),
                           dim3(kThreads), 0, ctx().stream, t->num_tiles,
                           t->tile_row_of.as<int>(), t->meta.as<int4>(), t->values_cm.as<double>(),
                           t->transpose_src.as<BRecord>());
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

int build_colmajor_values(DevTiles *t)
{
    RTT_TRY(t->values_cm.ensure(static_cast<size_t>(t->payload > 0 ? t->payload : 1) *
                                sizeof(double)));
    // Tiles of a few entries on average (prolongators, Galerkin products)
    // take one lane each, fuller tiles eight.
    if (t->num_tiles > 0 && t->payload < 6 * t->num_tiles)
        hipLaunchKernelGGL((colmajor_values_kernel<1>), dim3(blocks_for(t->num_tiles, kThreads)),
                           dim3(kThreads), 0, ctx().stream, t->num_tiles, t->meta.as<int4>(),
                           t->values.as<double>(), t->values_cm.as<double>());
    else if (t->num_tiles > 0)
        hipLaunchKernelGGL((colmajor_values_kernel<8>),
                           dim3(blocks_for(t->num_tiles * 8, kThreads)), dim3(kThreads), 0,
                           ctx().stream, t->num_tiles, t->meta.as<int4>(),
                           t->values.as<double>(), t->values_cm.as<double>());
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

int fill_tile_row_of(DevTiles *t)
{
    RTT_TRY(t->tile_row_of.ensure(t->num_tiles * sizeof(int)));
    if (t->tile_row_count > 0)
        hipLaunchKernelGGL(tile_row_of_kernel, dim3(blocks_for(t->tile_row_count, kThreads)),
                           dim3(kThreads), 0, ctx().stream, t->tile_row_count,
                           t->tile_row_ptr.as<int>(), t->tile_row_of.as<int>());
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

int upload_csr(const rtt_smatrix &host, DevCsr *out, BuildWork *w, double *h2d_ms,
               double *validate_ms, double *sort_ms, int *rows_were_sorted)
{
    RTT_TRY(require_ctx());
    hipStream_t s = ctx().stream;
    if (host.m < 0 || host.n < 0 || host.nnz < 0 || (host.m > 0 && !host.rowpointer) ||
        (host.nnz > 0 && (!host.columnindex || !host.value)))
        return fail(RTT_ERR_INVALID, "malformed host CSR");
    auto t0 = Clock::now();
    out->rows = host.m;
    out->cols = host.n;
    out->nnz = host.nnz;
    RTT_TRY(out->row_ptr.ensure((static_cast<size_t>(host.m) + 1) * sizeof(int)));
    RTT_TRY(out->col_idx.ensure(static_cast<size_t>(host.nnz) * sizeof(int)));
    RTT_TRY(out->values.ensure(static_cast<size_t>(host.nnz) * sizeof(double)));
    if (host.m > 0)
        RTT_HIP(hipMemcpyAsync(out->row_ptr.ptr, host.rowpointer,
                               (static_cast<size_t>(host.m) + 1) * sizeof(int),
                               hipMemcpyHostToDevice, s));
    else {
        int zero = 0;
        RTT_HIP(hipMemcpyAsync(out->row_ptr.ptr, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    }
    if (host.nnz > 0) {
        RTT_HIP(hipMemcpyAsync(out->col_idx.ptr, host.columnindex,
                               static_cast<size_t>(host.nnz) * sizeof(int),
                               hipMemcpyHostToDevice, s));
        RTT_HIP(hipMemcpyAsync(out->values.ptr, host.value,
                               static_cast<size_t>(host.nnz) * sizeof(double),
                               hipMemcpyHostToDevice, s));
    }
    RTT_HIP(hipStreamSynchronize(s));
    *h2d_ms = ms_since(t0);

    t0 = Clock::now();
    RTT_TRY(w->counters.ensure(sizeof(int64_t)));
    int *flags = w->counters.as<int>();
    RTT_HIP(hipMemsetAsync(flags, 0, sizeof(int), s));
    const int64_t work = std::max<int64_t>(host.m, host.nnz) + 1;
    hipLaunchKernelGGL(validate_csr_kernel, dim3(blocks_for(work, kThreads)), dim3(kThreads), 0,
                       s, host.m, host.n, static_cast<int64_t>(host.nnz),
                       out->row_ptr.as<int>(), out->col_idx.as<int>(), flags);
    RTT_HIP(hipGetLastError());
    int hflags = 0;
    RTT_HIP(hipMemcpyAsync(&hflags, flags, sizeof(int), hipMemcpyDeviceToHost, s));
    RTT_HIP(hipStreamSynchronize(s));
    *validate_ms = ms_since(t0);
    if (hflags & 3)
        return fail(RTT_ERR_INVALID, "CSR validation failed (flags %d)", hflags);
    *rows_were_sorted = (hflags & 4) ? 0 : 1;

    t0 = Clock::now();
    out->reordered = false;
    if (hflags & 4) {
        // Stable per-row sort by column; duplicates keep their input order.
        // input_of[x] = input position of sorted entry x, kept for value updates.
        DevBuf keys, ids_in, vals;
        RTT_TRY(keys.ensure(static_cast<size_t>(host.nnz) * sizeof(int)));
        RTT_TRY(ids_in.ensure(static_cast<size_t>(host.nnz) * sizeof(int)));
        RTT_TRY(vals.ensure(static_cast<size_t>(host.nnz) * sizeof(double)));
        RTT_TRY(out->input_of.ensure(static_cast<size_t>(host.nnz) * sizeof(int)));
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(host.nnz, kThreads)), dim3(kThreads), 0,
                           s, ids_in.as<int>(), static_cast<int64_t>(host.nnz));
        size_t bytes = 0;
        const int *rp = out->row_ptr.as<int>();
        RTT_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(
            nullptr, bytes, out->col_idx.as<int>(), keys.as<int>(), ids_in.as<int>(),
            out->input_of.as<int>(), host.nnz, host.m, rp, rp + 1, 0, bits_for(host.n), s));
        RTT_TRY(w->scan_tmp.ensure(bytes));
        RTT_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(
            w->scan_tmp.ptr, bytes, out->col_idx.as<int>(), keys.as<int>(), ids_in.as<int>(),
            out->input_of.as<int>(), host.nnz, host.m, rp, rp + 1, 0, bits_for(host.n), s));
        hipLaunchKernelGGL(gather_f64_kernel, dim3(blocks_for(host.nnz, kThreads)),
                           dim3(kThreads), 0, s, static_cast<int64_t>(host.nnz),
                           out->input_of.as<int>(), out->values.as<double>(), vals.as<double>());
        RTT_HIP(hipGetLastError());
        RTT_HIP(hipMemcpyAsync(out->col_idx.ptr, keys.ptr,
                               static_cast<size_t>(host.nnz) * sizeof(int),
                               hipMemcpyDeviceToDevice, s));
        RTT_HIP(hipMemcpyAsync(out->values.ptr, vals.ptr,
                               static_cast<size_t>(host.nnz) * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
        RTT_HIP(hipStreamSynchronize(s));
        out->reordered = true;
    }
    static std::atomic<uint64_t> next_structure_id{1};
    out->structure_id = next_structure_id.fetch_add(1);
    *sort_ms = ms_since(t0);
    return RTT_OK;
}

int build_tiles(const DevCsr &src, const int *row_map, int active_rows, int active_cols,
                int logical_rows, int logical_cols, bool col_major, int dense_threshold,
                bool want_csc, DevTiles *out, BuildWork *w, bool want_source_map)
{
    RTT_TRY(require_ctx());
    const int TR = col_major ? kTileK : kTileM;
    const int TC = col_major ? kTileN : kTileK;
    out->rows = logical_rows;
    out->cols = logical_cols;
    out->tile_rows = TR;
    out->tile_cols = TC;
    out->tile_row_count = (active_rows + TR - 1) / TR;
    out->tile_col_count = (active_cols + TC - 1) / TC;
    out->has_csc = false;
    out->arrays_ready = true;
    out->cols_ready = false;
    out->transpose_of = nullptr;
    hipStream_t s = ctx().stream;
    int2 *src_slot = nullptr;
    int *dup_flag = nullptr;
    out->map_source_id = 0;
    out->map_nnz = 0;
    out->map_dups = false;
    if (want_source_map) {
        RTT_TRY(out->src_slot.ensure(static_cast<size_t>(src.nnz > 0 ? src.nnz : 1) *
                                     sizeof(int2)));
        RTT_TRY(w->flag.ensure(sizeof(int)));
        src_slot = out->src_slot.as<int2>();
        dup_flag = w->flag.as<int>();
        hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(1), 0, s, dup_flag, int64_t(1), 0);
        if (src.nnz > 0)
            hipLaunchKernelGGL(fill_int_kernel, dim3(blocks_for(2 * src.nnz, kThreads)),
                               dim3(kThreads), 0, s, reinterpret_cast<int *>(src_slot),
                               2 * src.nnz, -1);
        RTT_HIP(hipGetLastError());
    }
    if (col_major)
        RTT_TRY((run_tile_build<kTileK, kTileN, true>(src, row_map, active_rows, active_cols,
                                                      dense_threshold, out, w, src_slot,
                                                      dup_flag)));
    else
        RTT_TRY((run_tile_build<kTileM, kTileK, false>(src, row_map, active_rows, active_cols,
                                                       dense_threshold, out, w, src_slot,
                                                       dup_flag)));
    if (want_source_map) {
        int dups = 0;
        RTT_HIP(hipMemcpyAsync(&dups, dup_flag, sizeof(int), hipMemcpyDeviceToHost, s));
        RTT_HIP(hipStreamSynchronize(s));
        out->map_source_id = src.structure_id;
        out->map_nnz = src.nnz;
        out->map_dups = dups != 0;
    }
    if (want_csc)
        RTT_TRY(build_csc(out, w));
    return RTT_OK;
}

int transpose_tiles(const DevTiles &a, int b_cols, DevTiles *b, BuildWork *w)
{
    RTT_TRY(require_ctx());
    hipStream_t s = ctx().stream;
    const int64_t n = a.num_tiles;
    b->rows = a.cols;
    b->cols = b_cols;
    b->tile_rows = kTileK;
    b->tile_cols = kTileN;
    b->tile_row_count = a.tile_col_count;
    b->tile_col_count = (b_cols + kTileN - 1) / kTileN;
    b->num_tiles = n;
    b->payload = a.payload;
    b->dense = a.dense;
    b->sparse = a.sparse;
    b->transpose_of = &a;
    b->arrays_ready = false;
    b->has_csc = false;
    b->cols_ready = false;
    const int R = b->tile_row_count;
    const int64_t slots = static_cast<int64_t>(kClassSlots) * R;
    RTT_TRY(b->tile_row_ptr.ensure((R + 1) * sizeof(int)));
    RTT_TRY(b->row_groups.ensure((slots + 1) * sizeof(int)));
    RTT_TRY(b->sorted_ids.ensure(n * sizeof(int)));
    RTT_TRY(w->ids_in.ensure(n * sizeof(int)));
    RTT_TRY(w->keys_in.ensure(n * sizeof(int)));
    RTT_TRY(w->keys_out.ensure(n * sizeof(int)));
    if (n > 0 && a.max_col_tiles >= 0 && a.max_col_tiles <= kShortCol &&
        n < (int64_t(1) << kOrderIdBits)) {
        // The same order by counting: B's rows are short.
        RTT_TRY(w->keys_out.ensure((std::max<int64_t>(n, R + 1)) * sizeof(int)));
        int *counts = w->keys_out.as<int>();
        int *rp = b->tile_row_ptr.as<int>();
        hipLaunchKernelGGL(fill_int_kernel, dim3(blocks_for(R + 1, kThreads)), dim3(kThreads), 0,
                           s, counts, static_cast<int64_t>(R) + 1, 0);
        hipLaunchKernelGGL(column_rank_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0,
                           s, n, a.tile_col_idx.as<int>(), a.masks.as<uint32_t>(), counts,
                           w->keys_in.as<int>());
        size_t bytes = 0;
        RTT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, counts, rp, R + 1, s));
        RTT_TRY(w->scan_tmp.ensure(bytes));
        RTT_HIP(hipcub::DeviceScan::ExclusiveSum(w->scan_tmp.ptr, bytes, counts, rp, R + 1, s));
        hipLaunchKernelGGL(column_scatter_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads),
                           0, s, n, a.tile_col_idx.as<int>(), w->keys_in.as<int>(), rp,
                           w->ids_in.as<uint32_t>());
        hipLaunchKernelGGL(order_rows_kernel, dim3(blocks_for(R, kOrderRows)), dim3(kThreads), 0,
                           s, R, n, rp, w->ids_in.as<uint32_t>(), b->sorted_ids.as<int>(),
                           b->row_groups.as<int>());
        RTT_TRY(b->records.ensure(n * sizeof(BRecord)));
        hipLaunchKernelGGL(transposed_records_kernel, dim3(blocks_for(n, kThreads)),
                           dim3(kThreads), 0, s, n, b->sorted_ids.as<int>(),
                           a.transpose_src.as<BRecord>(), b->records.as<int4>());
        RTT_HIP(hipGetLastError());
        return RTT_OK;
    }
    if (n > 0) {
        // Stable radix sort of A's tiles by (tile column, class): B's row
        // order with every row's heavy tiles first, then its light tiles by
        // the row k they occupy.
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                           w->ids_in.as<int>(), n);
        hipLaunchKernelGGL(class_keys_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                           n, a.tile_col_idx.as<int>(), a.meta.as<int4>(), w->keys_in.as<int>());
        RTT_TRY(sort_pairs(w->keys_in.as<int>(), w->keys_out.as<int>(), w->ids_in.as<int>(),
                           b->sorted_ids.as<int>(), n, bits_for(slots), &w->scan_tmp, s));
        RTT_TRY(b->records.ensure(n * sizeof(BRecord)));
        hipLaunchKernelGGL(transposed_records_kernel, dim3(blocks_for(n, kThreads)),
                           dim3(kThreads), 0, s, n, b->sorted_ids.as<int>(),
                           a.transpose_src.as<BRecord>(), b->records.as<int4>());
    }
    hipLaunchKernelGGL(row_ptr_from_sorted_kernel, dim3(blocks_for(n + 1, kThreads)),
                       dim3(kThreads), 0, s, w->keys_out.as<int>(), n, static_cast<int>(slots),
                       b->row_groups.as<int>());
    hipLaunchKernelGGL(group_row_ptr_kernel, dim3(blocks_for(R + 1, kThreads)), dim3(kThreads), 0,
                       s, R, b->row_groups.as<int>(), b->tile_row_ptr.as<int>());
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

int materialize_tiles(DevTiles *b, BuildWork *w)
{
    RTT_TRY(require_ctx());
    hipStream_t s = ctx().stream;
    const int64_t n = b->num_tiles;
    if (!b->arrays_ready) {
        const DevTiles &a = *b->transpose_of;
        const int R = b->tile_row_count;
        const int C = b->tile_col_count;
        RTT_TRY(b->tile_col_idx.ensure(n * sizeof(int)));
        RTT_TRY(b->masks.ensure(n * sizeof(uint32_t)));
        RTT_TRY(b->value_offsets.ensure((n + 1) * sizeof(int)));
        RTT_TRY(b->values.ensure(b->payload * sizeof(double)));
        RTT_TRY(b->tile_col_ptr.ensure((C + 1) * sizeof(int)));
        RTT_TRY(b->tile_row_idx.ensure(n * sizeof(int)));
        RTT_TRY(b->csc_tile_ids.ensure(n * sizeof(int)));
        RTT_TRY(w->spans.ensure((n + 1) * sizeof(int64_t)));
        RTT_TRY(w->counters.ensure((n + 1) * sizeof(int64_t)));
        if (n > 0) {
            // The timed update orders each B row heavy tiles first; the
            // standard arrays keep the tiles of a row in column order, which
            // a sort of A's tiles by column alone gives.
            RTT_TRY(w->ids_in.ensure(n * sizeof(int)));
            RTT_TRY(w->keys_out.ensure(n * sizeof(int)));
            RTT_TRY(w->ids_out.ensure(n * sizeof(int)));
            hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                               w->ids_in.as<int>(), n);
            RTT_TRY(sort_pairs(a.tile_col_idx.as<int>(), w->keys_out.as<int>(),
                               w->ids_in.as<int>(), w->ids_out.as<int>(), n, bits_for(R),
                               &w->scan_tmp, s));
            const int *col_order = w->ids_out.as<int>();
            int64_t *spans = w->spans.as<int64_t>();
            hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, s, spans + n, int64_t(1),
                               int64_t(0));
            hipLaunchKernelGGL(transpose_meta_kernel, dim3(blocks_for(n, kThreads)),
                               dim3(kThreads), 0, s, n, col_order,
                               a.tile_row_of.as<int>(), a.masks.as<uint32_t>(),
                               a.value_offsets.as<int>(), b->tile_col_idx.as<int>(),
                               b->masks.as<uint32_t>(), spans, b->csc_tile_ids.as<int>());
            int64_t *off64 = w->counters.as<int64_t>();
            RTT_TRY(exclusive_scan_i64(spans, off64, n, &w->scan_tmp, s));
            hipLaunchKernelGGL(transpose_payload_kernel, dim3(blocks_for(n * 32, kThreads)),
                               dim3(kThreads), 0, s, n, col_order,
                               a.value_offsets.as<int>(), a.values.as<double>(), off64,
                               b->value_offsets.as<int>(), b->values.as<double>());
            hipLaunchKernelGGL(narrow_i64_kernel, dim3(1), dim3(1), 0, s, off64 + n,
                               b->value_offsets.as<int>() + n, int64_t(1));
            // CSC of B = CSR of A: tile_row_idx = A.tile_col_idx and
            // tile_col_ptr = A.tile_row_ptr padded to C + 1 entries.
            RTT_HIP(hipMemcpyAsync(b->tile_row_idx.ptr, a.tile_col_idx.ptr, n * sizeof(int),
                                   hipMemcpyDeviceToDevice, s));
        } else {
            int zero = 0;
            RTT_HIP(hipMemcpyAsync(b->value_offsets.ptr, &zero, sizeof(int),
                                   hipMemcpyHostToDevice, s));
        }
        const int ar = a.tile_row_count;
        RTT_HIP(hipMemcpyAsync(b->tile_col_ptr.ptr, a.tile_row_ptr.ptr, (ar + 1) * sizeof(int),
                               hipMemcpyDeviceToDevice, s));
        if (C > ar) {
            std::vector<int> tail(C - ar, static_cast<int>(n));
            RTT_HIP(hipMemcpyAsync(b->tile_col_ptr.as<int>() + ar + 1, tail.data(),
                                   tail.size() * sizeof(int), hipMemcpyHostToDevice, s));
            RTT_HIP(hipStreamSynchronize(s));
        }
        b->arrays_ready = true;
        b->has_csc = true;
    }
    if (!b->has_csc)
        RTT_TRY(build_csc(b, w));
    RTT_HIP(hipGetLastError());
    RTT_HIP(hipStreamSynchronize(s));
    return RTT_OK;
}

int transpose_csr(const DevCsr &src, DevCsr *out, BuildWork *w)
{
    RTT_TRY(require_ctx());
    hipStream_t s = ctx().stream;
    const int64_t nnz = src.nnz;
    out->rows = src.cols;
    out->cols = src.rows;
    out->nnz = nnz;
    RTT_TRY(out->row_ptr.ensure((static_cast<size_t>(src.cols) + 1) * sizeof(int)));
    RTT_TRY(out->col_idx.ensure(nnz * sizeof(int)));
    RTT_TRY(out->values.ensure(nnz * sizeof(double)));
    RTT_TRY(w->row_tiles.ensure((static_cast<size_t>(src.cols) + 1) * sizeof(int64_t)));
    RTT_TRY(w->row_payload.ensure((static_cast<size_t>(src.cols) + 1) * sizeof(int64_t)));
    int64_t *cnt = w->row_tiles.as<int64_t>();
    hipLaunchKernelGGL(set_i64_kernel, dim3(blocks_for(src.cols + 1, kThreads)), dim3(kThreads),
                       0, s, cnt, static_cast<int64_t>(src.cols) + 1, int64_t(0));
    if (nnz > 0) {
        RTT_TRY(w->ids_in.ensure(nnz * sizeof(int)));
        RTT_TRY(w->ids_out.ensure(nnz * sizeof(int)));
        RTT_TRY(w->keys_out.ensure(nnz * sizeof(int)));
        RTT_TRY(w->spans.ensure(nnz * sizeof(int)));
        hipLaunchKernelGGL(histogram_kernel, dim3(blocks_for(nnz, kThreads)), dim3(kThreads), 0,
                           s, src.col_idx.as<int>(), nnz, cnt);
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(nnz, kThreads)), dim3(kThreads), 0, s,
                           w->ids_in.as<int>(), nnz);
        RTT_TRY(sort_pairs(src.col_idx.as<int>(), w->keys_out.as<int>(), w->ids_in.as<int>(),
                           w->ids_out.as<int>(), nnz, bits_for(src.cols), &w->scan_tmp, s));
        int *entry_row = w->spans.as<int>();
        hipLaunchKernelGGL(entry_rows_kernel, dim3(blocks_for(src.rows, kThreads)),
                           dim3(kThreads), 0, s, src.rows, src.row_ptr.as<int>(), entry_row);
        hipLaunchKernelGGL(gather_transposed_kernel, dim3(blocks_for(nnz, kThreads)),
                           dim3(kThreads), 0, s, nnz, w->ids_out.as<int>(), entry_row,
                           src.values.as<double>(), out->col_idx.as<int>(),
                           out->values.as<double>());
    }
    int64_t *ptr64 = w->row_payload.as<int64_t>();
    RTT_TRY(exclusive_scan_i64(cnt, ptr64, src.cols, &w->scan_tmp, s));
    hipLaunchKernelGGL(narrow_i64_kernel, dim3(blocks_for(src.cols + 1, kThreads)),
                       dim3(kThreads), 0, s, ptr64, out->row_ptr.as<int>(),
                       static_cast<int64_t>(src.cols) + 1);
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

int nnz_cub(const DevCsr &csr, bool aat, unsigned long long *result, BuildWork *w)
{
    RTT_TRY(require_ctx());
    hipStream_t s = ctx().stream;
    RTT_TRY(w->counters.ensure(sizeof(unsigned long long)));
    unsigned long long *acc = w->counters.as<unsigned long long>();
    RTT_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long), s));
    const int64_t *deg = nullptr;
    if (aat) {
        RTT_TRY(w->row_tiles.ensure((static_cast<size_t>(csr.cols) + 1) * sizeof(int64_t)));
        int64_t *d = w->row_tiles.as<int64_t>();
        hipLaunchKernelGGL(set_i64_kernel, dim3(blocks_for(csr.cols + 1, kThreads)),
                           dim3(kThreads), 0, s, d, static_cast<int64_t>(csr.cols) + 1,
                           int64_t(0));
        if (csr.nnz > 0)
            hipLaunchKernelGGL(histogram_kernel, dim3(blocks_for(csr.nnz, kThreads)),
                               dim3(kThreads), 0, s, csr.col_idx.as<int>(), csr.nnz, d);
        deg = d;
    }
    if (csr.nnz > 0)
        hipLaunchKernelGGL(nnz_cub_kernel,
                           dim3(static_cast<unsigned>(std::min<int64_t>(1024, blocks_for(csr.nnz, kThreads)))),
                           dim3(kThreads),
                           0, s, csr.nnz, csr.col_idx.as<int>(), csr.row_ptr.as<int>(), deg,
                           acc);
    RTT_HIP(hipGetLastError());
    RTT_HIP(hipMemcpyAsync(result, acc, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    RTT_HIP(hipStreamSynchronize(s));
    return RTT_OK;
}

int set_csr_values(DevCsr *csr, const double *host_values, BuildWork *w)
{
    RTT_TRY(require_ctx());
    hipStream_t s = ctx().stream;
    if (csr->nnz == 0)
        return RTT_OK;
    if (!host_values)
        return fail(RTT_ERR_INVALID, "NULL values for a CSR with %lld entries",
                    (long long)csr->nnz);
    const size_t bytes = static_cast<size_t>(csr->nnz) * sizeof(double);
    if (!csr->reordered) {
        RTT_HIP(hipMemcpyAsync(csr->values.ptr, host_values, bytes, hipMemcpyHostToDevice, s));
    } else {
        RTT_TRY(w->staging.ensure(bytes));
        RTT_HIP(hipMemcpyAsync(w->staging.ptr, host_values, bytes, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(gather_f64_kernel, dim3(blocks_for(csr->nnz, kThreads)),
                           dim3(kThreads), 0, s, csr->nnz, csr->input_of.as<int>(),
                           w->staging.as<double>(), csr->values.as<double>());
        RTT_HIP(hipGetLastError());
    }
    RTT_HIP(hipStreamSynchronize(s));
    return RTT_OK;
}

int update_b_values(const DevCsr &src, bool clear, DevTiles *b)
{
    RTT_TRY(require_ctx());
    hipStream_t s = ctx().stream;
    if (b->map_source_id == 0 || b->map_source_id != src.structure_id ||
        b->map_nnz != src.nnz || b->transpose_of)
        return fail(RTT_ERR_INVALID, "B has no source map for this CSR");
    // The k-major copy and the record value planes are kept current only
    // once a kernel has asked for them (ensure_record_values); otherwise they
    // are rebuilt from the payload when next needed.
    const bool kmajor = b->rec_values_ready;
    if (clear && b->payload > 0) {
        hipLaunchKernelGGL(fill_f64_kernel, dim3(blocks_for(b->payload, kThreads)),
                           dim3(kThreads), 0, s, b->values.as<double>(), b->payload, 0.0);
        if (kmajor)
            hipLaunchKernelGGL(fill_f64_kernel, dim3(blocks_for(b->payload, kThreads)),
                               dim3(kThreads), 0, s, b->values_cm.as<double>(), b->payload, 0.0);
    }
    if (src.nnz > 0)
        hipLaunchKernelGGL(scatter_runs_kernel, dim3(blocks_for(src.nnz, kThreads)),
                           dim3(kThreads), 0, s, src.nnz, b->src_slot.as<int2>(),
                           src.values.as<double>(), b->values.as<double>(),
                           kmajor ? b->values_cm.as<double>() : nullptr);
    if (kmajor && b->num_tiles > 0)
        hipLaunchKernelGGL(record_values_kernel, dim3(blocks_for(b->num_tiles, kThreads)),
                           dim3(kThreads), 0, s, b->num_tiles, b->meta.as<int4>(),
                           b->values_cm.as<double>(), b->records.as<int4>());
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

int nnz_cub_ab(const DevCsr &a, const DevCsr &b, unsigned long long *result, BuildWork *w)
{
    RTT_TRY(require_ctx());
    if (a.cols != b.rows)
        return fail(RTT_ERR_INVALID, "A has %d columns but B has %d rows", a.cols, b.rows);
    hipStream_t s = ctx().stream;
    RTT_TRY(w->counters.ensure(sizeof(unsigned long long)));
    unsigned long long *acc = w->counters.as<unsigned long long>();
    RTT_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long), s));
    if (a.nnz > 0)
        hipLaunchKernelGGL(nnz_cub_kernel,
                           dim3(static_cast<unsigned>(std::min<int64_t>(1024, blocks_for(a.nnz, kThreads)))),
                           dim3(kThreads), 0, s, a.nnz, a.col_idx.as<int>(), b.row_ptr.as<int>(),
                           static_cast<const int64_t *>(nullptr), acc);
    RTT_HIP(hipGetLastError());
    RTT_HIP(hipMemcpyAsync(result, acc, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    RTT_HIP(hipStreamSynchronize(s));
    return RTT_OK;
}

} // namespace rtt
