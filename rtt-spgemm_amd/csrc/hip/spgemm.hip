// SpGEMM core: symbolic -> allocate -> numeric for C = A * B with A in 8x4
// row-major tiles, B in 4x8 column-major tiles and C in 8x8 tiles.
//
// Work unit: one A tile row I (8 scalar rows), one 256-thread workgroup.
// Its (A tile, B tile) pairs are A(I,K) x B(K,J) for every A tile of row I
// and every B tile of B's tile row K -- a tile-level Gustavson expansion, so
// no K-list merge or candidate sort is needed (the reference reaches the same
// pairs through Super16 candidates + exact intersection,
// src/super16_symbolic.cuh:2406-2774, and a K-merge per C tile,
// src/dmma_spgemm.h:4307-4410).
//
// Pairs are enumerated per chunk of up to kChunk A tiles: every A tile's
// B-row extent comes from a per-A-tile table built with B, the pair count is
// scanned, and an LDS segment table maps pair index -> A tile, so a thread
// reaches its (A tile, B tile) with two LDS reads.
//
//  * symbolic: the 8x8 masks of the row's C tiles are ORed into an LDS hash
//    table, compacted, ranked by tile column and staged; after the size scan
//    a light pass writes the final structure.  Rows whose tile set outgrows
//    the table use a direct-mapped mask slab in HBM instead.
//  * numeric: the row's C layout (tile columns, masks, value offsets) and
//    its A payload (column-major per tile) are loaded into LDS.  Each thread
//    takes a pair, loads its 80-byte B record (packed meta + the first eight
//    values in k-major order), walks B's entries row by row with A's column
//    in registers, and adds each scalar product into an LDS FP64 accumulator
//    laid out exactly like the row's output segment, which is finally
//    streamed to HBM.  Rows that exceed the LDS budgets fall back to a
//    workgroup path with FP64 atomics in HBM.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "device.hpp"
#include "internal.hpp"

namespace rtt {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;

constexpr int kChunk = 128;        // A tiles per enumeration step
constexpr int kSeg = 1024;         // pairs per segment-table window

constexpr int kSymSlots = 1024;    // LDS hash slots (power of 2)
constexpr int kSymSlotBits = 10;
constexpr int kSymPerThread = kSymSlots / kThreads;
constexpr int kSymProbe = 32;      // longest probe before a row goes to the slab path
constexpr int kSymBatch = 4;       // pairs whose loads a thread issues together

constexpr int kWSlots = 256;       // hash slots of the one-wave symbolic table
constexpr int kWSlotBits = 8;
static_assert(kWSlots == 1 << kWSlotBits, "one-wave table size");
constexpr int kWStage = 3 * kWSlots / 4;   // C tiles a row may have there
constexpr int kWChunk = kWave;     // A tiles per one-wave enumeration step
constexpr int kWSeg = 512;         // pairs per one-wave segment window
constexpr int kWBatch = 6;         // pairs whose loads a lane of that kernel issues together

constexpr int kColRegs = 3;         // values of an A column a numeric lane holds in registers
constexpr int kExtVals = 3;         // B values past the record a numeric lane prefetches

// The numeric instances: threads per workgroup (four pair runs of kChunk A
// tiles each), pairs per segment-table window, and LDS budgets:
// accumulator slots, staged A values, C tiles, hash slots (at least twice
// the tiles).  TinyRow is one wave with 3 KB of LDS against SmallRow's
// 20 KB, so that a CU holds four times as many of the short rows of sparse
// operands (prolongators, Galerkin products) in flight.  A second one-wave
// instance with a 512-value accumulator measured no faster than SmallRow
// on the rows between the two.
struct TinyRow {
    static constexpr int kThreads = kWave, kChunk = kWave / 4, kSeg = 128;
    static constexpr bool kLeavesTiny = false;
    static constexpr int kAcc = 128, kAStage = 64, kTiles = 16, kSlots = 32, kSlotBits = 5;
};
// SmallRow, launched over every row, leaves the rows within TinyRow's
// budgets to that instance (launched over their list).
struct SmallRow {
    static constexpr int kThreads = 256, kChunk = 64, kSeg = 1024;
    static constexpr bool kLeavesTiny = true;
    static constexpr int kAcc = 1024, kAStage = 256, kTiles = 128, kSlots = 256, kSlotBits = 8;
};
// MediumRow takes rows past SmallRow's accumulator with few C tiles and A
// values (narrow bands squared) at a third of LargeRow's LDS.
struct MediumRow {
    static constexpr int kThreads = 256, kChunk = 64, kSeg = 1024;
    static constexpr bool kLeavesTiny = false;
    static constexpr int kAcc = 2048, kAStage = 256, kTiles = 64, kSlots = 128, kSlotBits = 7;
};
struct LargeRow {
    static constexpr int kThreads = 256, kChunk = 64, kSeg = 1024;
    static constexpr bool kLeavesTiny = false;
    static constexpr int kAcc = 4096, kAStage = 2048, kTiles = 512, kSlots = 1024, kSlotBits = 10;
};
static_assert(LargeRow::kAcc <= 65536 && LargeRow::kTiles <= 32768, "packed tile slots");
static_assert(TinyRow::kAcc <= SmallRow::kAcc && TinyRow::kAStage <= SmallRow::kAStage &&
                  TinyRow::kTiles <= SmallRow::kTiles,
              "nested row budgets");

// Whether a row of the LDS level with these sizes fits the instance Cfg.
template <class Cfg>
__device__ __forceinline__ bool fits_row(int64_t nt, int64_t nnz, int64_t a_len)
{
    return nt <= Cfg::kTiles && nnz <= Cfg::kAcc && a_len <= Cfg::kAStage;
}

constexpr int kWriteBuf = 2048;    // column bytes staged per row in the write pass

constexpr int kEmpty = -1;
constexpr uint32_t kDenseBit = 0x80000000u;

// kLevelLarge: the row outgrew the wave table and waits for the workgroup
// table; kLevelSlab: it outgrew that too.
enum RowLevel : int { kLevelEmpty = 0, kLevelLds = 1, kLevelLarge = 2, kLevelSlab = 3 };

struct Operands {
    int a_tile_rows;
    const int *__restrict__ a_trp;
    const int4 *__restrict__ a_meta;
    const int4 *__restrict__ a_brange;   // per A tile: first B tile, heavy, count
    const int2 *__restrict__ a_blight;   // per A tile: the light B tiles it can meet
    const int *__restrict__ a_off;
    const double *__restrict__ a_vals;
    const double *__restrict__ a_vals_cm;   // A payload, each tile column-major
    const int *__restrict__ a_row_off;      // A payload offset of each tile row
    const double *__restrict__ b_vals_km;   // B payload, each tile k-major
    const int *__restrict__ b_trp;
    const int4 *__restrict__ b_meta;        // B records' plane 0
    const int *__restrict__ b_groups;       // per B row: start of each class run (kClassSlots)
    const double *__restrict__ b_vals;
    int b_tile_cols;
    float tile_fill;   // rows whose A tiles hold this many values on average go to the
                       // tile-pair kernel (+inf: none)
};

// The tile-pair kernel takes the tile rows whose A tiles hold at least
// `fill` stored values on average; the LDS-accumulator kernels the others.
__device__ __forceinline__ bool tile_routed(float fill, int tiles, int payload)
{
    return tiles > 0 && static_cast<float>(payload) >= fill * static_cast<float>(tiles);
}

struct COut {
    const int64_t *__restrict__ tile_ptr;     // tilem + 1
    const int64_t *__restrict__ row_nnz_off;  // tilem + 1
    int *__restrict__ col;
    uint64_t *__restrict__ mask;
    int64_t *__restrict__ tile_nnz;
    uint8_t *__restrict__ csr_ptr;
    uint8_t *__restrict__ csr_col;
    double *__restrict__ val;
};

struct Stage {
    int *__restrict__ j;
    unsigned long long *__restrict__ m;
    int stride;
};


__device__ __forceinline__ uint32_t slot_of(int j, int bits)
{
    return hash_tile_col(j) >> (32 - bits);
}

__device__ __forceinline__ uint32_t byte_of(uint32_t v, int k) { return (v >> (8 * k)) & 0xFFu; }

// Structural product from the per-k row/column byte sets of the A tile
// (rows present in column k) and the B tile (columns present in row k).
__device__ __forceinline__ uint64_t bytes_product(uint32_t a_bytes, uint32_t b_bytes)
{
    uint64_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        out |= rows_to_bytes(byte_of(a_bytes, k)) &
               (static_cast<uint64_t>(byte_of(b_bytes, k)) * 0x0101010101010101ull);
    return out;
}

__device__ __forceinline__ int product_count(uint32_t a_bytes, uint32_t b_bytes)
{
    int n = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        n += __popc(byte_of(a_bytes, k)) * __popc(byte_of(b_bytes, k));
    return n;
}

// Byte r of the result is the number of set bits in bytes 0..r-1 of m (the
// u8 row offsets of a C tile): per-byte popcounts, then an exclusive byte
// prefix sum by multiplication (sums stay below 256, so no byte carries).
__device__ __forceinline__ uint64_t packed_row_offsets(uint64_t m)
{
    uint64_t x = m - ((m >> 1) & 0x5555555555555555ull);
    x = (x & 0x3333333333333333ull) + ((x >> 2) & 0x3333333333333333ull);
    x = (x + (x >> 4)) & 0x0F0F0F0F0F0F0F0Full;
    return (x * 0x0101010101010101ull) << 8;
}

__device__ __forceinline__ void write_tile_structure(const COut &c, int64_t t, int j,
                                                     uint64_t m, int64_t off)
{
    c.col[t] = j;
    c.mask[t] = m;
    c.tile_nnz[t] = off;
    reinterpret_cast<uint64_t *>(c.csr_ptr)[t] = packed_row_offsets(m);
    int i = 0;
    while (m) {
        const int pos = __builtin_ctzll(m);
        m &= m - 1;
        c.csr_col[off + i++] = static_cast<uint8_t>(pos & 7);
    }
}

// Exclusive scans of two values over a workgroup of NW waves with one
// barrier; s2 (2 * NW ints) must not be reused without a barrier between.
template <int NW = kWavesPerBlock>
__device__ __forceinline__ void dual_exclusive_scan(int a, int b, int *s2, int *pa, int *pb,
                                                    int *ta, int *tb)
{
    const int w = threadIdx.x / kWave;
    const int ia = wave_inclusive_scan(a);
    const int ib = wave_inclusive_scan(b);
    if (lane_id() == kWave - 1) {
        s2[w] = ia;
        s2[NW + w] = ib;
    }
    __syncthreads();
    int ba = 0, bb = 0, sa = 0, sb = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const int x = s2[i], y = s2[NW + i];
        ba += i < w ? x : 0;
        bb += i < w ? y : 0;
        sa += x;
        sb += y;
    }
    *pa = ba + ia - a;
    *pb = bb + ib - b;
    *ta = sa;
    *tb = sb;
}

// Exclusive scan over a workgroup of NW waves; *total receives the sum.
template <int NW = kWavesPerBlock>
__device__ __forceinline__ int block_exclusive_scan(int v, int *s_tmp, int *total)
{
    const int w = threadIdx.x / kWave;
    const int incl = wave_inclusive_scan(v);
    if (lane_id() == kWave - 1)
        s_tmp[w] = incl;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const int t = s_tmp[i];
        before += i < w ? t : 0;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *s_tmp)
{
    v = wave_sum(v);
    if (lane_id() == 0)
        s_tmp[threadIdx.x / kWave] = v;
    __syncthreads();
    T all = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; ++i)
        all += s_tmp[i];
    __syncthreads();
    return all;
}

// Segment table of pairs [r0, r1): s_seg[p - r0] = chunk-local A tile.
// Thread t < n owns A tile t with pairs [pre_t, pre_t + nb_t).
__device__ __forceinline__ void fill_segments(unsigned char *s_seg, int pre, int nb, int r0,
                                              int r1)
{
    const int lo = max(pre, r0), hi = min(pre + nb, r1);
    for (int p = lo; p < hi; ++p)
        s_seg[p - r0] = static_cast<unsigned char>(threadIdx.x);
}

// ---------------------------------------------------------------------------
// Symbolic, one wave per row: the same hash -> compact -> rank -> stage as the
// workgroup kernel below, with a kWSlots table, so that a CU holds four times
// as many rows and their load latencies overlap.  Pairs come in two runs per
// chunk of kWChunk A tiles, the heavy and then the light tiles of the B
// rows; lane t owns both runs of tile t.  A pair's mask sums only the k where
// both A's column and B's row are nonempty.  A row that outgrows the table is marked
// kLevelLarge for the workgroup kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t v)
{
    const uint32_t nz = (((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u;
    return ((nz >> 7) | (nz >> 14) | (nz >> 21) | (nz >> 28)) & 0xFu;
}

__global__ __launch_bounds__(kWave) void symbolic_wave_kernel(
    Operands g, Stage st, int64_t *__restrict__ row_tiles, int64_t *__restrict__ row_nnz,
    int *__restrict__ row_level)
{
    __shared__ __align__(16) int s_key[kWSlots];
    __shared__ unsigned long long s_msk[kWSlots];
    __shared__ unsigned long long s_spread[4][kWChunk];
    __shared__ int s_acls[kWChunk];   // nonempty columns (bits 0-3)
    __shared__ int s_pw[2 * kWChunk]; // per (run, tile): first B tile minus first pair
    __shared__ unsigned char s_seg[kWSeg];
    static_assert(2 * kWChunk <= 256, "segment entries are bytes");

    const int row = static_cast<int>(blockIdx.x);
    const int lane = threadIdx.x;
    const int a0 = g.a_trp[row], a1 = g.a_trp[row + 1];
    for (int i = lane; i < kWSlots; i += kWave) {
        s_key[i] = kEmpty;
        s_msk[i] = 0ull;
    }
    bool ovf = false;
    for (int c0 = a0; c0 < a1 && !ovf; c0 += kWChunk) {
        const int n = min(kWChunk, a1 - c0);
        int nh = 0, bh = 0, nl = 0, bls = 0;
        if (lane < n) {
            const int4 am = g.a_meta[c0 + lane];
            const int4 br = g.a_brange[c0 + lane];
            const int2 bl = g.a_blight[c0 + lane];
            const uint32_t ab = static_cast<uint32_t>(am.z);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                s_spread[k][lane] = rows_to_bytes(byte_of(ab, k));
            s_acls[lane] = static_cast<int>(nonzero_bytes(ab));
            nh = br.y;
            bh = br.x;
            nl = bl.y;
            bls = bl.x;
        }
        // Pair offsets: the heavy runs of all tiles, then the light runs.
        const int ih = wave_inclusive_scan(nh), il = wave_inclusive_scan(nl);
        const int heavy_total = __shfl(ih, kWave - 1, kWave);
        const int total = heavy_total + __shfl(il, kWave - 1, kWave);
        const int ph = ih - nh, pl = heavy_total + il - nl;
        s_pw[lane] = bh - ph;
        s_pw[kWChunk + lane] = bls - pl;
        for (int r0 = 0; r0 < total && !ovf; r0 += kWSeg) {
            const int r1 = min(total, r0 + kWSeg);
            for (int p = max(ph, r0); p < min(ph + nh, r1); ++p)
                s_seg[p - r0] = static_cast<unsigned char>(lane);
            for (int p = max(pl, r0); p < min(pl + nl, r1); ++p)
                s_seg[p - r0] = static_cast<unsigned char>(kWChunk + lane);
            __syncthreads();
            bool lane_ovf = false;
            for (int p0 = r0 + lane; p0 < r1; p0 += kWBatch * kWave) {
                int4 bm[kWBatch];
                int ai[kWBatch];
#pragma unroll
                for (int u = 0; u < kWBatch; ++u) {
                    const int p = p0 + u * kWave;
                    ai[u] = -1;
                    bm[u] = make_int4(0, 0, 0, 0);
                    if (p < r1) {
                        const int v = s_seg[p - r0];
                        ai[u] = v % kWChunk;
                        bm[u] = g.b_meta[s_pw[v] + p];
                    }
                }
#pragma unroll
                for (int u = 0; u < kWBatch; ++u) {
                    if (ai[u] < 0)
                        continue;
                    const uint32_t bb = static_cast<uint32_t>(bm[u].z);
                    uint32_t kk = static_cast<uint32_t>(s_acls[ai[u]]) & nonzero_bytes(bb);
                    uint64_t pm = 0;
                    while (kk) {
                        const int k = __builtin_ctz(kk);
                        kk &= kk - 1;
                        const uint32_t b4 = byte_of(bb, k) * 0x01010101u;
                        pm |= s_spread[k][ai[u]] & ((static_cast<uint64_t>(b4) << 32) | b4);
                    }
                    if (!pm)
                        continue;
                    const int j = bm[u].x;
                    uint32_t h = slot_of(j, kWSlotBits);
                    for (int probe = 0; probe < kSymProbe; ++probe) {
                        int k = s_key[h];
                        if (k == kEmpty) {
                            k = atomicCAS(&s_key[h], kEmpty, j);
                            if (k == kEmpty) {
                                atomicOr(&s_msk[h], static_cast<unsigned long long>(pm));
                                break;
                            }
                        }
                        if (k == j) {
                            atomicOr(&s_msk[h], static_cast<unsigned long long>(pm));
                            break;
                        }
                        h = (h + 1) & (kWSlots - 1);
                        if (probe == kSymProbe - 1)
                            lane_ovf = true;
                    }
                }
            }
            ovf = __any(lane_ovf);
            __syncthreads();
        }
    }
    // Compact the occupied slots to the front; the scan total is the tile count.
    constexpr int kPer = kWSlots / kWave;
    int rk[kPer];
    unsigned long long rm[kPer];
    int mine = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        rk[i] = s_key[lane * kPer + i];
        rm[i] = s_msk[lane * kPer + i];
        mine += rk[i] != kEmpty ? 1 : 0;
    }
    const int incl = wave_inclusive_scan(mine);
    const int nt = __shfl(incl, kWave - 1, kWave);
    if (ovf || nt > kWStage) {
        if (lane == 0)
            row_level[row] = kLevelLarge;
        return;
    }
    __syncthreads();
    int pos = incl - mine;
    long long nz = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i)
        if (rk[i] != kEmpty) {
            s_key[pos] = rk[i];
            s_msk[pos] = rm[i];
            nz += __popcll(rm[i]);
            ++pos;
        }
    __syncthreads();
    // Rank by counting: entry t goes to the number of smaller columns.
    const int64_t sb = static_cast<int64_t>(row) * st.stride;
    const int4 *k4 = reinterpret_cast<const int4 *>(s_key);
    const int full = nt & ~3;
    for (int t = lane; t < nt; t += kWave) {
        const int mj = s_key[t];
        int rank = 0;
        for (int f = 0; f < full / 4; ++f) {
            const int4 kk = k4[f];
            rank += (kk.x < mj ? 1 : 0) + (kk.y < mj ? 1 : 0) + (kk.z < mj ? 1 : 0) +
                    (kk.w < mj ? 1 : 0);
        }
        for (int f = full; f < nt; ++f)
            rank += s_key[f] < mj ? 1 : 0;
        st.j[sb + rank] = mj;
        st.m[sb + rank] = s_msk[t];
    }
    nz = wave_sum(nz);
    if (lane == 0) {
        row_level[row] = nt > 0 ? kLevelLds : kLevelEmpty;
        row_tiles[row] = nt;
        row_nnz[row] = nz;
    }
}

// ---------------------------------------------------------------------------
// Symbolic: per-row hash of C tile masks, compaction, ranking, staging.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void symbolic_count_kernel(
    Operands g, int stage_limit, Stage st, const int *__restrict__ rows,
    int64_t *__restrict__ row_tiles, int64_t *__restrict__ row_nnz, int *__restrict__ row_level)
{
    __shared__ __align__(16) int s_key[kSymSlots];
    __shared__ unsigned long long s_msk[kSymSlots];
    __shared__ unsigned long long s_spread[4][kChunk];
    __shared__ int s_b0[kChunk];
    __shared__ int s_pre[kChunk];
    __shared__ unsigned char s_seg[kSeg];
    __shared__ int s_tmp[kWavesPerBlock];
    __shared__ long long s_tmp64[kWavesPerBlock];
    __shared__ int s_ovf;

    const int row = rows ? rows[blockIdx.x] : static_cast<int>(blockIdx.x);
    const int tid = threadIdx.x;
    for (int i = tid; i < kSymSlots; i += kThreads) {
        s_key[i] = kEmpty;
        s_msk[i] = 0ull;
    }
    if (tid == 0)
        s_ovf = 0;
    const int a0 = g.a_trp[row], a1 = g.a_trp[row + 1];
    for (int c0 = a0; c0 < a1; c0 += kChunk) {
        const int n = min(kChunk, a1 - c0);
        int nb = 0;
        if (tid < n) {
            const uint32_t ab = static_cast<uint32_t>(g.a_meta[c0 + tid].z);
            const int4 br = g.a_brange[c0 + tid];
            nb = br.z;
            s_b0[tid] = br.x;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                s_spread[k][tid] = rows_to_bytes(byte_of(ab, k));
        }
        int total = 0;
        const int pre = block_exclusive_scan(nb, s_tmp, &total);
        if (tid < n)
            s_pre[tid] = pre;
        for (int r0 = 0; r0 < total; r0 += kSeg) {
            const int r1 = min(total, r0 + kSeg);
            if (tid < n)
                fill_segments(s_seg, pre, nb, r0, r1);
            __syncthreads();
            // Each thread takes kSymBatch pairs and issues their B-meta loads
            // together before using any of them.
            for (int p0 = r0 + tid; p0 < r1; p0 += kSymBatch * kThreads) {
                int4 bm[kSymBatch];
                int ai[kSymBatch];
#pragma unroll
                for (int u = 0; u < kSymBatch; ++u) {
                    const int p = p0 + u * kThreads;
                    ai[u] = -1;
                    bm[u] = make_int4(0, 0, 0, 0);
                    if (p < r1) {
                        ai[u] = s_seg[p - r0];
                        bm[u] = g.b_meta[s_b0[ai[u]] + (p - s_pre[ai[u]])];
                    }
                }
#pragma unroll
                for (int u = 0; u < kSymBatch; ++u) {
                    if (ai[u] < 0)
                        continue;
                    const uint32_t bb = static_cast<uint32_t>(bm[u].z);
                    uint64_t pm = 0;
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        pm |= s_spread[k][ai[u]] &
                              (static_cast<uint64_t>(byte_of(bb, k)) * 0x0101010101010101ull);
                    if (!pm)
                        continue;
                    const int j = bm[u].x;
                    uint32_t h = slot_of(j, kSymSlotBits);
                    // A probe longer than kSymProbe means the table is close
                    // to full: the row is sent to the slab path, which is
                    // always correct, instead of probing on.
                    for (int probe = 0; probe < kSymProbe; ++probe) {
                        int k = s_key[h];
                        if (k == kEmpty) {
                            k = atomicCAS(&s_key[h], kEmpty, j);
                            if (k == kEmpty) {
                                atomicOr(&s_msk[h], static_cast<unsigned long long>(pm));
                                break;
                            }
                        }
                        if (k == j) {
                            atomicOr(&s_msk[h], static_cast<unsigned long long>(pm));
                            break;
                        }
                        h = (h + 1) & (kSymSlots - 1);
                        if (probe == kSymProbe - 1)
                            s_ovf = 1;
                    }
                }
            }
            __syncthreads();
            if (s_ovf)
                break;
        }
        if (s_ovf)
            break;
    }
    __syncthreads();
    // Compact the occupied slots to the front; the scan total is the row's
    // tile count.  A probe that ran out (a table close to full) or more tiles
    // than the staging holds sends the row to the slab path.
    int rk[kSymPerThread];
    unsigned long long rm[kSymPerThread];
    int mine = 0;
#pragma unroll
    for (int i = 0; i < kSymPerThread; ++i) {
        rk[i] = s_key[tid * kSymPerThread + i];
        rm[i] = s_msk[tid * kSymPerThread + i];
        mine += rk[i] != kEmpty ? 1 : 0;
    }
    int nt = 0;
    int pos = block_exclusive_scan(mine, s_tmp, &nt);
    if (s_ovf || nt > stage_limit) {
        if (tid == 0) {
            row_level[row] = kLevelSlab;
            row_tiles[row] = 0;
            row_nnz[row] = 0;
        }
        return;
    }
    long long nz = 0;
#pragma unroll
    for (int i = 0; i < kSymPerThread; ++i)
        if (rk[i] != kEmpty) {
            s_key[pos] = rk[i];
            s_msk[pos] = rm[i];
            nz += __popcll(rm[i]);
            ++pos;
        }
    __syncthreads();
    const int64_t sb = static_cast<int64_t>(row) * st.stride;
    if (nt <= kThreads) {
        // Rank by counting: entry t goes to the number of smaller columns.
        // The keys are read four at a time, the last nt % 4 one by one.
        if (tid < nt) {
            const int mj = s_key[tid];
            const int4 *k4 = reinterpret_cast<const int4 *>(s_key);
            int rank = 0;
            const int full = nt & ~3;
            for (int f = 0; f < full / 4; ++f) {
                const int4 kk = k4[f];
                rank += (kk.x < mj ? 1 : 0) + (kk.y < mj ? 1 : 0) + (kk.z < mj ? 1 : 0) +
                        (kk.w < mj ? 1 : 0);
            }
            for (int f = full; f < nt; ++f)
                rank += s_key[f] < mj ? 1 : 0;
            st.j[sb + rank] = mj;
            st.m[sb + rank] = s_msk[tid];
        }
    } else {
        int npow = 1;
        while (npow < nt)
            npow <<= 1;
        for (int i = nt + tid; i < npow; i += kThreads)
            s_key[i] = INT_MAX;
        __syncthreads();
        for (int k = 2; k <= npow; k <<= 1) {
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                for (int i = tid; i < npow; i += kThreads) {
                    const int ixj = i ^ jj;
                    if (ixj > i) {
                        const int ki = s_key[i], kx = s_key[ixj];
                        const bool up = (i & k) == 0;
                        if ((ki > kx) == up) {
                            s_key[i] = kx;
                            s_key[ixj] = ki;
                            const unsigned long long t = s_msk[i];
                            s_msk[i] = s_msk[ixj];
                            s_msk[ixj] = t;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int i = tid; i < nt; i += kThreads) {
            st.j[sb + i] = s_key[i];
            st.m[sb + i] = s_msk[i];
        }
    }
    const long long row_nz = block_sum(nz, s_tmp64);
    if (tid == 0) {
        row_level[row] = nt > 0 ? kLevelLds : kLevelEmpty;
        row_tiles[row] = nt;
        row_nnz[row] = row_nz;
    }
}

// Staged (column, mask) lists -> final C structure; one wave per row.  The
// row's local-column bytes are assembled in LDS and written coalesced: a
// table gives, for a row byte x of a tile mask, the columns of its set bits
// packed in increasing order, and each tile row ORs them into the zeroed
// buffer with at most two 64-bit LDS ORs.
__global__ __launch_bounds__(kThreads) void symbolic_write_kernel(
    const int *__restrict__ rows, int count, Stage st, const int *__restrict__ row_level, COut c,
    float tile_fill = 0.f, const int *__restrict__ a_trp = nullptr,
    const int *__restrict__ a_row_off = nullptr)
{
    __shared__ unsigned long long s_buf[kWavesPerBlock][kWriteBuf / 8];
    __shared__ unsigned long long s_cols[256];
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    {
        const int x = threadIdx.x;
        unsigned long long v = 0;
        int nset = 0;
#pragma unroll
        for (int cl = 0; cl < 8; ++cl)
            if ((x >> cl) & 1)
                v |= static_cast<unsigned long long>(cl) << (8 * nset++);
        s_cols[x] = v;
    }
    __syncthreads();
    const int idx = blockIdx.x * kWavesPerBlock + w;
    if (idx >= count)
        return;
    const int row = rows ? rows[idx] : idx;
    if (row_level[row] != kLevelLds)
        return;
    // Without a list: the rows the tile-pair kernel takes.
    if (!rows && !tile_routed(tile_fill, a_trp[row + 1] - a_trp[row],
                              a_row_off[row + 1] - a_row_off[row]))
        return;
    const int64_t tb = c.tile_ptr[row];
    const int nt = static_cast<int>(c.tile_ptr[row + 1] - tb);
    const int64_t zb = c.row_nnz_off[row];
    const int64_t nnz_row = c.row_nnz_off[row + 1] - zb;
    const bool staged = nnz_row <= kWriteBuf;
    const int64_t sb = static_cast<int64_t>(row) * st.stride;
    unsigned long long *buf = s_buf[w];
    if (staged) {
        for (int i = lane; i < kWriteBuf / 8; i += kWave)
            buf[i] = 0ull;
        wave_sync();
    }
    int64_t carry = 0;
    for (int q0 = 0; q0 < nt; q0 += kWave) {
        const int q = q0 + lane;
        uint64_t m = q < nt ? st.m[sb + q] : 0ull;
        const int pc = __popcll(m);
        const int incl = wave_inclusive_scan(pc);
        if (q < nt) {
            const int64_t off = carry + (incl - pc);
            const int64_t t = tb + q;
            const uint64_t ro = packed_row_offsets(m);
            c.col[t] = st.j[sb + q];
            c.mask[t] = m;
            c.tile_nnz[t] = zb + off;
            reinterpret_cast<uint64_t *>(c.csr_ptr)[t] = ro;
            if (staged) {
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const uint32_t x = static_cast<uint32_t>(m >> (8 * r)) & 0xFFu;
                    if (!x)
                        continue;
                    const int p = static_cast<int>(off) + static_cast<int>((ro >> (8 * r)) & 0xFFu);
                    const unsigned long long v = s_cols[x];
                    const int sh = (p & 7) * 8;
                    atomicOr(&buf[p >> 3], v << sh);
                    if (sh && __popc(x) * 8 + sh > 64)
                        atomicOr(&buf[(p >> 3) + 1], v >> (64 - sh));
                }
            } else {
                int i = 0;
                while (m) {
                    const int pos = __builtin_ctzll(m);
                    m &= m - 1;
                    c.csr_col[zb + off + i] = static_cast<uint8_t>(pos & 7);
                    ++i;
                }
            }
        }
        carry += __shfl(incl, kWave - 1, kWave);
    }
    if (staged) {
        wave_sync();
        const unsigned char *bytes = reinterpret_cast<const unsigned char *>(buf);
        for (int64_t i = lane; i < nnz_row; i += kWave)
            c.csr_col[zb + i] = bytes[i];
    }
}

// Work queue of the slab kernels: the workgroup's next position in the row
// list, taken by one lane with a global atomic and shared through LDS.  The
// head only grows, so every workgroup leaves once it passes the list's end.
// The row bodies end with a barrier, so the slot is not rewritten while a
// lane still reads it.
__device__ __forceinline__ int next_row(int *queue)
{
    __shared__ int s_next;
    if (threadIdx.x == 0)
        s_next = atomicAdd(queue, 1);
    __syncthreads();
    return s_next;
}

// Cost of each listed row: the (A tile, B tile) pairs it visits.  One wave
// per row.
__global__ __launch_bounds__(kThreads) void row_pairs_kernel(const int *__restrict__ rows, int n,
                                                             const int *__restrict__ a_trp,
                                                             const int4 *__restrict__ a_brange,
                                                             unsigned int *__restrict__ cost)
{
    const int i = blockIdx.x * kWavesPerBlock + static_cast<int>(threadIdx.x) / kWave;
    if (i >= n)
        return;
    const int row = rows[i];
    unsigned long long sum = 0;
    for (int t = a_trp[row] + lane_id(); t < a_trp[row + 1]; t += kWave)
        sum += static_cast<unsigned int>(a_brange[t].z);
    sum = wave_sum(sum);
    if (lane_id() == 0)
        cost[i] = static_cast<unsigned int>(sum < 0xffffffffull ? sum : 0xffffffffull);
}

// Rows whose tile set outgrows the LDS table: one workgroup per row, one
// 64-bit mask per C tile column in an HBM slab, read back in column order.
__device__ __forceinline__ unsigned long long load_coherent(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The slab kernels walk a row in windows: each A tile keeps a cursor into
// every class run of its B row, each run sorted by column.
constexpr int kSlabRuns = 5;       // class runs of a B row: heavy, light k = 0..3
constexpr int kSlabWinTiles = 4096;  // A tiles up to which a slab row is walked in windows
static_assert(kSlabRuns + 1 <= kClassSlots, "the runs' ends come from the class slots");
constexpr int kSymWin = 2048;      // B tile columns a symbolic slab window covers

// First index after a, within [a, end), whose B tile column is >= j, given
// that a's column is below j (end when none): doubling steps from a, then a
// binary search in the last step.
__device__ __forceinline__ int gallop_col(const int4 *__restrict__ meta, int a, int end, int j)
{
    int step = 1;
    while (a + step < end && meta[a + step].x < j) {
        a += step;
        step <<= 1;
    }
    int lo = a + 1, hi = min(a + step, end);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (meta[mid].x < j)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// The count pass leaves each row's (column, mask) list, in column order, in
// a grow-only staging buffer when it fits (row_off[row] = its first entry,
// -1 otherwise); the fill pass then copies the list instead of repeating
// the row's atomics.  cursor sums every row's tile count, fitted or not, so
// that the caller learns the capacity the next call needs.
struct SlabStage {
    int *j = nullptr;
    unsigned long long *m = nullptr;
    int64_t *row_off = nullptr;
    unsigned long long *cursor = nullptr;
    unsigned long long capacity = 0;
};

template <bool FILL>
__global__ __launch_bounds__(kThreads) void symbolic_slab_kernel(
    Operands g, const int *__restrict__ rows, int nrows, int *__restrict__ queue,
    int static_rows, unsigned long long *__restrict__ slabs, int64_t *__restrict__ row_tiles,
    int64_t *__restrict__ row_nnz, SlabStage ss, COut c)
{
    __shared__ uint32_t s_ab[kThreads];
    __shared__ int s_tmp[kWavesPerBlock];
    __shared__ int s_lo, s_hi, s_min;
    __shared__ long long s_base;
    __shared__ unsigned long long s_wm[kSymWin];   // a window's masks
    __shared__ int s_wpw[kSlabRuns * kThreads];    // per (tile, run): first B tile minus first pair
    __shared__ int s_wpre[kSlabRuns * kThreads];   // per (tile, run): first pair
    unsigned long long *slab = slabs + static_cast<int64_t>(blockIdx.x) * g.b_tile_cols;
    const int tid = threadIdx.x;
    for (int i = tid; i < kSymWin; i += kThreads)
        s_wm[i] = 0;
    for (int step = 0;; ++step) {
        const int ri = static_rows ? static_cast<int>(blockIdx.x) + step * static_cast<int>(gridDim.x)
                                   : next_row(queue);
        if (ri >= nrows)
            break;
        const int row = rows[ri];
        if constexpr (FILL) {
            const int64_t base = ss.row_off ? ss.row_off[row] : -1;
            if (base >= 0) {
                // Staged: the structure in column order, value offsets by a scan.
                const int64_t tb = c.tile_ptr[row];
                const int nt = static_cast<int>(c.tile_ptr[row + 1] - tb);
                int64_t zb = c.row_nnz_off[row];
                for (int q0 = 0; q0 < nt; q0 += kThreads) {
                    const int q = q0 + tid;
                    unsigned long long m = 0;
                    int j = 0;
                    if (q < nt) {
                        j = ss.j[base + q];
                        m = ss.m[base + q];
                    }
                    int npc = 0;
                    const int zpos = block_exclusive_scan(__popcll(m), s_tmp, &npc);
                    if (q < nt)
                        write_tile_structure(c, tb + q, j, m, zb + zpos);
                    zb += npc;
                }
                __syncthreads();
                continue;
            }
        }
        const int a0 = g.a_trp[row], a1 = g.a_trp[row + 1];
        if (tid == 0) {
            s_lo = INT_MAX;
            s_hi = -1;
            s_min = INT_MAX;
        }
        __syncthreads();
        // The column range the row's masks land in, from the pairs
        // themselves (the tiles of a B row are not in column order).
        int jlo = INT_MAX, jhi = -1;
        // Each chunk's A tiles walk their B runs in windows of kSymWin
        // columns, starting each window at the lowest column a cursor holds
        // (so a chunk visits only windows it has pairs in).  A window's masks
        // are ORed in LDS, then its nonzero masks into the HBM table at
        // consecutive addresses.
        for (int c0 = a0; c0 < a1; c0 += kThreads) {
            const int n = min(kThreads, a1 - c0);
            int cur[kSlabRuns], end[kSlabRuns];
#pragma unroll
            for (int r = 0; r < kSlabRuns; ++r)
                cur[r] = end[r] = 0;
            if (tid < n) {
                const int4 am = g.a_meta[c0 + tid];
                const uint32_t ab = static_cast<uint32_t>(am.z);
                s_ab[tid] = ab;
                const int *gp = g.b_groups + static_cast<int64_t>(kClassSlots) * am.x;
#pragma unroll
                for (int r = 0; r < kSlabRuns; ++r) {
                    // Light run k meets only A tiles with entries in column k.
                    if (r == 0 || byte_of(ab, r - 1) != 0) {
                        cur[r] = gp[r];
                        end[r] = gp[r + 1];
                    }
                }
            }
            for (;;) {
                int mn = INT_MAX;
#pragma unroll
                for (int r = 0; r < kSlabRuns; ++r)
                    if (cur[r] < end[r])
                        mn = min(mn, g.b_meta[cur[r]].x);
                if (mn != INT_MAX)
                    atomicMin(&s_min, mn);
                __syncthreads();
                const int wlo = s_min;
                if (wlo == INT_MAX)
                    break;
                const int whi = wlo + kSymWin;
                int cnt[kSlabRuns], lo[kSlabRuns];
                int mine = 0;
#pragma unroll
                for (int r = 0; r < kSlabRuns; ++r) {
                    lo[r] = cur[r];
                    cnt[r] = 0;
                    if (cur[r] < end[r] && g.b_meta[cur[r]].x < whi) {
                        const int h = gallop_col(g.b_meta, cur[r], end[r], whi);
                        cnt[r] = h - cur[r];
                        cur[r] = h;
                    }
                    mine += cnt[r];
                }
                int total = 0;
                int pre = block_exclusive_scan(mine, s_tmp, &total);
#pragma unroll
                for (int r = 0; r < kSlabRuns; ++r) {
                    s_wpre[kSlabRuns * tid + r] = pre;
                    s_wpw[kSlabRuns * tid + r] = lo[r] - pre;
                    pre += cnt[r];
                }
                if (tid == 0)
                    s_min = INT_MAX;   // every thread read it before the scan's barriers
                __syncthreads();
                for (int p = tid; p < total; p += kThreads) {
                    const int e = upper_index(s_wpre, kSlabRuns * n, p);
                    const int4 bm = g.b_meta[s_wpw[e] + p];
                    const uint64_t pm = bytes_product(s_ab[e / kSlabRuns], static_cast<uint32_t>(bm.z));
                    if (pm && bm.x >= wlo && bm.x < whi)
                        atomicOr(&s_wm[bm.x - wlo], static_cast<unsigned long long>(pm));
                }
                __syncthreads();
                for (int i = tid; i < kSymWin; i += kThreads) {
                    const unsigned long long m = s_wm[i];
                    if (m) {
                        s_wm[i] = 0;
                        atomicOr(&slab[wlo + i], m);
                        jlo = min(jlo, wlo + i);
                        jhi = max(jhi, wlo + i);
                    }
                }
                __syncthreads();
            }
            // Every thread has read the final (empty) window start before the
            // next chunk lowers it again.
            __syncthreads();
        }
        if (jhi >= 0) {
            atomicMin(&s_lo, jlo);
            atomicMax(&s_hi, jhi);
        }
        __threadfence();
        __syncthreads();
        const int lo = s_lo, hi = s_hi;
        if constexpr (!FILL) {
            // Sizes first, without consuming the masks.
            int64_t tiles = 0, nz = 0;
            for (int j0 = lo; j0 <= hi; j0 += kThreads) {
                const int j = j0 + tid;
                const unsigned long long m = j <= hi ? load_coherent(&slab[j]) : 0ull;
                int ntile = 0, npc = 0;
                block_exclusive_scan(m ? 1 : 0, s_tmp, &ntile);
                block_exclusive_scan(__popcll(m), s_tmp, &npc);
                tiles += ntile;
                nz += npc;
            }
            if (tid == 0) {
                long long b = -1;
                if (ss.cursor && tiles > 0) {
                    const unsigned long long at =
                        atomicAdd(ss.cursor, static_cast<unsigned long long>(tiles));
                    if (at + static_cast<unsigned long long>(tiles) <= ss.capacity)
                        b = static_cast<long long>(at);
                }
                s_base = b;
                row_tiles[row] = tiles;
                row_nnz[row] = nz;
                if (ss.row_off)
                    ss.row_off[row] = b;
            }
            __syncthreads();
            // Then the list, staged when it fits, and the slab cleared.
            const long long base = s_base;
            int64_t done = 0;
            for (int j0 = lo; j0 <= hi; j0 += kThreads) {
                const int j = j0 + tid;
                unsigned long long m = 0;
                if (j <= hi) {
                    m = load_coherent(&slab[j]);
                    if (m)
                        slab[j] = 0;
                }
                int ntile = 0;
                const int tpos = block_exclusive_scan(m ? 1 : 0, s_tmp, &ntile);
                if (base >= 0 && m) {
                    ss.j[base + done + tpos] = j;
                    ss.m[base + done + tpos] = m;
                }
                done += ntile;
            }
        } else {
            // Not staged: the masks again, written straight to C's structure.
            int64_t tiles = 0, nz = 0;
            const int64_t tb = c.tile_ptr[row];
            const int64_t zb = c.row_nnz_off[row];
            for (int j0 = lo; j0 <= hi; j0 += kThreads) {
                const int j = j0 + tid;
                unsigned long long m = 0;
                if (j <= hi) {
                    m = load_coherent(&slab[j]);
                    if (m)
                        slab[j] = 0;
                }
                int ntile = 0, npc = 0;
                const int tpos = block_exclusive_scan(m ? 1 : 0, s_tmp, &ntile);
                const int zpos = block_exclusive_scan(__popcll(m), s_tmp, &npc);
                if (m)
                    write_tile_structure(c, tb + tiles + tpos, j, m, zb + nz + zpos);
                tiles += ntile;
                nz += npc;
            }
        }
        __threadfence();
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Numeric, LDS path: one workgroup per row, LDS accumulator.
// ---------------------------------------------------------------------------
// The row's output segment is an FP64 accumulator in LDS, every C tile of
// the row is in a hash table, and each thread takes the pairs p,
// p + kThreads, ... of a window.  A pair walks its B tile's entries in
// k-major order (its record holds them so): when the walk reaches row k of
// B, the lane loads the first values of column k of the A tile (staged
// column-major in LDS) into registers, and B entry (k, cl) adds
// A(r, k) * B(k, cl) into C(r, cl) for every row r of that column.
//
// The accumulator has the output's layout: a tile's entries row by row,
// columns ascending.  When every row of the C tile is one contiguous run of
// columns and every column of the A tile one contiguous run of rows, the
// slot of C(r, cl) is (first slot of row r - its first column) + cl: the
// lane takes those bases for the rows of A column k once per k, and each
// product costs one add.  Other tiles take the slot from a popcount of the
// tile mask.
// Hash slot of a C tile: {tile column or kEmpty, first value slot | local
// tile index << 16, row_runs low word, row_runs high word}, so that one
// LDS read gives a pair everything the walk needs of its C tile.
__device__ __forceinline__ int4 make_tile_slot(int j, int base, int q, uint64_t runs)
{
    return make_int4(j, base | (q << 16), static_cast<int>(static_cast<uint32_t>(runs)),
                     static_cast<int>(static_cast<uint32_t>(runs >> 32)));
}

// Byte value of a row whose columns are not one contiguous run.
constexpr uint32_t kNoRun = 0x80u;

__device__ __forceinline__ bool is_run(uint32_t bits)
{
    const uint32_t x = bits >> __builtin_ctz(bits);
    return (x & (x + 1)) == 0;
}

// Byte r: (slot of row r's first entry within the tile) - (its first
// column) when the row's columns are contiguous, kNoRun when they are not,
// 0 for an empty row (never addressed).
__device__ __forceinline__ uint64_t row_runs(uint64_t m)
{
    uint64_t out = 0;
    int first = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const uint32_t bits = static_cast<uint32_t>(m >> (8 * r)) & 0xFFu;
        uint32_t b = 0;
        if (bits)
            b = is_run(bits) ? static_cast<uint32_t>(first - __builtin_ctz(bits)) & 0xFFu : kNoRun;
        out |= static_cast<uint64_t>(b) << (8 * r);
        first += __popc(bits);
    }
    return out;
}

__device__ __forceinline__ int run_base(uint64_t runs, int r)
{
    return static_cast<int8_t>((runs >> (8 * r)) & 0xFFu);
}

// Byte k: rows of A column k as count | first row << 4, with bit 7 set when
// they are not one contiguous run of at most kColRegs rows.
__device__ __forceinline__ uint32_t a_column_runs(uint32_t ab)
{
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t rk = byte_of(ab, k);
        if (rk) {
            const bool regs = is_run(rk) && __popc(rk) <= kColRegs;
            out |= ((__popc(rk) | (__builtin_ctz(rk) << 4) | (regs ? 0u : kNoRun)) << (8 * k));
        }
    }
    return out;
}

// Sampled start/end times of the numeric kernel's waves (the timeline
// trace): workgroups b with b mod 2^shift == phase record, per wave, the
// compute unit and the device's constant-rate wall clock at entry and exit.
// Null pointers switch it off.
struct Timeline {
    unsigned long long *start = nullptr, *end = nullptr;
    unsigned int *cu = nullptr;
    unsigned int shift = 0, phase = 0;
};

__device__ __forceinline__ void trace_mark(const Timeline &tl, bool at_end)
{
    if (!tl.start || lane_id() != 0)
        return;
    const unsigned int b = blockIdx.x;
    if ((b & ((1u << tl.shift) - 1u)) != tl.phase)
        return;
    const size_t slot = static_cast<size_t>((b - tl.phase) >> tl.shift) * kWavesPerBlock +
                        threadIdx.x / kWave;
    const unsigned long long t = static_cast<unsigned long long>(wall_clock64());
    if (at_end) {
        tl.end[slot] = t;
    } else {
        tl.cu[slot] = __smid();
        tl.start[slot] = t;
    }
}

template <class Cfg, bool TRACE = false>
__global__ __launch_bounds__(Cfg::kThreads) void numeric_row_kernel(
    Operands g, const int4 *__restrict__ recs, int64_t rec_n, const int *__restrict__ rows,
    const int *__restrict__ row_level, Stage st, COut c, Timeline tl)
{
    constexpr int kT = Cfg::kThreads, kCh = Cfg::kChunk, kSg = Cfg::kSeg, kNW = kT / kWave;
    static_assert(kT == 4 * kCh && kT % kWave == 0, "one thread per (pair run, chunk tile)");
    __shared__ double s_acc[Cfg::kAcc];
    __shared__ double s_acol[Cfg::kAStage + kColRegs];
    __shared__ unsigned long long s_cmask[Cfg::kTiles];   // C tile masks, bit r * 8 + cl
    __shared__ int s_cbase[Cfg::kTiles];   // first slot of each tile within the row's segment
    __shared__ int4 s_slot[Cfg::kSlots];   // hash of the row's C tiles (make_tile_slot)
    // Per chunk A tile: its staged info and B runs (see stage_tile), and per
    // (run, tile) the first B record minus the run's first pair.
    __shared__ int4 s_at[kCh];
    __shared__ int4 s_br[kCh];
    __shared__ int s_pw[kT];
    __shared__ unsigned char s_seg[kSg];
    __shared__ int s_tmp[kNW];
    __shared__ int s_dual[2 * kNW];

    const int tid = threadIdx.x;
    if constexpr (TRACE)
        trace_mark(tl, false);
    // Rows (or list entries) in 64-row chunks per XCD: the B records that
    // neighbouring rows share stay in one L2 for a run of 64 rows, and the
    // window of rows in flight still spans the chip.  Against row b in
    // workgroup b: the same speed or 1 % better on the stencil step, a
    // third less L2 fetch traffic, and 0.5 % on the batched forms' lists.
    const int bx = xcd_chunk_row<64>(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    const int row = rows ? rows[bx] : bx;
    // The row's inputs come in two rounds of independent loads: its level,
    // C's row extents, A's tile and payload extents (the payload base per
    // tile row is precomputed); then the first staged C tile, A value and A
    // tile of each thread, before the first barrier.
    const int level = row_level[row];
    const int64_t t0 = c.tile_ptr[row];
    const int nt = static_cast<int>(c.tile_ptr[row + 1] - t0);
    const int64_t zb = c.row_nnz_off[row];
    const int64_t nnz_row = c.row_nnz_off[row + 1] - zb;
    const int a0 = g.a_trp[row], a1 = g.a_trp[row + 1];
    const int a_base = g.a_row_off[row];
    const int a_len = g.a_row_off[row + 1] - a_base;
    bool skip = (level != kLevelLds) | (nt > Cfg::kTiles) | (nnz_row > Cfg::kAcc) |
                (a_len > Cfg::kAStage) | tile_routed(g.tile_fill, a1 - a0, a_len);
    if constexpr (Cfg::kLeavesTiny)
        skip |= fits_row<TinyRow>(nt, nnz_row, a_len);
    if (skip) {
        if constexpr (TRACE)
            trace_mark(tl, true);
        return;
    }
    const int64_t sb = static_cast<int64_t>(row) * st.stride;
    int j0 = 0;
    uint64_t m0 = 0;
    if (tid < nt) {
        j0 = st.j[sb + tid];
        m0 = st.m[sb + tid];
    }
    double av0 = 0.0;
    if (tid < a_len)
        av0 = g.a_vals_cm[a_base + tid];
    int4 am0 = make_int4(0, 0, 0, 0), br0 = make_int4(0, 0, 0, 0);
    int2 bl0 = make_int2(0, 0);
    if (tid < min(kCh, a1 - a0)) {
        am0 = g.a_meta[a0 + tid];
        br0 = g.a_brange[a0 + tid];
        bl0 = g.a_blight[a0 + tid];
    }
    for (int i = tid; i < Cfg::kSlots; i += kT)
        s_slot[i] = make_int4(kEmpty, 0, 0, 0);
    for (int i = tid; i < nnz_row; i += kT)
        s_acc[i] = 0.0;
    if (tid < a_len)
        s_acol[tid] = av0;
    for (int i = tid + kT; i < a_len; i += kT)
        s_acol[i] = g.a_vals_cm[a_base + i];
    if (tid < kColRegs)
        s_acol[a_len + tid] = 0.0;   // the register loads of a column may run past it
    // Chunk tile info: {rows present in column k (byte k), heavy (bit 0) and
    // first staged value of column k (byte k, k > 0), first staged value,
    // row runs of the columns (a_column_runs)} and its B runs {first heavy
    // tile, heavy count, first light tile, light count}.
    auto stage_tile = [&](const int4 &am, const int4 &br, const int2 &bl) {
        const uint32_t ab = static_cast<uint32_t>(am.z);
        const uint32_t p1 = __popc(byte_of(ab, 0));
        const uint32_t p2 = p1 + __popc(byte_of(ab, 1));
        const uint32_t p3 = p2 + __popc(byte_of(ab, 2));
        const uint32_t heavy = __popc(static_cast<uint32_t>(am.y)) >= kHeavyMin ? 1u : 0u;
        s_at[tid] = make_int4(static_cast<int>(ab),
                              static_cast<int>(heavy | (p1 << 8) | (p2 << 16) | (p3 << 24)),
                              static_cast<int>(static_cast<uint32_t>(am.w) & ~kDenseBit) - a_base,
                              static_cast<int>(a_column_runs(ab)));
        s_br[tid] = make_int4(br.x, br.y, bl.x, bl.y);
    };
    // The chunk's pairs come in four runs, one per (A class, B class): heavy
    // and light A tiles times the heavy and light runs of their B rows, so
    // that the lanes of a wave take pairs of like cost.  Thread v stands for
    // run v / kCh of tile v % kCh.
    const int run = tid / kCh, vt = tid % kCh;
    auto chunk_runs = [&](int n, int &nb, int &bstart) {
        nb = 0;
        bstart = 0;
        if (vt < n && (s_at[vt].y & 1) == ((run & 1) == 0 ? 1 : 0)) {
            const int4 br = s_br[vt];
            nb = run < 2 ? br.y : br.w;
            bstart = run < 2 ? br.x : br.z;
        }
    };
    // C's structure for the row from the staged (column, mask) list, in
    // column order, with value offsets from a scan of the tile popcounts;
    // the row's tiles also go into LDS and the hash table.
    auto insert_tile = [&](int q, int j, uint64_t m, int base) {
        const int64_t t = t0 + q;
        c.col[t] = j;
        c.mask[t] = m;
        c.tile_nnz[t] = zb + base;
        reinterpret_cast<uint64_t *>(c.csr_ptr)[t] = packed_row_offsets(m);
        s_cmask[q] = m;
        s_cbase[q] = base;
        uint32_t h = slot_of(j, Cfg::kSlotBits);
        while (atomicCAS(&s_slot[h].x, kEmpty, j) != kEmpty)
            h = (h + 1) & (Cfg::kSlots - 1);
        s_slot[h] = make_tile_slot(j, base, q, row_runs(m));
    };
    if (tid < min(kCh, a1 - a0))
        stage_tile(am0, br0, bl0);
    __syncthreads();   // chunk 0 staged, table cleared
    // One scan gives the value offsets of the first kT tiles (one per
    // thread) and chunk 0's pair offsets.
    int nb = 0, bstart = 0, pre = 0, total = 0, base = 0, carry = 0;
    chunk_runs(min(kCh, a1 - a0), nb, bstart);
    dual_exclusive_scan<kNW>(tid < nt ? __popcll(m0) : 0, nb, s_dual, &base, &pre, &carry, &total);
    if (tid < nt)
        insert_tile(tid, j0, m0, base);
    for (int q0 = kT; q0 < nt; q0 += kT) {
        const int q = q0 + tid;
        int j = 0;
        uint64_t m = 0;
        if (q < nt) {
            j = st.j[sb + q];
            m = st.m[sb + q];
        }
        int tot = 0;
        const int qb = carry + block_exclusive_scan<kNW>(q < nt ? __popcll(m) : 0, s_tmp, &tot);
        carry += tot;
        if (q < nt)
            insert_tile(q, j, m, qb);
    }
    for (int c0 = a0; c0 < a1; c0 += kCh) {
        const int n = min(kCh, a1 - c0);
        if (c0 != a0) {
            if (tid < n)
                stage_tile(g.a_meta[c0 + tid], g.a_brange[c0 + tid], g.a_blight[c0 + tid]);
            __syncthreads();
            chunk_runs(n, nb, bstart);
            pre = block_exclusive_scan<kNW>(nb, s_tmp, &total);
        }
        s_pw[tid] = bstart - pre;
        {
            for (int r0 = 0; r0 < total; r0 += kSg) {
                const int r1 = min(total, r0 + kSg);
                if (nb > 0)
                    fill_segments(s_seg, pre, nb, r0, r1);
                __syncthreads();
                // Consecutive lanes take consecutive pairs, so a wave's record
                // loads are one contiguous stretch of B's records.
                for (int p = r0 + tid; p < r1; p += kT) {
                    const int v = s_seg[p - r0];
                    const int4 at = s_at[v % kCh];
                    // The pair's B record, part by part from the planes.
                    const int64_t rb_idx = s_pw[v] + p;
                    BRecord rec;
                    rec.meta = recs[rb_idx];
                    // Runs 2 and 3 hold the light B tiles (fewer than
                    // kHeavyMin entries, one row): their values fit in the
                    // first part.
                    static_assert(kHeavyMin - 1 <= 2, "light values fill one record part");
                    const bool light_b = v >= 2 * kCh;
#pragma unroll
                    for (int i = 0; i < kRecVals / 2; ++i) {
                        double2 d = make_double2(0.0, 0.0);
                        if (i < 1 || !light_b)
                            d = reinterpret_cast<const double2 *>(recs + (i + 1) * rec_n)[rb_idx];
                        rec.v[2 * i] = d.x;
                        rec.v[2 * i + 1] = d.y;
                    }
                    const uint32_t ab = static_cast<uint32_t>(at.x);
                    if (!product_count(ab, static_cast<uint32_t>(rec.meta.z)))
                        continue;
                    // C tile of the pair; every tile of the row is in the table and
                    // at least half of its slots are empty, so the probe ends.
                    uint32_t h = slot_of(rec.meta.x, Cfg::kSlotBits);
                    int4 sl;
                    for (;;) {
                        sl = s_slot[h];
                        if (sl.x == rec.meta.x || sl.x == kEmpty)
                            break;
                        h = (h + 1) & (Cfg::kSlots - 1);
                    }
                    if (sl.x != rec.meta.x)
                        continue;
                    const int q = sl.y >> 16;
                    const unsigned long long cruns =
                        (static_cast<unsigned long long>(static_cast<uint32_t>(sl.w)) << 32) |
                        static_cast<uint32_t>(sl.z);
                    double *acc = s_acc + (sl.y & 0xFFFF);
                    const uint32_t apf = static_cast<uint32_t>(at.y) & ~0xFFu;
                    const uint32_t aruns = static_cast<uint32_t>(at.w);
                    const double *acol = s_acol + at.z;
                    const uint32_t boff = static_cast<uint32_t>(rec.meta.w) & ~kDenseBit;
                    // B entries in k-major order: the record's values, then the
                    // payload's.  The first kExtVals payload values are loaded once
                    // kExtVals record values have been used, into the registers
                    // those held, and arrive while the rest of the record is walked.
                    auto walk = [&](auto &&entry) {
                        uint32_t mz = static_cast<uint32_t>(rec.meta.z);
                        const int nbv = __popc(mz);
                        double ext[kExtVals > 0 ? kExtVals : 1];
#pragma unroll
                        for (int j = 0; j < kRecVals; ++j) {
                            if (j == kExtVals) {
#pragma unroll
                                for (int i = 0; i < kExtVals; ++i)
                                    ext[i] = kRecVals + i < nbv ? g.b_vals_km[boff + kRecVals + i]
                                                                : 0.0;
                            }
                            if (mz) {
                                const int pos = __builtin_ctz(mz);
                                mz &= mz - 1;
                                entry(pos, rec.v[j]);
                            }
                        }
#pragma unroll
                        for (int i = 0; i < kExtVals; ++i) {
                            if (mz) {
                                const int pos = __builtin_ctz(mz);
                                mz &= mz - 1;
                                entry(pos, ext[i]);
                            }
                        }
                        int j = kRecVals + kExtVals;
                        while (mz) {
                            const int pos = __builtin_ctz(mz);
                            mz &= mz - 1;
                            entry(pos, g.b_vals_km[boff + j]);
                            ++j;
                        }
                    };
                    if (((cruns & 0x8080808080808080ull) | (aruns & 0x80808080u)) == 0) {
                        // All runs, and no A column longer than kColRegs: at each
                        // new k the lane takes the values and the row bases of the
                        // rows r0, r0 + 1, ... of A column k into registers.
                        static_assert(kColRegs <= 4, "row bases of one dword");
                        int kcur = -1;
                        uint32_t kd = 0;   // a_column_runs byte of column kcur
                        uint32_t rb = 0;   // byte i: row base of row r0 + i
                        double av[kColRegs];
#pragma unroll
                        for (int i = 0; i < kColRegs; ++i)
                            av[i] = 0.0;
                        walk([&](int pos, double vb) {
                            const int k = pos >> 3, cl = pos & 7;
                            if (k != kcur) {
                                kcur = k;
                                kd = byte_of(aruns, k);
                                rb = static_cast<uint32_t>(cruns >> (8 * (kd >> 4)));
                                const double *ap = acol + byte_of(apf, k);
#pragma unroll
                                for (int i = 0; i < kColRegs; ++i)
                                    av[i] = ap[i];
                            }
                            double *pc = acc + cl;
                            const int na = static_cast<int>(kd & 15u);
#pragma unroll
                            for (int i = 0; i < kColRegs; ++i)
                                if (i < na)
                                    atomicAdd(pc + run_base(rb, i), av[i] * vb);
                        });
                    } else {
                        // Slots from popcounts of the tile mask.
                        const uint64_t cm = s_cmask[q];
                        walk([&](int pos, double vb) {
                            const int k = pos >> 3, cl = pos & 7;
                            const double *ap = acol + byte_of(apf, k);
                            uint32_t rr = byte_of(ab, k);
                            int i = 0;
                            while (rr) {
                                const int r = __builtin_ctz(rr);
                                rr &= rr - 1;
                                atomicAdd(&acc[__popcll(cm & ((1ull << (r * 8 + cl)) - 1ull))],
                                          ap[i++] * vb);
                            }
                        });
                    }
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    // Column bytes of the row's entries, staged in the A buffer (free now),
    // then values and column bytes stream out together.
    // They are staged at the output's byte alignment, zb mod 4, so that
    // whole words go out as words.
    const int shift = static_cast<int>(zb & 3);
    unsigned char *s_colb = reinterpret_cast<unsigned char *>(s_acol);
    static_assert(sizeof(s_acol) >= Cfg::kAcc + 3, "column byte staging");
    for (int it = tid; it < nt * 8; it += kT) {
        const int q = it >> 3, r = it & 7;
        const uint64_t m = s_cmask[q];
        uint32_t x = static_cast<uint32_t>(m >> (8 * r)) & 0xFFu;
        int o = shift + s_cbase[q] + __popcll(m & ((1ull << (8 * r)) - 1ull));
        while (x) {
            s_colb[o++] = static_cast<unsigned char>(__builtin_ctz(x));
            x &= x - 1;
        }
    }
    __syncthreads();
    for (int64_t i = tid; i < nnz_row; i += kT)
        c.val[zb + i] = s_acc[i];
    // Words [0, nw) of the row's column bytes from the aligned word at zb -
    // shift; the first and last may hold bytes of the neighbouring rows and
    // are written byte by byte.
    unsigned char *colw = c.csr_col + (zb - shift);
    const int nw = static_cast<int>((shift + nnz_row + 3) >> 2);
    for (int i = tid; i < nw; i += kT) {
        const uint32_t v = reinterpret_cast<const uint32_t *>(s_colb)[i];
        if (i > 0 && i < nw - 1) {
            reinterpret_cast<uint32_t *>(colw)[i] = v;
        } else {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int idx = 4 * i + b;
                if (idx >= shift && idx < shift + nnz_row)
                    colw[idx] = static_cast<unsigned char>(v >> (8 * b));
            }
        }
    }
    if constexpr (TRACE)
        trace_mark(tl, true);
}

// ---------------------------------------------------------------------------
// Numeric for rows beyond the LDS budgets, one workgroup per row at a time.
// A direct-mapped HBM table, one int4 per B tile column {row, first value of
// the C tile within the row, mask low word, mask high word}, gives a pair its
// C tile in one load.  A's values are staged per chunk of A tiles in LDS
// when they fit; B's come from the record planes.
//
// Each chunk of A tiles walks the row in windows of consecutive C tiles
// that start within kSlabWinStep values of each other (at most kSlabWin
// values).  A window's products accumulate in LDS.  Each A tile keeps a
// cursor into every class run of its B row (each run is sorted by column)
// and gallops it to the window's last column, about one load per run and
// window.  A row of one chunk then stores each window to C with plain
// coalesced stores (no clearing pass, no HBM atomics); a row of up to
// kSlabWinTiles A tiles clears its C values first and each chunk adds its
// windows' nonzero sums with FP64 atomics on consecutive addresses.  Longer
// rows (hub rows, whose chunks would each walk every window) clear C and
// add every product with an FP64 atomic in HBM.
// ---------------------------------------------------------------------------
constexpr int kSlabStage = 1024;   // A values a slab chunk stages in LDS
constexpr int kSlabWin = 2048;     // C values a slab window accumulates in LDS
constexpr int kSlabWinStep = kSlabWin - 64;   // tile starts per window; a tile holds <= 64

// First tile in [lo, hi] of a row whose first value is at or after v (hi
// when none); nnz holds the row's tile starts (hi + 1 entries from the row's
// first tile, the last one the row's end).
__device__ __forceinline__ int first_tile_from(const int64_t *__restrict__ nnz, int lo, int hi,
                                               int64_t v)
{
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (nnz[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// The products of one pair (A tile at, B record b with meta) into the C tile
// whose values start at acc and whose mask is cm: into an LDS window (WIN)
// or with FP64 atomics on C's values in HBM.
template <bool WIN>
__device__ __forceinline__ void slab_products(const int4 *__restrict__ recs, int64_t rec_n,
                                              const double *__restrict__ b_vals_km, int64_t b,
                                              int4 meta, int4 at, const double *abase,
                                              uint64_t cm, double *acc)
{
    // Only the parts that hold the tile's values: most tiles of the rows that
    // come here have one or two entries.
    const int nbv = __popc(static_cast<uint32_t>(meta.z));
    double rv[kRecVals];
#pragma unroll
    for (int i = 0; i < kRecVals / 2; ++i) {
        double2 d = make_double2(0.0, 0.0);
        if (2 * i < nbv)
            d = reinterpret_cast<const double2 *>(recs + (i + 1) * rec_n)[b];
        rv[2 * i] = d.x;
        rv[2 * i + 1] = d.y;
    }
    const uint32_t ab = static_cast<uint32_t>(at.x);
    const double *acol = abase + at.z;
    const uint32_t apf = static_cast<uint32_t>(at.y);
    const uint32_t boff = static_cast<uint32_t>(meta.w) & ~kDenseBit;
    auto entry = [&](int pos, double vb) {
        const int k = pos >> 3, cl = pos & 7;
        const double *ap = acol + byte_of(apf, k);
        uint32_t rr = byte_of(ab, k);
        int i = 0;
        while (rr) {
            const int r = __builtin_ctz(rr);
            rr &= rr - 1;
            double *dst = &acc[__popcll(cm & ((1ull << (r * 8 + cl)) - 1ull))];
            if constexpr (WIN)
                atomicAdd(dst, ap[i++] * vb);
            else
                unsafeAtomicAdd(dst, ap[i++] * vb);
        }
    };
    uint32_t mz = static_cast<uint32_t>(meta.z);
#pragma unroll
    for (int j = 0; j < kRecVals; ++j) {
        if (mz) {
            const int pos = __builtin_ctz(mz);
            mz &= mz - 1;
            entry(pos, rv[j]);
        }
    }
    int j = kRecVals;
    while (mz) {
        const int pos = __builtin_ctz(mz);
        mz &= mz - 1;
        entry(pos, b_vals_km[boff + j]);
        ++j;
    }
}

// Per-chunk A tile info: {rows of column k (byte k), first value of column k
// (byte k), first value relative to the chunk's, unused}.
__device__ __forceinline__ int4 slab_a_tile(int4 am, int v0)
{
    const uint32_t ab = static_cast<uint32_t>(am.z);
    const uint32_t p1 = __popc(byte_of(ab, 0));
    const uint32_t p2 = p1 + __popc(byte_of(ab, 1));
    const uint32_t p3 = p2 + __popc(byte_of(ab, 2));
    return make_int4(static_cast<int>(ab), static_cast<int>((p1 << 8) | (p2 << 16) | (p3 << 24)),
                     static_cast<int>(static_cast<uint32_t>(am.w) & ~kDenseBit) - v0, 0);
}

__global__ __launch_bounds__(kThreads) void numeric_slab_kernel(
    Operands g, const int4 *__restrict__ recs, int64_t rec_n, const int *__restrict__ rows,
    int nrows, int *__restrict__ queue, int static_rows, int4 *__restrict__ slabs, COut c)
{
    __shared__ double s_win[kSlabWin];
    __shared__ double s_acol[kSlabStage + 8];
    __shared__ int4 s_at[kThreads];
    __shared__ int s_pw[kSlabRuns * kThreads];    // per (tile, run): first B tile minus first pair
    __shared__ int s_pre[kSlabRuns * kThreads];   // per (tile, run): first pair
    __shared__ int s_tmp[kWavesPerBlock];
    int4 *slab = slabs + static_cast<int64_t>(blockIdx.x) * g.b_tile_cols;
    const int tid = threadIdx.x;
    for (int i = tid; i < kSlabWin; i += kThreads)
        s_win[i] = 0.0;
    for (int step = 0;; ++step) {
        const int ri = static_rows ? static_cast<int>(blockIdx.x) + step * static_cast<int>(gridDim.x)
                                   : next_row(queue);
        if (ri >= nrows)
            break;
        const int row = rows[ri];
        const int64_t t0 = c.tile_ptr[row];
        const int nt = static_cast<int>(c.tile_ptr[row + 1] - t0);
        const int64_t zb = c.row_nnz_off[row];
        const int64_t nnz_row = c.row_nnz_off[row + 1] - zb;
        const int a0 = g.a_trp[row], a1 = g.a_trp[row + 1];
        // A row of one chunk stores its windows; a row of a few chunks clears
        // C first and each chunk adds its windows' sums; a longer row, where
        // chunks times windows would outweigh its pairs, adds every product.
        const bool single = a1 - a0 <= kThreads;
        const bool windowed = a1 - a0 <= kSlabWinTiles;
        // Columns are checked against the table's extent, so that a damaged
        // structure cannot send a store outside it.
        for (int q = tid; q < nt; q += kThreads) {
            const int j = c.col[t0 + q];
            const uint64_t m = c.mask[t0 + q];
            if (j >= 0 && j < g.b_tile_cols)
                slab[j] = make_int4(row, static_cast<int>(c.tile_nnz[t0 + q] - zb),
                                    static_cast<int>(static_cast<uint32_t>(m)),
                                    static_cast<int>(static_cast<uint32_t>(m >> 32)));
        }
        if (!single)
            for (int64_t i = tid; i < nnz_row; i += kThreads)
                c.val[zb + i] = 0.0;
        __threadfence();
        __syncthreads();
        double *cval = c.val + zb;
        const int64_t *tnz = c.tile_nnz + t0;
        for (int c0 = a0; windowed && c0 < a1; c0 += kThreads) {
            // The chunk's A tiles, their staged values and their run cursors
            // stay put for every window of the row.
            const int n = min(kThreads, a1 - c0);
            const int v0 = g.a_off[c0];
            const int span = g.a_off[c0 + n] - v0;
            const bool staged = span <= kSlabStage;
            int cur[kSlabRuns], end[kSlabRuns];
#pragma unroll
            for (int r = 0; r < kSlabRuns; ++r)
                cur[r] = end[r] = 0;
            if (tid < n) {
                const int4 am = g.a_meta[c0 + tid];
                s_at[tid] = slab_a_tile(am, v0);
                const uint32_t ab = static_cast<uint32_t>(am.z);
                const int *gp = g.b_groups + static_cast<int64_t>(kClassSlots) * am.x;
#pragma unroll
                for (int r = 0; r < kSlabRuns; ++r) {
                    // Light run k meets only A tiles with entries in column k.
                    if (r == 0 || byte_of(ab, r - 1) != 0) {
                        cur[r] = gp[r];
                        end[r] = gp[r + 1];
                    }
                }
            }
            if (staged)
                for (int i = tid; i < span; i += kThreads)
                    s_acol[i] = g.a_vals_cm[v0 + i];
            const double *abase = staged ? s_acol : g.a_vals_cm + v0;
            // Every thread finds the same window bounds, so they need no exchange.
            int t_lo = 0;
            for (int64_t wend = kSlabWinStep; t_lo < nt; wend += kSlabWinStep) {
                const int t_hi = first_tile_from(tnz, t_lo, nt, zb + wend);
                if (t_hi == t_lo)
                    continue;
                const int vb = static_cast<int>(tnz[t_lo] - zb);
                const int vs = static_cast<int>(tnz[t_hi] - zb) - vb;
                const int jhi = c.col[t0 + t_hi - 1] + 1;
                // B tiles before the window's first column that a cursor still
                // covers are not C tiles of the row: they meet nothing.
                int cnt[kSlabRuns], lo[kSlabRuns];
                int mine = 0;
#pragma unroll
                for (int r = 0; r < kSlabRuns; ++r) {
                    lo[r] = cur[r];
                    cnt[r] = 0;
                    if (cur[r] < end[r] && g.b_meta[cur[r]].x < jhi) {
                        const int h = gallop_col(g.b_meta, cur[r], end[r], jhi);
                        cnt[r] = h - cur[r];
                        cur[r] = h;
                    }
                    mine += cnt[r];
                }
                int total = 0;
                int pre = block_exclusive_scan(mine, s_tmp, &total);
#pragma unroll
                for (int r = 0; r < kSlabRuns; ++r) {
                    s_pre[kSlabRuns * tid + r] = pre;
                    s_pw[kSlabRuns * tid + r] = lo[r] - pre;
                    pre += cnt[r];
                }
                __syncthreads();
                // A window with no pairs in a longer row's chunk adds nothing.
                if (total == 0 && !single) {
                    t_lo = t_hi;
                    continue;
                }
                for (int p = tid; p < total; p += kThreads) {
                    const int e = upper_index(s_pre, kSlabRuns * n, p);
                    const int4 at = s_at[e / kSlabRuns];
                    const int64_t b = s_pw[e] + p;
                    const int4 meta = recs[b];
                    if (!product_count(static_cast<uint32_t>(at.x), static_cast<uint32_t>(meta.z)))
                        continue;
                    if (meta.x < 0 || meta.x >= g.b_tile_cols)
                        continue;
                    const int4 sl = slab[meta.x];
                    if (sl.x != row || static_cast<unsigned>(sl.y - vb) >= static_cast<unsigned>(vs))
                        continue;
                    const uint64_t cm = (static_cast<uint64_t>(static_cast<uint32_t>(sl.w)) << 32) |
                                        static_cast<uint32_t>(sl.z);
                    slab_products<true>(recs, rec_n, g.b_vals_km, b, meta, at, abase, cm,
                                        s_win + (sl.y - vb));
                }
                __syncthreads();
                // The window's values, and the accumulator cleared for the next
                // window by the thread that read it.  A single-chunk row
                // stores every value (zeros included); a longer row adds the
                // nonzero sums, consecutive addresses per wave.
                for (int i = tid; i < vs; i += kThreads) {
                    const double v = s_win[i];
                    s_win[i] = 0.0;
                    if (single)
                        cval[vb + i] = v;
                    else if (v != 0.0)
                        unsafeAtomicAdd(&cval[vb + i], v);
                }
                __syncthreads();
                t_lo = t_hi;
            }
        }
        for (int c0 = a0; !windowed && c0 < a1; c0 += kThreads) {
            const int n = min(kThreads, a1 - c0);
            const int v0 = g.a_off[c0];
            const int span = g.a_off[c0 + n] - v0;
            const bool staged = span <= kSlabStage;
            int nb = 0, bstart = 0;
            if (tid < n) {
                s_at[tid] = slab_a_tile(g.a_meta[c0 + tid], v0);
                const int4 br = g.a_brange[c0 + tid];
                nb = br.z;
                bstart = br.x;
            }
            if (staged)
                for (int i = tid; i < span; i += kThreads)
                    s_acol[i] = g.a_vals_cm[v0 + i];
            const double *abase = staged ? s_acol : g.a_vals_cm + v0;
            int total = 0;
            const int pre = block_exclusive_scan(nb, s_tmp, &total);
            s_pw[tid] = bstart - pre;
            s_pre[tid] = pre;
            __syncthreads();
            // A pair finds its A tile by a binary search over the chunk's
            // first pairs: a tile whose B row is a hub column would make one
            // thread fill most of a segment table.
            for (int p = tid; p < total; p += kThreads) {
                const int v = upper_index(s_pre, n, p);
                const int4 at = s_at[v];
                const int64_t b = s_pw[v] + p;
                const int4 meta = recs[b];
                if (!product_count(static_cast<uint32_t>(at.x), static_cast<uint32_t>(meta.z)))
                    continue;
                if (meta.x < 0 || meta.x >= g.b_tile_cols)
                    continue;
                const int4 sl = slab[meta.x];
                if (sl.x != row)
                    continue;
                const uint64_t cm = (static_cast<uint64_t>(static_cast<uint32_t>(sl.w)) << 32) |
                                    static_cast<uint32_t>(sl.z);
                slab_products<false>(recs, rec_n, g.b_vals_km, b, meta, at, abase, cm,
                                     cval + sl.y);
            }
            __syncthreads();
        }
        __threadfence();
        __syncthreads();
    }
}

// Appends v to out (count holds the length) for the threads where pred
// holds: one atomic per workgroup of NW waves (an atomic per wave on one
// counter serialised at the L2 on millions of rows), positions in thread
// order within the workgroup.  s holds NW + 1 ints; every thread calls it.
template <int NW>
__device__ __forceinline__ void block_append(bool pred, int v, int *__restrict__ out,
                                             int *__restrict__ count, int *s)
{
    const unsigned long long m = __ballot(pred);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0)
        s[w] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i)
            total += s[i];
        s[NW] = total ? atomicAdd(count, total) : 0;
    }
    __syncthreads();
    int base = s[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i)
        base += i < w ? s[i] : 0;
    if (pred)
        out[base + __popcll(m & ((1ull << lane_id()) - 1ull))] = v;
    __syncthreads();   // s is reused by the next call
}

// Lists rows for the other passes.  mode 0: symbolic slab rows; mode 2:
// rows for the workgroup symbolic table; mode 1: numeric rows off
// SmallRow, into the LargeRow list (out_lds), the slab list (out_slab),
// the TinyRow list (out_tiny) and the MediumRow list (out_mid).
__global__ __launch_bounds__(kThreads) void collect_rows_kernel(int n, int mode, const int *__restrict__ level,
                                    const int64_t *__restrict__ tile_ptr,
                                    const int64_t *__restrict__ nnz_off,
                                    const int *__restrict__ a_trp, const int *__restrict__ a_off,
                                    int *__restrict__ out_slab, int *__restrict__ n_slab,
                                    int *__restrict__ out_lds, int *__restrict__ n_lds,
                                    float tile_fill = INFINITY, int *__restrict__ out_tiny = nullptr,
                                    int *__restrict__ n_tiny = nullptr,
                                    int *__restrict__ out_mid = nullptr,
                                    int *__restrict__ n_mid = nullptr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool to_slab = false, to_lds = false, to_tiny = false, to_mid = false;
    if (i < n) {
        const int lv = level[i];
        if (mode == 0 || mode == 2) {
            to_slab = lv == (mode == 0 ? kLevelSlab : kLevelLarge);
        } else if (tile_routed(tile_fill, a_trp[i + 1] - a_trp[i],
                               a_off[a_trp[i + 1]] - a_off[a_trp[i]])) {
            // numeric on the tile-pair kernel
        } else if (lv == kLevelSlab) {
            to_slab = true;
        } else if (lv == kLevelLds) {
            const int64_t nt = tile_ptr[i + 1] - tile_ptr[i];
            const int64_t nz = nnz_off[i + 1] - nnz_off[i];
            const int64_t alen = a_off[a_trp[i + 1]] - a_off[a_trp[i]];
            if (nt > LargeRow::kTiles || nz > LargeRow::kAcc || alen > LargeRow::kAStage)
                to_slab = true;
            else if (fits_row<SmallRow>(nt, nz, alen))
                to_tiny = fits_row<TinyRow>(nt, nz, alen);
            else if (fits_row<MediumRow>(nt, nz, alen))
                to_mid = true;
            else
                to_lds = true;
        }
    }
    __shared__ int s_app[kWavesPerBlock + 1];
    block_append<kWavesPerBlock>(to_slab, i, out_slab, n_slab, s_app);
    if (mode == 1) {
        block_append<kWavesPerBlock>(to_lds, i, out_lds, n_lds, s_app);
        block_append<kWavesPerBlock>(to_tiny, i, out_tiny, n_tiny, s_app);
        block_append<kWavesPerBlock>(to_mid, i, out_mid, n_mid, s_app);
    }
}

// Number of (A tile, B tile) pairs: the sum of the per-A-tile extents, one
// atomic per workgroup.
__global__ void fill_i32_kernel(int *p, int64_t n, int v)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long *p, int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0ull;
}

__global__ void set_last_kernel(int64_t *p, const int64_t *v) { *p = *v; }

__global__ void set_last2_kernel(int64_t *p, const int64_t *v, int64_t *q, const int64_t *u)
{
    *p = *v;
    *q = *u;
}

// Zeroes the host-read counters except [0], which takes the pair total, the
// trailing entries of the per-row size arrays and the queue heads.
__global__ void symbolic_init_kernel(unsigned long long *ctr, int n_ctr,
                                     const unsigned long long *pair_total,
                                     unsigned long long *row_tiles_end,
                                     unsigned long long *row_nnz_end, int *queue, int n_queue)
{
    const int i = threadIdx.x;
    if (i < n_ctr)
        ctr[i] = i == 0 ? *pair_total : 0ull;
    if (i < n_queue)
        queue[i] = 0;
    if (i == 0) {
        *row_tiles_end = 0ull;
        *row_nnz_end = 0ull;
    }
}

double event_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0;
    }
    return ms;
}

} // namespace

// Timeline trace of the main numeric launch, switched on by the reference's
// environment variables: DMMA_TRACE_FILE (output path), DMMA_TRACE_SHIFT
// (sample every 2^shift-th workgroup, default 4, at most 10) and
// DMMA_TRACE_PHASE (which one, default min(7, 2^shift - 1)).
static unsigned int trace_env(const char *name, unsigned int fallback, unsigned int maximum)
{
    const char *text = std::getenv(name);
    if (!text || !*text)
        return fallback;
    char *end = nullptr;
    const unsigned long v = std::strtoul(text, &end, 10);
    return (end == text || *end || v > maximum) ? fallback : static_cast<unsigned int>(v);
}

// The reference's DMTL0001 layout (dmma_write_timeline_trace,
// src/dmma_spgemm.h:95-164): magic, version, compute units, waves per
// workgroup, shift, phase, output tiles, workgroups, slots, then the slot
// start times, end times (ns) and compute unit ids.
static int write_timeline(const char *path, const Timeline &tl, size_t slots, uint64_t blocks,
                          uint64_t output_tiles, int num_cus)
{
    std::vector<unsigned long long> start(slots), end(slots);
    std::vector<uint32_t> cu(slots);
    RTT_HIP(hipMemcpy(start.data(), tl.start, slots * sizeof(start[0]), hipMemcpyDeviceToHost));
    RTT_HIP(hipMemcpy(end.data(), tl.end, slots * sizeof(end[0]), hipMemcpyDeviceToHost));
    RTT_HIP(hipMemcpy(cu.data(), tl.cu, slots * sizeof(cu[0]), hipMemcpyDeviceToHost));
    int rate_khz = 0;
    RTT_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx().device));
    const double ns_per_tick = rate_khz > 0 ? 1e6 / rate_khz : 1.0;
    for (size_t i = 0; i < slots; ++i) {
        start[i] = static_cast<unsigned long long>(static_cast<double>(start[i]) * ns_per_tick);
        end[i] = static_cast<unsigned long long>(static_cast<double>(end[i]) * ns_per_tick);
    }
    FILE *f = std::fopen(path, "wb");
    if (!f)
        return fail(RTT_ERR_INVALID, "cannot open timeline output %s", path);
    const char magic[8] = {'D', 'M', 'T', 'L', '0', '0', '0', '1'};
    const uint32_t head[5] = {1u, static_cast<uint32_t>(num_cus),
                              static_cast<uint32_t>(kWavesPerBlock), tl.shift, tl.phase};
    const uint64_t tail[3] = {output_tiles, blocks, static_cast<uint64_t>(slots)};
    bool ok = std::fwrite(magic, sizeof(magic), 1, f) == 1 &&
              std::fwrite(head, sizeof(head), 1, f) == 1 &&
              std::fwrite(tail, sizeof(tail), 1, f) == 1;
    if (slots)
        ok = ok && std::fwrite(start.data(), sizeof(start[0]), slots, f) == slots &&
             std::fwrite(end.data(), sizeof(end[0]), slots, f) == slots &&
             std::fwrite(cu.data(), sizeof(cu[0]), slots, f) == slots;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok)
        return fail(RTT_ERR_INVALID, "write of timeline %s failed", path);
    return RTT_OK;
}

// Copies a slab row list into w->rows_sorted, heaviest row (most pairs)
// first, so that the work queue hands out the long rows before the short
// ones.  Stream-ordered: the previous user of rows_sorted has been queued.
static int order_rows_by_cost(const int *rows, int n, const Operands &g, SpgemmWork *w, hipStream_t s,
                       const int **sorted)
{
    RTT_TRY(w->row_cost.ensure(static_cast<size_t>(n) * sizeof(unsigned int)));
    RTT_TRY(w->row_cost2.ensure(static_cast<size_t>(n) * sizeof(unsigned int)));
    RTT_TRY(w->rows_sorted.ensure(static_cast<size_t>(n) * sizeof(int)));
    unsigned int *cost = w->row_cost.as<unsigned int>();
    unsigned int *cost2 = w->row_cost2.as<unsigned int>();
    int *out = w->rows_sorted.as<int>();
    hipLaunchKernelGGL(row_pairs_kernel, dim3(blocks_for(n, kWavesPerBlock)), dim3(kThreads), 0,
                       s, rows, n, g.a_trp, g.a_brange, cost);
    RTT_HIP(hipGetLastError());
    size_t bytes = 0;
    RTT_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, cost, cost2, rows, out,
                                                         n, 0, 32, s));
    RTT_TRY(w->scan_tmp.ensure(bytes));
    RTT_HIP(hipcub::DeviceRadixSort::SortPairsDescending(w->scan_tmp.ptr, bytes, cost, cost2,
                                                         rows, out, n, 0, 32, s));
    *sorted = out;
    return RTT_OK;
}

// Workgroups of a slab pass whose per-workgroup HBM table holds `cols`
// entries of `entry` bytes: up to four per CU, fewer when the tables would
// pass a quarter of the free device memory (plus what `table` already holds).
// The passes take rows from a work queue (or round robin over the grid), so
// any number of workgroups covers every row; one table that does not fit
// fails the allocation with RTT_ERR_OOM.
static int slab_grid(int rows, int num_cus, int64_t cols, size_t entry, const DevBuf &table)
{
    int blocks = std::max(1, std::min(rows, 4 * num_cus));
    size_t free_b = 0, total_b = 0;
    if (cols > 0 && hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        const uint64_t per = static_cast<uint64_t>(cols) * entry;
        const uint64_t budget = free_b / 4 + table.bytes;
        blocks = static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(blocks, budget / per)));
    }
    return blocks;
}

int spgemm_core(const DevTiles &a, const DevTiles &b, DevC *c, SpgemmWork *w, Events *ev,
                CoreTimes *times, int lb_mode, const TileOpts *tile)
{
    // tile: rows routed to the tile-pair kernel (all when fill is 0);
    // vector: rows left to the LDS-accumulator kernels.
    const bool any_tile = tile != nullptr;
    const bool any_vector = !tile || tile->fill > 0.f;
    RTT_TRY(require_ctx());
    RTT_TRY(ev->create());
    hipStream_t s = ctx().stream;
    const int R = a.tile_row_count;
    const int BC = b.tile_col_count;
    c->tilem = R;
    c->tilen = BC;
    c->rows = a.rows;
    c->cols = b.cols;

    Operands g;
    g.a_tile_rows = R;
    g.a_trp = a.tile_row_ptr.as<int>();
    g.a_meta = a.meta.as<int4>();
    g.a_brange = b.a_brange.as<int4>();
    g.a_blight = b.a_blight.as<int2>();
    g.a_off = a.value_offsets.as<int>();
    g.a_vals = a.values.as<double>();
    g.a_vals_cm = a.values_cm.as<double>();
    g.a_row_off = a.row_payload_off.as<int>();
    g.b_vals_km = b.payload_kmajor();
    g.b_trp = b.tile_row_ptr.as<int>();
    g.b_meta = b.records.as<int4>();   // part 0 of every record is the tile's meta
    g.b_groups = b.row_groups.as<int>();
    g.b_vals = b.payload_values();
    g.b_tile_cols = BC;
    g.tile_fill = tile ? tile->fill : INFINITY;

    const int num_cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
    // Counters read back by the host: [0] pairs, [1] symbolic slab rows,
    // [2] numeric slab rows, [3] numeric large-LDS rows, [4] C tiles,
    // [5] nnz(C), [6] rows for the workgroup symbolic table, [7] slab tiles
    // the staging cursor handed out, [8] numeric one-wave (TinyRow) rows,
    // [9] numeric MediumRow rows.
    constexpr int kCounters = 10;
    RTT_TRY(w->counters.ensure(kCounters * sizeof(unsigned long long)));
    RTT_TRY(w->host_counters.ensure(kCounters * sizeof(unsigned long long)));
    unsigned long long *ctr = w->counters.as<unsigned long long>();
    unsigned long long *hctr = w->host_counters.as<unsigned long long>();
    int *slab_count = reinterpret_cast<int *>(ctr + 1);
    int *num_slab_count = reinterpret_cast<int *>(ctr + 2);
    int *num_lds_count = reinterpret_cast<int *>(ctr + 3);
    int *large_count = reinterpret_cast<int *>(ctr + 6);
    int *num_tiny_count = reinterpret_cast<int *>(ctr + 8);
    int *num_mid_count = reinterpret_cast<int *>(ctr + 9);

    auto t_sym = Clock::now();
    // ---- symbolic: per-row tile sets, staged ------------------------------
    const int stage_limit = 3 * kSymSlots / 4;
    RTT_TRY(w->row_tiles.ensure((R + 1) * sizeof(int64_t)));
    RTT_TRY(w->row_nnz.ensure((R + 1) * sizeof(int64_t)));
    RTT_TRY(w->row_level.ensure((R + 1) * sizeof(int)));
    RTT_TRY(w->l3_rows.ensure((R + 1) * sizeof(int)));
    RTT_TRY(w->num_rows.ensure((R + 1) * sizeof(int)));
    RTT_TRY(w->lds_rows.ensure((R + 1) * sizeof(int)));
    RTT_TRY(w->tiny_rows.ensure((R + 1) * sizeof(int)));
    RTT_TRY(w->mid_rows.ensure((R + 1) * sizeof(int)));
    RTT_TRY(w->stage_j.ensure(static_cast<size_t>(R) * stage_limit * sizeof(int)));
    RTT_TRY(w->stage_m.ensure(static_cast<size_t>(R) * stage_limit * sizeof(uint64_t)));
    RTT_TRY(c->tile_ptr.ensure((R + 1) * sizeof(int64_t)));
    RTT_TRY(w->c_row_nnz_off.ensure((R + 1) * sizeof(int64_t)));
    Stage st{w->stage_j.as<int>(), w->stage_m.as<unsigned long long>(), stage_limit};
    // Queue heads: [0] symbolic slab count, [1] symbolic slab fill, [2] numeric slab.
    RTT_TRY(w->queue.ensure(4 * sizeof(int)));
    int *queue = w->queue.as<int>();
    // Counter [0] is the pair total, summed by the B update's pair-extent pass.
    hipLaunchKernelGGL(symbolic_init_kernel, dim3(1), dim3(kCounters), 0, s, ctr, kCounters,
                       b.pair_total.as<unsigned long long>(),
                       reinterpret_cast<unsigned long long *>(w->row_tiles.as<int64_t>() + R),
                       reinterpret_cast<unsigned long long *>(w->row_nnz.as<int64_t>() + R),
                       queue, 4);
    // One wave per row first; rows that outgrow its table go to the
    // workgroup table, and rows that outgrow that to the slab.
    RTT_HIP(hipEventRecord(ev->ev[0], s));
    if (R > 0)
        hipLaunchKernelGGL(symbolic_wave_kernel, dim3(R), dim3(kWave), 0, s, g, st,
                           w->row_tiles.as<int64_t>(), w->row_nnz.as<int64_t>(),
                           w->row_level.as<int>());
    RTT_HIP(hipGetLastError());
    int n_slab = 0;
    COut none{};
    // Sizes of C from the per-row counts: row offsets, the numeric routing
    // lists, and every counter read back with one copy.  Rows of the
    // tile-pair kernel are not listed.
    auto finish_sizes = [&]() -> int {
        RTT_TRY(exclusive_scan_i64(w->row_tiles.as<int64_t>(), c->tile_ptr.as<int64_t>(), R,
                                   &w->scan_tmp, s));
        RTT_TRY(exclusive_scan_i64(w->row_nnz.as<int64_t>(), w->c_row_nnz_off.as<int64_t>(), R,
                                   &w->scan_tmp, s));
        if (R > 0 && any_vector)
            hipLaunchKernelGGL(collect_rows_kernel, dim3(blocks_for(R, kThreads)),
                               dim3(kThreads), 0, s, R, 1, w->row_level.as<int>(),
                               c->tile_ptr.as<int64_t>(), w->c_row_nnz_off.as<int64_t>(),
                               g.a_trp, g.a_off, w->num_rows.as<int>(), num_slab_count,
                               w->lds_rows.as<int>(), num_lds_count, g.tile_fill,
                               w->tiny_rows.as<int>(), num_tiny_count, w->mid_rows.as<int>(),
                               num_mid_count);
        hipLaunchKernelGGL(set_last2_kernel, dim3(1), dim3(1), 0, s,
                           reinterpret_cast<int64_t *>(ctr + 4), c->tile_ptr.as<int64_t>() + R,
                           reinterpret_cast<int64_t *>(ctr + 5),
                           w->c_row_nnz_off.as<int64_t>() + R);
        RTT_HIP(hipMemcpyAsync(hctr, ctr, kCounters * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, s));
        RTT_HIP(hipStreamSynchronize(s));
        return RTT_OK;
    };
    bool sized = false;
    if (R > 0) {
        hipLaunchKernelGGL(collect_rows_kernel, dim3(blocks_for(R, kThreads)), dim3(kThreads), 0,
                           s, R, 2, w->row_level.as<int>(), nullptr, nullptr, nullptr, nullptr,
                           w->l3_rows.as<int>(), large_count, nullptr, nullptr);
        RTT_HIP(hipEventRecord(ev->ev[1], s));
        // Sized at once, as if every row fit the wave table: one readback
        // when none outgrew it (the usual case), redone below when some did.
        RTT_TRY(finish_sizes());
        sized = true;
        const int n_large = static_cast<int>(hctr[6] & 0xffffffffu);
        if (n_large > 0) {
            sized = false;
            hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(2), 0, s, ctr + 2, int64_t(2));
            hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(2), 0, s, ctr + 8, int64_t(2));
            hipLaunchKernelGGL(symbolic_count_kernel, dim3(n_large), dim3(kThreads), 0, s, g,
                               stage_limit, st, w->l3_rows.as<int>(),
                               w->row_tiles.as<int64_t>(), w->row_nnz.as<int64_t>(),
                               w->row_level.as<int>());
            RTT_HIP(hipGetLastError());
            hipLaunchKernelGGL(collect_rows_kernel, dim3(blocks_for(R, kThreads)), dim3(kThreads),
                               0, s, R, 0, w->row_level.as<int>(), nullptr, nullptr, nullptr,
                               nullptr, w->l3_rows.as<int>(), slab_count, nullptr, nullptr);
            RTT_HIP(hipMemcpyAsync(hctr, ctr, 2 * sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, s));
            RTT_HIP(hipStreamSynchronize(s));
            n_slab = static_cast<int>(hctr[1] & 0xffffffffu);
            RTT_HIP(hipEventRecord(ev->ev[1], s));
        }
    } else {
        RTT_HIP(hipEventRecord(ev->ev[1], s));
    }
    int slab_blocks = 0;
    const int *slab_rows = nullptr;
    const int static_rows = lb_mode == RTT_LB_STATIC ? 1 : 0;
    SlabStage ss;
    if (n_slab > 0) {
        // Staging of the slab rows' tile lists: grown to what the previous
        // calls needed, within a quarter of the free device memory; without
        // it every slab row is recomputed by the fill pass.
        if (w->slab_stage_need > w->slab_stage_cap) {
            size_t free_b = 0, total_b = 0;
            const uint64_t per = sizeof(int) + sizeof(unsigned long long);
            uint64_t want = w->slab_stage_need + w->slab_stage_need / 16;
            if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
                want = std::min<uint64_t>(want, free_b / 4 / per);
            w->slab_stage_cap = 0;
            if (want > 0 && w->slab_stage_j.ensure(want * sizeof(int)) == RTT_OK &&
                w->slab_stage_m.ensure(want * sizeof(unsigned long long)) == RTT_OK)
                w->slab_stage_cap = want;
            clear_error();
        }
        RTT_TRY(w->slab_row_off.ensure(static_cast<size_t>(R) * sizeof(int64_t)));
        ss.row_off = w->slab_row_off.as<int64_t>();
        ss.cursor = ctr + 7;
        if (w->slab_stage_cap > 0) {
            ss.j = w->slab_stage_j.as<int>();
            ss.m = w->slab_stage_m.as<unsigned long long>();
            ss.capacity = w->slab_stage_cap;
        }
        slab_blocks = slab_grid(n_slab, num_cus, BC, sizeof(uint64_t), w->slab);
        slab_rows = w->l3_rows.as<int>();
        if (lb_mode == RTT_LB_ADAPTIVE_TAIL)
            RTT_TRY(order_rows_by_cost(w->l3_rows.as<int>(), n_slab, g, w, s, &slab_rows));
        RTT_TRY(w->slab.ensure(static_cast<size_t>(slab_blocks) * BC * sizeof(uint64_t)));
        hipLaunchKernelGGL(fill_u64_kernel,
                           dim3(blocks_for(static_cast<int64_t>(slab_blocks) * BC, kThreads)),
                           dim3(kThreads), 0, s, w->slab.as<unsigned long long>(),
                           static_cast<int64_t>(slab_blocks) * BC);
        hipLaunchKernelGGL((symbolic_slab_kernel<false>), dim3(slab_blocks), dim3(kThreads), 0, s,
                           g, slab_rows, n_slab, queue, static_rows,
                           w->slab.as<unsigned long long>(), w->row_tiles.as<int64_t>(),
                           w->row_nnz.as<int64_t>(), ss, none);
        RTT_HIP(hipGetLastError());
    }
    if (!sized)
        RTT_TRY(finish_sizes());
    w->slab_stage_need = std::max<uint64_t>(w->slab_stage_need, hctr[7]);
    const unsigned long long pairs = hctr[0];
    const int n_num_slab = static_cast<int>(hctr[2] & 0xffffffffu);
    const int n_num_lds = static_cast<int>(hctr[3] & 0xffffffffu);
    const int n_num_tiny = static_cast<int>(hctr[8] & 0xffffffffu);
    const int n_num_mid = static_cast<int>(hctr[9] & 0xffffffffu);
    const int64_t total_tiles = static_cast<int64_t>(hctr[4]);
    const int64_t total_nnz = static_cast<int64_t>(hctr[5]);
    const double sym_count_ms = ms_since(t_sym);

    // ---- allocate C -----------------------------------------------------
    auto t_alloc = Clock::now();
    c->num_tiles = total_tiles;
    c->nnz = total_nnz;
    RTT_TRY(c->tile_col.ensure(c->num_tiles * sizeof(int)));
    RTT_TRY(c->tile_mask.ensure(c->num_tiles * sizeof(uint64_t)));
    RTT_TRY(c->tile_nnz.ensure((c->num_tiles + 1) * sizeof(int64_t)));
    RTT_TRY(c->tile_csr_ptr.ensure(c->num_tiles * 8));
    RTT_TRY(c->tile_csr_col.ensure(c->nnz));
    RTT_TRY(c->tile_csr_val.ensure(c->nnz * sizeof(double)));
    times->allocation_ms = ms_since(t_alloc);

    // ---- symbolic: final structure -------------------------------------------
    auto t_fill = Clock::now();
    COut co;
    co.tile_ptr = c->tile_ptr.as<int64_t>();
    co.row_nnz_off = w->c_row_nnz_off.as<int64_t>();
    co.col = c->tile_col.as<int>();
    co.mask = c->tile_mask.as<uint64_t>();
    co.tile_nnz = c->tile_nnz.as<int64_t>();
    co.csr_ptr = c->tile_csr_ptr.as<uint8_t>();
    co.csr_col = c->tile_csr_col.as<uint8_t>();
    co.val = c->tile_csr_val.as<double>();
    RTT_HIP(hipEventRecord(ev->ev[2], s));
    // The numeric LDS kernels write the structure of their own rows from the
    // staged lists; this pass writes it for the staged rows they cannot hold
    // (the numeric fallback list, whose slab rows it skips).  The tile-pair
    // path needs the whole structure first: every staged row.
    if (any_tile && R > 0 && c->num_tiles > 0)
        hipLaunchKernelGGL(symbolic_write_kernel, dim3(blocks_for(R, kWavesPerBlock)),
                           dim3(kThreads), 0, s, nullptr, R, st, w->row_level.as<int>(), co,
                           g.tile_fill, g.a_trp, g.a_row_off);
    if (n_num_slab > 0 && c->num_tiles > 0)
        hipLaunchKernelGGL(symbolic_write_kernel, dim3(blocks_for(n_num_slab, kWavesPerBlock)),
                           dim3(kThreads), 0, s, w->num_rows.as<int>(), n_num_slab, st,
                           w->row_level.as<int>(), co, 0.f, nullptr, nullptr);
    RTT_HIP(hipGetLastError());
    RTT_HIP(hipEventRecord(ev->ev[3], s));
    if (n_slab > 0)
        hipLaunchKernelGGL((symbolic_slab_kernel<true>), dim3(slab_blocks), dim3(kThreads), 0, s,
                           g, slab_rows, n_slab, queue + 1, static_rows,
                           w->slab.as<unsigned long long>(), nullptr, nullptr, ss, co);
    hipLaunchKernelGGL(set_last_kernel, dim3(1), dim3(1), 0, s,
                       c->tile_nnz.as<int64_t>() + c->num_tiles,
                       w->c_row_nnz_off.as<int64_t>() + R);
    RTT_HIP(hipGetLastError());
    RTT_HIP(hipEventRecord(ev->ev[6], s));
    // ---- numeric ----------------------------------------------------------
    Timeline tl;
    size_t trace_slots = 0;
    const char *trace_path = std::getenv("DMMA_TRACE_FILE");
    if (trace_path && *trace_path && R > 0 && c->nnz > 0) {
        tl.shift = trace_env("DMMA_TRACE_SHIFT", 4, 10);
        const unsigned int stride = 1u << tl.shift;
        tl.phase = trace_env("DMMA_TRACE_PHASE", std::min(7u, stride - 1u), stride - 1u);
        const size_t sampled = static_cast<unsigned int>(R) <= tl.phase
                                   ? 0
                                   : (static_cast<size_t>(R) - 1 - tl.phase) / stride + 1;
        trace_slots = sampled * kWavesPerBlock;
        if (trace_slots) {
            RTT_TRY(w->trace_start.ensure(trace_slots * sizeof(unsigned long long)));
            RTT_TRY(w->trace_end.ensure(trace_slots * sizeof(unsigned long long)));
            RTT_TRY(w->trace_cu.ensure(trace_slots * sizeof(unsigned int)));
            tl.start = w->trace_start.as<unsigned long long>();
            tl.end = w->trace_end.as<unsigned long long>();
            tl.cu = w->trace_cu.as<unsigned int>();
            hipLaunchKernelGGL(fill_u64_kernel, dim3(blocks_for(trace_slots, kThreads)),
                               dim3(kThreads), 0, s, tl.start, static_cast<int64_t>(trace_slots));
            hipLaunchKernelGGL(fill_u64_kernel, dim3(blocks_for(trace_slots, kThreads)),
                               dim3(kThreads), 0, s, tl.end, static_cast<int64_t>(trace_slots));
            hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(trace_slots, kThreads)),
                               dim3(kThreads), 0, s, reinterpret_cast<int *>(tl.cu),
                               static_cast<int64_t>(trace_slots), 0);
        }
    }
    RTT_HIP(hipEventRecord(ev->ev[4], s));
    if (any_tile) {
        RTT_TRY(tile_numeric(a, tile->bc, *c, *tile, w, s));
        RTT_HIP(hipEventRecord(ev->ev[5], s));
    }
    if (any_vector && R > 0 && c->nnz > 0) {
        if (trace_slots)
            hipLaunchKernelGGL((numeric_row_kernel<SmallRow, true>), dim3(R), dim3(kThreads), 0,
                               s, g, b.records.as<int4>(), b.num_tiles, nullptr,
                               w->row_level.as<int>(), st, co, tl);
        else
            hipLaunchKernelGGL((numeric_row_kernel<SmallRow>), dim3(R), dim3(kThreads), 0, s, g,
                               b.records.as<int4>(), b.num_tiles, nullptr,
                               w->row_level.as<int>(), st, co, tl);
    }
    RTT_HIP(hipGetLastError());
    if (!any_tile)
        RTT_HIP(hipEventRecord(ev->ev[5], s));
    if (any_vector && n_num_tiny > 0)
        hipLaunchKernelGGL((numeric_row_kernel<TinyRow>), dim3(n_num_tiny), dim3(TinyRow::kThreads),
                           0, s, g, b.records.as<int4>(), b.num_tiles, w->tiny_rows.as<int>(),
                           w->row_level.as<int>(), st, co, Timeline{});
    if (any_vector && n_num_mid > 0)
        hipLaunchKernelGGL((numeric_row_kernel<MediumRow>), dim3(n_num_mid), dim3(kThreads), 0,
                           s, g, b.records.as<int4>(), b.num_tiles, w->mid_rows.as<int>(),
                           w->row_level.as<int>(), st, co, Timeline{});
    if (n_num_lds > 0)
        hipLaunchKernelGGL((numeric_row_kernel<LargeRow>), dim3(n_num_lds), dim3(kThreads), 0,
                           s, g, b.records.as<int4>(), b.num_tiles, w->lds_rows.as<int>(),
                           w->row_level.as<int>(), st, co, Timeline{});
    if (n_num_slab > 0) {
        DevBuf &tab = w->slab;   // reused: the symbolic slab pass has been queued before
        const int blocks = slab_grid(n_num_slab, num_cus, BC, sizeof(int4), tab);
        const int64_t ents = static_cast<int64_t>(blocks) * BC;
        RTT_TRY(tab.ensure(static_cast<size_t>(ents) * sizeof(int4)));
        hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(4 * ents, kThreads)), dim3(kThreads), 0,
                           s, tab.as<int>(), 4 * ents, -1);
        // The symbolic fill pass above is queued before this sort reuses rows_sorted.
        const int *num_rows = w->num_rows.as<int>();
        if (lb_mode == RTT_LB_ADAPTIVE_TAIL)
            RTT_TRY(order_rows_by_cost(w->num_rows.as<int>(), n_num_slab, g, w, s, &num_rows));
        hipLaunchKernelGGL(numeric_slab_kernel, dim3(blocks), dim3(kThreads), 0, s, g,
                           b.records.as<int4>(), b.num_tiles, num_rows, n_num_slab, queue + 2,
                           static_rows, tab.as<int4>(), co);
        RTT_HIP(hipGetLastError());
    }
    RTT_HIP(hipEventRecord(ev->ev[7], s));
    RTT_HIP(hipStreamSynchronize(s));
    if (any_tile && c->num_tiles > 0) {
        RTT_HIP(hipMemcpy(&times->mfma_pairs, w->tile_pairs.ptr, sizeof(unsigned long long),
                          hipMemcpyDeviceToHost));
        if (tile->tail)
            RTT_TRY(tail_stats(w, &times->tail_reason, &times->tail_records,
                               &times->tail_chunks));
    }
    if (trace_slots)
        RTT_TRY(write_timeline(trace_path, tl, trace_slots, static_cast<uint64_t>(R),
                               static_cast<uint64_t>(c->num_tiles), num_cus));
    // Structure write and numeric share the stream without a host wait
    // between them; their split comes from the device events.
    const double fill_ms = event_ms(ev->ev[2], ev->ev[6]);
    const double tail_ms = ms_since(t_fill);
    times->symbolic_ms = sym_count_ms + fill_ms;
    times->numeric_ms = std::max(0.0, tail_ms - fill_ms);
    times->sym_count_kernel_ms = event_ms(ev->ev[0], ev->ev[1]);
    times->sym_fill_kernel_ms = event_ms(ev->ev[2], ev->ev[3]);
    times->numeric_kernel_ms = event_ms(ev->ev[4], ev->ev[5]);
    times->pairs = pairs;
    times->heavy_rows = static_cast<unsigned long long>(n_num_slab + n_num_lds + n_num_mid);
    return RTT_OK;
}

} // namespace rtt
