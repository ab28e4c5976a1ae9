// SpGEMM core: symbolic -> allocate -> numeric for C = A * B with A in 8x4
// row-major tiles, B in 4x8 column-major tiles and C in 8x8 tiles.
//
// Work unit: one A tile row I (8 scalar rows).  Its (A tile, B tile) pairs
// are the products A(I,K) x B(K,J) for every A tile in row I and every B tile
// in B's tile row K -- a tile-level Gustavson expansion, so no K-list merge
// or candidate sort is needed (the reference's Super16 candidate/exact steps,
// src/super16_symbolic.cuh:2406-2774, and its K-merge numeric,
// src/dmma_spgemm.h:4307-4410, do the equivalent through merges).
//
//  * symbolic: the 8x8 masks of all C tiles of row I are ORed into a hash
//    table in LDS (one wave per row); rows whose tile set outgrows the table
//    are redone by a workgroup over a direct-mapped mask slab in HBM.
//  * numeric: the exact C layout of row I (sorted tile columns, masks,
//    value offsets) is loaded into LDS; every scalar product is added into an
//    LDS accumulator laid out exactly like the row's output segment, which is
//    then streamed to HBM.  Rows whose segment does not fit go to a
//    workgroup path that accumulates with FP64 atomics directly in HBM.
#include <climits>

#include "device.hpp"
#include "internal.hpp"

namespace rtt {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;

constexpr int kSymSlots = 512;     // LDS hash slots per wave (power of 2)
constexpr int kSymSlotBits = 9;
constexpr int kSymLimit = 384;     // distinct C tiles before the slab path
constexpr int kSymPerLane = kSymLimit / kWave;   // 6

constexpr int kNumSlots = 512;     // numeric J -> local tile index hash
constexpr int kNumSlotBits = 9;
constexpr int kNumAcc = 1024;      // FP64 accumulator slots per wave

constexpr int kEmpty = -1;

enum RowLevel : int { kLevelEmpty = 0, kLevelWave = 1, kLevelSlab = 3 };

struct Operands {
    int a_tile_rows;
    const int *__restrict__ a_trp;
    const int *__restrict__ a_tci;
    const uint32_t *__restrict__ a_masks;
    const int *__restrict__ a_off;
    const double *__restrict__ a_vals;
    const int *__restrict__ b_trp;
    const int *__restrict__ b_tci;
    const uint32_t *__restrict__ b_masks;
    const int *__restrict__ b_off;
    const double *__restrict__ b_vals;
    int b_tile_cols;
};

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

__device__ __forceinline__ uint32_t slot_of(int j, int bits)
{
    return hash_tile_col(j) >> (32 - bits);
}

// Eight local row offsets of an 8x8 tile mask packed into one 64-bit word
// (byte r = number of entries in rows < r).
__device__ __forceinline__ uint64_t packed_row_offsets(uint64_t m)
{
    uint64_t out = 0;
#pragma unroll
    for (int r = 1; r < 8; ++r)
        out |= static_cast<uint64_t>(__popcll(m & ((1ull << (8 * r)) - 1ull))) << (8 * r);
    return out;
}

// Writes the structural part of one C tile: column, mask, value offset, row
// offsets and the local column of every entry.
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

// Loads up to 64 A tiles of one row into LDS with their B-row extents and an
// exclusive prefix of pair counts; returns the chunk's pair count.
__device__ __forceinline__ int load_a_chunk(const Operands &g, int c0, int n, uint32_t *s_am,
                                            int *s_b0, int *s_pre)
{
    const int lane = lane_id();
    int nb = 0;
    if (lane < n) {
        const int a = c0 + lane;
        const int k = g.a_tci[a];
        const int b0 = g.b_trp[k];
        nb = g.b_trp[k + 1] - b0;
        s_am[lane] = g.a_masks[a];
        s_b0[lane] = b0;
    }
    const int incl = wave_inclusive_scan(nb);
    if (lane < n)
        s_pre[lane] = incl - nb;
    return __shfl(incl, kWave - 1, kWave);
}

// ---------------------------------------------------------------------------
// Symbolic, one wave per A tile row, LDS hash of (tile column -> 8x8 mask).
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ __launch_bounds__(kThreads) void symbolic_wave_kernel(
    Operands g, int64_t *__restrict__ row_tiles, int64_t *__restrict__ row_nnz,
    int *__restrict__ row_level, unsigned long long *__restrict__ pair_counter, COut c)
{
    __shared__ int s_key[kWavesPerBlock][kSymSlots];
    __shared__ unsigned long long s_msk[kWavesPerBlock][kSymSlots];
    __shared__ uint32_t s_am[kWavesPerBlock][kWave];
    __shared__ int s_b0[kWavesPerBlock][kWave];
    __shared__ int s_pre[kWavesPerBlock][kWave];
    __shared__ int s_cnt[kWavesPerBlock];
    __shared__ int s_ovf[kWavesPerBlock];

    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    const int row = blockIdx.x * kWavesPerBlock + w;
    if (row >= g.a_tile_rows)
        return;
    if (FILL && row_level[row] != kLevelWave)
        return;
    int *key = s_key[w];
    unsigned long long *msk = s_msk[w];
    for (int i = lane; i < kSymSlots; i += kWave) {
        key[i] = kEmpty;
        msk[i] = 0;
    }
    if (lane == 0) {
        s_cnt[w] = 0;
        s_ovf[w] = 0;
    }
    wave_sync();

    const int a0 = g.a_trp[row], a1 = g.a_trp[row + 1];
    unsigned long long pairs = 0;
    for (int c0 = a0; c0 < a1; c0 += kWave) {
        const int n = min(kWave, a1 - c0);
        const int total = load_a_chunk(g, c0, n, s_am[w], s_b0[w], s_pre[w]);
        wave_sync();
        for (int p = lane; p < total; p += kWave) {
            const int ai = upper_index(s_pre[w], n, p);
            const int b = s_b0[w][ai] + (p - s_pre[w][ai]);
            const int j = g.b_tci[b];
            const uint64_t pm = tile_bool_product(s_am[w][ai], g.b_masks[b]);
            if (!pm)
                continue;
            uint32_t h = slot_of(j, kSymSlotBits);
            for (int probe = 0; probe < kSymSlots; ++probe) {
                int k = key[h];
                if (k == kEmpty) {
                    k = atomicCAS(&key[h], kEmpty, j);
                    if (k == kEmpty) {
                        atomicOr(&msk[h], static_cast<unsigned long long>(pm));
                        if (atomicAdd(&s_cnt[w], 1) >= kSymLimit)
                            s_ovf[w] = 1;
                        break;
                    }
                }
                if (k == j) {
                    atomicOr(&msk[h], static_cast<unsigned long long>(pm));
                    break;
                }
                h = (h + 1) & (kSymSlots - 1);
                if (probe == kSymSlots - 1)
                    s_ovf[w] = 1;
            }
        }
        pairs += total;
        wave_sync();
        if (s_ovf[w])
            break;
    }
    wave_sync();
    const bool overflow = s_ovf[w] != 0;
    const int nt = s_cnt[w];

    if constexpr (!FILL) {
        if (overflow) {
            if (lane == 0) {
                row_level[row] = kLevelSlab;
                row_tiles[row] = 0;
                row_nnz[row] = 0;
            }
            return;
        }
        long long nz = 0;
        for (int i = lane; i < kSymSlots; i += kWave)
            nz += __popcll(msk[i]);
        nz = wave_sum(nz);
        if (lane == 0) {
            row_level[row] = nt > 0 ? kLevelWave : kLevelEmpty;
            row_tiles[row] = nt;
            row_nnz[row] = nz;
            atomicAdd(pair_counter, pairs);
        }
    } else {
        // Compact the occupied slots (slot order), then rank by tile column.
        int rk[kSymSlots / kWave];
        unsigned long long rm[kSymSlots / kWave];
#pragma unroll
        for (int i = 0; i < kSymSlots / kWave; ++i) {
            rk[i] = key[lane + i * kWave];
            rm[i] = msk[lane + i * kWave];
        }
        wave_sync();
        const uint64_t lt = (1ull << lane) - 1ull;
        int base = 0;
#pragma unroll
        for (int i = 0; i < kSymSlots / kWave; ++i) {
            const bool occ = rk[i] != kEmpty;
            const uint64_t bal = __ballot(occ);
            if (occ) {
                const int pos = base + __popcll(bal & lt);
                key[pos] = rk[i];
                msk[pos] = rm[i];
            }
            base += __popcll(bal);
        }
        wave_sync();
        int ej[kSymPerLane], er[kSymPerLane];
        unsigned long long em[kSymPerLane];
#pragma unroll
        for (int i = 0; i < kSymPerLane; ++i) {
            const int e = lane + i * kWave;
            ej[i] = e < nt ? key[e] : INT_MAX;
            em[i] = e < nt ? msk[e] : 0ull;
            er[i] = 0;
        }
        for (int f = 0; f < nt; ++f) {
            const int jf = key[f];
#pragma unroll
            for (int i = 0; i < kSymPerLane; ++i)
                er[i] += jf < ej[i] ? 1 : 0;
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < kSymPerLane; ++i)
            if (lane + i * kWave < nt) {
                key[er[i]] = ej[i];
                msk[er[i]] = em[i];
            }
        wave_sync();
        const int64_t tb = c.tile_ptr[row];
        const int64_t zb = c.row_nnz_off[row];
        int64_t carry = 0;
        for (int q0 = 0; q0 < nt; q0 += kWave) {
            const int q = q0 + lane;
            const uint64_t m = q < nt ? msk[q] : 0ull;
            const int pc = __popcll(m);
            const int incl = wave_inclusive_scan(pc);
            if (q < nt)
                write_tile_structure(c, tb + q, key[q], m, zb + carry + (incl - pc));
            carry += __shfl(incl, kWave - 1, kWave);
        }
    }
}

// ---------------------------------------------------------------------------
// Symbolic for wide rows: one workgroup per row, direct-mapped mask slab in
// HBM (one 64-bit mask per C tile column), read back in column order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int block_exclusive_scan(int v, int *s_tmp, int *total)
{
    const int w = threadIdx.x / kWave;
    const int incl = wave_inclusive_scan(v);
    if (lane_id() == kWave - 1)
        s_tmp[w] = incl;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; ++i) {
        const int t = s_tmp[i];
        before += i < w ? t : 0;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

__device__ __forceinline__ unsigned long long load_coherent(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool FILL>
__global__ __launch_bounds__(kThreads) void symbolic_slab_kernel(
    Operands g, const int *__restrict__ rows, int nrows, unsigned long long *__restrict__ slabs,
    int64_t *__restrict__ row_tiles, int64_t *__restrict__ row_nnz,
    unsigned long long *__restrict__ pair_counter, COut c)
{
    __shared__ uint32_t s_am[kThreads];
    __shared__ int s_b0[kThreads];
    __shared__ int s_pre[kThreads];
    __shared__ int s_tmp[kWavesPerBlock];
    __shared__ int s_lo, s_hi;
    unsigned long long *slab = slabs + static_cast<int64_t>(blockIdx.x) * g.b_tile_cols;
    const int tid = threadIdx.x;
    for (int ri = blockIdx.x; ri < nrows; ri += gridDim.x) {
        const int row = rows[ri];
        const int a0 = g.a_trp[row], a1 = g.a_trp[row + 1];
        if (tid == 0) {
            s_lo = INT_MAX;
            s_hi = -1;
        }
        __syncthreads();
        unsigned long long pairs = 0;
        for (int c0 = a0; c0 < a1; c0 += kThreads) {
            const int n = min(kThreads, a1 - c0);
            int nb = 0;
            if (tid < n) {
                const int a = c0 + tid;
                const int k = g.a_tci[a];
                const int b0 = g.b_trp[k];
                nb = g.b_trp[k + 1] - b0;
                s_am[tid] = g.a_masks[a];
                s_b0[tid] = b0;
                if (nb > 0) {
                    atomicMin(&s_lo, g.b_tci[b0]);
                    atomicMax(&s_hi, g.b_tci[b0 + nb - 1]);
                }
            }
            int total = 0;
            const int excl = block_exclusive_scan(nb, s_tmp, &total);
            if (tid < n)
                s_pre[tid] = excl;
            __syncthreads();
            for (int p = tid; p < total; p += kThreads) {
                const int ai = upper_index(s_pre, n, p);
                const int b = s_b0[ai] + (p - s_pre[ai]);
                const uint64_t pm = tile_bool_product(s_am[ai], g.b_masks[b]);
                if (pm)
                    atomicOr(&slab[g.b_tci[b]], static_cast<unsigned long long>(pm));
            }
            pairs += total;
            __syncthreads();
        }
        __threadfence();
        __syncthreads();
        const int lo = s_lo, hi = s_hi;
        int64_t tiles = 0, nz = 0;
        int64_t tb = 0, zb = 0;
        if constexpr (FILL) {
            tb = c.tile_ptr[row];
            zb = c.row_nnz_off[row];
        }
        for (int j0 = lo; j0 <= hi; j0 += kThreads) {
            const int j = j0 + tid;
            unsigned long long m = 0;
            if (j <= hi) {
                m = load_coherent(&slab[j]);
                if (m)
                    slab[j] = 0;
            }
            const int occ = m ? 1 : 0;
            const int pc = __popcll(m);
            int ntile = 0, npc = 0;
            const int tpos = block_exclusive_scan(occ, s_tmp, &ntile);
            const int zpos = block_exclusive_scan(pc, s_tmp, &npc);
            if constexpr (FILL) {
                if (occ)
                    write_tile_structure(c, tb + tiles + tpos, j, m, zb + nz + zpos);
            }
            tiles += ntile;
            nz += npc;
        }
        if constexpr (!FILL) {
            if (tid == 0) {
                row_tiles[row] = tiles;
                row_nnz[row] = nz;
            }
            pairs = tid == 0 ? pairs : 0;
            if (tid == 0)
                atomicAdd(pair_counter, pairs);
        }
        __threadfence();
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Numeric, one wave per A tile row with an LDS accumulator.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void numeric_wave_kernel(
    Operands g, const int *__restrict__ row_level, COut c)
{
    __shared__ int s_hk[kWavesPerBlock][kNumSlots];
    __shared__ int s_hq[kWavesPerBlock][kNumSlots];
    __shared__ double s_acc[kWavesPerBlock][kNumAcc];
    __shared__ uint32_t s_am[kWavesPerBlock][kWave];
    __shared__ int s_b0[kWavesPerBlock][kWave];
    __shared__ int s_pre[kWavesPerBlock][kWave];
    __shared__ int s_ao[kWavesPerBlock][kWave];

    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    const int row = blockIdx.x * kWavesPerBlock + w;
    if (row >= g.a_tile_rows || row_level[row] != kLevelWave)
        return;
    const int64_t t0 = c.tile_ptr[row];
    const int nt = static_cast<int>(c.tile_ptr[row + 1] - t0);
    const int64_t zb = c.row_nnz_off[row];
    const int nnz_row = static_cast<int>(c.row_nnz_off[row + 1] - zb);
    if (nt > kSymLimit || nnz_row > kNumAcc)
        return;   // the workgroup path takes this row
    int *hk = s_hk[w];
    int *hq = s_hq[w];
    double *acc = s_acc[w];
    for (int i = lane; i < kNumSlots; i += kWave)
        hk[i] = kEmpty;
    for (int i = lane; i < nnz_row; i += kWave)
        acc[i] = 0.0;
    wave_sync();
    for (int q = lane; q < nt; q += kWave) {
        const int j = c.col[t0 + q];
        uint32_t h = slot_of(j, kNumSlotBits);
        while (atomicCAS(&hk[h], kEmpty, j) != kEmpty)
            h = (h + 1) & (kNumSlots - 1);
        hq[h] = q;
    }
    wave_sync();

    const int a0 = g.a_trp[row], a1 = g.a_trp[row + 1];
    for (int c0 = a0; c0 < a1; c0 += kWave) {
        const int n = min(kWave, a1 - c0);
        const int total = load_a_chunk(g, c0, n, s_am[w], s_b0[w], s_pre[w]);
        if (lane < n)
            s_ao[w][lane] = c0 + lane;
        wave_sync();
        for (int p = lane; p < total; p += kWave) {
            const int ai = upper_index(s_pre[w], n, p);
            const int b = s_b0[w][ai] + (p - s_pre[w][ai]);
            const uint32_t am = s_am[w][ai];
            const uint32_t bm = g.b_masks[b];
            const int j = g.b_tci[b];
            uint32_t h = slot_of(j, kNumSlotBits);
            int q = -1;
            for (;;) {
                const int k = hk[h];
                if (k == j) {
                    q = hq[h];
                    break;
                }
                if (k == kEmpty)
                    break;
                h = (h + 1) & (kNumSlots - 1);
            }
            if (q < 0)
                continue;   // structurally empty product
            const int64_t tq = t0 + q;
            const uint64_t cm = c.mask[tq];
            const int cb = static_cast<int>(c.tile_nnz[tq] - zb);
            const int at = s_ao[w][ai];
            const int aoff = g.a_off[at];
            const int aspan = g.a_off[at + 1] - aoff;
            const int boff = g.b_off[b];
            const int bspan = g.b_off[b + 1] - boff;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t rows = stride4_bits(am, k);
                const uint32_t cols = stride4_bits(bm, k);
                if (!cols)
                    continue;
                while (rows) {
                    const int r = __builtin_ctz(rows);
                    rows &= rows - 1;
                    const double va = g.a_vals[aoff + payload_slot(am, aspan, r * 4 + k)];
                    uint32_t cc = cols;
                    while (cc) {
                        const int cl = __builtin_ctz(cc);
                        cc &= cc - 1;
                        const double vb = g.b_vals[boff + payload_slot(bm, bspan, cl * 4 + k)];
                        const int pos = r * 8 + cl;
                        const int slot = cb + __popcll(cm & ((1ull << pos) - 1ull));
                        atomicAdd(&acc[slot], va * vb);
                    }
                }
            }
        }
        wave_sync();
    }
    wave_sync();
    for (int i = lane; i < nnz_row; i += kWave)
        c.val[zb + i] = acc[i];
}

// ---------------------------------------------------------------------------
// Numeric for rows that do not fit the wave path: one workgroup per row,
// direct-mapped (tile column -> local tile) slab, FP64 atomics in HBM.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void numeric_slab_kernel(
    Operands g, const int *__restrict__ rows, int nrows, int *__restrict__ slabs, COut c)
{
    __shared__ uint32_t s_am[kThreads];
    __shared__ int s_b0[kThreads];
    __shared__ int s_pre[kThreads];
    __shared__ int s_tmp[kWavesPerBlock];
    int *slab = slabs + static_cast<int64_t>(blockIdx.x) * g.b_tile_cols;
    const int tid = threadIdx.x;
    for (int ri = blockIdx.x; ri < nrows; ri += gridDim.x) {
        const int row = rows[ri];
        const int64_t t0 = c.tile_ptr[row];
        const int nt = static_cast<int>(c.tile_ptr[row + 1] - t0);
        const int64_t zb = c.row_nnz_off[row];
        const int64_t nnz_row = c.row_nnz_off[row + 1] - zb;
        for (int q = tid; q < nt; q += kThreads)
            slab[c.col[t0 + q]] = q;
        for (int64_t i = tid; i < nnz_row; i += kThreads)
            c.val[zb + i] = 0.0;
        __threadfence();
        __syncthreads();
        const int a0 = g.a_trp[row], a1 = g.a_trp[row + 1];
        for (int c0 = a0; c0 < a1; c0 += kThreads) {
            const int n = min(kThreads, a1 - c0);
            int nb = 0;
            if (tid < n) {
                const int a = c0 + tid;
                const int k = g.a_tci[a];
                const int b0 = g.b_trp[k];
                nb = g.b_trp[k + 1] - b0;
                s_am[tid] = g.a_masks[a];
                s_b0[tid] = b0;
            }
            int total = 0;
            const int excl = block_exclusive_scan(nb, s_tmp, &total);
            if (tid < n)
                s_pre[tid] = excl;
            __syncthreads();
            for (int p = tid; p < total; p += kThreads) {
                const int ai = upper_index(s_pre, n, p);
                const int b = s_b0[ai] + (p - s_pre[ai]);
                const uint32_t am = s_am[ai];
                const uint32_t bm = g.b_masks[b];
                if (!tile_bool_product(am, bm))
                    continue;
                const int q = __hip_atomic_load(&slab[g.b_tci[b]], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
                const int64_t tq = t0 + q;
                const uint64_t cm = c.mask[tq];
                const int64_t cb = c.tile_nnz[tq];
                const int at = c0 + ai;
                const int aoff = g.a_off[at];
                const int aspan = g.a_off[at + 1] - aoff;
                const int boff = g.b_off[b];
                const int bspan = g.b_off[b + 1] - boff;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    uint32_t rws = stride4_bits(am, k);
                    const uint32_t cols = stride4_bits(bm, k);
                    if (!cols)
                        continue;
                    while (rws) {
                        const int r = __builtin_ctz(rws);
                        rws &= rws - 1;
                        const double va = g.a_vals[aoff + payload_slot(am, aspan, r * 4 + k)];
                        uint32_t cc = cols;
                        while (cc) {
                            const int cl = __builtin_ctz(cc);
                            cc &= cc - 1;
                            const double vb =
                                g.b_vals[boff + payload_slot(bm, bspan, cl * 4 + k)];
                            const int pos = r * 8 + cl;
                            unsafeAtomicAdd(&c.val[cb + __popcll(cm & ((1ull << pos) - 1ull))],
                                            va * vb);
                        }
                    }
                }
            }
            __syncthreads();
        }
        __threadfence();
        __syncthreads();
        for (int q = tid; q < nt; q += kThreads)
            slab[c.col[t0 + q]] = -1;
        __threadfence();
        __syncthreads();
    }
}

__global__ void collect_rows_kernel(int n, const int *__restrict__ level, int want,
                                    const int64_t *__restrict__ tile_ptr,
                                    const int64_t *__restrict__ nnz_off, int *__restrict__ out,
                                    int *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    bool take;
    if (want == kLevelSlab && !tile_ptr) {
        take = level[i] == kLevelSlab;
    } else {
        // numeric fallback: slab-level rows and wave rows that do not fit
        const int64_t nt = tile_ptr[i + 1] - tile_ptr[i];
        const int64_t nz = nnz_off[i + 1] - nnz_off[i];
        take = level[i] == kLevelSlab ||
               (level[i] == kLevelWave && (nt > kSymLimit || nz > kNumAcc));
    }
    if (take)
        out[atomicAdd(count, 1)] = i;
}

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

int read_scalar(const void *d, void *h, size_t bytes, hipStream_t s)
{
    RTT_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
    RTT_HIP(hipStreamSynchronize(s));
    return RTT_OK;
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

int spgemm_core(const DevTiles &a, const DevTiles &b, DevC *c, SpgemmWork *w,
                Events *ev, CoreTimes *times)
{
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
    g.a_tci = a.tile_col_idx.as<int>();
    g.a_masks = a.masks.as<uint32_t>();
    g.a_off = a.value_offsets.as<int>();
    g.a_vals = a.values.as<double>();
    g.b_trp = b.tile_row_ptr.as<int>();
    g.b_tci = b.tile_col_idx.as<int>();
    g.b_masks = b.masks.as<uint32_t>();
    g.b_off = b.value_offsets.as<int>();
    g.b_vals = b.values.as<double>();
    g.b_tile_cols = BC;

    const int num_cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
    auto t_sym = Clock::now();
    // ---- symbolic: per-row counts ------------------------------------
    RTT_TRY(w->row_tiles.ensure((R + 1) * sizeof(int64_t)));
    RTT_TRY(w->row_nnz.ensure((R + 1) * sizeof(int64_t)));
    RTT_TRY(w->row_level.ensure((R + 1) * sizeof(int)));
    RTT_TRY(w->counters.ensure(4 * sizeof(unsigned long long)));
    RTT_TRY(w->l3_rows.ensure((R + 1) * sizeof(int)));
    RTT_TRY(c->tile_ptr.ensure((R + 1) * sizeof(int64_t)));
    RTT_TRY(w->c_row_nnz_off.ensure((R + 1) * sizeof(int64_t)));
    unsigned long long *ctr = w->counters.as<unsigned long long>();
    int *slab_count = reinterpret_cast<int *>(ctr + 1);
    int *num_count = reinterpret_cast<int *>(ctr + 2);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(4), 0, s, ctr, int64_t(4));
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(1), 0, s,
                       reinterpret_cast<unsigned long long *>(w->row_tiles.as<int64_t>() + R),
                       int64_t(1));
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(1), 0, s,
                       reinterpret_cast<unsigned long long *>(w->row_nnz.as<int64_t>() + R),
                       int64_t(1));
    COut none{};
    RTT_HIP(hipEventRecord(ev->ev[0], s));
    if (R > 0)
        hipLaunchKernelGGL((symbolic_wave_kernel<false>), dim3(blocks_for(R, kWavesPerBlock)),
                           dim3(kThreads), 0, s, g, w->row_tiles.as<int64_t>(),
                           w->row_nnz.as<int64_t>(), w->row_level.as<int>(), ctr, none);
    RTT_HIP(hipGetLastError());
    RTT_HIP(hipEventRecord(ev->ev[1], s));
    int n_slab = 0;
    if (R > 0) {
        hipLaunchKernelGGL(collect_rows_kernel, dim3(blocks_for(R, kThreads)), dim3(kThreads), 0,
                           s, R, w->row_level.as<int>(), int(kLevelSlab), nullptr, nullptr,
                           w->l3_rows.as<int>(), slab_count);
        RTT_TRY(read_scalar(slab_count, &n_slab, sizeof(int), s));
    }
    int slab_blocks = 0;
    if (n_slab > 0) {
        slab_blocks = std::min(n_slab, 2 * num_cus);
        RTT_TRY(w->slab.ensure(static_cast<size_t>(slab_blocks) * BC * sizeof(uint64_t)));
        hipLaunchKernelGGL(fill_u64_kernel,
                           dim3(blocks_for(static_cast<int64_t>(slab_blocks) * BC, kThreads)),
                           dim3(kThreads), 0, s, w->slab.as<unsigned long long>(),
                           static_cast<int64_t>(slab_blocks) * BC);
        hipLaunchKernelGGL((symbolic_slab_kernel<false>), dim3(slab_blocks), dim3(kThreads), 0, s,
                           g, w->l3_rows.as<int>(), n_slab, w->slab.as<unsigned long long>(),
                           w->row_tiles.as<int64_t>(), w->row_nnz.as<int64_t>(), ctr, none);
        RTT_HIP(hipGetLastError());
    }
    RTT_TRY(exclusive_scan_i64(w->row_tiles.as<int64_t>(), c->tile_ptr.as<int64_t>(), R,
                               &w->scan_tmp, s));
    RTT_TRY(exclusive_scan_i64(w->row_nnz.as<int64_t>(), w->c_row_nnz_off.as<int64_t>(), R,
                               &w->scan_tmp, s));
    int64_t totals[2] = {0, 0};
    RTT_TRY(read_scalar(c->tile_ptr.as<int64_t>() + R, &totals[0], sizeof(int64_t), s));
    RTT_TRY(read_scalar(w->c_row_nnz_off.as<int64_t>() + R, &totals[1], sizeof(int64_t), s));
    unsigned long long pairs = 0;
    RTT_TRY(read_scalar(ctr, &pairs, sizeof(pairs), s));
    const double sym_count_ms = ms_since(t_sym);

    // ---- allocate C ---------------------------------------------------
    auto t_alloc = Clock::now();
    c->num_tiles = totals[0];
    c->nnz = totals[1];
    RTT_TRY(c->tile_col.ensure(c->num_tiles * sizeof(int)));
    RTT_TRY(c->tile_mask.ensure(c->num_tiles * sizeof(uint64_t)));
    RTT_TRY(c->tile_nnz.ensure((c->num_tiles + 1) * sizeof(int64_t)));
    RTT_TRY(c->tile_csr_ptr.ensure(c->num_tiles * 8));
    RTT_TRY(c->tile_csr_col.ensure(c->nnz));
    RTT_TRY(c->tile_csr_val.ensure(c->nnz * sizeof(double)));
    times->allocation_ms = ms_since(t_alloc);

    // ---- symbolic: exact structure ---------------------------------------
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
    if (R > 0 && c->num_tiles > 0)
        hipLaunchKernelGGL((symbolic_wave_kernel<true>), dim3(blocks_for(R, kWavesPerBlock)),
                           dim3(kThreads), 0, s, g, nullptr, nullptr, w->row_level.as<int>(),
                           nullptr, co);
    RTT_HIP(hipGetLastError());
    RTT_HIP(hipEventRecord(ev->ev[3], s));
    if (n_slab > 0)
        hipLaunchKernelGGL((symbolic_slab_kernel<true>), dim3(slab_blocks), dim3(kThreads), 0, s,
                           g, w->l3_rows.as<int>(), n_slab, w->slab.as<unsigned long long>(),
                           nullptr, nullptr, nullptr, co);
    hipLaunchKernelGGL(set_last_kernel, dim3(1), dim3(1), 0, s,
                       c->tile_nnz.as<int64_t>() + c->num_tiles,
                       w->c_row_nnz_off.as<int64_t>() + R);
    RTT_HIP(hipGetLastError());
    RTT_HIP(hipStreamSynchronize(s));
    times->symbolic_ms = sym_count_ms + ms_since(t_fill);

    // ---- numeric --------------------------------------------------------
    auto t_num = Clock::now();
    int n_num_slab = 0;
    if (R > 0) {
        hipLaunchKernelGGL(collect_rows_kernel, dim3(blocks_for(R, kThreads)), dim3(kThreads), 0,
                           s, R, w->row_level.as<int>(), int(kLevelWave),
                           c->tile_ptr.as<int64_t>(), w->c_row_nnz_off.as<int64_t>(),
                           w->l3_rows.as<int>(), num_count);
        RTT_TRY(read_scalar(num_count, &n_num_slab, sizeof(int), s));
    }
    RTT_HIP(hipEventRecord(ev->ev[4], s));
    if (R > 0 && c->nnz > 0)
        hipLaunchKernelGGL(numeric_wave_kernel, dim3(blocks_for(R, kWavesPerBlock)),
                           dim3(kThreads), 0, s, g, w->row_level.as<int>(), co);
    RTT_HIP(hipGetLastError());
    RTT_HIP(hipEventRecord(ev->ev[5], s));
    if (n_num_slab > 0) {
        const int blocks = std::min(n_num_slab, 2 * num_cus);
        DevBuf &idx = w->slab;   // reused: numeric runs after symbolic finished
        RTT_TRY(idx.ensure(static_cast<size_t>(blocks) * BC * sizeof(int)));
        hipLaunchKernelGGL(fill_i32_kernel,
                           dim3(blocks_for(static_cast<int64_t>(blocks) * BC, kThreads)),
                           dim3(kThreads), 0, s, idx.as<int>(), static_cast<int64_t>(blocks) * BC,
                           -1);
        hipLaunchKernelGGL(numeric_slab_kernel, dim3(blocks), dim3(kThreads), 0, s, g,
                           w->l3_rows.as<int>(), n_num_slab, idx.as<int>(), co);
        RTT_HIP(hipGetLastError());
    }
    RTT_HIP(hipStreamSynchronize(s));
    times->numeric_ms = ms_since(t_num);
    times->sym_count_kernel_ms = event_ms(ev->ev[0], ev->ev[1]);
    times->sym_fill_kernel_ms = event_ms(ev->ev[2], ev->ev[3]);
    times->numeric_kernel_ms = event_ms(ev->ev[4], ev->ev[5]);
    times->pairs = pairs;
    times->heavy_rows = static_cast<unsigned long long>(n_num_slab);
    return RTT_OK;
}

} // namespace rtt
