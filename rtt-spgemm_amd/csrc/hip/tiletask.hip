// Tile-pair numeric path: C's values computed per 8x8 C tile with matrix
// cores, after the symbolic pass has written C's whole structure.
//
// A C tile (I, J) is the sum over the K where A's tile row I and B's tile
// column J both hold a tile of A(I,K) * B(K,J), an 8x4 by 4x8 product.  This
// is the reference's C8 task (a K-merge of A row I with B column J, then one
// tensor-core product per matched K: src/dmma_spgemm.h:4164-4300,
// 4307-4410); here one gfx950 v_mfma_f64_4x4x4f64 computes a whole pair --
// its four 4x4x4 blocks are the four 4x4 quarters of the 8x8 C tile -- with
// every lane holding one A entry, one B entry and one C entry.  Each C tile
// keeps its running sum in one register per lane and takes its pairs in
// increasing K, so C's values do not depend on scheduling: two calls give the
// same bits.
//
// Work unit: one wave per C tile row, taken from per-XCD row queues (queue x
// hands out chunks x, x + 8, x + 16, ... of 8 consecutive rows, so the XCDs
// work through one window of the rows together; a wave whose XCD queue is
// drained takes rows from the next XCD's).  The wave stages A row I's tile
// columns in LDS, then, for up to 64 C tiles at a time, enumerates the B
// column entries of all of them (one lane per entry), finds each entry's K
// among A's columns by binary search and compacts the matches, in (C tile,
// K) order, into an LDS pair list.  The pair loop then issues the operand
// loads of four pairs at once and runs their products into the running sum,
// storing each C tile's entries when its pairs end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace rtt {

namespace {

constexpr int kTThreads = 128;
constexpr int kTWaves = kTThreads / kWave;
constexpr int kAKeys = 128;     // A tiles of a row staged in LDS (keys, masks, offsets)
constexpr int kAVals = 512;     // A payload values of a row staged in LDS (0: read from HBM)
constexpr int kPairCap = 128;   // pairs a wave buffers before running them
constexpr int kBatch = 4;       // pairs whose operand loads are issued together (4 ran the
                                // block-FEM product 5 % faster than 8: fewer registers, no
                                // spills, and the next batch's loads are issued sooner)
constexpr int kEnumRounds = 4;  // rounds of 64 B column entries whose K are loaded together
constexpr int kRowGrab = 8;     // tile rows a wave takes from a row queue at once (one atomic:
                                // per-row atomics on 8 heads serialise at the L2); with the
                                // queues interleaved, 8 ran 6 % faster than 16 and 4 or 32
static_assert(kPairCap % kBatch == 0 && kPairCap >= 2 * kWave, "pair buffer");

struct TileTask {
    int rows;                            // tile rows of A and C
    const int *__restrict__ a_trp;       // A tile row pointer
    const int4 *__restrict__ a_meta;     // {K, mask, per-k bytes, offset | dense}
    const double *__restrict__ a_vals;   // A payload, each tile row-major
    const int *__restrict__ b_cptr;      // B tile column pointer
    const int4 *__restrict__ b_cent;     // {K, mask, per-k bytes, offset | dense}
    const double *__restrict__ b_vals;   // B payload, each tile column-major
    const int64_t *__restrict__ c_ptr;   // C tile row pointer
    const int *__restrict__ c_col;
    const unsigned long long *__restrict__ c_mask;
    const int64_t *__restrict__ c_off;   // value offset of each C tile
    double *__restrict__ c_val;
    const int *__restrict__ a_row_off;   // A payload offset of each tile row
    float fill;                          // rows taken: A tiles' mean stored values >= fill
    int *__restrict__ queue;             // kXcds row-queue heads, zero at launch
    unsigned long long *__restrict__ mfma_pairs;   // pairs multiplied (one MFMA each)
    // Heavy C tiles split into chunks (bit t of heavy: C tile t), used when
    // tail->enabled; null when no admission ran.
    const uint32_t *__restrict__ heavy;
    const struct TailState *__restrict__ tail;
};

// Tail queue (the reference's heavy-task split, src/dmma_spgemm.h:2534-2776,
// 4902-5254, 6003-6049): C tiles whose K-merge costs more than the admission
// threshold are split into chunks of their B-column range; each chunk's
// partial tile goes to a workspace slot, and a reduction sums the chunks of a
// tile in chunk order, so the split tiles are as reproducible as the rest.
struct TailRecord {          // DmmaTailRecord {id, scans, matches, chunks} + the tile row
    long long task;
    int row;
    int scans;
    int matches;
    int chunks;
};
struct TailState {
    unsigned int count;              // records admitted (may exceed capacity)
    int overflow;                    // an admission found the record list full
    unsigned long long chunks;       // chunks of the admitted records
    int reason;                      // gate decision (rtt_stats.tail_reason)
    int enabled;                     // 1: heavy tiles take the chunk kernels
    int chunk_head;                  // work queue of the chunk kernel
    int reduce_head;                 // work queue of the reduction
};

constexpr uint32_t kDense = 0x80000000u;

// Entry `phys` (bit position) of a hybrid tile: dense tiles hold all 32 slots
// (absent entries are zeros), sparse tiles their entries in mask-bit order.
__device__ __forceinline__ double tile_entry(const double *__restrict__ vals, uint32_t mask,
                                             uint32_t offw, int phys)
{
    const uint32_t off = offw & ~kDense;
    if (offw & kDense)
        return vals[off + phys];
    if (!((mask >> phys) & 1u))
        return 0.0;
    return vals[off + __popc(mask & ((1u << phys) - 1u))];
}

__device__ __forceinline__ int read_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Entry `phys` of a tile that is the same in every lane, split so that the
// load is unconditional: its mask and offset move to scalar registers, the
// load reads a slot inside the tile (the tile's first slot when the entry is
// absent), and *has says whether the value counts.  The caller applies `has`
// where it uses the value, so that the wait for the load lands there and not
// right after the load.
__device__ __forceinline__ double tile_entry_raw(const double *vals, uint32_t mask, uint32_t offw,
                                                 int phys, uint32_t below, bool *has)
{
    mask = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(mask)));
    offw = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(offw)));
    const uint32_t off = offw & ~kDense;
    const bool dense = (offw & kDense) != 0;
    const bool bit = ((mask >> phys) & 1u) != 0;
    *has = dense || bit;
    const uint32_t idx = dense ? static_cast<uint32_t>(phys) : (bit ? __popc(mask & below) : 0u);
    return vals[off + idx];
}

// tile_entry for a tile that is the same in every lane (the pair loops'
// operands): its mask and offset move to scalar registers, so the dense test
// is a scalar branch and a lane's own work is its mask bit and the popcount
// below it (`below` = (1 << phys) - 1, fixed per lane).
__device__ __forceinline__ double tile_entry_uniform(const double *__restrict__ vals,
                                                     uint32_t mask, uint32_t offw, int phys,
                                                     uint32_t below)
{
    mask = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(mask)));
    offw = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(offw)));
    const uint32_t off = offw & ~kDense;
    if (offw & kDense)
        return vals[off + phys];
    if (!((mask >> phys) & 1u))
        return 0.0;
    return vals[off + __popc(mask & below)];
}

// Lane roles of v_mfma_f64_4x4x4f64 (four independent 4x4x4 blocks; block b,
// measured on gfx950): A[b][i][k] in lane i + 4b + 16k, B[b][k][j] in lane
// j + 4b + 16k, D[b][i][j] in lane j + 4b + 16i.  Block b = 2h + v multiplies
// rows 4h..4h+3 of the A tile by columns 4v..4v+3 of the B tile, so D's four
// blocks are the quarters of the 8x8 C tile.
struct LaneRoles {
    int a_phys;   // A tile bit r*4+k this lane supplies
    int b_phys;   // B tile bit c*4+k
    int c_bit;    // C tile bit r*8+c this lane accumulates
    uint32_t a_below, b_below;   // (1 << phys) - 1
    __device__ LaneRoles(int lane)
    {
        const int blk = (lane >> 2) & 3, lo = lane & 3, hi = lane >> 4;
        a_phys = ((blk >> 1) * 4 + lo) * 4 + hi;
        b_phys = ((blk & 1) * 4 + lo) * 4 + hi;
        c_bit = ((blk >> 1) * 4 + hi) * 8 + (blk & 1) * 4 + lo;
        a_below = (1u << a_phys) - 1u;
        b_below = (1u << b_phys) - 1u;
    }
};

// 64-bit value held by lane `l` (l uniform across the wave).
__device__ __forceinline__ unsigned long long read_lane64(unsigned long long v, int l)
{
    return (static_cast<unsigned long long>(static_cast<uint32_t>(
                read_lane(static_cast<int>(v >> 32), l))) << 32) |
           static_cast<uint32_t>(read_lane(static_cast<int>(v), l));
}

// ---------------------------------------------------------------------------
// Plain FP64: one v_mfma_f64_4x4x4f64 per pair into a running sum of one
// register per lane.  A batch of kBatch pairs (which may span C tiles) has its
// operand loads issued together.
// ---------------------------------------------------------------------------
struct Fp64Product {
    struct Lane {
        LaneRoles roles;
        __device__ Lane(int lane) : roles(lane) {}
    };
    struct Acc {
        double v = 0.0;
    };
    static constexpr unsigned long long kMfmaPerPair = 1;
    // Waves per SIMD the tile kernel is compiled for: the kernel is bound by
    // memory latency, and 4 waves (at most 128 VGPRs) ran it 20 % faster
    // than the 3 the compiler picks unconstrained.
    static constexpr int kOccupancy = 4;

    __device__ static void store(const TileTask &g, const Lane &ln, int c_local,
                                 unsigned long long cm, long long co, const Acc &acc)
    {
        const unsigned long long m = read_lane64(cm, c_local);
        const long long off = static_cast<long long>(read_lane64(static_cast<unsigned long long>(co), c_local));
        const int bit = ln.roles.c_bit;
        if ((m >> bit) & 1ull)
            g.c_val[off + __popcll(m & ((1ull << bit) - 1ull))] = acc.v;
    }

    // Lane's entry of a finished tile into a 64-slot partial (bit order).
    __device__ static void dump(const Lane &ln, const Acc &acc, double *slot64)
    {
        slot64[ln.roles.c_bit] = acc.v;
    }

    // B entries of buffered pairs [p0, p0 + kBatch): one unconditional load
    // each (pairs past n repeat pair n - 1; their values are not used); bit
    // i of *has: the entry counts.
    __device__ static void load_b(const TileTask &g, const Lane &ln, const int4 *pm, int p0, int n,
                                  double (&bv)[kBatch], uint32_t *has)
    {
        uint32_t h = 0;
#pragma unroll
        for (int i = 0; i < kBatch; ++i) {
            const int4 m = pm[min(p0 + i, n - 1)];
            bool hi;
            bv[i] = tile_entry_raw(g.b_vals, static_cast<uint32_t>(m.z), static_cast<uint32_t>(m.w),
                                   ln.roles.b_phys, ln.roles.b_below, &hi);
            h |= hi ? (1u << i) : 0u;
        }
        *has = h;
    }

    // Runs buffered pairs [0, n) in order; a change of C tile stores the
    // finished one.  The B loads of the next batch are issued before the
    // products of the current one, so one batch's loads are always in flight
    // behind the matrix-core work.
    __device__ static void run(const TileTask &g, const Lane &ln, const double *a_vals,
                               const int4 *pm, const unsigned char *pc, int n, int &cur, Acc &acc,
                               unsigned long long cm, long long co)
    {
        if (n <= 0)
            return;
        double bv[kBatch];
        uint32_t bh = 0;
        load_b(g, ln, pm, 0, n, bv, &bh);
        for (int p0 = 0; p0 < n; p0 += kBatch) {
            double bn[kBatch];
            uint32_t bnh = 0;
            load_b(g, ln, pm, p0 + kBatch, n, bn, &bnh);
            const int nbat = min(kBatch, n - p0);
            double av[kBatch];
            uint32_t ah = 0;
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                const int4 m = pm[min(p0 + i, n - 1)];
                bool hi;
                av[i] = tile_entry_raw(a_vals, static_cast<uint32_t>(m.x), static_cast<uint32_t>(m.y),
                                       ln.roles.a_phys, ln.roles.a_below, &hi);
                ah |= hi ? (1u << i) : 0u;
            }
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                if (i < nbat) {
                    const int t = __builtin_amdgcn_readfirstlane(pc[p0 + i]);
                    if (t != cur) {
                        if (cur >= 0)
                            store(g, ln, cur, cm, co, acc);
                        cur = t;
                        acc.v = 0.0;
                    }
                    const double a = ((ah >> i) & 1u) ? av[i] : 0.0;
                    const double b = ((bh >> i) & 1u) ? bv[i] : 0.0;
                    acc.v = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc.v, 0, 0, 0);
                }
            }
#pragma unroll
            for (int i = 0; i < kBatch; ++i)
                bv[i] = bn[i];
            bh = bnh;
        }
    }
};

// ---------------------------------------------------------------------------
// Split precision (Ozaki scheme, error-free slicing): for a batch of at most
// eight pairs of one C tile, every A entry is written relative to the largest
// exponent E_r of its tile row r over the batch as a 64-bit fixed-point
// fraction, sign * sum_e q_e 2^(E_r - 8(e+1)) with eight integer slices
// q_e in [0, 255] (the bytes of the fraction), and every B entry likewise
// relative to its tile column's largest exponent.  Slices are exact in
// BF16, and a product of two slices is exact in FP32.  All 64 slice pairs
// are kept; grouped by d = e + f (0..14), the sums
//   P_d[r][c] = sum_{k, e+f=d} qa_e[r][k] qb_f[k][c]
// have at most 8 x 4 terms of at most 255^2 per pair, so eight pairs' sums
// stay below 2^24 and the FP32 accumulation of v_mfma_f32_16x16x32_bf16 is
// exact.  The reduction dimension of one instruction is 32 = (k, slice e):
// lane group k holds the eight slices of a(r, k); B supplies, for output
// column c + 8h, the slices qb_{d-e} with d = 2t + h, so instruction t gives
// P_{2t} and P_{2t+1} side by side (rows 8..15 of the 16x16 output are
// unused).  Each batch is then recombined in FP64 on the GPU,
//   C[r][c] += 2^(E_r + E_c) sum_d P_d 2^(-8(d+2)),
// and batches and C tiles accumulate in FP64.  The only approximation is the
// truncation of every operand 64 bits below its row (column) maximum, a
// relative error of at most 2^-63 of that maximum per entry.  A batch whose
// A rows or B columns span more than kSpreadBits binades (where the smallest
// entries would keep fewer than 43 bits) runs on the FP64 matrix cores
// instead, so every product keeps a relative error below 2^-42 and the
// forward error of C stays below about 2^-41 sum |a||b| whatever the data.
// Lane roles of v_mfma_f32_16x16x32_bf16 (measured on gfx950): A[i][k'] in
// lane i + 16(k' >> 3), element k' & 7; B[k'][j] in lane j + 16(k' >> 3),
// element k' & 7; D[i][j] in lane j + 16(i >> 2), element i & 3.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kSplitBatch = 8;   // pairs whose slice sums stay exact in FP32

// Biased exponent of |x| for alignment (subnormals count as 1, zero as 0).
__device__ __forceinline__ int align_exp(double x)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    const int ex = static_cast<int>((b >> 52) & 0x7FF);
    return ex ? ex : ((b << 1) ? 1 : 0);
}

// 64-bit fixed-point fraction of |x| / 2^(emax - 1022) (emax >= align_exp(x),
// emax >= 1): floor(|x| 2^(64 - (emax - 1022))).
__device__ __forceinline__ unsigned long long fixed_fraction(double x, int emax)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    int ex = static_cast<int>((b >> 52) & 0x7FF);
    unsigned long long m = b & ((1ull << 52) - 1ull);
    if (ex)
        m |= 1ull << 52;
    else
        ex = 1;
    const int sh = ex - emax + 11;   // |x| = m 2^(ex - 1075)
    if (sh >= 0)
        return m << sh;
    return sh > -64 ? (m >> -sh) : 0ull;
}

// BF16 bits of the integer byte (0..255) of v at byte position `byte`.
__device__ __forceinline__ uint32_t byte_bf16(uint32_t v, int byte)
{
    return __float_as_uint(static_cast<float>((v >> (8 * byte)) & 0xFFu)) >> 16;
}

// A batch whose A row (B column) exponents span more than this many binades
// is multiplied on the FP64 matrix cores instead: its smallest entries would
// keep fewer than 64 - 1 - kSpreadBits = 43 bits of the slices.
constexpr int kSpreadBits = 20;

struct SplitProduct {
    struct Lane {
        int a_phys;      // A tile bit this lane supplies (rows 0..7 only)
        bool a_live;     // lane's A row < 8
        int b_phys;      // B tile bit
        int h;           // output column half: d = 2t + h
        int out_row0;    // first of the four C rows this lane's outputs cover
        int out_col;     // C column of this lane's outputs
        bool out_live;   // lane holds C outputs (rows 0..7, h = 0) after the halves combine
        LaneRoles f64;   // roles in the FP64 fallback (v_mfma_f64_4x4x4f64)
        int f64_lane[4]; // lane whose FP64 entry is this lane's output v
        __device__ Lane(int lane) : f64(lane)
        {
            const int i = lane & 15, k = lane >> 4;
            a_live = i < 8;
            a_phys = (i & 7) * 4 + k;
            b_phys = (lane & 7) * 4 + k;
            h = (lane >> 3) & 1;
            out_row0 = 4 * (lane >> 4);
            out_col = lane & 7;
            out_live = lane < 32 && h == 0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = (out_row0 + v) & 7, c = out_col;
                const int b = ((r >> 2) << 1) | (c >> 2);
                f64_lane[v] = (c & 3) + 4 * b + 16 * (r & 3);
            }
        }
    };
    struct Acc {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        double f64 = 0.0;   // FP64-fallback partial, v_mfma_f64_4x4x4f64 layout
    };
    static constexpr unsigned long long kMfmaPerPair = 8;
    // No register cap: held to 4 waves per SIMD this product spills about 240
    // VGPRs to scratch, and that build gave wrong values; unconstrained it
    // needs about 207 VGPRs (2 waves) and spills nothing.
    static constexpr int kOccupancy = 1;

    // Lane's four outputs with the FP64-fallback partials added (every lane
    // takes part in the shuffles).
    __device__ static void finish(const Lane &ln, const Acc &acc, double out[4])
    {
#pragma unroll
        for (int v = 0; v < 4; ++v)
            out[v] = acc.v[v] + __shfl(acc.f64, ln.f64_lane[v], kWave);
    }

    __device__ static void store(const TileTask &g, const Lane &ln, int c_local,
                                 unsigned long long cm, long long co, const Acc &acc)
    {
        const unsigned long long m = read_lane64(cm, c_local);
        const long long off = static_cast<long long>(read_lane64(static_cast<unsigned long long>(co), c_local));
        double out[4];
        finish(ln, acc, out);
        if (!ln.out_live)
            return;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int bit = (ln.out_row0 + v) * 8 + ln.out_col;
            if ((m >> bit) & 1ull)
                g.c_val[off + __popcll(m & ((1ull << bit) - 1ull))] = out[v];
        }
    }

    __device__ static void dump(const Lane &ln, const Acc &acc, double *slot64)
    {
        double out[4];
        finish(ln, acc, out);
        if (!ln.out_live)
            return;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            slot64[(ln.out_row0 + v) * 8 + ln.out_col] = out[v];
    }

    // Smallest biased exponent of the nonzero entries (0x7FF when none).
    __device__ static int min_exp(double x)
    {
        const int e = align_exp(x);
        return e ? e : 0x7FF;
    }

    // Pairs [p0, p0 + n) of one C tile, n <= kSplitBatch, added into acc.
    __device__ static void batch(const TileTask &g, const Lane &ln, const double *a_vals,
                                 const int4 *pm, int p0, int n, Acc &acc)
    {
        double av[kSplitBatch], bv[kSplitBatch];
        int ea = 0, eb = 0, ea_lo = 0x7FF, eb_lo = 0x7FF;
#pragma unroll
        for (int i = 0; i < kSplitBatch; ++i) {
            av[i] = 0.0;
            bv[i] = 0.0;
            if (i < n) {
                const int4 m = pm[p0 + i];
                if (ln.a_live)
                    av[i] = tile_entry(a_vals, static_cast<uint32_t>(m.x),
                                       static_cast<uint32_t>(m.y), ln.a_phys);
                bv[i] = tile_entry(g.b_vals, static_cast<uint32_t>(m.z),
                                   static_cast<uint32_t>(m.w), ln.b_phys);
            }
            ea = max(ea, align_exp(av[i]));
            eb = max(eb, align_exp(bv[i]));
            ea_lo = min(ea_lo, min_exp(av[i]));
            eb_lo = min(eb_lo, min_exp(bv[i]));
        }
        // Row (column) maxima and minima over the four k of the tile: lanes
        // x, x+16, x+32, x+48 hold one row (column).
        ea = max(ea, __shfl_xor(ea, 16, kWave));
        ea = max(ea, __shfl_xor(ea, 32, kWave));
        eb = max(eb, __shfl_xor(eb, 16, kWave));
        eb = max(eb, __shfl_xor(eb, 32, kWave));
        ea_lo = min(ea_lo, __shfl_xor(ea_lo, 16, kWave));
        ea_lo = min(ea_lo, __shfl_xor(ea_lo, 32, kWave));
        eb_lo = min(eb_lo, __shfl_xor(eb_lo, 16, kWave));
        eb_lo = min(eb_lo, __shfl_xor(eb_lo, 32, kWave));
        const bool wide = (ea_lo != 0x7FF && ea - ea_lo > kSpreadBits) ||
                          (eb_lo != 0x7FF && eb - eb_lo > kSpreadBits);
        if (__ballot(wide)) {
            // Too wide a spread for the slices: these pairs on the FP64 cores.
#pragma unroll
            for (int i = 0; i < kSplitBatch; ++i) {
                if (i >= n)
                    break;
                const int4 m = pm[p0 + i];
                const double a = tile_entry(a_vals, static_cast<uint32_t>(m.x),
                                            static_cast<uint32_t>(m.y), ln.f64.a_phys);
                const double b = tile_entry(g.b_vals, static_cast<uint32_t>(m.z),
                                            static_cast<uint32_t>(m.w), ln.f64.b_phys);
                acc.f64 = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc.f64, 0, 0, 0);
            }
            return;
        }
        ea = max(ea, 1);
        eb = max(eb, 1);
        f32x4 acc_t[8];
#pragma unroll
        for (int t = 0; t < 8; ++t)
            acc_t[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < kSplitBatch; ++i) {
            if (i >= n)
                break;
            // A slices in element order e = 0..7 (most significant first).
            const unsigned long long fa = fixed_fraction(av[i], ea);
            const uint32_t ahi = static_cast<uint32_t>(fa >> 32), alo = static_cast<uint32_t>(fa);
            const uint32_t asg = av[i] < 0.0 ? 0x80008000u : 0u;
            u32x4 aw;
            aw[0] = (byte_bf16(ahi, 3) | (byte_bf16(ahi, 2) << 16)) | asg;
            aw[1] = (byte_bf16(ahi, 1) | (byte_bf16(ahi, 0) << 16)) | asg;
            aw[2] = (byte_bf16(alo, 3) | (byte_bf16(alo, 2) << 16)) | asg;
            aw[3] = (byte_bf16(alo, 1) | (byte_bf16(alo, 0) << 16)) | asg;
            // B slices q0..q7 laid out so that instruction t reads
            // qb_{2t+h-e} for element e from words 8-t..11-t.
            const unsigned long long fb = fixed_fraction(bv[i], eb);
            const uint32_t bhi = static_cast<uint32_t>(fb >> 32), blo = static_cast<uint32_t>(fb);
            const uint32_t bsg = bv[i] < 0.0 ? 0x80008000u : 0u;
            const uint32_t q0 = byte_bf16(bhi, 3), q1 = byte_bf16(bhi, 2), q2 = byte_bf16(bhi, 1),
                           q3 = byte_bf16(bhi, 0), q4 = byte_bf16(blo, 3), q5 = byte_bf16(blo, 2),
                           q6 = byte_bf16(blo, 1), q7 = byte_bf16(blo, 0);
            uint32_t wv[12];
#pragma unroll
            for (int x = 0; x < 12; ++x)
                wv[x] = 0u;
            if (ln.h == 0) {
                wv[4] = (q7 << 16) | bsg;
                wv[5] = (q6 | (q5 << 16)) | bsg;
                wv[6] = (q4 | (q3 << 16)) | bsg;
                wv[7] = (q2 | (q1 << 16)) | bsg;
                wv[8] = q0 | bsg;
            } else {
                wv[5] = (q7 | (q6 << 16)) | bsg;
                wv[6] = (q5 | (q4 << 16)) | bsg;
                wv[7] = (q3 | (q2 << 16)) | bsg;
                wv[8] = (q1 | (q0 << 16)) | bsg;
            }
            const bf16x8 aop = __builtin_bit_cast(bf16x8, aw);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                u32x4 bw;
                bw[0] = wv[8 - t];
                bw[1] = wv[9 - t];
                bw[2] = wv[10 - t];
                bw[3] = wv[11 - t];
                acc_t[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    aop, __builtin_bit_cast(bf16x8, bw), acc_t[t], 0, 0, 0);
            }
        }
        // Recombination in FP64: lane's outputs are rows out_row0 + v, column
        // (lane & 15) -> C column lane & 7 at d = 2t + h.
        const int lane = lane_id();
        const int ecol = __shfl(eb, lane & 7, kWave) - 1022;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int erow = __shfl(ea, (ln.out_row0 + v) & 7, kWave) - 1022;
            double sum = 0.0;
            double w = ln.h ? 0x1p-24 : 0x1p-16;   // 2^(-8(d+2)) at t = 0
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                sum += static_cast<double>(acc_t[t][v]) * w;
                w *= 0x1p-16;
            }
            sum = ldexp(sum, erow + ecol);
            sum += __shfl_xor(sum, 8, kWave);   // the d = 2t + 1 half
            acc.v[v] += sum;
        }
    }

    __device__ static void run(const TileTask &g, const Lane &ln, const double *a_vals,
                               const int4 *pm, const unsigned char *pc, int n, int &cur, Acc &acc,
                               unsigned long long cm, long long co)
    {
        int p = 0;
        while (p < n) {
            const int t = pc[p];
            if (t != cur) {
                if (cur >= 0)
                    store(g, ln, cur, cm, co, acc);
                cur = t;
                acc = Acc{};
            }
            int q = p + 1;
            while (q < n && q - p < kSplitBatch && pc[q] == t)
                ++q;
            batch(g, ln, a_vals, pm, p, q - p, acc);
            p = q;
        }
    }
};

// The C tiles of tile row `row` ([c0, c1)), up to 64 at a time: their B
// column entries are matched against A row I's K list and the pairs run.
// kStaged: A row I's keys, {mask, offset} and payload are in LDS (offsets
// relative to the row's first value); otherwise they are read from HBM.
template <class P, bool kStaged>
__device__ __forceinline__ unsigned long long tile_row(const TileTask &g,
                                                       const typename P::Lane &ln, int64_t c0,
                                                       int64_t c1, int a0, int na, const int *keys,
                                                       const int2 *am, const double *avals,
                                                       int4 *pm, unsigned char *pc, int *start,
                                                       int *cbeg, bool heavy_on)
{
    const int lane = lane_id();
    const double *a_vals = (kStaged && kAVals > 0) ? avals : g.a_vals;
    unsigned long long done = 0;
    for (int64_t cc = c0; cc < c1; cc += kWave) {
        const int nc = static_cast<int>(min<int64_t>(kWave, c1 - cc));
        // Lane c of the chunk holds C tile cc + c: mask, offset, B column run.
        unsigned long long cm = 0;
        long long co = 0;
        int nb = 0, cb = 0;
        if (lane < nc) {
            const int j = g.c_col[cc + lane];
            cm = g.c_mask[cc + lane];
            co = g.c_off[cc + lane];
            cb = g.b_cptr[j];
            nb = g.b_cptr[j + 1] - cb;
            // Heavy tiles are computed by the chunk kernels instead.
            const int64_t t = cc + lane;
            if (heavy_on && ((g.heavy[t >> 5] >> (t & 31)) & 1u))
                nb = 0;
        }
        const int incl = wave_inclusive_scan(nb);
        const int total = __shfl(incl, kWave - 1, kWave);
        start[lane] = incl - nb;
        cbeg[lane] = cb;
        if (lane == 0)
            start[kWave] = total;
        wave_sync();
        int npairs = 0;
        int cur = -1;
        typename P::Acc acc;
        const int rounds = (total + kWave - 1) / kWave;
        for (int r0 = 0; r0 < rounds; r0 += kEnumRounds) {
            // The K of kEnumRounds rounds of B column entries, loaded at once.
            int key[kEnumRounds], bix[kEnumRounds], cl[kEnumRounds];
#pragma unroll
            for (int i = 0; i < kEnumRounds; ++i) {
                const int q = (r0 + i) * kWave + lane;
                key[i] = -1;
                bix[i] = 0;
                cl[i] = 0;
                if (r0 + i < rounds && q < total) {
                    const int c = upper_index(start, nc, q);
                    cl[i] = c;
                    bix[i] = cbeg[c] + (q - start[c]);
                    key[i] = g.b_cent[bix[i]].x;
                }
            }
#pragma unroll
            for (int i = 0; i < kEnumRounds; ++i) {
                if (r0 + i >= rounds)
                    break;
                int ai = -1;
                if (key[i] >= 0) {
                    int l = 0, h = na;
                    while (l < h) {
                        const int mid = (l + h) >> 1;
                        const int k = kStaged ? keys[mid] : g.a_meta[a0 + mid].x;
                        if (k < key[i])
                            l = mid + 1;
                        else
                            h = mid;
                    }
                    if (l < na && (kStaged ? keys[l] : g.a_meta[a0 + l].x) == key[i])
                        ai = l;
                }
                const unsigned long long match = __ballot(ai >= 0);
                if (ai >= 0) {
                    int amask, aoff;
                    if (kStaged) {
                        const int2 m = am[ai];
                        amask = m.x;
                        aoff = m.y;
                    } else {
                        const int4 m = g.a_meta[a0 + ai];
                        amask = m.y;
                        aoff = m.w;
                    }
                    const int slot = npairs + __popcll(match & ((1ull << lane) - 1ull));
                    pm[slot] = make_int4(amask, aoff, bix[i], 0);
                    pc[slot] = static_cast<unsigned char>(cl[i]);
                }
                npairs += __popcll(match);
                // Run the buffered pairs once another round might not fit.
                const bool last = r0 + i + 1 >= rounds;
                if (npairs + kWave > kPairCap || last) {
                    wave_sync();
                    // The matched B entries' {mask, offset}: one round of
                    // independent loads for the whole buffer.
                    for (int p = lane; p < npairs; p += kWave) {
                        const int4 b = g.b_cent[pm[p].z];
                        pm[p].z = b.y;
                        pm[p].w = b.w;
                    }
                    wave_sync();
                    P::run(g, ln, a_vals, pm, pc, npairs, cur, acc, cm, co);
                    done += static_cast<unsigned long long>(npairs);
                    npairs = 0;
                    wave_sync();
                }
            }
        }
        if (cur >= 0)
            P::store(g, ln, cur, cm, co, acc);
        wave_sync();
    }
    return done;
}

template <class P>
__global__ __launch_bounds__(kTThreads) __attribute__((amdgpu_waves_per_eu(P::kOccupancy)))
void tile_numeric_kernel(TileTask g)
{
    __shared__ int s_keys[kTWaves][kAKeys];
    __shared__ int2 s_am[kTWaves][kAKeys];              // {mask, offset | dense} from the row's first value
    __shared__ double s_av[kTWaves][kAVals > 0 ? kAVals : 1];   // A row's payload
    __shared__ int4 s_pm[kTWaves][kPairCap];            // {A mask, A offset, B mask, B offset}
    __shared__ unsigned char s_pc[kTWaves][kPairCap];   // C tile (chunk-local) of each pair
    __shared__ int s_start[kTWaves][kWave + 1];
    __shared__ int s_cb[kTWaves][kWave];
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    int *keys = s_keys[w];
    int2 *am = s_am[w];
    double *av = s_av[w];
    const typename P::Lane ln(lane);
    unsigned long long done = 0;   // pairs this wave multiplied
    const bool heavy_on = g.heavy && g.tail->enabled;
    const int home = blockIdx.x % kXcds;
    for (int probe = 0; probe < kXcds; ++probe) {
        const int x = (home + probe) % kXcds;
        // Queue x hands out chunks x, x + 8, x + 16, ... of kRowGrab rows, so
        // that the XCDs work through one window of consecutive rows
        // together (faster than a contiguous eighth of the rows each).
        for (;;) {
            int r = 0;
            if (lane == 0)
                r = atomicAdd(&g.queue[x], 1);
            const int64_t base64 =
                (static_cast<int64_t>(read_lane(r, 0)) * kXcds + x) * kRowGrab;
            if (base64 >= g.rows)
                break;
            const int base = static_cast<int>(base64);
            const int rend = min(base + kRowGrab, g.rows);
            for (int row = base; row < rend; ++row) {
                const int64_t c0 = g.c_ptr[row], c1 = g.c_ptr[row + 1];
                if (c0 == c1)
                    continue;
                const int a0 = g.a_trp[row];
                const int na = g.a_trp[row + 1] - a0;
                const int v0 = g.a_row_off[row];
                const int nv = g.a_row_off[row + 1] - v0;
                // Rows below the fill threshold belong to the LDS-accumulator kernels.
                if (static_cast<float>(nv) < g.fill * static_cast<float>(na))
                    continue;
                if (na <= kAKeys && (kAVals == 0 || nv <= kAVals)) {
                    // Offsets relative to the row's first value when the
                    // payload is staged too.
                    const uint32_t base = kAVals > 0 ? static_cast<uint32_t>(v0) : 0u;
                    for (int i = lane; i < na; i += kWave) {
                        const int4 m = g.a_meta[a0 + i];
                        keys[i] = m.x;
                        const uint32_t ow = static_cast<uint32_t>(m.w);
                        am[i] = make_int2(m.y, static_cast<int>(((ow & ~kDense) - base) | (ow & kDense)));
                    }
                    if (kAVals > 0)
                        for (int i = lane; i < nv; i += kWave)
                            av[i] = g.a_vals[v0 + i];
                    wave_sync();
                    done += tile_row<P, true>(g, ln, c0, c1, a0, na, keys, am, av, s_pm[w], s_pc[w],
                                              s_start[w], s_cb[w], heavy_on);
                } else {
                    done += tile_row<P, false>(g, ln, c0, c1, a0, na, keys, am, av, s_pm[w],
                                               s_pc[w], s_start[w], s_cb[w], heavy_on);
                }
            }
        }
    }
    if (lane == 0 && done)
        atomicAdd(g.mfma_pairs, done);
}

// Chunks of a task whose merge takes `scans` steps with `matches` matched
// pairs: the reference's FP32 admission (cost = scans + 8 matches, strictly
// above the threshold) and chunk count (ceil(cost / target), clamped to
// [2, max_chunks] and to at most one chunk per scan step); 0 = not admitted
// (dmma_tail_chunk_count_from_counts, src/dmma_spgemm.h:2628-2647).
__device__ __forceinline__ int tail_chunks(int scans, int matches, float threshold, float target,
                                           int max_chunks)
{
    const float work = 1.0f * static_cast<float>(scans) + 8.0f * static_cast<float>(matches);
    if (!(work > threshold) || !(target > 0.0f) || scans <= 1 || max_chunks <= 1)
        return 0;
    const float want = ceilf(work / target);
    int c = want >= static_cast<float>(max_chunks) ? max_chunks : static_cast<int>(want);
    c = max(c, 2);
    c = min(c, scans);
    return c >= 2 ? c : 0;
}

struct TailArgs {
    float threshold, target;
    int max_chunks;
    double max_heavy_fraction;
    int force;
    unsigned int capacity;
    unsigned long long workspace_limit;   // bytes
    long long tiles;                      // C tiles of the call
    TailRecord *__restrict__ records;
    TailState *__restrict__ state;
    uint32_t *__restrict__ heavy;
    int *__restrict__ chunk_base;         // capacity + 1
    double *__restrict__ workspace;       // 64 doubles per chunk
};

__global__ void tail_reset_kernel(TailState *st)
{
    if (threadIdx.x == 0) {
        st->count = 0u;
        st->overflow = 0;
        st->chunks = 0ull;
        st->reason = 0;
        st->enabled = 0;
        st->chunk_head = 0;
        st->reduce_head = 0;
    }
}

// One wave per tile row of the tile-pair kernel: the exact merge counts of
// every C tile whose upper bound (all of both lists scanned, every pair of
// the shorter list matched) clears the threshold, and a record for each
// admitted tile.
__global__ __launch_bounds__(kTThreads) void tail_admit_kernel(TileTask g, TailArgs t)
{
    const int row = blockIdx.x * kTWaves + static_cast<int>(threadIdx.x) / kWave;
    const int lane = lane_id();
    if (row >= g.rows)
        return;
    const int64_t c0 = g.c_ptr[row], c1 = g.c_ptr[row + 1];
    const int a0 = g.a_trp[row];
    const int na = g.a_trp[row + 1] - a0;
    if (c0 == c1 || static_cast<float>(g.a_row_off[row + 1] - g.a_row_off[row]) <
                        g.fill * static_cast<float>(na))
        return;
    for (int64_t cc = c0; cc < c1; cc += kWave) {
        const int64_t tile = cc + lane;
        int cb = 0, nb = 0;
        bool cand = false;
        if (tile < c1) {
            const int j = g.c_col[tile];
            cb = g.b_cptr[j];
            nb = g.b_cptr[j + 1] - cb;
            const float ub = 1.0f * static_cast<float>(na + nb) +
                             8.0f * static_cast<float>(min(na, nb));
            cand = ub > t.threshold;
        }
        unsigned long long todo = __ballot(cand);
        while (todo) {
            const int l = __ffsll(static_cast<long long>(todo)) - 1;
            todo &= todo - 1;
            const int tcb = read_lane(cb, l), tnb = read_lane(nb, l);
            // Intersection of A row's K list with B column's K list: lanes
            // take the shorter list, binary search in the longer.
            const bool a_short = na <= tnb;
            const int ns = a_short ? na : tnb, nl = a_short ? tnb : na;
            int matches = 0;
            for (int x0 = 0; x0 < ns; x0 += kWave) {
                const int x = x0 + lane;
                bool hit = false;
                if (x < ns) {
                    const int key = a_short ? g.a_meta[a0 + x].x : g.b_cent[tcb + x].x;
                    int lo = 0, hi = nl;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        const int k = a_short ? g.b_cent[tcb + mid].x : g.a_meta[a0 + mid].x;
                        if (k < key)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    hit = lo < nl &&
                          (a_short ? g.b_cent[tcb + lo].x : g.a_meta[a0 + lo].x) == key;
                }
                matches += __popcll(__ballot(hit));
            }
            const int scans = na + tnb - matches;
            const int chunks = tail_chunks(scans, matches, t.threshold, t.target, t.max_chunks);
            if (lane == 0 && chunks > 0) {
                const unsigned int slot = atomicAdd(&t.state->count, 1u);
                if (slot < t.capacity) {
                    TailRecord r;
                    r.task = cc + l;
                    r.row = row;
                    r.scans = scans;
                    r.matches = matches;
                    r.chunks = chunks;
                    t.records[slot] = r;
                    atomicAdd(&t.state->chunks, static_cast<unsigned long long>(chunks));
                } else {
                    atomicExch(&t.state->overflow, 1);
                }
            }
        }
    }
}

// The gate (one thread): split only when some tile was admitted, the
// records fit, the heavy fraction is at most max_heavy_fraction (unless
// forced) and the chunk workspace fits its limit.  Reason codes: 1 split,
// 4 no heavy tile, 5 record capacity exceeded, 6 heavy fraction exceeded,
// 7 workspace limit exceeded.
__global__ void tail_gate_kernel(TailArgs t)
{
    if (threadIdx.x != 0)
        return;
    TailState *st = t.state;
    int reason;
    if (st->overflow || st->count > t.capacity)
        reason = 5;
    else if (st->count == 0u)
        reason = 4;
    else if (!t.force && static_cast<double>(st->count) / static_cast<double>(t.tiles) >
                             t.max_heavy_fraction)
        reason = 6;
    else if (st->chunks * 64ull * sizeof(double) > t.workspace_limit)
        reason = 7;
    else
        reason = 1;
    st->reason = reason;
    st->enabled = reason == 1 ? 1 : 0;
}

__global__ void tail_mark_kernel(TailArgs t)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    const TailState *st = t.state;
    const unsigned int n = min(st->count, t.capacity);
    if (i == 0)
        t.chunk_base[0] = 0;
    if (i >= n)
        return;
    const TailRecord r = t.records[i];
    if (st->enabled)
        atomicOr(&t.heavy[r.task >> 5], 1u << (r.task & 31));
}

// Exclusive prefix of the records' chunk counts (records in slot order; a
// tile's chunks are contiguous).  Single workgroup, sequential over the
// workgroup's partial sums: the record list is short.
__global__ __launch_bounds__(kTThreads) void tail_chunk_base_kernel(TailArgs t)
{
    __shared__ int s_sum[kTThreads];
    const TailState *st = t.state;
    const int n = static_cast<int>(min(st->count, t.capacity));
    const int per = (n + kTThreads - 1) / kTThreads;
    const int b = threadIdx.x * per, e = min(n, b + per);
    int sum = 0;
    for (int i = b; i < e; ++i)
        sum += t.records[i].chunks;
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int i = 0; i < kTThreads; ++i) {
            const int v = s_sum[i];
            s_sum[i] = run;
            run += v;
        }
    }
    __syncthreads();
    int run = s_sum[threadIdx.x];
    for (int i = b; i < e; ++i) {
        t.chunk_base[i] = run;
        run += t.records[i].chunks;
    }
    if (e == n && b < e)
        t.chunk_base[n] = run;
    if (n == 0 && threadIdx.x == 0)
        t.chunk_base[0] = 0;
}

// Persistent waves over the chunks: chunk j of record r covers B column
// entries [cb + nb j / chunks, cb + nb (j+1) / chunks) of its tile; the
// pairs found there are multiplied like the tile kernel's and the lane's
// partial entries go to the chunk's 64-slot workspace.
template <class P>
__global__ __launch_bounds__(kTThreads) void tail_chunk_kernel(TileTask g, TailArgs t)
{
    __shared__ int4 s_pm[kTWaves][kPairCap];
    __shared__ unsigned char s_pc[kTWaves][kPairCap];
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    int4 *pm = s_pm[w];
    unsigned char *pc = s_pc[w];
    const typename P::Lane ln(lane);
    const TailState *st = t.state;
    if (!st->enabled)
        return;
    const int n = static_cast<int>(min(st->count, t.capacity));
    const int total = t.chunk_base[n];
    for (;;) {
        int q = 0;
        if (lane == 0)
            q = atomicAdd(&t.state->chunk_head, 1);
        q = read_lane(q, 0);
        if (q >= total)
            break;
        const int r = upper_index(t.chunk_base, n, q);
        const TailRecord rec = t.records[r];
        const int j = q - t.chunk_base[r];
        const int row = rec.row;
        const int a0 = g.a_trp[row];
        const int na = g.a_trp[row + 1] - a0;
        const int jj = g.c_col[rec.task];
        const int cb = g.b_cptr[jj];
        const int nb = g.b_cptr[jj + 1] - cb;
        const int e0 = cb + static_cast<int>(static_cast<long long>(nb) * j / rec.chunks);
        const int e1 = cb + static_cast<int>(static_cast<long long>(nb) * (j + 1) / rec.chunks);
        int npairs = 0, cur = -1;
        typename P::Acc acc;
        for (int p0 = e0; p0 < e1; p0 += kWave) {
            const int p = p0 + lane;
            int ai = -1;
            int4 be = make_int4(0, 0, 0, 0);
            if (p < e1) {
                be = g.b_cent[p];
                int lo = 0, hi = na;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (g.a_meta[a0 + mid].x < be.x)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                if (lo < na && g.a_meta[a0 + lo].x == be.x)
                    ai = a0 + lo;
            }
            const unsigned long long match = __ballot(ai >= 0);
            if (ai >= 0) {
                const int4 am = g.a_meta[ai];
                const int slot = npairs + __popcll(match & ((1ull << lane) - 1ull));
                pm[slot] = make_int4(am.y, am.w, be.y, be.w);
                pc[slot] = 0;
            }
            npairs += __popcll(match);
            if (npairs + kWave > kPairCap || p0 + kWave >= e1) {
                wave_sync();
                P::run(g, ln, g.a_vals, pm, pc, npairs, cur, acc, 0ull, 0ll);
                npairs = 0;
                wave_sync();
            }
        }
        P::dump(ln, acc, t.workspace + static_cast<size_t>(q) * 64);
    }
}

// One wave per admitted tile: its chunks' partials summed in chunk order,
// stored at the tile's entries.
__global__ __launch_bounds__(kTThreads) void tail_reduce_kernel(TileTask g, TailArgs t)
{
    const TailState *st = t.state;
    if (!st->enabled)
        return;
    const int n = static_cast<int>(min(st->count, t.capacity));
    const int lane = lane_id();
    for (;;) {
        int r = 0;
        if (lane == 0)
            r = atomicAdd(&t.state->reduce_head, 1);
        r = read_lane(r, 0);
        if (r >= n)
            break;
        const TailRecord rec = t.records[r];
        const double *ws = t.workspace + static_cast<size_t>(t.chunk_base[r]) * 64;
        double sum = 0.0;
        for (int j = 0; j < rec.chunks; ++j)
            sum += ws[static_cast<size_t>(j) * 64 + lane];
        const unsigned long long m = g.c_mask[rec.task];
        if ((m >> lane) & 1ull)
            g.c_val[g.c_off[rec.task] + __popcll(m & ((1ull << lane) - 1ull))] = sum;
    }
}

__global__ void zero_u32_kernel(uint32_t *p, int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0u;
}

__global__ void zero_queue_kernel(int *q, int n, unsigned long long *pairs)
{
    if (static_cast<int>(threadIdx.x) < n)
        q[threadIdx.x] = 0;
    if (threadIdx.x == 0)
        *pairs = 0ull;
}

} // namespace

TileColumns b_columns(const DevTiles &b)
{
    TileColumns v;
    if (b.transpose_of) {
        // B tile (K,J) column-major is A tile (J,K) row-major, bit for bit:
        // B's column J is A's tile row J.
        v.ptr = b.transpose_of->tile_row_ptr.as<int>();
        v.ent = b.transpose_of->meta.as<int4>();
        v.vals = b.transpose_of->values.as<double>();
    } else {
        v.ptr = b.col_ptr.as<int>();
        v.ent = b.col_ent.as<int4>();
        v.vals = b.values.as<double>();
    }
    return v;
}

// count[0]: tile rows routed to the tile-pair kernel at `fill`; count[1]:
// tile rows holding any tile.  A grid-stride pass of at most kCountBlocks
// workgroups with one atomic pair per workgroup: an atomic pair per wave on
// the same two counters serialised at the L2 (0.3 ms for 2.6 M rows).
constexpr int kCountThreads = 256;
constexpr int kCountBlocks = 1024;
__global__ __launch_bounds__(kCountThreads) void count_tile_rows_kernel(
    int rows, const int *__restrict__ trp, const int *__restrict__ row_off, float fill,
    unsigned long long *__restrict__ count)
{
    __shared__ unsigned int s_m[kCountThreads / kWave], s_n[kCountThreads / kWave];
    unsigned int m = 0, n = 0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        const int t = trp[r + 1] - trp[r];
        const bool any = t > 0;
        const bool take =
            any && static_cast<float>(row_off[r + 1] - row_off[r]) >= fill * static_cast<float>(t);
        m += __popcll(__ballot(take));
        n += __popcll(__ballot(any));
    }
    // Lane 0 holds the wave's smallest row, so it takes part in every ballot
    // of its wave and its sums are the wave's totals.
    if (lane_id() == 0) {
        s_m[threadIdx.x / kWave] = m;
        s_n[threadIdx.x / kWave] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long bm = 0, bn = 0;
#pragma unroll
        for (int w = 0; w < kCountThreads / kWave; ++w) {
            bm += s_m[w];
            bn += s_n[w];
        }
        if (bm)
            atomicAdd(count, bm);
        if (bn)
            atomicAdd(count + 1, bn);
    }
}

int count_tile_rows(const DevTiles &a, float fill, DevBuf *scratch, int64_t *tile_rows,
                    int64_t *nonempty_rows)
{
    hipStream_t s = ctx().stream;
    RTT_TRY(scratch->ensure(2 * sizeof(unsigned long long)));
    unsigned long long *d = scratch->as<unsigned long long>();
    hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(4), 0, s, reinterpret_cast<int *>(d), 4,
                       d + 1);
    const int R = a.tile_row_count;
    if (R > 0)
        hipLaunchKernelGGL(count_tile_rows_kernel,
                           dim3(std::min<unsigned int>(kCountBlocks, blocks_for(R, kCountThreads))),
                           dim3(kCountThreads), 0, s, R, a.tile_row_ptr.as<int>(),
                           a.row_payload_off.as<int>(), fill, d);
    RTT_HIP(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    RTT_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    RTT_HIP(hipStreamSynchronize(s));
    *tile_rows = static_cast<int64_t>(h[0]);
    *nonempty_rows = static_cast<int64_t>(h[1]);
    return RTT_OK;
}

int tile_numeric(const DevTiles &a, const TileColumns &bc, const DevC &c, const TileOpts &opt,
                 SpgemmWork *w, hipStream_t s)
{
    if (c.tilem == 0 || c.num_tiles == 0)
        return RTT_OK;
    const int precision = opt.precision;
    if (precision != 0 && precision != 1)
        return fail(RTT_ERR_INVALID, "unknown tile numeric precision %d", precision);
    RTT_TRY(w->tile_queue.ensure(kXcds * sizeof(int)));
    RTT_TRY(w->tile_pairs.ensure(sizeof(unsigned long long)));
    TileTask g;
    g.rows = static_cast<int>(c.tilem);
    g.a_trp = a.tile_row_ptr.as<int>();
    g.a_meta = a.meta.as<int4>();
    g.a_vals = a.values.as<double>();
    g.b_cptr = bc.ptr;
    g.b_cent = bc.ent;
    g.b_vals = bc.vals;
    g.c_ptr = c.tile_ptr.as<int64_t>();
    g.c_col = c.tile_col.as<int>();
    g.c_mask = c.tile_mask.as<unsigned long long>();
    g.c_off = c.tile_nnz.as<int64_t>();
    g.c_val = c.tile_csr_val.as<double>();
    g.a_row_off = a.row_payload_off.as<int>();
    g.fill = opt.fill;
    g.queue = w->tile_queue.as<int>();
    g.mfma_pairs = w->tile_pairs.as<unsigned long long>();
    g.heavy = nullptr;
    g.tail = nullptr;
    hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(kXcds), 0, s, g.queue, kXcds,
                       g.mfma_pairs);
    const int cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
    TailArgs t{};
    if (opt.tail) {
        // Record capacity as the reference sizes it: one past the admissible
        // heavy count (floor(tiles * max_fraction)), or every tile when the
        // split is forced, and never more than tail_record_capacity
        // (src/dmma_spgemm.h:7000-7016).
        const int64_t T = c.num_tiles;
        const int64_t admissible = static_cast<int64_t>(
            std::floor(static_cast<double>(T) * opt.max_heavy_fraction));
        const int64_t cap = std::max<int64_t>(
            1, std::min<int64_t>(opt.force ? T : admissible + 1, opt.record_capacity));
        const unsigned long long limit = 1ull << 30;   // src/main.cu:89
        const unsigned long long ws_bytes = std::min<unsigned long long>(
            limit, static_cast<unsigned long long>(cap) * opt.max_chunks * 64 * sizeof(double));
        RTT_TRY(w->tail_records.ensure(static_cast<size_t>(cap) * sizeof(TailRecord)));
        RTT_TRY(w->tail_state.ensure(sizeof(TailState)));
        RTT_TRY(w->tail_chunk_base.ensure((static_cast<size_t>(cap) + 1) * sizeof(int)));
        RTT_TRY(w->tail_heavy.ensure(static_cast<size_t>((T + 31) / 32) * sizeof(uint32_t)));
        RTT_TRY(w->tail_workspace.ensure(ws_bytes));
        t.threshold = opt.threshold;
        t.target = opt.target;
        t.max_chunks = opt.max_chunks;
        t.max_heavy_fraction = opt.max_heavy_fraction;
        t.force = opt.force;
        t.capacity = static_cast<unsigned int>(cap);
        t.workspace_limit = ws_bytes;
        t.tiles = T;
        t.records = w->tail_records.as<TailRecord>();
        t.state = w->tail_state.as<TailState>();
        t.heavy = w->tail_heavy.as<uint32_t>();
        t.chunk_base = w->tail_chunk_base.as<int>();
        t.workspace = w->tail_workspace.as<double>();
        hipLaunchKernelGGL(tail_reset_kernel, dim3(1), dim3(1), 0, s, t.state);
        hipLaunchKernelGGL(zero_u32_kernel, dim3(blocks_for((T + 31) / 32, 256)), dim3(256), 0, s,
                           t.heavy, (T + 31) / 32);
        hipLaunchKernelGGL(tail_admit_kernel, dim3(blocks_for(g.rows, kTWaves)), dim3(kTThreads),
                           0, s, g, t);
        hipLaunchKernelGGL(tail_gate_kernel, dim3(1), dim3(1), 0, s, t);
        hipLaunchKernelGGL(tail_mark_kernel, dim3(blocks_for(cap, 256)), dim3(256), 0, s, t);
        hipLaunchKernelGGL(tail_chunk_base_kernel, dim3(1), dim3(kTThreads), 0, s, t);
        RTT_HIP(hipGetLastError());
        g.heavy = t.heavy;
        g.tail = t.state;
    }
    // Persistent: enough workgroups to fill every CU several times over.
    const int blocks = static_cast<int>(std::min<int64_t>(
        static_cast<int64_t>(cus) * 8, (static_cast<int64_t>(g.rows) + kTWaves - 1) / kTWaves));
    if (precision == 0)
        hipLaunchKernelGGL(tile_numeric_kernel<Fp64Product>, dim3(std::max(blocks, 1)),
                           dim3(kTThreads), 0, s, g);
    else
        hipLaunchKernelGGL(tile_numeric_kernel<SplitProduct>, dim3(std::max(blocks, 1)),
                           dim3(kTThreads), 0, s, g);
    RTT_HIP(hipGetLastError());
    if (opt.tail) {
        if (precision == 0)
            hipLaunchKernelGGL(tail_chunk_kernel<Fp64Product>, dim3(cus * 4), dim3(kTThreads), 0,
                               s, g, t);
        else
            hipLaunchKernelGGL(tail_chunk_kernel<SplitProduct>, dim3(cus * 4), dim3(kTThreads), 0,
                               s, g, t);
        hipLaunchKernelGGL(tail_reduce_kernel, dim3(cus * 2), dim3(kTThreads), 0, s, g, t);
        RTT_HIP(hipGetLastError());
    }
    return RTT_OK;
}

int tile_values_noalloc(const DevTiles &a, const TileColumns &bc, const DevC &c, int precision,
                        void *scratch, size_t scratch_bytes, double *c_values, hipStream_t s)
{
    // Scratch: the kXcds row-queue heads, then the pair counter.
    const size_t need = kXcds * sizeof(int) + sizeof(unsigned long long);
    if (need > scratch_bytes)
        return fail(RTT_ERR_INVALID, "tile values need %zu scratch bytes, %zu available", need,
                    scratch_bytes);
    if (c.tilem > INT_MAX)
        return fail(RTT_ERR_INVALID, "C has %lld tile rows: beyond the row queues' range",
                    (long long)c.tilem);
    if (precision != 0 && precision != 1)
        return fail(RTT_ERR_INVALID, "unknown tile numeric precision %d", precision);
    if (c.tilem == 0 || c.num_tiles == 0)
        return RTT_OK;
    TileTask g;
    g.rows = static_cast<int>(c.tilem);
    g.a_trp = a.tile_row_ptr.as<int>();
    g.a_meta = a.meta.as<int4>();
    g.a_vals = a.values.as<double>();
    g.b_cptr = bc.ptr;
    g.b_cent = bc.ent;
    g.b_vals = bc.vals;
    g.c_ptr = c.tile_ptr.as<int64_t>();
    g.c_col = c.tile_col.as<int>();
    g.c_mask = c.tile_mask.as<unsigned long long>();
    g.c_off = c.tile_nnz.as<int64_t>();
    g.c_val = c_values;
    g.a_row_off = a.row_payload_off.as<int>();
    g.fill = 0.f;
    g.queue = static_cast<int *>(scratch);
    g.mfma_pairs = reinterpret_cast<unsigned long long *>(static_cast<char *>(scratch) +
                                                          kXcds * sizeof(int));
    g.heavy = nullptr;
    g.tail = nullptr;
    // Cleared by a kernel (the scratch keeps the previous call's contents).
    hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(kXcds), 0, s, g.queue, kXcds,
                       g.mfma_pairs);
    const int cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
    const int blocks = static_cast<int>(std::min<int64_t>(
        static_cast<int64_t>(cus) * 8, (static_cast<int64_t>(g.rows) + kTWaves - 1) / kTWaves));
    if (precision == 0)
        hipLaunchKernelGGL(tile_numeric_kernel<Fp64Product>, dim3(std::max(blocks, 1)),
                           dim3(kTThreads), 0, s, g);
    else
        hipLaunchKernelGGL(tile_numeric_kernel<SplitProduct>, dim3(std::max(blocks, 1)),
                           dim3(kTThreads), 0, s, g);
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

int tail_stats(SpgemmWork *w, int *reason, unsigned long long *records,
               unsigned long long *chunks)
{
    TailState st{};
    RTT_HIP(hipMemcpy(&st, w->tail_state.ptr, sizeof(st), hipMemcpyDeviceToHost));
    *reason = st.reason;
    *records = st.count;
    *chunks = st.enabled ? st.chunks : 0ull;
    return RTT_OK;
}

int tail_records(SpgemmWork *w, int max_records, long long *task, int *scans, int *matches,
                 int *chunks, int *count)
{
    TailState st{};
    RTT_HIP(hipMemcpy(&st, w->tail_state.ptr, sizeof(st), hipMemcpyDeviceToHost));
    const int n = static_cast<int>(std::min<unsigned long long>(
        std::min<unsigned long long>(st.count, w->tail_records.bytes / sizeof(TailRecord)),
        static_cast<unsigned long long>(max_records)));
    std::vector<TailRecord> recs(static_cast<size_t>(n));
    if (n > 0)
        RTT_HIP(hipMemcpy(recs.data(), w->tail_records.ptr, n * sizeof(TailRecord),
                          hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        task[i] = recs[i].task;
        scans[i] = recs[i].scans;
        matches[i] = recs[i].matches;
        chunks[i] = recs[i].chunks;
    }
    *count = n;
    return RTT_OK;
}

} // namespace rtt
