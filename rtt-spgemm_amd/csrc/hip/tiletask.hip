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
// Work unit: one wave per C tile row, taken from per-XCD row queues (each XCD
// starts on its own contiguous eighth of the rows, so neighbouring rows, which
// share B columns, meet in one L2; a wave whose XCD queue is drained takes
// rows from the next XCD's).  The wave stages A row I's tile columns in LDS,
// then, for up to 64 C tiles at a time, enumerates the B column entries of
// all of them (one lane per entry), finds each entry's K among A's columns by
// binary search and compacts the matches, in (C tile, K) order, into an LDS
// pair list.  The pair loop then issues the operand loads of eight pairs at
// once and runs their products into the running sum, storing each C tile's
// entries when its pairs end.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"
#include "internal.hpp"

namespace rtt {

namespace {

constexpr int kTThreads = 256;
constexpr int kTWaves = kTThreads / kWave;
constexpr int kAKeys = 256;     // A tile columns of a row staged in LDS
constexpr int kPairCap = 256;   // pairs a wave buffers before running them
constexpr int kBatch = 8;       // pairs whose operand loads are issued together
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
    int *__restrict__ queue;             // kXcds row-queue heads, zero at launch
    unsigned long long *__restrict__ mfma_pairs;   // pairs multiplied (one MFMA each)
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

// Lane roles of v_mfma_f64_4x4x4f64 (four independent 4x4x4 blocks; block b,
// measured on gfx950): A[b][i][k] in lane i + 4b + 16k, B[b][k][j] in lane
// j + 4b + 16k, D[b][i][j] in lane j + 4b + 16i.  Block b = 2h + v multiplies
// rows 4h..4h+3 of the A tile by columns 4v..4v+3 of the B tile, so D's four
// blocks are the quarters of the 8x8 C tile.
struct LaneRoles {
    int a_phys;   // A tile bit r*4+k this lane supplies
    int b_phys;   // B tile bit c*4+k
    int c_bit;    // C tile bit r*8+c this lane accumulates
    __device__ LaneRoles(int lane)
    {
        const int blk = (lane >> 2) & 3, lo = lane & 3, hi = lane >> 4;
        a_phys = ((blk >> 1) * 4 + lo) * 4 + hi;
        b_phys = ((blk & 1) * 4 + lo) * 4 + hi;
        c_bit = ((blk >> 1) * 4 + hi) * 8 + (blk & 1) * 4 + lo;
    }
};

// Stores lane's entry of the finished C tile `c_local` (0..63 of the chunk,
// whose mask and value offset lanes hold as cm / co).
__device__ __forceinline__ void store_tile(const TileTask &g, int c_local, unsigned long long cm,
                                           long long co, int c_bit, double acc)
{
    const unsigned long long m =
        (static_cast<unsigned long long>(static_cast<uint32_t>(read_lane(static_cast<int>(cm >> 32), c_local))) << 32) |
        static_cast<uint32_t>(read_lane(static_cast<int>(cm), c_local));
    const long long off =
        (static_cast<long long>(read_lane(static_cast<int>(co >> 32), c_local)) << 32) |
        static_cast<uint32_t>(read_lane(static_cast<int>(co), c_local));
    if ((m >> c_bit) & 1ull)
        g.c_val[off + __popcll(m & ((1ull << c_bit) - 1ull))] = acc;
}

__global__ __launch_bounds__(kTThreads) void tile_numeric_f64_kernel(TileTask g)
{
    __shared__ int s_keys[kTWaves][kAKeys];
    __shared__ int4 s_pm[kTWaves][kPairCap];            // {A mask, A offset, B mask, B offset}
    __shared__ unsigned char s_pc[kTWaves][kPairCap];   // C tile (chunk-local) of each pair
    __shared__ int s_start[kTWaves][kWave + 1];
    __shared__ int s_cb[kTWaves][kWave];
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    int *keys = s_keys[w];
    int4 *pm = s_pm[w];
    unsigned char *pc = s_pc[w];
    int *start = s_start[w];
    int *cbeg = s_cb[w];
    const LaneRoles roles(lane);
    unsigned long long done = 0;   // pairs this wave multiplied
    const int home = blockIdx.x % kXcds;
    for (int probe = 0; probe < kXcds; ++probe) {
        const int x = (home + probe) % kXcds;
        const int lo = static_cast<int>(static_cast<int64_t>(g.rows) * x / kXcds);
        const int hi = static_cast<int>(static_cast<int64_t>(g.rows) * (x + 1) / kXcds);
        for (;;) {
            int r = 0;
            if (lane == 0)
                r = atomicAdd(&g.queue[x], 1);
            const int row = lo + read_lane(r, 0);
            if (row >= hi)
                break;
            const int64_t c0 = g.c_ptr[row], c1 = g.c_ptr[row + 1];
            if (c0 == c1)
                continue;
            const int a0 = g.a_trp[row];
            const int na = g.a_trp[row + 1] - a0;
            const bool lds_keys = na <= kAKeys;
            if (lds_keys) {
                for (int i = lane; i < na; i += kWave)
                    keys[i] = g.a_meta[a0 + i].x;
                wave_sync();
            }
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
                double acc = 0.0;
                for (int q0 = 0; q0 < total; q0 += kWave) {
                    const int q = q0 + lane;
                    int ai = -1;
                    int4 be = make_int4(0, 0, 0, 0);
                    int c_local = 0;
                    if (q < total) {
                        c_local = upper_index(start, nc, q);
                        be = g.b_cent[cbeg[c_local] + (q - start[c_local])];
                        int l = 0, h = na;
                        while (l < h) {
                            const int mid = (l + h) >> 1;
                            const int k = lds_keys ? keys[mid] : g.a_meta[a0 + mid].x;
                            if (k < be.x)
                                l = mid + 1;
                            else
                                h = mid;
                        }
                        if (l < na && (lds_keys ? keys[l] : g.a_meta[a0 + l].x) == be.x)
                            ai = a0 + l;
                    }
                    const unsigned long long match = __ballot(ai >= 0);
                    if (ai >= 0) {
                        const int4 am = g.a_meta[ai];
                        const int slot = npairs + __popcll(match & ((1ull << lane) - 1ull));
                        pm[slot] = make_int4(am.y, am.w, be.y, be.w);
                        pc[slot] = static_cast<unsigned char>(c_local);
                    }
                    npairs += __popcll(match);
                    // Run the buffered pairs once another round might not fit.
                    const bool last = q0 + kWave >= total;
                    if (npairs + kWave > kPairCap || last) {
                        wave_sync();
                        for (int p0 = 0; p0 < npairs; p0 += kBatch) {
                            const int nbat = min(kBatch, npairs - p0);
                            double av[kBatch], bv[kBatch];
#pragma unroll
                            for (int i = 0; i < kBatch; ++i) {
                                if (i < nbat) {
                                    const int4 m = pm[p0 + i];
                                    av[i] = tile_entry(g.a_vals, static_cast<uint32_t>(m.x),
                                                       static_cast<uint32_t>(m.y), roles.a_phys);
                                    bv[i] = tile_entry(g.b_vals, static_cast<uint32_t>(m.z),
                                                       static_cast<uint32_t>(m.w), roles.b_phys);
                                }
                            }
#pragma unroll
                            for (int i = 0; i < kBatch; ++i) {
                                if (i < nbat) {
                                    const int t = pc[p0 + i];
                                    if (t != cur) {
                                        if (cur >= 0)
                                            store_tile(g, cur, cm, co, roles.c_bit, acc);
                                        cur = t;
                                        acc = 0.0;
                                    }
                                    acc = __builtin_amdgcn_mfma_f64_4x4x4f64(av[i], bv[i], acc, 0,
                                                                            0, 0);
                                }
                            }
                        }
                        done += static_cast<unsigned long long>(npairs);
                        npairs = 0;
                        wave_sync();
                    }
                }
                if (cur >= 0)
                    store_tile(g, cur, cm, co, roles.c_bit, acc);
                wave_sync();
            }
            if (lds_keys)
                wave_sync();
        }
    }
    if (lane == 0 && done)
        atomicAdd(g.mfma_pairs, done);
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

int tile_numeric(const DevTiles &a, const TileColumns &bc, const DevC &c, int precision,
                 SpgemmWork *w, hipStream_t s)
{
    if (c.tilem == 0 || c.num_tiles == 0)
        return RTT_OK;
    if (precision != 0)
        return fail(RTT_ERR_UNSUPPORTED, "tile numeric precision %d is not built", precision);
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
    g.queue = w->tile_queue.as<int>();
    g.mfma_pairs = w->tile_pairs.as<unsigned long long>();
    hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(kXcds), 0, s, g.queue, kXcds,
                       g.mfma_pairs);
    const int cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
    // Persistent: enough workgroups to fill every CU several times over.
    const int blocks = static_cast<int>(std::min<int64_t>(
        static_cast<int64_t>(cus) * 8, (static_cast<int64_t>(g.rows) + kTWaves - 1) / kTWaves));
    hipLaunchKernelGGL(tile_numeric_f64_kernel, dim3(std::max(blocks, 1)), dim3(kTThreads), 0, s,
                       g);
    RTT_HIP(hipGetLastError());
    return RTT_OK;
}

} // namespace rtt
