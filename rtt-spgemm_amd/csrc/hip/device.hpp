// Device helpers shared by the kernels (wave64 primitives, tile bit algebra).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtt {

constexpr int kWave = 64;

// Orders the LDS traffic of one wavefront: every store issued before the
// barrier is visible to every lane after it.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on
// XCD b % 8).  The row kernels keep the rows in flight on the chip one
// window of consecutive rows spread over every XCD: measured against giving
// each XCD a contiguous eighth of the rows, the stencil A*A^T numeric pass
// ran 3 % and the banded batch 6 % faster.
constexpr int kXcds = 8;

// Row of workgroup b of n: chunks of C consecutive rows dealt to the XCDs
// round-robin (XCD x takes chunks x, x + 8, ...), so that the rows one L2
// holds come in runs of C; rows past the last whole round of 8 C rows map
// to themselves.  A bijection of [0, n).
template <int C>
__device__ __forceinline__ int xcd_chunk_row(int b, int n)
{
    if (b >= n / (kXcds * C) * (kXcds * C))
        return b;
    const int x = b % kXcds, i = b / kXcds;
    return ((i / C) * kXcds + x) * C + i % C;
}

// Gathers bits k, k+4, ..., k+28 of a 32-bit tile mask into 8 bits.  For an
// A tile (bit r*4+k) this is the set of rows present in column k; for a B
// tile (bit c*4+k) the set of columns present in row k.
__device__ __forceinline__ uint32_t stride4_bits(uint32_t mask, int k)
{
    uint32_t x = (mask >> k) & 0x11111111u;
    x = (x | (x >> 3)) & 0x03030303u;
    x = (x | (x >> 6)) & 0x000F000Fu;
    x = (x | (x >> 12)) & 0x000000FFu;
    return x;
}

// Bit r of `rows` -> byte r of the result is 0xFF.
__device__ __forceinline__ uint64_t rows_to_bytes(uint32_t rows)
{
    uint64_t s = rows;
    s = (s | (s << 28)) & 0x0000000F0000000Full;
    s = (s | (s << 14)) & 0x0003000300030003ull;
    s = (s | (s << 7)) & 0x0101010101010101ull;
    return s * 0xFFull;
}

// Structural product of an 8x4 A tile mask (bit r*4+k) and a 4x8 B tile
// mask (bit c*4+k): the 8x8 C tile mask (bit r*8+c).
__device__ __forceinline__ uint64_t tile_bool_product(uint32_t am, uint32_t bm)
{
    uint64_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t rows = stride4_bits(am, k);
        const uint32_t cols = stride4_bits(bm, k);
        out |= rows_to_bytes(rows) & (static_cast<uint64_t>(cols) * 0x0101010101010101ull);
    }
    return out;
}

// Payload slot of physical position `phys` of a hybrid tile.
__device__ __forceinline__ int payload_slot(uint32_t mask, int span, int phys)
{
    return span == 32 ? phys : __popc(mask & ((1u << phys) - 1u));
}

__device__ __forceinline__ uint32_t hash_tile_col(int j)
{
    return static_cast<uint32_t>(j) * 0x9E3779B1u;
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v)
{
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const T u = __shfl_up(v, o, kWave);
        if (lane >= o)
            v += u;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}

// Largest index i in [0, n) with pre[i] <= p (pre non-decreasing, pre[0] = 0).
__device__ __forceinline__ int upper_index(const int *pre, int n, int p)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

} // namespace rtt
