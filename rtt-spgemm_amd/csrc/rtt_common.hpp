// Shared host-side helpers of the rtt_spgemm library (error text, timing).
#pragma once

#include <chrono>
#include <cstdarg>
#include <cstdint>

#include "rtt_spgemm.h"

namespace rtt {

// Records a formatted message as the thread's last error and returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Zeroes *out and mallocs its CSR arrays (m + 1 row pointers, nnz columns and
// values); RTT_ERR_OVERFLOW beyond the int32 ABI, RTT_ERR_OOM on failure.
int alloc_host_csr(int m, int n, int64_t nnz, rtt_smatrix *out);

using Clock = std::chrono::steady_clock;
inline double ms_since(Clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// splitmix64 (Steele, Lea, Flood 2014): the counter-based generator every
// synthetic input is drawn from, so all hosts produce identical matrices.
struct SplitMix64 {
    uint64_t state;
    explicit SplitMix64(uint64_t seed) : state(seed) {}
    uint64_t next()
    {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // Uniform double in [0, 1) with 53 random bits.
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Stateless draw for position `i` of stream `seed` (parallel generation).
inline uint64_t mix_at(uint64_t seed, uint64_t i)
{
    SplitMix64 g(seed + i * 0x9E3779B97F4A7C15ull);
    return g.next();
}
inline double uniform_at(uint64_t seed, uint64_t i)
{
    return (mix_at(seed, i) >> 11) * (1.0 / 9007199254740992.0);
}

} // namespace rtt
