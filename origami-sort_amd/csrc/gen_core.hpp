// gen_core.hpp -- the counter-based input family shared by the device
// generator (gen_kernels.hip) and its host mirror (ogs_host.cpp).
//
// Item i of a distribution is a function of (seed, i) alone, built from a
// splitmix64 hash, so a GPU writes every item independently in one pass and
// the host reproduces the same bytes.  The distributions follow the inputs of
// the reference's datagen::Writer (Origami/writer.h:27-242, README D1-D9):
// uniform, all same, sorted, reverse sorted, almost sorted (every 7th item the
// largest key), Pareto keys min(ceil(7 (1/(1-u) - 1)), 100000), Pareto-length
// bursts of equal keys back to back and shuffled, and the wrapping Fibonacci
// sequence; plus Zipf(1.1) over 2^20 ranks, 16 distinct values and N(0,1)
// float32 from the project's own generator.  Every formula uses only integer
// arithmetic and correctly rounded double +, -, *, / and ceil, so host and
// device agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ogs {
namespace gen {

enum Dist {
    UNIFORM = 0,
    ZIPF = 1,
    SORTED = 2,
    REVERSE = 3,
    DISTINCT16 = 4,
    NORMAL = 5,
    SAME = 6,
    ALMOST_SORTED = 7,
    PARETO = 8,
    PARETO_BURSTS = 9,
    PARETO_SHUFFLED = 10,
    FIBONACCI = 11,
    COUNT = 12
};

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kZipfRanks = 1u << 20;
constexpr int kFibPeriod = 94;  // F(93) is the last Fibonacci number below 2^64

__host__ __device__ inline uint64_t mix64(uint64_t z) {  // splitmix64 step
    z += kGolden;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The hash of item i in stream s (independent streams for keys, burst lengths
// and shuffle tags).
__host__ __device__ inline uint64_t ctr(uint64_t seed, uint64_t s, uint64_t i) {
    return mix64(mix64(seed ^ (s * 0xD1B54A32D192ED03ull)) + i);
}

// (x >> 11) / 2^53: a double in [0, 1).
__host__ __device__ inline double unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

// Pareto length of the reference's writer (Origami/writer.h:86), beta = 7,
// capped at 100000.
__host__ __device__ inline uint64_t pareto(uint64_t x) {
    const double u = unit(x);
    const double f = ceil(7.0 * (1.0 / (1.0 - u) - 1.0));
    return f > 100000.0 ? 100000u : (uint64_t)f;
}

// N(0,1) by the Irwin-Hall sum of twelve 16-bit uniforms: exact in double and
// in float (at most 20 significant bits).
__host__ __device__ inline uint32_t normal_bits(uint64_t seed, uint64_t i) {
    uint64_t s = 0;
    for (uint64_t k = 0; k < 3; k++) {
        const uint64_t x = ctr(seed, k, i);
        s += (x & 0xFFFF) + ((x >> 16) & 0xFFFF) + ((x >> 32) & 0xFFFF) + (x >> 48);
    }
    const float f = (float)(((double)s - 393210.0) * (1.0 / 65536.0));
    union {
        float f;
        uint32_t u;
    } c;
    c.f = f;
    return c.u;
}

__host__ __device__ inline uint64_t fib(uint64_t i) {
    uint64_t a = 0, b = 1;
    for (uint64_t k = i % kFibPeriod; k; k--) {
        const uint64_t c = a + b;
        a = b;
        b = c;
    }
    return a;
}

// Shuffle tag of item i: random high bits over the index in the low `bits`
// bits, so tags are distinct and their order is a uniform permutation.
__host__ __device__ inline uint64_t shuffle_tag(uint64_t seed, uint64_t i, int bits) {
    const uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    return (ctr(seed, 2, i) & ~mask) | (i & mask);
}

// Key of item i for the distributions that need no table (all but ZIPF,
// PARETO_BURSTS and PARETO_SHUFFLED), as 64 bits; four-byte kinds take
// key32() of it.  `all_ones` is the largest unsigned key of the width.
__host__ __device__ inline uint64_t key64(int dist, uint64_t seed, uint64_t i, uint64_t n) {
    switch (dist) {
    case UNIFORM: return ctr(seed, 0, i);
    case SORTED: return i * (UINT64_MAX / n);
    case REVERSE: return (n - 1 - i) * (UINT64_MAX / n);
    case DISTINCT16: return mix64(ctr(seed, 0, i) & 15u);
    case NORMAL: return normal_bits(seed, i);
    case SAME: return ctr(seed, 0, 0);
    case ALMOST_SORTED: return i % 7 == 0 ? ~0ull : i * (UINT64_MAX / n);
    case PARETO: return pareto(ctr(seed, 0, i));
    case FIBONACCI: return fib(i);
    default: return 0;
    }
}

// The four-byte key of a distribution from its 64-bit form: hashed keys keep
// their high half; sorted/reverse/almost-sorted use their own 32-bit spacing;
// Pareto, Fibonacci and float bits are taken as values (Fibonacci wraps, as the
// reference's 32-bit store of its 64-bit sum does).
__host__ __device__ inline uint32_t key32(int dist, uint64_t k, uint64_t i, uint64_t n) {
    switch (dist) {
    case SORTED: return (uint32_t)(i * (UINT32_MAX / n));
    case REVERSE: return (uint32_t)((n - 1 - i) * (UINT32_MAX / n));
    case ALMOST_SORTED: return i % 7 == 0 ? 0xFFFFFFFFu : (uint32_t)(i * (UINT32_MAX / n));
    case NORMAL:
    case PARETO:
    case FIBONACCI: return (uint32_t)k;
    default: return (uint32_t)(k >> 32);
    }
}

}  // namespace gen
}  // namespace ogs
