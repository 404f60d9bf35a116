// gen_kernels.hip -- device form of the counter-based input family
// (gen_core.hpp): one grid-stride pass writes every item from (seed, index).
// Zipf reads the host's cumulative-weight table; Pareto bursts read the table
// of burst starts; the shuffled bursts are (tag, key) pairs put in tag order by
// this library's sort and then unpacked.
#include "gen_core.hpp"
#include "sort_kernels.hpp"

namespace ogs {

__device__ __forceinline__ uint64_t burst_of(const uint64_t *starts, uint64_t nb, uint64_t i) {
    uint64_t lo = 0, hi = nb;  // first start above i
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (starts[mid] <= i) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

// W: item layout, 4 / 8 (keys), 16 ({u64 key, u64 index}) or kGenKv32
// ({u32 key, u32 index}, the keys of the 4-byte layout)
template <int W>
__device__ __forceinline__ void gen_store(void *out, uint64_t i, int dist, uint64_t k, uint64_t n) {
    if constexpr (W == 4 || W == kGenKv32) {
        uint32_t v = gen::key32(dist, k, i, n);
        if (dist == gen::NORMAL) {
            if (i == n / 2) v = 0u;  // +0.0
            else if (i == 0) v = 0x80000000u;  // -0.0
        }
        if constexpr (W == 4) ((uint32_t *)out)[i] = v;
        else ((uint2 *)out)[i] = make_uint2(v, (uint32_t)i);
    } else if constexpr (W == 8) {
        ((uint64_t *)out)[i] = k;
    } else {
        ((ulonglong2 *)out)[i] = make_ulonglong2(k, i);
    }
}

template <int W>
__global__ void k_generate(int dist, uint64_t seed, uint64_t n, void *out, const double *__restrict__ cdf,
                           const uint64_t *__restrict__ starts, uint64_t nb) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t k;
        if (dist == gen::ZIPF) {
            const double u = gen::unit(gen::ctr(seed, 0, i)) * cdf[gen::kZipfRanks - 1];
            uint64_t lo = 0, hi = gen::kZipfRanks;  // first rank whose cumulative weight reaches u
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (cdf[mid] < u) lo = mid + 1;
                else hi = mid;
            }
            k = gen::mix64(lo);
        } else if (dist == gen::PARETO_BURSTS) {
            k = gen::ctr(seed, 0, burst_of(starts, nb, i));
        } else {
            k = gen::key64(dist, seed, i, n);
        }
        gen_store<W>(out, i, dist, k, n);
    }
}

__global__ void k_generate_tagged(uint64_t seed, uint64_t n, int bits, const uint64_t *__restrict__ starts,
                                  uint64_t nb, ulonglong2 *pairs) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        pairs[i] = make_ulonglong2(gen::shuffle_tag(seed, i, bits), gen::ctr(seed, 0, burst_of(starts, nb, i)));
}

template <int W>
__global__ void k_generate_untag(const ulonglong2 *__restrict__ pairs, uint64_t n, void *out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        gen_store<W>(out, i, gen::PARETO_SHUFFLED, pairs[i].y, n);
}

static dim3 gen_grid(uint64_t n) {
    const uint64_t b = (n + 255) / 256;
    return dim3((unsigned)(b < 65536 ? (b ? b : 1) : 65536));
}

hipError_t launch_generate(int width, int dist, uint64_t seed, uint64_t n, void *out, const double *cdf,
                           const uint64_t *starts, uint64_t nb, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (width == 4)
        hipLaunchKernelGGL(k_generate<4>, gen_grid(n), dim3(256), 0, s, dist, seed, n, out, cdf, starts, nb);
    else if (width == 8)
        hipLaunchKernelGGL(k_generate<8>, gen_grid(n), dim3(256), 0, s, dist, seed, n, out, cdf, starts, nb);
    else if (width == kGenKv32)
        hipLaunchKernelGGL(k_generate<kGenKv32>, gen_grid(n), dim3(256), 0, s, dist, seed, n, out, cdf, starts, nb);
    else
        hipLaunchKernelGGL(k_generate<16>, gen_grid(n), dim3(256), 0, s, dist, seed, n, out, cdf, starts, nb);
    return hipGetLastError();
}

hipError_t launch_generate_tagged(uint64_t seed, uint64_t n, int bits, const uint64_t *starts, uint64_t nb,
                                  void *pairs, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_generate_tagged, gen_grid(n), dim3(256), 0, s, seed, n, bits, starts, nb,
                       (ulonglong2 *)pairs);
    return hipGetLastError();
}

hipError_t launch_generate_untag(int width, const void *pairs, uint64_t n, void *out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const ulonglong2 *p = (const ulonglong2 *)pairs;
    if (width == 4) hipLaunchKernelGGL(k_generate_untag<4>, gen_grid(n), dim3(256), 0, s, p, n, out);
    else if (width == 8) hipLaunchKernelGGL(k_generate_untag<8>, gen_grid(n), dim3(256), 0, s, p, n, out);
    else if (width == kGenKv32) hipLaunchKernelGGL(k_generate_untag<kGenKv32>, gen_grid(n), dim3(256), 0, s, p, n, out);
    else hipLaunchKernelGGL(k_generate_untag<16>, gen_grid(n), dim3(256), 0, s, p, n, out);
    return hipGetLastError();
}

}  // namespace ogs
