// layout_kernels.hip -- the structure-of-arrays form of key-value sorts.
// Keys and values in two arrays are packed into the 16-byte {key, value}
// records the sort kernels move (the layout of KeyValue<i64,i64>,
// Origami/commons.h:51-72), and unpacked afterwards.  Each is one coalesced,
// grid-stride pass: 8-byte loads (or stores) per array, one 16-byte store (or
// load) per record.
#include "sort_kernels.hpp"

namespace ogs {

__global__ void k_pack_soa(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                           ulonglong2 *__restrict__ out, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = make_ulonglong2(keys[i], vals[i]);
}

__global__ void k_unpack_soa(const ulonglong2 *__restrict__ in, uint64_t *__restrict__ keys,
                             uint64_t *__restrict__ vals, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const ulonglong2 p = in[i];
        keys[i] = p.x;
        vals[i] = p.y;
    }
}

static dim3 layout_grid(uint64_t n) {
    const uint64_t b = (n + 255) / 256;
    return dim3((unsigned)(b < 32768 ? (b ? b : 1) : 32768));
}

hipError_t launch_pack_soa(const void *keys, const void *vals, void *out, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_soa, layout_grid(n), dim3(256), 0, s, (const uint64_t *)keys, (const uint64_t *)vals,
                       (ulonglong2 *)out, n);
    return hipGetLastError();
}

hipError_t launch_unpack_soa(const void *in, void *keys, void *vals, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack_soa, layout_grid(n), dim3(256), 0, s, (const ulonglong2 *)in, (uint64_t *)keys,
                       (uint64_t *)vals, n);
    return hipGetLastError();
}

}  // namespace ogs
