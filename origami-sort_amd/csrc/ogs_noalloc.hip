// ogs_noalloc.hip -- ogs_sort_noalloc: a plain sort that allocates nothing.
//
// ogs_sort draws its device workspace (the K-way pass's tagged samples and
// tile split tables, the 2-/4-way passes' co-rank tables) from a per-stream
// pool that grows with hipMalloc.  Callers that cannot allocate -- code under
// HIP graph capture, tight serving loops, harnesses that forbid it -- use this
// entry point instead: the workspace is the file-scope device array `scratch`
// below, of kNoallocScratchBytes, one copy per GPU for the life of the process.
//
// `scratch` is a device symbol: in host code its name is not a device address.
// Its address on each device is looked up once, with that device current,
// by hipGetSymbolAddress (noalloc_init, from ogs_ctx_create), and only that
// pointer is handed to kernels.  Nothing in the array must start at zero: the
// samples and split tables are written in full by the kernels of each pass
// before they are read.  Every call on one GPU shares the array, so calls that
// overlap on one GPU (different streams or threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/ogs.h"
#include "ogs_internal.hpp"

namespace ogs {

static constexpr size_t kNoallocScratchBytes = OGS_NOALLOC_SCRATCH_BYTES;
static constexpr int kNoallocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kNoallocMaxDevices];  // scratch's address per device (nullptr: not looked up)

int noalloc_init(int device) {
    if (device < 0 || device >= kNoallocMaxDevices) return OGS_OK;  // ogs_sort_noalloc reports it
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return OGS_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(no-allocation scratch)");
    if (!p) return fail(OGS_ERR_HIP, "hipGetSymbolAddress(no-allocation scratch) returned NULL");
    g_noalloc_ptr[device] = p;
    return OGS_OK;
}

static void *noalloc_ptr(int device) {
    if (device < 0 || device >= kNoallocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

}  // namespace ogs

using namespace ogs;

extern "C" int ogs_sort_noalloc(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch, uint64_t n,
                                void **d_sorted, void *stream) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    void *p = noalloc_ptr(c->device);
    if (!p) return fail(OGS_ERR_STATE, "no no-allocation scratch on device %d", c->device);
    const ExtScratch ext = {(char *)p, kNoallocScratchBytes};
    return c->sort(kt, nullptr, d_data, d_scratch, n, d_sorted, (hipStream_t)stream, nullptr, nullptr, nullptr, &ext);
}
