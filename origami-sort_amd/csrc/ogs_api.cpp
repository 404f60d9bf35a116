// ogs_api.cpp -- the C-ABI of include/ogs.h: contexts, sort drivers, batching.
//
// Driver structure mirrors origami_sorter::sort_single_thread
// (Origami/sorter.h:3851-3915): phase 1-2 = one block-sort launch over all
// tiles (each tile is the GPU analogue of a P2_N chunk, sorter.h:3862-3868),
// phase 3 = merge passes doubling the run length (merge_series,
// merger.h:4220-4228), returning whichever buffer holds the result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/ogs.h"
#include "ogs_internal.hpp"
#include "sort_kernels.hpp"

namespace ogs {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(OGS_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

bool key_info(ogs_key_t kt, KeyInfo *ki) {
    switch (kt) {
    case OGS_U32: *ki = {KC_32, 4, 0}; return true;
    case OGS_I32: *ki = {KC_32, 4, 1}; return true;
    case OGS_F32: *ki = {KC_32, 4, 2}; return true;
    case OGS_U64: *ki = {KC_64, 8, 0}; return true;
    case OGS_I64: *ki = {KC_64, 8, 1}; return true;
    case OGS_KV_U64U64: *ki = {KC_KV, 16, 0}; return true;
    case OGS_KV_I64I64: *ki = {KC_KV, 16, 1}; return true;
    default: return false;
    }
}

// Default tile table (overridden by the autotuner through ogs_tune_set).
static Tuning default_tuning(KeyClass kc) {
    Tuning t;
    (void)kc;
    t.v[TP_PLAIN_BLOCK] = 0;
    t.v[TP_PLAIN_MERGE] = 0;
    t.v[TP_SEG_SMALL_BLOCK] = 0;
    t.v[TP_SEG_SMALL_MAX] = 0;  // 0: the small class is the tile of its shape
    t.v[TP_SEG_LARGE_BLOCK] = 0;
    t.v[TP_SEG_MERGE] = 0;
    return t;
}

Workspace *Ctx::workspace(hipStream_t s) {
    std::lock_guard<std::mutex> g(mu);
    auto it = ws.find(s);
    if (it != ws.end()) return it->second;
    Workspace *w = new Workspace();
    ws[s] = w;
    return w;
}

int Workspace::reserve(size_t bytes) {
    if (bytes <= cap) return OGS_OK;
    size_t want = std::max(bytes, cap * 2);
    want = std::max<size_t>(want, 1 << 20);
    if (dev) {
        hipError_t e = hipFree(dev);  // implicit device sync: earlier users of the block are done
        dev = nullptr;
        cap = 0;
        if (e != hipSuccess) return hip_fail(e, "hipFree(workspace)");
    }
    hipError_t e = hipMalloc(&dev, want);
    if (e != hipSuccess) return fail(OGS_ERR_NOMEM, "hipMalloc(workspace %zu B): %s", want, hipGetErrorString(e));
    cap = want;
    return OGS_OK;
}

Workspace::~Workspace() {
    if (dev) (void)hipFree(dev);
}

const BlockShapeFns &Ctx::block_shape(KeyClass kc, int param) {
    int cnt = 0;
    const BlockShapeFns *t = block_shapes(kc, &cnt);
    int i = tune[kc].v[param];
    if (i < 0 || i >= cnt) i = 0;
    return t[i];
}

const MergeShapeFns &Ctx::merge_shape(KeyClass kc, int param, uint64_t block_tile) {
    int cnt = 0;
    const MergeShapeFns *t = merge_shapes(kc, &cnt);
    int i = tune[kc].v[param];
    if (i < 0 || i >= cnt) i = 0;
    // a merge tile must divide every pair span 2R (R >= block tile)
    if (t[i].tile() > 2 * block_tile) {
        for (int j = 0; j < cnt; j++)
            if (t[j].tile() <= 2 * block_tile) return t[j];
    }
    return t[i];
}

hipEvent_t Profiler::get() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

int Profiler::flush() {
    for (auto &r : pending) {
        hipError_t e = hipEventSynchronize(r.b);
        if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
        float ms = 0;
        e = hipEventElapsedTime(&ms, r.a, r.b);
        if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
        total_ms[r.kind] += ms;
        count[r.kind] += 1;
        items[r.kind] += r.items;
        if (r.own_a) pool.push_back(r.a);
        pool.push_back(r.b);
    }
    pending.clear();
    return OGS_OK;
}

void Profiler::reset() {
    (void)flush();
    for (int k = 0; k < PK_COUNT; k++) {
        total_ms[k] = 0;
        count[k] = 0;
        items[k] = 0;
    }
}

// Plain sort.  Returns the buffer holding the result in *out.
int Ctx::sort(ogs_key_t kt, const void *in, void *data, void *scratch, uint64_t n, void **out, hipStream_t s) {
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    if (!out) return fail(OGS_ERR_ARG, "d_sorted is NULL");
    if (n == 0) {
        *out = data;
        return OGS_OK;
    }
    if (!data || !scratch) return fail(OGS_ERR_ARG, "NULL data or scratch pointer");
    if (data == scratch) return fail(OGS_ERR_ARG, "data and scratch must be distinct buffers");
    if (in && (in == data || in == scratch)) return fail(OGS_ERR_ARG, "input must differ from both work buffers");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const BlockShapeFns &bs = block_shape(ki.kc, TP_PLAIN_BLOCK);
    const MergeShapeFns &ms = merge_shape(ki.kc, TP_PLAIN_MERGE, bs.tile());
    const uint64_t T = bs.tile();
    const bool one = n <= T;
    // passes after the block sort; start on the buffer that makes the last
    // pass land in `data` when reading from a separate input
    int merges = 0;
    for (uint64_t R = T; R < n; R *= 2) merges++;
    void *src = scratch, *dst = data;
    const void *bin = data;
    if (in) {
        bin = in;
        src = (merges & 1) ? scratch : data;
        dst = (src == data) ? scratch : data;
    }
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (prof.on) {
        ev[0] = prof.get();
        ev[1] = prof.get();
        (void)hipEventRecord(ev[0], s);
    }
    e = bs.run(bin, src, n, ki.codec, one ? ki.codec : 0, s);
    if (e != hipSuccess) return hip_fail(e, "block sort launch");
    if (prof.on) {
        (void)hipEventRecord(ev[1], s);
        prof.pending.push_back({PK_BLOCK, ev[0], ev[1], n, true});
    }
    if (!one) {
        for (uint64_t R = T; R < n; R *= 2) {
            const bool last = 2 * R >= n;
            if (prof.on) {
                ev[0] = prof.get();
                ev[1] = prof.get();
                (void)hipEventRecord(ev[0], s);
            }
            e = ms.run(src, dst, n, R, last ? ki.codec : 0, s);
            if (e != hipSuccess) return hip_fail(e, "merge pass launch");
            if (prof.on) {
                (void)hipEventRecord(ev[1], s);
                prof.pending.push_back({PK_MERGE, ev[0], ev[1], n, true});
            }
            std::swap(src, dst);
        }
    }
    *out = src;
    return OGS_OK;
}

int Ctx::pass_count(ogs_key_t kt, uint64_t n) {
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    if (n == 0) return 0;
    const uint64_t T = block_shape(ki.kc, TP_PLAIN_BLOCK).tile();
    int p = 1;
    for (uint64_t R = T; R < n; R *= 2) p++;
    return p;
}

// Segmented sort with host-visible offsets.  Every segment ends in `scratch`.
//  - segments of 1..small_max items: one tile-sort workgroup each (small shape)
//  - longer segments: TILE-sized pieces block-sorted (large shape), then
//    merge passes; a segment needing an odd number of passes block-sorts in
//    place so that its last pass lands in scratch.
int Ctx::segmented(ogs_key_t kt, void *data, void *scratch, const uint64_t *h_off, uint64_t nseg, void **out,
                   hipStream_t s) {
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    if (!out) return fail(OGS_ERR_ARG, "d_sorted is NULL");
    const uint64_t total = nseg ? h_off[nseg] : 0;
    *out = scratch;
    if (total == 0) {
        *out = data;
        return OGS_OK;
    }
    if (!data || !scratch || data == scratch) return fail(OGS_ERR_ARG, "bad data/scratch buffers");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const BlockShapeFns &sb = block_shape(ki.kc, TP_SEG_SMALL_BLOCK);
    const BlockShapeFns &lb = block_shape(ki.kc, TP_SEG_LARGE_BLOCK);
    const MergeShapeFns &mb = merge_shape(ki.kc, TP_SEG_MERGE, lb.tile());
    uint64_t small_max = (uint64_t)tune[ki.kc].v[TP_SEG_SMALL_MAX];
    if (small_max == 0 || small_max > sb.tile()) small_max = sb.tile();
    const uint64_t LT = lb.tile(), MT = mb.tile();

    std::vector<uint64_t> s_off, l_off;
    std::vector<uint32_t> s_cnt, l_cnt;
    struct Large { uint64_t seg, len; int passes; };
    std::vector<Large> large;
    for (uint64_t i = 0; i < nseg; i++) {
        const uint64_t a = h_off[i], b = h_off[i + 1];
        if (b < a || b > total) return fail(OGS_ERR_ARG, "offsets not non-decreasing at segment %llu",
                                            (unsigned long long)i);
        const uint64_t len = b - a;
        if (len == 0) continue;
        if (len <= small_max) {
            s_off.push_back(a);
            s_cnt.push_back((uint32_t)len | kTileFinalBit);
        } else {
            int p = 0;
            for (uint64_t R = LT; R < len; R *= 2) p++;
            large.push_back({a, len, p});
            const uint32_t flags = ((p & 1) ? kTileOddBit : 0u) | (p == 0 ? kTileFinalBit : 0u);
            for (uint64_t o = 0; o < len; o += LT) {
                l_off.push_back(a + o);
                l_cnt.push_back((uint32_t)std::min<uint64_t>(LT, len - o) | flags);
            }
        }
    }
    // merge-pass descriptors, pass by pass; segments with more passes left
    // read from the buffer that their parity dictates
    int max_p = 0;
    for (auto &L : large) max_p = std::max(max_p, L.passes);
    std::vector<SegTileDesc> desc;
    std::vector<uint64_t> pass_start(max_p + 1, 0), pass_cnt(max_p + 1, 0);
    for (int p = 0; p < max_p; p++) {
        pass_start[p] = desc.size();
        const uint64_t R = LT << p;
        for (auto &L : large) {
            if (L.passes <= p) continue;
            // after p passes the data sits in scratch iff (passes - p) is even
            const bool in_scratch = ((L.passes - p) & 1) == 0;
            const uint64_t flags = (in_scratch ? kSegDirBit : 0) | (p == L.passes - 1 ? kSegLastBit : 0);
            (void)R;
            for (uint64_t o = 0; o < L.len; o += MT) desc.push_back({L.seg, L.len, o | flags});
        }
        pass_cnt[p] = desc.size() - pass_start[p];
    }
    // one device upload for every descriptor table
    const size_t n_s = s_off.size(), n_l = l_off.size(), n_d = desc.size();
    size_t bytes = 0;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_soff = bytes; bytes = align(bytes + n_s * 8);
    const size_t o_scnt = bytes; bytes = align(bytes + n_s * 4);
    const size_t o_loff = bytes; bytes = align(bytes + n_l * 8);
    const size_t o_lcnt = bytes; bytes = align(bytes + n_l * 4);
    const size_t o_desc = bytes; bytes = align(bytes + n_d * sizeof(SegTileDesc));
    const size_t o_end = bytes;
    Workspace *w = workspace(s);
    int rc = w->reserve(bytes);
    if (rc) return rc;
    std::vector<uint8_t> &h = w->host;
    h.resize(o_end);
    if (n_s) {
        memcpy(h.data() + o_soff, s_off.data(), n_s * 8);
        memcpy(h.data() + o_scnt, s_cnt.data(), n_s * 4);
    }
    if (n_l) {
        memcpy(h.data() + o_loff, l_off.data(), n_l * 8);
        memcpy(h.data() + o_lcnt, l_cnt.data(), n_l * 4);
    }
    if (n_d) memcpy(h.data() + o_desc, desc.data(), n_d * sizeof(SegTileDesc));
    // pageable-source copy: returns once the source has been consumed
    e = hipMemcpyAsync(w->dev, h.data(), o_end, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "descriptor upload");
    uint8_t *D = (uint8_t *)w->dev;
    e = sb.run_seg(data, scratch, data, (const uint64_t *)(D + o_soff), (const uint32_t *)(D + o_scnt), n_s,
                   ki.codec, ki.codec, s);
    if (e != hipSuccess) return hip_fail(e, "small-segment launch");
    e = lb.run_seg(data, scratch, data, (const uint64_t *)(D + o_loff), (const uint32_t *)(D + o_lcnt), n_l,
                   ki.codec, ki.codec, s);
    if (e != hipSuccess) return hip_fail(e, "large-segment tile launch");
    for (int p = 0; p < max_p; p++) {
        e = mb.run_seg(data, scratch, D + o_desc + pass_start[p] * sizeof(SegTileDesc), pass_cnt[p], LT << p,
                       ki.codec, s);
        if (e != hipSuccess) return hip_fail(e, "segmented merge launch");
    }
    return OGS_OK;
}

}  // namespace ogs

using namespace ogs;

extern "C" {

int ogs_ctx_create(ogs_ctx **ctx, int device) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= ndev) return fail(OGS_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    Ctx *c = new Ctx();
    c->device = device;
    for (int k = 0; k < 3; k++) c->tune[k] = default_tuning((KeyClass)k);
    *ctx = reinterpret_cast<ogs_ctx *>(c);
    return OGS_OK;
}

int ogs_ctx_destroy(ogs_ctx *ctx) {
    if (!ctx) return OGS_OK;
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (auto &kv : c->ws) delete kv.second;
    if (c->host_buf[0]) (void)hipFree(c->host_buf[0]);
    if (c->host_buf[1]) (void)hipFree(c->host_buf[1]);
    delete c;
    return OGS_OK;
}

size_t ogs_item_size(ogs_key_t kt) {
    KeyInfo ki;
    return key_info(kt, &ki) ? (size_t)ki.width : 0;
}

int ogs_sort(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch, uint64_t n, void **d_sorted, void *stream) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    return reinterpret_cast<Ctx *>(ctx)->sort(kt, nullptr, d_data, d_scratch, n, d_sorted, (hipStream_t)stream);
}

int ogs_sort_keep(ogs_ctx *ctx, ogs_key_t kt, const void *d_in, void *d_data, void *d_scratch, uint64_t n,
                  void **d_sorted, void *stream) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    if (!d_in && n) return fail(OGS_ERR_ARG, "d_in is NULL");
    return reinterpret_cast<Ctx *>(ctx)->sort(kt, d_in, d_data, d_scratch, n, d_sorted, (hipStream_t)stream);
}

int ogs_profile_enable(ogs_ctx *ctx, int on) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    int rc = c->prof.flush();
    c->prof.on = on != 0;
    return rc;
}

int ogs_profile_read(ogs_ctx *ctx, int kind, double *total_ms, uint64_t *launches, uint64_t *items) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    if (kind < 0 || kind >= PK_COUNT) return fail(OGS_ERR_ARG, "unknown kernel kind %d", kind);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    int rc = c->prof.flush();
    if (rc) return rc;
    if (total_ms) *total_ms = c->prof.total_ms[kind];
    if (launches) *launches = c->prof.count[kind];
    if (items) *items = c->prof.items[kind];
    return OGS_OK;
}

int ogs_profile_reset(ogs_ctx *ctx) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    reinterpret_cast<Ctx *>(ctx)->prof.reset();
    return OGS_OK;
}

int ogs_sort_chunks(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch, uint64_t n, uint64_t chunk,
                    void **d_sorted, void *stream) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (chunk == 0 || chunk >= n) return c->sort(kt, nullptr, d_data, d_scratch, n, d_sorted, (hipStream_t)stream);
    const uint64_t nseg = (n + chunk - 1) / chunk;
    std::vector<uint64_t> off(nseg + 1);
    for (uint64_t i = 0; i <= nseg; i++) off[i] = std::min(n, i * chunk);
    return c->segmented(kt, d_data, d_scratch, off.data(), nseg, d_sorted, (hipStream_t)stream);
}

int ogs_segmented_sort(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch, const uint64_t *d_offsets,
                       uint64_t nseg, uint64_t n, void **d_sorted, void *stream) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    if (!d_sorted) return fail(OGS_ERR_ARG, "d_sorted is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (nseg == 0) {
        *d_sorted = d_data;
        return OGS_OK;
    }
    if (!d_offsets) return fail(OGS_ERR_ARG, "d_offsets is NULL");
    std::vector<uint64_t> off(nseg + 1);
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    e = hipMemcpyAsync(off.data(), d_offsets, (nseg + 1) * 8, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "offsets download");
    if (off[0] != 0 || off[nseg] != n)
        return fail(OGS_ERR_ARG, "offsets must start at 0 and end at n=%llu (got %llu..%llu)", (unsigned long long)n,
                    (unsigned long long)off[0], (unsigned long long)off[nseg]);
    return c->segmented(kt, d_data, d_scratch, off.data(), nseg, d_sorted, (hipStream_t)stream);
}

int ogs_sort_host(ogs_ctx *ctx, ogs_key_t kt, void *h_data, uint64_t n, uint64_t chunk) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    if (n == 0) return OGS_OK;
    if (!h_data) return fail(OGS_ERR_ARG, "h_data is NULL");
    std::lock_guard<std::mutex> g(c->host_mu);
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t bytes = (size_t)n * ki.width;
    if (bytes > c->host_cap) {
        for (int i = 0; i < 2; i++) {
            if (c->host_buf[i]) (void)hipFree(c->host_buf[i]);
            c->host_buf[i] = nullptr;
        }
        c->host_cap = 0;
        for (int i = 0; i < 2; i++) {
            e = hipMalloc(&c->host_buf[i], bytes);
            if (e != hipSuccess) return fail(OGS_ERR_NOMEM, "hipMalloc(%zu B): %s", bytes, hipGetErrorString(e));
        }
        c->host_cap = bytes;
    }
    e = hipMemcpy(c->host_buf[0], h_data, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "H2D copy");
    void *res = nullptr;
    int rc = ogs_sort_chunks(ctx, kt, c->host_buf[0], c->host_buf[1], n, chunk, &res, nullptr);
    if (rc) return rc;
    e = hipMemcpy(h_data, res, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "D2H copy");
    return OGS_OK;
}

int ogs_pass_count(ogs_ctx *ctx, ogs_key_t kt, uint64_t n) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    return reinterpret_cast<Ctx *>(ctx)->pass_count(kt, n);
}

const char *ogs_last_error(void) { return g_err.c_str(); }

}  // extern "C"
