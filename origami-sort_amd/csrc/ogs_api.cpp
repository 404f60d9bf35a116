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
#include "gen_core.hpp"
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
    case OGS_KV_U32U32: *ki = {KC_64, 8, 3}; return true;
    default: return false;
    }
}

// Shape indices by shape: the compiled lists' positions of a block-sort shape
// (block, ipt, wave levels), a merge shape and a K-way shape.  Absent shapes
// give 0, the list's first entry.
static int block_index(KeyClass kc, int block, int ipt, int wlev) {
    int cnt = 0;
    const BlockShapeFns *t = block_shapes(kc, &cnt);
    for (int i = 0; i < cnt; i++)
        if (t[i].block == block && t[i].ipt == ipt && t[i].wlev == wlev) return i;
    return 0;
}
static int merge_index(KeyClass kc, int block, int ipt) {
    int cnt = 0;
    const MergeShapeFns *t = merge_shapes(kc, &cnt);
    for (int i = 0; i < cnt; i++)
        if (t[i].block == block && t[i].ipt == ipt) return i;
    return 0;
}
static int kway_index(KeyClass kc, int block, int ipt) {
    int cnt = 0;
    const KwayShapeFns *t = kway_shapes(kc, &cnt);
    for (int i = 0; i < cnt; i++)
        if (t[i].block == block && t[i].ipt == ipt) return i;
    return 0;
}

// The shipped tile table, measured on MI355X (the run-time counterpart of the
// tuned constants of Origami/config.h:45-264).  A context starts from it;
// ogs_tune_set / ogs_tune_load ($OGS_TILES, written by autotune.py) override
// entries.
// The K-way shape of the sample class skc whose one tile holds a whole group of
// kw sample runs of 2^rs_log2 records (each padded to the shape's IPT), the
// smallest such; nullptr if none does.
static const KwayShapeFns *whole_group_shape(KeyClass skc, int kw, int rs_log2) {
    int cnt = 0;
    const KwayShapeFns *t = kway_shapes(skc, &cnt);
    const KwayShapeFns *best = nullptr;
    for (int j = 0; j < cnt; j++) {
        const uint64_t ipt = (uint64_t)t[j].ipt, need = (uint64_t)kw * (((1ull << rs_log2) + ipt - 1) / ipt * ipt);
        if (need <= t[j].tile() && (!best || t[j].tile() < best->tile())) best = &t[j];
    }
    return best;
}

Tuning default_tuning(KeyClass kc) {
    Tuning t;
    t.v[TP_PLAIN_CORANK] = 0;
    t.v[TP_SEG_TINY_MAX] = 0;    // no tiny class
    t.v[TP_SEG_SMALL_MAX] = 1024;
    t.v[TP_SEG_PLAIN_MIN] = 0;   // every batch segment runs through the batched kernels
    t.v[TP_PARTITION_LANES] = 4;
    t.v[TP_QUAD_LANES] = 1;
    t.v[TP_KWAY_SCORANK] = 1;
    t.v[TP_KWAY_SWHOLE] = 1;
    if (kc == KC_32) {
        t.v[TP_PLAIN_BLOCK] = block_index(kc, 512, 32, 6);
        t.v[TP_PLAIN_MERGE] = merge_index(kc, 512, 32);
        t.v[TP_SEG_TINY_BLOCK] = block_index(kc, 64, 16, 1);
        t.v[TP_SEG_SMALL_BLOCK] = block_index(kc, 256, 8, 1);
        t.v[TP_SEG_LARGE_BLOCK] = block_index(kc, 512, 32, 6);
        t.v[TP_SEG_MERGE] = merge_index(kc, 256, 16);
        t.v[TP_PLAIN_FANIN] = 2;
        t.v[TP_KWAY_FANIN] = 16;
        t.v[TP_KWAY_SHAPE] = kway_index(kc, 512, 16);
        t.v[TP_KWAY_SAMPLE] = 6;
    } else if (kc == KC_64) {
        t.v[TP_PLAIN_BLOCK] = block_index(kc, 512, 16, 3);
        t.v[TP_PLAIN_MERGE] = merge_index(kc, 512, 16);
        t.v[TP_SEG_TINY_BLOCK] = block_index(kc, 64, 8, 1);
        t.v[TP_SEG_SMALL_BLOCK] = block_index(kc, 256, 8, 1);
        t.v[TP_SEG_LARGE_BLOCK] = block_index(kc, 512, 16, 3);
        t.v[TP_SEG_MERGE] = merge_index(kc, 256, 16);
        t.v[TP_PLAIN_FANIN] = 2;
        t.v[TP_KWAY_FANIN] = 16;
        t.v[TP_KWAY_SHAPE] = kway_index(kc, 512, 16);
        t.v[TP_KWAY_SAMPLE] = 6;
    } else {
        t.v[TP_PLAIN_BLOCK] = block_index(kc, 512, 8, 3);
        t.v[TP_PLAIN_MERGE] = merge_index(kc, 512, 4);
        t.v[TP_SEG_TINY_BLOCK] = block_index(kc, 64, 8, 1);
        t.v[TP_SEG_SMALL_BLOCK] = block_index(kc, 256, 4, 1);
        t.v[TP_SEG_LARGE_BLOCK] = block_index(kc, 512, 8, 3);
        t.v[TP_SEG_MERGE] = merge_index(kc, 512, 4);
        t.v[TP_PLAIN_FANIN] = 4;
        t.v[TP_KWAY_FANIN] = 16;
        t.v[TP_KWAY_SHAPE] = kway_index(kc, 1024, 4);
        t.v[TP_KWAY_SAMPLE] = 5;
    }
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

// Frees stream s's workspace (the caller has finished all work on s).
void Ctx::drop_workspace(hipStream_t s) {
    std::lock_guard<std::mutex> g(mu);
    auto it = ws.find(s);
    if (it == ws.end()) return;
    delete it->second;
    ws.erase(it);
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

int Workspace::reserve_part(size_t bytes) {
    if (bytes <= part_cap) return OGS_OK;
    size_t want = std::max(bytes, part_cap * 2);
    want = std::max<size_t>(want, 1 << 20);
    if (part) {
        hipError_t e = hipFree(part);  // implicit device sync: earlier users of the block are done
        part = nullptr;
        part_cap = 0;
        if (e != hipSuccess) return hip_fail(e, "hipFree(co-rank table)");
    }
    hipError_t e = hipMalloc(&part, want);
    if (e != hipSuccess) return fail(OGS_ERR_NOMEM, "hipMalloc(co-rank table %zu B): %s", want, hipGetErrorString(e));
    part_cap = want;
    return OGS_OK;
}

Workspace::Workspace() { (void)hipEventCreateWithFlags(&staged, hipEventDisableTiming); }

Workspace::~Workspace() {
    if (dev) (void)hipFree(dev);
    if (part) (void)hipFree(part);
    if (staged) (void)hipEventDestroy(staged);
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
    // a merge tile must divide every pair span 2R (R >= block tile): otherwise
    // take the largest registered merge tile that does (every class registers
    // a 512-item shape, which fits the smallest block tile)
    if (t[i].tile() > 2 * block_tile) {
        int best = -1;
        for (int j = 0; j < cnt; j++)
            if (t[j].tile() <= 2 * block_tile && (best < 0 || t[j].tile() > t[best].tile())) best = j;
        if (best >= 0) return t[best];
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
int Ctx::sort(ogs_key_t kt, const void *in, void *data, void *scratch, uint64_t n, void **out, hipStream_t s,
              const void *vin, void *kout, void *vout, const ExtScratch *ext) {
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    if ((vin || vout) && ki.width != 16) return fail(OGS_ERR_ARG, "separate value arrays need a key-value type");
    if (vin && !in) return fail(OGS_ERR_ARG, "separate values need the keys as the input");
    if (!vout != !kout) return fail(OGS_ERR_ARG, "kout and vout go together");
    if (!out) return fail(OGS_ERR_ARG, "d_sorted is NULL");
    if (n == 0) {
        *out = data;
        return OGS_OK;
    }
    if (!data || !scratch) return fail(OGS_ERR_ARG, "NULL data or scratch pointer");
    if (data == scratch) return fail(OGS_ERR_ARG, "data and scratch must be distinct buffers");
    // `in` may alias data or scratch: the block sort reads each tile whole
    // before writing it back
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if (!vin && !vout && kway_on(ki, n)) return sort_kway(ki, in, data, scratch, n, out, s, ext);
    const bool prof_on = prof.on && !ext;  // profiling events are allocated
    const BlockShapeFns &bs = block_shape(ki.kc, TP_PLAIN_BLOCK);
    const MergeShapeFns &ms = merge_shape(ki.kc, TP_PLAIN_MERGE, bs.tile());
    const uint64_t T = bs.tile();
    const bool one = n <= T;
    // merge passes after the block sort: 4-way passes (runs R -> 4R, two
    // levels of the merge tree per HBM pass) while at least two levels remain
    // and the table asks for fan-in 4, else 2-way; start on the buffer that
    // makes the last pass land in `data` when reading from a separate input
    const bool quad = tune[ki.kc].v[TP_PLAIN_FANIN] == 4 && ms.quad_ok;
    int ways[kMaxPasses];
    int merges = 0;
    for (uint64_t R = T; R < n;) {
        const int wy = (quad && 2 * R < n) ? 4 : 2;
        ways[merges++] = wy;
        R *= (uint64_t)wy;
    }
    void *src = scratch, *dst = data;
    const void *bin = data;
    if (in) {
        bin = in;
        src = (merges & 1) ? scratch : data;
        dst = (src == data) ? scratch : data;
    }
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (prof_on) {
        ev[0] = prof.get();
        ev[1] = prof.get();
        (void)hipEventRecord(ev[0], s);
    }
    e = bs.run(bin, one && vout ? kout : src, n, ki.codec, one ? ki.codec : 0, s, vin, one ? vout : nullptr, nullptr);
    if (e != hipSuccess) return hip_fail(e, "block sort launch");
    if (prof_on) {
        (void)hipEventRecord(ev[1], s);
        prof.pending.push_back({PK_BLOCK, ev[0], ev[1], n, true});
    }
    // co-rank tables: 2-way tile co-ranks, and for 4-way passes the tile
    // split table (4 entries per slot; the first 4-way pass has the most)
    const uint64_t MT = ms.tile();
    uint64_t *part = nullptr;
    const bool table = tune[ki.kc].v[TP_PLAIN_CORANK] == 0;
    if (!one && (table || quad)) {
        uint64_t need = (n + MT - 1) / MT + 4;
        uint64_t R = T;
        for (int i = 0; i < merges; i++) {
            if (ways[i] == 4) need = std::max<uint64_t>(need, 4 * quad_table_slots(n, R, MT) + 8);
            R *= (uint64_t)ways[i];
        }
        if (ext) {
            if (need * sizeof(uint64_t) > ext->cap)
                return fail(OGS_ERR_ARG, "n=%llu needs %llu bytes of merge tables, above the %zu-byte scratch",
                            (unsigned long long)n, (unsigned long long)(need * sizeof(uint64_t)), ext->cap);
            part = (uint64_t *)ext->base;
        } else {
            Workspace *w = workspace(s);
            std::lock_guard<std::mutex> g(w->mu);
            int rc = w->reserve_part(need * sizeof(uint64_t));
            if (rc) return rc;
            part = (uint64_t *)w->part;
        }
    }
    if (!one) {
        uint64_t R = T;
        for (int pi = 0; pi < merges; pi++) {
            const int wy = ways[pi];
            const bool last = pi == merges - 1;
            if (prof_on) {
                for (int k = 0; k < 3; k++) ev[k] = prof.get();
                (void)hipEventRecord(ev[0], s);
            }
            if (wy == 4)
                e = ms.run_quad(src, last && vout ? kout : dst, n, R, last ? ki.codec : 0, part,
                                tune[ki.kc].v[TP_QUAD_LANES], s, prof_on ? ev[1] : nullptr, last ? vout : nullptr,
                                err);
            else
                e = ms.run(src, last && vout ? kout : dst, n, R, last ? ki.codec : 0, table ? part : nullptr,
                           tune[ki.kc].v[TP_PARTITION_LANES], s, prof_on ? ev[1] : nullptr, last ? vout : nullptr);
            R *= (uint64_t)wy;
            if (e != hipSuccess) return hip_fail(e, "merge pass launch");
            if (prof_on) {
                (void)hipEventRecord(ev[2], s);
                // ev[1] closes the partition and opens the merge; the
                // partition record owns it
                prof.pending.push_back({PK_PARTITION, ev[0], ev[1], n, true});
                prof.pending.push_back({wy == 4 ? PK_MERGE4 : PK_MERGE, ev[1], ev[2], n, false});
            }
            std::swap(src, dst);
        }
    }
    *out = vout ? kout : src;
    return OGS_OK;
}

static int ilog2(uint64_t x) { return 63 - __builtin_clzll(x); }

bool Ctx::kway_on(const KeyInfo &ki, uint64_t n) {
    if (tune[ki.kc].v[TP_KWAY_FANIN] < 2) return false;
    return n > block_shape(ki.kc, TP_PLAIN_BLOCK).tile();
}

// Merge levels per K-way pass: ceil(levels / log2(fan-in)) passes, the levels
// spread evenly over them, the larger passes first.
int Ctx::kway_plan(KeyClass kc, uint64_t n, uint64_t block_tile, int *lv) {
    int L = 0;
    for (uint64_t R = block_tile; R < n; R <<= 1) L++;
    const int lmax = ilog2((uint64_t)tune[kc].v[TP_KWAY_FANIN]);
    const int P = (L + lmax - 1) / lmax;
    for (int i = 0; i < P; i++) lv[i] = L / P + (i < L % P ? 1 : 0);
    return P;
}

// Plain sort with sample-split K-way passes.  Pass p merges groups of
// 2^lv[p] runs; its tile splits come from every 2^sl[p]-th item of each input
// run (written as tagged samples by the kernel that wrote the run), merged per
// group by 2-way merge passes over the sample array.  Sample buffers: S0 holds
// the unmerged samples of the current input; S1 and S2 take the sample merges.
int Ctx::sort_kway(const KeyInfo &ki, const void *in, void *data, void *scratch, uint64_t n, void **out,
                   hipStream_t s, const ExtScratch *ext) {
    const BlockShapeFns &bs = block_shape(ki.kc, TP_PLAIN_BLOCK);
    int kcnt = 0;
    const KwayShapeFns *kt = kway_shapes(ki.kc, &kcnt);
    int kidx = tune[ki.kc].v[TP_KWAY_SHAPE];
    if (kidx < 0 || kidx >= kcnt) kidx = 0;
    const KwayShapeFns &ks = kt[kidx];
    const uint64_t T = bs.tile();
    if (T & (T - 1)) return fail(OGS_ERR_ARG, "K-way passes need a power-of-two block tile");
    int lv[kMaxPasses], sl[kMaxPasses];
    const int P = kway_plan(ki.kc, n, T, lv);
    const bool prof_on = prof.on && !ext;  // profiling events are allocated
    // sample spacing per pass: at most the tuned one, and at most R / 256 so
    // the smallest 2-way sample merge tile (512) divides every pair span
    int rl = ilog2(T), smin = 63;
    for (int p = 0; p < P; p++) {
        sl[p] = std::min(tune[ki.kc].v[TP_KWAY_SAMPLE], rl - 8);
        if (sl[p] < 0 || rl - sl[p] > 27)
            return fail(OGS_ERR_ARG, "K-way pass %d: no sample spacing for runs of 2^%d", p, rl);
        smin = std::min(smin, sl[p]);
        rl += lv[p];
    }
    // sample records: u64 (32-bit keys) merged as u64 keys, {key, tag} (64-bit
    // keys) merged as stable key-value records
    const KeyClass skc = ki.kc == KC_32 ? KC_64 : KC_KV;
    const uint64_t sw = ki.kc == KC_32 ? 8 : 16;
    const uint64_t ns_max = (n + (1ull << smin) - 1) >> smin;
    uint64_t tab_n = 0;
    rl = ilog2(T);
    for (int p = 0; p < P; p++) {
        const uint64_t nt = kway_tiles(n, rl, sl[p], 1 << lv[p], ks.tile(), ks.ipt);
        if (nt == 0) return fail(OGS_ERR_ARG, "K-way pass %d: the tile shape cannot hold %d slices", p, 1 << lv[p]);
        tab_n = std::max(tab_n, nt * (uint64_t)(1 << lv[p]));
        rl += lv[p];
    }
    const uint64_t part_n = (ns_max + 511) / 512 + 4;
    auto al = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
    const uint64_t o_s1 = al(ns_max * sw), o_s2 = o_s1 + al(ns_max * sw), o_tab = o_s2 + al(ns_max * sw),
                   o_part = o_tab + al(tab_n * 8), bytes = o_part + al(part_n * 8);
    char *D = nullptr;
    std::unique_lock<std::mutex> g;
    if (ext) {
        if (bytes > ext->cap)
            return fail(OGS_ERR_ARG, "n=%llu needs %llu bytes of K-way samples and split tables, above the %zu-byte "
                        "scratch", (unsigned long long)n, (unsigned long long)bytes, ext->cap);
        D = ext->base;
    } else {
        Workspace *w = workspace(s);
        g = std::unique_lock<std::mutex>(w->mu);
        int rc = w->reserve(bytes);
        if (rc) return rc;
        D = (char *)w->dev;
    }
    void *S0 = D, *S1 = D + o_s1, *S2 = D + o_s2;
    uint64_t *tab = (uint64_t *)(D + o_tab), *mpart = (uint64_t *)(D + o_part);
    void *src = scratch, *dst = data;
    const void *bin = data;
    if (in) {
        bin = in;
        src = (P & 1) ? scratch : data;
        dst = (src == data) ? scratch : data;
    }
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (prof_on) {
        ev[0] = prof.get();
        ev[1] = prof.get();
        (void)hipEventRecord(ev[0], s);
    }
    rl = ilog2(T);
    const SampleSpec first = {S0, rl, 1 << lv[0], sl[0]};
    hipError_t e = bs.run(bin, src, n, ki.codec, 0, s, nullptr, nullptr, &first);
    if (e != hipSuccess) return hip_fail(e, "block sort launch");
    if (prof_on) {
        (void)hipEventRecord(ev[1], s);
        prof.pending.push_back({PK_BLOCK, ev[0], ev[1], n, true});
    }
    int mcnt = 0;
    const MergeShapeFns *mt = merge_shapes(skc, &mcnt);
    const int lanes = tune[skc].v[TP_PARTITION_LANES];
    for (int p = 0; p < P; p++) {
        const bool last = p == P - 1;
        if (prof_on) {
            for (int k = 0; k < 3; k++) ev[k] = prof.get();
            (void)hipEventRecord(ev[0], s);
        }
        // merge the samples of each group of 2^lv[p] runs
        const uint64_t ns = (n + (1ull << sl[p]) - 1) >> sl[p];
        const void *sa = S0;
        void *sb[2] = {S1, S2};
        // 32-bit keys: a u64 sample is its own total order (key, run, index)
        // and unique, so when a group's samples fill exactly one block-sort
        // tile of the 64-bit class, one block sort replaces the merge passes
        // a group's samples in one K-way tile: one launch merges every group
        const KwayShapeFns *kwh =
            tune[ki.kc].v[TP_KWAY_SWHOLE] ? whole_group_shape(skc, 1 << lv[p], rl - sl[p]) : nullptr;
        const BlockShapeFns *sbs = nullptr;
        if (ki.kc == KC_32 && !kwh) {
            const uint64_t gspan = 1ull << (rl - sl[p] + lv[p]);
            int bcnt = 0;
            const BlockShapeFns *bt = block_shapes(KC_64, &bcnt);
            for (int j = 0; j < bcnt && !sbs; j++)
                if (bt[j].tile() == gspan) sbs = &bt[j];
        }
        if (kwh) {
            e = kwh->whole[lv[p] - 1](S0, S1, ns, rl - sl[p], err, s);
            if (e != hipSuccess) return hip_fail(e, "sample merge launch");
            sa = S1;
        } else if (sbs) {
            e = sbs->run(S0, S1, ns, 0, 0, s, nullptr, nullptr, nullptr);
            if (e != hipSuccess) return hip_fail(e, "sample sort launch");
            sa = S1;
        }
        for (int q = 0; q < (sbs || kwh ? 0 : lv[p]); q++) {
            const uint64_t Rs = 1ull << (rl - sl[p] + q);
            int best = -1;
            for (int j = 0; j < mcnt; j++)
                if (mt[j].tile() <= 2 * Rs && (best < 0 || mt[j].tile() > mt[best].tile())) best = j;
            if (best < 0) return fail(OGS_ERR_ARG, "no sample merge tile divides 2^%d", (int)ilog2(2 * Rs));
            e = mt[best].run(sa, sb[q & 1], ns, Rs, 0, tune[ki.kc].v[TP_KWAY_SCORANK] ? nullptr : mpart, lanes, s,
                             nullptr, nullptr);
            if (e != hipSuccess) return hip_fail(e, "sample merge launch");
            sa = sb[q & 1];
        }
        SampleSpec next = {nullptr, 0, 0, 0};
        if (!last) next = {S0, rl + lv[p], 1 << lv[p + 1], sl[p + 1]};
        e = ks.run[lv[p] - 1](src, dst, n, rl, sl[p], last ? ki.codec : 0, tab, S0, sa, last ? nullptr : &next, s,
                              prof_on ? ev[1] : nullptr, err);
        if (e != hipSuccess) return hip_fail(e, "K-way pass launch");
        if (prof_on) {
            (void)hipEventRecord(ev[2], s);
            prof.pending.push_back({PK_PARTITION, ev[0], ev[1], n, true});
            prof.pending.push_back({PK_KWAY, ev[1], ev[2], n, false});
        }
        rl += lv[p];
        std::swap(src, dst);
    }
    *out = src;
    return OGS_OK;
}

int Ctx::pass_count(ogs_key_t kt, uint64_t n) {
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    if (n == 0) return 0;
    int lv[kMaxPasses];
    if (kway_on(ki, n)) return 1 + kway_plan(ki.kc, n, block_shape(ki.kc, TP_PLAIN_BLOCK).tile(), lv);
    const uint64_t T = block_shape(ki.kc, TP_PLAIN_BLOCK).tile();
    const bool quad = tune[ki.kc].v[TP_PLAIN_FANIN] == 4 && merge_shape(ki.kc, TP_PLAIN_MERGE, T).quad_ok;
    int p = 1;
    for (uint64_t R = T; R < n; p++) R *= (quad && 2 * R < n) ? 4 : 2;
    return p;
}

// Batched sort of independent segments; segment i is sorted from segs[i].in
// into segs[i].fin, using segs[i].other as its second buffer (`in` may alias
// either; fin and other must differ).  Size classes:
//  - up to tiny_max items: one workgroup of the tiny shape per segment
//  - up to small_max items: one workgroup of the small shape per segment
//  - longer: TILE-sized pieces block-sorted with the large shape, then merge
//    passes; the pieces start on the buffer that makes the segment's last
//    pass land in `fin` (a segment with P passes starts in fin when P is even).
// All segments of a class share one launch; all segments' p-th merge pass
// share one launch.
int Ctx::batch_build(ogs_key_t kt, const SegSpec *segs, size_t nseg, bool allow_plain, BatchBuild &b) {
    if (!key_info(kt, &b.ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    const KeyInfo &ki = b.ki;
    const BlockShapeFns &tb = block_shape(ki.kc, TP_SEG_TINY_BLOCK);
    const BlockShapeFns &sb = block_shape(ki.kc, TP_SEG_SMALL_BLOCK);
    const BlockShapeFns &lb = block_shape(ki.kc, TP_SEG_LARGE_BLOCK);
    b.mb = &merge_shape(ki.kc, TP_SEG_MERGE, lb.tile());
    b.cls[0] = &tb;
    b.cls[1] = &sb;
    b.cls[2] = &lb;
    uint64_t small_max = (uint64_t)tune[ki.kc].v[TP_SEG_SMALL_MAX];
    if (small_max == 0 || small_max > sb.tile()) small_max = sb.tile();
    uint64_t tiny_max = (uint64_t)tune[ki.kc].v[TP_SEG_TINY_MAX];
    if (tiny_max > tb.tile()) tiny_max = tb.tile();
    if (tiny_max > small_max) tiny_max = small_max;
    const uint64_t LT = lb.tile(), MT = b.mb->tile();
    b.LT = LT;
    const uint64_t w = (uint64_t)ki.width;
    const uint64_t plain_min = allow_plain && tune[ki.kc].v[TP_SEG_PLAIN_MIN] > 0
                                   ? (uint64_t)1 << tune[ki.kc].v[TP_SEG_PLAIN_MIN]
                                   : UINT64_MAX;
    b.seg_table = tune[ki.kc].v[TP_PLAIN_CORANK] == 0;
    b.lanes = tune[ki.kc].v[TP_PARTITION_LANES];
    (void)w;
    b.MT = MT;
    // The K-way plan for the large class: every pass of fan-in 2^lmax but a
    // segment's last, so every segment has runs of 2^(rl0 + lmax p) at pass p.
    // Samples of 32-bit keys are u64 records (key, run, index), merged as u64
    // keys; those of 64-bit keys {key, run << 32 | index} records, merged by
    // the stable key-value merge.
    const int lmax = tune[ki.kc].v[TP_KWAY_FANIN] >= 2 ? ilog2((uint64_t)tune[ki.kc].v[TP_KWAY_FANIN]) : 0;
    b.rl0 = (LT & (LT - 1)) == 0 ? ilog2(LT) : 0;
    b.kway = lmax > 0 && b.rl0 >= 11;
    b.skc = ki.kc == KC_32 ? KC_64 : KC_KV;
    b.sw = ki.kc == KC_32 ? 8 : 16;
    if (b.kway) {
        int kcnt = 0;
        const KwayShapeFns *kt = kway_shapes(ki.kc, &kcnt);
        int kidx = tune[ki.kc].v[TP_KWAY_SHAPE];
        if (kidx < 0 || kidx >= kcnt) kidx = 0;
        b.ks = &kt[kidx];
        const int sl0 = std::min(tune[ki.kc].v[TP_KWAY_SAMPLE], b.rl0 - 8);
        b.kway = kway_tiles(1, b.rl0, sl0, 1 << lmax, b.ks->tile(), b.ks->ipt) > 0;
        b.lmax = lmax;
    }
    std::vector<size_t> large;  // indices into b.recs
    for (size_t i = 0; i < nseg; i++) {
        const SegSpec &g = segs[i];
        if (g.len == 0) continue;
        if (!g.in || !g.fin || !g.other || g.fin == g.other)
            return fail(OGS_ERR_ARG, "segment %zu: NULL buffer or fin == other", i);
        if (g.len >= plain_min) {
            b.plain.push_back(i);
            continue;
        }
        SegRecH r{g.in, g.fin, g.other, g.len, 0, 0, 0, 0, 0};
        if (g.len <= tiny_max) {
            r.cls = 0;
            r.tidx = b.ntd[0]++;
        } else if (g.len <= small_max) {
            r.cls = 1;
            r.tidx = b.ntd[1]++;
        } else {
            int p = 0;
            for (uint64_t R = LT; R < g.len; R *= 2) p++;
            r.cls = 2;
            r.tidx = b.ntd[2];
            b.ntd[2] += (g.len + LT - 1) / LT;
            if (b.kway && p > 0) {
                BatchBuild::KSeg k{};
                k.seg = b.recs.size();
                k.passes = (p + lmax - 1) / lmax;
                int rl = b.rl0, smin = 63;
                for (int q = 0; q < k.passes; q++) {
                    k.lv[q] = (uint8_t)(q + 1 < k.passes ? lmax : p - lmax * (k.passes - 1));
                    k.sl[q] = (uint8_t)std::min(tune[ki.kc].v[TP_KWAY_SAMPLE], rl - 8);
                    if (rl - k.sl[q] > 27)
                        return fail(OGS_ERR_ARG, "segment %zu: no sample spacing for runs of 2^%d", i, rl);
                    smin = std::min(smin, (int)k.sl[q]);
                    const uint64_t nt = kway_tiles(g.len, rl, k.sl[q], 1 << k.lv[q], b.ks->tile(), b.ks->ipt);
                    k.tab = std::max<uint64_t>(k.tab, nt << k.lv[q]);
                    rl += k.lv[q];
                }
                k.ns = (g.len + (1ull << smin) - 1) >> smin;  // records of sw bytes
                r.passes = (uint32_t)k.passes;
                r.kseg1 = (uint32_t)b.kseg.size() + 1;
                b.kpasses = std::max(b.kpasses, k.passes);
                b.kseg.push_back(k);
            } else {
                r.passes = (uint32_t)p;
                r.lidx = (uint32_t)large.size();
                b.max_p = std::max(b.max_p, p);
                large.push_back(b.recs.size());
            }
        }
        b.recs.push_back(r);
    }
    b.nlarge = large.size();
    b.pass_start.assign(b.max_p + 1, 0);
    b.pass_cnt.assign(b.max_p + 1, 0);
    b.poff.assign((size_t)b.max_p * b.nlarge, 0);
    for (int p = 0; p < b.max_p; p++) {
        b.pass_start[p] = b.ndesc;
        for (size_t k = 0; k < large.size(); k++) {
            const SegRecH &r = b.recs[large[k]];
            if ((int)r.passes <= p) continue;
            b.poff[(size_t)p * b.nlarge + k] = b.ndesc;
            b.ndesc += (r.len + MT - 1) / MT;
        }
        b.pass_cnt[p] = b.ndesc - b.pass_start[p];
        b.max_pass = std::max<uint64_t>(b.max_pass, b.pass_cnt[p]);
    }
    // one device table: class tables, merge table, then the uploaded part
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    b.bytes = 0;
    for (int c = 0; c < 3; c++) {
        b.o_td[c] = b.bytes;
        b.bytes = align(b.bytes + b.ntd[c] * sizeof(TileDescH));
    }
    b.o_desc = b.bytes;
    b.bytes = align(b.bytes + b.ndesc * sizeof(MergeDescH));
    b.o_seg = b.bytes;
    b.bytes = align(b.bytes + b.recs.size() * sizeof(SegRecH));
    b.o_poff = b.bytes;
    b.bytes = align(b.bytes + b.poff.size() * sizeof(uint64_t));
    if (!b.kseg.empty()) {
        b.o_kseg = b.bytes;
        b.bytes = align(b.bytes + b.kseg.size() * sizeof(KSegH));
        // per pass: the sample merges, then one split + merge launch per fan-in
        const uint64_t C = b.ks->tile();
        int rl = b.rl0;
        for (int p = 0; p < b.kpasses; p++) {
            int lvmax = 0, sl = 0;
            for (const auto &k : b.kseg)
                if (k.passes > p) {
                    lvmax = std::max(lvmax, (int)k.lv[p]);
                    sl = k.sl[p];  // the same for every segment: it depends on rl alone
                }
            // group samples merged by one K-way launch per fan-in when a group's
            // samples fit one K-way tile, else sorted by one u64 tile sort when a
            // tile holds a group (the first pass: groups of 2^lv runs of
            // 2^(rl - sl) samples), else lv levels of 2-way merges
            const uint64_t span = 1ull << (rl - sl + lvmax);
            const KwayShapeFns *kwh =
                tune[ki.kc].v[TP_KWAY_SWHOLE] ? whole_group_shape(b.skc, 1 << lvmax, rl - sl) : nullptr;
            const BlockShapeFns *sbs = nullptr;
            if (!kwh && p == 0 && ki.kc == KC_32) {  // u64 samples are unique: a sort is the merge
                int bcnt = 0;
                const BlockShapeFns *bt = block_shapes(KC_64, &bcnt);
                for (int j = 0; j < bcnt; j++)
                    if (bt[j].tile() >= span && (!sbs || bt[j].tile() < sbs->tile())) sbs = &bt[j];
            }
            if (kwh) {
                for (int lvv = 1; lvv <= lvmax; lvv++) {
                    BatchBuild::KLaunch W;
                    W.pass = p;
                    W.lv = lvv;
                    W.smi = 1;
                    // the smallest shape that holds this fan-in's groups (every
                    // fan-in up to lvmax fits, as the largest does)
                    W.whole = whole_group_shape(b.skc, 1 << lvv, rl - sl);
                    if (!W.whole) W.whole = kwh;
                    W.tfirst.push_back(0);
                    const uint64_t gs = 1ull << (rl - sl + lvv);  // sample records per group
                    for (size_t i = 0; i < b.kseg.size(); i++) {
                        const auto &k = b.kseg[i];
                        if (k.passes <= p || k.lv[p] != lvv) continue;
                        const uint64_t nsp = (b.recs[k.seg].len + (1ull << sl) - 1) >> sl;
                        const uint64_t ng = (nsp + gs - 1) / gs;
                        if (W.tfirst.back() + ng >= (1ull << 32)) return fail(OGS_ERR_ARG, "batch too large for one sample merge launch");
                        W.list.push_back((uint32_t)i);
                        W.tfirst.push_back(W.tfirst.back() + (uint32_t)ng);
                    }
                    if (W.list.empty()) continue;
                    W.o_list = b.bytes;
                    b.bytes = align(b.bytes + W.list.size() * sizeof(uint32_t));
                    W.o_tfirst = b.bytes;
                    b.bytes = align(b.bytes + W.tfirst.size() * sizeof(uint32_t));
                    b.klaunch.push_back(std::move(W));
                }
            } else if (sbs) {
                BatchBuild::SMerge m;
                m.pass = p;
                m.level = -1;
                m.bs = sbs;
                for (const auto &k : b.kseg) {
                    if (k.passes <= p) continue;
                    const uint64_t g = (uint64_t)1 << (rl - sl + k.lv[p]);
                    const uint64_t nsp = (b.recs[k.seg].len + (1ull << sl) - 1) >> sl;
                    m.ndesc += (nsp + g - 1) / g;
                }
                m.o_desc = b.bytes;
                b.bytes = align(b.bytes + m.ndesc * sizeof(TileDescH));
                b.smerge.push_back(m);
            } else {
                int mcnt = 0;
                const MergeShapeFns *mt = merge_shapes(b.skc, &mcnt);
                for (int q = 0; q < lvmax; q++) {
                    BatchBuild::SMerge m;
                    m.pass = p;
                    m.level = q;
                    m.R = 1ull << (rl - sl + q);
                    for (int j = 0; j < mcnt; j++)
                        if (mt[j].tile() <= 2 * m.R && (!m.ms || mt[j].tile() > m.ms->tile())) m.ms = &mt[j];
                    if (!m.ms) return fail(OGS_ERR_ARG, "no sample merge tile divides 2^%d", rl - sl + q + 1);
                    for (const auto &k : b.kseg) {
                        if (k.passes <= p || k.lv[p] <= q) continue;
                        const uint64_t nsp = (b.recs[k.seg].len + (1ull << sl) - 1) >> sl;
                        m.ndesc += (nsp + m.ms->tile() - 1) / m.ms->tile();
                    }
                    m.o_desc = b.bytes;
                    b.bytes = align(b.bytes + m.ndesc * sizeof(MergeDescH));
                    b.smerge.push_back(m);
                }
            }
            for (int lvv = 1; lvv <= lvmax; lvv++) {
                BatchBuild::KLaunch L;
                L.pass = p;
                L.lv = lvv;
                L.smi = kwh || sbs ? 1 : 1 + ((lvv - 1) & 1);
                L.tfirst.push_back(0);
                for (size_t i = 0; i < b.kseg.size(); i++) {
                    const auto &k = b.kseg[i];
                    if (k.passes <= p || k.lv[p] != lvv) continue;
                    const uint64_t nt = kway_tiles(b.recs[k.seg].len, rl, sl, 1 << lvv, C, b.ks->ipt);
                    if (L.tfirst.back() + nt >= (1ull << 32)) return fail(OGS_ERR_ARG, "batch too large for one K-way launch");
                    L.list.push_back((uint32_t)i);
                    L.tfirst.push_back(L.tfirst.back() + (uint32_t)nt);
                }
                if (L.list.empty()) continue;
                L.o_list = b.bytes;
                b.bytes = align(b.bytes + L.list.size() * sizeof(uint32_t));
                L.o_tfirst = b.bytes;
                b.bytes = align(b.bytes + L.tfirst.size() * sizeof(uint32_t));
                b.klaunch.push_back(std::move(L));
            }
            rl += lmax;  // every pass before a segment's last is of the full fan-in
        }
    }
    b.up_end = b.bytes;
    // device-only regions: tile maps, then each segment's samples and split table
    for (auto &L : b.klaunch) {
        L.o_tmap = b.bytes;
        b.bytes = align(b.bytes + (size_t)L.tfirst.back() * sizeof(uint2));
    }
    for (auto &k : b.kseg) {
        k.o_data = b.bytes;
        b.bytes = align(b.bytes + 3 * align(k.ns * b.sw) + k.tab * sizeof(uint64_t));
    }
    if (b.recs.empty()) b.bytes = b.up_end = 0;
    return OGS_OK;
}

void BatchBuild::pack(uint8_t *h, uint8_t *D, const SegSpec *segs) const {
    (void)segs;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    if (!recs.empty()) memcpy(h, recs.data(), recs.size() * sizeof(SegRecH));
    if (!poff.empty()) memcpy(h + (o_poff - o_seg), poff.data(), poff.size() * sizeof(uint64_t));
    if (kseg.empty()) return;
    // the K-way records, with device pointers into the table at D
    std::vector<KSegH> ks(kseg.size());
    for (size_t i = 0; i < kseg.size(); i++) {
        const KSeg &k = kseg[i];
        const SegRecH &r = recs[k.seg];
        KSegH &x = ks[i];
        memset(&x, 0, sizeof(x));
        x.buf[0] = (k.passes & 1) ? r.other : r.fin;
        x.buf[1] = (k.passes & 1) ? r.fin : r.other;
        x.len = r.len;
        const size_t sb = align(k.ns * sw);
        for (int j = 0; j < 3; j++) x.smp[j] = D + k.o_data + j * sb;
        x.tab = (uint64_t *)(D + k.o_data + 3 * sb);
        x.passes = (uint32_t)k.passes;
        x.cout = (uint32_t)ki.codec;
        memcpy(x.lv, k.lv, 8);
        memcpy(x.sl, k.sl, 8);
    }
    memcpy(h + (o_kseg - o_seg), ks.data(), ks.size() * sizeof(KSegH));
    for (const KLaunch &L : klaunch) {
        memcpy(h + (L.o_list - o_seg), L.list.data(), L.list.size() * sizeof(uint32_t));
        memcpy(h + (L.o_tfirst - o_seg), L.tfirst.data(), L.tfirst.size() * sizeof(uint32_t));
    }
    // the sample merges' descriptors
    int rl = rl0;
    for (int p = 0; p < kpasses; p++) {
        for (const SMerge &m : smerge) {
            if (m.pass != p) continue;
            uint8_t *dst = h + (m.o_desc - o_seg);
            size_t o = 0;
            for (size_t i = 0; i < kseg.size(); i++) {
                const KSeg &k = kseg[i];
                if (k.passes <= p) continue;
                const int sl = k.sl[p];
                const uint64_t nsp = (recs[k.seg].len + (1ull << sl) - 1) >> sl;
                char *smp = (char *)ks[i].smp[0];
                const size_t sb = align(k.ns * sw);
                if (m.level < 0) {
                    // the tile sort of each group's samples, smp[0] -> smp[1]
                    const uint64_t g = (uint64_t)1 << (rl - sl + k.lv[p]);
                    for (uint64_t g0 = 0; g0 < nsp; g0 += g) {
                        const uint64_t c = nsp - g0 < g ? nsp - g0 : g;
                        TileDescH t{smp + g0 * sw, smp + sb + g0 * sw, (uint32_t)c, 0};
                        memcpy(dst + o * sizeof(TileDescH), &t, sizeof(t));
                        o++;
                    }
                } else {
                    if (k.lv[p] <= m.level) continue;
                    const int q = m.level;
                    const char *src = smp + (q == 0 ? 0 : (1 + ((q - 1) & 1)) * sb);
                    char *dd = smp + (1 + (q & 1)) * sb;
                    const uint64_t MT = m.ms->tile();
                    for (uint64_t o0 = 0; o0 < nsp; o0 += MT) {
                        MergeDescH md{src, dd, nsp, o0};
                        memcpy(dst + o * sizeof(MergeDescH), &md, sizeof(md));
                        o++;
                    }
                }
            }
        }
        rl += lmax;  // every pass before a segment's last is of the full fan-in
    }
}

// Writes the batch's descriptors into the device table at D (whose records
// and offsets are uploaded).
int Ctx::batch_expand(const BatchBuild &b, void *D, hipStream_t s) {
    uint8_t *d = (uint8_t *)D;
    hipError_t e = launch_expand((const SegRecH *)(d + b.o_seg), b.recs.size(), (TileDescH *)(d + b.o_td[0]),
                                 (TileDescH *)(d + b.o_td[1]), (TileDescH *)(d + b.o_td[2]),
                                 (MergeDescH *)(d + b.o_desc), (const uint64_t *)(d + b.o_poff), b.nlarge, b.LT, b.MT,
                                 (uint32_t)b.ki.width, s);
    if (e != hipSuccess) return hip_fail(e, "descriptor expansion launch");
    return OGS_OK;
}

// Launches a built batch whose descriptor table sits at D (device) and whose
// co-rank table (max_pass entries, when the batch uses one) is `part`.  Issues
// kernel launches only, so it can be captured into a graph.
int Ctx::batch_enqueue(const BatchBuild &b, void *D, uint64_t *part, hipStream_t s) {
    uint8_t *d = (uint8_t *)D;
    const KSegH *kseg = b.kseg.empty() ? nullptr : (const KSegH *)(d + b.o_kseg);
    for (int c = 0; c < 3; c++) {
        if (!b.ntd[c]) continue;
        hipError_t e = b.cls[c]->run_seg(d + b.o_td[c], b.ntd[c], b.ki.codec, b.ki.codec, c == 2 ? kseg : nullptr, s);
        if (e != hipSuccess) return hip_fail(e, "batched tile-sort launch");
    }
    for (int p = 0; p < b.max_p; p++) {
        hipError_t e = b.mb->run_seg(d + b.o_desc + b.pass_start[p] * sizeof(MergeDescH), b.pass_cnt[p], b.LT << p,
                                     b.ki.codec, b.seg_table ? part : nullptr, b.lanes, s);
        if (e != hipSuccess) return hip_fail(e, "batched merge launch");
    }
    // K-way passes of the large class: per pass the group samples merged, then
    // per fan-in the tile map, the splits and the merge
    int rl = b.rl0;
    for (int p = 0; p < b.kpasses; p++) {
        for (const auto &m : b.smerge) {
            if (m.pass != p || !m.ndesc) continue;
            hipError_t e = m.level < 0 ? m.bs->run_seg(d + m.o_desc, m.ndesc, 0, 0, nullptr, s)
                                       : m.ms->run_seg(d + m.o_desc, m.ndesc, m.R, 0, nullptr, b.lanes, s);
            if (e != hipSuccess) return hip_fail(e, "batched sample merge launch");
        }
        for (const auto &L : b.klaunch) {
            if (L.pass != p) continue;
            int sl = 0;
            for (const auto &k : b.kseg)
                if (k.passes > p) sl = k.sl[p];
            hipError_t e = launch_kway_tilemap((const uint32_t *)(d + L.o_list), (const uint32_t *)(d + L.o_tfirst),
                                               (uint32_t)L.list.size(), (uint2 *)(d + L.o_tmap), s);
            if (e == hipSuccess && L.whole)  // the pass's sample merge (listed before its K-way launches)
                e = L.whole->whole_seg[L.lv - 1](kseg, (const uint2 *)(d + L.o_tmap), L.tfirst.back(), p, rl - sl,
                                                 err, s);
            else if (e == hipSuccess)
                e = b.ks->run_seg[L.lv - 1](kseg, (const uint2 *)(d + L.o_tmap), L.tfirst.back(), p, rl, sl, L.smi,
                                            err, s);
            if (e != hipSuccess) return hip_fail(e, "batched K-way pass launch");
        }
        rl += b.lmax;
    }
    return OGS_OK;
}

// Batched sort of independent segments; see batch_build for the size classes.
int Ctx::batch(ogs_key_t kt, const SegSpec *segs, size_t nseg, hipStream_t s) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    BatchBuild b;
    int rc = batch_build(kt, segs, nseg, true, b);
    if (rc) return rc;
    // plain-class segments first (each a full plain sort on this stream)
    for (size_t i : b.plain) {
        const SegSpec &g = segs[i];
        void *res = nullptr;
        rc = sort(kt, g.in, g.fin, g.other, g.len, &res, s);
        if (rc) return rc;
    }
    if (b.bytes == 0) return OGS_OK;
    Workspace *ws_ = workspace(s);
    std::lock_guard<std::mutex> g(ws_->mu);
    rc = ws_->reserve(b.bytes);
    if (rc) return rc;
    if (b.seg_table && b.max_pass) {
        rc = ws_->reserve_part(b.max_pass * sizeof(uint64_t));
        if (rc) return rc;
    }
    // the previous upload from this staging buffer must have been consumed
    e = hipEventSynchronize(ws_->staged);
    if (e != hipSuccess) return hip_fail(e, "staging wait");
    std::vector<uint8_t> &h = ws_->host;
    h.resize(b.up_end - b.o_seg);
    b.pack(h.data(), (uint8_t *)ws_->dev, segs);
    e = hipMemcpyAsync((uint8_t *)ws_->dev + b.o_seg, h.data(), b.up_end - b.o_seg, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(ws_->staged, s);
    if (e != hipSuccess) return hip_fail(e, "segment table upload");
    rc = batch_expand(b, ws_->dev, s);
    if (rc) return rc;
    return batch_enqueue(b, ws_->dev, (uint64_t *)ws_->part, s);
}

// CSR form: segment i = [off[i], off[i+1]) of data; results land in scratch.
int Ctx::segmented(ogs_key_t kt, void *data, void *scratch, const uint64_t *h_off, uint64_t nseg, void **out,
                   hipStream_t s) {
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    if (!out) return fail(OGS_ERR_ARG, "d_sorted is NULL");
    const uint64_t total = nseg ? h_off[nseg] : 0;
    if (total == 0) {
        *out = data;
        return OGS_OK;
    }
    if (!data || !scratch || data == scratch) return fail(OGS_ERR_ARG, "bad data/scratch buffers");
    std::vector<SegSpec> segs;
    int rc = csr_segments(ki, data, scratch, h_off, nseg, segs);
    if (rc) return rc;
    *out = scratch;
    return batch(kt, segs.data(), segs.size(), s);
}

// CSR offsets -> segments sorted from data into scratch (data is the second
// buffer).
int csr_segments(const KeyInfo &ki, void *data, void *scratch, const uint64_t *h_off, uint64_t nseg,
                 std::vector<SegSpec> &segs) {
    const uint64_t total = nseg ? h_off[nseg] : 0;
    const uint64_t w = (uint64_t)ki.width;
    segs.resize(nseg);
    for (uint64_t i = 0; i < nseg; i++) {
        const uint64_t a = h_off[i], b = h_off[i + 1];
        if (b < a || b > total)
            return fail(OGS_ERR_ARG, "offsets not non-decreasing at segment %llu", (unsigned long long)i);
        segs[i] = {(char *)data + a * w, (char *)scratch + a * w, (char *)data + a * w, b - a};
    }
    return OGS_OK;
}

// A captured segmented sort (ogs_segmented_plan).
struct Plan {
    int device = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *dev = nullptr;   // descriptor table
    void *part = nullptr;  // co-rank table
    ~Plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (dev) (void)hipFree(dev);
        if (part) (void)hipFree(part);
    }
};

}  // namespace ogs

using namespace ogs;

extern "C" {

int ogs_ctx_create(ogs_ctx **ctx, int device) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= ndev) return fail(OGS_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    uint32_t *err = nullptr;
    e = hipMalloc(&err, 256);
    if (e == hipSuccess) e = hipMemset(err, 0, 256);
    if (e != hipSuccess) {
        if (err) (void)hipFree(err);
        return fail(OGS_ERR_NOMEM, "device error word: %s", hipGetErrorString(e));
    }
    const int nrc = noalloc_init(device);
    if (nrc) {
        (void)hipFree(err);
        return nrc;
    }
    Ctx *c = new Ctx();
    c->device = device;
    c->err = err;
    for (int k = 0; k < 3; k++) c->tune[k] = default_tuning((KeyClass)k);
    *ctx = reinterpret_cast<ogs_ctx *>(c);
    const int rc = load_env_tiles(*ctx);  // a bad $OGS_TILES fails the call
    if (rc) {
        const std::string msg = ogs_last_error();
        ogs_ctx_destroy(*ctx);
        *ctx = nullptr;
        return fail(rc, "%s", msg.c_str());
    }
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
    if (c->err) (void)hipFree(c->err);
    delete c;
    return OGS_OK;
}

int ogs_device_errors(ogs_ctx *ctx, uint32_t *flags) {
    if (!ctx || !flags) return fail(OGS_ERR_ARG, "NULL argument");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    *flags = 0;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(flags, c->err, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && *flags) {
        const uint32_t zero = 0;
        e = hipMemcpy(c->err, &zero, sizeof(uint32_t), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return hip_fail(e, "device error word");
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

int ogs_sort_soa(ogs_ctx *ctx, ogs_key_t kt, void *d_keys, void *d_values, void *d_work, uint64_t n,
                 void *stream) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    if (ki.width != 16) return fail(OGS_ERR_ARG, "ogs_sort_soa takes a key-value type (got %d)", (int)kt);
    if (n == 0) return OGS_OK;
    if (!d_keys || !d_values || !d_work) return fail(OGS_ERR_ARG, "d_keys, d_values and d_work must be set");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    // the block sort reads the two arrays and the last pass writes them back;
    // the passes between run on 16-byte records in the work space
    char *wa = (char *)d_work, *wb = wa + 16 * n;
    void *res = nullptr;
    return c->sort(kt, d_keys, wa, wb, n, &res, s, d_values, d_keys, d_values);
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

int ogs_segmented_plan(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch, const uint64_t *d_offsets,
                       uint64_t nseg, uint64_t n, ogs_plan **plan, void **d_sorted) {
    if (!ctx || !plan || !d_sorted) return fail(OGS_ERR_ARG, "NULL argument");
    *plan = nullptr;
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    std::vector<uint64_t> off(nseg + 1, 0);
    if (nseg) {
        if (!d_offsets) return fail(OGS_ERR_ARG, "d_offsets is NULL");
        e = hipMemcpy(off.data(), d_offsets, (nseg + 1) * 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "offsets download");
        if (off[0] != 0 || off[nseg] != n)
            return fail(OGS_ERR_ARG, "offsets must start at 0 and end at n=%llu (got %llu..%llu)",
                        (unsigned long long)n, (unsigned long long)off[0], (unsigned long long)off[nseg]);
    }
    Plan *p = new Plan();
    p->device = c->device;
    *d_sorted = n ? d_scratch : d_data;
    if (n == 0) {
        *plan = reinterpret_cast<ogs_plan *>(p);
        return OGS_OK;
    }
    if (!d_data || !d_scratch || d_data == d_scratch) {
        delete p;
        return fail(OGS_ERR_ARG, "bad data/scratch buffers");
    }
    std::vector<SegSpec> segs;
    BatchBuild b;
    int rc = csr_segments(ki, d_data, d_scratch, off.data(), nseg, segs);
    if (!rc) rc = c->batch_build(kt, segs.data(), segs.size(), false, b);
    if (rc) {
        delete p;
        return rc;
    }
    auto bail = [&](hipError_t err, const char *what) {
        delete p;
        return hip_fail(err, what);
    };
    if (b.bytes) {
        e = hipMalloc(&p->dev, b.bytes);
        if (e != hipSuccess) return bail(e, "hipMalloc(plan descriptors)");
        std::vector<uint8_t> h(b.up_end - b.o_seg);
        b.pack(h.data(), (uint8_t *)p->dev, segs.data());
        e = hipMemcpy((uint8_t *)p->dev + b.o_seg, h.data(), b.up_end - b.o_seg, hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(e, "plan segment table upload");
        if (b.seg_table && b.max_pass) {
            e = hipMalloc(&p->part, b.max_pass * sizeof(uint64_t));
            if (e != hipSuccess) return bail(e, "hipMalloc(plan co-rank table)");
        }
        e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking);
        if (e != hipSuccess) return bail(e, "plan stream");
        // the descriptors are written once, before capture
        rc = c->batch_expand(b, p->dev, p->cap);
        if (rc) {
            delete p;
            return rc;
        }
        e = hipStreamSynchronize(p->cap);
        if (e != hipSuccess) return bail(e, "plan descriptor expansion");
        e = hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return bail(e, "hipStreamBeginCapture");
        rc = c->batch_enqueue(b, p->dev, (uint64_t *)p->part, p->cap);
        e = hipStreamEndCapture(p->cap, &p->graph);
        if (rc) {
            delete p;
            return rc;
        }
        if (e != hipSuccess) return bail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return bail(e, "hipGraphInstantiate");
    }
    *plan = reinterpret_cast<ogs_plan *>(p);
    return OGS_OK;
}

int ogs_plan_launch(ogs_plan *plan, void *stream) {
    if (!plan) return fail(OGS_ERR_ARG, "plan is NULL");
    Plan *p = reinterpret_cast<Plan *>(plan);
    if (!p->exec) return OGS_OK;
    hipError_t e = hipSetDevice(p->device);
    if (e == hipSuccess) e = hipGraphLaunch(p->exec, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return OGS_OK;
}

int ogs_plan_destroy(ogs_plan *plan) {
    if (!plan) return OGS_OK;
    Plan *p = reinterpret_cast<Plan *>(plan);
    (void)hipSetDevice(p->device);
    delete p;
    return OGS_OK;
}

int ogs_generate(ogs_ctx *ctx, ogs_key_t kt, int dist, uint64_t seed, uint64_t n, void *d_out, void *stream) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    const int d = dist & ~OGS_GEN_COUNTER;
    if (d < 0 || d >= gen::COUNT) return fail(OGS_ERR_ARG, "unknown distribution %d", dist);
    if (n == 0) return OGS_OK;
    if (!d_out) return fail(OGS_ERR_ARG, "d_out is NULL");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    // temporaries: freed once the stream has finished with them
    std::vector<void *> tmp;
    auto done = [&](int rc) {
        if (hipStreamSynchronize(s) != hipSuccess && rc == OGS_OK) rc = hip_fail(hipGetLastError(), "generate");
        for (void *p : tmp) (void)hipFree(p);
        return rc;
    };
    auto upload = [&](const void *h, size_t bytes, void **dp) {
        hipError_t err = hipMalloc(dp, bytes);
        if (err != hipSuccess) return err;
        tmp.push_back(*dp);
        return hipMemcpy(*dp, h, bytes, hipMemcpyHostToDevice);
    };
    double *d_cdf = nullptr;
    uint64_t *d_starts = nullptr;
    uint64_t nb = 0;
    if (d == gen::ZIPF) {
        std::vector<double> cdf;
        zipf_cdf(cdf);
        e = upload(cdf.data(), cdf.size() * sizeof(double), (void **)&d_cdf);
        if (e != hipSuccess) return done(hip_fail(e, "Zipf table upload"));
    }
    if (d == gen::PARETO_BURSTS || d == gen::PARETO_SHUFFLED) {
        std::vector<uint64_t> st;
        burst_starts(seed, n, st);
        nb = st.size();
        e = upload(st.data(), st.size() * sizeof(uint64_t), (void **)&d_starts);
        if (e != hipSuccess) return done(hip_fail(e, "burst table upload"));
    }
    const int layout = kt == OGS_KV_U32U32 ? kGenKv32 : ki.width;
    if (d != gen::PARETO_SHUFFLED) {
        e = launch_generate(layout, d, seed, n, d_out, d_cdf, d_starts, nb, s);
        return done(e == hipSuccess ? OGS_OK : hip_fail(e, "generator launch"));
    }
    // shuffled bursts: (tag, key) pairs sorted by tag with this library's sort
    void *p0 = nullptr, *p1 = nullptr;
    e = hipMalloc(&p0, n * 16);
    if (e == hipSuccess) {
        tmp.push_back(p0);
        e = hipMalloc(&p1, n * 16);
        if (e == hipSuccess) tmp.push_back(p1);
    }
    if (e != hipSuccess) return done(fail(OGS_ERR_NOMEM, "shuffle pairs (%llu B): %s", (unsigned long long)(n * 32),
                                          hipGetErrorString(e)));
    e = launch_generate_tagged(seed, n, shuffle_bits(n), d_starts, nb, p0, s);
    if (e != hipSuccess) return done(hip_fail(e, "generator launch"));
    void *sorted = nullptr;
    int rc = c->sort(OGS_KV_U64U64, nullptr, p0, p1, n, &sorted, s);
    if (rc) return done(rc);
    e = launch_generate_untag(layout, sorted, n, d_out, s);
    return done(e == hipSuccess ? OGS_OK : hip_fail(e, "generator launch"));
}

int ogs_sort_batch(ogs_ctx *ctx, ogs_key_t kt, const ogs_batch_item *items, uint64_t count, void *stream) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    if (count && !items) return fail(OGS_ERR_ARG, "items is NULL");
    std::vector<SegSpec> segs(count);
    for (uint64_t i = 0; i < count; i++) {
        const ogs_batch_item &it = items[i];
        // a segment whose scratch is NULL may use its input as the second buffer
        void *other = it.scratch ? it.scratch : (void *)it.in;
        if (it.n && other == it.out)
            return fail(OGS_ERR_ARG, "item %llu: scratch (or, without scratch, in) must differ from out",
                        (unsigned long long)i);
        segs[i] = {it.in, it.out, other, it.n};
    }
    return reinterpret_cast<Ctx *>(ctx)->batch(kt, segs.data(), segs.size(), (hipStream_t)stream);
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
    uint32_t flags = 0;
    rc = ogs_device_errors(ctx, &flags);
    if (rc) return rc;
    if (flags)
        return fail(OGS_ERR_STATE, "a merge kernel reported an inconsistent tile split table (device error word 0x%x)",
                    flags);
    return OGS_OK;
}

int ogs_pass_count(ogs_ctx *ctx, ogs_key_t kt, uint64_t n) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    return reinterpret_cast<Ctx *>(ctx)->pass_count(kt, n);
}

const char *ogs_last_error(void) { return g_err.c_str(); }

}  // extern "C"
