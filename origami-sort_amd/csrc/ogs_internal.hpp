// ogs_internal.hpp -- host-side objects behind the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>

#include <map>
#include <mutex>
#include <vector>

#include "../../include/ogs.h"
#include "sort_kernels.hpp"

namespace ogs {

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

struct KeyInfo {
    KeyClass kc;
    int width;  // bytes per item
    int codec;  // 0 identity, 1 sign flip, 2 float totalOrder, 3 swap the 32-bit halves (8-byte records)
};
bool key_info(ogs_key_t kt, KeyInfo *ki);

// Tile table per key class: indices into the compiled shape lists (the
// analogue of the tuned constants in Origami/config.h:45-264).
// Segment size classes: tiny (one wavefront-sized workgroup per segment),
// small (one workgroup per segment), large (TILE-sized pieces + merge passes).
enum TuneParam {
    TP_PLAIN_BLOCK = 0,     // block-sort shape of plain sorts
    TP_PLAIN_MERGE = 1,     // merge shape of plain sorts
    TP_SEG_SMALL_BLOCK = 2, // block-sort shape of the small segment class
    TP_SEG_SMALL_MAX = 3,   // longest segment in the small class (0 = its tile)
    TP_SEG_LARGE_BLOCK = 4, // block-sort shape of segments above the small class
    TP_SEG_MERGE = 5,       // merge shape of segmented passes
    TP_SEG_TINY_BLOCK = 6,  // block-sort shape of the tiny segment class
    TP_SEG_TINY_MAX = 7,    // longest segment in the tiny class (0 = no tiny class)
    TP_PLAIN_CORANK = 8,    // merges: 0 partition kernel + table, 1 co-ranks inside the merge kernel
    TP_SEG_PLAIN_MIN = 9,   // log2 of the shortest batch segment run as a plain sort (0 = never)
    TP_PARTITION_LANES = 10,// lanes per merge tile in the partition kernel: 1, 4, 8 or 16
    TP_PLAIN_FANIN = 11,    // plain merges: 2 (one level per HBM pass) or 4 (two levels)
    TP_QUAD_LANES = 12,     // lanes per tile slot in the 4-way split kernel: 1, 2, 4 or 8
    TP_KWAY_FANIN = 13,     // key-only plain merges: largest fan-in of the sample-split K-way pass (0 = off)
    TP_KWAY_SHAPE = 14,     // tile shape of the K-way merge kernel
    TP_KWAY_SAMPLE = 15,    // log2 of the sample spacing of the K-way pass
    TP_KWAY_SCORANK = 16,   // sample merges of the K-way pass: 0 co-rank table kernel, 1 co-ranks inside the merge kernel
    TP_KWAY_SWHOLE = 17,    // 1: a pass's sample merge is one K-way launch when a group's samples fit one tile
    TP_COUNT = 18
};
struct Tuning {
    int v[TP_COUNT];
};

// Device workspace bound to one stream (co-rank tables, descriptors).  Work
// on one stream is ordered, so reusing the block across calls on the same
// stream is safe; growing it frees the old block (hipFree synchronises).
struct Workspace {
    std::mutex mu;  // one batched call at a time per stream
    void *dev = nullptr;   // descriptor tables
    size_t cap = 0;
    void *part = nullptr;  // co-rank tables
    size_t part_cap = 0;
    int reserve_part(size_t bytes);
    std::vector<uint8_t> host;       // staging for descriptor uploads
    hipEvent_t staged = nullptr;     // recorded after the last upload from `host`
    int reserve(size_t bytes);
    Workspace();
    ~Workspace();
};

// One segment of a batched sort: sorted from `in` into `fin`, with `other`
// as the second ping-pong buffer (`in` may alias either).
struct SegSpec {
    const void *in;
    void *fin;
    void *other;
    uint64_t len;
};

// Per-kernel timing with HIP events recorded on the launching stream
// (ogs_profile_*): kernel kinds 0 block sort, 1 merge partition (both the
// 2-way table, the 4-way split table, and the K-way sample merges + split
// table), 2 2-way merge pass, 3 4-way merge pass, 4 K-way merge pass.
enum ProfKind { PK_BLOCK = 0, PK_PARTITION = 1, PK_MERGE = 2, PK_MERGE4 = 3, PK_KWAY = 4, PK_COUNT = 5 };
struct Profiler {
    bool on = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
        uint64_t items;
        bool own_a;  // false when `a` is another record's end event
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double total_ms[PK_COUNT] = {};
    uint64_t count[PK_COUNT] = {};
    uint64_t items[PK_COUNT] = {};
    hipEvent_t get();
    int flush();  // waits for pending events and accumulates
    void reset();
};

// The launches of one batched sort, worked out on the host (Ctx::batch_build):
// descriptor counts per size class (tiny, small, large) and per merge pass,
// the layout of one device table holding the class tables, the merge table,
// the per-segment records and the per-pass offsets, and the segments that run
// as plain sorts instead.  Only the records and offsets are uploaded (`pack`
// writes them, from byte o_seg on); the descriptor-expansion kernel writes the
// descriptors on the device.
struct BatchBuild {
    KeyInfo ki;
    const BlockShapeFns *cls[3] = {nullptr, nullptr, nullptr};
    const MergeShapeFns *mb = nullptr;
    uint64_t LT = 0, MT = 0;
    uint64_t ntd[3] = {0, 0, 0};
    uint64_t ndesc = 0;
    std::vector<SegRecH> recs;
    std::vector<uint64_t> poff;  // [pass][large segment]
    uint64_t nlarge = 0;
    std::vector<uint64_t> pass_start, pass_cnt;
    int max_p = 0;
    size_t o_td[3] = {0, 0, 0}, o_desc = 0, o_seg = 0, o_poff = 0, bytes = 0;
    uint64_t max_pass = 0;
    bool seg_table = true;
    int lanes = 4;
    std::vector<size_t> plain;
    // Large segments on the sample-split K-way plan (32-bit keys): their
    // records, and per pass the launches by fan-in, the sample merges and the
    // device regions (tile maps, sample arrays, split tables).
    bool kway = false;
    const KwayShapeFns *ks = nullptr;
    struct KSeg {
        size_t seg;              // index into recs
        int passes;
        uint8_t lv[8], sl[8];
        uint64_t ns;             // sample records at the finest spacing
        uint64_t tab;            // split-table entries
        size_t o_data;           // its sample arrays and split table (device-only region)
    };
    std::vector<KSeg> kseg;
    struct KLaunch {             // pass p, fan-in 2^lv: its segments (indices into kseg)
        int pass, lv, smi;
        const KwayShapeFns *whole = nullptr;  // set: the pass's sample merge, one workgroup per group
        std::vector<uint32_t> list, tfirst;
        size_t o_list = 0, o_tfirst = 0, o_tmap = 0;
    };
    std::vector<KLaunch> klaunch;
    struct SMerge {              // one launch of the sample merges: a tile sort or a 2-way level
        int pass, level;         // level -1: the tile sort of each group's samples
        uint64_t R = 0;          // 2-way: run length in samples
        const BlockShapeFns *bs = nullptr;
        const MergeShapeFns *ms = nullptr;
        uint64_t ndesc = 0;
        size_t o_desc = 0;
    };
    std::vector<SMerge> smerge;
    int kpasses = 0;             // the most K-way passes of any segment
    int lmax = 0;                // log2 of the full fan-in
    KeyClass skc = KC_64;        // key class the samples are merged as
    uint64_t sw = 8;             // bytes per sample record
    int rl0 = 0;                 // log2 of the large tile
    size_t o_kseg = 0, up_end = 0;  // uploaded bytes are [o_seg, up_end); the rest is device-only
    // bytes [o_seg, up_end) of the table into h[0 ..), device pointers relative to D
    void pack(uint8_t *h, uint8_t *D, const SegSpec *segs) const;
};

// Caller-provided device workspace for a plain sort (ogs_sort_noalloc): the
// co-rank tables, K-way samples and split tables are carved from it instead of
// the stream's pooled workspace, and nothing is allocated.
struct ExtScratch {
    char *base;
    size_t cap;
};

struct Ctx {
    int device = 0;
    uint32_t *err = nullptr;  // device error word (kErr* bits), read and cleared by ogs_device_errors
    Profiler prof;
    Tuning tune[3];
    std::mutex mu;  // guards ws
    std::map<hipStream_t, Workspace *> ws;
    std::mutex host_mu;  // serialises ogs_sort_host
    void *host_buf[2] = {nullptr, nullptr};
    size_t host_cap = 0;

    Workspace *workspace(hipStream_t s);
    void drop_workspace(hipStream_t s);
    const BlockShapeFns &block_shape(KeyClass kc, int param);
    const MergeShapeFns &merge_shape(KeyClass kc, int param, uint64_t block_tile);
    // in == nullptr: sort data in place (data/scratch ping-pong); otherwise read
    // the input from `in` (left intact) and ping-pong between data and scratch.
    // vin: KV values apart from the keys in `in`; kout/vout: the sorted keys and
    // values written to separate arrays by the last pass (structure of arrays)
    // ext: the workspace comes from ext (no allocation, no profiling events)
    int sort(ogs_key_t kt, const void *in, void *data, void *scratch, uint64_t n, void **out, hipStream_t s,
             const void *vin = nullptr, void *kout = nullptr, void *vout = nullptr, const ExtScratch *ext = nullptr);
    // plain sort with sample-split K-way merge passes (n > block tile)
    int sort_kway(const KeyInfo &ki, const void *in, void *data, void *scratch, uint64_t n, void **out, hipStream_t s,
                  const ExtScratch *ext);
    // merge passes of a K-way plain sort: fills lv[0 .. P) with the levels per
    // pass (largest first) and returns P (at most kMaxPasses)
    static constexpr int kMaxPasses = 64;
    int kway_plan(KeyClass kc, uint64_t n, uint64_t block_tile, int *lv);
    bool kway_on(const KeyInfo &ki, uint64_t n);
    int segmented(ogs_key_t kt, void *data, void *scratch, const uint64_t *h_off, uint64_t nseg, void **out,
                  hipStream_t s);
    int batch(ogs_key_t kt, const SegSpec *segs, size_t nseg, hipStream_t s);
    // allow_plain: long segments may run as plain sorts (not in a captured plan)
    int batch_build(ogs_key_t kt, const SegSpec *segs, size_t nseg, bool allow_plain, BatchBuild &b);
    int batch_expand(const BatchBuild &b, void *D, hipStream_t s);
    int batch_enqueue(const BatchBuild &b, void *D, uint64_t *part, hipStream_t s);
    int pass_count(ogs_key_t kt, uint64_t n);
};

Tuning default_tuning(KeyClass kc);
// The no-allocation scratch array's device address on `device` (the current
// device), looked up once per device (ogs_noalloc.hip); called by
// ogs_ctx_create.
int noalloc_init(int device);
// Applies the tile-table file named by $OGS_TILES, when set (all or nothing).
int load_env_tiles(ogs_ctx *ctx);
int csr_segments(const KeyInfo &ki, void *data, void *scratch, const uint64_t *h_off, uint64_t nseg,
                 std::vector<SegSpec> &segs);
// Counter-based generator tables and host mirror (ogs_host.cpp).
void zipf_cdf(std::vector<double> &cdf);
void burst_starts(uint64_t seed, uint64_t n, std::vector<uint64_t> &starts);
int shuffle_bits(uint64_t n);
int generate_counter_host(const KeyInfo &ki, int dist, uint64_t seed, uint64_t n, void *out);

}  // namespace ogs
