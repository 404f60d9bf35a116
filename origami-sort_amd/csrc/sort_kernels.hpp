// sort_kernels.hpp -- host-visible launch table of the sort kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ogs {

// Key classes: the kernels only see ordered unsigned keys.
enum KeyClass { KC_32 = 0, KC_64 = 1, KC_KV = 2 };

// Bits of a context's device error word (ogs_device_errors): set by a merge
// kernel that finds its tile split table inconsistent and leaves that tile
// unmerged.  Neither can happen with the tables the split kernels write.
static constexpr uint32_t kErrKwaySplit = 1u;  // K-way pass
static constexpr uint32_t kErrQuadSplit = 2u;  // 4-way pass

// Batched-sort descriptors (layouts shared with the device structs).
// Tile sort: sort `cnt & 0x3FFFFFFF` items from `in` to `out`; bit 30 of cnt
// = final output of its segment (undo the key codec).
struct TileDescH {
    const void *in;
    void *out;
    uint32_t cnt;
    uint32_t pad;
};
// Merge pass: one output tile of a segment: src/dst buffers of the segment,
// segment length, tile start (bit 62 = the segment's last pass).
struct MergeDescH {
    const void *src;
    void *dst;
    uint64_t len;
    uint64_t o0;
};
static_assert(sizeof(TileDescH) == 24 && sizeof(MergeDescH) == 32, "descriptor layouts are shared with the kernels");
static constexpr uint64_t kSegLastBit = 1ull << 62;
static constexpr uint32_t kTileFinalBit = 1u << 30;
// A large-class tile of a segment on the K-way plan: the tile sort also writes
// its samples; `pad` is the segment's index in the batch's KSegH array.
static constexpr uint32_t kTileKwayBit = 1u << 29;
static constexpr uint32_t kTileCountMask = (1u << 29) - 1;

// A large segment of a batched sort on the sample-split K-way plan (device
// record).  Pass p reads buf[p & 1] and writes buf[(p + 1) & 1]; buf[0] is the
// tile sort's output, so the last pass lands in the segment's result buffer.
// smp[0] holds the tagged samples of the current runs (written by the kernel
// that wrote the runs), smp[1] and smp[2] their group merges; tab is the split
// table (tiles * fan-in entries of the segment's largest pass).  Passes run
// 2^lv[p] runs of 2^(tile + sum of earlier lv) items into one, with samples
// every 2^sl[p] items.
struct KSegH {
    void *buf[2];
    uint64_t len;
    uint64_t *tab;
    void *smp[3];
    uint32_t passes;
    uint32_t cout;  // output codec of the last pass
    uint8_t lv[8];
    uint8_t sl[8];
};
static_assert(sizeof(KSegH) == 80, "record layout is shared with the kernels");

// One segment of a batched sort as the descriptor-expansion kernel reads it:
// its buffers and length, its size class (0 tiny, 1 small, 2 large), its
// first tile descriptor in that class's table, and for the large class its
// merge-pass count and either its row in the per-pass offset table of 2-way
// merge descriptors or, on the K-way plan, its K-way record (kseg1 = index + 1;
// 0 = 2-way plan).
struct SegRecH {
    const void *in;
    void *fin;
    void *other;
    uint64_t len;
    uint64_t tidx;
    uint32_t cls;
    uint32_t passes;
    uint32_t lidx;
    uint32_t kseg1;
};
// Writes every tile and merge descriptor of a batch on the device (one
// wavefront per segment): td[c] are the class tables, md the merge table,
// poff[q * nlarge + k] the first merge descriptor of large segment k in pass
// q; LT/MT the large-class block and merge tiles, w the item size in bytes.
hipError_t launch_expand(const SegRecH *segs, uint64_t nseg, TileDescH *td0, TileDescH *td1, TileDescH *td2,
                         MergeDescH *md, const uint64_t *poff, uint64_t nlarge, uint64_t LT, uint64_t MT, uint32_t w,
                         hipStream_t s);

// Samples for a sample-split K-way pass: every 2^slog2-th item of each output
// run, tagged with the index of its run within the next pass's group of `kw`
// runs of 2^rlog2 items and with its index in that run over 2^slog2.  32-bit
// keys are stored as u64 key << 32 | run << 28 | index (so u64 order is the
// total order (key, run, index)); 64-bit keys as {key, run << 32 | index}
// records.  smp == nullptr: no samples.
struct SampleSpec {
    void *smp;
    int rlog2;
    int kw;
    int slog2;
};

// vin / vout (KV only; else nullptr): values read from or written to a separate
// array, keys from `in` / to `out` (structure of arrays).  sp (nullable, key-only
// classes): samples of the sorted tiles for the first K-way pass.
typedef hipError_t (*BlockLaunch)(const void *in, void *out, uint64_t n, int cin, int cout, hipStream_t s,
                                  const void *vin, void *vout, const SampleSpec *sp);
// kseg: the batch's K-way segment records, for tiles flagged kTileKwayBit
// (nullptr when there are none).
typedef hipError_t (*SegBlockLaunch)(const void *tiles, uint64_t ntiles, int cin, int cout, const KSegH *kseg,
                                     hipStream_t s);
// part == nullptr: co-ranks found inside the merge kernel; else a table of
// ceil(n / tile) entries that the partition kernel fills first.
// `mid` (nullable) is recorded between the partition and the merge kernel.
// `lanes` = lanes per tile of the partition kernel (1, 4, 8, 16).
typedef hipError_t (*MergeLaunch)(const void *in, void *out, uint64_t n, uint64_t R, int cout, uint64_t *part,
                                  int lanes, hipStream_t s, hipEvent_t mid, void *vout);
// part: ntiles entries for the descriptors' start co-ranks (filled by a
// partition kernel first), or nullptr to search inside the merge kernel.
typedef hipError_t (*SegMergeLaunch)(const void *tiles, uint64_t ntiles, uint64_t R, int cout, uint64_t *part,
                                     int lanes, hipStream_t s);

struct BlockShapeFns {
    int block, ipt, wlev;
    BlockLaunch run;
    SegBlockLaunch run_seg;
    uint64_t tile() const { return (uint64_t)block * ipt; }
};
// One 4-way pass (runs of R -> 4R, tile/2 dividing R): the tile split table
// tab (4 * quad_table_slots entries) is filled, then the merge runs; `mid` is
// recorded between the two; `err` is the context's device error word.
typedef hipError_t (*QuadLaunch)(const void *in, void *out, uint64_t n, uint64_t R, int cout, uint64_t *tab,
                                 int quad_lanes, hipStream_t s, hipEvent_t mid, void *vout, uint32_t *err);
// Tile slots of one 4-way pass (groups of 4R, 1 + 8R/tile slots each).
inline uint64_t quad_table_slots(uint64_t n, uint64_t R, uint64_t tile) {
    return (n + 4 * R - 1) / (4 * R) * (1 + 8 * R / tile);
}

struct MergeShapeFns {
    int block, ipt;
    MergeLaunch run;
    SegMergeLaunch run_seg;
    QuadLaunch run_quad;
    bool quad_ok;  // the 4-way tile (2 * tile items) fits in LDS
    uint64_t tile() const { return (uint64_t)block * ipt; }
};

// Counter-based input generator (gen_kernels.hip): width = the item layout, 4,
// 8 or 16 (u64 KV) bytes, or kGenKv32 for 8-byte {u32 key, u32 value} records;
// cdf: the Zipf table (dist 1), starts/nb: Pareto burst starts (dist 9).
static constexpr int kGenKv32 = 12;
hipError_t launch_generate(int width, int dist, uint64_t seed, uint64_t n, void *out, const double *cdf,
                           const uint64_t *starts, uint64_t nb, hipStream_t s);
// Shuffled bursts: (tag, burst key) pairs, then, once sorted by tag, the keys.
hipError_t launch_generate_tagged(uint64_t seed, uint64_t n, int bits, const uint64_t *starts, uint64_t nb,
                                  void *pairs, hipStream_t s);
hipError_t launch_generate_untag(int width, const void *pairs, uint64_t n, void *out, hipStream_t s);

// One sample-split K-way pass over runs of 2^rlog2 items (kw of them per
// group): the split kernel reads the unmerged samples s0 and the merged
// samples sm (both spaced 2^slog2) and fills tab (kway_tiles * kw entries),
// then the merge kernel writes the output and, when `next` names a buffer, the
// samples of the output runs for the pass after this one.  `mid` is recorded
// between the two kernels; `err` is the context's device error word.
typedef hipError_t (*KwayLaunch)(const void *in, void *out, uint64_t n, int rlog2, int slog2, int cout,
                                 uint64_t *tab, const void *s0, const void *sm, const SampleSpec *next, hipStream_t s,
                                 hipEvent_t mid, uint32_t *err);
// Pass `pass` of the K-way plan for many segments at once, all of fan-in KW
// at that pass (the launch's kernel): tmap[i] = {segment index into segs, tile
// index within the segment} for each of the launch's ntiles tiles; runs of
// 2^rlog2 items and samples every 2^slog2 items in every listed segment; the
// merged samples are in smp[smi].  The split kernel, then the merge kernel.
typedef hipError_t (*SegKwayLaunch)(const KSegH *segs, const uint2 *tmap, uint64_t ntiles, int pass, int rlog2,
                                    int slog2, int smi, uint32_t *err, hipStream_t s);
// Every group of kw runs of 2^rlog2 items merged by one workgroup (no split
// table; the padded group must fit one tile).
typedef hipError_t (*KwayWholeLaunch)(const void *in, void *out, uint64_t n, int rlog2, uint32_t *err,
                                      hipStream_t s);
// Batched form: the samples of every listed segment's groups merged, one
// workgroup per group (tmap: {segment, group}).
typedef hipError_t (*SegKwayWholeLaunch)(const KSegH *segs, const uint2 *tmap, uint64_t ngroups, int pass,
                                         int rlog2, uint32_t *err, hipStream_t s);
struct KwayShapeFns {
    int block, ipt;
    KwayLaunch run[5];  // fan-in 2, 4, 8, 16, 32
    SegKwayLaunch run_seg[5];
    KwayWholeLaunch whole[5];
    SegKwayWholeLaunch whole_seg[5];
    uint64_t tile() const { return (uint64_t)block * ipt; }
};
// Writes tmap[tfirst[i] + k] = {list[i], k} for k < tfirst[i + 1] - tfirst[i].
hipError_t launch_kway_tilemap(const uint32_t *list, const uint32_t *tfirst, uint32_t nlist, uint2 *tmap,
                               hipStream_t s);
// Tiles of one K-way pass (0 when the shape cannot hold kw slices at this
// sample spacing).
uint64_t kway_tiles(uint64_t n, int rlog2, int slog2, int kw, uint64_t tile, int ipt);
const KwayShapeFns *kway_shapes(KeyClass kc, int *count);

const BlockShapeFns *block_shapes(KeyClass kc, int *count);
const MergeShapeFns *merge_shapes(KeyClass kc, int *count);

}  // namespace ogs
