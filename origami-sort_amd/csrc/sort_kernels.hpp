// sort_kernels.hpp -- host-visible launch table of the sort kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ogs {

// Key classes: the kernels only see ordered unsigned keys.
enum KeyClass { KC_32 = 0, KC_64 = 1, KC_KV = 2 };

// Per-tile descriptor of the segmented merge pass (must match the device
// struct): segment base, segment length, tile start; bit 63 of o0 = direction
// (1: scratch -> data), bit 62 = the segment's last pass.
struct SegTileDesc {
    uint64_t seg, len, o0;
};
static constexpr uint64_t kSegDirBit = 1ull << 63;
static constexpr uint64_t kSegLastBit = 1ull << 62;
// Tile-sort descriptor count word: bit 31 = write to the odd buffer, bit 30 =
// final output of its segment, bits 0..29 = item count.
static constexpr uint32_t kTileOddBit = 1u << 31;
static constexpr uint32_t kTileFinalBit = 1u << 30;

typedef hipError_t (*BlockLaunch)(const void *in, void *out, uint64_t n, int cin, int cout, hipStream_t s);
typedef hipError_t (*SegBlockLaunch)(const void *in, void *out_even, void *out_odd, const uint64_t *tile_off,
                                     const uint32_t *tile_cnt, uint64_t ntiles, int cin, int cout, hipStream_t s);
typedef hipError_t (*MergeLaunch)(const void *in, void *out, uint64_t n, uint64_t R, int cout, hipStream_t s);
typedef hipError_t (*SegMergeLaunch)(void *buf0, void *buf1, const void *tiles, uint64_t ntiles, uint64_t R, int cout,
                                     hipStream_t s);

struct BlockShapeFns {
    int block, ipt, wlev;
    BlockLaunch run;
    SegBlockLaunch run_seg;
    uint64_t tile() const { return (uint64_t)block * ipt; }
};
struct MergeShapeFns {
    int block, ipt;
    MergeLaunch run;
    SegMergeLaunch run_seg;
    uint64_t tile() const { return (uint64_t)block * ipt; }
};

const BlockShapeFns *block_shapes(KeyClass kc, int *count);
const MergeShapeFns *merge_shapes(KeyClass kc, int *count);

}  // namespace ogs
