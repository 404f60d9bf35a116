// sort_kernels.hip -- gfx950 kernels of the mergesort and their launchers.
//
//   k_block_sort      phase 1-2 of the reference (phase2_sort, sorter.h:3831-3849):
//                     one workgroup sorts a tile in registers + LDS.
//   k_merge_pass      one 2-way pass of merge_series (merger.h:4079-4228) as a
//                     merge-path merge staged in LDS; each workgroup finds its own
//                     co-ranks with a wave-wide search (Partition_Ptrs,
//                     Partition.h:117-166, specialised to two runs).
//   k_seg_*           the same kernels driven by per-tile descriptors, for batches
//                     of independent segments (uniform chunks in the reference:
//                     sort_single_thread's sort_n, sorter.h:3857-3868).
#include "sort_core.hpp"
#include "sort_kernels.hpp"

namespace ogs {

// ----------------------------------------------------------------------------
// Tile sort: sorts cnt (<= TILE) items in[base .. base+cnt) into out[base ..].
// The tile is padded with maximum keys.  In KV mode a partial tile whose real
// items include the maximum key would let pads displace real values, so those
// items are written straight to the tail of the tile first (their key is the
// largest there is, so they belong at the end in any order).
// ----------------------------------------------------------------------------
template <typename K, bool HV, int BLOCK, int IPT, int WLEV>
__device__ __forceinline__ void tile_sort(const void *in, void *out, uint64_t base, int cnt, int cin, int cout,
                                          K *sk, uint64_t *sv, int *smisc) {
    constexpr int TILE = BLOCK * IPT;
    constexpr K KMAX = KeyMax<K>::v;
    const int tid = threadIdx.x, lane = tid & 63;
    K r[IPT];
    uint64_t v[IPT];
    int nmax = 0;
#pragma unroll
    for (int i = 0; i < IPT; i++) {
        const int e = i * BLOCK + tid;
        if (e < cnt) {
            K k;
            uint64_t vv = 0;
            GMem<K, HV>::load(in, base + e, k, vv);
            r[i] = enc_key(k, cin);
            v[i] = vv;
            if constexpr (HV) nmax += (r[i] == KMAX);
        } else {
            r[i] = KMAX;
            v[i] = 0;
        }
    }
    int lim = cnt;
    if constexpr (HV) {
        if (cnt < TILE) {
            if (tid == 0) { smisc[0] = 0; smisc[1] = 0; }
            __syncthreads();
            if (nmax) atomicAdd(&smisc[0], nmax);
            __syncthreads();
            const int total = smisc[0];
            lim = cnt - total;
            if (total) {
#pragma unroll
                for (int i = 0; i < IPT; i++) {
                    const int e = i * BLOCK + tid;
                    if (e < cnt && r[i] == KMAX) {
                        int slot = atomicAdd(&smisc[1], 1);
                        GMem<K, HV>::store(out, base + lim + slot, dec_key(r[i], cout), v[i]);
                    }
                }
            }
        }
    }
    lane_network<IPT, HV>(r, v);
    wave_sort_levels<WLEV, IPT, HV>(r, v, lane);
#pragma unroll 1
    for (int R = IPT << WLEV; R < TILE; R <<= 1) {
        __syncthreads();
        lds_put<IPT, HV>(sk, sv, tid * IPT, r, v);
        __syncthreads();
        const int o = tid * IPT;
        const int pb = o & ~(2 * R - 1);
        const int diag = o - pb;
        lds_merge_chains<IPT, OGS_MERGE_CHAINS, HV>(sk, sv, pb, R, pb + R, R, diag, IPT, TILE - 1, r, v);
    }
    __syncthreads();
    lds_put<IPT, HV>(sk, sv, tid * IPT, r, v);
    __syncthreads();
    lds_store_tile<BLOCK, TILE, HV>(sk, sv, out, base, lim, cout);
}

template <typename K, bool HV, int BLOCK, int IPT, int WLEV>
__global__ __launch_bounds__(BLOCK) void k_block_sort(const void *in, void *out, uint64_t n, int cin, int cout) {
    constexpr int TILE = BLOCK * IPT;
    __shared__ __attribute__((aligned(16))) K sk[TILE];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? TILE : 2];
    __shared__ int smisc[2];
    const uint64_t base = (uint64_t)blockIdx.x * TILE;
    const int cnt = (int)(n - base < (uint64_t)TILE ? n - base : (uint64_t)TILE);
    tile_sort<K, HV, BLOCK, IPT, WLEV>(in, out, base, cnt, cin, cout, sk, sv, smisc);
}

// Batched tile sort: descriptor i sorts one tile -- a whole small segment, or
// one TILE-sized piece of a large segment -- from `in` to `out` (they may be
// the same buffer: a tile is read whole before it is written).  Bit 30 of
// cnt marks a tile that is its segment's final output (undo the codec).
struct TileDesc {
    const void *in;
    void *out;
    uint32_t cnt;
    uint32_t pad;
};
static_assert(sizeof(TileDesc) == 24, "descriptor layout is shared with the host");

template <typename K, bool HV, int BLOCK, int IPT, int WLEV>
__global__ __launch_bounds__(BLOCK) void k_seg_tile_sort(const TileDesc *__restrict__ tiles, int cin, int cout) {
    constexpr int TILE = BLOCK * IPT;
    __shared__ __attribute__((aligned(16))) K sk[TILE];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? TILE : 2];
    __shared__ int smisc[2];
    const TileDesc t = tiles[blockIdx.x];
    const int cnt = (int)(t.cnt & 0x3FFFFFFFu);
    tile_sort<K, HV, BLOCK, IPT, WLEV>(t.in, t.out, 0, cnt, cin, ((t.cnt >> 30) & 1u) ? cout : 0, sk, sv, smisc);
}

// ----------------------------------------------------------------------------
// 2-way merge pass.  Runs of length R at multiples of 2R (relative to a segment
// base) are merged pairwise; output tile j of a segment covers items
// [j*TILE, j*TILE+TILE).
// ----------------------------------------------------------------------------
struct PairGeom {
    uint64_t pb, la, lb;  // pair base (relative), run lengths
};
__device__ __forceinline__ PairGeom pair_geom(uint64_t o, uint64_t n, uint64_t R) {
    PairGeom g;
    g.pb = (o / (2 * R)) * (2 * R);
    const uint64_t rest = n - g.pb;
    g.la = rest < R ? rest : R;
    g.lb = rest > R ? (rest - R < R ? rest - R : R) : 0;
    return g;
}

// Co-rank of every merge tile's start, LANES lanes per tile: each step tests
// LANES evenly spaced candidates of the merge-path predicate at once and keeps
// the bracket the group's ballot bits point to (a 2^27-item run resolves in
// seven dependent steps with 16 lanes).  part[j] = A items among the first
// j*tile outputs of tile j's pair.
template <typename K, bool HV, int LANES>
__device__ __forceinline__ uint64_t group_corank(const void *in, uint64_t abase, uint64_t la, uint64_t bbase,
                                                 uint64_t lb, uint64_t diag, int l, int gshift) {
    uint64_t lo = diag > lb ? diag - lb : 0;
    uint64_t hi = diag < la ? diag : la;
    while (hi - lo > (uint64_t)LANES) {
        const uint64_t span = hi - lo;
        const uint64_t m = lo + (span * (uint64_t)(l + 1)) / (LANES + 1);
        const K a = GMem<K, HV>::load_key(in, abase + m);
        const K b = GMem<K, HV>::load_key(in, bbase + diag - 1 - m);
        const uint64_t bits = (__ballot(b < a) >> gshift) & ((1ull << LANES) - 1);
        if (!bits) {
            lo = lo + (span * LANES) / (LANES + 1) + 1;
        } else {
            const int f = __ffsll((unsigned long long)bits) - 1;
            const uint64_t mf = lo + (span * (uint64_t)(f + 1)) / (LANES + 1);
            lo = f ? lo + (span * (uint64_t)f) / (LANES + 1) + 1 : lo;
            hi = mf;
        }
    }
    bool p = false;
    if ((uint64_t)l < hi - lo) {
        const uint64_t m = lo + l;
        p = GMem<K, HV>::load_key(in, bbase + diag - 1 - m) < GMem<K, HV>::load_key(in, abase + m);
    }
    const uint64_t bits = (__ballot(p) >> gshift) & ((1ull << LANES) - 1);
    return bits ? lo + (uint64_t)(__ffsll((unsigned long long)bits) - 1) : hi;
}

template <typename K, bool HV, int LANES>
__global__ void k_merge_partition(const void *__restrict__ in, uint64_t n, uint64_t R, uint32_t tile,
                                  uint64_t ntiles, uint64_t *__restrict__ part) {
    const uint64_t gt = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = (int)(gt % LANES);
    const int gshift = (int)(threadIdx.x & 63) & ~(LANES - 1);
    uint64_t j = gt / LANES;
    const bool live = j < ntiles;
    if (!live) j = ntiles - 1;
    const uint64_t o = j * tile;
    const PairGeom g = pair_geom(o, n, R);
    const uint64_t a = group_corank<K, HV, LANES>(in, g.pb, g.la, g.pb + g.la, g.lb, o - g.pb, l, gshift);
    if (live && l == 0) part[j] = a;
}

// Merges one output tile: the co-ranks of the tile's start and end come from
// the partition table (part != nullptr: part[0] start, part[1] next tile's
// start) or, without a table, from waves 0 and 1 searching in place.  The
// tile's A and B slices are loaded into LDS as one coalesced stream, each
// thread merges IPT outputs along its merge path, and the tile is written back
// through LDS with lane-consecutive stores.
template <typename K, bool HV, int BLOCK, int IPT>
__device__ __forceinline__ void merge_tile(const void *__restrict__ in, void *__restrict__ out, uint64_t seg,
                                           uint64_t n, uint64_t R, uint64_t o0, int cout, K *sk, uint64_t *sv,
                                           uint64_t *sco, const uint64_t *part) {
    static_assert(BLOCK >= 128, "the co-rank searches need two waves");
    constexpr int TILE = BLOCK * IPT;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int cnt = (int)(n - o0 < (uint64_t)TILE ? n - o0 : (uint64_t)TILE);
    const PairGeom g = pair_geom(o0, n, R);
    const uint64_t d0 = o0 - g.pb, d1 = d0 + cnt;
    const uint64_t abase = seg + g.pb, bbase = seg + g.pb + g.la;
    uint64_t a0, a1;
    if (part) {
        a0 = part[0];
        a1 = (d1 == g.la + g.lb) ? g.la : part[1];
    } else {
        if (wave == 0) {
            const uint64_t a = wave_corank<K, HV>(in, abase, g.la, bbase, g.lb, d0, lane);
            if (lane == 0) sco[0] = a;
        } else if (wave == 1) {
            const uint64_t a =
                (d1 == g.la + g.lb) ? g.la : wave_corank<K, HV>(in, abase, g.la, bbase, g.lb, d1, lane);
            if (lane == 0) sco[1] = a;
        }
        __syncthreads();
        a0 = sco[0];
        a1 = sco[1];
    }
    const int la = (int)(a1 - a0);
    const int lb = cnt - la;
    const uint64_t A = abase + a0;
    const uint64_t B = bbase + (d0 - a0);
#pragma unroll 4
    for (int e = tid; e < cnt; e += BLOCK) {
        const uint64_t src = e < la ? A + e : B + (e - la);
        K k;
        uint64_t vv = 0;
        GMem<K, HV>::load(in, src, k, vv);
        sk[lsw<K>(e)] = k;
        if constexpr (HV) sv[lsw<uint64_t>(e)] = vv;
    }
    __syncthreads();
    const int d = tid * IPT;
    K r[IPT];
    uint64_t v[IPT];
    if (d < cnt) {
        const int c = cnt - d < IPT ? cnt - d : IPT;
        lds_merge_chains<IPT, OGS_MERGE_CHAINS, HV>(sk, sv, 0, la, la, lb, d, c, cnt - 1, r, v);
    }
    __syncthreads();
    if (d < cnt) lds_put<IPT, HV>(sk, sv, d, r, v);
    __syncthreads();
    lds_store_tile<BLOCK, TILE, HV>(sk, sv, out, seg + o0, cnt, cout);
}

template <typename K, bool HV, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void k_merge_pass(const void *__restrict__ in, void *__restrict__ out, uint64_t n,
                                                      uint64_t R, int cout, const uint64_t *__restrict__ part) {
    constexpr int TILE = BLOCK * IPT;
    __shared__ __attribute__((aligned(16))) K sk[TILE];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? TILE : 2];
    __shared__ uint64_t sco[2];
    merge_tile<K, HV, BLOCK, IPT>(in, out, 0, n, R, (uint64_t)blockIdx.x * TILE, cout, sk, sv, sco,
                                  part ? part + blockIdx.x : nullptr);
}

// Batched form: descriptor t = one output tile of one segment that still has
// runs of length R to merge: the segment's source and destination buffers,
// its length, and the tile start (bit 62 marks the segment's last pass, when
// the output codec is undone).  Segments with different pass counts simply
// carry different buffers.
struct MergeDesc {
    const void *src;
    void *dst;
    uint64_t len;
    uint64_t o0;
};
static_assert(sizeof(MergeDesc) == 32, "descriptor layout is shared with the host");
static constexpr uint64_t kLastBit = 1ull << 62;

template <typename K, bool HV, int LANES>
__global__ void k_seg_merge_partition(const MergeDesc *__restrict__ tiles, uint64_t ntiles, uint64_t R,
                                      uint64_t *__restrict__ part) {
    const uint64_t gt = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = (int)(gt % LANES);
    const int gshift = (int)(threadIdx.x & 63) & ~(LANES - 1);
    uint64_t t = gt / LANES;
    const bool live = t < ntiles;
    if (!live) t = ntiles - 1;
    const MergeDesc d = tiles[t];
    const uint64_t o0 = d.o0 & ~kLastBit;
    const PairGeom g = pair_geom(o0, d.len, R);
    const uint64_t a = group_corank<K, HV, LANES>(d.src, g.pb, g.la, g.pb + g.la, g.lb, o0 - g.pb, l, gshift);
    if (live && l == 0) part[t] = a;
}

// part[t] = co-rank of descriptor t's tile start (the next descriptor of the
// same segment supplies the end), or nullptr to search inside the kernel.
template <typename K, bool HV, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void k_seg_merge_pass(const MergeDesc *__restrict__ tiles, uint64_t R, int cout,
                                                          const uint64_t *__restrict__ part) {
    constexpr int TILE = BLOCK * IPT;
    __shared__ __attribute__((aligned(16))) K sk[TILE];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? TILE : 2];
    __shared__ uint64_t sco[2];
    const MergeDesc d = tiles[blockIdx.x];
    const int co = (d.o0 & kLastBit) ? cout : 0;
    const uint64_t o0 = d.o0 & ~kLastBit;
    merge_tile<K, HV, BLOCK, IPT>(d.src, d.dst, 0, d.len, R, o0, co, sk, sv, sco,
                                  part ? part + blockIdx.x : nullptr);
}

// ----------------------------------------------------------------------------
// Launch table.  One instantiation per (key class, tile shape); the host picks
// the tile shape per size class from the tuning table.
// ----------------------------------------------------------------------------
template <typename K, bool HV, int BLOCK, int IPT, int WLEV>
static hipError_t launch_block(const void *in, void *out, uint64_t n, int cin, int cout, hipStream_t s) {
    constexpr uint64_t TILE = (uint64_t)BLOCK * IPT;
    const uint64_t grid = (n + TILE - 1) / TILE;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((k_block_sort<K, HV, BLOCK, IPT, WLEV>), dim3((unsigned)grid), dim3(BLOCK), 0, s, in, out, n,
                       cin, cout);
    return hipGetLastError();
}

template <typename K, bool HV, int BLOCK, int IPT, int WLEV>
static hipError_t launch_seg_block(const void *tiles, uint64_t ntiles, int cin, int cout, hipStream_t s) {
    if (ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL((k_seg_tile_sort<K, HV, BLOCK, IPT, WLEV>), dim3((unsigned)ntiles), dim3(BLOCK), 0, s,
                       (const TileDesc *)tiles, cin, cout);
    return hipGetLastError();
}

// part == nullptr: co-ranks found inside the merge kernel; else `part` (ntiles
// entries) is filled by the partition kernel first.
// Partition launch with `lanes` lanes per tile (1, 4, 8 or 16): fewer lanes
// issue fewer loads, more lanes take fewer dependent steps.
template <typename K, bool HV, int LANES>
static void launch_partition_l(const void *in, uint64_t n, uint64_t R, uint32_t tile, uint64_t ntiles,
                               uint64_t *part, hipStream_t s) {
    const uint64_t threads = ntiles * LANES;
    hipLaunchKernelGGL((k_merge_partition<K, HV, LANES>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s,
                       in, n, R, tile, ntiles, part);
}
template <typename K, bool HV>
static void launch_partition(int lanes, const void *in, uint64_t n, uint64_t R, uint32_t tile, uint64_t ntiles,
                             uint64_t *part, hipStream_t s) {
    switch (lanes) {
    case 1: launch_partition_l<K, HV, 1>(in, n, R, tile, ntiles, part, s); break;
    case 4: launch_partition_l<K, HV, 4>(in, n, R, tile, ntiles, part, s); break;
    case 8: launch_partition_l<K, HV, 8>(in, n, R, tile, ntiles, part, s); break;
    default: launch_partition_l<K, HV, 16>(in, n, R, tile, ntiles, part, s); break;
    }
}
template <typename K, bool HV, int LANES>
static void launch_seg_partition_l(const void *tiles, uint64_t ntiles, uint64_t R, uint64_t *part, hipStream_t s) {
    const uint64_t threads = ntiles * LANES;
    hipLaunchKernelGGL((k_seg_merge_partition<K, HV, LANES>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       s, (const MergeDesc *)tiles, ntiles, R, part);
}
template <typename K, bool HV>
static void launch_seg_partition(int lanes, const void *tiles, uint64_t ntiles, uint64_t R, uint64_t *part,
                                 hipStream_t s) {
    switch (lanes) {
    case 1: launch_seg_partition_l<K, HV, 1>(tiles, ntiles, R, part, s); break;
    case 4: launch_seg_partition_l<K, HV, 4>(tiles, ntiles, R, part, s); break;
    case 8: launch_seg_partition_l<K, HV, 8>(tiles, ntiles, R, part, s); break;
    default: launch_seg_partition_l<K, HV, 16>(tiles, ntiles, R, part, s); break;
    }
}

template <typename K, bool HV, int BLOCK, int IPT>
static hipError_t launch_merge(const void *in, void *out, uint64_t n, uint64_t R, int cout, uint64_t *part, int lanes,
                               hipStream_t s, hipEvent_t mid) {
    constexpr uint64_t TILE = (uint64_t)BLOCK * IPT;
    const uint64_t ntiles = (n + TILE - 1) / TILE;
    if (ntiles == 0) return hipSuccess;
    if (part) launch_partition<K, HV>(lanes, in, n, R, (uint32_t)TILE, ntiles, part, s);
    if (mid) (void)hipEventRecord(mid, s);
    hipLaunchKernelGGL((k_merge_pass<K, HV, BLOCK, IPT>), dim3((unsigned)ntiles), dim3(BLOCK), 0, s, in, out, n, R,
                       cout, (const uint64_t *)part);
    return hipGetLastError();
}

template <typename K, bool HV, int BLOCK, int IPT>
static hipError_t launch_seg_merge(const void *tiles, uint64_t ntiles, uint64_t R, int cout, uint64_t *part,
                                   int lanes, hipStream_t s) {
    if (ntiles == 0) return hipSuccess;
    if (part) launch_seg_partition<K, HV>(lanes, tiles, ntiles, R, part, s);
    hipLaunchKernelGGL((k_seg_merge_pass<K, HV, BLOCK, IPT>), dim3((unsigned)ntiles), dim3(BLOCK), 0, s,
                       (const MergeDesc *)tiles, R, cout, (const uint64_t *)part);
    return hipGetLastError();
}


// ----------------------------------------------------------------------------
// 4-way merge pass: groups of four runs A, B, C, D (length R each) become one
// run of 4R in a single HBM pass -- two levels of the reference's merge tree
// (MergeTreeEven, merge_tree.h:2576-3010) per read and write of the data.
// The output tile [r, r+T) of a group needs the exact split of all four runs.
// With X = merge(A, B) and Y = merge(C, D), the split is the merge-path
// co-rank x of (X, Y) at r followed by the co-ranks of (A, B) at x and of
// (C, D) at r - x.  X and Y are never materialised: a 2-way co-rank table Q
// sampled every S outputs (the 2-way partition kernel run with tile S) gives
// X[jS] and Y[qS-1] directly, so a coarse search over multiples of S brackets
// x to one S-window, and a fine search inside it evaluates X[x] and Y[y]
// through co-ranks confined to one sample interval (<= log2(S)+1 steps).
// ----------------------------------------------------------------------------
struct QuadGeom {
    uint64_t g4;              // group base
    uint64_t la, lb, lc, ld;  // run lengths (ragged in the last group)
};
__device__ __forceinline__ QuadGeom quad_geom(uint64_t o, uint64_t n, uint64_t R) {
    QuadGeom q;
    q.g4 = (o / (4 * R)) * (4 * R);
    const uint64_t L = n - q.g4 < 4 * R ? n - q.g4 : 4 * R;
    auto clip = [&](uint64_t s) { return L > s ? (L - s < R ? L - s : R) : 0; };
    q.la = clip(0);
    q.lb = clip(R);
    q.lc = clip(2 * R);
    q.ld = clip(3 * R);
    return q;
}

// co-rank(x) of merge(A, B) (ties to A), confined by the sample table:
// Q[qbase + j] = co-rank at j*S for j*S < la + lb.
template <typename K, bool HV>
__device__ __forceinline__ uint64_t pair_corank_win(const void *in, uint64_t abase, uint64_t la, uint64_t bbase,
                                                    uint64_t lb, const uint64_t *Q, uint64_t qbase, uint32_t S,
                                                    uint64_t x) {
    const uint64_t js = x / S;
    const uint64_t alo = Q[qbase + js];
    uint64_t u = (js + 1) * S, au;
    if (u >= la + lb) {
        u = la + lb;
        au = la;
    } else {
        au = Q[qbase + js + 1];
    }
    // co-rank grows by 0 or 1 per output
    uint64_t lo = alo, hi = au;
    if (au > u - x && au - (u - x) > lo) lo = au - (u - x);
    if (alo + (x - js * S) < hi) hi = alo + (x - js * S);
    if (x > lb && x - lb > lo) lo = x - lb;
    if (x < hi) hi = x;
    if (la < hi) hi = la;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        const K a = GMem<K, HV>::load_key(in, abase + mid);
        const K b = GMem<K, HV>::load_key(in, bbase + x - 1 - mid);
        if (!(b < a)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename K, bool HV>
__device__ __forceinline__ K key_after_split(const void *in, uint64_t abase, uint64_t la, uint64_t bbase, uint64_t lb,
                                             uint64_t a, uint64_t b) {
    // the next output after taking a items of A and b items of B
    if (a >= la) return GMem<K, HV>::load_key(in, bbase + b);
    if (b >= lb) return GMem<K, HV>::load_key(in, abase + a);
    const K ka = GMem<K, HV>::load_key(in, abase + a);
    const K kb = GMem<K, HV>::load_key(in, bbase + b);
    return (kb < ka) ? kb : ka;
}

template <typename K, bool HV>
__device__ __forceinline__ K key_before_split(const void *in, uint64_t abase, uint64_t bbase, uint64_t a, uint64_t b) {
    // the last output after taking a items of A and b items of B (a + b > 0)
    if (a == 0) return GMem<K, HV>::load_key(in, bbase + b - 1);
    if (b == 0) return GMem<K, HV>::load_key(in, abase + a - 1);
    const K ka = GMem<K, HV>::load_key(in, abase + a - 1);
    const K kb = GMem<K, HV>::load_key(in, bbase + b - 1);
    return (kb < ka) ? ka : kb;
}

// Smallest index in [lo, hi) whose predicate holds, else hi, searched by a
// group of LANES lanes (monotone predicate; all lanes of the group call it).
template <int LANES, typename Pred>
__device__ __forceinline__ uint64_t group_first_true(uint64_t lo, uint64_t hi, int l, int gshift, Pred pred) {
    while (hi - lo > (uint64_t)LANES) {
        const uint64_t span = hi - lo;
        const uint64_t m = lo + (span * (uint64_t)(l + 1)) / (LANES + 1);
        const uint64_t bits = (__ballot(pred(m)) >> gshift) & ((1ull << LANES) - 1);
        if (!bits) {
            lo = lo + (span * LANES) / (LANES + 1) + 1;
        } else {
            const int f = __ffsll((unsigned long long)bits) - 1;
            const uint64_t mf = lo + (span * (uint64_t)(f + 1)) / (LANES + 1);
            lo = f ? lo + (span * (uint64_t)f) / (LANES + 1) + 1 : lo;
            hi = mf;
        }
    }
    bool p = false;
    if ((uint64_t)l < hi - lo) p = pred(lo + l);
    const uint64_t bits = (__ballot(p) >> gshift) & ((1ull << LANES) - 1);
    return bits ? lo + (uint64_t)(__ffsll((unsigned long long)bits) - 1) : hi;
}

// One group of LANES lanes per 4-way output tile: tab[3t..3t+2] = (a, x, c),
// the A items, X = A+B items and C items before the tile's first output.
template <typename K, bool HV, int LANES>
__global__ void k_quad_partition(const void *__restrict__ in, uint64_t n, uint64_t R, uint32_t T, uint32_t S,
                                 uint64_t ntiles, const uint64_t *__restrict__ Q, uint64_t *__restrict__ tab) {
    const uint64_t gt = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = (int)(gt % LANES);
    const int gshift = (int)(threadIdx.x & 63) & ~(LANES - 1);
    uint64_t t = gt / LANES;
    const bool live = t < ntiles;
    if (!live) t = ntiles - 1;
    const uint64_t o = t * T;
    const QuadGeom q = quad_geom(o, n, R);
    const uint64_t r = o - q.g4;
    const uint64_t lx = q.la + q.lb, ly = q.lc + q.ld;
    const uint64_t A = q.g4, B = q.g4 + q.la, C = q.g4 + lx, D = q.g4 + lx + q.lc;
    const uint64_t qab = q.g4 / S, qcd = (q.g4 + 2 * R) / S;
    const uint64_t xlo = r > ly ? r - ly : 0, xhi = r < lx ? r : lx;
    uint64_t x = xhi;
    if (xlo < xhi) {
        // coarse: candidates x = jS in [xlo, xhi); P(jS): X[jS] > Y[r - jS - 1]
        const uint64_t jlo = (xlo + S - 1) / S, jhi1 = (xhi - 1) / S + 1;  // j in [jlo, jhi1)
        uint64_t wlo = xlo, whi = xhi;
        if (jlo < jhi1) {
            auto coarse = [&](uint64_t j) {
                const uint64_t xs = j * S;
                const uint64_t a = Q[qab + j];
                const K kx = key_after_split<K, HV>(in, A, q.la, B, q.lb, a, xs - a);
                const uint64_t ys = r - xs;  // multiple of S, 1 <= ys <= ly
                const uint64_t c = ys >= ly ? q.lc : Q[qcd + ys / S];
                const K ky = key_before_split<K, HV>(in, C, D, c, ys - c);
                return ky < kx;
            };
            const uint64_t jt = group_first_true<LANES>(jlo, jhi1, l, gshift, coarse);
            if (jt < jhi1) {
                whi = jt * S;
                wlo = jt > jlo ? (jt - 1) * S + 1 : xlo;
            } else {
                wlo = (jhi1 - 1) * S + 1;
            }
        }
        // fine: smallest x in [wlo, whi) with X[x] > Y[r - 1 - x], else whi
        auto fine = [&](uint64_t xx) {
            const uint64_t a = pair_corank_win<K, HV>(in, A, q.la, B, q.lb, Q, qab, S, xx);
            const K kx = key_after_split<K, HV>(in, A, q.la, B, q.lb, a, xx - a);
            const uint64_t yy = r - 1 - xx;  // Y[yy] = the key after yy outputs of Y
            const uint64_t c = pair_corank_win<K, HV>(in, C, q.lc, D, q.ld, Q, qcd, S, yy);
            const K ky = key_after_split<K, HV>(in, C, q.lc, D, q.ld, c, yy - c);
            return ky < kx;
        };
        x = wlo < whi ? group_first_true<LANES>(wlo, whi, l, gshift, fine) : wlo;
    }
    const uint64_t a = lx ? pair_corank_win<K, HV>(in, A, q.la, B, q.lb, Q, qab, S, x < lx ? x : lx - 1) : 0;
    const uint64_t a_fix = (x == lx) ? q.la : a;
    const uint64_t y = r - x;
    const uint64_t c = ly ? pair_corank_win<K, HV>(in, C, q.lc, D, q.ld, Q, qcd, S, y < ly ? y : ly - 1) : 0;
    const uint64_t c_fix = (y == ly) ? q.lc : c;
    if (live && l == 0) {
        tab[3 * t + 0] = a_fix;
        tab[3 * t + 1] = x;
        tab[3 * t + 2] = c_fix;
    }
}

// Serial merge of `count` outputs of merge(A, B) starting at diagonal diag,
// written to LDS (dk/dv) at out_pos.. (runtime count).
template <bool HV, typename K>
__device__ __forceinline__ void lds_merge_to_lds(const K *sk, const uint64_t *sv, int a_off, int la, int b_off, int lb,
                                                 int diag, int count, int limit, K *dk, uint64_t *dv, int out_pos) {
    const int a = lds_corank<K>(sk, a_off, la, b_off, lb, diag);
    int ia = a_off + a, ib = b_off + (diag - a);
    const int ea = a_off + la, eb = b_off + lb;
    K ka = sk[lsw<K>(min(ia, limit))], kb = sk[lsw<K>(min(ib, limit))];
    uint64_t va = 0, vb = 0;
    if constexpr (HV) {
        va = sv[lsw<uint64_t>(min(ia, limit))];
        vb = sv[lsw<uint64_t>(min(ib, limit))];
    }
    for (int i = 0; i < count; i++) {
        const bool takeA = ib >= eb || (ia < ea && !(kb < ka));
        dk[lsw<K>(out_pos + i)] = takeA ? ka : kb;
        if constexpr (HV) dv[lsw<uint64_t>(out_pos + i)] = takeA ? va : vb;
        ia += takeA;
        ib += !takeA;
        const int nx = min(takeA ? ia : ib, limit);
        const K nk = sk[lsw<K>(nx)];
        ka = takeA ? nk : ka;
        kb = takeA ? kb : nk;
        if constexpr (HV) {
            const uint64_t nv = sv[lsw<uint64_t>(nx)];
            va = takeA ? nv : va;
            vb = takeA ? vb : nv;
        }
    }
}

template <typename K, bool HV, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void k_quad_merge_pass(const void *__restrict__ in, void *__restrict__ out,
                                                           uint64_t n, uint64_t R, int cout,
                                                           const uint64_t *__restrict__ tab) {
    constexpr int TILE = BLOCK * IPT;
    __shared__ __attribute__((aligned(16))) K sk[2 * TILE];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? 2 * TILE : 2];
    K *k1 = sk, *k2 = sk + TILE;
    uint64_t *v1 = sv, *v2 = HV ? sv + TILE : sv;
    const int tid = threadIdx.x;
    const uint64_t t = blockIdx.x;
    const uint64_t o0 = t * TILE;
    const QuadGeom q = quad_geom(o0, n, R);
    const uint64_t r0 = o0 - q.g4;
    const int cnt = (int)(n - o0 < (uint64_t)TILE ? n - o0 : (uint64_t)TILE);
    const uint64_t r1 = r0 + cnt;
    const uint64_t lx = q.la + q.lb, ly = q.lc + q.ld;
    const uint64_t a0 = tab[3 * t], x0 = tab[3 * t + 1], c0 = tab[3 * t + 2];
    uint64_t a1, x1, c1;
    if (r1 == lx + ly) {
        a1 = q.la;
        x1 = lx;
        c1 = q.lc;
    } else {
        a1 = tab[3 * t + 3];
        x1 = tab[3 * t + 4];
        c1 = tab[3 * t + 5];
    }
    const uint64_t b0 = x0 - a0, b1 = x1 - a1, y0 = r0 - x0, y1 = r1 - x1, d0 = y0 - c0, d1 = y1 - c1;
    const int na = (int)(a1 - a0), nb = (int)(b1 - b0), nc = (int)(c1 - c0), nd = (int)(d1 - d0);
    const uint64_t sA = q.g4 + a0, sB = q.g4 + q.la + b0, sC = q.g4 + lx + c0, sD = q.g4 + lx + q.lc + d0;
    const int e1 = na, e2 = na + nb, e3 = na + nb + nc;
#pragma unroll 4
    for (int e = tid; e < cnt; e += BLOCK) {
        const uint64_t src = e < e1 ? sA + e : e < e2 ? sB + (e - e1) : e < e3 ? sC + (e - e2) : sD + (e - e3);
        K k;
        uint64_t vv = 0;
        GMem<K, HV>::load(in, src, k, vv);
        k1[lsw<K>(e)] = k;
        if constexpr (HV) v1[lsw<uint64_t>(e)] = vv;
    }
    __syncthreads();
    // level 1: X part = merge(A, B) -> k2[0, nx), Y part = merge(C, D) -> k2[nx, cnt)
    const int nx = na + nb;
    const int d = tid * IPT;
    if (d < cnt) {
        const int c = cnt - d < IPT ? cnt - d : IPT;
        if (d < nx) {
            const int k = nx - d < c ? nx - d : c;
            lds_merge_to_lds<HV>(k1, v1, 0, na, na, nb, d, k, cnt - 1, k2, v2, d);
        }
        if (d + c > nx) {
            const int s0 = d > nx ? d : nx;
            lds_merge_to_lds<HV>(k1, v1, e2, nc, e3, nd, s0 - nx, d + c - s0, cnt - 1, k2, v2, s0);
        }
    }
    __syncthreads();
    // level 2: merge(X, Y) -> registers -> LDS -> coalesced stores
    K r[IPT];
    uint64_t v[IPT];
    if (d < cnt) {
        const int c = cnt - d < IPT ? cnt - d : IPT;
        lds_merge_chains<IPT, OGS_MERGE_CHAINS, HV>(k2, v2, 0, nx, nx, cnt - nx, d, c, cnt - 1, r, v);
    }
    __syncthreads();
    if (d < cnt) lds_put<IPT, HV>(k1, v1, d, r, v);
    __syncthreads();
    lds_store_tile<BLOCK, TILE, HV>(k1, v1, out, o0, cnt, cout);
}

// Launches the three kernels of one 4-way pass.  Q: ceil(n/S) entries,
// tab: 3*ceil(n/TILE) entries.
template <typename K, bool HV, int BLOCK, int IPT>
static hipError_t launch_quad(const void *in, void *out, uint64_t n, uint64_t R, int cout, uint64_t *Q, uint64_t *tab,
                              uint32_t S, int lanes, hipStream_t s, hipEvent_t mid) {
    constexpr uint64_t TILE = (uint64_t)BLOCK * IPT;
    constexpr int QL = 16;
    const uint64_t ntiles = (n + TILE - 1) / TILE;
    if (ntiles == 0) return hipSuccess;
    const uint64_t nq = (n + S - 1) / S;
    launch_partition<K, HV>(lanes, in, n, R, S, nq, Q, s);
    const uint64_t threads = ntiles * QL;
    hipLaunchKernelGGL((k_quad_partition<K, HV, QL>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, in, n,
                       R, (uint32_t)TILE, S, ntiles, (const uint64_t *)Q, tab);
    if (mid) (void)hipEventRecord(mid, s);
    hipLaunchKernelGGL((k_quad_merge_pass<K, HV, BLOCK, IPT>), dim3((unsigned)ntiles), dim3(BLOCK), 0, s, in, out, n,
                       R, cout, (const uint64_t *)tab);
    return hipGetLastError();
}

// Registry of compiled shapes per key class.  Block-sort shapes are
// (BLOCK, IPT, WLEV), merge shapes (BLOCK, IPT).  The first entry of each list
// is the default; the tuning table may pick any other.
#define OGS_BLOCK_SHAPES_32(X) X(256, 16, 1) X(256, 16, 0) X(256, 16, 2) X(256, 16, 3) X(256, 32, 1) \
    X(512, 16, 1) X(512, 32, 1) X(1024, 16, 1) X(1024, 32, 1) X(128, 16, 1) X(256, 8, 1) X(64, 16, 1) X(64, 8, 1)
#define OGS_MERGE_SHAPES_32(X) X(256, 16) X(256, 8) X(512, 16) X(128, 16) X(256, 32) X(512, 8) X(1024, 16) \
    X(512, 32) X(128, 4)
#define OGS_BLOCK_SHAPES_64(X) X(256, 16, 1) X(256, 8, 1) X(128, 16, 1) X(512, 8, 1) X(512, 16, 1) \
    X(1024, 16, 1) X(64, 16, 1) X(64, 8, 1)
#define OGS_MERGE_SHAPES_64(X) X(256, 16) X(256, 8) X(512, 8) X(128, 16) X(512, 16) X(1024, 8) X(128, 4)
#define OGS_BLOCK_SHAPES_KV(X) X(256, 8, 1) X(128, 8, 1) X(256, 4, 1) X(512, 4, 1) X(512, 8, 1) X(64, 8, 1) \
    X(64, 4, 1)
#define OGS_MERGE_SHAPES_KV(X) X(256, 8) X(128, 8) X(256, 4) X(512, 4) X(512, 8) X(1024, 4) X(128, 4)

#define OGS_BX(K, HV, B, I, W) {B, I, W, &launch_block<K, HV, B, I, W>, &launch_seg_block<K, HV, B, I, W>},
#define OGS_MX(K, HV, B, I) {B, I, &launch_merge<K, HV, B, I>, &launch_seg_merge<K, HV, B, I>, &launch_quad<K, HV, B, I>},

static const BlockShapeFns kBlock32[] = {
#define X(B, I, W) OGS_BX(uint32_t, false, B, I, W)
    OGS_BLOCK_SHAPES_32(X)
#undef X
};
static const BlockShapeFns kBlock64[] = {
#define X(B, I, W) OGS_BX(uint64_t, false, B, I, W)
    OGS_BLOCK_SHAPES_64(X)
#undef X
};
static const BlockShapeFns kBlockKV[] = {
#define X(B, I, W) OGS_BX(uint64_t, true, B, I, W)
    OGS_BLOCK_SHAPES_KV(X)
#undef X
};
static const MergeShapeFns kMerge32[] = {
#define X(B, I) OGS_MX(uint32_t, false, B, I)
    OGS_MERGE_SHAPES_32(X)
#undef X
};
static const MergeShapeFns kMerge64[] = {
#define X(B, I) OGS_MX(uint64_t, false, B, I)
    OGS_MERGE_SHAPES_64(X)
#undef X
};
static const MergeShapeFns kMergeKV[] = {
#define X(B, I) OGS_MX(uint64_t, true, B, I)
    OGS_MERGE_SHAPES_KV(X)
#undef X
};

#define OGS_COUNT(a) ((int)(sizeof(a) / sizeof((a)[0])))

const BlockShapeFns *block_shapes(KeyClass kc, int *count) {
    switch (kc) {
    case KC_32: *count = OGS_COUNT(kBlock32); return kBlock32;
    case KC_64: *count = OGS_COUNT(kBlock64); return kBlock64;
    default: *count = OGS_COUNT(kBlockKV); return kBlockKV;
    }
}

const MergeShapeFns *merge_shapes(KeyClass kc, int *count) {
    switch (kc) {
    case KC_32: *count = OGS_COUNT(kMerge32); return kMerge32;
    case KC_64: *count = OGS_COUNT(kMerge64); return kMerge64;
    default: *count = OGS_COUNT(kMergeKV); return kMergeKV;
    }
}

}  // namespace ogs
