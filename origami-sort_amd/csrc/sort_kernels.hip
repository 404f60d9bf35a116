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
                                          K *sk, uint64_t *sv, int *smisc, const void *vin = nullptr,
                                          void *vout = nullptr, const SampleSpec *sp = nullptr,
                                          uint64_t sbase = ~0ull) {
    constexpr int TILE = BLOCK * IPT;
    constexpr K KMAX = KeyMax<K>::v;
    const int tid = threadIdx.x, lane = tid & 63;
    K r[IPT];
    uint64_t v[IPT];
    int nmax = 0;
    bool loaded = false;
    if constexpr (!HV) {
        if (cnt == TILE) {
            // full key-only tile: loads off one base pointer, codec without
            // per-key branches
            const K *ip = (const K *)in + base;
#pragma unroll
            for (int i = 0; i < IPT; i++) {
                r[i] = ip[(uint32_t)(i * BLOCK + tid)];
                v[i] = 0;
            }
            if (cin) {
                const CodecMask<K> m = codec_mask<K>(cin);
#pragma unroll
                for (int i = 0; i < IPT; i++) r[i] = enc_bf(r[i], m);
            }
            loaded = true;
        }
    }
    // all IPT loads are issued before any is used (indices past the end are
    // clamped to the last item, then replaced by pads)
    if (loaded) {
    } else if (HV && vin) {
        // structure of arrays: keys from in[], values from vin[]
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const uint64_t e = base + (uint64_t)(i * BLOCK + tid < cnt ? i * BLOCK + tid : cnt - 1);
            r[i] = ((const K *)in)[e];
            v[i] = ((const uint64_t *)vin)[e];
        }
    } else {
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int e = i * BLOCK + tid;
            v[i] = 0;
            GMem<K, HV>::load(in, base + (e < cnt ? e : cnt - 1), r[i], v[i]);
        }
    }
    if (!loaded) {
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const bool live = i * BLOCK + tid < cnt;
            r[i] = live ? enc_key(r[i], cin) : KMAX;
            v[i] = live ? v[i] : 0;
            if constexpr (HV) nmax += live && r[i] == KMAX;
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
                        if (vout) {
                            ((K *)out)[base + lim + slot] = dec_key(r[i], cout);
                            ((uint64_t *)vout)[base + lim + slot] = v[i];
                        } else {
                            GMem<K, HV>::store(out, base + lim + slot, dec_key(r[i], cout), v[i]);
                        }
                    }
                }
            }
        }
    }
    lane_network<IPT, HV>(r, v);
    wave_sort_levels<WLEV, IPT, HV>(r, v, lane);
    // key-only: the first and last key of every IPT-item block of the current
    // runs (thread tid's block, in its run's ascending order, as pair tid),
    // for the co-rank's steps of a block or more; a guard pair either side
    __shared__ __attribute__((aligned(16))) K s_spb[HV ? 2 : 2 * (BLOCK + 2)];
    K *s_spFL = s_spb + (HV ? 0 : 2);
#pragma unroll 1
    for (int R = IPT << WLEV; R < TILE; R <<= 1) {
        const int o = tid * IPT;
        const int pb = o & ~(2 * R - 1);
        const int diag = o - pb;
        __syncthreads();
        if constexpr (!HV) {
            // odd runs of length R are the B side of this round: descending
            if (o & R) lds_put_desc<IPT>(sk, 2 * (o & ~(R - 1)) + R - 1 - o, r);
            else lds_put<IPT, HV>(sk, sv, o, r, v);
            put_pair(s_spFL + 2 * tid, r[0], r[IPT - 1]);
            __syncthreads();
            const int a = lds_corank_split_rev<IPT>(sk, s_spFL, pb, R, pb + R, R, pb + 2 * R - 1, diag, R);
            lds_merge_rev<IPT, true>(sk, pb + a, pb + 2 * R - 1 - (diag - a), 0, IPT, r);
        } else {
            lds_put<IPT, HV>(sk, sv, o, r, v);
            __syncthreads();
            lds_merge_slice<IPT, HV>(sk, sv, pb, R, pb + R, R, diag, IPT, R, r, v);
        }
    }
    __syncthreads();
    lds_put<IPT, HV>(sk, sv, tid * IPT, r, v);
    __syncthreads();
    lds_store_tile<BLOCK, TILE, HV>(sk, sv, out, base, lim, cout, vout);
    // samples of the whole tile: in KV tiles the items past lim (maximum keys
    // written straight to the tail) sit in LDS as maximum keys too
    // (the tile's first item is item sbase of the sampled array; base by default)
    if (sp && sp->smp)
        write_samples<BLOCK>(sk, cnt, sbase == ~0ull ? base : sbase, sp->smp, sp->rlog2, sp->kw, sp->slog2);
}

template <typename K, bool HV, int BLOCK, int IPT, int WLEV>
__global__ __launch_bounds__(BLOCK) void k_block_sort(const void *in, void *out, uint64_t n, int cin, int cout,
                                                      const void *vin, void *vout, SampleSpec sp) {
    constexpr int TILE = BLOCK * IPT;
    __shared__ __attribute__((aligned(16))) K skb[TILE + lds_extra<K, IPT>()];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? TILE + lds_back<K, IPT>() : 2];
    __shared__ int smisc[2];
    K *sk = skb + lds_front<K>();
    const uint64_t base = (uint64_t)blockIdx.x * TILE;
    const int cnt = (int)(n - base < (uint64_t)TILE ? n - base : (uint64_t)TILE);
    tile_sort<K, HV, BLOCK, IPT, WLEV>(in, out, base, cnt, cin, cout, sk, sv, smisc, vin, vout, &sp);
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
__global__ __launch_bounds__(BLOCK) void k_seg_tile_sort(const TileDesc *__restrict__ tiles, int cin, int cout,
                                                         const KSegH *__restrict__ kseg) {
    constexpr int TILE = BLOCK * IPT;
    __shared__ __attribute__((aligned(16))) K skb[TILE + lds_extra<K, IPT>()];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? TILE + lds_back<K, IPT>() : 2];
    __shared__ int smisc[2];
    K *sk = skb + lds_front<K>();
    const TileDesc t = tiles[blockIdx.x];
    const int cnt = (int)(t.cnt & kTileCountMask);
    const int co = ((t.cnt >> 30) & 1u) ? cout : 0;
    if (t.cnt & kTileKwayBit) {
        // a tile of a segment on the K-way plan: also the samples of its run
        // for the segment's first K-way pass (the tile sits in the segment's
        // buf[0] at item (out - buf[0]) / item size)
        const KSegH &g = kseg[t.pad];
        constexpr uint64_t W = HV ? 16 : sizeof(K);
        const uint64_t g0 = (uint64_t)((const char *)t.out - (const char *)g.buf[0]) / W;
        const SampleSpec sp = {g.smp[0], __builtin_ctz((unsigned)TILE), 1 << g.lv[0], (int)g.sl[0]};
        tile_sort<K, HV, BLOCK, IPT, WLEV>(t.in, t.out, 0, cnt, cin, co, sk, sv, smisc, nullptr, nullptr, &sp, g0);
        return;
    }
    tile_sort<K, HV, BLOCK, IPT, WLEV>(t.in, t.out, 0, cnt, cin, co, sk, sv, smisc);
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
                                           uint64_t *sco, const uint64_t *part, void *vout = nullptr) {
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
    {
        // all loads in flight at once (indices past the end clamped)
        K kk[IPT];
        uint64_t vv[IPT];
        const uint64_t Bs = B - (uint64_t)la;  // item e >= la is at Bs + e
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int e = i * BLOCK + tid;
            const int ec = e < cnt ? e : cnt - 1;
            vv[i] = 0;
            GMem<K, HV>::load(in, (ec < la ? A : Bs) + (uint64_t)ec, kk[i], vv[i]);
        }
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int e = i * BLOCK + tid;
            if (e < cnt) {
                // key-only: the B slice descending after A (bitonic layout)
                if constexpr (!HV) sk[lsw<K>(e < la ? e : cnt - 1 + la - e)] = kk[i];
                if constexpr (HV) {
                    sk[lsw<K>(e)] = kk[i];
                    sv[lsw<uint64_t>(e)] = vv[i];
                }
            }
        }
    }
    __syncthreads();
    const int d = tid * IPT;
    K r[IPT];
    uint64_t v[IPT];
    if constexpr (!HV) {
        if (d < cnt)
            lds_merge_slice_rev<IPT>(sk, 0, la, cnt - 1, lb, d, cnt - d < IPT ? cnt - d : IPT,
                                     pow2_floor(la < lb ? la : lb), r);
#pragma unroll
        for (int i = 0; i < IPT; i++) v[i] = 0;
    } else {
        if (d < cnt)
            lds_merge_slice<IPT, HV>(sk, sv, 0, la, la, lb, d, cnt - d < IPT ? cnt - d : IPT,
                                     pow2_floor(la < lb ? la : lb), r, v);
    }
    __syncthreads();
    if (d < cnt) lds_put<IPT, HV>(sk, sv, d, r, v);
    __syncthreads();
    lds_store_tile<BLOCK, TILE, HV>(sk, sv, out, seg + o0, cnt, cout, vout);
}

template <typename K, bool HV, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void k_merge_pass(const void *__restrict__ in, void *__restrict__ out, uint64_t n,
                                                      uint64_t R, int cout, const uint64_t *__restrict__ part,
                                                      void *vout) {
    constexpr int TILE = BLOCK * IPT;
    __shared__ __attribute__((aligned(16))) K skb[TILE + lds_extra<K, IPT>()];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? TILE + lds_back<K, IPT>() : 2];
    __shared__ uint64_t sco[2];
    K *sk = skb + lds_front<K>();
    merge_tile<K, HV, BLOCK, IPT>(in, out, 0, n, R, (uint64_t)blockIdx.x * TILE, cout, sk, sv, sco,
                                  part ? part + blockIdx.x : nullptr, vout);
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
    __shared__ __attribute__((aligned(16))) K skb[TILE + lds_extra<K, IPT>()];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? TILE + lds_back<K, IPT>() : 2];
    __shared__ uint64_t sco[2];
    K *sk = skb + lds_front<K>();
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
static hipError_t launch_block(const void *in, void *out, uint64_t n, int cin, int cout, hipStream_t s,
                              const void *vin, void *vout, const SampleSpec *sp) {
    constexpr uint64_t TILE = (uint64_t)BLOCK * IPT;
    const uint64_t grid = (n + TILE - 1) / TILE;
    if (grid == 0) return hipSuccess;
    const SampleSpec none = {nullptr, 0, 0, 0};
    hipLaunchKernelGGL((k_block_sort<K, HV, BLOCK, IPT, WLEV>), dim3((unsigned)grid), dim3(BLOCK), 0, s, in, out, n,
                       cin, cout, vin, vout, sp ? *sp : none);
    return hipGetLastError();
}

template <typename K, bool HV, int BLOCK, int IPT, int WLEV>
static hipError_t launch_seg_block(const void *tiles, uint64_t ntiles, int cin, int cout, const KSegH *kseg,
                                   hipStream_t s) {
    if (ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL((k_seg_tile_sort<K, HV, BLOCK, IPT, WLEV>), dim3((unsigned)ntiles), dim3(BLOCK), 0, s,
                       (const TileDesc *)tiles, cin, cout, kseg);
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
                               hipStream_t s, hipEvent_t mid, void *vout) {
    constexpr uint64_t TILE = (uint64_t)BLOCK * IPT;
    const uint64_t ntiles = (n + TILE - 1) / TILE;
    if (ntiles == 0) return hipSuccess;
    if (part) launch_partition<K, HV>(lanes, in, n, R, (uint32_t)TILE, ntiles, part, s);
    if (mid) (void)hipEventRecord(mid, s);
    hipLaunchKernelGGL((k_merge_pass<K, HV, BLOCK, IPT>), dim3((unsigned)ntiles), dim3(BLOCK), 0, s, in, out, n, R,
                       cout, (const uint64_t *)part, vout);
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
// 4-way merge pass: groups of four runs A, B, C, D (length R each, ragged in
// the last group) become one run of 4R in a single HBM pass -- two levels of
// the reference's merge tree (MergeTreeEven, merge_tree.h:2576-3010) per read
// and write of the data.  With X = merge(A, B) and Y = merge(C, D) (ties go
// A, B, C, D), the output is merge(X, Y).
//
// Tiles are cut at points of the (X, Y) merge path: one at every h-th item of
// X, one at every h-th item of Y (h = BLOCK*IPT/2), and the group start.
// Between consecutive points X and Y each advance by at most h, so an interval
// holds at most 2h items, and every point splits all four runs exactly:
//  - the point at X index x: (a, b) = co-rank of (A, B) at x, v = X[x], and Y
//    contributes its items below v: c = #C < v, d = #D < v;
//  - the point at Y index y: (c, d) = co-rank of (C, D) at y, w = Y[y], and X
//    contributes its items up to w: a = #A <= w, b = #B <= w.
// Along the path the X point j is preceded by exactly ceil((c+d)/h) Y points
// (a Y point at the same y follows it), and the Y point i by ceil((a+b)/h) X
// points, so each point's slot in its group's tile list is known in closed
// form and no sort of the points is needed.  Every X and Y index is found with
// one 2-way co-rank and two independent binary searches.
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

// Tile slots per group: the group start, 2R/h X points and 2R/h Y points
// (h divides R).  The last group uses the first 1 + ceil(lx/h) + ceil(ly/h).
__host__ __device__ inline uint64_t quad_slots(uint64_t R, uint64_t h) { return 1 + 4 * R / h; }

// Per-thread merge-path co-rank in global memory (ties to A).
template <typename K, bool HV>
__device__ __forceinline__ uint64_t thread_corank(const void *in, uint64_t abase, uint64_t la, uint64_t bbase,
                                                  uint64_t lb, uint64_t diag) {
    uint64_t lo = diag > lb ? diag - lb : 0;
    uint64_t hi = diag < la ? diag : la;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        const K a = GMem<K, HV>::load_key(in, abase + mid);
        const K b = GMem<K, HV>::load_key(in, bbase + diag - 1 - mid);
        if (!(b < a)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The item after taking a items of run P and b items of run Q in merge(P, Q)
// (ties to P); at least one of them has an item left.
template <typename K, bool HV>
__device__ __forceinline__ K next_merged(const void *in, uint64_t pbase, uint64_t lp, uint64_t qbase, uint64_t lq,
                                         uint64_t a, uint64_t b) {
    if (b >= lq) return GMem<K, HV>::load_key(in, pbase + a);
    if (a >= lp) return GMem<K, HV>::load_key(in, qbase + b);
    const K kp = GMem<K, HV>::load_key(in, pbase + a);
    const K kq = GMem<K, HV>::load_key(in, qbase + b);
    return kq < kp ? kq : kp;
}

// Items of runs P and Q below v (upper = false) or up to v (upper = true):
// two independent binary searches advanced in lockstep.
template <typename K, bool HV>
__device__ __forceinline__ void rank_pair(const void *in, uint64_t pbase, uint64_t lp, uint64_t qbase, uint64_t lq,
                                          K v, bool upper, uint64_t &rp, uint64_t &rq) {
    uint64_t lo0 = 0, hi0 = lp, lo1 = 0, hi1 = lq;
    while (lo0 < hi0 || lo1 < hi1) {
        const uint64_t m0 = (lo0 + hi0) >> 1, m1 = (lo1 + hi1) >> 1;
        const bool on0 = lo0 < hi0, on1 = lo1 < hi1;
        K k0 = v, k1 = v;
        if (on0) k0 = GMem<K, HV>::load_key(in, pbase + m0);
        if (on1) k1 = GMem<K, HV>::load_key(in, qbase + m1);
        if (on0) {
            if (upper ? !(v < k0) : k0 < v) lo0 = m0 + 1;
            else hi0 = m0;
        }
        if (on1) {
            if (upper ? !(v < k1) : k1 < v) lo1 = m1 + 1;
            else hi1 = m1;
        }
    }
    rp = lo0;
    rq = lo1;
}

// One thread per tile slot (the form with the fewest loads): writes the slot's split (a, b, c, d) of its group
// to tab[4*slot ..].
template <typename K, bool HV>
__global__ void k_quad_bounds1(const void *__restrict__ in, uint64_t n, uint64_t R, uint64_t h, uint64_t nslots,
                              uint64_t *__restrict__ tab) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const uint64_t per = quad_slots(R, h);
    const uint64_t g = s / per, k = s % per;
    const QuadGeom q = quad_geom(g * 4 * R, n, R);
    const uint64_t lx = q.la + q.lb, ly = q.lc + q.ld;
    const uint64_t A = q.g4, B = A + q.la, C = A + lx, D = C + q.lc;
    const uint64_t half = 2 * R / h;
    uint64_t a = 0, b = 0, c = 0, d = 0, slot = 0;
    if (k > 0 && k - 1 < half) {
        const uint64_t j = k - 1, x = j * h;
        if (x >= lx) return;
        a = thread_corank<K, HV>(in, A, q.la, B, q.lb, x);
        b = x - a;
        const K v = next_merged<K, HV>(in, A, q.la, B, q.lb, a, b);
        rank_pair<K, HV>(in, C, q.lc, D, q.ld, v, false, c, d);
        slot = 1 + j + (c + d + h - 1) / h;
    } else if (k > 0) {
        const uint64_t i = k - 1 - half, y = i * h;
        if (y >= ly) return;
        c = thread_corank<K, HV>(in, C, q.lc, D, q.ld, y);
        d = y - c;
        const K w = next_merged<K, HV>(in, C, q.lc, D, q.ld, c, d);
        rank_pair<K, HV>(in, A, q.la, B, q.lb, w, true, a, b);
        slot = 1 + i + (a + b + h - 1) / h;
    }
    ulonglong2 *t = (ulonglong2 *)(tab + 4 * (g * per + slot));
    t[0] = make_ulonglong2(a, b);
    t[1] = make_ulonglong2(c, d);
}

// Items of the sorted run P[0, len) below v (upper = false) or up to v
// (upper = true), found by the LANES lanes of one group with a (LANES+1)-ary
// search: every step probes LANES evenly spaced items and keeps the gap in
// which the predicate turns true.
template <typename K, bool HV, int LANES>
__device__ __forceinline__ uint64_t group_bound(const void *in, uint64_t base, uint64_t len, K v, bool upper, int l,
                                                int gshift) {
    uint64_t lo = 0, hi = len;
    while (hi - lo > (uint64_t)LANES) {
        const uint64_t span = hi - lo;
        const uint64_t m = lo + (span * (uint64_t)(l + 1)) / (LANES + 1);
        const K k = GMem<K, HV>::load_key(in, base + m);
        const bool p = upper ? v < k : !(k < v);
        const uint64_t bits = (__ballot(p) >> gshift) & ((1ull << LANES) - 1);
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
        const K k = GMem<K, HV>::load_key(in, base + lo + l);
        p = upper ? v < k : !(k < v);
    }
    const uint64_t bits = (__ballot(p) >> gshift) & ((1ull << LANES) - 1);
    return bits ? lo + (uint64_t)(__ffsll((unsigned long long)bits) - 1) : hi;
}

// LANES lanes per tile slot: write the slot's split (a, b, c, d) of its group
// to tab[4*slot ..].  The 2-way co-rank uses all LANES lanes; the two ranks in
// the other pair of runs use LANES/2 lanes each, side by side.  Groups past
// the end search a valid point and write nothing, so every lane of a wave
// takes part in every ballot of its group.
template <typename K, bool HV, int LANES>
__global__ void k_quad_bounds(const void *__restrict__ in, uint64_t n, uint64_t R, uint64_t h, uint64_t nslots,
                              uint64_t *__restrict__ tab) {
    static_assert(LANES >= 2 && LANES <= 64 && (LANES & (LANES - 1)) == 0, "LANES: a power of two");
    constexpr int HL = LANES / 2;
    const uint64_t gt = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = (int)(gt % LANES);
    const int gshift = (int)(threadIdx.x & 63) & ~(LANES - 1);
    uint64_t s = gt / LANES;
    bool live = s < nslots;
    if (!live) s = 0;
    const uint64_t per = quad_slots(R, h);
    const uint64_t g = s / per, k = s % per;
    const QuadGeom q = quad_geom(g * 4 * R, n, R);
    const uint64_t lx = q.la + q.lb, ly = q.lc + q.ld;
    const uint64_t A = q.g4, B = A + q.la, C = A + lx, D = C + q.lc;
    const uint64_t half = 2 * R / h;
    const bool isx = k > 0 && k - 1 < half;
    const uint64_t idx = k == 0 ? 0 : (isx ? k - 1 : k - 1 - half) * h;
    if (k == 0 || idx >= (isx ? lx : ly)) {
        // the group start, or a point past a ragged group's end
        if (live && k == 0 && l == 0) {
            ulonglong2 *t = (ulonglong2 *)(tab + 4 * s);
            t[0] = make_ulonglong2(0, 0);
            t[1] = make_ulonglong2(0, 0);
        }
        return;
    }
    // the split's own pair (P, Q) and the other pair (U, V)
    const uint64_t P = isx ? A : C, lp = isx ? q.la : q.lc, Q = isx ? B : D, lq = isx ? q.lb : q.ld;
    const uint64_t U = isx ? C : A, lu = isx ? q.lc : q.la, V = isx ? D : B, lv = isx ? q.ld : q.lb;
    const uint64_t p = group_corank<K, HV, LANES>(in, P, lp, Q, lq, idx, l, gshift);
    const K w = next_merged<K, HV>(in, P, lp, Q, lq, p, idx - p);
    // X points count Y's items below X[x]; Y points count X's items up to Y[y]
    const bool second = l >= HL;
    const uint64_t r = group_bound<K, HV, HL>(in, second ? V : U, second ? lv : lu, w, !isx, l - (second ? HL : 0),
                                              gshift + (second ? HL : 0));
    const int base = (int)(threadIdx.x & 63) & ~(LANES - 1);
    const uint64_t ru = __shfl(r, base), rv = __shfl(r, base + HL);
    if (!live || l != 0) return;
    const uint64_t slot = 1 + (idx / h) + (ru + rv + h - 1) / h;
    ulonglong2 *t = (ulonglong2 *)(tab + 4 * (g * per + slot));
    if (isx) {
        t[0] = make_ulonglong2(p, idx - p);
        t[1] = make_ulonglong2(ru, rv);
    } else {
        t[0] = make_ulonglong2(ru, rv);
        t[1] = make_ulonglong2(p, idx - p);
    }
}

// Level 1 for the outputs [d, d + IPT) of the concatenation [X | Y] of one
// tile, X = merge(A, B) at LDS [0, na + nb) and Y = merge(C, D) after it: a
// thread whose outputs straddle the X/Y boundary merges both parts.
template <int IPT, bool HV, typename K>
__device__ __forceinline__ void quad_level1(const K *sk, const uint64_t *sv, int na, int nb, int nc, int nd, int d,
                                            K (&r)[IPT], uint64_t (&v)[IPT]) {
    const int nx = na + nb, cnt = nx + nc + nd;
    if (d < nx) {
        const int cx = nx - d < IPT ? nx - d : IPT;
        const int a = lds_corank_lift<K>(sk, 0, na, na, nb, d, pow2_floor(na < nb ? na : nb));
        if (cx == IPT) lds_merge_serial<IPT, HV, true>(sk, sv, 0, na, na, nb, a, d - a, 0, IPT, r, v);
        else lds_merge_serial<IPT, HV, false>(sk, sv, 0, na, na, nb, a, d - a, 0, cx, r, v);
    }
    if (d + IPT > nx && nx < cnt) {
        const int st = d > nx ? d : nx;
        const int yd = st - nx;
        const int cy = (d + IPT < cnt ? d + IPT : cnt) - st;
        const int e3 = nx + nc;
        const int c = lds_corank_lift<K>(sk, nx, nc, e3, nd, yd, pow2_floor(nc < nd ? nc : nd));
        if (cy == IPT) lds_merge_serial<IPT, HV, true>(sk, sv, nx, nc, e3, nd, c, yd - c, 0, IPT, r, v);
        else lds_merge_serial<IPT, HV, false>(sk, sv, nx, nc, e3, nd, c, yd - c, st - d, cy, r, v);
    }
}

// One workgroup per two consecutive tile slots of a group (the intervals from
// point 2w to point 2w + 2).  Along the merge path the X and Y points come in
// near pairs, so one interval is often almost empty and two consecutive ones
// hold close to BLOCK * IPT items; each interval holds at most 2h = BLOCK * IPT,
// so the pair fits the 2 * BLOCK * IPT items of LDS.  The workgroup loads the
// four slices into LDS, merges A+B and C+D (level 1), merges the two results
// (level 2), and writes the tile.  Each level covers the tile in one or two
// rounds of BLOCK * IPT outputs; both rounds are merged into registers before
// either is written back, since the merge reads what the write overwrites.
// Slots past the end of a ragged group, and empty tiles, exit at once.
template <typename K, bool HV, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void k_quad_merge_pass(const void *__restrict__ in, void *__restrict__ out,
                                                           uint64_t n, uint64_t R, int cout,
                                                           const uint64_t *__restrict__ tab, void *vout,
                                                           uint32_t *__restrict__ err) {
    constexpr int C = BLOCK * IPT;  // outputs per round
    constexpr int CAP = 2 * C;      // LDS items: two intervals of at most C
    constexpr uint64_t h = C / 2;
    __shared__ __attribute__((aligned(16))) K skb[CAP + lds_extra<K, IPT>()];
    __shared__ __attribute__((aligned(16))) uint64_t sv[HV ? CAP + lds_back<K, IPT>() : 2];
    K *sk = skb + lds_front<K>();
    const int tid = threadIdx.x;
    const uint64_t per = quad_slots(R, h);
    const uint64_t wpg = (per + 1) / 2;
    const uint64_t g = blockIdx.x / wpg, s0 = 2 * (blockIdx.x % wpg);
    const QuadGeom q = quad_geom(g * 4 * R, n, R);
    const uint64_t lx = q.la + q.lb, ly = q.lc + q.ld;
    const uint64_t used = 1 + (lx + h - 1) / h + (ly + h - 1) / h;
    if (s0 >= used) return;
    const uint64_t *t0 = tab + 4 * (g * per + s0);
    const uint64_t a0 = t0[0], b0 = t0[1], c0 = t0[2], d0 = t0[3];
    uint64_t a1 = q.la, b1 = q.lb, c1 = q.lc, d1 = q.ld;
    if (s0 + 2 < used) {
        a1 = t0[8];
        b1 = t0[9];
        c1 = t0[10];
        d1 = t0[11];
    }
    // a split table that is not monotone or overfills the tile is never read
    // past: such a tile is not merged and the error word reports it (never
    // with a consistent table)
    if (a0 > a1 || a1 > q.la || b0 > b1 || b1 > q.lb || c0 > c1 || c1 > q.lc || d0 > d1 || d1 > q.ld ||
        (a1 - a0) + (b1 - b0) + (c1 - c0) + (d1 - d0) > (uint64_t)CAP) {
        if (tid == 0) atomicOr(err, kErrQuadSplit);
        return;
    }
    const int na = (int)(a1 - a0), nb = (int)(b1 - b0), nc = (int)(c1 - c0), nd = (int)(d1 - d0);
    const int cnt = na + nb + nc + nd;
    if (cnt == 0) return;
    const uint64_t sA = q.g4 + a0, sB = q.g4 + q.la + b0, sC = q.g4 + lx + c0, sD = q.g4 + lx + q.lc + d0;
    const int e1 = na, e2 = na + nb, e3 = na + nb + nc;
    {
        // item e of the tile is at (base of its slice) + e
        const uint64_t bA = sA, bB = sB - (uint64_t)e1, bC = sC - (uint64_t)e2, bD = sD - (uint64_t)e3;
#pragma unroll
        for (int half = 0; half < 2; half++) {
            if (half * C >= cnt) break;
            K kk[IPT];
            uint64_t vv[IPT];
#pragma unroll
            for (int i = 0; i < IPT; i++) {
                const int e0 = half * C + i * BLOCK + tid;
                const int e = e0 < cnt ? e0 : cnt - 1;
                const uint64_t src = (e < e1 ? bA : e < e2 ? bB : e < e3 ? bC : bD) + (uint64_t)e;
                vv[i] = 0;
                GMem<K, HV>::load(in, src, kk[i], vv[i]);
            }
#pragma unroll
            for (int i = 0; i < IPT; i++) {
                const int e = half * C + i * BLOCK + tid;
                if (e < cnt) {
                    sk[lsw<K>(e)] = kk[i];
                    if constexpr (HV) sv[lsw<uint64_t>(e)] = vv[i];
                }
            }
        }
    }
    __syncthreads();
    const bool two = cnt > C;  // uniform over the workgroup
    const int dA = tid * IPT, dB = C + tid * IPT;
    K r0[IPT], r1[IPT];
    uint64_t v0[IPT], v1[IPT];
    if (dA < cnt) quad_level1<IPT, HV>(sk, sv, na, nb, nc, nd, dA, r0, v0);
    if (two && dB < cnt) quad_level1<IPT, HV>(sk, sv, na, nb, nc, nd, dB, r1, v1);
    __syncthreads();
    if (dA < cnt) lds_put<IPT, HV>(sk, sv, dA, r0, v0);
    if (two && dB < cnt) lds_put<IPT, HV>(sk, sv, dB, r1, v1);
    __syncthreads();
    const int nx = e2, ny = cnt - e2;
    const int top = pow2_floor(nx < ny ? nx : ny);
    if (dA < cnt) lds_merge_slice<IPT, HV>(sk, sv, 0, nx, nx, ny, dA, cnt - dA < IPT ? cnt - dA : IPT, top, r0, v0);
    if (two && dB < cnt)
        lds_merge_slice<IPT, HV>(sk, sv, 0, nx, nx, ny, dB, cnt - dB < IPT ? cnt - dB : IPT, top, r1, v1);
    __syncthreads();
    if (dA < cnt) lds_put<IPT, HV>(sk, sv, dA, r0, v0);
    if (two && dB < cnt) lds_put<IPT, HV>(sk, sv, dB, r1, v1);
    __syncthreads();
    lds_store_tile<BLOCK, CAP, HV>(sk, sv, out, q.g4 + a0 + b0 + c0 + d0, cnt, cout, vout);
}

// Slots of one 4-way pass over n items with runs of R and this tile size.
__host__ __device__ inline uint64_t quad_pass_slots(uint64_t n, uint64_t R, uint64_t tile) {
    return (n + 4 * R - 1) / (4 * R) * quad_slots(R, tile / 2);
}

// Launches the two kernels of one 4-way pass; tab holds 4 * quad_pass_slots
// entries.  The split kernel runs with quad_lanes lanes per slot (1: one
// thread per slot, the fewest loads; 2, 4, 8: fewer dependent steps, more
// loads).  `mid` is recorded between the two.
template <typename K, bool HV, int BLOCK, int IPT>
static hipError_t launch_quad(const void *in, void *out, uint64_t n, uint64_t R, int cout, uint64_t *tab,
                              int quad_lanes, hipStream_t s, hipEvent_t mid, void *vout, uint32_t *err) {
    constexpr uint64_t TILE = (uint64_t)BLOCK * IPT;
    // the 4-way tile holds 2 * TILE items (keys, and values for KV) in LDS
    constexpr uint64_t lds = 2 * TILE * (sizeof(K) + (HV ? sizeof(uint64_t) : 0));
    if constexpr (lds > 160 * 1024) {
        (void)in, (void)out, (void)n, (void)R, (void)cout, (void)tab, (void)quad_lanes, (void)s, (void)mid, (void)vout,
            (void)err;
        return hipErrorInvalidConfiguration;  // this merge shape has no 4-way form
    } else {
        if (n == 0) return hipSuccess;
        if (R % (TILE / 2)) return hipErrorInvalidValue;
        const uint64_t nslots = quad_pass_slots(n, R, TILE);
        const uint64_t per = quad_slots(R, TILE / 2);
        const uint64_t nwg = (n + 4 * R - 1) / (4 * R) * ((per + 1) / 2);
        if (quad_lanes <= 1)
            hipLaunchKernelGGL((k_quad_bounds1<K, HV>), dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s, in, n, R,
                               TILE / 2, nslots, tab);
        else if (quad_lanes == 2)
            hipLaunchKernelGGL((k_quad_bounds<K, HV, 2>), dim3((unsigned)((nslots * 2 + 255) / 256)), dim3(256), 0, s, in,
                               n, R, TILE / 2, nslots, tab);
        else if (quad_lanes == 4)
            hipLaunchKernelGGL((k_quad_bounds<K, HV, 4>), dim3((unsigned)((nslots * 4 + 255) / 256)), dim3(256), 0, s, in,
                               n, R, TILE / 2, nslots, tab);
        else
            hipLaunchKernelGGL((k_quad_bounds<K, HV, 8>), dim3((unsigned)((nslots * 8 + 255) / 256)), dim3(256), 0, s, in,
                               n, R, TILE / 2, nslots, tab);
        if (mid) (void)hipEventRecord(mid, s);
        hipLaunchKernelGGL((k_quad_merge_pass<K, HV, BLOCK, IPT>), dim3((unsigned)nwg), dim3(BLOCK), 0, s, in, out, n, R,
                           cout, (const uint64_t *)tab, vout, err);
        return hipGetLastError();
    }
}

// Registry of compiled shapes per key class.  Block-sort shapes are
// (BLOCK, IPT, WLEV), merge shapes (BLOCK, IPT).  The first entry of each list
// is the default; the tuning table may pick any other.
#define OGS_BLOCK_SHAPES_32(X) X(256, 16, 1) X(256, 16, 0) X(256, 16, 2) X(256, 16, 3) X(256, 32, 1) \
    X(512, 16, 1) X(512, 32, 1) X(1024, 16, 1) X(1024, 32, 1) X(128, 16, 1) X(256, 8, 1) X(64, 16, 1) X(64, 8, 1) \
    X(512, 16, 0) X(1024, 16, 0) X(512, 32, 0) X(512, 32, 2) X(512, 32, 3) X(512, 32, 4) X(512, 16, 3) X(512, 16, 5) X(512, 32, 5) X(512, 32, 6)
#define OGS_MERGE_SHAPES_32(X) X(256, 16) X(256, 8) X(512, 16) X(128, 16) X(256, 32) X(512, 8) X(1024, 16) \
    X(512, 32) X(128, 4)
#define OGS_BLOCK_SHAPES_64(X) X(256, 16, 1) X(256, 8, 1) X(128, 16, 1) X(512, 8, 1) X(512, 16, 1) \
    X(1024, 16, 1) X(64, 16, 1) X(64, 8, 1) X(512, 16, 2) X(512, 16, 3)
#define OGS_MERGE_SHAPES_64(X) X(256, 16) X(256, 8) X(512, 8) X(128, 16) X(512, 16) X(1024, 8) X(128, 4)
#define OGS_BLOCK_SHAPES_KV(X) X(256, 8, 1) X(128, 8, 1) X(256, 4, 1) X(512, 4, 1) X(512, 8, 1) X(64, 8, 1) \
    X(64, 4, 1) X(512, 8, 2) X(512, 8, 3)
#define OGS_MERGE_SHAPES_KV(X) X(256, 8) X(128, 8) X(256, 4) X(512, 4) X(512, 8) X(1024, 4) X(128, 4)

#define OGS_BX(K, HV, B, I, W) {B, I, W, &launch_block<K, HV, B, I, W>, &launch_seg_block<K, HV, B, I, W>},
#define OGS_MX(K, HV, B, I)                                                                                   \
    {B, I, &launch_merge<K, HV, B, I>, &launch_seg_merge<K, HV, B, I>, &launch_quad<K, HV, B, I>,              \
     2ull * B * I * (sizeof(K) + (HV ? sizeof(uint64_t) : 0)) <= 160 * 1024},

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

// ----------------------------------------------------------------------------
// Descriptor expansion of a batched sort: the host uploads one record per
// segment, and one wavefront per segment writes that segment's tile
// descriptors (its block-sort tiles, in the buffer that makes its last pass
// land in `fin`) and its merge descriptors for every pass.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_expand(const SegRecH *__restrict__ segs, TileDescH *td0, TileDescH *td1,
                                               TileDescH *td2, MergeDescH *md, const uint64_t *__restrict__ poff,
                                               uint64_t nlarge, uint64_t LT, uint64_t MT, uint32_t w) {
    const SegRecH g = segs[blockIdx.x];
    const uint64_t lane = threadIdx.x;
    if (g.len == 0) return;
    if (g.cls < 2) {
        if (lane == 0) (g.cls == 0 ? td0 : td1)[g.tidx] = TileDescH{g.in, g.fin, (uint32_t)g.len | kTileFinalBit, 0};
        return;
    }
    const uint32_t p = g.passes;
    char *dst = (char *)((p & 1) ? g.other : g.fin);
    const uint32_t fin = p == 0 ? kTileFinalBit : 0u;
    const uint32_t kw = g.kseg1 ? kTileKwayBit : 0u;
    const uint64_t nt = (g.len + LT - 1) / LT;
    for (uint64_t k = lane; k < nt; k += 64) {
        const uint64_t o = k * LT;
        const uint64_t c = g.len - o < LT ? g.len - o : LT;
        td2[g.tidx + k] = TileDescH{(const char *)g.in + o * w, dst + o * w, (uint32_t)c | fin | kw,
                                    g.kseg1 ? g.kseg1 - 1 : 0u};
    }
    if (g.kseg1) return;  // the K-way passes take it from here
    const uint64_t nm = (g.len + MT - 1) / MT;
    for (uint32_t q = 0; q < p; q++) {
        // after q passes the data sits in fin iff (passes - q) is even
        const bool in_fin = ((p - q) & 1) == 0;
        const void *src = in_fin ? g.fin : g.other;
        void *dd = in_fin ? g.other : g.fin;
        const uint64_t last = q == p - 1 ? kSegLastBit : 0;
        const uint64_t m0 = poff[(uint64_t)q * nlarge + g.lidx];
        for (uint64_t k = lane; k < nm; k += 64) md[m0 + k] = MergeDescH{src, dd, g.len, (k * MT) | last};
    }
}

hipError_t launch_expand(const SegRecH *segs, uint64_t nseg, TileDescH *td0, TileDescH *td1, TileDescH *td2,
                         MergeDescH *md, const uint64_t *poff, uint64_t nlarge, uint64_t LT, uint64_t MT, uint32_t w,
                         hipStream_t s) {
    if (nseg == 0) return hipSuccess;
    hipLaunchKernelGGL(k_expand, dim3((unsigned)nseg), dim3(64), 0, s, segs, td0, td1, td2, md, poff, nlarge, LT, MT,
                       w);
    return hipGetLastError();
}

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
