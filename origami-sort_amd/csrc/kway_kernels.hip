// kway_kernels.hip -- the sample-split K-way merge pass.
//
// One HBM pass merges groups of KW sorted runs (KW = 2 ... 32): log2(KW)
// levels of the reference's merge tree (MergeTreeEven / Merge4Way,
// merge_tree.h:319-3010) per read and write of the data, with the tiles cut by
// the role of Partition_Ptrs (Partition.h:117-166) for KW runs at once.
//
// Tiles.  The pass before this one wrote every s-th item of each run (s =
// 2^slog2) as a tagged sample (key, run, index); the tagged samples of each
// group are merged (a few 2-way passes over n/s samples, run by the host), so
// the group's samples are in the total order (key, run, index).  The tile
// boundaries of a group are its merged samples number 0, Q, 2Q, ...: for a
// boundary item b = (v, r, m*s) every run j is split exactly at the number of
// its items that precede b in the total order -- the items <= v for j < r, the
// items < v for j > r, and m*s for j = r.  Those counts are found by a binary
// search over run j's own samples, then one over the s-item window of run j
// between two samples.  Between two boundaries Q samples apart each run
// contributes at most (its samples there + 1) * s - 1 items, so a tile holds at
// most (Q + KW) * s - KW items and at least (Q - KW) * s; Q is chosen so the
// tile, with each of its KW slices padded to a multiple of IPT, fits the
// BLOCK * IPT items of LDS.
//
// Merge.  The workgroup loads the KW slices into LDS, each padded with maximum
// keys to a multiple of IPT (so every pair of runs at every level starts on a
// thread boundary and no thread's IPT outputs straddle two pairs), merges the
// pairs level by level through registers, and writes the first cnt items (the
// real ones) to the output at the group base plus the sum of the tile's start
// splits.  Key-only: the B side of every pair is stored descending (the
// bitonic layout of the block sort's rounds, so merges test no exhaustion);
// pads are maximum keys and sort last, and equal keys are identical bits.
// Key-value (16-byte {key, value} records, keys and values in two LDS arrays):
// every run ascending, and merges test exhaustion against each run's real
// length, so the merged pair is its real items then its pads and a pad never
// displaces a real maximum-key item.
#include "sort_core.hpp"
#include "sort_kernels.hpp"

namespace ogs {

// Items of run j of the group starting at gbase (KW runs of R, ragged at n).
__host__ __device__ inline uint64_t kw_run_len(uint64_t n, uint64_t gbase, uint64_t R, int j) {
    const uint64_t st = gbase + (uint64_t)j * R;
    return n > st ? (n - st < R ? n - st : R) : 0;
}

// Merged samples between two tile boundaries (0 if the shape cannot hold KW
// padded slices at this spacing).
__host__ __device__ inline uint64_t kw_q(uint64_t tile, int ipt, int kw, int slog2) {
    const int64_t q = ((int64_t)tile - (int64_t)kw * ipt) / ((int64_t)1 << slog2) - kw;
    return q > 0 ? (uint64_t)q : 0;
}

// Tiles of a group with `ns` merged samples.
__host__ __device__ inline uint64_t kw_group_tiles(uint64_t ns, uint64_t q) { return (ns + q - 1) / q; }

// Merged samples of the last group of n items (KW runs of R).
__host__ __device__ inline uint64_t kw_last_samples(uint64_t n, int rlog2, int slog2, int kw) {
    const uint64_t R = 1ull << rlog2, GR = R * (uint64_t)kw;
    const uint64_t gbase = (n - 1) / GR * GR;
    uint64_t ns = 0;
    for (int j = 0; j < kw; j++) ns += (kw_run_len(n, gbase, R, j) + (1ull << slog2) - 1) >> slog2;
    return ns;
}

// Division of a tile index t < 2^32 by a tiles-per-group count d (1 <= d <
// 2^32) with m = min(floor(2^32 / d), 2^32 - 1), computed once on the host:
// mulhi(t, m) is the quotient or one below it, so one correction step makes it
// exact.  With t, d and m uniform this is scalar code; the compiler's own
// division is about 20 vector instructions.
__host__ __device__ inline uint32_t kw_div_magic(uint64_t d) {
    const uint64_t m = (1ull << 32) / d;
    return m > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)m;
}
__device__ __forceinline__ uint32_t kw_div(uint32_t t, uint32_t d, uint32_t m, uint32_t &r) {
    uint32_t q = __umulhi(t, m);
    r = t - q * d;
    if (r >= d) {
        q++;
        r -= d;
    }
    return q;
}

uint64_t kway_tiles(uint64_t n, int rlog2, int slog2, int kw, uint64_t tile, int ipt) {
    if (n == 0) return 0;
    const uint64_t q = kw_q(tile, ipt, kw, slog2);
    if (q == 0) return 0;
    const uint64_t GR = (1ull << rlog2) * (uint64_t)kw;
    const uint64_t groups = (n + GR - 1) / GR;
    const uint64_t ntg = kw_group_tiles(GR >> slog2, q);
    return (groups - 1) * ntg + kw_group_tiles(kw_last_samples(n, rlog2, slog2, kw), q);
}

// ----------------------------------------------------------------------------
// Tile splits: one wave per tile (two tiles of 32 lanes for KW <= 8);
// tab[t * KW + j] = the split of run j at tile t's first item, the boundary
// sample b = (v, r, m) at merged position p.  The number of run j's samples
// that precede b is the index of run j's first sample at or after p in the
// merged order (run j's samples appear there in index order), so the tile's
// lanes read the 8 * KW merged samples from p on, and its lane j takes run j's
// first one by ballot.  A run with no sample in that window has all its
// samples before p when the window reaches the group's end, else the count
// comes from a search over run j's own samples.  A second search then places b
// among the s items that separate two samples of run j.  Each search is run by
// the G = (lanes per tile) / KW lanes of run j together (group_search).
// ----------------------------------------------------------------------------

// The first position in [lo, hi) whose item does not satisfy pre() (hi if all
// do), for a predicate that holds on a prefix of the range: a (G+1)-ary search
// by the G lanes of each lane group (lanes [q*G, q*G + G) search one range), one
// probe per lane per step.  Every lane of the wave calls it (a group with
// lo == hi has nothing to do); each step narrows a group's range to the gap
// between its last probe that holds and its first that does not.
template <int G, typename F>
__device__ __forceinline__ uint64_t group_search(uint64_t lo, uint64_t hi, int lane, F pre) {
    const int q = lane / G, g = lane % G;
    const uint64_t gmask = (G == 64 ? ~0ull : ((1ull << G) - 1)) << (q * G);
    // w * (G + 1) < 2^32 for every range below 2^26 (G < 64): then the
    // probe offsets take 32-bit arithmetic (a wave-uniform choice)
    const bool narrow = __ballot(hi - lo >= (1ull << 26)) == 0;
    while (__ballot(lo < hi) != 0) {
        const uint64_t w = hi - lo;
        const auto off = [&](int k) -> uint64_t {
            return narrow ? (uint64_t)((uint32_t)w * (uint32_t)k / (uint32_t)(G + 1))
                          : w * (uint64_t)k / (G + 1);
        };
        const uint64_t mid = lo + off(g + 1);
        const bool ok = lo < hi && pre(mid);
        const int c = __popcll(__ballot(ok) & gmask);  // probes that hold: a prefix of the group's
        if (lo < hi) {
            const uint64_t m_lo = lo + off(c);      // probe c - 1 (c > 0)
            const uint64_t m_hi = lo + off(c + 1);  // probe c (c < G)
            if (c < G) hi = m_hi;
            if (c > 0) lo = m_lo + 1;
        }
    }
    return lo;
}

// Tiles per wave of the split kernel: two for fan-in 8 or less (32 lanes per
// tile, at least 4 per run), else one.
template <int KW> constexpr int kw_bounds_tpw() { return KW <= 8 ? 2 : 1; }

// The splits of tile t of the pass over the n items at `in` (t < that pass's
// tiles when `valid`); every lane of the wave calls it, the lanes of half h
// (when two tiles share a wave) for their own tile.
template <typename K, bool HV, int KW>
__device__ __forceinline__ void kway_bounds_tile(const void *__restrict__ in, uint64_t n, int rlog2, int slog2,
                                                 uint64_t Q, uint64_t ntg, uint64_t t, bool valid,
                                                 const void *__restrict__ s0, const void *__restrict__ sm,
                                                 uint64_t *__restrict__ tab) {
    constexpr int TPW = kw_bounds_tpw<KW>();    // tiles per wave
    constexpr int LPT = 64 / TPW;               // lanes per tile: [h * LPT, (h + 1) * LPT) take tile h
    constexpr int NW = KW * 8 / LPT > 1 ? KW * 8 / LPT : 1;  // window samples per lane
    constexpr int G = LPT / KW;                 // lanes per run
    const int lane = threadIdx.x & 63;
    const int h = lane / LPT, hl = lane % LPT;
    const uint64_t R = 1ull << rlog2;
    const uint64_t g = t / ntg, tl = t % ntg;
    const uint64_t gbase = g * KW * R;
    const bool active = valid && tl > 0;  // the tile's splits need a search (tile 0 of a group starts at 0)
    const uint64_t gs0 = gbase >> slog2;  // the group's first merged sample
    const uint64_t gs_end =
        active ? gs0 + (gbase + KW * R >= n ? kw_last_samples(n, rlog2, slog2, KW) : (KW * R) >> slog2) : 0;
    const uint64_t p = gs0 + tl * Q;
    int rr[NW];
    uint32_t mm[NW];
    K kk0 = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        const uint64_t idx = p + (uint64_t)(hl + LPT * i);
        rr[i] = KW;  // no run
        mm[i] = 0;
        if (idx < gs_end) {
            K v;
            int r;
            uint64_t m;
            Sample<K>::get(sm, idx, v, r, m);
            rr[i] = r;
            mm[i] = (uint32_t)m;
            if (i == 0) kk0 = v;
        }
    }
    // lane h * LPT + j: the window position of run j's first sample in tile
    // h's window (-1: none)
    int src = -1;
    const uint64_t tmask = (LPT == 64 ? ~0ull : ((1ull << LPT) - 1)) << (h * LPT);
#pragma unroll
    for (int i = 0; i < NW; i++) {
#pragma unroll
        for (int j = 0; j < KW; j++) {
            const uint64_t b = __ballot(rr[i] == j) & tmask;
            if (hl == j && src < 0 && b != 0) src = i * LPT + (__ffsll((unsigned long long)b) - 1 - h * LPT);
        }
    }
    uint32_t mfound = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        const int sl = h * LPT + (src >= 0 ? src % LPT : 0);
        const uint32_t x = (uint32_t)__builtin_amdgcn_ds_bpermute(sl << 2, (int)mm[i]);
        if (src >= 0 && src / LPT == i) mfound = x;
    }
    const bool to_end = p + (uint64_t)(LPT * NW) >= gs_end;
    // lanes [h * LPT + j * G, ... + G) take run j of tile h: the window result
    // of lane h * LPT + j to its group, and the boundary sample (v, r, m),
    // window item 0 of the tile, from lane h * LPT
    const int j = hl / G;
    const int has = __builtin_amdgcn_ds_bpermute((h * LPT + j) << 2, src >= 0 ? 1 : 0);
    const uint32_t mf = (uint32_t)__builtin_amdgcn_ds_bpermute((h * LPT + j) << 2, (int)mfound);
    K v;
    if constexpr (sizeof(K) == 4) {
        v = (K)__builtin_amdgcn_ds_bpermute((h * LPT) << 2, (int)kk0);
    } else {
        v = ((K)(uint32_t)__builtin_amdgcn_ds_bpermute((h * LPT) << 2, (int)(kk0 >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_ds_bpermute((h * LPT) << 2, (int)(uint32_t)kk0);
    }
    const int r = __builtin_amdgcn_ds_bpermute((h * LPT) << 2, rr[0]);
    const uint64_t m = (uint32_t)__builtin_amdgcn_ds_bpermute((h * LPT) << 2, (int)mm[0]);
    const uint64_t lenj = active ? kw_run_len(n, gbase, R, j) : 0;
    // run j's items precede b: those <= v when j < r, those < v when j > r
    const bool up = j < r;
    const uint64_t ns_j = (lenj + (1ull << slog2) - 1) >> slog2;
    const bool search = lenj > 0 && j != r;
    // the number of run j's samples that precede b
    const uint64_t sb = (gbase + (uint64_t)j * R) >> slog2;
    const bool need = search && !has && !to_end;
    uint64_t lo = group_search<G>(0, need ? ns_j : 0, lane, [&](uint64_t i) {
        const K k = Sample<K>::key(s0, sb + i);
        return up ? !(v < k) : k < v;
    });
    if (!need) lo = has ? (uint64_t)mf : ns_j;
    // sample lo-1 precedes b and sample lo (if any) does not: b's place among
    // the items between them
    const uint64_t X = gbase + (uint64_t)j * R;
    const bool win = search && lo > 0;
    uint64_t a0 = win ? ((lo - 1) << slog2) + 1 : 0;
    uint64_t b0 = win ? (lo << slog2 < lenj ? lo << slog2 : lenj) : 0;
    const uint64_t a = group_search<G>(a0, b0, lane, [&](uint64_t i) {
        const K k = GMem<K, HV>::load_key(in, X + i);
        return up ? !(v < k) : k < v;
    });
    uint64_t split = 0;
    if (lenj > 0) split = j == r ? m << slog2 : (lo > 0 ? a : 0);
    if (!valid || lane % G != 0) return;
    tab[t * KW + j] = split;
}

template <typename K, bool HV, int KW>
__global__ __launch_bounds__(256) void k_kway_bounds(const void *__restrict__ in, uint64_t n, int rlog2, int slog2,
                                                     uint64_t Q, uint64_t ntg, uint64_t ntiles,
                                                     const void *__restrict__ s0, const void *__restrict__ sm,
                                                     uint64_t *__restrict__ tab) {
    constexpr int TPW = kw_bounds_tpw<KW>();
    const int h = (int)(threadIdx.x & 63) / (64 / TPW);
    const uint64_t t0 = ((uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)) * TPW;
    if (t0 >= ntiles) return;  // uniform in the wave
    const uint64_t t = t0 + (uint64_t)h;
    kway_bounds_tile<K, HV, KW>(in, n, rlog2, slog2, Q, ntg, t, t < ntiles, s0, sm, tab);
}

// Inclusive sums across lanes 0 .. KW-1 of a wave (KW <= 32): DPP row shifts
// give the sums within each row of 16 lanes (lanes with no source add 0); the
// lanes of the second row then add the first row's total, lane 15's sum.
template <int D> __device__ __forceinline__ uint32_t lane_shr(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x110 + D, 0xF, 0xF, true);
}
template <int KW> __device__ __forceinline__ uint32_t lane_scan(uint32_t x) {
    if constexpr (KW > 1) x += lane_shr<1>(x);
    if constexpr (KW > 2) x += lane_shr<2>(x);
    if constexpr (KW > 4) x += lane_shr<4>(x);
    if constexpr (KW > 8) x += lane_shr<8>(x);
    if constexpr (KW > 16) {
        const uint32_t u = (uint32_t)__shfl((int)x, 15, 64);
        x += (threadIdx.x & 63) >= 16 ? u : 0u;
    }
    return x;
}
template <int D> __device__ __forceinline__ uint64_t lane_shr64(uint64_t x) {
    return ((uint64_t)lane_shr<D>((uint32_t)(x >> 32)) << 32) | lane_shr<D>((uint32_t)x);
}
template <int KW> __device__ __forceinline__ uint64_t lane_scan64(uint64_t x) {
    if constexpr (KW > 1) x += lane_shr64<1>(x);
    if constexpr (KW > 2) x += lane_shr64<2>(x);
    if constexpr (KW > 4) x += lane_shr64<4>(x);
    if constexpr (KW > 8) x += lane_shr64<8>(x);
    if constexpr (KW > 16) {
        const uint64_t u = ((uint64_t)(uint32_t)__shfl((int)(x >> 32), 15, 64) << 32) |
                           (uint32_t)__shfl((int)(uint32_t)x, 15, 64);
        x += (threadIdx.x & 63) >= 16 ? u : 0ull;
    }
    return x;
}

// Items [0, cnt) of a merged LDS tile (position p is output item p of ob) to
// ob with 16-byte stores, applying the output codec CODEC: chunk c = tid + k *
// BLOCK (k < IPT / E, E = 16 / sizeof(K)) is LDS positions E*c .. E*c + E - 1,
// one aligned 16-byte LDS read (the swizzle keeps a 16-byte chunk whole, and
// lane-consecutive chunks hit distinct banks), and one 16-byte global store to
// ob + E*c, whose address is aligned to the item size (global vector accesses
// need dword alignment only).  The last cnt % E items go one per thread, and so
// does every item of a buffer that is not aligned to its item size.
template <int BLOCK, int IPT, int CODEC, typename K>
__device__ __forceinline__ void store_chunks(const K *sk, K *ob, int cnt, int tid) {
    constexpr int E = chunk_items<K>();
    constexpr int NK = IPT / E;
    static_assert(IPT % E == 0 && (E * BLOCK) % (sizeof(K) == 4 ? 256 : 128) == 0, "chunks: whole swizzle periods");
    if (((uintptr_t)ob & (sizeof(K) - 1)) != 0) {
        // a buffer not aligned to its items (uniform): one item per store
#pragma unroll 4
        for (int p = tid; p < cnt; p += BLOCK) ob[p] = dec_key(sk[lsw<K>(p)], CODEC);
        return;
    }
    const int nb = cnt / E;
    const int tail = cnt - nb * E;
    if (tid < tail) {
        const int p = nb * E + tid;
        ob[p] = dec_key(sk[lsw<K>(p)], CODEC);
    }
    // chunk k's LDS address is the first chunk's plus k * E * BLOCK (a
    // multiple of the swizzle period)
    const K *s0 = sk + lsw<K>(E * tid);
    const CodecMask<K> m = codec_mask<K>(CODEC);
    auto put = [&](int k, const K *x) {
        K *d = ob + (uint32_t)(E * (tid + k * BLOCK));
        // non-temporal: the pass's output is next read by the following pass,
        // after far more data than the caches hold
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
        if constexpr (E == 4)
            __builtin_nontemporal_store(u32x4{(uint32_t)dec_bf(x[0], m), (uint32_t)dec_bf(x[1], m),
                                              (uint32_t)dec_bf(x[2], m), (uint32_t)dec_bf(x[3], m)},
                                        (u32x4 *)d);
        else
            __builtin_nontemporal_store(u64x2{(uint64_t)dec_bf(x[0], m), (uint64_t)dec_bf(x[1], m)}, (u64x2 *)d);
    };
    K v[IPT];
    if (tid + (NK - 1) * BLOCK < nb) {
        // every chunk of this thread is whole: all reads, then all stores
#pragma unroll
        for (int k = 0; k < NK; k++) {
            if constexpr (E == 4) {
                const uint4 q = *(const uint4 *)(s0 + k * E * BLOCK);
                v[4 * k] = q.x, v[4 * k + 1] = q.y, v[4 * k + 2] = q.z, v[4 * k + 3] = q.w;
            } else {
                const ulonglong2 q = *(const ulonglong2 *)(s0 + k * E * BLOCK);
                v[2 * k] = (K)q.x, v[2 * k + 1] = (K)q.y;
            }
        }
#pragma unroll
        for (int k = 0; k < NK; k++) put(k, v + E * k);
    } else {
#pragma unroll
        for (int k = 0; k < NK; k++) {
            if (tid + k * BLOCK < nb) {
                if constexpr (E == 4) {
                    const uint4 q = *(const uint4 *)(s0 + k * E * BLOCK);
                    v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
                } else {
                    const ulonglong2 q = *(const ulonglong2 *)(s0 + k * E * BLOCK);
                    v[0] = (K)q.x, v[1] = (K)q.y;
                }
                put(k, v);
            }
        }
    }
}

// ----------------------------------------------------------------------------
// The merge: one workgroup per tile.
// ----------------------------------------------------------------------------
// Merges tile t of the pass over the n items at `in` into `out` (one
// workgroup).
template <typename K, bool HV, int KW, int BLOCK, int IPT>
__device__ __forceinline__ void kway_merge_tile(const void *__restrict__ in, void *__restrict__ out, uint64_t n,
                                                int rlog2, uint64_t ntg, uint32_t ntg_m, uint64_t nt_last,
                                                int cout, const uint64_t *__restrict__ tab, const SampleSpec &sp,
                                                uint32_t *__restrict__ err, const uint32_t t) {
    static_assert(KW == 2 || KW == 4 || KW == 8 || KW == 16 || KW == 32, "fan-in 2, 4, 8, 16 or 32");
    static_assert((IPT & (IPT - 1)) == 0 && IPT >= 4, "IPT: a power of two >= 4");
    constexpr int C = BLOCK * IPT;
    constexpr int LOGKW = KW == 2 ? 1 : KW == 4 ? 2 : KW == 8 ? 3 : KW == 16 ? 4 : 5;
    constexpr K KMAX = KeyMax<K>::v;
    __shared__ __attribute__((aligned(16))) K skb[C + lds_extra<K, IPT>()];
    __shared__ __attribute__((aligned(16))) uint64_t svb[HV ? C + lds_extra<K, IPT>() : 2];  // values (KV)
    __shared__ uint64_t s_src[KW];  // global index of each slice's first item
    __shared__ int s_n[KW];         // items per slice
    __shared__ int s_P[KW + 1];     // prefix sums of the padded slice lengths
    __shared__ int s_R[KW + 1];     // prefix sums of the real slice lengths
    __shared__ uint64_t s_o0;       // output index of the tile's first item
    __shared__ int s_cnt;           // items in the tile (-1: inconsistent table, reported in *err)
    __shared__ int s_top[LOGKW + 1];  // per level: co-rank step bound over all pairs
    // key-only: the first and last item of every IPT-item block of the current
    // runs (block k of the run at LDS offset o as pair o / IPT + k, in the
    // run's ascending order), for the block-granular co-rank steps; one guard
    // pair either side (a step clamped to its current answer reads pair -1 or
    // BLOCK, and its result does not depend on the value).
    __shared__ __attribute__((aligned(16))) K s_spb[HV ? 2 : 2 * (BLOCK + 2)];
    K *s_spFL = s_spb + (HV ? 0 : 2);
    // key-only, per level lev and pair p of the level (slices [p * 2^lev,
    // (p + 1) * 2^lev)), at entry KW - 2 * (KW >> lev) + p: the pair's LDS
    // offset, its second run's offset, its end and the level's co-rank step
    // bound (the largest over the level's pairs) -- one 16-byte read per level
    __shared__ int4 s_pair[HV ? 1 : KW];
    K *sk = skb + lds_front<K>();
    uint64_t *sv = svb + (HV ? lds_front<K>() : 0);
    const int tid = threadIdx.x, lane = tid & 63;
    const uint64_t R = 1ull << rlog2;
    uint32_t tl32;
    const uint64_t g = kw_div(t, (uint32_t)ntg, ntg_m, tl32), tl = tl32;
    const uint64_t gbase = g * KW * R;
    const uint64_t nt_here = gbase + KW * R >= n ? nt_last : ntg;
    if (tid < 64) {
        // wave 0: lane j < KW takes slice j, and the prefix sums over the
        // slices are inclusive scans across the lanes
        const int j = tid;
        int nj = 0;
        uint64_t lo = 0;
        if (j < KW) {
            const uint64_t lenj = kw_run_len(n, gbase, R, j);
            // no table: one tile per group, every run from its start
            lo = tab ? tab[t * KW + j] : 0;
            const uint64_t hi = tab && tl + 1 < nt_here ? tab[(t + 1) * KW + j] : lenj;
            s_src[j] = gbase + (uint64_t)j * R + lo;
            // a split table that is not monotone gives a negative count: the
            // tile is not merged and the error word reports it
            nj = hi >= lo && hi <= lenj ? (int)(hi - lo) : -1;
            s_n[j] = nj;
        }
        const bool bad = __ballot(nj < 0) != 0;
        const uint32_t p = lane_scan<KW>(nj > 0 ? (uint32_t)((nj + IPT - 1) & ~(IPT - 1)) : 0u);
        const uint32_t c = lane_scan<KW>(nj > 0 ? (uint32_t)nj : 0u);
        const uint64_t o = lane_scan64<KW>(lo);  // o0 = gbase + the sum of the start splits
        if (j < KW) {
            s_P[j + 1] = (int)p;
            s_R[j + 1] = (int)c;
        }
        if (j == 0) {
            s_P[0] = 0;
            s_R[0] = 0;
        }
        if (j == KW - 1) {
            s_o0 = gbase + o;
            s_cnt = bad || (int)p > C ? -1 : (int)c;  // cannot happen with a consistent table
        }
        // The per-level data below reads the prefix sums this wave has just
        // written: a wave's LDS accesses complete in order, so no barrier --
        // only a wavefront-scope fence, which keeps the compiler from moving
        // a read of another lane's entry above this lane's writes.
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (!HV && tid < 64) {
        // wave 0: lane e < KW - 1 builds entry e (level lev, pair p); the
        // level's bound is a max over its pairs, which are an aligned group of
        // KW >> lev lanes: a butterfly within the group
        const int e = tid;
        const int u = KW - 1 - e;
        const int lev = u >= 1 ? LOGKW - (31 - __clz(u)) : LOGKW + 1;  // lanes e >= KW - 1: none
        int a_off = 0, mid = 0, end = 0, mx = 0;
        if (lev <= LOGKW) {
            const int hv = 1 << (lev - 1), p = e - (KW - 2 * (KW >> lev));
            a_off = s_P[p * 2 * hv];
            mid = s_P[p * 2 * hv + hv];
            end = s_P[p * 2 * hv + 2 * hv];
            mx = min(mid - a_off, end - mid);
        }
        const int gsz = lev <= LOGKW ? KW >> lev : 1;
        if constexpr (KW >= 4) { const int o = (int)xshfl<1>((uint32_t)mx); mx = gsz > 1 ? max(mx, o) : mx; }
        if constexpr (KW >= 8) { const int o = (int)xshfl<2>((uint32_t)mx); mx = gsz > 2 ? max(mx, o) : mx; }
        if constexpr (KW >= 16) { const int o = (int)xshfl<4>((uint32_t)mx); mx = gsz > 4 ? max(mx, o) : mx; }
        if constexpr (KW >= 32) { const int o = (int)xshfl<8>((uint32_t)mx); mx = gsz > 8 ? max(mx, o) : mx; }
        if (lev <= LOGKW) s_pair[e] = make_int4(a_off, mid, end, pow2_floor(mx));
    }
    if (HV && tid >= 1 && tid <= LOGKW) {
        // the co-rank of a pair (la, lb) needs pow2_floor(min(la, lb)) lifting
        // steps (KV: the real lengths); every thread of a level takes the
        // largest, so the step loop is uniform (more steps than a pair needs
        // are harmless).  Lane lev computes level lev's bound.
        // Pair k of the level spans slices [k * 2hv, (k + 1) * 2hv); indices
        // are clamped to KW, so the pairs past the last are empty and every
        // read is independent of the others.
        const int lev = tid, hv = 1 << (lev - 1);
        const int *Q = HV ? s_R : s_P;
        int mx = 0;
#pragma unroll
        for (int k = 0; k < KW / 2; k++) {
            const int a0 = min(k * 2 * hv, KW), a1 = min(a0 + hv, KW), a2 = min(a0 + 2 * hv, KW);
            const int la = Q[a1] - Q[a0], lb = Q[a2] - Q[a1];
            mx = max(mx, min(la, lb));
        }
        s_top[lev] = pow2_floor(mx);
    }
    __syncthreads();
    const int cnt = s_cnt;
    if (cnt < 0) {  // uniform: an inconsistent split table (never with a consistent one)
        if (tid == 0) atomicOr(err, kErrKwaySplit);
        return;
    }
    const int Np = s_P[KW];
    const uint64_t o0 = s_o0;
    // this thread's chunk [c0, c0 + IPT) lies in slice j0: the last j with
    // s_P[j] <= c0 (binary search over the prefix sums)
    const int c0 = tid * IPT;
    const bool live = c0 < Np;
    int j0 = 0;
#pragma unroll
    for (int b = KW / 2; b >= 1; b >>= 1) j0 += s_P[j0 + b] <= c0 ? b : 0;
    // chunk source: position c0 + r holds item base + r (ascending slices) or
    // base - r (key-only odd slices, the B side of the first level, stored
    // descending); real items are r in [v0, v1), the rest are pads.
    // Narrow form (every item of the group lies within 4 GiB of its first):
    // loads take the group's address as a uniform base plus a 32-bit byte
    // offset; a chunk is sent as cb_lo = its byte offset B minus d, cb_hi = d
    // (-1 descending, else 0) and cmeta = the mask of its pad positions, so
    // item r sits at ((r * ES) ^ d) + (B - d) = B -/+ r * ES.  Wide form:
    // cb_lo / cb_hi = the 64-bit item index base, cmeta = v0 | v1 << 8 |
    // desc << 16.
    constexpr uint32_t ES = HV ? 16u : (uint32_t)sizeof(K);  // bytes per record
    const uint64_t gend = gbase + KW * R < n ? gbase + KW * R : n;
    const bool narrow = (gend - gbase) * ES <= (1ull << 32);  // uniform
    uint32_t cb_lo = 0, cb_hi = 0, cmeta = narrow ? ~0u : 0u;
    if (live) {
        const int pj = s_P[j0], npj = s_P[j0 + 1] - pj, nj = s_n[j0];
        const uint64_t src = s_src[j0];
        const int q = c0 - pj;
        const bool desc = !HV && (j0 & 1) != 0;
        int v0, v1;
        uint64_t base;
        if (desc) {
            base = src + (uint64_t)(npj - 1 - q);
            v0 = npj - q - nj;
            v0 = v0 < 0 ? 0 : v0 > IPT ? IPT : v0;
            v1 = IPT;
        } else {
            base = src + (uint64_t)q;
            v0 = 0;
            v1 = nj - q;
            v1 = v1 < 0 ? 0 : v1 > IPT ? IPT : v1;
        }
        if (narrow) {
            const uint32_t d = desc ? ~0u : 0u;
            cb_lo = (uint32_t)((base - gbase) * ES) - d;
            cb_hi = d;
            cmeta = ~(uint32_t)(((1ull << v1) - 1) & ~((1ull << v0) - 1));
        } else {
            cb_lo = (uint32_t)base;
            cb_hi = (uint32_t)(base >> 32);
            cmeta = (uint32_t)(v0 | (v1 << 8) | ((int)desc << 16));
        }
    }
    const uint64_t gspan = (gend - gbase) * ES;  // the group's bytes
    const bool vec16 = !HV && narrow && gspan >= 16 && gspan <= (1ull << 31);  // uniform
    if constexpr (!HV) {
        if (vec16) {
            // Key-only groups of 16 bytes to 2 GiB: 16-byte loads.  Wave w
            // loads LDS positions [w * 64 * IPT, (w + 1) * 64 * IPT); load c
            // of a lane covers the E = 16 / ES positions wbase + 64 * E * c +
            // E * lane .. + E - 1 (lane-consecutive 16-byte windows), which lie
            // in the chunk of wave lane (64 * E * c + E * lane) / IPT: that
            // lane's (cb_lo, cmeta) pair is read back from LDS as in the item
            // form below.  An ascending chunk's window is the E items from its
            // first position's item; a descending chunk's is the E items below
            // and including it, reversed in registers.  A window of pads only
            // reads the group's first 16 bytes.  A window that is not inside
            // the group (at most one per slice end, or its first or last
            // items) loads its real items one at a time, its pads reading the
            // group's first item; pads become the maximum key.  The group's
            // byte span is at most 2^31, so a window starting below the group
            // (a descending chunk at its first items) wraps to an offset above
            // the limit.
            constexpr int E = chunk_items<K>();
            constexpr int NL = IPT / E;         // 16-byte loads per thread
            constexpr int CPL = 64 * E / IPT;   // chunks per load of a wave
            static_assert(IPT % E == 0 && (64 * E) % IPT == 0, "whole 16-byte windows, whole chunks per load");
            typedef uint32_t u32x4a __attribute__((ext_vector_type(4), aligned(4)));
            typedef uint64_t u64x2a __attribute__((ext_vector_type(2), aligned(8)));
            const int wbase = (tid >> 6) * 64 * IPT;
            uint2 *s_src2 = (uint2 *)(sk + wbase);
            s_src2[lane] = make_uint2(cb_lo, cmeta);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const uint2 *sq = s_src2 + (E * lane) / IPT;
            const char *gb = (const char *)in + gbase * ES;  // uniform: scalar registers
            const int r0 = (E * lane) & (IPT - 1);
            const uint32_t rb = (uint32_t)r0 * ES;
            const uint32_t lim = (uint32_t)(gspan - 16);
            constexpr uint32_t WIN = (uint32_t)((E - 1) * ES);
            constexpr uint32_t EM = (1u << E) - 1;
            K x[NL][E];
            uint32_t tb[NL];
#pragma unroll
            for (int c = 0; c < NL; c++) {
                const uint2 src = sq[CPL * c];
                const uint32_t bb = src.x;
                const uint32_t t = src.y >> r0;  // bit j: position r0 + j is a pad
                // 0 ascending, ~0 descending (bb = B + 1, odd)
                const uint32_t d = (uint32_t)__builtin_amdgcn_sbfe((int)bb, 0u, 1u);
                // ascending: B + r0 * ES; descending: B - r0 * ES - WIN
                const uint32_t start = ((rb ^ d) + bb) - (d & WIN);
                const uint32_t off = (t & EM) == EM ? 0u : start;
                tb[c] = (t & EM) | (d & 0x80000000u);  // the window's pad bits; the top bit: descending
                if (off <= lim) {
                    if constexpr (E == 4) {
                        const u32x4a q = *(const u32x4a *)(gb + off);
                        x[c][0] = (K)q.x; x[c][1] = (K)q.y; x[c][2] = (K)q.z; x[c][3] = (K)q.w;
                    } else {
                        const u64x2a q = *(const u64x2a *)(gb + off);
                        x[c][0] = (K)q.x; x[c][1] = (K)q.y;
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < E; i++) {
                        // memory item i of the window is position r0 + (d ? E - 1 - i : i)
                        const int j = d ? E - 1 - i : i;
                        const bool pad = (t >> j) & 1u;
                        x[c][i] = *(const K *)(gb + (pad ? 0u : start + (uint32_t)i * ES));
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < NL; c++) {
                if ((int)tb[c] < 0) {
                    // descending: memory order reversed into position order
                    asm volatile("");
#pragma unroll
                    for (int i = 0; i < E / 2; i++) swap_regs(x[c][i], x[c][E - 1 - i]);
                }
                if (tb[c] & EM) {
                    // a window with pads (at most two per slice: a branch that
                    // most waves skip)
                    asm volatile("");
#pragma unroll
                    for (int i = 0; i < E; i++) {
                        // the builtin returns an unsigned int: through int, so
                        // the pad mask sign-extends to 64-bit keys
                        const int p = (int)__builtin_amdgcn_sbfe((int)tb[c], (uint32_t)i, 1u);
                        x[c][i] |= (K)(typename std::make_signed<K>::type)p;
                    }
                }
            }
            // every pair read of this wave precedes the item writes that
            // overwrite the pairs
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            // window c of the lane sits at lds_off(wbase + E * lane) ^
            // lds_off(64 * E * c) = + 1024 * c bytes (XOR-linear swizzle,
            // disjoint bits)
            char *wp = (char *)sk + lds_off<K>(wbase + E * lane);
#pragma unroll
            for (int c = 0; c < NL; c++) {
                if constexpr (E == 4) {
                    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                    *(u32x4 *)(wp + 1024 * c) = u32x4{(uint32_t)x[c][0], (uint32_t)x[c][1], (uint32_t)x[c][2], (uint32_t)x[c][3]};
                } else {
                    typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
                    *(u64x2 *)(wp + 1024 * c) = u64x2{(uint64_t)x[c][0], (uint64_t)x[c][1]};
                }
            }
        }
    }
    if (HV || !vec16) {
        // Wave w loads LDS positions [w * 64 * IPT, (w + 1) * 64 * IPT), the
        // chunks of its own lanes: load i covers positions wbase + BLK * (i %
        // NB) + 64 * (i / NB) + lane (BLK = the index span the LDS swizzle
        // repeats over, so a lane's LDS index is its group's swizzled base plus
        // a multiple of BLK), and the lane fetches its chunk's source from the
        // owning lane.  All loads are in flight at once; pads read the group's
        // first item and keep the maximum key.
        constexpr int BLK = sizeof(K) == 4 ? 256 : 128;
        constexpr int NB = 64 * IPT / BLK;
        static_assert(NB >= 1 && 64 * IPT % BLK == 0, "a wave's load range is whole swizzle blocks");
        const int r = lane & (IPT - 1);
        const int wbase = (tid >> 6) * 64 * IPT;
        K kk[IPT];
        uint64_t vv[IPT];
        if (narrow) {
            // every thread leaves its chunk's (cb_lo, cmeta) in LDS and the
            // lanes of each load read their chunk's pair back (one 8-byte read
            // that 64 / IPT lanes share, instead of three ds_bpermute); a
            // descending chunk's cb_lo = B + 1 is odd (B is a multiple of the
            // item size), so d is its low bit.  The pairs sit at the start of
            // the wave's own tile range, which no other wave touches during the
            // load phase and which this wave overwrites with items only after
            // its last pair read (a wave's LDS accesses complete in order; the
            // fences keep the compiler from moving a read of another lane's
            // pair above the writes, or an item write above the reads).
            uint2 *s_src2 = (uint2 *)(sk + (tid & ~63) * IPT);
            s_src2[lane] = make_uint2(cb_lo, cmeta);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const uint2 *sq = s_src2 + lane / IPT;
            const char *gb = (const char *)in + gbase * ES;  // uniform: scalar registers
            const uint32_t rb = (uint32_t)r * ES;
#pragma unroll
            for (int i = 0; i < IPT; i++) {
                // position within the wave's range: BLK * (i % NB) + 64 * (i / NB) + lane
                const uint2 src = sq[(BLK / IPT) * (i % NB) + (64 / IPT) * (i / NB)];
                const uint32_t bb = src.x, pm = src.y;
                // -1 for a pad, else 0: a pad reads the group's first item and
                // becomes the maximum key (value 0)
                const int p = __builtin_amdgcn_sbfe((int)pm, (uint32_t)r, 1u);
                if constexpr (HV) {
                    // KV slices are ascending (d = 0)
                    const uint32_t off = (rb + bb) & ~(uint32_t)p;
                    const ulonglong2 t = *(const ulonglong2 *)(gb + off);
                    kk[i] = (K)t.x | (K)(int64_t)p;
                    vv[i] = t.y & ~(uint64_t)(int64_t)p;
                } else {
                    const uint32_t d = (uint32_t)__builtin_amdgcn_sbfe((int)bb, 0u, 1u);
                    const uint32_t off = ((rb ^ d) + bb) & ~(uint32_t)p;
                    kk[i] = *(const K *)(gb + off) | (K)(typename std::make_signed<K>::type)p;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < IPT; i++) {
                const int pw = BLK * (i % NB) + 64 * (i / NB) + lane;
                const int sl = (pw / IPT) << 2;
                const uint32_t blo = (uint32_t)__builtin_amdgcn_ds_bpermute(sl, (int)cb_lo);
                const uint32_t bhi = (uint32_t)__builtin_amdgcn_ds_bpermute(sl, (int)cb_hi);
                const uint32_t m = (uint32_t)__builtin_amdgcn_ds_bpermute(sl, (int)cmeta);
                const bool real = r >= (int)(m & 255) && r < (int)((m >> 8) & 255);
                // branch-free: offset +r or -r (conditional negate by the sign
                // mask), then one select of the address
                const int neg = -(int)(m >> 16);
                const int64_t off = (int64_t)((r ^ neg) - neg);
                const uint64_t b = ((uint64_t)bhi << 32) | blo;
                const uint64_t mask = (uint64_t)0 - (uint64_t)real;
                const uint64_t src = gbase + ((b + (uint64_t)off - gbase) & mask);
                K x;
                uint64_t xv = 0;
                GMem<K, HV>::load(in, src, x, xv);
                kk[i] = real ? x : KMAX;
                vv[i] = real ? xv : 0;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const int e = lsw<K>(wbase + 64 * (i / NB) + lane) + BLK * (i % NB);
            sk[e] = kk[i];
            if constexpr (HV) sv[e] = vv[i];
        }
    }
    if constexpr (HV) __syncthreads();
    if constexpr (!HV) {
        // the slices' block splitters: an ascending slice's block is this
        // thread's chunk; an odd slice is stored descending, so the chunk at
        // c0 is its block (P[j0+1] - c0) / IPT - 1 read backwards.  A wave's
        // threads' chunks are the tile range this wave has just written, so a
        // wavefront fence (no barrier) orders the reads after the writes.
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (live) {
            const K first = sk[lsw<K>(c0)], last = sk[lsw<K>(c0 + IPT - 1)];
            if (j0 & 1) {
                const int idx = (s_P[j0] + s_P[j0 + 1] - c0) / IPT - 1;
                put_pair(s_spFL + 2 * idx, last, first);
            } else {
                put_pair(s_spFL + 2 * tid, first, last);
            }
        }
        __syncthreads();
    }
    K r[IPT];
    uint64_t v[IPT];
#pragma unroll 1
    for (int lev = 1; lev <= LOGKW; lev++) {
        int end = 0, diag = 0;
        if (live) {
            if constexpr (!HV) {
                const int4 pr = s_pair[KW - 2 * (KW >> lev) + (j0 >> lev)];
                const int a_off = pr.x, mid = pr.y;
                end = pr.z;
                diag = c0 - a_off;
                const int top = __builtin_amdgcn_readfirstlane(pr.w);
                const int a = lds_corank_split_rev<IPT>(sk, s_spFL, a_off, mid - a_off, mid, end - mid, end - 1, diag,
                                                        top);
                lds_merge_rev<IPT, true>(sk, a_off + a, end - 1 - (diag - a), 0, IPT, r);
            } else {
                const int hv = 1 << (lev - 1);
                const int s0 = (j0 >> lev) << lev;
                const int a_off = s_P[s0], mid = s_P[s0 + hv];
                end = s_P[s0 + 2 * hv];
                diag = c0 - a_off;
                const int top = __builtin_amdgcn_readfirstlane(s_top[lev]);
                // KV: ascending runs; merges test exhaustion against the real
                // lengths, so a pad never displaces a real maximum-key item and
                // the merged pair is its real items, then pads
                const int la = s_R[s0 + hv] - s_R[s0], lb = s_R[s0 + 2 * hv] - s_R[s0 + hv];
                const int nreal = la + lb;
                int c = nreal - diag;
                c = c < 0 ? 0 : c > IPT ? IPT : c;
                if (c > 0) {
                    const int a = lds_corank_lift<K>(sk, a_off, la, mid, lb, diag, top);
                    if (c == IPT) lds_merge_serial<IPT, true, true>(sk, sv, a_off, la, mid, lb, a, diag - a, 0, IPT, r, v);
                    else lds_merge_serial<IPT, true, false>(sk, sv, a_off, la, mid, lb, a, diag - a, 0, c, r, v);
                }
                if (c < IPT) {
                    // outputs past the pair's real items are pads (threads at
                    // a pair's real end or in its pad region: a branch most
                    // waves skip)
                    asm volatile("");
#pragma unroll
                    for (int i = 0; i < IPT; i++) {
                        r[i] = i < c ? r[i] : KMAX;
                        v[i] = i < c ? v[i] : 0;
                    }
                }
            }
        }
        __syncthreads();
        if (live) {
            if constexpr (!HV) {
                // this thread's outputs are block (c0 - pair start) / IPT of
                // the merged run at the pair's offset: splitter index tid
                if (lev < LOGKW) put_pair(s_spFL + 2 * tid, r[0], r[IPT - 1]);
            }
            // the next level's B sides are stored descending (key-only): a
            // descending thread reverses its registers (exec-masked moves, a
            // branch that waves holding only ascending runs skip) and writes
            // them ascending from end - diag - IPT, so a wave holding both
            // orientations issues one set of 16-byte LDS writes, not two
            const bool dsc = !HV && lev < LOGKW && ((j0 >> lev) & 1);
            if (dsc) {
                asm volatile("");
                // in place, one v_swap_b32 per pair (the compiler otherwise
                // copies through temporaries, or pairs the moves into packed
                // 64-bit moves)
#pragma unroll
                for (int i = 0; i < IPT / 2; i++) swap_regs(r[i], r[IPT - 1 - i]);
            }
            lds_put<IPT, HV>(sk, sv, dsc ? end - diag - IPT : c0, r, v);
        }
        __syncthreads();
    }
    // item tid + k * BLOCK sits at LDS index lsw(tid) + k * BLOCK: BLOCK is a
    // multiple of the span the swizzle repeats over
    static_assert(!HV || BLOCK % (sizeof(K) == 4 ? 256 : 128) == 0, "BLOCK: whole swizzle blocks");
    // o0 is the same in every lane: kept in scalar registers, the stores take
    // it as their base and a 32-bit lane offset
    const uint64_t o0u = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(o0 >> 32)) << 32) |
                         (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)o0);
    if constexpr (HV) {
        const K *sp0 = sk + lsw<K>(tid);
        ulonglong2 *ob = (ulonglong2 *)out + o0u;
        const uint64_t *vp0 = sv + lsw<uint64_t>(tid);
        const CodecMask<K> m = codec_mask<K>(cout);
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            if (tid + k * BLOCK < cnt) {
                ulonglong2 q;
                q.x = (unsigned long long)dec_bf(sp0[k * BLOCK], m);
                q.y = (unsigned long long)vp0[k * BLOCK];
                ob[(uint32_t)(tid + k * BLOCK)] = q;
            }
        }
    } else {
        K *ob = (K *)out + o0u;
        // the codec is uniform: one branch picks the store loop
        if (cout == 0) store_chunks<BLOCK, IPT, 0>(sk, ob, cnt, tid);
        else if (cout == 1) store_chunks<BLOCK, IPT, 1>(sk, ob, cnt, tid);
        else store_chunks<BLOCK, IPT, sizeof(K) == 4 ? 2 : 3>(sk, ob, cnt, tid);
    }
    if (sp.smp) write_samples<BLOCK>(sk, cnt, o0, sp.smp, sp.rlog2, sp.kw, sp.slog2);
}

template <typename K, bool HV, int KW, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void k_kway_merge(const void *__restrict__ in, void *__restrict__ out, uint64_t n,
                                                      int rlog2, uint64_t ntg, uint32_t ntg_m, uint64_t nt_last,
                                                      int cout, const uint64_t *__restrict__ tab, SampleSpec sp,
                                                      uint32_t *__restrict__ err) {
    kway_merge_tile<K, HV, KW, BLOCK, IPT>(in, out, n, rlog2, ntg, ntg_m, nt_last, cout, tab, sp, err, blockIdx.x);
}

template <typename K, bool HV, int KW, int BLOCK, int IPT>
static hipError_t launch_kway(const void *in, void *out, uint64_t n, int rlog2, int slog2, int cout, uint64_t *tab,
                              const void *s0, const void *sm, const SampleSpec *next, hipStream_t s, hipEvent_t mid,
                              uint32_t *err) {
    constexpr uint64_t C = (uint64_t)BLOCK * IPT;
    if (n == 0) return hipSuccess;
    const uint64_t Q = kw_q(C, IPT, KW, slog2);
    if (Q == 0 || rlog2 - slog2 > 27 || rlog2 < slog2) return hipErrorInvalidValue;
    const uint64_t GR = (1ull << rlog2) * KW;
    const uint64_t ntg = kw_group_tiles(GR >> slog2, Q);
    const uint64_t nt_last = kw_group_tiles(kw_last_samples(n, rlog2, slog2, KW), Q);
    const uint64_t ntiles = (n + GR - 1) / GR * ntg - ntg + nt_last;
    const uint64_t tpb = 4 * (uint64_t)kw_bounds_tpw<KW>();  // tiles per 256-thread block
    hipLaunchKernelGGL((k_kway_bounds<K, HV, KW>), dim3((unsigned)((ntiles + tpb - 1) / tpb)), dim3(256), 0, s, in, n,
                       rlog2, slog2, Q, ntg, ntiles, s0, sm, tab);
    if (mid) (void)hipEventRecord(mid, s);
    const SampleSpec none = {nullptr, 0, 0, 0};
    hipLaunchKernelGGL((k_kway_merge<K, HV, KW, BLOCK, IPT>), dim3((unsigned)ntiles), dim3(BLOCK), 0, s, in, out, n,
                       rlog2, ntg, kw_div_magic(ntg), nt_last, cout, (const uint64_t *)tab, next ? *next : none, err);
    return hipGetLastError();
}

// Every group of KW runs of 2^rlog2 items (ragged at n) merged by one
// workgroup, with no split table: the group, its slices padded to IPT, must fit
// the BLOCK * IPT items of LDS.  Merges the samples of a K-way pass when a
// group's samples fit one tile.
template <typename K, bool HV, int KW, int BLOCK, int IPT>
static hipError_t launch_kway_whole(const void *in, void *out, uint64_t n, int rlog2, uint32_t *err, hipStream_t s) {
    constexpr uint64_t C = (uint64_t)BLOCK * IPT;
    if (n == 0) return hipSuccess;
    const uint64_t R = 1ull << rlog2;
    if (rlog2 > 30 || (uint64_t)KW * ((R + IPT - 1) / IPT * IPT) > C) return hipErrorInvalidValue;
    const uint64_t groups = (n + R * KW - 1) / (R * KW);
    const SampleSpec none = {nullptr, 0, 0, 0};
    hipLaunchKernelGGL((k_kway_merge<K, HV, KW, BLOCK, IPT>), dim3((unsigned)groups), dim3(BLOCK), 0, s, in, out, n,
                       rlog2, (uint64_t)1, kw_div_magic(1), (uint64_t)1, 0, (const uint64_t *)nullptr, none, err);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Batched form: one pass of many segments' K-way plans (segments of a batched
// sort), all of fan-in KW at this pass.  Runs are 2^rlog2 items in every
// listed segment (each pass before a segment's last is of the full fan-in), so
// the tile geometry -- Q, tiles per full group -- is the same for all; a tile
// map gives each global tile its segment and its index there.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_kway_tilemap(const uint32_t *__restrict__ list,
                                                      const uint32_t *__restrict__ tfirst, uint2 *__restrict__ tmap) {
    const uint32_t i = blockIdx.x;
    const uint32_t a = tfirst[i], b = tfirst[i + 1], seg = list[i];
    for (uint32_t k = threadIdx.x; a + k < b; k += blockDim.x) tmap[a + k] = make_uint2(seg, k);
}

hipError_t launch_kway_tilemap(const uint32_t *list, const uint32_t *tfirst, uint32_t nlist, uint2 *tmap,
                               hipStream_t s) {
    if (nlist == 0) return hipSuccess;
    hipLaunchKernelGGL(k_kway_tilemap, dim3(nlist), dim3(256), 0, s, list, tfirst, tmap);
    return hipGetLastError();
}

template <typename K, bool HV, int KW>
__global__ __launch_bounds__(256) void k_seg_kway_bounds(const KSegH *__restrict__ segs, const uint2 *__restrict__ tmap,
                                                         uint64_t ntiles, int pass, int rlog2, int slog2, uint64_t Q,
                                                         uint64_t ntg, int smi) {
    constexpr int TPW = kw_bounds_tpw<KW>();
    const int h = (int)(threadIdx.x & 63) / (64 / TPW);
    const uint64_t t0 = ((uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)) * TPW;
    if (t0 >= ntiles) return;  // uniform in the wave
    const uint64_t tg = t0 + (uint64_t)h;
    const bool valid = tg < ntiles;
    const uint2 m = tmap[valid ? tg : ntiles - 1];
    const KSegH &g = segs[m.x];
    kway_bounds_tile<K, HV, KW>(g.buf[pass & 1], g.len, rlog2, slog2, Q, ntg, m.y, valid, g.smp[0], g.smp[smi],
                                g.tab);
}

template <typename K, bool HV, int KW, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void k_seg_kway_merge(const KSegH *__restrict__ segs,
                                                          const uint2 *__restrict__ tmap, int pass, int rlog2,
                                                          int slog2, uint64_t Q, uint64_t ntg, uint32_t ntg_m,
                                                          uint32_t *__restrict__ err) {
    const uint2 m = tmap[blockIdx.x];
    const KSegH &g = segs[m.x];
    const uint64_t nt_last = kw_group_tiles(kw_last_samples(g.len, rlog2, slog2, KW), Q);
    const bool last = pass + 1 >= (int)g.passes;
    SampleSpec sp = {nullptr, 0, 0, 0};
    if (!last) sp = {g.smp[0], rlog2 + (int)g.lv[pass], 1 << g.lv[pass + 1], (int)g.sl[pass + 1]};
    kway_merge_tile<K, HV, KW, BLOCK, IPT>(g.buf[pass & 1], g.buf[(pass + 1) & 1], g.len, rlog2, ntg, ntg_m, nt_last,
                                           last ? (int)g.cout : 0, g.tab, sp, err, m.y);
}

template <typename K, bool HV, int KW, int BLOCK, int IPT>
static hipError_t launch_seg_kway(const KSegH *segs, const uint2 *tmap, uint64_t ntiles, int pass, int rlog2,
                                  int slog2, int smi, uint32_t *err, hipStream_t s) {
    constexpr uint64_t C = (uint64_t)BLOCK * IPT;
    if (ntiles == 0) return hipSuccess;
    const uint64_t Q = kw_q(C, IPT, KW, slog2);
    if (Q == 0 || rlog2 - slog2 > 27 || rlog2 < slog2) return hipErrorInvalidValue;
    const uint64_t ntg = kw_group_tiles(((1ull << rlog2) * KW) >> slog2, Q);
    const uint64_t tpb = 4 * (uint64_t)kw_bounds_tpw<KW>();  // tiles per 256-thread block
    hipLaunchKernelGGL((k_seg_kway_bounds<K, HV, KW>), dim3((unsigned)((ntiles + tpb - 1) / tpb)), dim3(256), 0, s,
                       segs, tmap, ntiles, pass, rlog2, slog2, Q, ntg, smi);
    hipLaunchKernelGGL((k_seg_kway_merge<K, HV, KW, BLOCK, IPT>), dim3((unsigned)ntiles), dim3(BLOCK), 0, s, segs,
                       tmap, pass, rlog2, slog2, Q, ntg, kw_div_magic(ntg), err);
    return hipGetLastError();
}

// The samples of every listed segment's groups of KW runs of 2^rlog2 sample
// records (runs of the segment's sample array smp[0], whose length follows
// from the segment's length and this pass's spacing) merged into smp[1], one
// workgroup per group (tmap: {segment, group}); the padded group must fit one
// tile, which the host checks.
template <typename K, bool HV, int KW, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void k_seg_kway_whole(const KSegH *__restrict__ segs,
                                                          const uint2 *__restrict__ tmap, int pass, int rlog2,
                                                          uint32_t *__restrict__ err) {
    const uint2 m = tmap[blockIdx.x];
    const KSegH &g = segs[m.x];
    const int sl = (int)g.sl[pass];
    const uint64_t ns = (g.len + (1ull << sl) - 1) >> sl;
    const SampleSpec none = {nullptr, 0, 0, 0};
    kway_merge_tile<K, HV, KW, BLOCK, IPT>(g.smp[0], g.smp[1], ns, rlog2, 1, kw_div_magic(1), 1, 0, nullptr, none, err,
                                           m.y);
}

template <typename K, bool HV, int KW, int BLOCK, int IPT>
static hipError_t launch_seg_kway_whole(const KSegH *segs, const uint2 *tmap, uint64_t ngroups, int pass, int rlog2,
                                        uint32_t *err, hipStream_t s) {
    constexpr uint64_t C = (uint64_t)BLOCK * IPT;
    if (ngroups == 0) return hipSuccess;
    const uint64_t R = 1ull << rlog2;
    if (rlog2 > 30 || (uint64_t)KW * ((R + IPT - 1) / IPT * IPT) > C) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_seg_kway_whole<K, HV, KW, BLOCK, IPT>), dim3((unsigned)ngroups), dim3(BLOCK), 0, s, segs,
                       tmap, pass, rlog2, err);
    return hipGetLastError();
}

#define OGS_KWAY_SHAPES_32(X) X(512, 32) X(256, 32) X(512, 16) X(256, 16) X(1024, 32) X(1024, 16)
#define OGS_KWAY_SHAPES_64(X) X(256, 16) X(512, 16) X(256, 8) X(512, 8) X(1024, 16) X(1024, 8)
#define OGS_KWAY_SHAPES_KV(X) X(512, 8) X(256, 16) X(256, 8) X(512, 4) X(1024, 4)
#define OGS_KX(K, HV, B, I)                                                                                 \
    {B, I, {&launch_kway<K, HV, 2, B, I>, &launch_kway<K, HV, 4, B, I>, &launch_kway<K, HV, 8, B, I>,       \
            &launch_kway<K, HV, 16, B, I>, &launch_kway<K, HV, 32, B, I>},                                  \
     {&launch_seg_kway<K, HV, 2, B, I>, &launch_seg_kway<K, HV, 4, B, I>, &launch_seg_kway<K, HV, 8, B, I>, \
      &launch_seg_kway<K, HV, 16, B, I>, &launch_seg_kway<K, HV, 32, B, I>},                                \
     {&launch_kway_whole<K, HV, 2, B, I>, &launch_kway_whole<K, HV, 4, B, I>,                               \
      &launch_kway_whole<K, HV, 8, B, I>, &launch_kway_whole<K, HV, 16, B, I>,                              \
      &launch_kway_whole<K, HV, 32, B, I>},                                                                 \
     {&launch_seg_kway_whole<K, HV, 2, B, I>, &launch_seg_kway_whole<K, HV, 4, B, I>,                       \
      &launch_seg_kway_whole<K, HV, 8, B, I>, &launch_seg_kway_whole<K, HV, 16, B, I>,                      \
      &launch_seg_kway_whole<K, HV, 32, B, I>}},

static const KwayShapeFns kKway32[] = {
#define X(B, I) OGS_KX(uint32_t, false, B, I)
    OGS_KWAY_SHAPES_32(X)
#undef X
};
static const KwayShapeFns kKway64[] = {
#define X(B, I) OGS_KX(uint64_t, false, B, I)
    OGS_KWAY_SHAPES_64(X)
#undef X
};
static const KwayShapeFns kKwayKV[] = {
#define X(B, I) OGS_KX(uint64_t, true, B, I)
    OGS_KWAY_SHAPES_KV(X)
#undef X
};

const KwayShapeFns *kway_shapes(KeyClass kc, int *count) {
    switch (kc) {
    case KC_32: *count = (int)(sizeof(kKway32) / sizeof(kKway32[0])); return kKway32;
    case KC_64: *count = (int)(sizeof(kKway64) / sizeof(kKway64[0])); return kKway64;
    default: *count = (int)(sizeof(kKwayKV) / sizeof(kKwayKV[0])); return kKwayKV;
    }
}

}  // namespace ogs
