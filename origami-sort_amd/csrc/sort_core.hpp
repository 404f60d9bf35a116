// sort_core.hpp -- device building blocks of the gfx950 mergesort.
//
// Reference pipeline (arif-arman/origami-sort, Origami/):
//   phase1_sort / in_register_sort*  (sorter.h:33-1556, 3405-3828)  -> per-lane network + wave bitonic
//   mergebl2_* / merge_series        (merger.h:45-4228)             -> merge-path co-rank merges in LDS
//   Partition_Ptrs                   (Partition.h:117-166)          -> co-rank searches (2-way and k-way)
//   origami_utils::store<stream>     (merge_utils.h:61-81)          -> coalesced / non-temporal stores
//
// Keys travel in an "ordered" unsigned form (u32 or u64) whose unsigned order is
// the requested order; the codec is applied on the first load and undone on the
// last store, so signed and float keys cost no extra pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ogs {

// ----------------------------------------------------------------------------
// Key codecs.  0: identity, 1: flip the sign bit (two's-complement order),
// 2: float32 IEEE totalOrder, 3 (8-byte items only): swap the two 32-bit
// halves, so an 8-byte {u32 key, u32 value} record (little-endian: the key in
// the low half) sorts as the u64 key << 32 | value.  The reference compares
// u32 unsigned (avx2_utils.h:129-133) and i64 signed (avx2_utils.h:140-145).
// ----------------------------------------------------------------------------
__device__ __forceinline__ uint64_t swap_halves(uint64_t x) { return (x << 32) | (x >> 32); }
__device__ __forceinline__ uint32_t enc_key(uint32_t x, int c) {
    if (c == 1) return x ^ 0x80000000u;
    if (c == 2) return x ^ ((uint32_t)((int32_t)x >> 31) | 0x80000000u);
    return x;
}
__device__ __forceinline__ uint32_t dec_key(uint32_t x, int c) {
    if (c == 1) return x ^ 0x80000000u;
    if (c == 2) return x ^ (~(uint32_t)((int32_t)x >> 31) | 0x80000000u);
    return x;
}
__device__ __forceinline__ uint64_t enc_key(uint64_t x, int c) {
    return c == 1 ? x ^ (1ull << 63) : c == 3 ? swap_halves(x) : x;
}
__device__ __forceinline__ uint64_t dec_key(uint64_t x, int c) { return enc_key(x, c); }

// The same codecs without branches, for loops over many keys with one codec:
// enc(x) = x ^ ((sx & a) | b), dec(x) = x ^ ((~sx & a) | b), with sx all ones
// when the top bit of x is set; (a, b) = (0, 0) identity, (0, top) sign flip,
// (~0, top) float32 totalOrder; `rot` swaps the halves of an 8-byte key
// (codec 3, an involution).
template <typename K> struct CodecMask {
    K a, b;
    bool rot;
};
template <typename K> __device__ __forceinline__ CodecMask<K> codec_mask(int c) {
    constexpr K TOP = (K)1 << (sizeof(K) * 8 - 1);
    CodecMask<K> m;
    m.a = (sizeof(K) == 4 && c == 2) ? (K)~(K)0 : (K)0;
    m.b = (c == 1 || c == 2) ? TOP : (K)0;
    m.rot = sizeof(K) == 8 && c == 3;
    return m;
}
template <typename K> __device__ __forceinline__ K sign_all(K x) {
    if constexpr (sizeof(K) == 4) return (K)((int32_t)x >> 31);
    else return (K)((int64_t)x >> 63);
}
template <typename K> __device__ __forceinline__ K rot_if(K x, bool rot) {
    if constexpr (sizeof(K) == 8) return rot ? (K)swap_halves((uint64_t)x) : x;
    else return x;
}
template <typename K> __device__ __forceinline__ K enc_bf(K x, CodecMask<K> m) {
    return rot_if(x, m.rot) ^ ((sign_all(x) & m.a) | m.b);
}
template <typename K> __device__ __forceinline__ K dec_bf(K x, CodecMask<K> m) {
    return rot_if(x ^ ((~sign_all(x) & m.a) | m.b), m.rot);
}

template <typename K> struct KeyMax;
template <> struct KeyMax<uint32_t> { static constexpr uint32_t v = 0xFFFFFFFFu; };
template <> struct KeyMax<uint64_t> { static constexpr uint64_t v = ~0ull; };

// ----------------------------------------------------------------------------
// Global memory items.  Key-only arrays hold K; KV arrays hold 16-byte
// {u64 key, u64 value} records (KeyValue<i64,i64>, commons.h:51-72).
// ----------------------------------------------------------------------------
template <typename K, bool HV>
struct GMem {
    static __device__ __forceinline__ void load(const void *p, uint64_t i, K &k, uint64_t &v) {
        if constexpr (HV) {
            const ulonglong2 *q = (const ulonglong2 *)p;
            ulonglong2 t = q[i];
            k = (K)t.x;
            v = t.y;
        } else {
            k = ((const K *)p)[i];
            (void)v;
        }
    }
    static __device__ __forceinline__ K load_key(const void *p, uint64_t i) {
        if constexpr (HV) return (K)((const ulonglong2 *)p)[i].x;
        else return ((const K *)p)[i];
    }
    static __device__ __forceinline__ void store(void *p, uint64_t i, K k, uint64_t v) {
        if constexpr (HV) {
            ulonglong2 t;
            t.x = (unsigned long long)k;
            t.y = v;
            ((ulonglong2 *)p)[i] = t;
        } else {
            ((K *)p)[i] = k;
            (void)v;
        }
    }
};

// LDS layout: an XOR swizzle of the 16-byte chunk bits of the byte offset by
// higher bits.  A lane writing its IPT consecutive items as 16-byte stores
// then hits distinct banks from its neighbours (plain layout: lanes 0 and 4
// collide), and lane-consecutive accesses stay a permutation of the banks.
//   4-byte items: byte bits 4..7 ^= byte bits 6..9 (item bits 2..5 ^= 4..7);
//   8-byte items: byte bits 4..6 ^= byte bits 7..9 (item bits 1..3 ^= 4..6).
// Either way the byte offset of item i is (i * W) ^ (i & M) -- a shift and one
// v_bitop3 from the item index, which merge loops keep as their pointers --
// the map is XOR-linear (so an aligned block of items is its first item's
// offset XOR constants), it repeats every 256 (4-byte) or 128 (8-byte) items,
// and it never moves an item out of its 16-byte chunk.
template <typename T> __device__ __forceinline__ int lsw(int i) {
    if constexpr (sizeof(T) == 4) return i ^ ((i >> 2) & 0x3C);
    else return i ^ (((i >> 4) & 7) << 1);
}
// items per 16-byte chunk
template <typename T> constexpr int chunk_items() { return 16 / (int)sizeof(T); }

// Swizzled byte offset of LDS item i (= lsw<K>(i) * sizeof(K)).
template <typename K> __device__ __forceinline__ int lds_off(int i) {
    if constexpr (sizeof(K) == 4) return (i << 2) ^ (i & 0xF0);
    else return (i << 3) ^ (i & 0x70);
}
// The swizzle on byte offsets b = i * sizeof(K).
template <typename K> __device__ __forceinline__ int lsw_bytes(int b) {
    if constexpr (sizeof(K) == 4) return b ^ ((b >> 2) & 0xF0);
    else return b ^ ((b >> 3) & 0x70);
}
template <typename K> __device__ __forceinline__ K lds_at_bytes(const K *sk, int b) {
    return *(const K *)((const char *)sk + lsw_bytes<K>(b));
}
template <typename K> __device__ __forceinline__ K lds_at(const K *sk, int i) {
    return *(const K *)((const char *)sk + lds_off<K>(i));
}

// ----------------------------------------------------------------------------
// Compare-exchange (SWAP / SWAPKV2, commons.h:111-149): a <- smaller, b <- larger.
// ----------------------------------------------------------------------------
template <bool HV, typename K>
__device__ __forceinline__ void cex(K &a, K &b, uint64_t &va, uint64_t &vb) {
    if constexpr (!HV) {
        K lo = a < b ? a : b;
        b = a < b ? b : a;
        a = lo;
        (void)va; (void)vb;
    } else {
        bool sw = b < a;
        K ta = sw ? b : a, tb = sw ? a : b;
        uint64_t ua = sw ? vb : va, ub = sw ? va : vb;
        a = ta; b = tb; va = ua; vb = ub;
    }
}

// Batcher odd-even merge sort network over N registers, fully unrolled
// (the reference uses fixed tables, e.g. in_register_sort8/16, sorter.h:40-66).
template <int N, bool HV, typename K>
__device__ __forceinline__ void lane_network(K (&r)[N], uint64_t (&v)[N]) {
#pragma unroll
    for (int p = 1; p < N; p <<= 1)
#pragma unroll
        for (int k = p; k >= 1; k >>= 1)
#pragma unroll
            for (int j = k % p; j + k < N; j += 2 * k)
#pragma unroll
                for (int i = 0; i < k; i++)
                    if (i + j + k < N && (i + j) / (2 * p) == (i + j + k) / (2 * p))
                        cex<HV>(r[i + j], r[i + j + k], v[i + j], v[i + j + k]);
}

// ----------------------------------------------------------------------------
// Cross-lane exchange with a constant XOR mask.  DPP quad_perm for 1/2/3,
// DPP row_half_mirror (xor 7) and row_mirror (xor 15), ds_swizzle bitmask mode
// for the other masks below 32, ds_bpermute beyond.
// ----------------------------------------------------------------------------
// The DPP forms are written as update_dpp with bound_ctrl set: every lane has
// a source under these patterns, so the zero fill never applies, and the
// compiler can then fold the move into the min / max that consumes it
// (v_min_u32_dpp), one VALU instruction fewer per item per exchange.
template <int CTRL> __device__ __forceinline__ uint32_t dpp32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, true);
}
template <int M>
__device__ __forceinline__ uint32_t xshfl32(uint32_t x) {
    if constexpr (M == 1) return dpp32<0xB1>(x);
    else if constexpr (M == 2) return dpp32<0x4E>(x);
    else if constexpr (M == 3) return dpp32<0x1B>(x);
    else if constexpr (M == 7) return dpp32<0x141>(x);
    else if constexpr (M == 15) return dpp32<0x140>(x);
    else if constexpr (M < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, (M << 10) | 0x1F);
    else return (uint32_t)__shfl_xor((int)x, M, 64);
}
template <int M> __device__ __forceinline__ uint32_t xshfl(uint32_t x) { return xshfl32<M>(x); }
template <int M> __device__ __forceinline__ uint64_t xshfl(uint64_t x) {
    uint32_t lo = xshfl32<M>((uint32_t)x), hi = xshfl32<M>((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// One exchange step between lane pairs: the lane whose `upper` flag is clear
// keeps the smaller item, its partner the larger; on equal keys both keep their
// own item, so no item is duplicated or lost.
template <bool HV, typename K>
__device__ __forceinline__ void lane_pair_keep(K &mine, uint64_t &vm, K other, uint64_t vo, bool upper) {
    if constexpr (!HV && sizeof(K) == 4) {
        // key-only 32-bit: the median of (mine, other, 0 or ~0) is the min
        // for the lower lane and the max for the upper one -- one v_med3_u32
        // (the partner's item is one DPP move or swizzle before it)
        const K c = upper ? KeyMax<K>::v : (K)0;
        const K lo = other < mine ? other : mine;
        const K hi = other < mine ? mine : other;
        const K t = hi < c ? hi : c;
        mine = lo < t ? t : lo;
        (void)vm; (void)vo;
    } else if constexpr (!HV) {
        // key-only 64-bit: one compare; a lane takes its partner's item when
        // that is smaller (lower lane) or not smaller (upper lane) -- equal
        // keys are identical bits, so either choice keeps every item once
        const bool take = (other < mine) != upper;
        mine = take ? other : mine;
        (void)vm; (void)vo;
    } else {
        bool keep = upper ? !(mine < other) : !(other < mine);
        mine = keep ? mine : other;
        vm = keep ? vm : vo;
    }
}

// In-lane bitonic half-cleaners: distances IPT/2 .. 1.
template <int IPT, bool HV, typename K>
__device__ __forceinline__ void lane_cleaners(K (&r)[IPT], uint64_t (&v)[IPT]) {
#pragma unroll
    for (int d = IPT / 2; d >= 1; d >>= 1)
#pragma unroll
        for (int i = 0; i < IPT; i++)
            if (!(i & d)) cex<HV>(r[i], r[i + d], v[i], v[i + d]);
}

// Cross-lane half-cleaner at lane distance D, then the smaller distances.
template <int D, int IPT, bool HV, typename K>
__device__ __forceinline__ void cross_cleaners(K (&r)[IPT], uint64_t (&v)[IPT], int lane) {
    if constexpr (D >= 1) {
        const bool up = (lane & D) != 0;
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            K o = xshfl<D>(r[i]);
            uint64_t vo = 0;
            if constexpr (HV) vo = xshfl<D>(v[i]);
            lane_pair_keep<HV>(r[i], v[i], o, vo, up);
        }
        cross_cleaners<D / 2, IPT, HV>(r, v, lane);
    }
}

// Merges sorted runs held by lane groups of G/2 lanes (lane-major: lane L holds
// run positions L*IPT .. L*IPT+IPT-1) into runs held by G lanes.  Bitonic merge
// in the "flip" form: first stage pairs item i of lane L with item IPT-1-i of
// lane L^(G-1), then half-cleaners across lanes, then within each lane.  This is
// the role of the reference's mcmerge / mrmerge matrix merges (sorter.h:1628-2319).
template <int G, int IPT, bool HV, typename K>
__device__ __forceinline__ void wave_merge_level(K (&r)[IPT], uint64_t (&v)[IPT], int lane) {
    const bool upper = (lane & (G / 2)) != 0;
    // each exchange reads the partner's item from a copy of the registers and
    // is issued next to the min / max that use it, so the compiler can fold
    // the DPP move into them
    K q[IPT];
    uint64_t qv[IPT];
#pragma unroll
    for (int i = 0; i < IPT; i++) {
        q[i] = r[i];
        qv[i] = HV ? v[i] : 0;
    }
#pragma unroll
    for (int i = 0; i < IPT; i++) {
        const K pk = xshfl<G - 1>(q[IPT - 1 - i]);
        const uint64_t pv = HV ? xshfl<G - 1>(qv[IPT - 1 - i]) : 0;
        lane_pair_keep<HV>(r[i], v[i], pk, pv, upper);
    }
    cross_cleaners<G / 4, IPT, HV>(r, v, lane);
    lane_cleaners<IPT, HV>(r, v);
}

template <int LEVELS, int IPT, bool HV, typename K>
__device__ __forceinline__ void wave_sort_levels(K (&r)[IPT], uint64_t (&v)[IPT], int lane) {
    if constexpr (LEVELS >= 1) wave_merge_level<2, IPT, HV>(r, v, lane);
    if constexpr (LEVELS >= 2) wave_merge_level<4, IPT, HV>(r, v, lane);
    if constexpr (LEVELS >= 3) wave_merge_level<8, IPT, HV>(r, v, lane);
    if constexpr (LEVELS >= 4) wave_merge_level<16, IPT, HV>(r, v, lane);
    if constexpr (LEVELS >= 5) wave_merge_level<32, IPT, HV>(r, v, lane);
    if constexpr (LEVELS >= 6) wave_merge_level<64, IPT, HV>(r, v, lane);
}

// ----------------------------------------------------------------------------
// Merge path in LDS (the co-rank form of mergebl2: merger.h:964-1169).
// A = sk[a_off .. a_off+la), B = sk[b_off .. b_off+lb).  The co-rank of a
// diagonal is the number of A items among its first `diag` outputs; ties go
// to A (stable).  It is found by binary lifting: a fixed, wave-uniform number
// of steps (one per bit of `top`, a power of two with 2*top > min(la, lb)) and
// no data-dependent branch, so every lane issues the same instructions.  Step s
// tries to advance the answer by s (clamped to the upper bound) and keeps the
// move when the last A item it would take still precedes the B item at the
// same diagonal.  Reads A[-1] at most (guard space before the buffer) and
// B[lb] at most (guard space after it).
// ----------------------------------------------------------------------------
template <typename K>
__device__ __forceinline__ int lds_corank_lift(const K *sk, int a_off, int la, int b_off, int lb, int diag, int top) {
    const int lo = diag > lb ? diag - lb : 0;
    const int hi = diag < la ? diag : la;
    // in bytes: A[t-1] at pa + t*W, B[diag-t] at pb - t*W
    constexpr int W = (int)sizeof(K);
    const int pa = (a_off - 1) * W, pb = (b_off + diag) * W, hib = hi * W;
    int ab = lo * W;
    for (int s = top * W; s >= W; s >>= 1) {
        const int tb = min(ab + s, hib);
        const K ka = lds_at_bytes(sk, pa + tb);
        const K kb = lds_at_bytes(sk, pb - tb);
        ab = kb < ka ? ab : tb;
    }
    return (int)((uint32_t)ab / (uint32_t)W);  // ab >= 0
}

// Items of LDS guard space around a merge buffer: one 128-byte row before it
// (the co-rank reads A[-1], which the swizzle maps up to one row back) and
// IPT + 32 items after it (merges read up to IPT-1 items past a run, and the
// swizzle moves an index by at most 31 within its row).
template <typename K> constexpr int lds_front() { return 128 / (int)sizeof(K); }
template <typename K, int IPT> constexpr int lds_back() { return IPT + 32; }
template <typename K, int IPT> constexpr int lds_extra() { return lds_front<K>() + lds_back<K, IPT>(); }

// Serial branch-free merge of outputs into registers, starting at A index ai
// and B index bi: all IPT outputs when FULL, else r[i0 .. i0+cnt) (i0 + cnt <=
// IPT).  Reads one LDS item per output, with no clamp: reading one item past
// either run stays inside the guard space after the buffer.  Key-only runs
// read a side that is exhausted as the maximum key, so one compare picks the
// output (a maximum-key tie then takes the exhausted side, which emits the same
// bits); KV runs test exhaustion explicitly, so a pad never displaces a value.
template <int IPT, bool HV, bool FULL, typename K>
__device__ __forceinline__ void lds_merge_serial(const K *sk, const uint64_t *sv, int a_off, int la, int b_off,
                                                 int lb, int ai, int bi, int i0, int cnt, K (&r)[IPT],
                                                 uint64_t (&v)[IPT]) {
    int ia = a_off + ai, ib = b_off + bi;
    const int ea = a_off + la, eb = b_off + lb;
    if constexpr (!HV) {
        constexpr K KMAX = KeyMax<K>::v;
        K ka = sk[lsw<K>(ia)], kb = sk[lsw<K>(ib)];
        ka = ia < ea ? ka : KMAX;
        kb = ib < eb ? kb : KMAX;
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            if (FULL || (i >= i0 && i < i0 + cnt)) {
                const bool takeA = !(kb < ka);
                r[i] = takeA ? ka : kb;
                ia += takeA;
                ib += !takeA;
                const int nx = takeA ? ia : ib;
                const K nk0 = sk[lsw<K>(nx)];
                const K nk = nx < (takeA ? ea : eb) ? nk0 : KMAX;
                ka = takeA ? nk : ka;
                kb = takeA ? kb : nk;
                v[i] = 0;
            }
        }
    } else {
        // the merge carries each output's LDS position; values are gathered
        // once afterwards, off the dependency chain of the key merge
        if constexpr (FULL) {
            // pointers and positions as byte offsets (keys and values are
            // both 8 bytes, so one offset addresses both arrays)
            constexpr int W = (int)sizeof(K);
            int pa = ia * W, pb = ib * W;
            const int eab = ea * W, ebb = eb * W;
            K ka = lds_at_bytes(sk, pa), kb = lds_at_bytes(sk, pb);
            int pos[IPT];
#pragma unroll
            for (int i = 0; i < IPT; i++) {
                const bool takeA = pb >= ebb || (pa < eab && !(kb < ka));
                r[i] = takeA ? ka : kb;
                pos[i] = takeA ? pa : pb;
                pa += takeA ? W : 0;
                pb += takeA ? 0 : W;
                const K nk = lds_at_bytes(sk, takeA ? pa : pb);
                ka = takeA ? nk : ka;
                kb = takeA ? kb : nk;
            }
#pragma unroll
            for (int i = 0; i < IPT; i++) v[i] = lds_at_bytes(sv, pos[i]);
            return;
        }
        K ka = sk[lsw<K>(ia)], kb = sk[lsw<K>(ib)];
        int pos[IPT];
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            if (FULL || (i >= i0 && i < i0 + cnt)) {
                const bool takeA = ib >= eb || (ia < ea && !(kb < ka));
                r[i] = takeA ? ka : kb;
                pos[i] = takeA ? ia : ib;
                ia += takeA;
                ib += !takeA;
                const int nx = takeA ? ia : ib;
                const K nk = sk[lsw<K>(nx)];
                ka = takeA ? nk : ka;
                kb = takeA ? kb : nk;
            }
        }
#pragma unroll
        for (int i = 0; i < IPT; i++)
            if (FULL || (i >= i0 && i < i0 + cnt)) v[i] = sv[lsw<uint64_t>(pos[i])];
    }
}

// ----------------------------------------------------------------------------
// Key-only merges over a bitonic LDS layout: run A ascending at
// [a_off, a_off + la), run B descending (B[j] at LDS index b_top - j).  A
// pointer that runs off the end of its run lands on the other run's largest
// items: once A is spent its pointer reads B from the top, and such an item
// wins a compare only against an equal remaining B item (the same bits are
// emitted, and the two pointers never meet inside one slice); once B is spent
// its pointer reads A's largest item, which never wins.  So neither the co-rank
// nor the serial merge tests for exhaustion, which saves three instructions per
// output over lds_merge_serial.  Values would be duplicated by such a take, so
// KV runs keep the ascending layout.
// ----------------------------------------------------------------------------
template <typename K>
__device__ __forceinline__ int lds_corank_lift_rev(const K *sk, int a_off, int la, int b_top, int lb, int diag,
                                                   int top) {
    const int lo = diag > lb ? diag - lb : 0;
    const int hi = diag < la ? diag : la;
    constexpr int W = (int)sizeof(K);
    // in bytes: A[t-1] at pa + t*W, B[diag-t] (stored descending) at pb + t*W;
    // the answer and the step are kept in bytes too, so each step is one
    // min, two adds, two swizzles, a compare and a select
    const int pa = (a_off - 1) * W, pb = (b_top - diag) * W, hib = hi * W;
    int ab = lo * W;
    for (int s = top * W; s >= W; s >>= 1) {
        const int tb = min(ab + s, hib);
        const K ka = lds_at_bytes(sk, pa + tb);
        const K kb = lds_at_bytes(sk, pb + tb);
        ab = kb < ka ? ab : tb;
    }
    return (int)((uint32_t)ab / (uint32_t)W);  // ab >= 0
}

// Serial merge on the bitonic layout from A's LDS index ia (moving up) and B's
// LDS index ib (moving down): all IPT outputs when FULL, else r[i0 .. i0+cnt).
template <int IPT, bool FULL, typename K>
__device__ __forceinline__ void lds_merge_rev(const K *sk, int ia, int ib, int i0, int cnt, K (&r)[IPT]) {
    K ka = sk[lsw<K>(ia)], kb = sk[lsw<K>(ib)];
    if constexpr (FULL) {
        // Key-only, so equal keys are identical bits and ties may go either
        // way: the registers hold the head that lost the last compare (L) and
        // the item read last (N, from run A when nA), the output is min(N, L),
        // the loser max(N, L), and the winner's run is a lane mask (scalar
        // instructions).  The pointers are LDS item indices: the winner's
        // pointer moves by the mask itself (one add with carry-in each:
        // ia + fromA, ib - 1 + fromA), and the next read's byte offset is a
        // shift and a v_bitop3 (lds_off): eight VALU instructions per output.
        K N = ka, L = kb;
        bool nA = true;
        const int dlt = ib - ia;
#pragma unroll
        for (int i = 0; i < IPT; i++) {
            const bool lN = L < N;  // L wins (ties: N)
            const K lo = lN ? L : N;
            L = lN ? N : L;
            r[i] = lo;
            const bool fromA = lN != nA;  // the winner's run
            ia += (int)fromA;
            asm("" : "+v"(ia));
            int nb = ia + dlt - (i + 1);
            asm("" : "+v"(nb));
            N = lds_at(sk, fromA ? ia : nb);
            nA = fromA;
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < IPT; i++) {
        if (FULL || (i >= i0 && i < i0 + cnt)) {
            const bool takeA = !(kb < ka);
            r[i] = takeA ? ka : kb;
            ia += takeA ? 1 : 0;
            ib -= takeA ? 0 : 1;
            const K nk = sk[lsw<K>(takeA ? ia : ib)];
            ka = takeA ? nk : ka;
            kb = takeA ? kb : nk;
        }
    }
}

// Outputs [d, d + c) of merge(A, B) on the bitonic layout.
template <int IPT, typename K>
__device__ __forceinline__ void lds_merge_slice_rev(const K *sk, int a_off, int la, int b_top, int lb, int d, int c,
                                                    int top, K (&r)[IPT]) {
    const int a = lds_corank_lift_rev<K>(sk, a_off, la, b_top, lb, d, top);
    if (c == IPT) lds_merge_rev<IPT, true>(sk, a_off + a, b_top - (d - a), 0, IPT, r);
    else lds_merge_rev<IPT, false>(sk, a_off + a, b_top - (d - a), 0, c, r);
}

// Co-rank on the bitonic layout with block splitters: runs whose lengths and
// offsets are multiples of IPT, a diagonal d that is one too, and for every
// IPT-item block of a run (block k of the run at LDS offset o: splitter pair
// o / IPT + k, in the run's ascending order whatever its storage direction)
// its first item at spFL[2 * pair] and its last at spFL[2 * pair + 1] (one
// 2-item store per block).  The lifting steps of IPT items and more compare
// splitters -- A[t-1] is the last item of A's block t/IPT - 1, B[d-t] the first
// of B's block (d-t)/IPT -- on a small unswizzled array, six instructions per
// step instead of ten; the last log2(IPT) steps run on the items themselves.
// The same answer as lds_corank_lift_rev for any top with 2 * top > min(la, lb).
template <int IPT, typename K>
__device__ __forceinline__ int lds_corank_split_rev(const K *sk, const K *spFL, int a_off, int la, int b_off, int lb,
                                                    int b_top, int diag, int top) {
    static_assert((IPT & (IPT - 1)) == 0 && IPT >= (int)sizeof(K), "IPT: a power of two, at least the key bytes");
    constexpr uint32_t W = (uint32_t)sizeof(K);
    // every length and offset here is a non-negative multiple of IPT: an
    // item count x is x * 2W / IPT bytes of splitter pairs
    constexpr bool DN = IPT >= 2 * (int)W;
    constexpr uint32_t SH = DN ? (uint32_t)__builtin_ctz((unsigned)IPT / (2 * W))
                               : (uint32_t)__builtin_ctz(2 * W / (unsigned)IPT);
    const auto cv = [](uint32_t x) -> uint32_t { return DN ? x >> SH : x << SH; };
    const uint32_t lo = (uint32_t)(diag > lb ? diag - lb : 0);
    const uint32_t hi = (uint32_t)(diag < la ? diag : la);
    // block-granular steps: A[t-1] at pA + cv(t), B[diag-t] at pB - cv(t)
    const char *pA = (const char *)spFL + cv((uint32_t)a_off) - W;
    const char *pB = (const char *)spFL + cv((uint32_t)(b_off + diag));
    const uint32_t hiq = cv(hi);
    uint32_t abq = cv(lo);
    for (uint32_t s = cv((uint32_t)top); s >= 2 * W; s >>= 1) {
        const uint32_t tq = min(abq + s, hiq);
        const K ka = *(const K *)(pA + tq);
        const K kb = *(const K *)(pB - tq);
        abq = kb < ka ? abq : tq;
    }
    // the last steps on the items (in bytes), from the block-granular answer;
    // steps of WS = min(IPT, 16) items or more on any offsets
    constexpr int WS = IPT < 16 ? IPT : 16;
    const int pa = (a_off - 1) * (int)W, pb = (b_top - diag) * (int)W, hib = (int)(hi * W);
    int ab = (int)((DN ? abq << SH : abq >> SH) * W);
#pragma unroll
    for (int s = IPT / 2 * (int)W; s >= WS * (int)W; s >>= 1) {
        const int tb = min(ab + s, hib);
        const K ka = lds_at_bytes(sk, pa + tb);
        const K kb = lds_at_bytes(sk, pb + tb);
        ab = kb < ka ? ab : tb;
    }
    // Steps below WS items stay inside two WS-aligned item windows: with a0 =
    // the answer so far (a multiple of WS), the probe a0 + v + 1 compares A
    // item xa + v with B item yb + v (xa = a_off + a0, yb = b_top - diag + a0
    // + 1, both multiples of WS, as every run offset, length and diagonal here
    // is a multiple of IPT); for x a multiple of WS and v < WS <= 16, x + v =
    // x ^ v and the swizzle is XOR-linear with lds_off(v) = v * W, so item
    // x + v sits at lds_off(x) ^ (v * W): each step is an add, a min, two
    // v_bitop3 (offset XOR (v*W masked to the window)), a compare and a
    // select.  v is kept as v * W from -W (no A item past a0 taken); a probe
    // clamped to -W reads inside the window (the mask) and cannot move.
    const int a0 = (int)((uint32_t)ab / W);
    const int swa = lds_off<K>(a_off + a0), swb = lds_off<K>(b_top - diag + a0 + 1);
    const int hv = (hi - (int)a0 - 1) * (int)W;  // largest probe v * W
    constexpr int MW = (WS - 1) * (int)W;
    const char *sb = (const char *)sk;
    int vb = -(int)W;
#pragma unroll
    for (int s = WS / 2 * (int)W; s >= (int)W; s >>= 1) {
        const int tb = min(vb + s, hv);
        const K ka = *(const K *)(sb + (swa ^ (tb & MW)));
        const K kb = *(const K *)(sb + (swb ^ (tb & MW)));
        vb = kb < ka ? vb : tb;
    }
    // vb is a multiple of W, at least -W: the arithmetic shift divides exactly
    return a0 + (vb >> __builtin_ctz(W)) + 1;
}

// A block's splitter pair (its first and last key) as one 8- or 16-byte store.
// Exchanges two registers in place with v_swap_b32 (one per 32-bit half).
template <typename K> __device__ __forceinline__ void swap_regs(K &a, K &b) {
    static_assert(sizeof(K) == 4 || sizeof(K) == 8, "32- or 64-bit keys");
    if constexpr (sizeof(K) == 4) {
        uint32_t x = (uint32_t)a, y = (uint32_t)b;
        asm("v_swap_b32 %0, %1" : "+v"(x), "+v"(y));
        a = (K)x;
        b = (K)y;
    } else {
        uint32_t xl = (uint32_t)a, xh = (uint32_t)((uint64_t)a >> 32);
        uint32_t yl = (uint32_t)b, yh = (uint32_t)((uint64_t)b >> 32);
        asm("v_swap_b32 %0, %1" : "+v"(xl), "+v"(yl));
        asm("v_swap_b32 %0, %1" : "+v"(xh), "+v"(yh));
        a = (K)(((uint64_t)xh << 32) | xl);
        b = (K)(((uint64_t)yh << 32) | yl);
    }
}

template <typename K> __device__ __forceinline__ void put_pair(K *p, K first, K last) {
    if constexpr (sizeof(K) == 4) *(uint2 *)p = make_uint2(first, last);
    else *(ulonglong2 *)p = make_ulonglong2((unsigned long long)first, (unsigned long long)last);
}

// Largest power of two <= x (0 for 0): the co-rank step bound for runs whose
// shorter side has x items.
__device__ __forceinline__ int pow2_floor(int x) { return x > 0 ? 1 << (31 - __clz(x)) : 0; }

// Merges outputs [d, d + c) of merge(A, B) into registers: binary-lifting
// co-rank, then a serial merge (full-slice fast path when c == IPT).
template <int IPT, bool HV, typename K>
__device__ __forceinline__ void lds_merge_slice(const K *sk, const uint64_t *sv, int a_off, int la, int b_off,
                                                int lb, int d, int c, int top, K (&r)[IPT], uint64_t (&v)[IPT]) {
    const int a = lds_corank_lift<K>(sk, a_off, la, b_off, lb, d, top);
    if (c == IPT) lds_merge_serial<IPT, HV, true>(sk, sv, a_off, la, b_off, lb, a, d - a, 0, IPT, r, v);
    else lds_merge_serial<IPT, HV, false>(sk, sv, a_off, la, b_off, lb, a, d - a, 0, c, r, v);
}

// Store IPT consecutive registers to LDS at item index base (a multiple of
// IPT, a power of two) as 16-byte stores: chunk c of the block is at
// lds_off(base) ^ lds_off(c), the swizzle being XOR-linear and base + c = base
// ^ c, so one address per thread and a constant XOR per chunk.
template <int IPT, bool HV, typename K>
__device__ __forceinline__ void lds_put(K *sk, uint64_t *sv, int base, const K (&r)[IPT], const uint64_t (&v)[IPT]) {
    constexpr int E = chunk_items<K>();
    static_assert(IPT % E == 0 && (IPT & (IPT - 1)) == 0, "IPT: a power of two of whole 16-byte chunks");
    const int sw = lds_off<K>(base);
#pragma unroll
    for (int c = 0; c < IPT; c += E) {
        char *p = (char *)sk + (sw ^ lds_off<K>(c));
        if constexpr (E == 4) {
            uint4 q = {(uint32_t)r[c], (uint32_t)r[c + 1], (uint32_t)r[c + 2], (uint32_t)r[c + 3]};
            *(uint4 *)p = q;
        } else {
            ulonglong2 q;
            q.x = (unsigned long long)r[c];
            q.y = (unsigned long long)r[c + 1];
            *(ulonglong2 *)p = q;
        }
    }
    if constexpr (HV) {
        const int swv = lds_off<uint64_t>(base);
#pragma unroll
        for (int c = 0; c < IPT; c += 2) {
            ulonglong2 q;
            q.x = v[c];
            q.y = v[c + 1];
            *(ulonglong2 *)((char *)sv + (swv ^ lds_off<uint64_t>(c))) = q;
        }
    }
}

// Stores IPT registers, items [o, o + IPT) of a run that the next merge reads
// descending, at LDS indices top, top - 1, ..., top - IPT + 1 (top + 1 a
// multiple of IPT, which every caller's block alignment gives) as 16-byte
// stores: the block starts at base = top + 1 - IPT, and the chunk of LDS items
// base + x .. base + x + E - 1 (x = IPT - E - c) is at lds_off(base) ^
// lds_off(x) (XOR-linear swizzle, base + x = base ^ x).
template <int IPT, typename K>
__device__ __forceinline__ void lds_put_desc(K *sk, int top, const K (&r)[IPT]) {
    constexpr int E = chunk_items<K>();
    static_assert(IPT % E == 0 && (IPT & (IPT - 1)) == 0, "IPT: a power of two of whole 16-byte chunks");
    const int sw = lds_off<K>(top + 1 - IPT);
#pragma unroll
    for (int c = 0; c < IPT; c += E) {
        char *p = (char *)sk + (sw ^ lds_off<K>(IPT - E - c));
        if constexpr (E == 4) {
            uint4 q = {(uint32_t)r[c + 3], (uint32_t)r[c + 2], (uint32_t)r[c + 1], (uint32_t)r[c]};
            *(uint4 *)p = q;
        } else {
            ulonglong2 q;
            q.x = (unsigned long long)r[c + 1];
            q.y = (unsigned long long)r[c];
            *(ulonglong2 *)p = q;
        }
    }
}

// Writes items [0, cnt) of an LDS tile to out[dst ..] with lane-consecutive,
// fully coalesced stores (16-byte stores when the destination is aligned),
// applying the output codec.  KV tiles go to {key, value} records, or, with
// vout set, keys to out[] and values to vout[] (structure of arrays).
template <int BLOCK, int TILE, bool HV, typename K>
__device__ __forceinline__ void lds_store_tile(const K *sk, const uint64_t *sv, void *out, uint64_t dst, int cnt,
                                               int cout, void *vout = nullptr) {
    const int tid = threadIdx.x;
    if constexpr (HV) {
        if (vout) {
            K *ko = (K *)out + dst;
            uint64_t *vo = (uint64_t *)vout + dst;
#pragma unroll 4
            for (int e = tid; e < cnt; e += BLOCK) {
                ko[e] = dec_key(sk[lsw<K>(e)], cout);
                vo[e] = sv[lsw<uint64_t>(e)];
            }
            return;
        }
        ulonglong2 *o = (ulonglong2 *)out + dst;
#pragma unroll 4
        for (int e = tid; e < cnt; e += BLOCK) {
            ulonglong2 q;
            q.x = (unsigned long long)dec_key(sk[lsw<K>(e)], cout);
            q.y = sv[lsw<uint64_t>(e)];
            o[e] = q;
        }
    } else {
        constexpr int E = chunk_items<K>();
        K *o = (K *)out + dst;
        const bool al16 = ((uintptr_t)o & 15) == 0;
        if (al16 && cnt == TILE) {
            // uniform codec: no per-key branch; stores off one base pointer
            const CodecMask<K> m = codec_mask<K>(cout);
#pragma unroll
            for (int q = tid; q < TILE / E; q += BLOCK) {
                const uint32_t e = (uint32_t)(q * E);
                // non-temporal: a pass's output is next read after far more
                // data than the caches hold
                typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
                if constexpr (E == 4) {
                    u32x4 x = *(const u32x4 *)&sk[lsw<K>(e)];
                    if (cout) {
                        x.x = dec_bf<uint32_t>(x.x, m);
                        x.y = dec_bf<uint32_t>(x.y, m);
                        x.z = dec_bf<uint32_t>(x.z, m);
                        x.w = dec_bf<uint32_t>(x.w, m);
                    }
                    __builtin_nontemporal_store(x, (u32x4 *)&o[e]);
                } else {
                    u64x2 x = *(const u64x2 *)&sk[lsw<K>(e)];
                    if (cout) {
                        x.x = dec_bf<uint64_t>(x.x, m);
                        x.y = dec_bf<uint64_t>(x.y, m);
                    }
                    __builtin_nontemporal_store(x, (u64x2 *)&o[e]);
                }
            }
        } else {
#pragma unroll 4
            for (int e = tid; e < cnt; e += BLOCK) o[e] = dec_key(sk[lsw<K>(e)], cout);
        }
    }
}

// ----------------------------------------------------------------------------
// Tagged samples for the sample-split K-way pass.  Items [0, cnt) of an LDS
// tile are output items g0 .. g0+cnt-1; every item whose output index is a
// multiple of 2^slog2 is written to smp[index >> slog2] with its run r within
// the next pass's group of kw runs of 2^rlog2 items and its index m in that
// run over 2^slog2: 32-bit keys as u64 key << 32 | r << 27 | m (r < 32, m <
// 2^27; u64 order is then the total order (key, run, index)), 64-bit keys as
// {key, r << 32 | m}.
// ----------------------------------------------------------------------------
template <int BLOCK, typename K>
__device__ __forceinline__ void write_samples(const K *sk, int cnt, uint64_t g0, void *smp, int rlog2, int kw,
                                              int slog2) {
    const int s = 1 << slog2;
    const int first = (int)((0 - g0) & (uint64_t)(s - 1));
    for (int e = first + (int)threadIdx.x * s; e < cnt; e += BLOCK * s) {
        const uint64_t gi = g0 + (uint64_t)e;
        const uint64_t run = gi >> rlog2;
        const uint64_t r = run & (uint64_t)(kw - 1);
        const uint64_t m = (gi - (run << rlog2)) >> slog2;
        const K key = sk[lsw<K>(e)];
        if constexpr (sizeof(K) == 4)
            ((uint64_t *)smp)[gi >> slog2] = ((uint64_t)key << 32) | (r << 27) | m;
        else
            ((ulonglong2 *)smp)[gi >> slog2] = make_ulonglong2((unsigned long long)key, (r << 32) | m);
    }
}

// Reading samples back: the key, and the key with its run and index.
template <typename K> struct Sample;
template <> struct Sample<uint32_t> {
    static __device__ __forceinline__ uint32_t key(const void *p, uint64_t i) {
        return (uint32_t)(((const uint64_t *)p)[i] >> 32);
    }
    static __device__ __forceinline__ void get(const void *p, uint64_t i, uint32_t &k, int &r, uint64_t &m) {
        const uint64_t x = ((const uint64_t *)p)[i];
        k = (uint32_t)(x >> 32);
        r = (int)((x >> 27) & 31);
        m = x & ((1ull << 27) - 1);
    }
};
template <> struct Sample<uint64_t> {
    static __device__ __forceinline__ uint64_t key(const void *p, uint64_t i) {
        return (uint64_t)((const ulonglong2 *)p)[i].x;
    }
    static __device__ __forceinline__ void get(const void *p, uint64_t i, uint64_t &k, int &r, uint64_t &m) {
        const ulonglong2 x = ((const ulonglong2 *)p)[i];
        k = (uint64_t)x.x;
        r = (int)(x.y >> 32);
        m = x.y & 0xFFFFFFFFull;
    }
};

// ----------------------------------------------------------------------------
// Wave-cooperative co-rank in global memory (merge path, Partition.h:117-166
// specialised to two runs): returns the number of A items among the first
// `diag` outputs of merge(A, B), ties to A.  Each step evaluates 64 evenly
// spaced candidates at once and keeps the bracket the ballot points to, so a
// run of 2^27 items resolves in five dependent steps.  Must be called by all
// 64 lanes of a wave with uniform arguments.
// ----------------------------------------------------------------------------
template <typename K, bool HV>
__device__ __forceinline__ uint64_t wave_corank(const void *in, uint64_t a_base, uint64_t la, uint64_t b_base,
                                                uint64_t lb, uint64_t diag, int lane) {
    uint64_t lo = diag > lb ? diag - lb : 0;
    uint64_t hi = diag < la ? diag : la;
    while (hi > lo) {
        const uint64_t span = hi - lo;
        if (span <= 64) {
            bool p = false;
            if ((uint64_t)lane < span) {
                const uint64_t m = lo + lane;
                const K a = GMem<K, HV>::load_key(in, a_base + m);
                const K b = GMem<K, HV>::load_key(in, b_base + diag - 1 - m);
                p = b < a;
            }
            const uint64_t mask = __ballot(p);
            return mask ? lo + (uint64_t)(__ffsll((unsigned long long)mask) - 1) : hi;
        }
        const uint64_t m = lo + (span * (uint64_t)(lane + 1)) / 65;
        const K a = GMem<K, HV>::load_key(in, a_base + m);
        const K b = GMem<K, HV>::load_key(in, b_base + diag - 1 - m);
        const uint64_t mask = __ballot(b < a);
        if (!mask) {
            lo = lo + (span * 64) / 65 + 1;
        } else {
            const int f = __ffsll((unsigned long long)mask) - 1;
            const uint64_t mf = lo + (span * (uint64_t)(f + 1)) / 65;
            const uint64_t nlo = f ? lo + (span * (uint64_t)f) / 65 + 1 : lo;
            hi = mf;
            lo = nlo;
        }
    }
    return lo;
}

}  // namespace ogs
