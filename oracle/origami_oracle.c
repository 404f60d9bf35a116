/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of Origami's CPU mergesort (arif-arman/origami-sort,
 * directory Origami/).  It is the checker for the GPU library and the CPU
 * baseline timed beside it; it is never linked into, called by or shipped with
 * the product library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.
 *
 * Every function names the reference lines it restates.  Keys are handled in
 * an "ordered" unsigned representation (u32 or u64) whose unsigned order is
 * the order the reference (or the ABI, for types the reference lacks) defines:
 *   u32  identity                      (avx2_utils.h:129-133 unsigned min/max)
 *   i32  x ^ 0x80000000                (signed order; ABI extension)
 *   f32  IEEE totalOrder bijection     (ABI extension: the reference has no float path)
 *   u64  identity                      (ABI extension: the reference has only signed i64)
 *   i64  x ^ 2^63                      (avx2_utils.h:140-145 signed cmpgt)
 *   KV   16-byte {key, value} AoS, compared on the key only
 *        (commons.h:51-72, avx2_utils.h:151-157); key signed for KV_I64I64.
 *
 * Parity pins (tests/test_oracle.py): std-library sort of the same bytes
 * (numpy), 0-1-principle exhaustive checks of every sorting network, brute
 * force co-rank on tiny inputs, and the C++ standard's known-answer values for
 * mt19937 / mt19937_64.  Deviations from the reference, all documented at the
 * function: arbitrary n (the reference needs powers of two, sorter.h:3857-3868),
 * and Partition_Ptrs terminates on lmax <= rmin with -inf/+inf sentinels
 * (the reference's strict '<' and zero sentinel, Partition.h:120-148, loop or
 * mis-split on duplicate and negative keys).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

enum { ORC_U32 = 0, ORC_I32 = 1, ORC_U64 = 2, ORC_I64 = 3, ORC_F32 = 4,
       ORC_KV_U64U64 = 5, ORC_KV_I64I64 = 6, ORC_KV_U32U32 = 7 };

typedef struct { uint64_t k, v; } kv_t;
/* the reference's 8-byte KeyValue<ui, ui> (Origami/commons.h:51-71): packed
 * {u32 key, u32 value}, compared on the key only like every KeyValue */
typedef struct { uint32_t k, v; } kv32_t;

/* ------------------------------------------------------------------------ */
/* Seeded generators.  datagen::Writer (writer.h:47-242) draws with a        */
/* default-seeded mt19937 + uniform_int_distribution, whose output is        */
/* implementation-defined; both sides here use the raw engine output, which  */
/* the C++ standard fixes ([rand.predef]).                                   */
/* ------------------------------------------------------------------------ */
void orc_mt19937(uint32_t seed, uint32_t *out, uint64_t n)
{
    uint32_t mt[624];
    mt[0] = seed;
    for (int i = 1; i < 624; i++)
        mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    int idx = 624;
    for (uint64_t j = 0; j < n; j++) {
        if (idx >= 624) {
            for (int i = 0; i < 624; i++) {
                uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            idx = 0;
        }
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        out[j] = y;
    }
}

void orc_mt19937_64(uint64_t seed, uint64_t *out, uint64_t n)
{
    uint64_t mt[312];
    mt[0] = seed;
    for (int i = 1; i < 312; i++)
        mt[i] = 6364136223846793005ull * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
    int idx = 312;
    for (uint64_t j = 0; j < n; j++) {
        if (idx >= 312) {
            for (int i = 0; i < 312; i++) {
                uint64_t y = (mt[i] & 0xFFFFFFFF80000000ull) | (mt[(i + 1) % 312] & 0x7FFFFFFFull);
                mt[i] = mt[(i + 156) % 312] ^ (y >> 1) ^ ((y & 1u) ? 0xB5026F5AA96619E9ull : 0ull);
            }
            idx = 0;
        }
        uint64_t y = mt[idx++];
        y ^= (y >> 29) & 0x5555555555555555ull;
        y ^= (y << 17) & 0x71D67FFFEDA60000ull;
        y ^= (y << 37) & 0xFFF7EEE000000000ull;
        y ^= y >> 43;
        out[j] = y;
    }
}

/* ------------------------------------------------------------------------ */
/* Key codecs (ordered representation).                                     */
/* ------------------------------------------------------------------------ */
static inline uint32_t f32_enc(uint32_t b) { return b ^ ((b >> 31) ? 0xFFFFFFFFu : 0x80000000u); }
static inline uint32_t f32_dec(uint32_t u) { return u ^ ((u >> 31) ? 0x80000000u : 0xFFFFFFFFu); }

static void encode(int kt, void *d, uint64_t n, int dir)
{
    uint64_t i;
    switch (kt) {
    case ORC_I32: { uint32_t *p = (uint32_t *)d; for (i = 0; i < n; i++) p[i] ^= 0x80000000u; } break;
    case ORC_F32: { uint32_t *p = (uint32_t *)d;
                    for (i = 0; i < n; i++) p[i] = dir > 0 ? f32_enc(p[i]) : f32_dec(p[i]); } break;
    case ORC_I64: { uint64_t *p = (uint64_t *)d; for (i = 0; i < n; i++) p[i] ^= 1ull << 63; } break;
    case ORC_KV_I64I64: { kv_t *p = (kv_t *)d; for (i = 0; i < n; i++) p[i].k ^= 1ull << 63; } break;
    default: break;
    }
}

/* ------------------------------------------------------------------------ */
/* Sorting networks restated from sorter.h.                                 */
/* ------------------------------------------------------------------------ */
/* in_register_sort8, sorter.h:40-48 (19 compare-exchanges). */
static const uint8_t NET8[19][2] = {
    {0,4},{1,5},{2,6},{3,7}, {0,2},{1,3},{4,6},{5,7}, {2,4},{3,5},{0,1},{6,7},
    {2,3},{4,5}, {1,4},{3,6}, {1,2},{3,4},{5,6}};
/* in_register_sort16 / in_register_sort16_kv_scalar, sorter.h:60-66 (60 CEs). */
static const uint8_t NET16[60][2] = {
    {0,1},{2,3},{4,5},{6,7},{8,9},{10,11},{12,13},{14,15},
    {0,2},{4,6},{8,10},{12,14},{1,3},{5,7},{9,11},{13,15},
    {0,4},{8,12},{1,5},{9,13},{2,6},{10,14},{3,7},{11,15},
    {0,8},{1,9},{2,10},{3,11},{4,12},{5,13},{6,14},{7,15},
    {5,10},{6,9},{3,12},{13,14},{7,11},{1,2},{4,8},
    {1,4},{7,13},{2,8},{11,14},{2,4},{5,6},{9,10},{11,13},{3,8},{7,12},
    {6,8},{10,12},{3,5},{7,9},{3,4},{5,6},{7,8},{9,10},{11,12},{6,7},{8,9}};
/* Register-merge networks of mergebl2_scalar, merger.h:83-176:
 * nreg=4 loop body (merger.h:122-124) and its tail form (merger.h:187-190),
 * which adds the (5,6) exchange the loop body leaves out. */
static const uint8_t MRG4_LOOP[8][2] = {{0,4},{1,5},{2,6},{3,7},{2,4},{3,5},{1,2},{3,4}};
static const uint8_t MRG4_TAIL[9][2] = {{0,4},{1,5},{2,6},{3,7},{2,4},{3,5},{1,2},{3,4},{5,6}};
/* drain of the upper register block, merger.h:219-222 */
static const uint8_t MRG4_UPPER[5][2] = {{4,6},{5,7},{4,5},{6,7},{5,6}};

/* ------------------------------------------------------------------------ */
/* Type-generic body.  T is the ordered item; LT is the strict order.        */
/* ------------------------------------------------------------------------ */
#define ORC_GEN(T, S, LT)                                                                     \
/* SWAP / SWAPKV2 (commons.h:111-149): x <- smaller, y <- larger; equal keys exchange */      \
static inline void ce_##S(T *x, T *y) {                                                       \
    int less = LT((*x), (*y));                                                                \
    T lo = less ? *x : *y; T hi = less ? *y : *x; *x = lo; *y = hi; }                         \
static void net_##S(T *a, const uint8_t (*net)[2], int m) {                                   \
    for (int i = 0; i < m; i++) ce_##S(&a[net[i][0]], &a[net[i][1]]); }                        \
                                                                                              \
/* phase1_sort, scalar build (sorter.h:3636-3700, _P1_N = 8 or 16, config.h:46-95):         \
 * every full block of p1n items is sorted by the in-register network; a ragged tail         \
 * (absent in the reference) is insertion-sorted. */                                          \
static void phase1_##S(T *d, uint64_t n, int p1n) {                                           \
    uint64_t i = 0;                                                                           \
    for (; i + (uint64_t)p1n <= n; i += p1n) {                                                \
        if (p1n == 8) net_##S(d + i, NET8, 19); else net_##S(d + i, NET16, 60); }             \
    for (uint64_t j = i + 1; j < n; j++) {                                                    \
        T x = d[j]; uint64_t k = j;                                                           \
        while (k > i && LT(x, d[k - 1])) { d[k] = d[k - 1]; k--; }                            \
        d[k] = x; } }                                                                         \
                                                                                              \
/* mergebl2_scalar (merger.h:45-224): branchless 2-way merge that keeps nreg items of each   \
 * stream in registers, exchanges them through a merge network, emits the low nreg and      \
 * reloads from the stream whose next head is smaller ('<=' favours the current stream,      \
 * merger.h:84).  nreg = 1 works for any lengths; nreg = 4 needs lengths that are multiples  \
 * of 4 and is used only then. */                                                             \
static void mergebl_##S(const T *A, uint64_t la, const T *B, uint64_t lb, T *C, int nreg) {   \
    if (la == 0) { memcpy(C, B, lb * sizeof(T)); return; }                                    \
    if (lb == 0) { memcpy(C, A, la * sizeof(T)); return; }                                    \
    if (nreg == 4 && ((la | lb) & 3)) nreg = 1;                                               \
    const T *endA = A + la, *endB = B + lb;                                                   \
    T r[8];                                                                                   \
    for (int i = 0; i < nreg; i++) { r[i] = A[i]; r[nreg + i] = B[i]; }                       \
    const T *loadFrom = A + nreg, *opposite = B + nreg;                                       \
    while (loadFrom != endA && loadFrom != endB) {                                            \
        /* *loadFrom <= *opposite; an exhausted opposite reads as +inf here, where the       \
         * reference reads one item past its end (only reachable when a run is a single     \
         * register block, which the reference's equal power-of-two runs never produce) */  \
        int oppDone = opposite == endA || opposite == endB;                                   \
        int first = oppDone || !LT((*opposite), (*loadFrom));                                 \
        const T *t = first ? loadFrom : opposite;                                             \
        opposite = first ? opposite : loadFrom; loadFrom = t;                                 \
        if (nreg == 1) ce_##S(&r[0], &r[1]); else net_##S(r, MRG4_LOOP, 8);                   \
        for (int i = 0; i < nreg; i++) { C[i] = r[i]; r[i] = loadFrom[i]; }                   \
        C += nreg; loadFrom += nreg; }                                                        \
    if (nreg == 1) ce_##S(&r[0], &r[1]); else net_##S(r, MRG4_TAIL, 9);                       \
    for (int i = 0; i < nreg; i++) C[i] = r[i];                                               \
    C += nreg;                                                                                \
    const T *endOp = (loadFrom == endA) ? endB : endA;                                        \
    while (opposite != endOp) {                                                               \
        for (int i = 0; i < nreg; i++) r[i] = opposite[i];                                    \
        if (nreg == 1) ce_##S(&r[0], &r[1]); else net_##S(r, MRG4_LOOP, 8);                   \
        for (int i = 0; i < nreg; i++) C[i] = r[i];                                           \
        opposite += nreg; C += nreg; }                                                        \
    if (nreg == 4) net_##S(r, MRG4_UPPER, 5);                                                 \
    for (int i = 0; i < nreg; i++) C[i] = r[nreg + i]; }                                      \
                                                                                              \
/* merge / merge_series (merger.h:4079-4228): one pass merges neighbouring runs of length    \
 * `start` (the last pair may be ragged here); passes double `start` until sort_n.           \
 * Returns the buffer holding the result (ping-pong, merger.h:4225). */                      \
static T *merge_series_##S(T *src, T *dst, uint64_t n, uint64_t start, uint64_t sort_n,      \
                           int nreg) {                                                        \
    while (start < sort_n) {                                                                  \
        uint64_t next = start << 1;                                                           \
        for (uint64_t off = 0; off < n; off += next) {                                        \
            uint64_t la = n - off < start ? n - off : start;                                  \
            uint64_t rest = n - off - la;                                                     \
            uint64_t lb = rest < start ? rest : start;                                        \
            mergebl_##S(src + off, la, src + off + la, lb, dst + off, nreg); }                \
        T *t = src; src = dst; dst = t; start = next; }                                       \
    return src; }                                                                             \
                                                                                              \
/* MergeTreeEven/Odd::merge with RefillNode (merge_tree.h:2505-2558, 2792-3010): a k-way     \
 * merge as a tree of buffered 2-way nodes; a node whose buffer runs dry is refilled from    \
 * its children before the parent continues.  Ties go to the right child ('<' in            \
 * merge_leaf_to_internal, merge_tree.h:40).  Node buffers hold `buf` items. */              \
typedef struct { const T *p, *e; T *buf; uint64_t head, cnt; int leaf; } node_##S;           \
static int node_avail_##S(node_##S *nd, int i, uint64_t bufn);                                \
static int node_peek_##S(node_##S *nd, int i, uint64_t bufn, T *out) {                        \
    if (nd[i].leaf) { if (nd[i].p == nd[i].e) return 0; *out = *nd[i].p; return 1; }          \
    if (!node_avail_##S(nd, i, bufn)) return 0;                                               \
    *out = nd[i].buf[nd[i].head]; return 1; }                                                 \
static void node_pop_##S(node_##S *nd, int i) {                                               \
    if (nd[i].leaf) nd[i].p++; else { nd[i].head++; nd[i].cnt--; } }                          \
static int node_avail_##S(node_##S *nd, int i, uint64_t bufn) {                               \
    if (nd[i].cnt) return 1;                                                                  \
    nd[i].head = 0;                                                                           \
    while (nd[i].cnt < bufn) {                                                                \
        T a, b; int ha = node_peek_##S(nd, 2 * i + 1, bufn, &a);                              \
        int hb = node_peek_##S(nd, 2 * i + 2, bufn, &b);                                      \
        if (!ha && !hb) break;                                                                \
        int takeA = ha && (!hb || LT(a, b));                                                  \
        nd[i].buf[nd[i].cnt++] = takeA ? a : b;                                               \
        node_pop_##S(nd, takeA ? 2 * i + 1 : 2 * i + 2); }                                    \
    return nd[i].cnt != 0; }                                                                  \
static void kway_##S(const T *const *runs, const uint64_t *lens, int K, T *out, uint64_t bufn) {\
    if (K == 1) { memcpy(out, runs[0], lens[0] * sizeof(T)); return; }                        \
    int L = 1; while (L < K) L <<= 1;                                                         \
    node_##S *nd = (node_##S *)calloc(2 * L - 1, sizeof(node_##S));                           \
    T *bufs = (T *)malloc((size_t)(L - 1) * bufn * sizeof(T));                                \
    for (int i = 0; i < L - 1; i++) nd[i].buf = bufs + (size_t)i * bufn;                      \
    for (int i = 0; i < L; i++) {                                                             \
        node_##S *lf = &nd[L - 1 + i]; lf->leaf = 1;                                          \
        if (i < K) { lf->p = runs[i]; lf->e = runs[i] + lens[i]; } }                          \
    T x;                                                                                      \
    while (node_peek_##S(nd, 0, bufn, &x)) { *out++ = x; node_pop_##S(nd, 0); }               \
    free(bufs); free(nd); }                                                                   \
                                                                                              \
/* Partition_Ptrs (Partition.h:117-166): K-way co-rank.  Start from any guess S with        \
 * sum(S) = rank inside bounds [L,R]; while the largest item left of the split exceeds the   \
 * smallest item right of it, exclude the offending halves and move both pointers by the     \
 * same delta in opposite directions, keeping sum(S) constant.  Here the empty side is      \
 * treated as -inf/+inf, and equal keys are ordered by run index so the order is total: the  \
 * loop ends on lmax < rmin at the one split of that order, which keeps the splits of        \
 * increasing ranks monotone in every run (with ties left open, two searches can split a     \
 * run of equal keys differently and cut negative-length slices). */                         \
static void partition_##S(const T *const *runs, const uint64_t *lens, int K, uint64_t rank,   \
                          uint64_t *Sx) {                                                     \
    uint64_t *Lb = (uint64_t *)malloc(K * sizeof(uint64_t));                                  \
    uint64_t *Rb = (uint64_t *)malloc(K * sizeof(uint64_t));                                  \
    uint64_t total = 0, left = rank;                                                          \
    for (int k = 0; k < K; k++) { Lb[k] = 0; Rb[k] = lens[k]; total += lens[k]; }             \
    /* initial guess: proportional split, then fill greedily so sum == rank */                \
    for (int k = 0; k < K; k++) {                                                             \
        uint64_t g = total ? (uint64_t)((double)rank * lens[k] / (double)total) : 0;          \
        if (g > lens[k]) { g = lens[k]; } if (g > left) { g = left; } Sx[k] = g; left -= g; }   \
    for (int k = 0; k < K && left; k++) {                                                     \
        uint64_t room = lens[k] - Sx[k]; uint64_t g = room < left ? room : left;              \
        Sx[k] += g; left -= g; }                                                              \
    for (;;) {                                                                                \
        int smax = -1, smin = -1; T lmax, rmin;                                               \
        for (int k = 0; k < K; k++) {                                                         \
            if (Sx[k] > Lb[k]) { T x = runs[k][Sx[k] - 1];                                    \
                if (smax < 0 || LT(lmax, x) || (!LT(x, lmax) && smax < k)) { lmax = x; smax = k; } }\
            if (Sx[k] < Rb[k]) { T x = runs[k][Sx[k]];                                        \
                if (smin < 0 || LT(x, rmin) || (!LT(rmin, x) && k < smin)) { rmin = x; smin = k; } } }\
        if (smax < 0 || smin < 0 || !(LT(rmin, lmax) || (!LT(lmax, rmin) && smin < smax))) break;\
        Rb[smax] = Sx[smax] - 1; Lb[smin] = Sx[smin] + 1;                                     \
        uint64_t dmax = (Rb[smax] - Lb[smax]) / 2, dmin = (Rb[smin] - Lb[smin]) / 2;          \
        uint64_t dl = dmax < dmin ? dmax : dmin;                                              \
        Sx[smax] = Rb[smax] - dl; Sx[smin] = Lb[smin] + dl; }                                 \
    free(Lb); free(Rb); }                                                                     \
                                                                                              \
/* sort_single_thread (sorter.h:3851-3915) / sort_multi_thread (sorter.h:4194-4287):         \
 * phases 1-2 sort independent P2_N chunks (phase2_sort, sorter.h:3831-3849); phase 3 merges \
 * runs min_k at a time (merge_series when min_k == 2, else merge trees); with threads > 1   \
 * every merge group is split by Partition_Ptrs so all threads share it (sorter.h:4086-4191).\
 * Returns d or tmp, whichever holds the result. */                                          \
static T *sort_##S(T *d, T *tmp, uint64_t n, uint64_t sort_n, int p1n, uint64_t p2n,          \
                   int nreg, int threads, int min_k) {                                        \
    if (n == 0) return d;                                                                     \
    if (sort_n == 0 || sort_n > n) sort_n = n;                                                \
    uint64_t Y = p2n < sort_n ? p2n : sort_n;                                                 \
    int64_t nchunks = (int64_t)((n + Y - 1) / Y);                                             \
    T **res = (T **)malloc(nchunks * sizeof(T *));                                            \
    _Pragma("omp parallel for num_threads(threads) schedule(dynamic, 1)")                     \
    for (int64_t c = 0; c < nchunks; c++) {                                                   \
        uint64_t off = (uint64_t)c * Y, len = n - off < Y ? n - off : Y;                      \
        phase1_##S(d + off, len, p1n);                                                        \
        res[c] = merge_series_##S(d + off, tmp + off, len, (uint64_t)p1n, Y, nreg); }         \
    /* every chunk takes the same number of passes, so all land in one buffer */               \
    T *src = (res[0] == d) ? d : tmp, *dst = (src == d) ? tmp : d;                            \
    free(res);                                                                                \
    uint64_t run = Y;                                                                         \
    while (run < sort_n) {                                                                    \
        uint64_t way = (uint64_t)min_k;                                                       \
        while (way > 2 && run * (way / 2) >= sort_n) way >>= 1;                               \
        uint64_t span = run * way;                                                            \
        if (min_k == 2 && threads == 1 && sort_n == n) {                                      \
            merge_series_##S(src, dst, n, run, run * 2, nreg);                                \
        } else {                                                                              \
            /* merge groups: `way` runs of one sort_n chunk at a time */                      \
            uint64_t groups = 0;                                                              \
            for (uint64_t c0 = 0; c0 < n; c0 += sort_n) {                                     \
                uint64_t clen = n - c0 < sort_n ? n - c0 : sort_n;                            \
                groups += (clen + span - 1) / span; }                                         \
            uint64_t *gs = (uint64_t *)malloc(groups * sizeof(uint64_t));                     \
            uint64_t *ge = (uint64_t *)malloc(groups * sizeof(uint64_t));                     \
            uint64_t gi = 0;                                                                  \
            for (uint64_t c0 = 0; c0 < n; c0 += sort_n) {                                     \
                uint64_t cend = n - c0 < sort_n ? n : c0 + sort_n;                            \
                for (uint64_t g0 = c0; g0 < cend; g0 += span) {                               \
                    gs[gi] = g0; ge[gi] = cend - g0 < span ? cend : g0 + span; gi++; } }       \
            int parts = threads;                                                              \
            int64_t jobs = (int64_t)groups * parts;                                           \
            _Pragma("omp parallel for num_threads(threads) schedule(dynamic, 1)")             \
            for (int64_t j = 0; j < jobs; j++) {                                              \
                uint64_t g = (uint64_t)j / parts; int part = (int)(j % parts);                \
                uint64_t s0 = gs[g], s1 = ge[g], glen = s1 - s0;                              \
                const T *rp[64]; uint64_t rl[64]; int K = 0;                                  \
                for (uint64_t o = s0; o < s1; o += run) {                                     \
                    rp[K] = src + o; rl[K] = s1 - o < run ? s1 - o : run; K++; }              \
                uint64_t r0 = glen * part / parts, r1 = glen * (part + 1) / parts;            \
                if (r0 == r1) continue;                                                       \
                uint64_t sp0[64], sp1[64];                                                    \
                partition_##S(rp, rl, K, r0, sp0);                                            \
                partition_##S(rp, rl, K, r1, sp1);                                            \
                const T *sub[64]; uint64_t sl[64];                                            \
                for (int k = 0; k < K; k++) { sub[k] = rp[k] + sp0[k]; sl[k] = sp1[k] - sp0[k]; }\
                if (K == 2) mergebl_##S(sub[0], sl[0], sub[1], sl[1], dst + s0 + r0, nreg);      \
                else kway_##S(sub, sl, K, dst + s0 + r0, 1024); }                             \
            free(gs); free(ge); }                                                             \
        T *t = src; src = dst; dst = t; run = span; }                                         \
    return src; }

#define LT_PLAIN(a, b) ((a) < (b))
#define LT_KV(a, b) ((a).k < (b).k)
ORC_GEN(uint32_t, u32, LT_PLAIN)
ORC_GEN(uint64_t, u64, LT_PLAIN)
ORC_GEN(kv_t, kv, LT_KV)
ORC_GEN(kv32_t, kv32, LT_KV)

static int width_of(int kt)
{
    switch (kt) {
    case ORC_U32: case ORC_I32: case ORC_F32: return 4;
    case ORC_U64: case ORC_I64: case ORC_KV_U32U32: return 8;
    case ORC_KV_U64U64: case ORC_KV_I64I64: return 16;
    default: return 0;
    }
}

/* ------------------------------------------------------------------------ */
/* Exported checker entry points (ctypes).                                   */
/* ------------------------------------------------------------------------ */

/* Full sort of n items of type kt in `data` (tmp: n items of scratch).  The
 * result is written back into `data`.  sort_n < n sorts every sort_n chunk
 * independently (sort_single_thread's sort_n argument).  threads > 1 runs the
 * multithreaded driver; min_k >= 2 is the merge fan-in of phase 3.
 * Parameters follow config.h's scalar build (REG_TYPE 0, config.h:46-95):
 * P1_N = 8 (u32/i64) or 16 (KV); P2_N = (1 MB / 2) / sizeof(Item)
 * (config.h:265); merge registers nreg = 4 (u32/i64) or 1 (KV). */
int orc_sort(int kt, void *data, void *tmp, uint64_t n, uint64_t sort_n, int threads, int min_k)
{
    int w = width_of(kt);
    if (!w) return -1;
    if (threads < 1) threads = 1;
    if (min_k < 2) min_k = 2;
    if (min_k > 64) min_k = 64;
    encode(kt, data, n, +1);
    void *res;
    uint64_t p2n = (512u * 1024u) / (uint64_t)w;
    if (w == 4) res = sort_u32((uint32_t *)data, (uint32_t *)tmp, n, sort_n, 8, p2n, 4, threads, min_k);
    else if (kt == ORC_KV_U32U32)  /* a KV type: P1_N 16, one merge register (config.h:82-95) */
        res = sort_kv32((kv32_t *)data, (kv32_t *)tmp, n, sort_n, 16, p2n, 1, threads, min_k);
    else if (w == 8) res = sort_u64((uint64_t *)data, (uint64_t *)tmp, n, sort_n, 8, p2n, 4, threads, min_k);
    else res = sort_kv((kv_t *)data, (kv_t *)tmp, n, sort_n, 16, p2n, 1, threads, min_k);
    if (res != data) memcpy(data, res, n * (uint64_t)w);
    encode(kt, data, n, -1);
    return 0;
}

/* Sort each CSR segment [off[i], off[i+1]) independently (no reference
 * counterpart beyond uniform chunks; segments are sorted with orc_sort's
 * single-threaded driver, segments spread over threads). */
int orc_segmented_sort(int kt, void *data, const uint64_t *off, uint64_t nseg, int threads)
{
    int w = width_of(kt);
    if (!w) return -1;
    if (threads < 1) threads = 1;
    int err = 0;
    #pragma omp parallel for num_threads(threads) schedule(dynamic, 16) reduction(|:err)
    for (int64_t s = 0; s < (int64_t)nseg; s++) {
        uint64_t n = off[s + 1] - off[s];
        if (!n) continue;
        char *p = (char *)data + off[s] * (uint64_t)w;
        void *tmp = malloc(n * (uint64_t)w);
        if (!tmp) { err |= 1; continue; }
        err |= orc_sort(kt, p, tmp, n, 0, 1, 2) != 0;
        free(tmp);
    }
    return err ? -1 : 0;
}

/* Apply one of the reference's sorting networks to nb consecutive groups of
 * u32 values: which = 8 (in_register_sort8), 16 (in_register_sort16),
 * 4 / 5 (mergebl2 nreg=4 loop / tail merge network on 8 values),
 * 6 (the upper-block drain network on values 4..7). */
int orc_network_u32(int which, uint32_t *a, uint64_t nb)
{
    for (uint64_t b = 0; b < nb; b++) {
        switch (which) {
        case 8: net_u32(a + 8 * b, NET8, 19); break;
        case 16: net_u32(a + 16 * b, NET16, 60); break;
        case 4: net_u32(a + 8 * b, MRG4_LOOP, 8); break;
        case 5: net_u32(a + 8 * b, MRG4_TAIL, 9); break;
        case 6: net_u32(a + 8 * b, MRG4_UPPER, 5); break;
        default: return -1;
        }
    }
    return 0;
}

/* mergebl2_scalar on two sorted runs (ordered representation expected for
 * signed/float kinds: callers pass u32/u64/KV_U64U64 kinds). */
int orc_merge(int kt, int nreg, const void *A, uint64_t la, const void *B, uint64_t lb, void *C)
{
    int w = width_of(kt);
    if (w == 4) mergebl_u32((const uint32_t *)A, la, (const uint32_t *)B, lb, (uint32_t *)C, nreg);
    else if (kt == ORC_KV_U32U32) mergebl_kv32((const kv32_t *)A, la, (const kv32_t *)B, lb, (kv32_t *)C, 1);
    else if (w == 8) mergebl_u64((const uint64_t *)A, la, (const uint64_t *)B, lb, (uint64_t *)C, nreg);
    else if (w == 16) mergebl_kv((const kv_t *)A, la, (const kv_t *)B, lb, (kv_t *)C, 1);
    else return -1;
    return 0;
}

/* K-way merge tree over runs stored back to back in `base` (lengths lens[]). */
int orc_kway_merge(int kt, const void *base, const uint64_t *lens, int K, void *out, uint64_t bufn)
{
    int w = width_of(kt);
    if (!w || K < 1 || K > 4096) return -1;
    const void **rp = (const void **)malloc(K * sizeof(void *));
    uint64_t o = 0;
    for (int k = 0; k < K; k++) { rp[k] = (const char *)base + o * (uint64_t)w; o += lens[k]; }
    if (bufn < 1) bufn = 1;
    if (w == 4) kway_u32((const uint32_t *const *)rp, lens, K, (uint32_t *)out, bufn);
    else if (kt == ORC_KV_U32U32) kway_kv32((const kv32_t *const *)rp, lens, K, (kv32_t *)out, bufn);
    else if (w == 8) kway_u64((const uint64_t *const *)rp, lens, K, (uint64_t *)out, bufn);
    else kway_kv((const kv_t *const *)rp, lens, K, (kv_t *)out, bufn);
    free(rp);
    return 0;
}

/* Partition_Ptrs restated: splits[k] = items of run k left of output rank. */
int orc_partition(int kt, const void *base, const uint64_t *lens, int K, uint64_t rank, uint64_t *splits)
{
    int w = width_of(kt);
    if (!w || K < 1 || K > 4096) return -1;
    const void **rp = (const void **)malloc(K * sizeof(void *));
    uint64_t o = 0;
    for (int k = 0; k < K; k++) { rp[k] = (const char *)base + o * (uint64_t)w; o += lens[k]; }
    if (rank > o) { free(rp); return -1; }
    if (w == 4) partition_u32((const uint32_t *const *)rp, lens, K, rank, splits);
    else if (kt == ORC_KV_U32U32) partition_kv32((const kv32_t *const *)rp, lens, K, rank, splits);
    else if (w == 8) partition_u64((const uint64_t *const *)rp, lens, K, rank, splits);
    else partition_kv((const kv_t *const *)rp, lens, K, rank, splits);
    free(rp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Seeded input distributions (the writer.h analogue the tests regenerate).  */
/* ------------------------------------------------------------------------ */
static inline uint64_t mix64(uint64_t z)
{   /* splitmix64 finaliser */
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t orc_mix64(uint64_t z) { return mix64(z); }

/* Counter-based family: item i is a function of (seed, i).  The distributions
 * follow Origami/writer.h:27-242 (README D1-D9): all same (writer.h:142-155),
 * almost sorted with every 7th item the largest key (:167-173), Pareto keys
 * min(ceil(beta (1/(1-u) - 1)), cap) with beta = 7 (:73-90; cap 100000),
 * Pareto-length bursts back to back (:91-120) and shuffled (:133-138, here with
 * a seeded order), and the wrapping Fibonacci sequence (:175-220, restated as
 * its loop); u and the keys come from a splitmix64 hash of (seed, stream, i)
 * instead of the reference's LCG / unseeded engines. */
static uint64_t o_ctr(uint64_t seed, uint64_t s, uint64_t i)
{
    return mix64(mix64(seed ^ (s * 0xD1B54A32D192ED03ull)) + i);
}
static double o_unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }
static uint64_t o_pareto(uint64_t x)
{
    double u = o_unit(x);
    double f = ceil(7.0 * (1.0 / (1.0 - u) - 1.0));
    return f > 100000.0 ? 100000u : (uint64_t)f;
}
typedef struct { uint64_t tag, key; } o_tagged;
static int o_tag_cmp(const void *a, const void *b)
{
    uint64_t x = ((const o_tagged *)a)->tag, y = ((const o_tagged *)b)->tag;
    return x < y ? -1 : x > y;
}

static int orc_generate_counter(int kt, int dist, uint64_t seed, uint64_t n, void *out)
{
    int w = width_of(kt);
    uint64_t *key = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    if (!key) return -1;
    double *cdf = NULL;
    const uint64_t R = 1u << 20;
    if (dist == 1) {
        cdf = (double *)malloc(R * sizeof(double));
        double acc = 0;
        for (uint64_t r = 0; r < R; r++) { acc += pow((double)(r + 1), -1.1); cdf[r] = acc; }
    }
    if (dist == 9 || dist == 10) {
        uint64_t pos = 0, b = 0;
        while (pos < n) {
            uint64_t L = o_pareto(o_ctr(seed, 1, b));
            if (L == 0) L = 1;
            uint64_t k = o_ctr(seed, 0, b);
            for (uint64_t j = pos; j < n && j < pos + L; j++) key[j] = k;
            pos += L;
            b++;
        }
        if (dist == 10 && n) {
            int bits = 1;
            while (bits < 64 && (1ull << bits) < n) bits++;
            uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
            o_tagged *t = (o_tagged *)malloc(n * sizeof(o_tagged));
            for (uint64_t i = 0; i < n; i++) { t[i].tag = (o_ctr(seed, 2, i) & ~mask) | (i & mask); t[i].key = key[i]; }
            qsort(t, n, sizeof(o_tagged), o_tag_cmp);
            for (uint64_t i = 0; i < n; i++) key[i] = t[i].key;
            free(t);
        }
    } else if (dist == 11) {
        /* Origami/writer.h:175-220: 0, 1, then sums until the 64-bit sum wraps,
         * when the sequence restarts at 0, 1 */
        uint64_t a = 0, b = 1, c, i = 2;
        if (n > 0) key[0] = 0;
        if (n > 1) key[1] = 1;
        while (i < n) {
            c = a + b;
            if (c < b) {
                a = 0; b = 1;
                key[i++] = 0;
                if (i < n) key[i++] = 1;
            } else {
                a = b; b = c;
                key[i++] = b;
            }
        }
    } else {
        for (uint64_t i = 0; i < n; i++) {
            switch (dist) {
            case 0: key[i] = o_ctr(seed, 0, i); break;
            case 1: {
                double u = o_unit(o_ctr(seed, 0, i)) * cdf[R - 1];
                uint64_t lo = 0, hi = R - 1;
                while (lo < hi) { uint64_t m = (lo + hi) / 2; if (cdf[m] < u) lo = m + 1; else hi = m; }
                key[i] = mix64(lo);
            } break;
            case 2: key[i] = (uint64_t)i * (UINT64_MAX / n); break;
            case 3: key[i] = (uint64_t)(n - 1 - i) * (UINT64_MAX / n); break;
            case 4: key[i] = mix64(o_ctr(seed, 0, i) & 15u); break;
            case 5: {
                uint64_t sum = 0;
                for (uint64_t s = 0; s < 3; s++) {
                    uint64_t x = o_ctr(seed, s, i);
                    for (int q = 0; q < 4; q++) sum += (x >> (16 * q)) & 0xFFFF;
                }
                float f = (float)(((double)sum - 393210.0) / 65536.0);
                uint32_t bits32; memcpy(&bits32, &f, 4);
                key[i] = bits32;
            } break;
            case 6: key[i] = o_ctr(seed, 0, 0); break;
            case 7: key[i] = (i % 7 == 0) ? UINT64_MAX : (uint64_t)i * (UINT64_MAX / n); break;
            case 8: key[i] = o_pareto(o_ctr(seed, 0, i)); break;
            default: free(key); free(cdf); return -1;
            }
        }
    }
    for (uint64_t i = 0; i < n; i++) {
        if (w == 4) {
            uint32_t k32;
            switch (dist) {
            case 2: k32 = (uint32_t)((uint64_t)i * (UINT32_MAX / n)); break;
            case 3: k32 = (uint32_t)((uint64_t)(n - 1 - i) * (UINT32_MAX / n)); break;
            case 7: k32 = (i % 7 == 0) ? UINT32_MAX : (uint32_t)((uint64_t)i * (UINT32_MAX / n)); break;
            case 5: case 8: case 11: k32 = (uint32_t)key[i]; break;
            default: k32 = (uint32_t)(key[i] >> 32); break;
            }
            ((uint32_t *)out)[i] = k32;
        } else if (w == 8) {
            ((uint64_t *)out)[i] = key[i];
        } else {
            ((kv_t *)out)[i].k = key[i];
            ((kv_t *)out)[i].v = i;
        }
    }
    if (w == 4 && dist == 5 && n) {
        ((uint32_t *)out)[0] = 0x80000000u;
        ((uint32_t *)out)[n / 2] = 0u;
    }
    free(key);
    free(cdf);
    return 0;
}

/* dist: 0 uniform, 1 zipf(1.1, 2^20 ranks), 2 sorted, 3 reverse, 4 sixteen
 * distinct, 5 normal(0,1) (f32 only, with -0.0 at 0 and +0.0 at n/2), drawn
 * from mt19937 / mt19937_64; with the 0x100 flag, or for dist 6-11, the
 * counter-based family above.  Writes n items of kind kt; KV values are the
 * original index. */
int orc_generate(int kt, int dist, uint64_t seed, uint64_t n, void *out)
{
    int w = width_of(kt);
    if (!w) return -1;
    if (kt == ORC_KV_U32U32) {
        /* {u32 key, u32 value}: the u32 keys of the same distribution, value =
         * original index */
        uint32_t *k = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
        if (!k) return -1;
        int rc = orc_generate(ORC_U32, dist, seed, n, k);
        for (uint64_t i = 0; rc == 0 && i < n; i++) { ((kv32_t *)out)[i].k = k[i]; ((kv32_t *)out)[i].v = (uint32_t)i; }
        free(k);
        return rc;
    }
    if ((dist & 0x100) || (dist >= 6 && dist < 12)) {
        int base = dist & ~0x100;
        if (base < 0 || base >= 12) return -1;
        return orc_generate_counter(kt, base, seed, n, out);
    }
    uint64_t *raw = NULL;
    uint32_t *raw32 = NULL;
    if (w == 4 && dist == 0) {
        raw32 = (uint32_t *)out;
        orc_mt19937((uint32_t)seed, raw32, n);
        return 0;
    }
    raw = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    if (!raw) return -1;
    if (dist == 0 || dist == 1 || dist == 4 || dist == 5) orc_mt19937_64(seed, raw, n);
    double *cdf = NULL;
    if (dist == 1) {
        const uint64_t R = 1u << 20;
        cdf = (double *)malloc(R * sizeof(double));
        double acc = 0;
        for (uint64_t r = 0; r < R; r++) { acc += pow((double)(r + 1), -1.1); cdf[r] = acc; }
    }
    for (uint64_t i = 0; i < n; i++) {
        uint64_t key;
        switch (dist) {
        case 0: key = raw[i]; break;
        case 1: {
            const uint64_t R = 1u << 20;
            double u = (double)(raw[i] >> 11) * (1.0 / 9007199254740992.0) * cdf[R - 1];
            uint64_t lo = 0, hi = R - 1;
            while (lo < hi) { uint64_t m = (lo + hi) / 2; if (cdf[m] < u) lo = m + 1; else hi = m; }
            key = mix64(lo);
        } break;
        case 2: key = n ? (uint64_t)i * (UINT64_MAX / n) : 0; break;
        case 3: key = n ? (uint64_t)(n - 1 - i) * (UINT64_MAX / n) : 0; break;
        case 4: key = mix64(raw[i] & 15u); break;
        case 5: {
            double u1 = ((double)(raw[i] >> 11) + 1.0) * (1.0 / 9007199254740992.0);
            double u2 = (double)((raw[i] * 0x9E3779B97F4A7C15ull) >> 11) * (1.0 / 9007199254740992.0);
            double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
            float f = (float)z;
            uint32_t b; memcpy(&b, &f, 4);
            key = b;
        } break;
        default: free(raw); free(cdf); return -1;
        }
        if (w == 4) {
            uint32_t k32 = (dist == 5) ? (uint32_t)key : (uint32_t)(key >> 32);
            if (dist == 2 || dist == 3) k32 = n ? (uint32_t)((uint64_t)(dist == 2 ? i : n - 1 - i) * (UINT32_MAX / n)) : 0;
            ((uint32_t *)out)[i] = k32;
        } else if (w == 8) {
            ((uint64_t *)out)[i] = key;
        } else {
            ((kv_t *)out)[i].k = key;
            ((kv_t *)out)[i].v = i;
        }
    }
    if (w == 4 && dist == 5 && n) {
        uint32_t nz = 0x80000000u, pz = 0;
        ((uint32_t *)out)[0] = nz;
        ((uint32_t *)out)[n / 2] = pz;
    }
    free(raw);
    free(cdf);
    return 0;
}
