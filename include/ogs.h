/*
 * ogs.h -- C-ABI of the MI355X (gfx950) Origami mergesort library (libogs.so).
 *
 * The reference (arif-arman/origami-sort) is a header-only C++ template
 * library whose sort entry points are
 *   origami_sorter::sort_single_thread(Item* d, Item* tmp, Item* end, ui64 sort_n,
 *                                      ui min_k, Item* interim)   Origami/sorter.h:3851-3852
 *   origami_sorter::sort_multi_thread(Item* d, Item* out, ui64 n, ui n_threads,
 *                                     ui n_cores, ui min_k, Item* kway_buf)
 *                                                                  Origami/sorter.h:4194-4195
 * with the item type fixed at compile time (Origami/config.h:23-24): uint32,
 * signed int64, or a packed 16-byte {int64 key, int64 value}
 * (Origami/commons.h:51-72); the reference also declares the packed 8-byte
 * KeyValue<ui, ui> (Origami/commons.h:71).  This ABI takes the same shapes at
 * run time and adds the types the reference lacks (int32, uint64, float32,
 * unsigned 16-byte KV).
 * include/origami_gpu_compat.h re-creates the two template entry points on
 * top of it, so reference host code recompiles unchanged.
 *
 * Conventions
 *  - Device pointers unless a name says _host.  Buffers hold n items; items
 *    are 4 bytes (U32/I32/F32), 8 bytes (U64/I64, and KV_U32U32: key at
 *    offset 0, value at offset 4, like KeyValue<ui,ui>) or 16 bytes
 *    (KV_U64U64/KV_I64I64: key at offset 0, value at offset 8, like
 *    KeyValue<i64,i64>).
 *  - Sorting is ascending.  The result lands in *sorted, which is d_data or
 *    d_scratch (the other one is clobbered), as the reference returns d or
 *    tmp (Origami/merger.h:4225, Origami/sorter.h:3909-3914).
 *  - Key-only results are the unique sorted permutation.  KV sorts are not
 *    stable (neither is the reference): keys come out sorted and each key
 *    keeps its multiset of values.  KV_U32U32 sorts each record as the
 *    64-bit number key * 2^32 + value, so equal keys come out with their
 *    values ascending (one of the orders the reference may produce).
 *  - F32 keys sort in IEEE-754 totalOrder (-NaN < -inf < ... < -0 < +0 < ... <
 *    +inf < +NaN).  I32/I64/KV_I64I64 keys are signed; U32/U64/KV_U64U64 are
 *    unsigned.
 *  - Every call returns OGS_OK (0) or a negative ogs_status; the message of
 *    the last failure on the calling thread is ogs_last_error().  No call
 *    exits the process (the reference's ReportError does, commons.h:104).
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Calls are stream-ordered; a context may be shared by threads that use
 *    different streams.
 */
#ifndef OGS_H
#define OGS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    OGS_U32 = 0,        /* uint32  (Origami KEY_TYPE 0, config.h:23) */
    OGS_I32 = 1,        /* int32   (no reference counterpart) */
    OGS_U64 = 2,        /* uint64  (no reference counterpart) */
    OGS_I64 = 3,        /* int64   (Origami KEY_TYPE 1) */
    OGS_F32 = 4,        /* float32 (no reference counterpart) */
    OGS_KV_U64U64 = 5,  /* {uint64 key, uint64 value} (unsigned form of KEY_TYPE 2) */
    OGS_KV_I64I64 = 6,  /* {int64 key, int64 value}   (Origami KEY_TYPE 2) */
    OGS_KV_U32U32 = 7   /* {uint32 key, uint32 value} (KeyValue<ui, ui>, Origami/commons.h:71) */
} ogs_key_t;

typedef enum {
    OGS_OK = 0,
    OGS_ERR_ARG = -1,       /* bad argument (null pointer, unknown key type, ...) */
    OGS_ERR_HIP = -2,       /* a HIP runtime call failed */
    OGS_ERR_NOMEM = -3,     /* device or host allocation failed / over budget */
    OGS_ERR_STATE = -4,     /* call not valid in the object's current state */
    OGS_ERR_TIMEOUT = -5    /* a wait ran past its deadline */
} ogs_status;

typedef struct ogs_ctx ogs_ctx;

/* Context: binds a device, owns pooled workspace (co-rank tables, staging).
 * Replaces the reference's process globals (Origami/sorter.h:3941-3959). */
int ogs_ctx_create(ogs_ctx **ctx, int device);
int ogs_ctx_destroy(ogs_ctx *ctx);

/* Device error word of a context: waits for the device, returns the bits set
 * since the last call in *flags and clears them.  Bit 0: a K-way merge tile
 * found its split table inconsistent, bit 1: the same for a 4-way merge tile;
 * such a tile is left unmerged, so its output range is not the sort.  Neither
 * happens with the tables the library's split kernels write; stream-ordered
 * calls cannot report it in their return code, ogs_sort_host does.  No
 * reference counterpart (its ReportError exits, Origami/commons.h:104). */
int ogs_device_errors(ogs_ctx *ctx, uint32_t *flags);

/* Bytes per item of a key type (0 for an unknown type). */
size_t ogs_item_size(ogs_key_t kt);

/* Full sort of n items in d_data, d_scratch = n items of scratch.
 * Replaces sort_single_thread / sort_multi_thread (Origami/sorter.h:3851,4194)
 * for sort_n == n; any n >= 0 (the reference needs a power of two >= P2_N). */
int ogs_sort(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch, uint64_t n,
             void **d_sorted, void *stream);

/* ogs_sort without allocation: the same sort, with its device workspace (the
 * K-way pass's samples and split tables, the merge passes' co-rank tables)
 * taken from a static device array of OGS_NOALLOC_SCRATCH_BYTES (32 MiB) that
 * each GPU holds for the life of the process, instead of the context's
 * per-stream pool, which grows with hipMalloc.  No reference counterpart: it
 * serves callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The call makes no hipMalloc or hipFree and
 * records no profiling events.  Sizes whose workspace exceeds the array fail
 * with OGS_ERR_ARG before anything is launched: above about 2^26 keys for
 * 4-byte kinds, 2^24 for 8-byte kinds and 2^23 pairs for 16-byte KV with the
 * default tile table.  Calls that overlap on one GPU share the array and race. */
#define OGS_NOALLOC_SCRATCH_BYTES (32u << 20)
int ogs_sort_noalloc(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch, uint64_t n,
                     void **d_sorted, void *stream);

/* Out-of-place form: reads n items from d_in (left intact) and sorts them
 * using d_data and d_scratch (n items each) as ping-pong buffers; the result
 * lands in *d_sorted (d_data or d_scratch).  The reference always sorts in
 * place (sort_single_thread's d, sorter.h:3851); this form serves callers
 * that must keep their input (ORDER BY over a table column). */
int ogs_sort_keep(ogs_ctx *ctx, ogs_key_t kt, const void *d_in, void *d_data, void *d_scratch,
                  uint64_t n, void **d_sorted, void *stream);

/* Key-value sort of a structure of arrays: n keys in d_keys and n values in
 * d_values, 8 bytes each (kt: OGS_KV_U64U64 or OGS_KV_I64I64).  d_work is
 * 32 * n bytes of device scratch.  The block sort reads the two arrays, the
 * passes between run on the 16-byte records of KeyValue<i64,i64>
 * (Origami/commons.h:51-72) in d_work, and the last pass writes the sorted
 * keys and values back to d_keys and d_values, all ordered on `stream`: the
 * same pass count as the AoS sort.  The reference has no such form (its KV
 * type is AoS only); this serves callers that hold keys and values apart. */
int ogs_sort_soa(ogs_ctx *ctx, ogs_key_t kt, void *d_keys, void *d_values, void *d_work, uint64_t n,
                 void *stream);

/* Sorts every chunk of `chunk` items independently (the last may be shorter):
 * the reference's sort_single_thread with sort_n < n (Origami/sorter.h:3857-3868,
 * and the chunked benchmark Origami/bench_sort_phases.cpp:1307-1425). */
int ogs_sort_chunks(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch, uint64_t n,
                    uint64_t chunk, void **d_sorted, void *stream);

/* Sorts each CSR segment [d_offsets[i], d_offsets[i+1]) of d_data
 * independently, all segments in one batched launch sequence.  d_offsets is a
 * device array of nseg+1 non-decreasing uint64 offsets with d_offsets[0] == 0
 * and d_offsets[nseg] == n (the total item count, given in n).  No reference
 * counterpart beyond uniform chunks (see ogs_sort_chunks). */
int ogs_segmented_sort(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch,
                       const uint64_t *d_offsets, uint64_t nseg, uint64_t n,
                       void **d_sorted, void *stream);

/* Batched sort of independent arrays in one launch sequence: item i sorts
 * items[i].n items from items[i].in into items[i].out, using items[i].scratch
 * (n items) as the second buffer.  scratch may be NULL when `in` may be
 * clobbered (then `in` is the second buffer); in may equal out.  Arrays are
 * binned by length into the tile table's size classes (see ogs_tune_set).
 * The serving loop (ogs_server_*) submits its coalesced requests this way. */
typedef struct {
    const void *in;
    void *out;
    void *scratch;
    uint64_t n;
} ogs_batch_item;
int ogs_sort_batch(ogs_ctx *ctx, ogs_key_t kt, const ogs_batch_item *items, uint64_t count, void *stream);

/* A segmented sort captured once as a HIP graph and replayed by one
 * hipGraphLaunch per call.  It sorts the CSR segments of d_data (as
 * ogs_segmented_sort: results land in d_scratch, d_data is consumed).
 * - Bound buffers: the plan holds its own descriptor and co-rank tables. The
 *   pointers d_data and d_scratch, and the offsets' values, must stay as they
 *   were at creation. The data's contents may change between launches.
 * - Every segment runs through the batched kernels; there is no plain-sort
 *   class.
 * - Launches of one plan must not overlap.
 * No reference counterpart (the reference's uniform chunks, sorter.h:3857-3868,
 * are re-planned on every call). It is for callers that sort one layout
 * repeatedly, such as a serving loop's fixed request shapes. */
typedef struct ogs_plan ogs_plan;
int ogs_segmented_plan(ogs_ctx *ctx, ogs_key_t kt, void *d_data, void *d_scratch, const uint64_t *d_offsets,
                       uint64_t nseg, uint64_t n, ogs_plan **plan, void **d_sorted);
int ogs_plan_launch(ogs_plan *plan, void *stream);
int ogs_plan_destroy(ogs_plan *plan);

/* Host-memory convenience: copies h_data (n items) to the device, sorts, and
 * copies the result back into h_data; synchronous.  This is what
 * origami_gpu_compat.h calls for the reference's host-pointer signatures.
 * Returns OGS_ERR_STATE when the device error word (ogs_device_errors) was
 * set during the call. */
int ogs_sort_host(ogs_ctx *ctx, ogs_key_t kt, void *h_data, uint64_t n, uint64_t chunk);

/* Analytic HBM pass count P of a plain sort of n items of type kt with the
 * context's current tiles (the roofline denominator: bytes = P * 2 * n * w). */
int ogs_pass_count(ogs_ctx *ctx, ogs_key_t kt, uint64_t n);

/* Seeded input generators: the datagen::Writer analogue (Origami/writer.h:27-242,
 * README D1-D9), reproducible.  KV values are the original index.
 *
 * Distributions (dist):
 *  0 uniform                    1 Zipf(s=1.1) over 2^20 ranks
 *  2 sorted (evenly spaced)     3 reverse sorted
 *  4 sixteen distinct values    5 N(0,1) float32 bits, -0.0 at 0 and +0.0 at n/2
 *  6 all the same key (D2)      7 almost sorted: every 7th item the largest key (D5)
 *  8 Pareto keys min(ceil(7 (1/(1-u) - 1)), 100000) (D6)
 *  9 Pareto-length bursts of equal keys, back to back (D7)
 * 10 the same bursts shuffled (D8)
 * 11 Fibonacci, wrapping past 2^64 (D9; 4-byte kinds keep the low half)
 *
 * ogs_generate_host with dist 0-5 draws from std::mt19937 (4-byte kinds,
 * dist 0) or std::mt19937_64: the parity inputs.
 * With OGS_GEN_COUNTER | dist, or any dist >= 6, item i is a function of
 * (seed, i) alone (a splitmix64 hash): that is the family ogs_generate writes
 * on the device, and ogs_generate_host mirrors it bit for bit. */
#define OGS_GEN_COUNTER 0x100
int ogs_generate_host(ogs_key_t kt, int dist, uint64_t seed, uint64_t n, void *h_out);
/* Device form of the counter family (dist 0-11, OGS_GEN_COUNTER optional),
 * written to d_out (n items) on `stream` without a host round trip.  Pareto
 * bursts upload their start table; the shuffled form sorts (tag, key) pairs with
 * this library's sort in a temporary device buffer. */
int ogs_generate(ogs_ctx *ctx, ogs_key_t kt, int dist, uint64_t seed, uint64_t n, void *d_out, void *stream);

/* Tile table (the run-time counterpart of the tuned constants in
 * Origami/config.h:45-264).  param: "plain_block", "plain_merge",
 * "seg_tiny_block", "seg_tiny_max", "seg_small_block", "seg_small_max",
 * "seg_large_block", "seg_merge", "plain_corank" (0: a partition kernel
 * computes every merge tile's co-rank first, 1: each merge workgroup searches
 * its own), "seg_plain_min" (log2 of the shortest batch segment that runs as
 * a plain sort instead; 0 = never), "partition_lanes" (lanes per merge tile
 * in the partition kernel: 1, 4, 8 or 16), "plain_fanin" (2: one merge level per
 * HBM pass; 4: two levels per pass, the last pass 2-way when one level is
 * left), "quad_lanes" (lanes per tile split in a 4-way pass: 1, 2, 4 or 8),
 * "kway_fanin" (the largest fan-in of the sample-split K-way merge pass, 2,
 * 4, 8, 16 or 32, which replaces the 2- and 4-way passes of a plain sort; 0 =
 * off; structure-of-arrays KV sorts keep the 2- and 4-way passes), "kway_shape" (index into ogs_shape_info's K-way
 * list), "kway_sample" (log2 of its sample spacing, 3-10), "kway_scorank" (its
 * sample merges: 0 a co-rank table kernel per merge, 1 co-ranks searched inside
 * the merge kernel), "kway_swhole" (1: when a group's samples fit one K-way
 * tile, a pass merges them in one K-way launch; 0: 2-way merges).  Block, merge and
 * K-way shape parameters take an index into ogs_shape_info's lists; *_max
 * take an item count.  Segments of up to
 * seg_tiny_max items (0 = no tiny class) get one workgroup of the tiny shape
 * each, segments up to seg_small_max (0 = the small shape's tile) one
 * workgroup of the small shape, longer ones are cut into tiles of the large
 * shape and merged by passes of the seg_merge shape. */
int ogs_tune_set(ogs_ctx *ctx, ogs_key_t kt, const char *param, int value);
int ogs_tune_get(ogs_ctx *ctx, ogs_key_t kt, const char *param, int *value);

/* Applies a tile-table file: lines "kind param value" (kind u32, i32, f32,
 * u64, i64, kv_u64u64, kv_i64i64 or kv_u32u32), '#' starts a comment; all
 * lines or none (a bad line leaves the table unchanged and names itself in
 * ogs_last_error).  A context starts from the compiled-in table, measured on
 * MI355X; ogs_ctx_create then applies the file named by $OGS_TILES, when set,
 * and fails if it does not apply.  The autotuner
 * (origami-sort_amd/autotune.py) writes such files. */
int ogs_tune_load(ogs_ctx *ctx, const char *path);
/* Restores every key class's compiled-in tile table. */
int ogs_tune_reset(ogs_ctx *ctx);

/* Compiled kernel shapes of a key type: merge = 0 lists block-sort shapes
 * (workgroup size, items per lane, in-wave bitonic levels), merge = 1 merge
 * shapes, merge = 2 K-way merge shapes (both wlev = -1).  Fills entry idx when it exists; returns the count. */
int ogs_shape_info(ogs_key_t kt, int merge, int idx, int *block, int *ipt, int *wlev);

/* Per-kernel timing with HIP events recorded on the launching stream, for
 * plain sorts issued while enabled.  kind: 0 block sort, 1 merge partition
 * (co-rank tables; for K-way passes the sample merges and the tile split
 * table), 2 2-way merge pass, 3 4-way merge pass, 4 K-way merge pass.  read
 * returns the summed milliseconds, the launch count and the items those
 * launches covered (waits for the recorded events). */
int ogs_profile_enable(ogs_ctx *ctx, int on);
int ogs_profile_read(ogs_ctx *ctx, int kind, double *total_ms, uint64_t *launches, uint64_t *items);
int ogs_profile_reset(ogs_ctx *ctx);

/* ---- Serving loop: continuous batching of sort requests -------------------
 * A server owns a device arena of hbm_budget_bytes (the pooled allocator that
 * admission draws from) and `streams` HIP streams.  Submitted requests wait
 * in a FIFO; a scheduler thread coalesces pending requests of one key type
 * into a batch (ogs_sort_batch) as soon as a stream is free and the arena
 * holds the batch's footprint (device request: one scratch copy; host
 * request: staging + scratch = 2x payload), so requests arriving while
 * others run are admitted as capacity frees up and their work overlaps on
 * the device.  The batch's key type is the oldest request's, unless a batch of
 * that type is already running and a request of a type with none waits among
 * the next 64 (then that type goes first).  Each stream has a waiter thread
 * that finishes its batch (host requests are copied back) and timestamps the
 * requests.  No reference counterpart: the reference is a library call
 * (Origami/sorter.h:3851, 4194). */
typedef struct ogs_server ogs_server;
typedef struct {
    uint64_t hbm_budget_bytes;   /* device arena size; admission limit */
    int streams;                 /* batches in flight at once (1..16) */
    uint32_t max_batch_requests; /* 0 = 1024 */
    uint64_t max_batch_items;    /* 0 = unlimited */
    uint64_t coalesce_max_items; /* longer requests run alone as plain sorts; 0 = 2^22 */
} ogs_server_config;
typedef struct {
    uint64_t submitted, completed, rejected, batches, batched_requests, items;
} ogs_server_stats;
#define OGS_SUBMIT_HOST 1 /* src/dst are host pointers (else device pointers) */

int ogs_server_create(ogs_server **srv, ogs_ctx *ctx, const ogs_server_config *cfg);
/* Drains every queued request, then stops the threads and frees the arena. */
int ogs_server_destroy(ogs_server *srv);
/* Queues a sort of n items of type kt from src into dst (dst may equal src;
 * src is left intact otherwise).  Returns a ticket; the buffers must stay
 * valid until the ticket is done. */
int ogs_server_submit(ogs_server *srv, ogs_key_t kt, const void *src, void *dst, uint64_t n, int flags,
                      uint64_t *ticket);
/* Blocks until the ticket is done (timeout_s < 0: no limit); returns the
 * request's status, or OGS_ERR_TIMEOUT. */
int ogs_server_wait(ogs_server *srv, uint64_t ticket, double timeout_s);
int ogs_server_poll(ogs_server *srv, uint64_t ticket, int *done);
/* Seconds since server creation: submitted, admitted into a batch, done. */
int ogs_server_times(ogs_server *srv, uint64_t ticket, double *t_submit, double *t_start, double *t_done);
/* Forgets a finished ticket. */
int ogs_server_release(ogs_server *srv, uint64_t ticket);
int ogs_server_get_stats(ogs_server *srv, ogs_server_stats *stats);

/* Message of the last failed call on this thread ("" if none). */
const char *ogs_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* OGS_H */
