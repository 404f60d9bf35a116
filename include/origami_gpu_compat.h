/*
 * origami_gpu_compat.h -- drop-in for Origami's sort entry points on the GPU.
 *
 * Host code written against the reference (arif-arman/origami-sort) calls
 *   origami_sorter::sort_single_thread<Item, Reg>(d, tmp, end, sort_n, min_k, interim)
 *       Origami/sorter.h:3851-3852
 *   origami_sorter::sort_multi_thread<Item, Reg>(d, out, n, n_threads, n_cores, min_k, kway_buf)
 *       Origami/sorter.h:4194-4195
 * on host arrays of ui (uint32), i64, KeyValue<i64, i64> or KeyValue<ui, ui>
 * (Origami/commons.h:47-72), and also of int32, float, ui64 and
 * KeyValue<ui64, ui64>; any other item type fails to compile.
 * Including this header instead of sorter.h keeps those calls compiling; they
 * run on the GPU through libogs.so (include/ogs.h).  The SIMD register type,
 * thread counts, min_k and scratch buffers only tune the CPU algorithm and are
 * accepted and ignored.  Differences a caller can observe:
 *   - the result is always written back into d, and d is returned (the
 *     reference returns d or tmp depending on its pass count);
 *   - n need not be a power of two, and a failure returns nullptr after
 *     printing ogs_last_error() instead of exiting (ReportError, commons.h:104);
 *   - sort_n < n sorts every sort_n chunk independently, as the reference does.
 * The device is OGS_DEVICE from the environment (default 0).
 */
#ifndef ORIGAMI_GPU_COMPAT_H
#define ORIGAMI_GPU_COMPAT_H

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <type_traits>

#include "ogs.h"

#ifndef ORIGAMI_GPU_COMPAT_NO_TYPES
typedef int64_t i64;
typedef uint64_t ui64;
typedef uint32_t ui;

/* Packed {key, value} record, laid out like the reference's KeyValue
 * (Origami/commons.h:51-72): 16 bytes for 64-bit keys, 8 for KeyValue<ui, ui>;
 * comparisons look at the key only. */
#pragma pack(push, 1)
template <typename Keytype, typename Valuetype>
struct KeyValue {
    Keytype key;
    Valuetype value;
    bool operator<(const KeyValue &kv) const { return key < kv.key; }
    bool operator>(const KeyValue &kv) const { return key > kv.key; }
    bool operator!=(const KeyValue &kv) const { return key != kv.key; }
    bool operator<=(const KeyValue &kv) const { return key <= kv.key; }
};
#pragma pack(pop)
#endif

namespace origami_sorter {
namespace gpu_detail {

inline ogs_ctx *context() {
    static ogs_ctx *ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        const char *dev = std::getenv("OGS_DEVICE");
        if (ogs_ctx_create(&ctx, dev ? std::atoi(dev) : 0) != OGS_OK) {
            std::fprintf(stderr, "origami_gpu_compat: %s\n", ogs_last_error());
            ctx = nullptr;
        }
    });
    return ctx;
}

template <typename T> struct always_false : std::false_type {};

// The ABI kind of an item type.  Only the types below map; anything else
// (unsigned long long, double, KeyValue<i32, i32>, ...) stops the build here,
// so no item is ever read with another type's width.
template <typename Item> constexpr ogs_key_t key_type() {
    if constexpr (std::is_same<Item, uint32_t>::value) return OGS_U32;
    else if constexpr (std::is_same<Item, int32_t>::value) return OGS_I32;
    else if constexpr (std::is_same<Item, float>::value) return OGS_F32;
    else if constexpr (std::is_same<Item, uint64_t>::value) return OGS_U64;
    else if constexpr (std::is_same<Item, int64_t>::value) return OGS_I64;
    else if constexpr (std::is_same<Item, KeyValue<int64_t, int64_t>>::value) return OGS_KV_I64I64;
    else if constexpr (std::is_same<Item, KeyValue<uint64_t, uint64_t>>::value) return OGS_KV_U64U64;
    else if constexpr (std::is_same<Item, KeyValue<uint32_t, uint32_t>>::value) {
        static_assert(sizeof(Item) == 8, "KeyValue<ui, ui> must be the packed 8-byte record");
        return OGS_KV_U32U32;
    } else {
        static_assert(always_false<Item>::value,
                      "origami_gpu_compat: unsupported item type (ui, i32, float, ui64, i64, KeyValue<i64, i64>, "
                      "KeyValue<ui64, ui64> or KeyValue<ui, ui>)");
        return OGS_U32;
    }
}

template <typename Item> Item *run(Item *d, uint64_t n, uint64_t chunk) {
    ogs_ctx *ctx = context();
    if (!ctx) return nullptr;
    if (ogs_sort_host(ctx, key_type<Item>(), (void *)d, n, chunk < n ? chunk : 0) != OGS_OK) {
        std::fprintf(stderr, "origami_gpu_compat: %s\n", ogs_last_error());
        return nullptr;
    }
    return d;
}

}  // namespace gpu_detail

template <typename Item, typename Reg = Item, unsigned... Tuned>
Item *sort_single_thread(Item *d, Item *tmp_buf, Item *end, uint64_t sort_n, unsigned min_k = 2,
                         Item *interim_buf = nullptr) {
    (void)tmp_buf; (void)min_k; (void)interim_buf;
    const uint64_t n = (uint64_t)(end - d);
    return gpu_detail::run(d, n, sort_n);
}

template <typename Item, typename Reg = Item, unsigned... Tuned>
Item *sort_multi_thread(Item *d, Item *out, uint64_t n_items, unsigned n_threads, unsigned n_cores, unsigned min_k,
                        Item *kway_buf, Item *d_back = nullptr) {
    (void)out; (void)n_threads; (void)n_cores; (void)min_k; (void)kway_buf; (void)d_back;
    return gpu_detail::run(d, n_items, n_items);
}

}  // namespace origami_sorter

#endif /* ORIGAMI_GPU_COMPAT_H */
