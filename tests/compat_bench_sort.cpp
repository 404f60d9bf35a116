// Drop-in check of include/origami_gpu_compat.h: the call shapes of the
// reference's harness (Origami/bench_sort.cpp:95-101) compiled against the GPU
// library instead of sorter.h.  Exit status = number of failed cases.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "origami_gpu_compat.h"

static int report(const char *name, bool ok) {
    std::printf("%-44s %s\n", name, ok ? "ok" : "FAILED");
    return ok ? 0 : 1;
}

int main(int argc, char **argv) {
    const ui64 n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : (1ull << 20);
    int fails = 0;
    {   // uint32, single-thread entry, sort_n == n (bench_sort.cpp:96-98)
        std::vector<ui> d(n), tmp(n);
        std::mt19937 g(42);
        for (auto &x : d) x = (ui)g();
        std::vector<ui> ref = d;
        std::sort(ref.begin(), ref.end());
        ui *o = origami_sorter::sort_single_thread<ui, ui>(d.data(), tmp.data(), d.data() + n, n, 2, nullptr);
        fails += report("sort_single_thread<ui>", o && std::equal(ref.begin(), ref.end(), o));
    }
    {   // int64 (signed, negative keys), multi-thread entry (bench_sort.cpp:99-101)
        std::vector<i64> d(n), out(n);
        std::mt19937_64 g(7);
        for (auto &x : d) x = (i64)g();
        std::vector<i64> ref = d;
        std::sort(ref.begin(), ref.end());
        i64 *o = origami_sorter::sort_multi_thread<i64, i64>(d.data(), out.data(), n, 8, 8, 4, nullptr);
        fails += report("sort_multi_thread<i64>", o && std::equal(ref.begin(), ref.end(), o));
    }
    {   // KeyValue<i64, i64>: keys sorted, value = original index under its own key
        typedef KeyValue<i64, i64> KV;
        std::vector<KV> d(n), orig;
        std::mt19937_64 g(9);
        for (ui64 i = 0; i < n; i++) { d[i].key = (i64)(g() % 1000) - 500; d[i].value = (i64)i; }
        orig = d;
        KV *o = origami_sorter::sort_single_thread<KV, KV>(d.data(), nullptr, d.data() + n, n);
        bool ok = o != nullptr;
        std::vector<char> seen(n, 0);
        for (ui64 i = 0; ok && i < n; i++) {
            if (i && o[i].key < o[i - 1].key) ok = false;
            ui64 v = (ui64)o[i].value;
            if (v >= n || seen[v] || orig[v].key != o[i].key) ok = false;
            else seen[v] = 1;
        }
        fails += report("sort_single_thread<KeyValue<i64,i64>>", ok);
    }
    {   // KeyValue<ui, ui>, the reference's 8-byte record (commons.h:71)
        typedef KeyValue<ui, ui> KV;
        static_assert(sizeof(KV) == 8, "packed 8-byte record");
        std::vector<KV> d(n), orig;
        std::mt19937 g(11);
        for (ui64 i = 0; i < n; i++) { d[i].key = (ui)(g() % 5000); d[i].value = (ui)i; }
        orig = d;
        KV *o = origami_sorter::sort_multi_thread<KV, KV>(d.data(), nullptr, n, 8, 8, 4, nullptr);
        bool ok = o != nullptr;
        std::vector<char> seen(n, 0);
        for (ui64 i = 0; ok && i < n; i++) {
            if (i && o[i].key < o[i - 1].key) ok = false;
            ui64 v = (ui64)o[i].value;
            if (v >= n || seen[v] || orig[v].key != o[i].key) ok = false;
            else seen[v] = 1;
        }
        fails += report("sort_multi_thread<KeyValue<ui,ui>>", ok);
    }
    {   // sort_n < n: every chunk sorted independently (sorter.h:3857-3868)
        const ui64 chunk = n / 8 ? n / 8 : 1;
        std::vector<ui> d(n), tmp(n);
        std::mt19937 g(3);
        for (auto &x : d) x = (ui)g();
        std::vector<ui> ref = d;
        for (ui64 c = 0; c < n; c += chunk) std::sort(ref.begin() + c, ref.begin() + std::min(n, c + chunk));
        ui *o = origami_sorter::sort_single_thread<ui, ui>(d.data(), tmp.data(), d.data() + n, chunk);
        fails += report("sort_single_thread<ui> chunked", o && std::equal(ref.begin(), ref.end(), o));
    }
    std::printf("%d failure(s)\n", fails);
    return fails;
}
