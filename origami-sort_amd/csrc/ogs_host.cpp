// ogs_host.cpp -- host-side parts of the C-ABI that launch nothing: the seeded
// input generator (the datagen::Writer analogue, Origami/writer.h:47-242) and
// the tile-table accessors used by the autotuner.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "gen_core.hpp"
#include "ogs_internal.hpp"

using namespace ogs;

namespace {

uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

const char *kParamNames[TP_COUNT] = {"plain_block",     "plain_merge", "seg_small_block",
                                     "seg_small_max",   "seg_large_block", "seg_merge",
                                     "seg_tiny_block",  "seg_tiny_max", "plain_corank", "seg_plain_min",
                                     "partition_lanes", "plain_fanin", "quad_lanes", "kway_fanin",
                                     "kway_shape", "kway_sample", "kway_scorank", "kway_swhole"};

int param_index(const char *name) {
    if (!name) return -1;
    for (int i = 0; i < TP_COUNT; i++)
        if (!strcmp(name, kParamNames[i])) return i;
    return -1;
}

}  // namespace

namespace ogs {

// Zipf(1.1) cumulative weights over 2^20 ranks, summed in rank order.
void zipf_cdf(std::vector<double> &cdf) {
    cdf.resize(gen::kZipfRanks);
    double acc = 0;
    for (uint64_t r = 0; r < gen::kZipfRanks; r++) {
        acc += std::pow((double)(r + 1), -1.1);
        cdf[r] = acc;
    }
}

// First item index of every Pareto burst until n items are covered
// (burst b: max(1, pareto(ctr(seed, 1, b))) items of key ctr(seed, 0, b)).
void burst_starts(uint64_t seed, uint64_t n, std::vector<uint64_t> &starts) {
    starts.clear();
    for (uint64_t pos = 0, b = 0; pos < n; b++) {
        starts.push_back(pos);
        const uint64_t L = gen::pareto(gen::ctr(seed, 1, b));
        pos += L ? L : 1;
    }
}

int shuffle_bits(uint64_t n) {
    int b = 1;
    while (b < 64 && (1ull << b) < n) b++;
    return b;
}

// Host mirror of the counter-based family (ogs_generate's device output).
int generate_counter_host(const KeyInfo &ki, int dist, uint64_t seed, uint64_t n, void *out) {
    const int w = ki.width;
    std::vector<uint64_t> keys;  // 64-bit keys, for the table-driven distributions
    if (dist == gen::ZIPF) {
        std::vector<double> cdf;
        zipf_cdf(cdf);
        keys.resize(n);
        for (uint64_t i = 0; i < n; i++) {
            const double u = gen::unit(gen::ctr(seed, 0, i)) * cdf[gen::kZipfRanks - 1];
            keys[i] = gen::mix64((uint64_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()));
        }
    } else if (dist == gen::PARETO_BURSTS || dist == gen::PARETO_SHUFFLED) {
        std::vector<uint64_t> st;
        burst_starts(seed, n, st);
        keys.resize(n);
        for (uint64_t b = 0; b < st.size(); b++) {
            const uint64_t e = b + 1 < st.size() ? st[b + 1] : n;
            const uint64_t k = gen::ctr(seed, 0, b);
            for (uint64_t i = st[b]; i < e; i++) keys[i] = k;
        }
        if (dist == gen::PARETO_SHUFFLED) {
            const int bits = shuffle_bits(n);
            std::vector<std::pair<uint64_t, uint64_t>> t(n);
            for (uint64_t i = 0; i < n; i++) t[i] = {gen::shuffle_tag(seed, i, bits), keys[i]};
            std::sort(t.begin(), t.end());
            for (uint64_t i = 0; i < n; i++) keys[i] = t[i].second;
        }
    }
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t k = keys.empty() ? gen::key64(dist, seed, i, n) : keys[i];
        if (w == 4) {
            ((uint32_t *)out)[i] = gen::key32(dist, k, i, n);
        } else if (w == 8) {
            ((uint64_t *)out)[i] = k;
        } else {
            ((uint64_t *)out)[2 * i] = k;
            ((uint64_t *)out)[2 * i + 1] = i;
        }
    }
    if (w == 4 && dist == gen::NORMAL && n) {
        ((uint32_t *)out)[0] = 0x80000000u;
        ((uint32_t *)out)[n / 2] = 0u;
    }
    return OGS_OK;
}

}  // namespace ogs

extern "C" {

// Distributions as documented in include/ogs.h.  4-byte kinds draw raw
// std::mt19937 output for the uniform case (the 2^20 / 2^28 u32 workloads);
// every other case draws std::mt19937_64 and narrows to the top 32 bits.
// The engines' outputs are fixed by the C++ standard, so any conforming
// library reproduces these bytes.
int ogs_generate_host(ogs_key_t kt, int dist, uint64_t seed, uint64_t n, void *h_out) {
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    const int base = dist & ~OGS_GEN_COUNTER;
    if (base < 0 || base >= gen::COUNT) return fail(OGS_ERR_ARG, "unknown distribution %d", dist);
    if (n && !h_out) return fail(OGS_ERR_ARG, "h_out is NULL");
    if (kt == OGS_KV_U32U32) {
        // {u32 key, u32 value}: the four-byte keys of the same distribution,
        // value = original index
        std::vector<uint32_t> k(n);
        const int rc = ogs_generate_host(OGS_U32, dist, seed, n, k.data());
        if (rc) return rc;
        uint32_t *o = (uint32_t *)h_out;
        for (uint64_t i = 0; i < n; i++) {
            o[2 * i] = k[i];
            o[2 * i + 1] = (uint32_t)i;
        }
        return OGS_OK;
    }
    if ((dist & OGS_GEN_COUNTER) || base >= gen::SAME) return generate_counter_host(ki, base, seed, n, h_out);
    const int w = ki.width;
    if (w == 4 && dist == 0) {
        std::mt19937 g((uint32_t)seed);
        uint32_t *o = (uint32_t *)h_out;
        for (uint64_t i = 0; i < n; i++) o[i] = (uint32_t)g();
        return OGS_OK;
    }
    std::mt19937_64 g(seed);
    std::vector<double> cdf;
    std::vector<uint32_t> guide;  // guide[j] = lower_bound of cdf at j/G of the total mass
    const uint64_t R = 1u << 20, G = 1u << 16;
    if (dist == 1) {
        cdf.resize(R);
        double acc = 0;
        for (uint64_t r = 0; r < R; r++) {
            acc += std::pow((double)(r + 1), -1.1);
            cdf[r] = acc;
        }
        guide.resize(G + 1);
        for (uint64_t j = 0; j <= G; j++)
            guide[j] = (uint32_t)(std::lower_bound(cdf.begin(), cdf.end(), (double)j / (double)G * cdf[R - 1]) -
                                  cdf.begin());
    }
    for (uint64_t i = 0; i < n; i++) {
        uint64_t key = 0;
        uint64_t raw = (dist == 0 || dist == 1 || dist == 4 || dist == 5) ? g() : 0;
        switch (dist) {
        case 0: key = raw; break;
        case 1: {
            const double u = (double)(raw >> 11) * (1.0 / 9007199254740992.0) * cdf[R - 1];
            // the exact lower_bound of u, searched inside its guide-table bracket
            // (one bracket of slack either side absorbs rounding in j)
            uint64_t j = (uint64_t)(u / cdf[R - 1] * (double)G);
            if (j >= G) j = G - 1;
            const uint64_t lo = j ? guide[j - 1] : 0;
            const uint64_t hi = std::min<uint64_t>(R, (uint64_t)guide[std::min<uint64_t>(G, j + 2)] + 1);
            key = mix64((uint64_t)(std::lower_bound(cdf.begin() + lo, cdf.begin() + hi, u) - cdf.begin()));
        } break;
        case 2: key = (uint64_t)i * (UINT64_MAX / n); break;
        case 3: key = (uint64_t)(n - 1 - i) * (UINT64_MAX / n); break;
        case 4: key = mix64(raw & 15u); break;
        case 5: {
            const double u1 = ((double)(raw >> 11) + 1.0) * (1.0 / 9007199254740992.0);
            const double u2 = (double)((raw * 0x9E3779B97F4A7C15ull) >> 11) * (1.0 / 9007199254740992.0);
            const float f = (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            uint32_t b;
            memcpy(&b, &f, 4);
            key = b;
        } break;
        }
        if (w == 4) {
            uint32_t k32 = dist == 5 ? (uint32_t)key : (uint32_t)(key >> 32);
            if (dist == 2 || dist == 3) k32 = (uint32_t)((uint64_t)(dist == 2 ? i : n - 1 - i) * (UINT32_MAX / n));
            ((uint32_t *)h_out)[i] = k32;
        } else if (w == 8) {
            ((uint64_t *)h_out)[i] = key;
        } else {
            ((uint64_t *)h_out)[2 * i] = key;
            ((uint64_t *)h_out)[2 * i + 1] = i;
        }
    }
    if (w == 4 && dist == 5 && n) {
        ((uint32_t *)h_out)[0] = 0x80000000u;
        ((uint32_t *)h_out)[n / 2] = 0u;
    }
    return OGS_OK;
}

}  // extern "C"

namespace ogs {
// Sets one tile-table entry of t (the table of ki's key class) after checking
// the value against the parameter's range.
static int tune_set_into(Tuning *t, const KeyInfo &ki, const char *param, int value) {
    const int p = param_index(param);
    if (p < 0) return fail(OGS_ERR_ARG, "unknown tuning parameter '%s'", param ? param : "(null)");
    int cnt = 0;
    const bool is_count = p == TP_SEG_SMALL_MAX || p == TP_SEG_TINY_MAX || p == TP_SEG_PLAIN_MIN;
    bool ok = true;
    if (p == TP_PLAIN_MERGE || p == TP_SEG_MERGE) merge_shapes(ki.kc, &cnt);
    else if (p == TP_PLAIN_CORANK || p == TP_KWAY_SCORANK || p == TP_KWAY_SWHOLE) cnt = 2;
    else if (p == TP_PLAIN_FANIN) ok = value == 2 || value == 4;
    else if (p == TP_QUAD_LANES) ok = value == 1 || value == 2 || value == 4 || value == 8;
    else if (p == TP_KWAY_FANIN) ok = value == 0 || value == 2 || value == 4 || value == 8 || value == 16 || value == 32;
    else if (p == TP_KWAY_SAMPLE) ok = value >= 3 && value <= 10;
    else if (p == TP_KWAY_SHAPE) kway_shapes(ki.kc, &cnt);
    else if (p == TP_PARTITION_LANES) ok = value == 1 || value == 4 || value == 8 || value == 16;
    else if (!is_count) block_shapes(ki.kc, &cnt);
    if (cnt) ok = value >= 0 && value < cnt;
    else if (is_count) ok = value >= 0;
    if (!ok) return fail(OGS_ERR_ARG, "value %d out of range for '%s'", value, param);
    t[ki.kc].v[p] = value;
    return OGS_OK;
}
}  // namespace ogs

extern "C" {

// Tile-table access (no reference counterpart beyond the compile-time
// constants of Origami/config.h).
int ogs_tune_set(ogs_ctx *ctx, ogs_key_t kt, const char *param, int value) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    return tune_set_into(reinterpret_cast<Ctx *>(ctx)->tune, ki, param, value);
}

int ogs_tune_reset(ogs_ctx *ctx) {
    if (!ctx) return fail(OGS_ERR_ARG, "ctx is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    for (int k = 0; k < 3; k++) c->tune[k] = default_tuning((KeyClass)k);
    return OGS_OK;
}

int ogs_tune_get(ogs_ctx *ctx, ogs_key_t kt, const char *param, int *value) {
    if (!ctx || !value) return fail(OGS_ERR_ARG, "NULL argument");
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    const int p = param_index(param);
    if (p < 0) return fail(OGS_ERR_ARG, "unknown tuning parameter '%s'", param ? param : "(null)");
    *value = reinterpret_cast<Ctx *>(ctx)->tune[ki.kc].v[p];
    return OGS_OK;
}

// Tile-table files: one "kind param value" triple per line, '#' comments.
// kind is u32/i32/f32/u64/i64/kv_u64u64/kv_i64i64/kv_u32u32 (the classes share
// tables: u32/i32/f32, u64/i64/kv_u32u32, kv_u64u64/kv_i64i64).  The file is
// applied all or nothing: any bad line leaves the table as it was.
int ogs_tune_load(ogs_ctx *ctx, const char *path) {
    if (!ctx || !path) return fail(OGS_ERR_ARG, "NULL argument");
    FILE *f = fopen(path, "r");
    if (!f) return fail(OGS_ERR_ARG, "cannot open tile table '%s'", path);
    static const char *kinds[] = {"u32", "i32", "u64", "i64", "f32", "kv_u64u64", "kv_i64i64", "kv_u32u32"};
    constexpr int NK = (int)(sizeof(kinds) / sizeof(kinds[0]));
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    Tuning t[3] = {c->tune[0], c->tune[1], c->tune[2]};
    char line[256];
    int lineno = 0, rc = OGS_OK;
    while (fgets(line, sizeof line, f)) {
        lineno++;
        char *h = strchr(line, '#');
        if (h) *h = 0;
        char k[64], p[64];
        int v;
        const int got = sscanf(line, "%63s %63s %d", k, p, &v);
        if (got <= 0) continue;
        if (got != 3) { rc = fail(OGS_ERR_ARG, "%s:%d: expected 'kind param value'", path, lineno); break; }
        int kt = -1;
        for (int i = 0; i < NK; i++)
            if (!strcmp(k, kinds[i])) kt = i;
        if (kt < 0) { rc = fail(OGS_ERR_ARG, "%s:%d: unknown kind '%s'", path, lineno, k); break; }
        KeyInfo ki;
        key_info((ogs_key_t)kt, &ki);
        rc = tune_set_into(t, ki, p, v);
        if (rc) {
            std::string msg = ogs_last_error();
            rc = fail(rc, "%s:%d: %s", path, lineno, msg.c_str());
            break;
        }
    }
    fclose(f);
    if (rc == OGS_OK)
        for (int i = 0; i < 3; i++) c->tune[i] = t[i];
    return rc;
}

// Compiled shapes: merge = 0 lists block-sort shapes (block, ipt, wave levels),
// merge = 1 merge shapes (block, ipt, wave levels = -1).  Returns the count.
int ogs_shape_info(ogs_key_t kt, int merge, int idx, int *block, int *ipt, int *wlev) {
    KeyInfo ki;
    if (!key_info(kt, &ki)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    int cnt = 0;
    if (merge == 2) {
        const KwayShapeFns *t = kway_shapes(ki.kc, &cnt);
        if (idx >= 0 && idx < cnt) {
            if (block) *block = t[idx].block;
            if (ipt) *ipt = t[idx].ipt;
            if (wlev) *wlev = -1;
        }
    } else if (merge) {
        const MergeShapeFns *t = merge_shapes(ki.kc, &cnt);
        if (idx >= 0 && idx < cnt) {
            if (block) *block = t[idx].block;
            if (ipt) *ipt = t[idx].ipt;
            if (wlev) *wlev = -1;
        }
    } else {
        const BlockShapeFns *t = block_shapes(ki.kc, &cnt);
        if (idx >= 0 && idx < cnt) {
            if (block) *block = t[idx].block;
            if (ipt) *ipt = t[idx].ipt;
            if (wlev) *wlev = t[idx].wlev;
        }
    }
    return cnt;
}

}  // extern "C"

namespace ogs {
// Applies $OGS_TILES when it is set: a context starts from the compiled-in
// table, and a tile file is an explicit override (autotune.py writes one).
int load_env_tiles(ogs_ctx *ctx) {
    const char *env = getenv("OGS_TILES");
    if (!env || !*env) return OGS_OK;
    const int rc = ogs_tune_load(ctx, env);
    if (rc) std::fprintf(stderr, "libogs: OGS_TILES not applied: %s\n", ogs_last_error());
    return rc;
}
}  // namespace ogs
