"""GPU parity: the HIP library (through its C-ABI) against the oracle on the same
seeded inputs.  Key-only outputs must be bit-exact; KV outputs must have the
bit-exact key sequence and, per key, the same multiset of values (Origami is
not stable, Origami/merger.h:84 vs 1002 vs merge_tree.h:40)."""
import sys

import numpy as np
import pytest

import ogs
import oracle

pytestmark = pytest.mark.gpu

KV = (ogs.KV_U64U64, ogs.KV_I64I64, ogs.KV_U32U32)
KV16 = (ogs.KV_U64U64, ogs.KV_I64I64)  # the kinds ogs_sort_soa takes


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    c = ogs.Context(0)
    yield c
    c.close()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_sort(ctx, kind, a):
    import torch
    n = a.size
    w = ogs.WIDTH[kind]
    d = ogs.to_device(a) if n else torch.zeros(16, dtype=torch.uint8, device="cuda")
    s = torch.empty_like(d)
    res = ctx.sort_tensor(kind, d, s, n)
    torch.cuda.synchronize()
    return ogs.from_device(res, kind, n) if n else a.copy()


def assert_same(kind, got, a):
    ref = oracle.sort(kind, a)
    if kind in KV:
        assert np.array_equal(got["key"], ref["key"]), "key sequence differs"
        # values are a permutation of the input's and each sits under its own key
        assert np.array_equal(np.sort(got["value"]), np.sort(a["value"]))
        order_in = np.lexsort((a["value"], a["key"]))
        order_out = np.lexsort((got["value"], got["key"]))
        assert np.array_equal(a[order_in].view(np.uint8), got[order_out].view(np.uint8))
    else:
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


SIZES = [0, 1, 2, 63, 64, 65, 1000, 2047, 2048, 2049, 4095, 4096, 4097, 8191, 8192, 12289,
         65536, 100003, 1 << 18, (1 << 20) + 17]


@pytest.mark.parametrize("n", SIZES)
def test_u32_uniform_sizes(ctx, n):
    say(f"case u32 uniform n={n}")
    a = ogs.generate(ogs.U32, ogs.UNIFORM, n, seed=42 + n)
    assert_same(ogs.U32, gpu_sort(ctx, ogs.U32, a), a)


def test_u32_config1_2pow20_seed42(ctx):
    # 2^20 uint32 keys, raw mt19937(42): the reference's single-threaded case
    say("case u32 2^20 seed 42")
    a = ogs.generate(ogs.U32, ogs.UNIFORM, 1 << 20, seed=42)
    got = gpu_sort(ctx, ogs.U32, a)
    assert np.array_equal(got, np.sort(a))
    assert_same(ogs.U32, got, a)


@pytest.mark.parametrize("kind", [ogs.U32, ogs.I32, ogs.F32, ogs.U64, ogs.I64, ogs.KV_U64U64, ogs.KV_I64I64,
                                  ogs.KV_U32U32])
@pytest.mark.parametrize("dist", [ogs.UNIFORM, ogs.ZIPF, ogs.SORTED, ogs.REVERSE, ogs.DISTINCT16])
@pytest.mark.parametrize("n", [777, 5000, 300007])
def test_kinds_and_distributions(ctx, kind, dist, n):
    say(f"case kind={ogs.KIND_NAMES[kind]} dist={dist} n={n}")
    a = ogs.generate(kind, ogs.NORMAL if kind == ogs.F32 and dist == ogs.UNIFORM else dist, n, seed=7 + n)
    assert_same(kind, gpu_sort(ctx, kind, a), a)


@pytest.mark.parametrize("kind", [ogs.U32, ogs.U64, ogs.KV_U64U64, ogs.I32, ogs.I64, ogs.KV_I64I64, ogs.KV_U32U32])
@pytest.mark.parametrize("n", [5, 300, 3001, 70001])
def test_extreme_keys(ctx, kind, n):
    # maximum and minimum keys mixed with pads in partial tiles
    say(f"case extremes kind={ogs.KIND_NAMES[kind]} n={n}")
    rng = np.random.default_rng(n)
    a = ogs.generate(kind, ogs.UNIFORM, n, seed=n)
    raw = a.view(np.uint8).reshape(n, -1)
    pick = rng.random(n)
    w = ogs.WIDTH[kind]
    kw = 4 if kind == ogs.KV_U32U32 else min(w, 8)
    raw[pick < 0.2, :kw] = 0xFF
    raw[(pick >= 0.2) & (pick < 0.3), :kw] = 0x00
    if kind in (ogs.I32, ogs.I64, ogs.KV_I64I64):
        raw[(pick >= 0.3) & (pick < 0.4), kw - 1] = 0x80
        raw[(pick >= 0.3) & (pick < 0.4), : kw - 1] = 0
        raw[(pick >= 0.4) & (pick < 0.5), kw - 1] = 0x7F
    assert_same(kind, gpu_sort(ctx, kind, a), a)


@pytest.mark.parametrize("n", [1000, 70001, (1 << 20) + 5, 3 << 20])
def test_kv32_records(ctx, n):
    # KeyValue<ui, ui>: 8-byte records sorted on the key (the oracle's bar: key
    # sequence bit-exact, each key's values the same multiset); this library
    # orders equal keys by value, so the output is also the lexicographic
    # (key, value) order exactly
    say(f"case kv32 n={n}")
    rng = np.random.default_rng(n)
    a = np.empty(n, dtype=ogs.NP_DTYPE[ogs.KV_U32U32])
    a["key"] = rng.integers(0, max(2, n // 8), n, dtype=np.uint32)
    a["value"] = rng.permutation(n).astype(np.uint32)
    got = gpu_sort(ctx, ogs.KV_U32U32, a)
    assert_same(ogs.KV_U32U32, got, a)
    assert np.array_equal(got.view(np.uint8), a[np.lexsort((a["value"], a["key"]))].view(np.uint8))


def test_f32_specials(ctx):
    say("case f32 specials")
    vals = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1e-45, -1e-45, 3.4e38, -3.4e38, 1.0, -1.0],
                    np.float32)
    a = np.tile(vals, 1000)
    np.random.default_rng(1).shuffle(a)
    assert_same(ogs.F32, gpu_sort(ctx, ogs.F32, a), a)


@pytest.mark.parametrize("kind", [ogs.U32, ogs.U64, ogs.KV_U64U64, ogs.F32, ogs.KV_U32U32])
def test_chunked_sort(ctx, kind):
    import torch
    n, chunk = 300000, 4096 * 3
    say(f"case chunks kind={ogs.KIND_NAMES[kind]}")
    a = ogs.generate(kind, ogs.NORMAL if kind == ogs.F32 else ogs.UNIFORM, n, seed=3)
    d = ogs.to_device(a)
    s = torch.empty_like(d)
    p = ctx.sort_chunks_ptr(kind, d.data_ptr(), s.data_ptr(), n, chunk)
    torch.cuda.synchronize()
    got = ogs.from_device(d if p == d.data_ptr() else s, kind, n)
    ref = oracle.sort(kind, a, sort_n=chunk)
    if kind in KV:
        assert np.array_equal(got["key"], ref["key"])
    else:
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("kind", [ogs.U32, ogs.I64, ogs.KV_U64U64, ogs.F32, ogs.KV_U32U32])
def test_segmented_sort(ctx, kind):
    import torch
    say(f"case segmented kind={ogs.KIND_NAMES[kind]}")
    rng = np.random.default_rng(11)
    lens = np.floor(2.0 ** (rng.random(600) * 17)).astype(np.int64)
    lens[:5] = [0, 1, 0, 2, 4096]
    lens[5:9] = [4097, 8192, 8193, 1 << 17]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n = int(off[-1])
    a = ogs.generate(kind, ogs.NORMAL if kind == ogs.F32 else ogs.ZIPF, n, seed=5)
    d = ogs.to_device(a)
    s = torch.empty_like(d)
    doff = torch.from_numpy(off.view(np.int64).copy()).cuda()
    p = ctx.segmented_ptr(kind, d.data_ptr(), s.data_ptr(), doff.data_ptr(), lens.size, n)
    torch.cuda.synchronize()
    got = ogs.from_device(d if p == d.data_ptr() else s, kind, n)
    ref = oracle.segmented_sort(kind, a, off, threads=4)
    if kind in KV:
        assert np.array_equal(got["key"], ref["key"])
        for i in range(lens.size):
            g = got[off[i]:off[i + 1]]
            x = a[off[i]:off[i + 1]]
            assert np.array_equal(np.sort(g["value"]), np.sort(x["value"]))
    else:
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("kind", [ogs.U32, ogs.KV_U64U64, ogs.F32, ogs.U64])
def test_segmented_plan_graph_replays(ctx, kind):
    # the graph-captured segmented sort, replayed several times with eager
    # work (refilling the input, an eager sort) and synchronisation between
    # replays; every replay must equal the oracle and the eager result
    import torch
    rng = np.random.default_rng(11 + kind)
    lens = np.concatenate([[0, 1, 2, 33, 64, 1000, 4096, 4097],
                           rng.integers(0, 70000, 40), [1 << 18, 300001]])
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n = int(off[-1])
    doff = torch.from_numpy(off.view(np.int64).copy()).cuda()
    d = torch.empty(n * ogs.WIDTH[kind], dtype=torch.uint8, device="cuda")
    s = torch.empty_like(d)
    plan, res = ctx.segmented_plan(kind, d.data_ptr(), s.data_ptr(), doff.data_ptr(), lens.size, n)
    try:
        assert res == s.data_ptr()
        for rep in range(3):
            say(f"case plan kind={ogs.KIND_NAMES[kind]} replay={rep} n={n}")
            dist = ogs.NORMAL if kind == ogs.F32 else (ogs.ZIPF if rep % 2 else ogs.UNIFORM)
            a = ogs.generate(kind, dist, n, seed=100 * kind + rep)
            d.copy_(ogs.to_device(a))
            torch.cuda.synchronize()
            plan.launch()
            torch.cuda.synchronize()
            got = ogs.from_device(s, kind, n)
            ref = oracle.segmented_sort(kind, a, off, threads=4)
            if kind in KV:
                assert np.array_equal(got["key"], ref["key"])
                # the same (key, value) records, only reordered within ties
                oi, oo = np.lexsort((a["value"], a["key"])), np.lexsort((got["value"], got["key"]))
                assert np.array_equal(a[oi].view(np.uint8), got[oo].view(np.uint8))
            else:
                assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
            # eager path on the same buffers between replays
            d.copy_(ogs.to_device(a))
            p = ctx.segmented_ptr(kind, d.data_ptr(), s.data_ptr(), doff.data_ptr(), lens.size, n)
            torch.cuda.synchronize()
            eager = ogs.from_device(s if p == s.data_ptr() else d, kind, n)
            if kind in KV:
                assert np.array_equal(eager["key"], got["key"])
            else:
                assert np.array_equal(eager.view(np.uint8), got.view(np.uint8))
    finally:
        plan.close()


def test_host_entry_point(ctx):
    say("case ogs_sort_host")
    a = ogs.generate(ogs.I64, ogs.UNIFORM, 123457, seed=9)
    got = ctx.sort_host(ogs.I64, a.copy())
    assert np.array_equal(got, np.sort(a))


@pytest.mark.parametrize("kind", KV16)
@pytest.mark.parametrize("n", [0, 1, 2, 1000, 100003, (1 << 20) + 5])
def test_structure_of_arrays_kv(ctx, kind, n):
    # ogs_sort_soa: keys and values in separate arrays, packed, sorted and
    # unpacked on the device; checked like the AoS sort
    import torch
    say(f"case soa kind={ogs.KIND_NAMES[kind]} n={n}")
    a = ogs.generate(kind, ogs.ZIPF if n % 2 else ogs.UNIFORM, n, seed=n + 3)
    keys = np.ascontiguousarray(a["key"])
    vals = np.ascontiguousarray(a["value"]) * np.uint64(7) if kind == ogs.KV_U64U64 else \
        np.ascontiguousarray(a["value"]) * np.int64(-7)
    a = a.copy()
    a["value"] = vals
    dk = torch.from_numpy(keys.view(np.int64).copy()).cuda() if n else torch.zeros(2, dtype=torch.int64, device="cuda")
    dv = torch.from_numpy(vals.view(np.int64).copy()).cuda() if n else torch.zeros(2, dtype=torch.int64, device="cuda")
    work = torch.empty(max(32 * n, 16), dtype=torch.uint8, device="cuda")
    ctx.sort_soa_ptr(kind, dk.data_ptr(), dv.data_ptr(), work.data_ptr(), n)
    torch.cuda.synchronize()
    got = np.empty(n, dtype=ogs.NP_DTYPE[kind])
    got["key"] = dk[:n].cpu().numpy().view(got["key"].dtype)
    got["value"] = dv[:n].cpu().numpy().view(got["value"].dtype)
    if n:
        assert_same(kind, got, a)


def test_sort_soa_rejects_key_only_kinds(ctx):
    with pytest.raises(ogs.OgsError):
        ctx.sort_soa_ptr(ogs.U64, 16, 16, 16, 4)


def test_sort_is_repeatable(ctx):
    # the same input sorted twice gives the same bytes (also for KV ties)
    say("case repeatability")
    a = ogs.generate(ogs.KV_U64U64, ogs.DISTINCT16, 200003, seed=1)
    g1 = gpu_sort(ctx, ogs.KV_U64U64, a)
    g2 = gpu_sort(ctx, ogs.KV_U64U64, a)
    assert np.array_equal(g1.view(np.uint8), g2.view(np.uint8))


@pytest.mark.parametrize("kind", [ogs.U32, ogs.U64, ogs.KV_U64U64])
def test_every_compiled_shape(ctx, kind):
    # every block-sort and merge shape the tuning table may select is parity-checked
    bshapes = ogs.shapes(kind, 0)
    mshapes = ogs.shapes(kind, 1)
    try:
        for bi, (b, ipt, w) in enumerate(bshapes):
            ctx.tune_set(kind, "plain_block", bi)
            mi = bi % len(mshapes)
            ctx.tune_set(kind, "plain_merge", mi)
            for n in (b * ipt - 3, 5 * b * ipt + 77):
                say(f"case shape kind={ogs.KIND_NAMES[kind]} block={bshapes[bi]} merge={mshapes[mi]} n={n}")
                a = ogs.generate(kind, ogs.UNIFORM if n % 2 else ogs.DISTINCT16, n, seed=n)
                assert_same(kind, gpu_sort(ctx, kind, a), a)
        # the 2-way (and, where the table asks for it, 4-way) merge passes of
        # every merge shape: the K-way plan off
        ctx.tune_set(kind, "kway_fanin", 0)
        for mi, (b, ipt, w) in enumerate(mshapes):
            for corank in (0, 1):
                ctx.tune_set(kind, "plain_block", 0)
                ctx.tune_set(kind, "plain_merge", mi)
                ctx.tune_set(kind, "plain_corank", corank)
                n = 200003
                say(f"case merge shape kind={ogs.KIND_NAMES[kind]} merge={mshapes[mi]} corank={corank} n={n}")
                a = ogs.generate(kind, ogs.ZIPF, n, seed=mi)
                assert_same(kind, gpu_sort(ctx, kind, a), a)
    finally:
        ctx.tune_reset()


@pytest.mark.parametrize("plain_min", [22, 16])
@pytest.mark.parametrize("kind", [ogs.U32, ogs.KV_I64I64, ogs.F32])
def test_pointer_batch(ctx, kind, plain_min):
    # ogs_sort_batch: independent arrays, out-of-place, in-place (out == in) and
    # without scratch (in clobbered), all in one launch sequence
    import torch
    say(f"case pointer batch kind={ogs.KIND_NAMES[kind]}")
    rng = np.random.default_rng(2)
    lens = [0, 1, 17, 600, 4096, 4097, 70000, 300001] + [int(x) for x in rng.integers(1, 40000, 20)]
    arrays, items, keep = [], [], []
    for j, n in enumerate(lens):
        a = ogs.generate(kind, ogs.NORMAL if kind == ogs.F32 else ogs.UNIFORM, n, seed=100 + j)
        arrays.append(a)
        din = ogs.to_device(a) if n else torch.zeros(16, dtype=torch.uint8, device="cuda")
        mode = j % 3
        if mode == 0:
            out, scr = torch.empty_like(din), torch.empty_like(din)
            items.append((din.data_ptr(), out.data_ptr(), scr.data_ptr(), n))
        elif mode == 1:
            out, scr = din, torch.empty_like(din)
            items.append((din.data_ptr(), din.data_ptr(), scr.data_ptr(), n))
        else:
            out, scr = torch.empty_like(din), None
            items.append((din.data_ptr(), out.data_ptr(), 0, n))
        keep.append((din, out, scr))
    ctx.tune_set(kind, "seg_plain_min", plain_min)  # 16: the largest arrays run as plain sorts
    try:
        ctx.sort_batch_ptrs(kind, items)
        torch.cuda.synchronize()
    finally:
        ctx.tune_reset()
    for a, (din, out, scr), n in zip(arrays, keep, lens):
        if n == 0:
            continue
        assert_same(kind, ogs.from_device(out, kind, n), a)


def test_segmented_tiny_class(ctx):
    # the wavefront-sized class (one 64-lane workgroup per segment) with a cut-off
    import torch
    kind = ogs.U32
    bshapes = ogs.shapes(kind, 0)
    tiny = [i for i, s in enumerate(bshapes) if s[0] == 64][0]
    ctx.tune_set(kind, "seg_tiny_block", tiny)
    ctx.tune_set(kind, "seg_tiny_max", 300)
    try:
        say("case segmented tiny class")
        rng = np.random.default_rng(4)
        lens = rng.integers(0, 2000, 3000)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
        n = int(off[-1])
        a = ogs.generate(kind, ogs.UNIFORM, n, seed=8)
        d = ogs.to_device(a)
        s = torch.empty_like(d)
        doff = torch.from_numpy(off.view(np.int64).copy()).cuda()
        p = ctx.segmented_ptr(kind, d.data_ptr(), s.data_ptr(), doff.data_ptr(), lens.size, n)
        torch.cuda.synchronize()
        got = ogs.from_device(d if p == d.data_ptr() else s, kind, n)
        ref = oracle.segmented_sort(kind, a, off, threads=4)
        assert np.array_equal(got, ref)
    finally:
        ctx.tune_reset()


@pytest.mark.parametrize("kind", [ogs.U32, ogs.U64, ogs.KV_U64U64, ogs.F32, ogs.I64])
@pytest.mark.parametrize("dist", [ogs.UNIFORM, ogs.DISTINCT16, ogs.SORTED, ogs.REVERSE, ogs.ZIPF])
def test_four_way_passes(ctx, kind, dist):
    # plain_fanin 4: two merge levels per HBM pass (tiles cut at merge-path
    # points every TILE/2 items of each pair's merge, two LDS merge levels);
    # odd level counts end with a 2-way pass
    bshapes = ogs.shapes(kind, 0)
    old = {p: ctx.tune_get(kind, p) for p in ("plain_fanin", "plain_block", "plain_merge")}
    try:
        ctx.tune_set(kind, "plain_fanin", 4)
        for bi in (0, len(bshapes) - 1):
            ctx.tune_set(kind, "plain_block", bi)
            tile = bshapes[bi][0] * bshapes[bi][1]
            for n in (2 * tile + 1, 3 * tile, 4 * tile - 1, 5 * tile + 7, 16 * tile + 3, 33 * tile + 5,
                      (1 << 20) + 5):
                say(f"case 4-way kind={ogs.KIND_NAMES[kind]} dist={dist} block={bshapes[bi]} n={n}")
                d = ogs.NORMAL if kind == ogs.F32 and dist == ogs.UNIFORM else dist
                a = ogs.generate(kind, d, n, seed=n + dist)
                assert_same(kind, gpu_sort(ctx, kind, a), a)
    finally:
        for p, v in old.items():
            ctx.tune_set(kind, p, v)


@pytest.mark.parametrize("kind", [ogs.U32, ogs.KV_U64U64, ogs.I64])
@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
def test_four_way_split_searches(ctx, kind, lanes):
    # every form of the 4-way split table (1, 2, 4 or 8 lanes per split)
    # gives the same sort
    old = {p: ctx.tune_get(kind, p) for p in ("plain_fanin", "quad_lanes")}
    try:
        ctx.tune_set(kind, "plain_fanin", 4)
        ctx.tune_set(kind, "quad_lanes", lanes)
        for n, dist in ((300007, ogs.ZIPF), ((1 << 21) + 9, ogs.DISTINCT16), (1 << 22, ogs.UNIFORM),
                        ((1 << 20) + 1, ogs.SORTED), (777777, ogs.REVERSE)):
            say(f"case 4-way split kind={ogs.KIND_NAMES[kind]} lanes={lanes} n={n} dist={dist}")
            a = ogs.generate(kind, dist, n, seed=n + lanes)
            assert_same(kind, gpu_sort(ctx, kind, a), a)
    finally:
        for p, v in old.items():
            ctx.tune_set(kind, p, v)


@pytest.mark.parametrize("kind", [ogs.U32, ogs.U64, ogs.KV_U64U64])
def test_four_way_every_merge_shape(ctx, kind):
    mshapes = ogs.shapes(kind, 1)
    old = {p: ctx.tune_get(kind, p) for p in ("plain_fanin", "plain_block", "plain_merge")}
    try:
        ctx.tune_set(kind, "plain_fanin", 4)
        ctx.tune_set(kind, "plain_block", 0)
        for mi, ms in enumerate(mshapes):
            ctx.tune_set(kind, "plain_merge", mi)
            for n in (300007, 1 << 21):
                say(f"case 4-way merge shape kind={ogs.KIND_NAMES[kind]} merge={ms} n={n}")
                a = ogs.generate(kind, ogs.ZIPF if n % 2 else ogs.UNIFORM, n, seed=mi)
                assert_same(kind, gpu_sort(ctx, kind, a), a)
    finally:
        for p, v in old.items():
            ctx.tune_set(kind, p, v)
