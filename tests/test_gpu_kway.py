"""GPU parity of the sample-split K-way merge pass (kway_kernels.hip) against the
oracle: key-only kinds, every fan-in (2, 4, 8, 16, 32), sample spacings, tile shapes,
distributions with heavy ties, ragged sizes, and a large size checked against an
independent device sort of the same keys."""
import sys

import numpy as np
import pytest

import ogs
import oracle

pytestmark = pytest.mark.gpu

KEY_ONLY = (ogs.U32, ogs.I32, ogs.F32, ogs.U64, ogs.I64)


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


def configure(ctx, kind, fanin, sample=5, shape=0, small_block=True):
    # small block tiles give many merge levels at test sizes
    if small_block:
        ctx.tune_set(kind, "plain_block", 10 if ogs.WIDTH[kind] == 4 else 1)  # 256 x 8 keys
    ctx.tune_set(kind, "kway_fanin", fanin)
    ctx.tune_set(kind, "kway_sample", sample)
    ctx.tune_set(kind, "kway_shape", shape)


def restore(ctx, kind):
    ctx.tune_reset()


def gpu_sort(ctx, kind, a):
    import torch
    n = a.size
    d = ogs.to_device(a)
    s = torch.empty_like(d)
    res = ctx.sort_tensor(kind, d, s, n)
    torch.cuda.synchronize()
    return ogs.from_device(res, kind, n)


def check(kind, got, a):
    ref = oracle.sort(kind, a)
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("kind", KEY_ONLY)
@pytest.mark.parametrize("fanin", [2, 4, 8, 16, 32])
@pytest.mark.parametrize("dist", [ogs.UNIFORM, ogs.ZIPF, ogs.SORTED, ogs.REVERSE, ogs.DISTINCT16])
def test_kway_kinds(ctx, kind, fanin, dist):
    n = 300007
    say(f"case kway kind={ogs.KIND_NAMES[kind]} fanin={fanin} dist={dist} n={n}")
    configure(ctx, kind, fanin)
    try:
        a = ogs.generate(kind, ogs.NORMAL if (kind == ogs.F32 and dist == ogs.UNIFORM) else dist, n, seed=11 + fanin)
        check(kind, gpu_sort(ctx, kind, a), a)
    finally:
        restore(ctx, kind)


@pytest.mark.parametrize("n", [2049, 4096, 4097, 16385, 65536, 65537, 262143, 1000003])
@pytest.mark.parametrize("fanin", [4, 16, 32])
def test_kway_sizes(ctx, n, fanin):
    say(f"case kway u32 sizes fanin={fanin} n={n}")
    configure(ctx, ogs.U32, fanin)
    try:
        a = ogs.generate(ogs.U32, ogs.UNIFORM, n, seed=n)
        check(ogs.U32, gpu_sort(ctx, ogs.U32, a), a)
    finally:
        restore(ctx, ogs.U32)


@pytest.mark.parametrize("kind", [ogs.U32, ogs.U64])
@pytest.mark.parametrize("sample", [3, 4, 6, 7])
@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5])
def test_kway_shapes_and_spacing(ctx, kind, sample, shape):
    n = 777777
    say(f"case kway kind={ogs.KIND_NAMES[kind]} sample={sample} shape={shape} n={n}")
    configure(ctx, kind, 16, sample=sample, shape=shape, small_block=False)
    try:
        a = ogs.generate(kind, ogs.DISTINCT16, n, seed=sample * 10 + shape)
        check(kind, gpu_sort(ctx, kind, a), a)
        b = ogs.generate(kind, ogs.UNIFORM, n, seed=sample * 10 + shape)
        check(kind, gpu_sort(ctx, kind, b), b)
    finally:
        restore(ctx, kind)


def test_kway_extremes(ctx):
    # maximum keys (the pad value) mixed with ordinary keys, all-equal runs
    n = 200003
    rng = np.random.default_rng(3)
    for kind in (ogs.U32, ogs.U64):
        say(f"case kway extremes kind={ogs.KIND_NAMES[kind]}")
        configure(ctx, kind, 16)
        try:
            dt = np.uint32 if kind == ogs.U32 else np.uint64
            mx = np.iinfo(dt).max
            a = rng.choice(np.array([0, 1, mx - 1, mx], dtype=dt), n)
            check(kind, gpu_sort(ctx, kind, a), a)
            a = np.full(n, mx, dtype=dt)
            check(kind, gpu_sort(ctx, kind, a), a)
        finally:
            restore(ctx, kind)


def test_kway_keep_input_and_pass_count(ctx):
    import torch
    n = (1 << 20) + 3
    configure(ctx, ogs.U32, 16, small_block=False)
    try:
        a = ogs.generate(ogs.U32, ogs.UNIFORM, n, seed=9)
        d_in = ogs.to_device(a)
        d0, d1 = torch.empty_like(d_in), torch.empty_like(d_in)
        res = ctx.sort_keep_ptr(ogs.U32, d_in.data_ptr(), d0.data_ptr(), d1.data_ptr(), n)
        torch.cuda.synchronize()
        assert res == d0.data_ptr()  # the result lands in the data buffer
        check(ogs.U32, ogs.from_device(d0, ogs.U32, n), a)
        assert np.array_equal(ogs.from_device(d_in, ogs.U32, n), a)
        T = 512 * 32
        levels = int(np.ceil(np.log2(n / T)))
        assert ctx.pass_count(ogs.U32, n) == 1 + (levels + 3) // 4
    finally:
        restore(ctx, ogs.U32)


def test_kway_large_checksums(ctx):
    import torch
    n = (1 << 27) + 12345
    say(f"case kway u32 large n={n}")
    configure(ctx, ogs.U32, 16, small_block=False)
    try:
        g = torch.Generator(device="cuda").manual_seed(5)
        d = torch.randint(-2**31, 2**31, (n,), dtype=torch.int32, device="cuda", generator=g)
        # an independent sort of the same keys (the unsigned order as int64)
        ref = torch.sort(d.to(torch.int64) & 0xFFFFFFFF).values
        s = torch.empty_like(d)
        out = ctx.sort_tensor(ogs.U32, d, s, n)
        torch.cuda.synchronize()
        o = out.view(torch.int32)[:n].to(torch.int64) & 0xFFFFFFFF
        assert bool((o[1:] >= o[:-1]).all())
        assert bool(torch.equal(o, ref))
    finally:
        restore(ctx, ogs.U32)


KV = (ogs.KV_U64U64, ogs.KV_I64I64)


def check_kv(kind, got, a):
    """Key sequence bit-exact; per key the same multiset of values (Origami is
    not stable)."""
    ref = oracle.sort(kind, a)
    assert np.array_equal(got["key"], ref["key"]), "key sequence differs"
    order_in = np.lexsort((a["value"], a["key"]))
    order_out = np.lexsort((got["value"], got["key"]))
    assert np.array_equal(a[order_in].view(np.uint8), got[order_out].view(np.uint8))


@pytest.mark.parametrize("kind", KV)
@pytest.mark.parametrize("fanin", [2, 4, 8, 16])
@pytest.mark.parametrize("dist", [ogs.UNIFORM, ogs.ZIPF, ogs.SORTED, ogs.REVERSE, ogs.DISTINCT16])
def test_kway_kv_kinds(ctx, kind, fanin, dist):
    n = 200003
    say(f"case kway kv kind={ogs.KIND_NAMES[kind]} fanin={fanin} dist={dist} n={n}")
    ctx.tune_set(kind, "plain_block", 1)  # 128 x 8 pairs: many merge levels
    ctx.tune_set(kind, "kway_fanin", fanin)
    ctx.tune_set(kind, "kway_sample", 5)
    ctx.tune_set(kind, "kway_shape", 0)
    try:
        a = ogs.generate(kind, dist, n, seed=21 + fanin)
        check_kv(kind, gpu_sort(ctx, kind, a), a)
    finally:
        restore(ctx, kind)


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("sample", [4, 6])
def test_kway_kv_shapes_and_max_keys(ctx, shape, sample):
    # maximum keys (the pad key) and ties inside and across runs, values = index
    n = 654321
    say(f"case kway kv shape={shape} sample={sample} n={n}")
    ctx.tune_set(ogs.KV_U64U64, "kway_fanin", 16)
    ctx.tune_set(ogs.KV_U64U64, "kway_sample", sample)
    ctx.tune_set(ogs.KV_U64U64, "kway_shape", shape)
    try:
        rng = np.random.default_rng(shape * 10 + sample)
        a = np.zeros(n, dtype=ogs.KV_DTYPE)
        a["key"] = rng.choice(np.array([0, 7, 2**64 - 2, 2**64 - 1], dtype=np.uint64), n)
        a["value"] = np.arange(n, dtype=np.uint64)
        check_kv(ogs.KV_U64U64, gpu_sort(ctx, ogs.KV_U64U64, a), a)
        b = ogs.generate(ogs.KV_U64U64, ogs.UNIFORM, n, seed=shape + sample)
        check_kv(ogs.KV_U64U64, gpu_sort(ctx, ogs.KV_U64U64, b), b)
    finally:
        restore(ctx, ogs.KV_U64U64)


@pytest.mark.parametrize("kind", [ogs.U32, ogs.F32, ogs.U64, ogs.KV_U64U64])
@pytest.mark.parametrize("dist", [ogs.UNIFORM, ogs.ZIPF, ogs.SORTED, ogs.REVERSE, ogs.DISTINCT16])
def test_kway_default_tiles(ctx, kind, dist):
    # the tile table's own plan (block shape, K-way shape, fan-in, sample
    # spacing) on every distribution: sorted and reverse runs do not interleave,
    # so the split kernel takes its search path, uniform keys its sample window
    n = (3 << 20) + 777 if kind != ogs.KV_U64U64 else (3 << 19) + 555
    say(f"case kway default tiles kind={ogs.KIND_NAMES[kind]} dist={dist} n={n}")
    ctx.tune_reset()
    assert ctx.tune_get(kind, "kway_fanin") > 0
    a = ogs.generate(kind, dist, n, seed=77 + dist)
    if kind == ogs.KV_U64U64:
        check_kv(kind, gpu_sort(ctx, kind, a), a)
    else:
        check(kind, gpu_sort(ctx, kind, a), a)


# Batched segments on the K-way plan: segments longer than one large-class
# tile take K-way passes of fan-in kway_fanin (the last of a segment's passes
# may be narrower), all segments of a pass in one split and one merge launch.
BATCH_KINDS = (ogs.U32, ogs.F32, ogs.U64, ogs.I64, ogs.KV_U64U64, ogs.KV_I64I64, ogs.KV_U32U32)
TILE2048 = {4: 10, 8: 1, 16: 0}  # a 2048-item block shape per item width (256 x 8)


def batch_case(kind, dist, lens, seed):
    import torch
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n = int(off[-1])
    d0 = ogs.NORMAL if kind == ogs.F32 and dist == ogs.UNIFORM else dist
    a = ogs.generate(kind, d0, n, seed=seed)
    return a, off, torch.from_numpy(off.view(np.int64).copy()).cuda()


def check_segments(kind, got, a, off):
    ref = oracle.segmented_sort(kind, a, off, threads=4)
    if kind in (ogs.KV_U64U64, ogs.KV_I64I64, ogs.KV_U32U32):
        assert np.array_equal(got["key"], ref["key"]), "key sequence differs"
        seg = np.repeat(np.arange(off.size - 1), np.diff(off).astype(np.int64))
        oi = np.lexsort((a["value"], a["key"], seg))
        oo = np.lexsort((got["value"], got["key"], seg))
        assert np.array_equal(a[oi].view(np.uint8), got[oo].view(np.uint8)), "values moved across keys"
    else:
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("kind", BATCH_KINDS)
@pytest.mark.parametrize("fanin", [4, 16])
@pytest.mark.parametrize("dist", [ogs.UNIFORM, ogs.ZIPF, ogs.SORTED, ogs.REVERSE, ogs.DISTINCT16])
def test_batched_kway(ctx, kind, fanin, dist):
    # 2048-item large-class tiles: with fan-in 4 a segment of 2^17 + 5 items
    # takes four K-way passes, the last of fan-in 2
    import torch
    rng = np.random.default_rng(fanin * 7 + dist)
    lens = np.floor(2.0 ** (rng.random(300) * 17)).astype(np.int64)
    lens = np.concatenate([lens, [2049, 4096, 8192, 8193, 1 << 15, (1 << 15) + 3, 1 << 17, (1 << 17) + 5,
                                  300001, 0, 1]])
    rng.shuffle(lens)
    a, off, doff = batch_case(kind, dist, lens, seed=fanin + 10 * dist)
    n = a.size
    say(f"case batched kway kind={ogs.KIND_NAMES[kind]} fanin={fanin} dist={dist} n={n} segs={lens.size}")
    try:
        ctx.tune_set(kind, "seg_large_block", TILE2048[ogs.WIDTH[kind]])
        ctx.tune_set(kind, "kway_fanin", fanin)
        ctx.tune_set(kind, "kway_sample", 5)
        d = ogs.to_device(a)
        s = torch.empty_like(d)
        p = ctx.segmented_ptr(kind, d.data_ptr(), s.data_ptr(), doff.data_ptr(), lens.size, n)
        torch.cuda.synchronize()
        got = ogs.from_device(d if p == d.data_ptr() else s, kind, n)
    finally:
        ctx.tune_reset()
    check_segments(kind, got, a, off)


@pytest.mark.parametrize("kind", [ogs.U32, ogs.I64, ogs.KV_U64U64, ogs.KV_U32U32])
def test_batched_kway_default_tiles(ctx, kind):
    # the tile table's own large-class plan, segments of up to 2^22 items,
    # eager and as a replayed graph
    import torch
    ctx.tune_reset()
    rng = np.random.default_rng(kind)
    lens = np.floor(2.0 ** (8 + rng.random(60) * 14)).astype(np.int64)
    lens = np.concatenate([lens, [1 << 22, (1 << 20) + 1, 65536, 65537]])
    a, off, doff = batch_case(kind, ogs.ZIPF, lens, seed=kind)
    n = a.size
    say(f"case batched kway default tiles kind={ogs.KIND_NAMES[kind]} n={n}")
    d = ogs.to_device(a)
    s = torch.empty_like(d)
    p = ctx.segmented_ptr(kind, d.data_ptr(), s.data_ptr(), doff.data_ptr(), lens.size, n)
    torch.cuda.synchronize()
    got = ogs.from_device(d if p == d.data_ptr() else s, kind, n)
    check_segments(kind, got, a, off)
    plan, res = ctx.segmented_plan(kind, d.data_ptr(), s.data_ptr(), doff.data_ptr(), lens.size, n)
    try:
        for rep in range(2):
            d.copy_(ogs.to_device(a))
            torch.cuda.synchronize()
            plan.launch()
            torch.cuda.synchronize()
            g2 = ogs.from_device(d if res == d.data_ptr() else s, kind, n)
            if kind in (ogs.KV_U64U64, ogs.KV_U32U32):
                assert np.array_equal(g2["key"], got["key"])
            else:
                assert np.array_equal(g2.view(np.uint8), got.view(np.uint8))
    finally:
        plan.close()


@pytest.mark.parametrize("kind", [ogs.U32, ogs.F32, ogs.U64, ogs.I64, ogs.KV_U64U64, ogs.KV_U32U32])
@pytest.mark.parametrize("swhole", [0, 1])
def test_kway_sample_merge_forms(ctx, kind, swhole):
    # a pass's sample merge as one K-way launch (a group's samples in one tile)
    # or as 2-way merges / a block sort: the same sort either way
    try:
        ctx.tune_set(kind, "kway_swhole", swhole)
        for n, dist in (((1 << 21) + 77, ogs.ZIPF), ((3 << 20) + 5, ogs.UNIFORM), (70001, ogs.DISTINCT16)):
            say(f"case kway sample merge kind={ogs.KIND_NAMES[kind]} swhole={swhole} n={n}")
            a = ogs.generate(kind, dist, n, seed=n + swhole)
            if kind in (ogs.KV_U64U64, ogs.KV_U32U32):
                check_kv(kind, gpu_sort(ctx, kind, a), a)
            else:
                check(kind, gpu_sort(ctx, kind, a), a)
    finally:
        ctx.tune_reset()


@pytest.mark.parametrize("kind", [ogs.U32, ogs.U64, ogs.KV_U64U64, ogs.KV_U32U32])
@pytest.mark.parametrize("swhole", [0, 1])
def test_batched_kway_sample_merge_forms(ctx, kind, swhole):
    # batched segments with the samples of each pass merged by one K-way launch
    # per fan-in (swhole 1) or by 2-way merges / a block sort (swhole 0)
    import torch
    rng = np.random.default_rng(40 + swhole)
    lens = np.floor(2.0 ** (rng.random(120) * 18)).astype(np.int64)
    lens = np.concatenate([lens, [1 << 18, (1 << 20) + 3, 2049, 65536]])
    a, off, doff = batch_case(kind, ogs.ZIPF, lens, seed=kind + swhole)
    n = a.size
    say(f"case batched kway sample merge kind={ogs.KIND_NAMES[kind]} swhole={swhole} n={n}")
    try:
        ctx.tune_set(kind, "kway_swhole", swhole)
        d = ogs.to_device(a)
        s = torch.empty_like(d)
        p = ctx.segmented_ptr(kind, d.data_ptr(), s.data_ptr(), doff.data_ptr(), lens.size, n)
        torch.cuda.synchronize()
        got = ogs.from_device(d if p == d.data_ptr() else s, kind, n)
    finally:
        ctx.tune_reset()
    check_segments(kind, got, a, off)
