"""Plain sorts at full and beyond-4G sizes, checked through size-independent
properties: the output is in key order (compared chunk by chunk, across chunk
edges too) and is a permutation of the input (sum and sum of squares of the
items, wrapping in 64 bits); for KV the value is the original index, so
every output pair must match the input at that index.  Inputs are generated on
the GPU (torch's generator), which is enough for properties that do not need a
reference output."""
import sys

import pytest

import ogs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    return ogs.Context(0)


CHUNK = 1 << 27


def u32_as_i64(t):
    import torch
    return t.view(torch.int32).to(torch.int64) & 0xFFFFFFFF


def check_u32(src, out, n):
    s_in = s_out = q_in = q_out = 0
    prev = None
    for o in range(0, n, CHUNK):
        a = u32_as_i64(src[o:min(n, o + CHUNK)])
        b = u32_as_i64(out[o:min(n, o + CHUNK)])
        assert bool((b[1:] >= b[:-1]).all()), f"not sorted inside chunk at {o}"
        if prev is not None:
            assert int(b[0]) >= prev, f"not sorted across chunk edge at {o}"
        prev = int(b[-1])
        s_in += int(a.sum()); s_out += int(b.sum())
        q_in += int((a * a).sum()); q_out += int((b * b).sum())
        del a, b
    assert (s_in - s_out) % (1 << 64) == 0, "sum differs"
    assert (q_in - q_out) % (1 << 64) == 0, "sum of squares differs"


# (1 << 29) + 7: the last K-way pass merges one group of more than 2 GiB (the
# per-item load path); (1 << 30) + 17: a group of more than 4 GiB (64-bit
# chunk sources); (1 << 32) + 5: beyond 32-bit item counts
@pytest.mark.parametrize("n", [(1 << 29) + 7, (1 << 30) + 17, (1 << 32) + 5])
def test_u32_large(ctx, n):
    import torch
    need = 3 * 4 * n + (2 << 30)
    free, _ = torch.cuda.mem_get_info()
    if free < need:
        pytest.skip(f"needs {need >> 30} GiB of free HBM, {free >> 30} GiB free")
    say(f"case large u32 n={n}")
    g = torch.Generator(device="cuda")
    g.manual_seed(n)
    src = torch.empty(n, dtype=torch.int32, device="cuda")
    for o in range(0, n, CHUNK):
        c = min(n, o + CHUNK) - o
        src[o:o + c] = torch.randint(-(1 << 31), 1 << 31, (c,), dtype=torch.int64, device="cuda",
                                     generator=g).to(torch.int32)
    a = torch.empty_like(src)
    b = torch.empty_like(src)
    p = ctx.sort_keep_ptr(ogs.U32, src.data_ptr(), a.data_ptr(), b.data_ptr(), n)
    torch.cuda.synchronize()
    check_u32(src, a if p == a.data_ptr() else b, n)


def test_u64_large(ctx):
    """2^28 + 3 u64 keys: the last K-way pass merges one group of more than
    2 GiB, whose chunks load item by item."""
    import torch
    n = (1 << 28) + 3
    need = 3 * 8 * n + (2 << 30)
    free, _ = torch.cuda.mem_get_info()
    if free < need:
        pytest.skip(f"needs {need >> 30} GiB of free HBM")
    say(f"case large u64 n={n}")
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    src = torch.empty(n, dtype=torch.int64, device="cuda")
    for o in range(0, n, CHUNK):
        c = min(n, o + CHUNK) - o
        src[o:o + c] = torch.randint(-(1 << 63), (1 << 63) - 1, (c,), dtype=torch.int64, device="cuda", generator=g)
    a = torch.empty_like(src)
    b = torch.empty_like(src)
    p = ctx.sort_keep_ptr(ogs.U64, src.data_ptr(), a.data_ptr(), b.data_ptr(), n)
    torch.cuda.synchronize()
    out = a if p == a.data_ptr() else b
    flip = -(1 << 63)
    s_in = s_out = q_in = q_out = 0
    prev = None
    for o in range(0, n, CHUNK):
        x = src[o:min(n, o + CHUNK)]
        k = out[o:min(n, o + CHUNK)] ^ flip  # unsigned order as signed
        assert bool((k[1:] >= k[:-1]).all()), f"not sorted inside chunk at {o}"
        if prev is not None:
            assert int(k[0]) >= prev, f"not sorted across chunk edge at {o}"
        prev = int(k[-1])
        y = out[o:min(n, o + CHUNK)]
        # wrapping 64-bit sums of the items and of their squares
        s_in += int(x.sum()); s_out += int(y.sum())
        q_in += int((x * x).sum()); q_out += int((y * y).sum())
        del x, y, k
    assert (s_in - s_out) % (1 << 64) == 0, "sum differs"
    assert (q_in - q_out) % (1 << 64) == 0, "sum of squares differs"


def test_kv_large(ctx):
    import torch
    n = (1 << 28) + 3
    need = 3 * 16 * n + (4 << 30)
    free, _ = torch.cuda.mem_get_info()
    if free < need:
        pytest.skip(f"needs {need >> 30} GiB of free HBM")
    say(f"case large kv n={n}")
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    src = torch.empty((n, 2), dtype=torch.int64, device="cuda")
    src[:, 0] = torch.randint(-(1 << 62), 1 << 62, (n,), dtype=torch.int64, device="cuda", generator=g) * 2
    src[:, 1] = torch.arange(n, dtype=torch.int64, device="cuda")
    a = torch.empty_like(src)
    b = torch.empty_like(src)
    p = ctx.sort_keep_ptr(ogs.KV_U64U64, src.data_ptr(), a.data_ptr(), b.data_ptr(), n)
    torch.cuda.synchronize()
    out = a if p == a.data_ptr() else b
    flip = -(1 << 63)
    prev = None
    for o in range(0, n, CHUNK):
        k = out[o:min(n, o + CHUNK), 0] ^ flip  # unsigned order as signed
        assert bool((k[1:] >= k[:-1]).all()), f"keys not sorted at {o}"
        if prev is not None:
            assert int(k[0]) >= prev
        prev = int(k[-1])
        idx = out[o:min(n, o + CHUNK), 1]
        assert bool((src[idx, 0] == out[o:min(n, o + CHUNK), 0]).all()), "pair does not match its input index"
    # every index exactly once
    seen = torch.zeros(n, dtype=torch.int8, device="cuda")
    seen[out[:, 1]] = 1
    assert int(seen.sum()) == n, "values are not a permutation"
