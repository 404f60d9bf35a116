"""ogs_sort_noalloc (the sort with its workspace in the library's static device
array) against ogs_sort through the C-ABI, on the same seeded inputs: two
calls in a row per key kind, the second smaller; the outputs must be the same
bytes (the two entry points run the same kernels with the same tiles, so even
the order of equal-key values is the same).  Sizes beyond the array are
refused before anything runs, and a fresh context sorting through this entry
point leaves the device's free memory unchanged."""
import ctypes
import sys

import numpy as np
import pytest

import ogs

pytestmark = pytest.mark.gpu


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


def run(ctx, kind, a, noalloc):
    import torch
    n = a.size
    d = ogs.to_device(a)
    s = torch.empty_like(d)
    if noalloc:
        p = ctx.sort_noalloc_ptr(kind, d.data_ptr(), s.data_ptr(), n)
    else:
        p = ctx.sort_ptr(kind, d.data_ptr(), s.data_ptr(), n)
    torch.cuda.synchronize()
    return ogs.from_device(d if p == d.data_ptr() else s, kind, n)


# (kind, first n, second n): the first near each kind's largest size
CASES = [(ogs.U32, (1 << 26) - 1000, 100003), (ogs.F32, (3 << 23) + 7, 65536),
         (ogs.U64, (1 << 24) - 77, 5000), (ogs.KV_U64U64, (1 << 23) - 5, 300007),
         (ogs.KV_U32U32, (1 << 22) + 9, 4097), (ogs.I32, 1 << 20, 1000)]


@pytest.mark.parametrize("kind,n1,n2", CASES)
def test_noalloc_matches_sort(ctx, kind, n1, n2):
    for call, n in enumerate((n1, n2)):
        say(f"case noalloc kind={ogs.KIND_NAMES[kind]} call={call} n={n}")
        dist = ogs.NORMAL if kind == ogs.F32 else (ogs.ZIPF if call else ogs.UNIFORM)
        a = ogs.generate(kind, dist, n, seed=1000 + 17 * call + kind)
        got = run(ctx, kind, a, True)
        ref = run(ctx, kind, a, False)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
        keys = got["key"] if got.dtype.names else got
        if kind in (ogs.U32, ogs.U64, ogs.KV_U64U64, ogs.KV_U32U32):
            assert np.all(keys[1:] >= keys[:-1])


@pytest.mark.parametrize("kind,n", [(ogs.U32, 1 << 28), (ogs.U64, 1 << 26), (ogs.KV_U64U64, 1 << 25)])
def test_noalloc_refuses_sizes_beyond_the_scratch(ctx, kind, n):
    # no kernel runs: the size check comes first (the pointers are never read)
    say(f"case noalloc refuse kind={ogs.KIND_NAMES[kind]} n={n}")
    out = ctypes.c_void_p()
    rc = ogs.lib().ogs_sort_noalloc(ctx.h, kind, ctypes.c_void_p(1 << 20), ctypes.c_void_p(2 << 20), n,
                                    ctypes.byref(out), None)
    assert rc == -1
    assert b"scratch" in ogs.lib().ogs_last_error()


def test_noalloc_allocates_nothing():
    import torch
    say("case noalloc free memory")
    kind, n = ogs.U32, (1 << 24) + 3
    a = ogs.generate(kind, ogs.UNIFORM, n, seed=5)
    d = ogs.to_device(a)
    s = torch.empty_like(d)
    c = ogs.Context(0)  # fresh: its per-stream workspace pool is empty
    try:
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for _ in range(3):
            p = c.sort_noalloc_ptr(kind, d.data_ptr(), s.data_ptr(), n)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 == free0, (free0, free1)
        got = ogs.from_device(d if p == d.data_ptr() else s, kind, n)
        assert np.array_equal(got, np.sort(a))
    finally:
        c.close()
