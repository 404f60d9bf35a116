"""The device generator (ogs_generate) against its host mirror
(ogs_generate_host), bit for bit, for every kind and distribution; the host
mirror is pinned to the oracle in test_generators.py."""
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


def device_generate(ctx, kind, dist, n, seed):
    import torch
    w = ogs.WIDTH[kind]
    d = torch.empty(max(n * w, 16), dtype=torch.uint8, device="cuda")
    ctx.generate_ptr(kind, dist, seed, n, d.data_ptr())
    torch.cuda.synchronize()
    return d[: n * w].cpu().numpy().view(ogs.NP_DTYPE[kind])


@pytest.mark.parametrize("kind", range(7))
def test_device_matches_host_mirror(ctx, kind):
    for dist in range(12):
        for n in (1, 2, 95, 4099, (1 << 18) + 3):
            print(f"generate kind={kind} dist={dist} n={n}", flush=True)
            got = device_generate(ctx, kind, dist | ogs.GEN_COUNTER, n, seed=n + 7)
            want = ogs.generate(kind, dist | ogs.GEN_COUNTER, n, seed=n + 7)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (kind, dist, n)


@pytest.mark.parametrize("dist", [ogs.ZIPF, ogs.PARETO_BURSTS, ogs.PARETO_SHUFFLED, ogs.NORMAL])
def test_device_large(ctx, dist):
    kind = ogs.F32 if dist == ogs.NORMAL else ogs.U64
    n = (1 << 22) + 5
    sys.stdout.flush()
    got = device_generate(ctx, kind, dist | ogs.GEN_COUNTER, n, seed=11)
    want = ogs.generate(kind, dist | ogs.GEN_COUNTER, n, seed=11)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_device_generate_then_sort(ctx):
    # the generator feeds the sort without a host round trip
    import torch
    n = 1 << 20
    d = torch.empty(n * 8, dtype=torch.uint8, device="cuda")
    s = torch.empty_like(d)
    ctx.generate_ptr(ogs.U64, ogs.PARETO_SHUFFLED, 3, n, d.data_ptr())
    res = ctx.sort_tensor(ogs.U64, d, s, n)
    torch.cuda.synchronize()
    got = ogs.from_device(res, ogs.U64, n)
    want = np.sort(ogs.generate(ogs.U64, ogs.PARETO_SHUFFLED, n, seed=3))
    assert np.array_equal(got, want)


def test_zero_items_is_a_no_op(ctx):
    ctx.generate_ptr(ogs.U32, ogs.UNIFORM | ogs.GEN_COUNTER, 1, 0, 0)
