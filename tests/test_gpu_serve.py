"""Serving loop: every request's result in a concurrent mix equals the oracle's
and the same request sorted alone; requests up to 2^24 keys, which run alone
as plain sorts beside the coalesced batches; host and device requests;
admission under a tight HBM budget; oversize requests are rejected with a
status."""
import sys

import numpy as np
import pytest

import ogs
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    c = ogs.Context(0)
    yield c
    c.close()


def _check(kind, got, a):
    ref = oracle.sort(kind, a, threads=8)
    if kind in (ogs.KV_U64U64, ogs.KV_I64I64):
        assert np.array_equal(got["key"], ref["key"])
        assert np.array_equal(np.sort(got["value"]), np.sort(a["value"]))
        o1 = np.lexsort((a["value"], a["key"]))
        o2 = np.lexsort((got["value"], got["key"]))
        assert np.array_equal(a[o1].view(np.uint8), got[o2].view(np.uint8))
    else:
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


def test_mixed_requests_same_alone_and_in_mix(ctx):
    import torch
    rng = np.random.default_rng(21)
    kinds = [ogs.U32, ogs.U64, ogs.KV_U64U64, ogs.F32, ogs.I64]
    reqs = []
    for j in range(60):
        kind = kinds[j % len(kinds)]
        n = int(2 ** rng.uniform(0, 20))
        dist = ogs.NORMAL if kind == ogs.F32 else (ogs.ZIPF if j % 2 else ogs.UNIFORM)
        a = ogs.generate(kind, dist, n, seed=300 + j)
        reqs.append((kind, n, a))
    say("case serve mix: 60 requests, 4 streams")
    srv = ogs.Server(ctx, hbm_budget_bytes=64 << 20, streams=4)
    bufs, tickets = [], []
    for j, (kind, n, a) in enumerate(reqs):
        src = ogs.to_device(a) if n else torch.zeros(16, dtype=torch.uint8, device="cuda")
        dst = src if j % 3 == 0 else torch.empty_like(src)
        bufs.append((src, dst))
    torch.cuda.synchronize()
    # one burst of submissions: requests queue behind the busy streams and
    # are coalesced
    for (kind, n, a), (src, dst) in zip(reqs, bufs):
        tickets.append(srv.submit(kind, src.data_ptr(), dst.data_ptr(), n))
    for t in tickets:
        srv.wait(t, 60.0)
    st = srv.stats()
    say(f"stats {st}")
    assert st["completed"] == len(reqs) and st["batches"] < len(reqs)  # requests were coalesced
    alone = []
    for j, ((kind, n, a), (src, dst)) in enumerate(zip(reqs, bufs)):
        got = ogs.from_device(dst, kind, n)
        _check(kind, got, a)
        # the same request sorted alone (plain entry point) gives the same bytes for key-only kinds
        if kind not in (ogs.KV_U64U64, ogs.KV_I64I64) and n:
            d = ogs.to_device(a)
            s = torch.empty_like(d)
            res = ctx.sort_tensor(kind, d, s, n)
            torch.cuda.synchronize()
            assert np.array_equal(ogs.from_device(res, kind, n).view(np.uint8), got.view(np.uint8))
    for t in tickets:
        srv.release(t)
    srv.close()


def test_host_requests_and_budget(ctx):
    say("case serve host requests, tight budget")
    srv = ogs.Server(ctx, hbm_budget_bytes=8 << 20, streams=2)
    arrays = [ogs.generate(ogs.U64, ogs.UNIFORM, n, seed=n) for n in (1000, 300000, 200000, 5, 400000)]
    outs = [np.empty_like(a) for a in arrays]
    tickets = [srv.submit(ogs.U64, a.ctypes.data, o.ctypes.data, a.size, host=True) for a, o in zip(arrays, outs)]
    for t in tickets:
        srv.wait(t, 60.0)
    for a, o in zip(arrays, outs):
        assert np.array_equal(o, np.sort(a))
    # a request bigger than the whole budget is rejected, not hung
    big = np.zeros(1 << 20, np.uint64)
    t = srv.submit(ogs.U64, big.ctypes.data, big.ctypes.data, big.size, host=True)
    with pytest.raises(ogs.OgsError):
        srv.wait(t, 30.0)
    srv.close()


def _serve_and_check(ctx, reqs, streams, coalesce=0, budget=None, gaps_ms=None):
    """Submits every (kind, n, host array) request (with the given gaps between
    submissions), waits, and checks each result against the oracle and against
    the same request sorted alone by the plain entry point."""
    import time

    import torch
    bufs = []
    for kind, n, a in reqs:
        src = ogs.to_device(a)
        bufs.append((src, torch.empty_like(src)))
    need = sum(n * ogs.WIDTH[k] for k, n, _ in reqs)
    srv = ogs.Server(ctx, hbm_budget_bytes=int(budget or need + (64 << 20)), streams=streams,
                     coalesce_max_items=coalesce)
    torch.cuda.synchronize()
    tickets = []
    try:
        for j, ((kind, n, a), (src, dst)) in enumerate(zip(reqs, bufs)):
            if gaps_ms is not None:
                time.sleep(gaps_ms[j] / 1e3)
            tickets.append(srv.submit(kind, src.data_ptr(), dst.data_ptr(), n))
        for t in tickets:
            srv.wait(t, 120.0)
        st = srv.stats()
    finally:
        srv.close()
    say(f"stats {st}")
    assert st["completed"] == len(reqs)
    for (kind, n, a), (src, dst) in zip(reqs, bufs):
        say(f"check kind={ogs.KIND_NAMES[kind]} n={n}")
        got = ogs.from_device(dst, kind, n)
        _check(kind, got, a)
        if kind not in (ogs.KV_U64U64, ogs.KV_I64I64):
            d = ogs.to_device(a)
            s = torch.empty_like(d)
            res = ctx.sort_tensor(kind, d, s, n)
            torch.cuda.synchronize()
            assert np.array_equal(ogs.from_device(res, kind, n).view(np.uint8), got.view(np.uint8))
    return st


def test_large_requests_beside_batches(ctx):
    # requests above the coalescing limit (2^22) up to 2^24 keys run alone as
    # plain sorts on their own streams while the short ones are coalesced
    rng = np.random.default_rng(33)
    kinds = [ogs.U32, ogs.U64, ogs.KV_U64U64, ogs.F32, ogs.I64]
    sizes = [1 << 24, (1 << 23) + 5, (1 << 22) + 7, 1 << 22, (1 << 22) + 1, 3000001]
    sizes += [int(2 ** rng.uniform(10, 20)) for _ in range(30)]
    reqs = []
    for j, n in enumerate(sizes):
        kind = kinds[j % len(kinds)] if n < (1 << 24) else ogs.U32
        if kind == ogs.KV_U64U64 and n > (1 << 23):
            kind = ogs.U64
        dist = ogs.NORMAL if kind == ogs.F32 else (ogs.ZIPF if j % 2 else ogs.UNIFORM)
        reqs.append((kind, n, ogs.generate(kind, dist, n, seed=700 + j)))
    order = rng.permutation(len(reqs))
    reqs = [reqs[i] for i in order]
    say(f"case serve large: {len(reqs)} requests up to 2^24 keys, 4 streams, one burst")
    st = _serve_and_check(ctx, reqs, streams=4)
    assert st["batches"] < len(reqs)  # the short ones were coalesced
    # the same requests arriving spread out (Poisson gaps of 0.2 ms on average)
    say("case serve large: spread arrivals")
    _serve_and_check(ctx, reqs, streams=4, gaps_ms=rng.exponential(0.2, len(reqs)))


def test_every_request_alone_on_many_streams(ctx):
    # coalescing off (limit 1000 items): every longer request is its own plain
    # sort, up to 8 of them on the GPU at once
    rng = np.random.default_rng(34)
    kinds = [ogs.U32, ogs.KV_U64U64, ogs.U64, ogs.I32]
    reqs = []
    for j in range(24):
        kind = kinds[j % len(kinds)]
        n = int(2 ** rng.uniform(8, 21))
        reqs.append((kind, n, ogs.generate(kind, ogs.ZIPF if j % 3 else ogs.UNIFORM, n, seed=900 + j)))
    say("case serve alone: 24 requests, coalescing off, 8 streams")
    _serve_and_check(ctx, reqs, streams=8, coalesce=1000)
