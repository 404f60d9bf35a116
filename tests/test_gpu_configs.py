"""GPU parity at the benchmark configurations' full sizes, on the same seeded
inputs as bench.py, against the oracle's multithreaded restatement of the
reference's sort (oracle.sort, Origami/sorter.h:4194-4287, 16 host threads):

  - 2^28 uint32 keys, raw std::mt19937(42): bit-exact;
  - 2^28 float32 keys, N(0,1) with -0.0 and +0.0 present: bit-exact;
  - 2^27 {uint64 key, uint64 value = index} pairs for each of uniform, Zipf(1.1),
    sorted, reverse and 16 distinct keys: the key sequence bit-exact, and the
    values a permutation of 0..n-1 with key_in[value] == key_out at every
    position, which makes each key's value multiset the input's (the oracle's
    too, since it keeps each key's values by the same check);
  - the batched segmented sort of bench.py's batch: log-uniform lengths in
    [1, 2^20] drawn until 2^30 uniform uint32 keys, against the oracle's
    per-segment sort.
"""
import sys

import numpy as np
import pytest

import ogs
import oracle

pytestmark = pytest.mark.gpu

THREADS = 16


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


def gpu_sort_keep(ctx, kind, a):
    """ogs_sort_keep of host array a (the bench's step); returns the host result."""
    import torch
    n = a.size
    d_in = ogs.to_device(a)
    d_a = torch.empty_like(d_in)
    d_b = torch.empty_like(d_in)
    p = ctx.sort_keep_ptr(kind, d_in.data_ptr(), d_a.data_ptr(), d_b.data_ptr(), n)
    torch.cuda.synchronize()
    out = d_a if p == d_a.data_ptr() else d_b
    del d_in
    got = ogs.from_device(out, kind, n)
    del d_a, d_b, out
    torch.cuda.empty_cache()
    return got


def test_u32_2pow28_mt19937_seed42(ctx):
    n = 1 << 28
    say("case C2 u32 2^28 mt19937(42)")
    a = ogs.generate(ogs.U32, ogs.UNIFORM, n, seed=42)
    got = gpu_sort_keep(ctx, ogs.U32, a)
    ref = oracle.sort(oracle.U32, a, threads=THREADS, min_k=4)
    assert np.array_equal(got, ref)


def test_f32_2pow28_normal_signed_zeros(ctx):
    n = 1 << 28
    say("case C2 f32 2^28 N(0,1) with +-0")
    a = ogs.generate(ogs.F32, ogs.NORMAL, n, seed=42)
    bits = a.view(np.uint32)
    assert bits[0] == 0x80000000 and bits[n // 2] == 0
    got = gpu_sort_keep(ctx, ogs.F32, a)
    ref = oracle.sort(oracle.F32, a, threads=THREADS, min_k=4)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("dist", [ogs.UNIFORM, ogs.ZIPF, ogs.SORTED, ogs.REVERSE, ogs.DISTINCT16])
def test_kv_2pow27(ctx, dist):
    n = 1 << 27
    say(f"case C3 KV u64+u64 2^27 dist={dist}")
    a = ogs.generate(ogs.KV_U64U64, dist, n, seed=42)
    got = gpu_sort_keep(ctx, ogs.KV_U64U64, a)
    ref = oracle.sort(oracle.KV_U64U64, a, threads=THREADS, min_k=4)
    assert np.array_equal(got["key"], ref["key"])
    v = got["value"]
    assert int(v.max()) < n
    seen = np.zeros(n, dtype=bool)
    seen[v] = True
    assert seen.all(), "values are not a permutation of the input's"
    assert np.array_equal(a["key"][v], got["key"]), "a value moved away from its key"


def batch_lengths(total_log2=30, max_log2=20, seed=4):
    """bench.py's C4 batch: log-uniform lengths until 2^total_log2 keys."""
    rng = np.random.default_rng(seed)
    cap, acc, lens = 1 << total_log2, 0, []
    while acc < cap:
        ln = min(max(1, int(2.0 ** (rng.random() * max_log2))), cap - acc)
        lens.append(ln)
        acc += ln
    return np.array(lens, dtype=np.int64)


def test_segmented_batch_2pow30(ctx):
    import torch
    lens = batch_lengths()
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n = int(off[-1])
    say(f"case C4 batched {lens.size} segments, {n} keys")
    a = ogs.generate(ogs.U32, ogs.UNIFORM, n, seed=4)
    d = ogs.to_device(a)
    s = torch.empty_like(d)
    doff = torch.from_numpy(off.view(np.int64).copy()).cuda()
    p = ctx.segmented_ptr(ogs.U32, d.data_ptr(), s.data_ptr(), doff.data_ptr(), lens.size, n)
    torch.cuda.synchronize()
    got = ogs.from_device(d if p == d.data_ptr() else s, ogs.U32, n)
    del d, s
    torch.cuda.empty_cache()
    ref = oracle.segmented_sort(oracle.U32, a, off, threads=THREADS)
    assert np.array_equal(got, ref)
