"""Pins the oracle (oracle/origami_oracle.c) before anything is checked against it.

The reference ships no golden vectors or known-answer tests (its checks are
std::sort comparisons in Origami/bench_sort.cpp:68-78 and utils.h:38-58), so
every oracle function is pinned here by what the mathematics fixes: numpy's
sort of the same bytes, exhaustive 0-1-principle runs of every network,
brute-force co-ranks, and the C++ standard's mt19937 known answers.
"""
import itertools

import numpy as np
import pytest

import oracle


def ordered(kind, a):
    """Total-order key of each element as an unsigned integer (numpy side)."""
    if kind == oracle.F32:
        b = a.view(np.uint32)
        return np.where(b >> 31 == 1, ~b, b | np.uint32(0x80000000))
    if kind == oracle.I32:
        return a.view(np.uint32) ^ np.uint32(0x80000000)
    if kind == oracle.I64:
        return a.view(np.uint64) ^ np.uint64(1 << 63)
    if kind in (oracle.KV_U64U64, oracle.KV_U32U32):
        return a["key"]
    if kind == oracle.KV_I64I64:
        return a["key"].view(np.uint64) ^ np.uint64(1 << 63)
    return a


def numpy_sorted(kind, a):
    if kind in (oracle.KV_U64U64, oracle.KV_I64I64):
        return a[np.argsort(ordered(kind, a), kind="stable")]
    return a[np.argsort(ordered(kind, a), kind="stable")]


def test_mt19937_known_answers():
    # [rand.predef]: the 10000th output of a default-constructed engine
    assert oracle.mt19937(5489, 10000)[-1] == 4123659995
    assert oracle.mt19937_64(5489, 10000)[-1] == 9981545732273789042


def test_mt19937_matches_numpy_legacy_seeding():
    # numpy's RandomState(seed) uses the same init_genrand seeding as std::mt19937(seed)
    rs = np.random.RandomState(42)
    ref = rs.randint(0, 2**32, size=5000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(oracle.mt19937(42, 5000), ref)


@pytest.mark.parametrize("which,width", [(8, 8), (16, 16)])
def test_sorting_networks_zero_one_principle(which, width):
    # 0-1 principle: a comparator network sorts every input iff it sorts all 2^w 0-1 inputs
    vals = np.arange(2 ** width, dtype=np.uint64)
    bits = ((vals[:, None] >> np.arange(width, dtype=np.uint64)) & 1).astype(np.uint32)
    out = oracle.network_u32(which, bits.reshape(-1)).reshape(-1, width)
    assert np.all(out[:, :-1] <= out[:, 1:])
    assert np.array_equal(out.sum(1), bits.sum(1))


def test_merge_network_loop_keeps_low_block_exact():
    # mergebl2 nreg=4 loop network: inputs a0..a3 sorted, a4..a7 sorted -> outputs 0..3 are the
    # 4 smallest in order (0-1 principle over all sorted 0-1 pairs), the tail form sorts all 8
    outs = []
    for za in range(5):
        for zb in range(5):
            a = [0] * za + [1] * (4 - za)
            b = [0] * zb + [1] * (4 - zb)
            outs.append(a + b)
    g = np.array(outs, dtype=np.uint32)
    lo = oracle.network_u32(4, g.reshape(-1)).reshape(-1, 8)
    zeros = 8 - g.sum(1).astype(np.int64)
    assert np.all(lo[:, :3] <= lo[:, 1:4])
    assert np.array_equal(lo[:, :4].sum(1), np.maximum(0, 4 - zeros))  # the 4 smallest
    full = oracle.network_u32(5, g.reshape(-1)).reshape(-1, 8)
    assert np.all(full[:, :-1] <= full[:, 1:])


@pytest.mark.parametrize("nreg", [1, 4])
def test_merge_matches_sorted_concat(nreg):
    rng = np.random.default_rng(7)
    for _ in range(400):
        la, lb = 4 * int(rng.integers(1, 40)), 4 * int(rng.integers(1, 40))
        hi = int(rng.choice([3, 1000, 2**32]))
        A = np.sort(rng.integers(0, hi, la, dtype=np.uint64).astype(np.uint32))
        B = np.sort(rng.integers(0, hi, lb, dtype=np.uint64).astype(np.uint32))
        C = oracle.merge(oracle.U32, nreg, A, B)
        assert np.array_equal(C, np.sort(np.concatenate([A, B])))


def test_merge_plants_detected():
    # a merge that drops the tie rule or an element must fail this comparison
    A = np.array([1, 3, 5, 7], np.uint32)
    B = np.array([2, 3, 6, 8], np.uint32)
    C = oracle.merge(oracle.U32, 4, A, B)
    assert C.tolist() == [1, 2, 3, 3, 5, 6, 7, 8]


@pytest.mark.parametrize("K", [1, 2, 3, 4, 7, 16])
def test_kway_merge_tree(K):
    rng = np.random.default_rng(K)
    runs = [np.sort(rng.integers(0, 50, int(rng.integers(0, 300)), dtype=np.uint64)) for _ in range(K)]
    out = oracle.kway_merge(oracle.U64, runs, bufn=16)
    assert np.array_equal(out, np.sort(np.concatenate(runs)))


def test_partition_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(300):
        K = int(rng.integers(1, 6))
        hi = int(rng.choice([2, 5, 1000]))
        runs = [np.sort(rng.integers(0, hi, int(rng.integers(0, 12)), dtype=np.uint64).astype(np.uint32)) for _ in range(K)]
        total = sum(r.size for r in runs)
        for rank in range(total + 1):
            s = oracle.partition(oracle.U32, runs, rank)
            assert int(s.sum()) == rank
            left = [r[: int(x)] for r, x in zip(runs, s)]
            right = [r[int(x):] for r, x in zip(runs, s)]
            lmax = max((int(l[-1]) for l in left if l.size), default=-1)
            rmin = min((int(r[0]) for r in right if r.size), default=1 << 40)
            assert lmax <= rmin


def test_partition_splits_are_monotone_under_ties():
    # the threads of one merge group cut their slices at the splits of
    # consecutive ranks, so every run's split must be non-decreasing in rank
    rng = np.random.default_rng(1)
    for _ in range(300):
        K = int(rng.integers(2, 6))
        runs = [np.sort(rng.integers(0, 3, int(rng.integers(1, 50)))).astype(np.uint32) for _ in range(K)]
        prev = None
        for rank in range(sum(r.size for r in runs) + 1):
            s = oracle.partition(oracle.U32, runs, rank).astype(np.int64)
            assert prev is None or np.all(s >= prev), (rank, prev, s)
            prev = s


@pytest.mark.parametrize("dist", [oracle.SAME, oracle.PARETO_BURSTS, oracle.PARETO])
def test_threaded_sort_of_bursty_input(dist):
    a = oracle.generate(oracle.U32, dist, (1 << 20) + 3, seed=42)
    for threads, min_k in ((8, 2), (4, 4)):
        assert np.array_equal(oracle.sort(oracle.U32, a, threads=threads, min_k=min_k), np.sort(a))


KINDS = [oracle.U32, oracle.I32, oracle.F32, oracle.U64, oracle.I64, oracle.KV_U64U64, oracle.KV_I64I64,
         oracle.KV_U32U32]
KV_KINDS = (oracle.KV_U64U64, oracle.KV_I64I64, oracle.KV_U32U32)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 1000, (1 << 17) + 5, 300001])
@pytest.mark.parametrize("threads,min_k", [(1, 2), (1, 4), (4, 4)])
def test_sort_matches_numpy(kind, n, threads, min_k):
    dist = oracle.NORMAL if kind == oracle.F32 else oracle.UNIFORM
    a = oracle.generate(kind, dist, n, seed=42 + n)
    if kind in (oracle.I32, oracle.I64):
        a = a.view(a.dtype)  # full range: negative keys present
    s = oracle.sort(kind, a, threads=threads, min_k=min_k)
    ref = numpy_sorted(kind, a)
    if kind in KV_KINDS:
        assert np.array_equal(s["key"], ref["key"])
        # values = original index: a permutation, each pointing at its own key
        assert np.array_equal(np.sort(s["value"]), np.arange(n, dtype=s["value"].dtype))
        assert np.array_equal(a["key"][s["value"].astype(np.int64)], s["key"])
    else:
        assert np.array_equal(s.view(np.uint8), ref.view(np.uint8))


def test_kv32_sorts_on_the_key_alone():
    # KeyValue<ui, ui> compares keys only (Origami/commons.h:55-57): with the
    # values in descending order under every key, a key-only sort must still
    # put the keys in order and keep each key's values (a sort on the whole
    # 8-byte record would also pass; this pins the key's position in the record)
    n = 50000
    a = np.empty(n, dtype=oracle.DTYPES[oracle.KV_U32U32])
    rng = np.random.default_rng(3)
    a["key"] = rng.integers(0, 300, n, dtype=np.uint32)
    a["value"] = np.arange(n, 0, -1, dtype=np.uint32)
    s = oracle.sort(oracle.KV_U32U32, a, threads=4, min_k=4)
    assert np.array_equal(s["key"], np.sort(a["key"]))
    order_in = np.lexsort((a["value"], a["key"]))
    order_out = np.lexsort((s["value"], s["key"]))
    assert np.array_equal(a[order_in].view(np.uint8), s[order_out].view(np.uint8))
    # the key sits in the low four bytes: sorting the raw 8-byte words as u64
    # would order by the value instead
    assert not np.array_equal(np.sort(a.view(np.uint64)).view(a.dtype)["key"], s["key"])


@pytest.mark.parametrize("dist", [oracle.ZIPF, oracle.SORTED, oracle.REVERSE, oracle.DISTINCT16])
def test_sort_kv_distributions(dist):
    n = 200003
    a = oracle.generate(oracle.KV_U64U64, dist, n)
    s = oracle.sort(oracle.KV_U64U64, a, threads=4, min_k=4)
    assert np.array_equal(s["key"], np.sort(a["key"]))
    assert np.array_equal(np.sort(s["value"]), np.arange(n, dtype=np.uint64))
    assert np.array_equal(a["key"][s["value"].astype(np.int64)], s["key"])


def test_f32_total_order_zero_signs():
    a = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 0.0, -0.0], np.float32)
    s = oracle.sort(oracle.F32, a)
    bits = s.view(np.uint32).tolist()
    assert bits == [0xFF800000, 0xBF800000, 0x80000000, 0x80000000, 0, 0, 0x3F800000, 0x7F800000]


def test_chunked_sort_sort_n():
    a = oracle.generate(oracle.U32, oracle.UNIFORM, 1 << 16)
    s = oracle.sort(oracle.U32, a, sort_n=1 << 12, threads=2, min_k=4)
    ref = np.sort(a.reshape(-1, 1 << 12), axis=1).reshape(-1)
    assert np.array_equal(s, ref)


def test_segmented_sort():
    rng = np.random.default_rng(5)
    lens = rng.integers(0, 3000, 50)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    a = oracle.generate(oracle.U64, oracle.UNIFORM, int(off[-1]))
    s = oracle.segmented_sort(oracle.U64, a, off, threads=4)
    for i in range(lens.size):
        assert np.array_equal(s[off[i]:off[i + 1]], np.sort(a[off[i]:off[i + 1]]))


def test_generators_shapes():
    z = oracle.generate(oracle.U64, oracle.ZIPF, 100000)
    assert np.unique(z).size < 100000  # heavy repetition
    d = oracle.generate(oracle.U64, oracle.DISTINCT16, 10000)
    assert np.unique(d).size == 16
    s = oracle.generate(oracle.U32, oracle.SORTED, 1000)
    assert np.all(s[:-1] < s[1:])
    r = oracle.generate(oracle.U64, oracle.REVERSE, 1000)
    assert np.all(r[:-1] > r[1:])
    f = oracle.generate(oracle.F32, oracle.NORMAL, 1001)
    assert f.view(np.uint32)[0] == 0x80000000 and f.view(np.uint32)[500] == 0
    assert abs(float(f.astype(np.float64).std()) - 1.0) < 0.1
