"""The counter-based input family (the datagen::Writer analogue,
Origami/writer.h:27-242, README D1-D9) on the CPU.

The oracle's C restatement is pinned by what the mathematics fixes rather than
by its own formulas: splitmix64's published outputs, the exact Fibonacci
sequence, the Pareto law P(L <= k) = k / (k + 7) for beta = 7 (writer.h:86),
the Zipf(1.1) head frequency 1 / H(2^20, 1.1), the Irwin-Hall moments of the
normal keys, and the shape of sorted / reverse / almost-sorted / bursty input.
The host mirror in libogs.so (ogs_generate_host, the code the device generator
shares) must then equal the oracle bit for bit; the device output is compared
with the host mirror in test_gpu_generate.py.
"""
import numpy as np
import pytest

import ogs
import oracle

KINDS = range(8)
DISTS = range(12)
SIZES = [0, 1, 2, 7, 94, 95, 1000, 4099, 65537]


def gen(kind, dist, n, seed=5):
    return oracle.generate(kind, dist | oracle.GEN_COUNTER, n, seed=seed)


def keys(a):
    return a["key"] if a.dtype.names else a


def test_splitmix64_known_answers():
    # Vigna's splitmix64 seeded with 0: the state advances by the golden gamma
    # and each output is mix64(state before the step).
    g = 0x9E3779B97F4A7C15
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert [oracle.mix64((k * g) % 2**64) for k in range(3)] == want


@pytest.mark.parametrize("kind", KINDS)
def test_host_mirror_matches_oracle(kind):
    for dist in DISTS:
        for n in SIZES:
            for seed in (5, 2**63 + 11):
                a = ogs.generate(kind, dist | ogs.GEN_COUNTER, n, seed=seed)
                b = gen(kind, dist, n, seed)
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (kind, dist, n, seed)


def test_dists_6_to_11_need_no_flag():
    for dist in range(6, 12):
        a = ogs.generate(ogs.U64, dist, 1000, seed=3)
        assert np.array_equal(a, ogs.generate(ogs.U64, dist | ogs.GEN_COUNTER, 1000, seed=3))


def test_bad_dist_is_an_error():
    out = np.empty(4, dtype=np.uint64)
    import ctypes
    for d in (12 | ogs.GEN_COUNTER, -1, 12):
        assert ogs.lib().ogs_generate_host(ogs.U64, d, 1, 4, ctypes.c_void_p(out.ctypes.data)) < 0


def test_kv_values_are_the_index():
    for dist in DISTS:
        a = gen(ogs.KV_U64U64, dist, 1000)
        assert np.array_equal(a["value"], np.arange(1000, dtype=np.uint64))
        assert np.array_equal(a["key"], gen(ogs.U64, dist, 1000))


def test_fibonacci_is_the_exact_wrapping_sequence():
    # F(93) is the last Fibonacci number below 2^64; the reference's writer
    # restarts at 0, 1 when the 64-bit sum wraps (writer.h:175-220).
    F = [0, 1]
    while F[-1] + F[-2] < 2**64:
        F.append(F[-1] + F[-2])
    assert len(F) == 94
    n = 3 * 94 + 5
    want = np.array([F[i % 94] for i in range(n)], dtype=np.uint64)
    assert np.array_equal(gen(oracle.U64, oracle.FIBONACCI, n), want)
    assert np.array_equal(gen(oracle.U32, oracle.FIBONACCI, n), (want & 0xFFFFFFFF).astype(np.uint32))


@pytest.mark.parametrize("kind", [oracle.U32, oracle.U64])
def test_ordered_shapes(kind):
    n = 10007
    top = np.iinfo(oracle.DTYPES[kind]).max
    s = gen(kind, oracle.SORTED, n)
    assert s[0] == 0 and np.all(s[:-1] < s[1:]) and int(s[-1]) > top - 2 * (top // n)
    r = gen(kind, oracle.REVERSE, n)
    assert np.array_equal(r, s[::-1])
    a = gen(kind, oracle.ALMOST_SORTED, n)
    hot = np.arange(n) % 7 == 0
    assert np.all(a[hot] == top)
    assert np.array_equal(a[~hot], s[~hot])
    same = gen(kind, oracle.SAME, n)
    assert np.unique(same).size == 1
    assert gen(kind, oracle.SAME, 4, seed=6)[0] != same[0]


def test_pareto_law():
    # u uniform, L = ceil(7 (1/(1-u) - 1)): P(L <= k) = k / (k + 7), capped at 1e5
    n = 1 << 20
    p = gen(oracle.U64, oracle.PARETO, n).astype(np.int64)
    assert p.min() >= 0 and p.max() == 100000
    for k in (1, 7, 70, 700):
        frac = np.count_nonzero(p <= k) / n
        want = k / (k + 7)
        assert abs(frac - want) < 5 * np.sqrt(want * (1 - want) / n), (k, frac, want)
    capped = np.count_nonzero(p == 100000)
    assert 30 < capped < 130  # expectation 7/100007 * 2^20 = 73
    assert np.array_equal(gen(oracle.U32, oracle.PARETO, 1000), p[:1000].astype(np.uint32))


def _runs(a):
    cut = np.flatnonzero(a[1:] != a[:-1]) + 1
    edges = np.concatenate(([0], cut, [a.size]))
    return np.diff(edges), a[edges[:-1]]


def test_pareto_bursts():
    n = 1 << 20
    b = gen(oracle.U64, oracle.PARETO_BURSTS, n)
    lens, ks = _runs(b)
    assert np.unique(ks).size == ks.size  # each burst has its own key
    full = lens[:-1]  # the last burst is cut at n
    assert full.max() <= 100000
    # a burst is max(1, L) items: P(len <= k) = k / (k + 7)
    for k in (1, 7, 70):
        frac = np.count_nonzero(full <= k) / full.size
        want = k / (k + 7)
        assert abs(frac - want) < 5 * np.sqrt(want * (1 - want) / full.size) + 1e-3, (k, frac)
    s = gen(oracle.U64, oracle.PARETO_SHUFFLED, n)
    assert np.array_equal(np.sort(s), np.sort(b))
    assert not np.array_equal(s, b)
    # a uniform shuffle leaves each adjacent pair equal with probability
    # sum c (c - 1) / (n (n - 1)) over the burst sizes c
    _, counts = np.unique(b, return_counts=True)
    c = counts.astype(np.float64)
    p_eq = float((c * (c - 1)).sum() / (n * (n - 1.0)))
    eq = np.count_nonzero(s[1:] == s[:-1]) / (n - 1)
    assert abs(eq - p_eq) < 0.1 * p_eq + 1e-3, (eq, p_eq)
    assert np.array_equal(gen(oracle.U32, oracle.PARETO_SHUFFLED, 5000),
                          (gen(oracle.U64, oracle.PARETO_SHUFFLED, 5000) >> 32).astype(np.uint32))


def test_zipf_head():
    n = 1 << 20
    z = gen(oracle.U64, oracle.ZIPF, n)
    _, counts = np.unique(z, return_counts=True)
    counts = np.sort(counts)[::-1]
    H = float(np.sum(np.arange(1, (1 << 20) + 1, dtype=np.float64) ** -1.1))
    p0 = 1.0 / H
    assert abs(counts[0] / n - p0) < 5 * np.sqrt(p0 * (1 - p0) / n)
    assert abs(counts[0] / counts[1] - 2 ** 1.1) < 0.05


def test_normal_moments_and_zeros():
    n = 1 << 20
    f = gen(oracle.F32, oracle.NORMAL, n).astype(np.float64)
    u = gen(oracle.F32, oracle.NORMAL, n).view(np.uint32)
    assert u[0] == 0x80000000 and u[n // 2] == 0
    # Irwin-Hall of twelve 16-bit uniforms: mean 0, variance 1 - 2^-32,
    # values on the 2^-16 grid within [-6, 6]
    assert abs(f.mean()) < 5 / np.sqrt(n)
    assert abs(f.var() - 1.0) < 0.01
    assert np.all(np.abs(f) <= 6.0) and np.all(f * 65536 == np.round(f * 65536))
    assert abs(np.mean(np.abs(f) < 1.0) - 0.6827) < 0.005


def test_uniform_and_distinct16():
    n = 1 << 18
    u = gen(oracle.U64, oracle.UNIFORM, n)
    hist = np.bincount((u >> 60).astype(np.int64), minlength=16)
    chi2 = float(((hist - n / 16) ** 2 / (n / 16)).sum())
    assert chi2 < 50  # 15 degrees of freedom
    d = gen(oracle.U64, oracle.DISTINCT16, n)
    vals, cnt = np.unique(d, return_counts=True)
    assert vals.size == 16 and cnt.min() > n / 16 * 0.9
    assert np.array_equal(gen(oracle.U32, oracle.UNIFORM, n), (u >> 32).astype(np.uint32))
    assert not np.array_equal(gen(oracle.U64, oracle.UNIFORM, 100, seed=6), u[:100])
