"""CPU model of the sample-split K-way merge pass (origami-sort_amd/csrc/kway_kernels.hip).

The model follows the kernels step by step -- tagged samples every s items of each
run, the per-group sample merge, k_kway_bounds' splits (each run's sample count
from the window of merged samples after the boundary, else a binary search over
the run's samples, then a search among s items; checked against the binary
search alone), and
k_kway_merge's padded LDS layout (odd slices descending), per-thread co-ranks on
the bitonic layout and level-by-level pair merges -- and checks the invariants
the pass relies on: splits are monotone and exact (every item of a tile precedes
every item of the next in the total order (key, run, index)), every tile with
its padded slices fits the LDS tile, and the tiles laid end to end are the
merge of the group.  This is the role of Partition_Ptrs (Origami/Partition.h:
117-166) for KW runs, restated on the tiles of one pass.
"""
import numpy as np
import pytest


def kw_q(C, ipt, kw, slog2):
    q = (C - kw * ipt) // (1 << slog2) - kw
    return max(q, 0)


def group_samples(runs, slog2):
    """Tagged samples (key, run, index/s) of each run, merged in the total order."""
    s = 1 << slog2
    tags = []
    for j, X in enumerate(runs):
        for m in range(0, len(X), s):
            tags.append((int(X[m]), j, m >> slog2))
    tags.sort()
    return tags


def split(X, v, up, slog2, samples_j):
    """k_kway_bounds for one run: the items of X that precede the boundary."""
    s = 1 << slog2
    lo, hi = 0, len(samples_j)
    while lo < hi:
        mid = (lo + hi) // 2
        k = samples_j[mid]
        if (not (v < k)) if up else (k < v):
            lo = mid + 1
        else:
            hi = mid
    if lo == 0:
        return 0
    a, b = ((lo - 1) << slog2) + 1, min(lo << slog2, len(X))
    while a < b:
        mid = (a + b) // 2
        k = int(X[mid])
        if (not (v < k)) if up else (k < v):
            a = mid + 1
        else:
            b = mid
    return a


def bounds_window(kw):
    """The merged samples k_kway_bounds reads per tile: LPT lanes per tile (32
    when two tiles share a wave, fan-in <= 8; else 64), NW = max(1, 8 * kw / LPT)
    samples per lane (kw_bounds_tpw and NW in kway_kernels.hip)."""
    lpt = 32 if kw <= 8 else 64
    return lpt * max(1, 8 * kw // lpt)


def test_bounds_window_matches_kernel_constants():
    # fan-in 2 and 4 read 32 samples (one per lane of a half wave), 8 reads
    # 64, 16 reads 128, 32 reads 256
    assert [bounds_window(k) for k in (2, 4, 8, 16, 32)] == [32, 32, 64, 128, 256]


def split_window(X, v, r, j, slog2, samples_j, tags, p, kw):
    """k_kway_bounds as the kernel runs it: the count of run j's samples that
    precede the boundary (merged position p) is the index of run j's first
    sample in the window of merged samples from p; with none there, all of them
    when the window reaches the group's end, else the binary search."""
    kw_window = bounds_window(kw)
    window = tags[p:p + kw_window]
    lo = None
    for (_, rj, mj) in window:
        if rj == j:
            lo = mj
            break
    if lo is None and p + kw_window >= len(tags):
        lo = len(samples_j)
    if lo is None:
        return split(X, v, j < r, slog2, samples_j), "search"
    if lo == 0:
        return 0, "window"
    up = j < r
    a, b = ((lo - 1) << slog2) + 1, min(lo << slog2, len(X))
    while a < b:
        mid = (a + b) // 2
        k = int(X[mid])
        if (not (v < k)) if up else (k < v):
            a = mid + 1
        else:
            b = mid
    return a, "window"


def pow2_floor(x):
    return 1 << (x.bit_length() - 1) if x > 0 else 0


def corank_rev(L, a_off, la, b_top, lb, diag, top):
    """lds_corank_lift_rev: A items among the first diag outputs (ties to A);
    top is the level's largest step bound over all pairs (any top with
    2 * top > min(la, lb) gives the same answer)."""
    lo = max(0, diag - lb)
    hi = min(diag, la)
    a = lo
    s = top
    while s > 0:
        t = min(a + s, hi)
        ka = L[a_off + t - 1]
        kb = L[b_top - diag + t]
        a = a if kb < ka else t
        s >>= 1
    return a


def merge_rev(L, ia, ib, ipt):
    """lds_merge_rev: A moving up from ia, B moving down from ib."""
    out = []
    ka, kb = L[ia], L[ib]
    for _ in range(ipt):
        take_a = not (kb < ka)
        out.append(ka if take_a else kb)
        if take_a:
            ia += 1
            ka = L[ia]
        else:
            ib -= 1
            kb = L[ib]
    return out


KMAX = 2**32 - 1


def corank_split_rev(L, spF, spL, ipt, a_off, la, b_off, lb, b_top, diag, top):
    """lds_corank_split_rev: lifting steps of ipt items and more on the block
    splitters (A[t-1] = last item of A's block t/ipt - 1 = spL[a_off/ipt + t/ipt - 1],
    B[diag-t] = first item of B's block (diag-t)/ipt = spF[(b_off + diag - t)/ipt]),
    then the last log2(ipt) steps on the items."""
    lo = max(0, diag - lb)
    hi = min(diag, la)
    ab = lo
    s = top
    while s >= ipt:
        t = min(ab + s, hi)
        ka = spL[a_off // ipt + t // ipt - 1]
        kb = spF[(b_off + diag - t) // ipt]
        ab = ab if kb < ka else t
        s >>= 1
    s = ipt // 2
    while s > 0:
        t = min(ab + s, hi)
        ka = L[a_off + t - 1]
        kb = L[b_top - diag + t]
        ab = ab if kb < ka else t
        s >>= 1
    return ab


def model_tile(runs, lo, hi, block, ipt):
    """k_kway_merge on one tile: returns the tile's output (first cnt items)."""
    kw = len(runs)
    C = block * ipt
    n = [hi[j] - lo[j] for j in range(kw)]
    assert all(x >= 0 for x in n)
    P = [0]
    for j in range(kw):
        P.append(P[-1] + (n[j] + ipt - 1) // ipt * ipt)
    Np = P[-1]
    assert Np <= C, (Np, C)
    # LDS starts as garbage (guard reads such as A[-1] happen only where they
    # cannot change the result; the output check catches any that would)
    L = [int(x) for x in np.random.default_rng(Np).integers(0, 2**32, C + 2 * ipt + 2)]
    # load phase: chunk table per thread, then every LDS position e
    cb, cm = {}, {}
    for tid in range(block):
        c0 = tid * ipt
        if c0 >= Np:
            continue
        j0 = sum(1 for j in range(1, kw) if P[j] <= c0)
        pj, npj, nj = P[j0], P[j0 + 1] - P[j0], n[j0]
        q = c0 - pj
        desc = j0 & 1
        if desc:
            base = lo[j0] + (npj - 1 - q)
            v0, v1 = min(max(npj - q - nj, 0), ipt), ipt
        else:
            base = lo[j0] + q
            v0, v1 = 0, min(max(nj - q, 0), ipt)
        cb[tid] = (j0, base, desc)
        cm[tid] = (v0, v1)
    for e in range(Np):
        c, r = e // ipt, e % ipt
        j0, base, desc = cb[c]
        v0, v1 = cm[c]
        if v0 <= r < v1:
            L[e] = int(runs[j0][base - r if desc else base + r])
        else:
            L[e] = KMAX
    # the slices' block splitters, as the kernel writes them after the loads;
    # entries no thread wrote (the guards at -1 and block, blocks past the tile's
    # items) read as garbage, as LDS would
    class Garbage(dict):
        def __missing__(self, k):
            return (k * 2654435761 + Np) % 2**32
    spF, spL = Garbage(), Garbage()
    for tid in range(block):
        c0 = tid * ipt
        if c0 >= Np:
            continue
        j0 = sum(1 for j in range(1, kw) if P[j] <= c0)
        first, last = L[c0], L[c0 + ipt - 1]
        if j0 & 1:
            idx = (P[j0] + P[j0 + 1] - c0) // ipt - 1
            spF[idx], spL[idx] = last, first
        else:
            spF[tid], spL[tid] = first, last
    logkw = kw.bit_length() - 1
    for lev in range(1, logkw + 1):
        regs = {}
        hv = 1 << (lev - 1)
        top = pow2_floor(max(min(P[s0 + hv] - P[s0], P[s0 + 2 * hv] - P[s0 + hv])
                             for s0 in range(0, kw, 2 * hv)))
        for tid in range(block):
            c0 = tid * ipt
            if c0 >= Np:
                continue
            j0 = sum(1 for j in range(1, kw) if P[j] <= c0)
            hv = 1 << (lev - 1)
            s0 = (j0 >> lev) << lev
            a_off, mid, end = P[s0], P[s0 + hv], P[s0 + 2 * hv]
            la, lb, diag = mid - a_off, end - mid, c0 - a_off
            a = corank_split_rev(L, spF, spL, ipt, a_off, la, mid, lb, end - 1, diag, top)
            assert a == corank_rev(L, a_off, la, end - 1, lb, diag, top), (lev, tid)
            regs[tid] = (merge_rev(L, a_off + a, end - 1 - (diag - a), ipt), j0, end, diag)
        for tid, (r, j0, end, diag) in regs.items():
            c0 = tid * ipt
            if lev < logkw and ((j0 >> lev) & 1):
                top = end - 1 - diag
                for i in range(ipt):
                    L[top - i] = r[i]
            else:
                for i in range(ipt):
                    L[c0 + i] = r[i]
            if lev < logkw:
                spF[tid], spL[tid] = r[0], r[ipt - 1]
    cnt = sum(n)
    return L[:cnt]


PATHS = {}  # how the model's splits were found: by the window or the fallback search


def model_pass(keys, R, kw, slog2, block, ipt):
    """One K-way pass over keys (runs of R sorted): returns the output."""
    nitems = len(keys)
    C = block * ipt
    Q = kw_q(C, ipt, kw, slog2)
    assert Q > 0
    out = np.empty_like(keys)
    for gbase in range(0, nitems, kw * R):
        runs = [keys[gbase + j * R: min(gbase + (j + 1) * R, nitems)] if gbase + j * R < nitems
                else keys[:0] for j in range(kw)]
        for X in runs:
            assert np.all(X[1:] >= X[:-1])
        tags = group_samples(runs, slog2)
        ns = len(tags)
        nt = (ns + Q - 1) // Q
        smp = [[int(X[m]) for m in range(0, len(X), 1 << slog2)] for X in runs]
        bounds = []
        for tl in range(nt):
            if tl == 0:
                bounds.append([0] * kw)
                continue
            v, r, m = tags[tl * Q]
            row = []
            for j in range(kw):
                if len(runs[j]) == 0:
                    row.append(0)
                elif j == r:
                    row.append(m << slog2)
                else:
                    ref = split(runs[j], v, j < r, slog2, smp[j])
                    got, how = split_window(runs[j], v, r, j, slog2, smp[j], tags, tl * Q, kw)
                    assert got == ref, (tl, j, got, ref, how)
                    PATHS[how] = PATHS.get(how, 0) + 1
                    row.append(got)
            bounds.append(row)
        bounds.append([len(X) for X in runs])
        o = 0
        for tl in range(nt):
            lo, hi = bounds[tl], bounds[tl + 1]
            assert all(lo[j] <= hi[j] for j in range(kw)), (lo, hi)
            # exactness: every item of the tile precedes every item after it
            if tl + 1 < nt:
                pre = [(int(runs[j][i]), j, i) for j in range(kw) for i in range(lo[j], hi[j])]
                post = [(int(runs[j][hi[j]]), j, hi[j]) for j in range(kw) if hi[j] < len(runs[j])]
                if pre and post:
                    assert max(pre) < min(post)
            size = sum(hi[j] - lo[j] for j in range(kw))
            assert size <= (Q + kw) * (1 << slog2) - kw
            assert sum(lo) == o
            tile = model_tile(runs, lo, hi, block, ipt)
            out[gbase + o: gbase + o + size] = tile
            o += size
        assert o == sum(len(X) for X in runs)
    return out


def runs_of(keys, R):
    k = keys.copy()
    for b in range(0, len(k), R):
        k[b:b + R] = np.sort(k[b:b + R])
    return k


@pytest.mark.parametrize("kw", [2, 4, 8, 16, 32])
@pytest.mark.parametrize("dist", ["uniform", "few", "sorted", "reverse"])
def test_kway_pass_model(kw, dist):
    rng = np.random.default_rng(kw * 7 + len(dist))
    # 32 padded slices need a larger tile; shorter runs keep the model quick
    block, ipt, slog2 = (64, 8, 3) if kw < 32 else (128, 8, 3)
    R = 1024 if kw < 32 else 256
    n = R * kw * 2 + 3 * R + 77  # two full groups and a ragged one
    if dist == "uniform":
        keys = rng.integers(0, 2**32, n, dtype=np.uint64)
    elif dist == "few":
        keys = rng.integers(0, 4, n, dtype=np.uint64) * 1000
    elif dist == "sorted":
        keys = np.arange(n, dtype=np.uint64) // 3
    else:
        keys = (n - np.arange(n, dtype=np.uint64)) // 5
    keys = runs_of(keys, R)
    out = model_pass(keys, R, kw, slog2, block, ipt)
    ref = np.concatenate([np.sort(keys[b:b + kw * R]) for b in range(0, n, kw * R)])
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("kw,ipt,slog2", [(16, 8, 2), (4, 16, 4), (8, 4, 3)])
def test_kway_pass_model_shapes(kw, ipt, slog2):
    rng = np.random.default_rng(5)
    block = 1024 // ipt
    R = 512
    n = R * kw + R // 2 + 5
    keys = runs_of(rng.integers(0, 50, n, dtype=np.uint64), R)
    out = model_pass(keys, R, kw, slog2, block, ipt)
    ref = np.concatenate([np.sort(keys[b:b + kw * R]) for b in range(0, n, kw * R)])
    assert np.array_equal(out, ref)


def test_kway_tile_bounds_hold_at_headline_shapes():
    """Q leaves room for the padded slices at the shapes the library registers."""
    for C, ipt in ((32768, 32), (16384, 32), (16384, 16), (8192, 32), (8192, 16), (4096, 16), (4096, 8), (2048, 8)):
        for kw in (2, 4, 8, 16, 32):
            for slog2 in range(3, 8):
                q = kw_q(C, ipt, kw, slog2)
                if q == 0:
                    continue
                worst = (q + kw) * (1 << slog2) - kw + kw * (ipt - 1)
                assert worst <= C


# ---- key-value form: ascending runs, exhaustion tested against real lengths

def corank_asc(K, a_off, la, b_off, lb, diag, top):
    """lds_corank_lift: A items among the first diag outputs of merge(A, B) (ties to A)."""
    lo = max(0, diag - lb)
    hi = min(diag, la)
    a = lo
    s = top
    while s > 0:
        t = min(a + s, hi)
        ka = K[a_off + t - 1]
        kb = K[b_off + diag - t]
        a = a if kb < ka else t
        s >>= 1
    return a


def merge_serial_kv(K, V, a_off, la, b_off, lb, ai, bi, cnt, ipt):
    """lds_merge_serial<HV=true>: outputs r[0, cnt), values by position."""
    ia, ib = a_off + ai, b_off + bi
    ea, eb = a_off + la, b_off + lb
    rk, rv = [KMAX64] * ipt, [0] * ipt
    for i in range(cnt):
        take_a = ib >= eb or (ia < ea and not (K[ib] < K[ia]))
        p = ia if take_a else ib
        rk[i], rv[i] = K[p], V[p]
        if take_a:
            ia += 1
        else:
            ib += 1
    return rk, rv


KMAX64 = 2**64 - 1


def model_tile_kv(runs, lo, hi, block, ipt):
    kw = len(runs)
    C = block * ipt
    n = [hi[j] - lo[j] for j in range(kw)]
    P, Rr = [0], [0]
    for j in range(kw):
        P.append(P[-1] + (n[j] + ipt - 1) // ipt * ipt)
        Rr.append(Rr[-1] + n[j])
    Np = P[-1]
    assert Np <= C
    rng = np.random.default_rng(Np)
    K = [int(x) for x in rng.integers(0, 2**63, C + 2 * ipt + 2)]
    V = [int(x) for x in rng.integers(0, 2**63, C + 2 * ipt + 2)]
    for j in range(kw):
        for q in range(P[j + 1] - P[j]):
            if q < n[j]:
                K[P[j] + q], V[P[j] + q] = runs[j][lo[j] + q]
            else:
                K[P[j] + q], V[P[j] + q] = KMAX64, 0
    logkw = kw.bit_length() - 1
    for lev in range(1, logkw + 1):
        hv = 1 << (lev - 1)
        top = pow2_floor(max(min(Rr[s0 + hv] - Rr[s0], Rr[s0 + 2 * hv] - Rr[s0 + hv])
                             for s0 in range(0, kw, 2 * hv)))
        regs = {}
        for tid in range(block):
            c0 = tid * ipt
            if c0 >= Np:
                continue
            j0 = sum(1 for j in range(1, kw) if P[j] <= c0)
            s0 = (j0 >> lev) << lev
            a_off, mid = P[s0], P[s0 + hv]
            la, lb = Rr[s0 + hv] - Rr[s0], Rr[s0 + 2 * hv] - Rr[s0 + hv]
            diag = c0 - a_off
            c = min(max(la + lb - diag, 0), ipt)
            rk, rv = [KMAX64] * ipt, [0] * ipt
            if c > 0:
                a = corank_asc(K, a_off, la, mid, lb, diag, top)
                rk, rv = merge_serial_kv(K, V, a_off, la, mid, lb, a, diag - a, c, ipt)
            regs[tid] = (rk, rv)
        for tid, (rk, rv) in regs.items():
            for i in range(ipt):
                K[tid * ipt + i], V[tid * ipt + i] = rk[i], rv[i]
    cnt = sum(n)
    return list(zip(K[:cnt], V[:cnt]))


@pytest.mark.parametrize("kw", [2, 4, 8, 16])
@pytest.mark.parametrize("dist", ["uniform", "few", "max"])
def test_kway_kv_tile_model(kw, dist):
    """KV tiles: keys sorted, (key, value) multiset kept, maximum keys never
    displaced by pads (the 'max' case mixes real maximum keys into every run)."""
    rng = np.random.default_rng(kw + len(dist))
    block, ipt = 64, 8
    runs, lo, hi = [], [], []
    for j in range(kw):
        m = int(rng.integers(0, min(60, block * ipt // kw - ipt)))  # padded slices fit the tile
        if dist == "uniform":
            keys = rng.integers(0, 2**63, m)
        elif dist == "few":
            keys = rng.integers(0, 3, m)
        else:
            keys = np.where(rng.random(m) < 0.5, KMAX64, rng.integers(0, 5, m).astype(object))
        keys = sorted(int(k) for k in keys)
        runs.append([(k, 1000 * j + i) for i, k in enumerate(keys)])
        lo.append(0)
        hi.append(m)
    out = model_tile_kv(runs, lo, hi, block, ipt)
    allp = sorted(p for run in runs for p in run)
    assert [k for k, _ in out] == [k for k, _ in allp]
    assert sorted(out) == allp


def group_search_model(lo, hi, G, pre):
    """k_kway_bounds' group_search for one lane group: a (G+1)-ary search for the
    first position in [lo, hi) where pre fails, one probe per lane per step."""
    steps = 0
    while lo < hi:
        w = hi - lo
        mids = [lo + w * (g + 1) // (G + 1) for g in range(G)]
        oks = [pre(x) for x in mids]
        c = sum(oks)
        assert oks == [True] * c + [False] * (G - c)  # the holding probes are a prefix
        m_lo = lo + w * c // (G + 1)
        m_hi = lo + w * (c + 1) // (G + 1)
        nlo, nhi = (m_lo + 1 if c > 0 else lo), (m_hi if c < G else hi)
        assert nlo <= nhi and (nlo > lo or nhi < hi)  # progress every step
        lo, hi = nlo, nhi
        steps += 1
    return lo, steps


@pytest.mark.parametrize("G", [2, 4, 8, 16, 32])
def test_group_search_matches_bisect(G):
    import bisect
    rng = np.random.default_rng(G)
    for trial in range(400):
        n = int(rng.integers(0, 5000)) if trial % 4 else int(rng.integers(0, 70))
        keys = np.sort(rng.integers(0, max(1, n // 3 + 1), n))
        v = int(rng.integers(-1, n // 3 + 2))
        lo = int(rng.integers(0, n + 1))
        hi = int(rng.integers(lo, n + 1))
        for up in (True, False):
            pre = (lambda i: not (v < keys[i])) if up else (lambda i: keys[i] < v)
            got, steps = group_search_model(lo, hi, G, pre)
            sub = keys[lo:hi]
            ref = lo + (bisect.bisect_right(sub, v) if up else bisect.bisect_left(sub, v))
            assert got == ref, (G, lo, hi, v, up)
            # (G+1)-ary: about log_{G+1} of the range
            if hi > lo:
                assert steps <= int(np.ceil(np.log(hi - lo + 1) / np.log(G + 1))) + 2


@pytest.mark.parametrize("kw", [2, 4, 8])
def test_window_and_search_paths_both_taken(kw):
    # with the kernel's window (32 merged samples for fan-in 2 and 4) uniform
    # runs resolve in the window and sorted runs, which do not interleave,
    # need the search: both paths of k_kway_bounds are modelled at each fan-in
    block, ipt, slog2 = 64, 8, 3
    R = 1024
    n = R * kw * 2 + 3 * R + 77
    taken = {}
    for name, keys in (("uniform", np.random.default_rng(kw).integers(0, 2**32, n, dtype=np.uint64)),
                       ("sorted", np.arange(n, dtype=np.uint64) // 3)):
        PATHS.clear()
        keys = runs_of(keys, R)
        out = model_pass(keys, R, kw, slog2, block, ipt)
        ref = np.concatenate([np.sort(keys[b:b + kw * R]) for b in range(0, n, kw * R)])
        assert np.array_equal(out, ref)
        taken[name] = dict(PATHS)
    assert taken["uniform"].get("window", 0) > 0
    assert taken["sorted"].get("search", 0) > 0
