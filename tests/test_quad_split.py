"""The tile split of the 4-way merge pass, checked on a CPU model.

k_quad_bounds (origami-sort_amd/csrc/sort_kernels.hip) cuts each group of four
sorted runs A, B, C, D at points of the merge path of X = merge(A, B) and
Y = merge(C, D): one at every h-th item of X, one at every h-th item of Y, and
the group start; each point's slot in the group's tile list is computed in
closed form.  This model states the same rule and checks what the kernel
relies on, for random runs with heavy ties, sorted and reverse-ordered runs,
and ragged last groups:
  - the slots are a bijection onto [0, 1 + ceil(lx/h) + ceil(ly/h));
  - consecutive points are monotone in all four runs, at most h apart in X and
    in Y (so an interval holds at most 2h items, and the two intervals a
    workgroup of k_quad_merge_pass takes, from point 2w to point 2w + 2, fit its
    4h-item LDS tile);
  - merging each tile's four slices (A, B, C, D tie order) and concatenating
    the tiles gives the 4-way merge of the group.
The merge tree it is checked against is the reference's order of merging
(MergeTreeEven, Origami/merge_tree.h:2576-3010): merge(merge(A, B), merge(C, D)).
"""
import bisect
import random

import pytest


def merge2(p, q):
    """Stable merge, ties to p (merge path, Origami/merger.h:964-1169)."""
    i = j = 0
    out = []
    while i < len(p) or j < len(q):
        if j >= len(q) or (i < len(p) and not (q[j] < p[i])):
            out.append(p[i])
            i += 1
        else:
            out.append(q[j])
            j += 1
    return out


def corank(p, q, x):
    """Items of p among the first x outputs of merge2(p, q)."""
    lo, hi = max(0, x - len(q)), min(x, len(p))
    while lo < hi:
        m = (lo + hi) // 2
        if not (q[x - 1 - m] < p[m]):
            lo = m + 1
        else:
            hi = m
    return lo


def next_merged(p, q, a, b):
    if b >= len(q):
        return p[a]
    if a >= len(p):
        return q[b]
    return q[b] if q[b] < p[a] else p[a]


def tile_points(A, B, C, D, h):
    la, lb, lc, ld = map(len, (A, B, C, D))
    lx, ly = la + lb, lc + ld
    used = 1 + -(-lx // h) + -(-ly // h)
    slots = [None] * used
    slots[0] = (0, 0, 0, 0)
    for j in range(-(-lx // h)):
        x = j * h
        a = corank(A, B, x)
        v = next_merged(A, B, a, x - a)
        c, d = bisect.bisect_left(C, v), bisect.bisect_left(D, v)
        s = 1 + j + -(-(c + d) // h)
        assert slots[s] is None, f"X point {j} collides at slot {s}"
        slots[s] = (a, x - a, c, d)
    for i in range(-(-ly // h)):
        y = i * h
        c = corank(C, D, y)
        w = next_merged(C, D, c, y - c)
        a, b = bisect.bisect_right(A, w), bisect.bisect_right(B, w)
        s = 1 + i + -(-(a + b) // h)
        assert slots[s] is None, f"Y point {i} collides at slot {s}"
        slots[s] = (a, b, c, y - c)
    assert all(t is not None for t in slots), "slots are not a bijection"
    return slots


def check_group(A, B, C, D, h):
    slots = tile_points(A, B, C, D, h) + [tuple(map(len, (A, B, C, D)))]
    out = []
    for s0, s1 in zip(slots[:-1], slots[1:]):
        assert all(u <= v for u, v in zip(s0, s1)), (s0, s1)
        nx = (s1[0] - s0[0]) + (s1[1] - s0[1])
        ny = (s1[2] - s0[2]) + (s1[3] - s0[3])
        assert nx <= h and ny <= h, (nx, ny, h)
        X = merge2(A[s0[0]:s1[0]], B[s0[1]:s1[1]])
        Y = merge2(C[s0[2]:s1[2]], D[s0[3]:s1[3]])
        out += merge2(X, Y)
    # a workgroup's two intervals: from point 2w to point 2w + 2 (or the end)
    for p0, p2 in zip(slots[:-1:2], slots[2::2] + [slots[-1]]):
        assert sum(p2) - sum(p0) <= 4 * h
    assert out == merge2(merge2(A, B), merge2(C, D))


def make_runs(rng, R, kind, ragged):
    lens = [R] * 4
    if ragged:
        tot = rng.randint(1, 4 * R)
        lens = [max(0, min(R, tot - k * R)) for k in range(4)]
    if kind == "stacked":  # A < B < C < D, or the reverse
        runs, s = [], 0
        for m in lens:
            runs.append(list(range(s, s + m)))
            s += m
        if rng.random() < 0.5:
            runs = [list(range(1000 - k * R, 1000 - k * R + m)) for k, m in enumerate(lens)]
        return runs
    hi = {"uniform": 1000, "ties": 3, "few": 6}[kind]
    return [sorted(rng.randrange(hi) for _ in range(m)) for m in lens]


@pytest.mark.parametrize("kind", ["uniform", "ties", "few", "stacked"])
def test_quad_split_model(kind):
    rng = random.Random(hash(kind) & 0xFFFF)
    for _ in range(150):
        R = rng.choice([4, 8, 16, 32])
        h = min(R, rng.choice([1, 2, 4, 8]))
        check_group(*make_runs(rng, R, kind, rng.random() < 0.4), h)


def test_quad_split_catches_a_wrong_slot_rule():
    # a plausible slip -- ranking X points by floor instead of ceil -- must
    # break the bijection on some input
    def bad(A, B, C, D, h):
        la, lb, lc, ld = map(len, (A, B, C, D))
        used = 1 + -(-(la + lb) // h) + -(-(lc + ld) // h)
        seen = set()
        for j in range(-(-(la + lb) // h)):
            x = j * h
            a = corank(A, B, x)
            v = next_merged(A, B, a, x - a)
            c, d = bisect.bisect_left(C, v), bisect.bisect_left(D, v)
            seen.add(1 + j + (c + d) // h)
        for i in range(-(-(lc + ld) // h)):
            y = i * h
            c = corank(C, D, y)
            w = next_merged(C, D, c, y - c)
            a, b = bisect.bisect_right(A, w), bisect.bisect_right(B, w)
            seen.add(1 + i + -(-(a + b) // h))
        return len(seen) == used - 1 and max(seen) < used

    rng = random.Random(7)
    assert any(not bad(*make_runs(rng, 8, "uniform", False), 4) for _ in range(50))
