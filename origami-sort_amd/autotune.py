#!/usr/bin/env python3
"""Autotuner for libogs.so's tile table (the run-time counterpart of the tuned
constants the reference fixes at compile time, Origami/config.h:45-264, and
measures with its phase benchmarks, Origami/bench_sort_phases.cpp).

For each key class (u32 / u64 / KV) it times, on the GPU:
  - plain sorts: every compatible (block-sort shape, merge shape) pair;
  - segmented sorts of a log-uniform batch: the large-class block and merge
    shapes, then the small class (one workgroup per segment), then the tiny
    class (one wavefront-sized workgroup per segment) and its cut-off;
and writes the winners as a tile-table file (ogs_tune_load format); a context
applies it when $OGS_TILES names it (the compiled-in table is the default).

  python origami-sort_amd/autotune.py [--quick] [--out PATH]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import ogs  # noqa: E402

CLASSES = [("u32", ogs.U32), ("u64", ogs.U64), ("kv_u64u64", ogs.KV_U64U64)]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def seg_batch(kind, total, max_log2, seed):
    rng = np.random.default_rng(seed)
    lens = []
    acc = 0
    while acc < total:
        ln = max(1, int(2.0 ** (rng.random() * max_log2)))
        ln = min(ln, total - acc)
        lens.append(ln)
        acc += ln
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    return off


def tune_class(ctx, name, kind, quick):
    import torch
    w = ogs.WIDTH[kind]
    bs = ogs.shapes(kind, 0)
    ms = ogs.shapes(kind, 1)
    reps = 3 if quick else 5
    result = {}
    # ---- plain sort
    n = (1 << 28) // (w // 4) if not quick else (1 << 24) // (w // 4)
    host = ogs.generate(kind, ogs.UNIFORM, n, seed=42)
    d_in = ogs.to_device(host)
    a = torch.empty_like(d_in)
    b = torch.empty_like(d_in)
    best = None
    # shapes x co-rank source for 2-way passes; shapes for 4-way passes (which
    # always use the split table)
    for bi, (bb, bp, bw) in enumerate(bs):
        for mi, (mb, mp, _) in enumerate(ms):
            if mb * mp > 2 * bb * bp or bb * bp < 1024 or mb * mp < 2048:
                continue
            for fanin, corank in ((2, 0), (2, 1), (4, 0)):
                ctx.tune_set(kind, "plain_block", bi)
                ctx.tune_set(kind, "plain_merge", mi)
                ctx.tune_set(kind, "plain_corank", corank)
                ctx.tune_set(kind, "plain_fanin", fanin)
                t = timed(lambda: ctx.sort_keep_ptr(kind, d_in.data_ptr(), a.data_ptr(), b.data_ptr(), n), reps)
                log(f"  {name} plain n={n} block={bs[bi]} merge={ms[mi]} fanin={fanin} corank={corank}: "
                    f"{t * 1e3:.3f} ms")
                if best is None or t < best[0]:
                    best = (t, bi, mi, corank, fanin)
    ctx.tune_set(kind, "plain_block", best[1])
    ctx.tune_set(kind, "plain_merge", best[2])
    ctx.tune_set(kind, "plain_corank", best[3])
    ctx.tune_set(kind, "plain_fanin", best[4])
    result["plain_block"], result["plain_merge"], result["plain_corank"] = best[1], best[2], best[3]
    result["plain_fanin"] = best[4]
    if best[3] == 0:  # the partition kernel runs (2-way table or 4-way split points)
        lanes0 = ctx.tune_get(kind, "partition_lanes")  # the width `best` was timed with
        for lanes in (1, 4, 8, 16):
            ctx.tune_set(kind, "partition_lanes", lanes)
            t = timed(lambda: ctx.sort_keep_ptr(kind, d_in.data_ptr(), a.data_ptr(), b.data_ptr(), n), reps)
            log(f"  {name} plain partition_lanes={lanes}: {t * 1e3:.3f} ms")
            if t < best[0]:
                best = (t,) + best[1:]
                result["partition_lanes"] = lanes
        # keep that width unless another one measured faster
        result.setdefault("partition_lanes", lanes0)
        ctx.tune_set(kind, "partition_lanes", result["partition_lanes"])
    if best[4] == 4:  # the 4-way split kernel's lanes per slot
        qlanes0 = ctx.tune_get(kind, "quad_lanes")
        for lanes in (1, 2, 4, 8):
            ctx.tune_set(kind, "quad_lanes", lanes)
            t = timed(lambda: ctx.sort_keep_ptr(kind, d_in.data_ptr(), a.data_ptr(), b.data_ptr(), n), reps)
            log(f"  {name} plain quad_lanes={lanes}: {t * 1e3:.3f} ms")
            if t < best[0]:
                best = (t,) + best[1:]
                result["quad_lanes"] = lanes
        result.setdefault("quad_lanes", qlanes0)
        ctx.tune_set(kind, "quad_lanes", result["quad_lanes"])
    result["kway_fanin"] = 0
    if w <= 8:
        # key-only kinds: sample-split K-way passes (fan-in x shape x sample
        # spacing x block-sort shape); kept only when faster than the best
        # 2- and 4-way plan above
        ks = ogs.shapes(kind, 2)
        kbest = None
        for bi, (bb, bp, bw) in enumerate(bs):
            if bb * bp < 4096 or bb * bp & (bb * bp - 1):
                continue
            for ki in range(len(ks)):
                for fanin in (8, 16, 32):
                    for sample in (5, 6, 7):
                        ctx.tune_set(kind, "plain_block", bi)
                        ctx.tune_set(kind, "kway_fanin", fanin)
                        ctx.tune_set(kind, "kway_shape", ki)
                        ctx.tune_set(kind, "kway_sample", sample)
                        try:
                            t = timed(lambda: ctx.sort_keep_ptr(kind, d_in.data_ptr(), a.data_ptr(), b.data_ptr(), n),
                                      reps)
                        except ogs.OgsError as e:  # a shape that cannot hold the slices
                            log(f"  {name} kway block={bs[bi]} shape={ks[ki]} fanin={fanin}: {e}")
                            continue
                        log(f"  {name} kway block={bs[bi]} shape={ks[ki]} fanin={fanin} sample={sample}: "
                            f"{t * 1e3:.3f} ms")
                        if kbest is None or t < kbest[0]:
                            kbest = (t, bi, ki, fanin, sample)
        if kbest is not None and kbest[0] < best[0]:
            best = (kbest[0],) + best[1:]
            result["plain_block"] = kbest[1]
            result["kway_shape"], result["kway_fanin"], result["kway_sample"] = kbest[2], kbest[3], kbest[4]
        ctx.tune_set(kind, "plain_block", result["plain_block"])
        ctx.tune_set(kind, "kway_fanin", result["kway_fanin"])
        if result["kway_fanin"]:
            ctx.tune_set(kind, "kway_shape", result["kway_shape"])
            ctx.tune_set(kind, "kway_sample", result["kway_sample"])
    result["_plain"] = {"n": n, "ms": best[0] * 1e3, "gkeys": n / best[0] / 1e9}
    del d_in, a, b
    # ---- segmented batch: log-uniform lengths in [1, 2^20]
    total = (1 << 25) // (w // 4) if not quick else (1 << 23) // (w // 4)
    off = seg_batch(kind, total, 20, seed=1)
    nseg = off.size - 1
    host = ogs.generate(kind, ogs.UNIFORM, int(off[-1]), seed=7)
    d = ogs.to_device(host)
    s = torch.empty_like(d)
    d0 = d.clone()
    doff = torch.from_numpy(off.view(np.int64).copy()).cuda()
    N = int(off[-1])

    def run_seg():
        d.copy_(d0)
        ctx.segmented_ptr(kind, d.data_ptr(), s.data_ptr(), doff.data_ptr(), nseg, N)

    copy_t = timed(lambda: d.copy_(d0), reps)

    def measure(param, value):
        ctx.tune_set(kind, param, value)
        return timed(run_seg, reps) - copy_t

    def descend(param, values):
        bestv = None
        for v in values:
            t = measure(param, v)
            log(f"  {name} segmented {param}={v}: {t * 1e3:.3f} ms")
            if bestv is None or t < bestv[0]:
                bestv = (t, v)
        ctx.tune_set(kind, param, bestv[1])
        result[param] = bestv[1]
        return bestv[0]

    big = [i for i, (bb, bp, _) in enumerate(bs) if bb >= 128]
    descend("seg_large_block", big)
    lt = bs[result["seg_large_block"]][0] * bs[result["seg_large_block"]][1]
    descend("seg_merge", [i for i, (mb, mp, _) in enumerate(ms) if mb * mp <= 2 * lt])
    descend("seg_small_block", big)
    st = bs[result["seg_small_block"]][0] * bs[result["seg_small_block"]][1]
    descend("seg_small_max", sorted({st // 4, st // 2, st}))
    waves = [i for i, (bb, bp, _) in enumerate(bs) if bb == 64]
    if waves:
        descend("seg_tiny_block", waves)
        tt = bs[result["seg_tiny_block"]][0] * bs[result["seg_tiny_block"]][1]
        t_best = descend("seg_tiny_max", sorted({0, tt // 8, tt // 4, tt // 2, tt}))
    else:
        t_best = measure("seg_tiny_max", 0)
    t_best = descend("seg_plain_min", [0, 18, 20, 22])
    result["_segmented"] = {"segments": nseg, "items": N, "ms": t_best * 1e3, "gkeys": N / t_best / 1e9}
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--out", default=os.path.join(HERE, "ogs_tiles.autotuned.txt"),
                    help="tile file to write; a context applies it when $OGS_TILES names it")
    args = ap.parse_args()
    os.environ.pop("OGS_TILES", None)  # start from the compiled-in table
    ctx = ogs.Context(0)
    lines = ["# libogs.so tile table written by autotune.py (kind param value)"]
    summary = {}
    for name, kind in CLASSES:
        log(f"tuning {name}")
        res = tune_class(ctx, name, kind, args.quick)
        summary[name] = res
        for p in ogs.Context.TUNE_PARAMS:
            if p in res:
                lines.append(f"{name} {p} {res[p]}")
        lines.append(f"# {name} plain: {json.dumps(res['_plain'])}")
        lines.append(f"# {name} segmented: {json.dumps(res['_segmented'])}")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
