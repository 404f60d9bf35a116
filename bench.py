#!/usr/bin/env python3
"""Headline benchmark: plain ascending sort of 2^28 uint32 keys (1 GiB), raw
std::mt19937(42) output, on one MI355X through libogs.so's C-ABI.

A step is one full sort of the same resident input (ogs_sort_keep reads the
pristine input buffer and ping-pongs between two work buffers, so no restore
copy runs inside the timed region).  Prints one JSON line:
  value        whole-job G keys/s over the K timed steps
  roofline     the dominant kernel (by summed time) vs the 8 TB/s HBM peak,
               algorithmic bytes per launch (2 * n * w) / its mean launch time,
               timed with HIP events recorded on the launching stream over
               --profile-steps sorts run after the timed region (events between
               every kernel add a few microseconds per kernel, so the timed
               steps run without them)
  pass_roofline the pass-count roofline of the whole sort: P * 2 * n * w / t
  cpu_baseline the oracle (a scalar C restatement of Origami's mergesort, the
               reference's REG_TYPE 0 build) on the same full input, on a
               power-of-two number of this host's cores (at most 16), min_k = 4
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "origami-sort_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    calibrated against a copy kernel), committed under profiles/; None unless
    the committed measurement is of this workload."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        if d.get("bench_workload") != workload:
            return None
        return d.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def host_cpu():
    """CPU model (name, family/model numbers, AVX-512 presence), physical cores and
    logical CPUs of this host (/proc/cpuinfo)."""
    model, phys = "unknown", set()
    first = {}
    logical = os.cpu_count() or 1
    cur = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f.read().splitlines() + [""]:
                if not line.strip():
                    if "core id" in cur:
                        phys.add((cur.get("physical id", "0"), cur["core id"]))
                    if not first:
                        first = dict(cur)
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
    except OSError:
        pass
    if first.get("model name"):
        model = first["model name"]
        if "cpu family" in first and "model" in first:
            model += f" (family {first['cpu family']}, model {first['model']})"
        flags = first.get("flags", "").split()
        model += ", AVX-512" if "avx512f" in flags else (", AVX2" if "avx2" in flags else "")
    return model, (len(phys) or logical), logical


def cpu_baseline(kind_oracle, host_keys, max_threads=16):
    """The oracle's multithreaded driver (sort_multi_thread restated,
    Origami/sorter.h:4194-4287, scalar compare-exchanges as in the reference's
    REG_TYPE 0 build) on the full input, timed like Origami/bench_sort.cpp:94-129:
    the sort call only, 3 repeats, mean and median; n_threads = the largest
    power of two <= the physical cores, at most max_threads (the host share of
    one GPU), min_k = 4 (Origami/README.md:557)."""
    import oracle
    model, phys, logical = host_cpu()
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else logical
    threads = 1
    while threads * 2 <= min(phys, avail, max_threads, 64):
        threads *= 2
    n = host_keys.size
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        out = oracle.sort(kind_oracle, host_keys, threads=threads, min_k=4)
        times.append(time.perf_counter() - t0)
    del out
    mean, med = sum(times) / 3, sorted(times)[1]
    return {"value": round(n / med / 1e9, 6), "unit": "G keys/s", "cores": threads, "kind": "port",
            "mean_value": round(n / mean / 1e9, 6), "cpu_model": model, "physical_cores": phys,
            "logical_cpus": logical,
            "sample": f"the full input (2^{n.bit_length() - 1} keys), median of 3 sort calls; "
                      f"oracle/origami_oracle.c sort_multi_thread restatement, scalar restatement (REG_TYPE 0), "
                      f"{threads} threads, min_k=4; {model}, {phys} physical cores"}


def result_checks(kind, d_in, d_out, n):
    """Size-independent checks of a sort's device output against its input:
    (sorted in the key order, the same multiset of items).  The multiset check
    compares the wrapping sum, the xor and the wrapping sum of squares of the
    items as 64-bit words (4-byte kinds: of the keys widened to 64 bits)."""
    import torch

    import ogs
    w = ogs.WIDTH[kind]

    def words(t):
        if w == 4:
            return t[: n * 4].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        return t[: n * w].view(torch.int64)

    def xor_all(t):
        t = t.clone()
        m = t.numel()
        while m > 1:  # halving xor-reduction
            h = m // 2
            t[:h] ^= t[h:2 * h]
            if m & 1:
                t[0] ^= t[m - 1]
            m = h
        return int(t[0]) if t.numel() else 0

    a, b = words(d_in), words(d_out)
    same = (int(a.sum()) == int(b.sum()) and xor_all(a) == xor_all(b)
            and int((a * a).sum()) == int((b * b).sum()))
    # each item's key in the sort order as an order-preserving int64
    if kind == ogs.F32:
        k = torch.where(b >= 0x80000000, 0xFFFFFFFF - b, b + 0x80000000)
    elif w == 4:
        k = b
    else:
        k = (b[0::2] if w == 16 else b) ^ torch.tensor(-(1 << 63), dtype=torch.int64, device=b.device)
    ok_sorted = bool((k[1:] >= k[:-1]).all())
    del a, b, k
    torch.cuda.empty_cache()
    return ok_sorted, same


def bench_batched(ctx, total_log2, max_seg_log2, seed=4, nseg_fixed=0):
    """Batched segmented sort (one ogs_segmented_sort launch sequence over CSR
    segments with log-uniform lengths in [1, 2^max_seg_log2], and the same sort
    captured as a HIP graph and replayed) against one plain ogs_sort launch per
    segment, same u32 keys, all on the GPU.  Segments are drawn until
    2^total_log2 keys, or exactly nseg_fixed of them when that is set (then the
    keys come from torch's GPU generator: the total is tens of GB)."""
    import numpy as np
    import torch

    import ogs
    rng = np.random.default_rng(seed)
    lens = []
    if nseg_fixed:
        lens = [max(1, int(2.0 ** (u * max_seg_log2))) for u in rng.random(nseg_fixed)]
    else:
        total_cap = 1 << total_log2
        acc = 0
        while acc < total_cap:
            ln = max(1, int(2.0 ** (rng.random() * max_seg_log2)))
            ln = min(ln, total_cap - acc)
            lens.append(ln)
            acc += ln
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n = int(off[-1])
    log(f"bench: batched {len(lens)} segments, {n} keys")
    if nseg_fixed:
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        d0 = torch.empty(n * 4, dtype=torch.uint8, device="cuda")
        d0i = d0.view(torch.int32)
        for o in range(0, n, 1 << 28):
            c = min(n, o + (1 << 28)) - o
            d0i[o:o + c] = torch.randint(-(1 << 31), 1 << 31, (c,), dtype=torch.int64, device="cuda",
                                         generator=g).to(torch.int32)
    else:
        host = ogs.generate(ogs.U32, ogs.UNIFORM, n, seed=seed)
        d0 = ogs.to_device(host)
    d = torch.empty_like(d0)
    s = torch.empty_like(d0)
    doff = torch.from_numpy(off.view(np.int64).copy()).cuda()
    d.copy_(d0)
    ctx.segmented_ptr(ogs.U32, d.data_ptr(), s.data_ptr(), doff.data_ptr(), len(lens), n)  # warm-up
    torch.cuda.synchronize()
    reps = 3
    t_b = 0.0
    for _ in range(reps):
        d.copy_(d0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p = ctx.segmented_ptr(ogs.U32, d.data_ptr(), s.data_ptr(), doff.data_ptr(), len(lens), n)
        torch.cuda.synchronize()
        t_b += time.perf_counter() - t0
    t_b /= reps
    res_b = s if p == s.data_ptr() else d
    batched_out = res_b.clone()
    # the same sort captured once as a HIP graph (ogs_segmented_plan) and
    # replayed: one hipGraphLaunch per sort
    plan, pres = ctx.segmented_plan(ogs.U32, d.data_ptr(), s.data_ptr(), doff.data_ptr(), len(lens), n)
    t_g = 0.0
    graph_same = True
    for r in range(reps + 1):
        d.copy_(d0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.launch()
        torch.cuda.synchronize()
        if r:
            t_g += time.perf_counter() - t0
        graph_same = graph_same and torch.equal(s if pres == s.data_ptr() else d, batched_out)
    t_g /= reps
    plan.close()
    # one plain launch per segment
    d.copy_(d0)
    torch.cuda.synchronize()
    w = 4
    base_d, base_s = d.data_ptr(), s.data_ptr()
    t0 = time.perf_counter()
    outs = []
    for i, ln in enumerate(lens):
        o = int(off[i]) * w
        outs.append(ctx.sort_ptr(ogs.U32, base_d + o, base_s + o, ln) - o)
    torch.cuda.synchronize()
    t_p = time.perf_counter() - t0
    # every segment equals the batched result
    same = True
    for i, ln in enumerate(lens):
        o = int(off[i]) * w
        src = d if outs[i] == base_d else s
        if not torch.equal(src[o:o + ln * w], batched_out[o:o + ln * w]):
            same = False
            break
    ok_sorted = True
    return {"segments": len(lens), "keys": n, "max_segment": int(max(lens)),
            "lengths": (f"log-uniform [1, 2^{max_seg_log2}], {nseg_fixed} segments" if nseg_fixed else
                        f"log-uniform [1, 2^{max_seg_log2}], drawn until 2^{total_log2} keys"),
            "batched_ms": round(t_b * 1e3, 3), "batched_keys_per_s": round(n / t_b / 1e9, 4),
            "graph_ms": round(t_g * 1e3, 3), "graph_keys_per_s": round(n / t_g / 1e9, 4),
            "graph_identical": bool(graph_same),
            "plain_per_segment_ms": round(t_p * 1e3, 3), "plain_keys_per_s": round(n / t_p / 1e9, 4),
            "unit": "G keys/s", "speedup": round(t_p / t_b, 2), "graph_speedup": round(t_p / t_g, 2),
            "identical_results": bool(same and ok_sorted)}


def bench_generator(ctx, buf, reps=5):
    """Device input generation (ogs_generate) into an already allocated
    buffer: GB/s of keys written, median of `reps`.  The shuffled Pareto
    bursts include their host burst table and the (tag, key) sort."""
    import torch

    import ogs
    res = {}
    nbytes = buf.numel()
    cases = [("u32", ogs.U32, ogs.UNIFORM | ogs.GEN_COUNTER), ("u64", ogs.U64, ogs.PARETO),
             ("u64", ogs.U64, ogs.ZIPF | ogs.GEN_COUNTER), ("u64", ogs.U64, ogs.PARETO_SHUFFLED)]
    names = {ogs.UNIFORM: "uniform", ogs.PARETO: "pareto", ogs.ZIPF: "zipf", ogs.PARETO_SHUFFLED: "pareto_shuffled"}
    for kname, kind, dist in cases:
        w = ogs.WIDTH[kind]
        n = nbytes // w
        if dist == ogs.PARETO_SHUFFLED:
            n = min(n, 1 << 26)  # pairs take 32 bytes per key
        times = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.generate_ptr(kind, dist, 42, n, buf.data_ptr())
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        t = sorted(times)[reps // 2]
        res[f"{kname}_{names[dist & 0xFF]}"] = {"n": n, "ms": round(t * 1e3, 3),
                                                 "GB_per_s": round(n * w / t / 1e9, 1)}
        log(f"bench: generator {kname} {names[dist & 0xFF]} n={n}: {t * 1e3:.3f} ms")
    return res


def bench_serving(ctx, nreq, rate_factor, streams=4, seed=5, coalesce=0):
    """Serving loop: Poisson arrivals of ORDER-BY requests (key type uniform over
    u32/u64/u64+u64, length log-uniform in [2^10, 2^26], keys uniform or Zipf 50/50,
    device-resident), served by ogs_server (continuous batching, 4 streams) and,
    for comparison, by serial plain launches at the same arrival times."""
    import threading

    import numpy as np
    import torch

    import ogs
    rng = np.random.default_rng(seed)
    kinds = [ogs.U32, ogs.U64, ogs.KV_U64U64]
    reqs = []
    for j in range(nreq):
        kind = kinds[int(rng.integers(0, 3))]
        n = int(2.0 ** (10 + 16 * rng.random()))
        dist = ogs.ZIPF if rng.random() < 0.5 else ogs.UNIFORM
        a = ogs.generate(kind, dist, n, seed=1000 + j)
        src = ogs.to_device(a)
        reqs.append((kind, n, src))
    total_keys = sum(r[1] for r in reqs)
    log(f"bench: serving {nreq} requests, {total_keys} keys")
    # serial service times (one plain launch per request) set the offered rate;
    # HIP events around each launch split a request's wall time into kernel
    # time and the fixed host cost (issue and synchronisation)
    outs_serial = []
    svc, ker = [], []
    for kind, n, src in reqs:
        a_ = torch.empty_like(src)
        b_ = torch.empty_like(src)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        p = ctx.sort_keep_ptr(kind, src.data_ptr(), a_.data_ptr(), b_.data_ptr(), n)
        e1.record()
        torch.cuda.synchronize()
        svc.append(time.perf_counter() - t0)
        ker.append(e0.elapsed_time(e1) / 1e3)
        outs_serial.append(a_ if p == a_.data_ptr() else b_)
    mean_svc = float(np.mean(svc))
    buckets = {}
    for (kind, n, _), w_, k_ in zip(reqs, svc, ker):
        b = "< 2^20" if n < (1 << 20) else ("2^20 .. 2^22" if n <= (1 << 22) else "> 2^22")
        e = buckets.setdefault(b, {"requests": 0, "keys": 0, "wall_ms": 0.0, "kernel_ms": 0.0})
        e["requests"] += 1
        e["keys"] += n
        e["wall_ms"] += w_ * 1e3
        e["kernel_ms"] += k_ * 1e3
    for e in buckets.values():
        e["wall_ms"] = round(e["wall_ms"], 3)
        e["kernel_ms"] = round(e["kernel_ms"], 3)
    # the coalescing bound: every request of a key type in one batch, the
    # types back to back (no arrivals to wait for)
    bound_s = 0.0
    for k in kinds:
        sel = [(n, src) for kk, n, src in reqs if kk == k]
        if not sel:
            continue
        outs_b = [torch.empty_like(src) for _, src in sel]
        scr = [torch.empty_like(src) for _, src in sel]
        items = [(src.data_ptr(), o.data_ptr(), sc.data_ptr(), n) for (n, src), o, sc in zip(sel, outs_b, scr)]
        ctx.sort_batch_ptrs(k, items)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.sort_batch_ptrs(k, items)
        torch.cuda.synchronize()
        bound_s += time.perf_counter() - t0
        del outs_b, scr
    rate = rate_factor / mean_svc
    gaps = rng.exponential(1.0 / rate, nreq)
    arrivals = np.cumsum(gaps) - gaps[0]

    def pct(x, q):
        return float(np.percentile(np.array(x) * 1e3, q))

    # serial plain launches at those arrival times (FIFO, one at a time)
    lat_s = []
    scratch = [(torch.empty_like(src), torch.empty_like(src)) for _, _, src in reqs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for (kind, n, src), at, (a_, b_) in zip(reqs, arrivals, scratch):
        now = time.perf_counter() - t0
        if now < at:
            time.sleep(at - now)
        ctx.sort_keep_ptr(kind, src.data_ptr(), a_.data_ptr(), b_.data_ptr(), n)
        torch.cuda.synchronize()
        lat_s.append(time.perf_counter() - t0 - at)
    span_s = time.perf_counter() - t0
    del scratch
    # the server at the same arrival times
    dsts = [torch.empty_like(src) for _, _, src in reqs]
    budget = 3 * max(n * ogs.WIDTH[k] for k, n, _ in reqs) + (256 << 20)
    srv = ogs.Server(ctx, hbm_budget_bytes=int(budget), streams=streams, coalesce_max_items=coalesce)
    torch.cuda.synchronize()
    tickets = []
    t0 = time.perf_counter()
    for (kind, n, src), at, dst in zip(reqs, arrivals, dsts):
        now = time.perf_counter() - t0
        if now < at:
            time.sleep(at - now)
        tickets.append(srv.submit(kind, src.data_ptr(), dst.data_ptr(), n))
    for t in tickets:
        srv.wait(t, 120.0)
    span_v = time.perf_counter() - t0
    lat_v = []
    for t in tickets:
        ts, _, td = srv.times(t)
        lat_v.append(td - ts)
    st = srv.stats()
    srv.close()
    same = all(torch.equal(dst[: n * ogs.WIDTH[k]], so[: n * ogs.WIDTH[k]]) if k != ogs.KV_U64U64 else
               torch.equal(dst[: n * 16].view(torch.int64)[0::2], so[: n * 16].view(torch.int64)[0::2])
               for (k, n, _), dst, so in zip(reqs, dsts, outs_serial))
    return {"requests": nreq, "keys": total_keys, "offered_rate_per_s": round(rate, 2),
            "offered_load_vs_serial": rate_factor,
            "server": {"keys_per_s": round(total_keys / span_v / 1e9, 4), "p50_ms": round(pct(lat_v, 50), 3),
                       "p99_ms": round(pct(lat_v, 99), 3), "batches": st["batches"]},
            "serial_plain": {"keys_per_s": round(total_keys / span_s / 1e9, 4), "p50_ms": round(pct(lat_s, 50), 3),
                             "p99_ms": round(pct(lat_s, 99), 3)},
            "server_vs_serial": round(span_s / span_v, 3),
            "serial_time_by_size": buckets,
            "serial_wall_ms": round(sum(svc) * 1e3, 3), "serial_kernel_ms": round(sum(ker) * 1e3, 3),
            "coalesced_bound": {"ms": round(bound_s * 1e3, 3), "keys_per_s": round(total_keys / bound_s / 1e9, 4),
                                "vs_serial": round(span_s / bound_s, 3),
                                "vs_serial_wall": round(sum(svc) / bound_s, 3),
                                "server_vs_bound": round(bound_s / span_v, 3),
                                "note": "each key type's requests in one ogs_sort_batch, types back to back"},
            "streams": streams, "unit": "G keys/s", "identical_results": bool(same),
            "note": "device-resident requests; latency = submit (arrival) to result ready"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=10,
                    help="sorts run with per-kernel HIP events after the timed region (0: --steps)")
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--kind", default="u32", choices=["u32", "f32", "u64", "kv_u64u64"])
    ap.add_argument("--dist", default=None,
                    choices=["uniform", "zipf", "sorted", "reverse", "distinct16", "normal", "same",
                             "almost_sorted", "pareto", "pareto_bursts", "pareto_shuffled", "fibonacci"],
                    help="key distribution (default: uniform; normal for f32); the last six are written "
                         "on the device by ogs_generate")
    ap.add_argument("--no-generator", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batched-log2-total", type=int, default=30)
    ap.add_argument("--batched-segments", type=int, default=0,
                    help="draw exactly this many segments instead (100000 = the literal C4 batch, ~7.6e9 keys)")
    ap.add_argument("--serving-requests", type=int, default=48)
    ap.add_argument("--serving-streams", type=int, default=4)
    ap.add_argument("--serving-load", type=float, default=3.0,
                    help="offered request rate as a multiple of the serial plain-launch capacity")
    ap.add_argument("--serving-coalesce", type=int, default=0,
                    help="longest request the server batches with others (0: the server's default)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="after timing, run three 1-input-sized device copies (torch copy_) as the "
                         "known-size reference kernel that calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--tune", action="append", default=[], metavar="PARAM=VALUE",
                    help="override one tile-table entry of the benchmarked key class (repeatable)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "the sort shards by one all-to-all after local sorts, which is design only here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import ogs
    kind = ogs.KINDS[args.kind]
    n = 1 << args.log2n
    w = ogs.WIDTH[kind]
    dists = {"uniform": ogs.UNIFORM, "zipf": ogs.ZIPF, "sorted": ogs.SORTED, "reverse": ogs.REVERSE,
             "distinct16": ogs.DISTINCT16, "normal": ogs.NORMAL, "same": ogs.SAME,
             "almost_sorted": ogs.ALMOST_SORTED, "pareto": ogs.PARETO, "pareto_bursts": ogs.PARETO_BURSTS,
             "pareto_shuffled": ogs.PARETO_SHUFFLED, "fibonacci": ogs.FIBONACCI}
    dname = args.dist or ("normal" if kind == ogs.F32 else "uniform")
    dist = dists[dname]
    ctx = ogs.Context(0)
    for t in args.tune:
        pname, pval = t.split("=")
        ctx.tune_set(kind, pname, int(pval))
    if dist < ogs.SAME:
        log(f"bench: generating {n} {args.kind} keys, {dname} (seed 42)")
        host = ogs.generate(kind, dist, n, seed=42)
        d_in = ogs.to_device(host)
    else:
        # counter-based family, written on the device (the host copy is the
        # CPU baseline's input)
        log(f"bench: generating {n} {args.kind} keys on the device, {dname} (seed 42)")
        d_in = torch.empty(n * w, dtype=torch.uint8, device="cuda")
        ctx.generate_ptr(kind, dist, 42, n, d_in.data_ptr())
        host = d_in.cpu().numpy().view(ogs.NP_DTYPE[kind])
    d_a = torch.empty_like(d_in)
    d_b = torch.empty_like(d_in)
    torch.cuda.synchronize()

    def step():
        return ctx.sort_keep_ptr(kind, d_in.data_ptr(), d_a.data_ptr(), d_b.data_ptr(), n)

    # size-independent checks of one sort's result, before the warmup steps
    # (so the timed steps follow the warmup directly): sorted in the key
    # order, and the same multiset of items (wrapping sum, xor and sum of
    # squares of the raw bytes as 64-bit words; for 4-byte keys, of the keys)
    res = step()
    torch.cuda.synchronize()
    out_t = d_a if res == d_a.data_ptr() else d_b
    ok_sorted, ok_multiset = result_checks(kind, d_in, out_t, n)
    log(f"bench: result sorted={ok_sorted} multiset={ok_multiset}")
    assert ok_sorted and ok_multiset
    del out_t

    log(f"bench: warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_step / 1e3) / 1e9

    # per-kernel times: HIP events on the launch stream around every kernel of
    # prof_steps further sorts (outside the timed region)
    prof_steps = args.profile_steps or args.steps
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(prof_steps):
        step()
    torch.cuda.synchronize()
    ctx.profile(False)

    kernels = {}
    for kid, name in ((0, "block_sort"), (1, "merge_partition"), (2, "merge_pass"), (3, "merge4_pass"),
                      (4, "kway_pass")):
        ms, cnt, items = ctx.profile_read(kid)
        if cnt:
            kernels[name] = {"total_ms": ms, "launches": cnt, "mean_ms": ms / cnt, "items": items}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    dk = kernels[dom]
    bytes_per_launch = 2.0 * (dk["items"] / dk["launches"]) * w
    achieved = bytes_per_launch / (dk["mean_ms"] / 1e3) / 1e9
    P = ctx.pass_count(kind, n)
    pass_gbs = P * 2.0 * n * w / (ms_step / 1e3) / 1e9
    workload = (f"plain sort, 2^{args.log2n} {args.kind} " + ("key+value pairs" if kind == ogs.KV_U64U64 else "keys")
                + f", {dname}, ascending" + ("" if kind == ogs.KV_U64U64 else ", key-only"))
    line = {
        "metric": "G keys/s + % HBM roofline, 2^28 u32 sort; batched keys/s and p50/p99 vs plain",
        "value": round(value, 4),
        "unit": "G keys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"u32": "u32", "f32": "f32", "u64": "u64", "kv_u64u64": "u64+u64"}[args.kind],
        "data": ("synthetic: raw std::mt19937(42) keys" if kind == ogs.U32 and dname == "uniform"
                 else f"synthetic: seeded generator, {dname} keys"
                 + (" (counter family, written on the device)" if dist >= ogs.SAME else "")),
        "config": {"workload": workload, "n": n, "item_bytes": w},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(dom, workload),
                     "timing": f"HIP events on the launch stream, mean over {dk['launches']} launches in "
                               f"{prof_steps} sorts after the timed region"},
        "pass_roofline": {"P": P, "bytes_per_key": P * 2 * w, "achieved_GBs": round(pass_gbs, 2),
                          "frac": round(pass_gbs / HBM_PEAK_GBS, 4),
                          "single_pass_GBs": round(2.0 * n * w / (ms_step / 1e3) / 1e9, 2)},
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
    }
    if args.copy_calib:
        for _ in range(3):
            d_b.copy_(d_in)
        torch.cuda.synchronize()
        line["copy_calib_bytes"] = int(d_in.numel())
    if not args.no_generator:
        line["generator"] = bench_generator(ctx, d_a)
    if not args.no_batched:
        del d_a, d_b
        torch.cuda.empty_cache()
        line["batched"] = bench_batched(ctx, args.batched_log2_total, 20, nseg_fixed=args.batched_segments)
    if not args.no_serving:
        torch.cuda.empty_cache()
        line["serving"] = bench_serving(ctx, args.serving_requests, args.serving_load, args.serving_streams,
                                        coalesce=args.serving_coalesce)
    if not args.no_cpu_baseline:
        import oracle
        okind = {"u32": oracle.U32, "f32": oracle.F32, "u64": oracle.U64, "kv_u64u64": oracle.KV_U64U64}[args.kind]
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(okind, host)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
