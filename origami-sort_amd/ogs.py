"""ctypes binding of libogs.so (include/ogs.h) for the tests, the bench and the
serving loop.  Device memory comes from torch tensors (plumbing only: the
sort itself runs in the HIP kernels of libogs.so).

The library must be built (``build.sh``); loading fails loudly otherwise --
there is no CPU fallback.
"""
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("OGS_LIB") or os.path.join(HERE, "libogs.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ogs.h")

U32, I32, U64, I64, F32, KV_U64U64, KV_I64I64, KV_U32U32 = range(8)
KINDS = {"u32": U32, "i32": I32, "u64": U64, "i64": I64, "f32": F32,
         "kv_u64u64": KV_U64U64, "kv_i64i64": KV_I64I64, "kv_u32u32": KV_U32U32}
KIND_NAMES = {v: k for k, v in KINDS.items()}
WIDTH = {U32: 4, I32: 4, F32: 4, U64: 8, I64: 8, KV_U64U64: 16, KV_I64I64: 16, KV_U32U32: 8}
KV_KINDS = (KV_U64U64, KV_I64I64, KV_U32U32)
KV_DTYPE = np.dtype([("key", "<u8"), ("value", "<u8")])
KV_I_DTYPE = np.dtype([("key", "<i8"), ("value", "<i8")])
KV32_DTYPE = np.dtype([("key", "<u4"), ("value", "<u4")])
NP_DTYPE = {U32: np.dtype("<u4"), I32: np.dtype("<i4"), U64: np.dtype("<u8"), I64: np.dtype("<i8"),
            F32: np.dtype("<f4"), KV_U64U64: KV_DTYPE, KV_I64I64: KV_I_DTYPE, KV_U32U32: KV32_DTYPE}
UNIFORM, ZIPF, SORTED, REVERSE, DISTINCT16, NORMAL = range(6)
SAME, ALMOST_SORTED, PARETO, PARETO_BURSTS, PARETO_SHUFFLED, FIBONACCI = range(6, 12)
GEN_COUNTER = 0x100  # the counter-based family for dists 0-5 (6-11 always are)

_lib = None
_live_contexts = weakref.WeakSet()


class OgsError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OgsError(f"{LIB_PATH} is not built; run origami-sort_amd/build.sh")
        # One HIP runtime per process: torch ships its own libamdhip64 (soname
        # libamdhip64.so.7, file libamdhip64.so).  Importing torch first makes
        # libogs.so bind to that copy; loading libogs.so first would pull in
        # /opt/rocm's copy as well, and torch's device init then fails.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i, cp = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_char_p
        L.ogs_ctx_create.argtypes = [ctypes.POINTER(vp), i]
        L.ogs_ctx_destroy.argtypes = [vp]
        L.ogs_device_errors.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32)]
        L.ogs_tune_reset.argtypes = [vp]
        L.ogs_item_size.argtypes = [i]
        L.ogs_item_size.restype = ctypes.c_size_t
        L.ogs_sort.argtypes = [vp, i, vp, vp, u64, ctypes.POINTER(vp), vp]
        L.ogs_sort_chunks.argtypes = [vp, i, vp, vp, u64, u64, ctypes.POINTER(vp), vp]
        L.ogs_sort_keep.argtypes = [vp, i, vp, vp, vp, u64, ctypes.POINTER(vp), vp]
        L.ogs_sort_noalloc.argtypes = [vp, i, vp, vp, u64, ctypes.POINTER(vp), vp]
        L.ogs_sort_soa.argtypes = [vp, i, vp, vp, vp, u64, vp]
        L.ogs_profile_enable.argtypes = [vp, i]
        L.ogs_profile_read.argtypes = [vp, i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64),
                                       ctypes.POINTER(u64)]
        L.ogs_profile_reset.argtypes = [vp]
        L.ogs_segmented_sort.argtypes = [vp, i, vp, vp, vp, u64, u64, ctypes.POINTER(vp), vp]
        L.ogs_sort_host.argtypes = [vp, i, vp, u64, u64]
        L.ogs_sort_batch.argtypes = [vp, i, vp, u64, vp]
        L.ogs_segmented_plan.argtypes = [vp, i, vp, vp, vp, u64, u64, ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L.ogs_plan_launch.argtypes = [vp, vp]
        L.ogs_plan_destroy.argtypes = [vp]
        L.ogs_pass_count.argtypes = [vp, i, u64]
        L.ogs_generate_host.argtypes = [i, i, u64, u64, vp]
        L.ogs_generate.argtypes = [vp, i, i, u64, u64, vp, vp]
        L.ogs_tune_set.argtypes = [vp, i, cp, i]
        L.ogs_tune_get.argtypes = [vp, i, cp, ctypes.POINTER(i)]
        L.ogs_tune_load.argtypes = [vp, cp]
        L.ogs_shape_info.argtypes = [i, i, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]
        L.ogs_server_create.argtypes = [ctypes.POINTER(vp), vp, ctypes.POINTER(ServerConfig)]
        L.ogs_server_destroy.argtypes = [vp]
        L.ogs_server_submit.argtypes = [vp, i, vp, vp, u64, i, ctypes.POINTER(u64)]
        L.ogs_server_wait.argtypes = [vp, u64, ctypes.c_double]
        L.ogs_server_poll.argtypes = [vp, u64, ctypes.POINTER(i)]
        L.ogs_server_times.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
        L.ogs_server_release.argtypes = [vp, u64]
        L.ogs_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.ogs_last_error.restype = cp
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise OgsError(f"ogs error {rc}: {lib().ogs_last_error().decode()}")
    return rc


def generate(kind, dist, n, seed=42):
    """Seeded host input (ogs_generate_host)."""
    out = np.empty(n, dtype=NP_DTYPE[kind])
    check(lib().ogs_generate_host(kind, dist, seed, n, ctypes.c_void_p(out.ctypes.data)))
    return out


def shapes(kind, merge):
    L = lib()
    b, p, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    cnt = check(L.ogs_shape_info(kind, merge, 0, ctypes.byref(b), ctypes.byref(p), ctypes.byref(w)))
    res = []
    for idx in range(cnt):
        L.ogs_shape_info(kind, merge, idx, ctypes.byref(b), ctypes.byref(p), ctypes.byref(w))
        res.append((b.value, p.value, w.value))
    return res


class Plan:
    """A captured segmented sort (ogs_segmented_plan); launch() replays it."""

    def __init__(self, h):
        self.h = h

    def launch(self, stream=0):
        check(lib().ogs_plan_launch(self.h, ctypes.c_void_p(stream)))

    def close(self):
        if self.h is not None:
            check(lib().ogs_plan_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("streams", ctypes.c_int),
                ("max_batch_requests", ctypes.c_uint32), ("max_batch_items", ctypes.c_uint64),
                ("coalesce_max_items", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in
                ("submitted", "completed", "rejected", "batches", "batched_requests", "items")]


SUBMIT_HOST = 1


class Server:
    """ogs_server: continuous-batching sort server on a context's device."""

    def __init__(self, ctx, hbm_budget_bytes, streams=4, max_batch_requests=0, max_batch_items=0,
                 coalesce_max_items=0):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        cfg = ServerConfig(hbm_budget_bytes, streams, max_batch_requests, max_batch_items, coalesce_max_items)
        check(lib().ogs_server_create(ctypes.byref(self.h), ctx.h, ctypes.byref(cfg)))

    def submit(self, kind, src, dst, n, host=False):
        t = ctypes.c_uint64()
        check(lib().ogs_server_submit(self.h, kind, ctypes.c_void_p(src), ctypes.c_void_p(dst), n,
                                      SUBMIT_HOST if host else 0, ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout=-1.0):
        return check(lib().ogs_server_wait(self.h, ticket, timeout))

    def poll(self, ticket):
        d = ctypes.c_int()
        check(lib().ogs_server_poll(self.h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def times(self, ticket):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().ogs_server_times(self.h, ticket, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def release(self, ticket):
        check(lib().ogs_server_release(self.h, ticket))

    def stats(self):
        st = ServerStats()
        check(lib().ogs_server_get_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().ogs_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchItem(ctypes.Structure):
    _fields_ = [("in_", ctypes.c_void_p), ("out", ctypes.c_void_p), ("scratch", ctypes.c_void_p),
                ("n", ctypes.c_uint64)]


class Context:
    """An ogs_ctx bound to one device."""

    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        check(lib().ogs_ctx_create(ctypes.byref(self.h), device))
        _live_contexts.add(self)

    def device_errors(self):
        """ogs_device_errors: waits for the device and returns (and clears) the
        context's error bits (1: K-way split table, 2: 4-way split table)."""
        f = ctypes.c_uint32()
        check(lib().ogs_device_errors(self.h, ctypes.byref(f)))
        return f.value

    def close(self):
        if self.h:
            lib().ogs_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- raw device-pointer entry points ------------------------------------
    def sort_ptr(self, kind, data, scratch, n, stream=0):
        out = ctypes.c_void_p()
        check(lib().ogs_sort(self.h, kind, ctypes.c_void_p(data), ctypes.c_void_p(scratch), n,
                             ctypes.byref(out), ctypes.c_void_p(stream)))
        return out.value

    def sort_noalloc_ptr(self, kind, data, scratch, n, stream=0):
        """ogs_sort_noalloc: ogs_sort with its workspace in the library's static
        device array (no allocation)."""
        out = ctypes.c_void_p()
        check(lib().ogs_sort_noalloc(self.h, kind, ctypes.c_void_p(data), ctypes.c_void_p(scratch), n,
                                     ctypes.byref(out), ctypes.c_void_p(stream)))
        return out.value

    def sort_keep_ptr(self, kind, src, data, scratch, n, stream=0):
        out = ctypes.c_void_p()
        check(lib().ogs_sort_keep(self.h, kind, ctypes.c_void_p(src), ctypes.c_void_p(data),
                                  ctypes.c_void_p(scratch), n, ctypes.byref(out), ctypes.c_void_p(stream)))
        return out.value

    def sort_soa_ptr(self, kind, keys, values, work, n, stream=0):
        """KV sort of separate key and value arrays (ogs_sort_soa); work: 32*n bytes."""
        check(lib().ogs_sort_soa(self.h, kind, ctypes.c_void_p(keys), ctypes.c_void_p(values), ctypes.c_void_p(work),
                                 n, ctypes.c_void_p(stream)))

    def profile(self, on):
        check(lib().ogs_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        check(lib().ogs_profile_reset(self.h))

    def profile_read(self, kind):
        """kind 0 block sort, 1 merge partition, 2 2-way merge pass, 3 4-way merge pass
        -> (ms, launches, items)."""
        ms, cnt, items = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().ogs_profile_read(self.h, kind, ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(items)))
        return ms.value, cnt.value, items.value

    def sort_chunks_ptr(self, kind, data, scratch, n, chunk, stream=0):
        out = ctypes.c_void_p()
        check(lib().ogs_sort_chunks(self.h, kind, ctypes.c_void_p(data), ctypes.c_void_p(scratch), n, chunk,
                                    ctypes.byref(out), ctypes.c_void_p(stream)))
        return out.value

    def segmented_ptr(self, kind, data, scratch, offsets, nseg, n, stream=0):
        out = ctypes.c_void_p()
        check(lib().ogs_segmented_sort(self.h, kind, ctypes.c_void_p(data), ctypes.c_void_p(scratch),
                                       ctypes.c_void_p(offsets), nseg, n, ctypes.byref(out), ctypes.c_void_p(stream)))
        return out.value

    def segmented_plan(self, kind, data, scratch, offsets, nseg, n):
        """ogs_segmented_plan: a graph-captured segmented sort bound to these
        buffers; returns (Plan, pointer of the result buffer)."""
        h = ctypes.c_void_p()
        out = ctypes.c_void_p()
        check(lib().ogs_segmented_plan(self.h, kind, ctypes.c_void_p(data), ctypes.c_void_p(scratch),
                                       ctypes.c_void_p(offsets), nseg, n, ctypes.byref(h), ctypes.byref(out)))
        return Plan(h), out.value

    def sort_batch_ptrs(self, kind, items, stream=0):
        """items: list of (in_ptr, out_ptr, scratch_ptr_or_0, n)."""
        arr = (BatchItem * len(items))(*[BatchItem(a, b, c or None, n) for a, b, c, n in items])
        check(lib().ogs_sort_batch(self.h, kind, ctypes.cast(arr, ctypes.c_void_p), len(items),
                                   ctypes.c_void_p(stream)))

    def sort_host(self, kind, arr, chunk=0):
        a = np.ascontiguousarray(arr)
        check(lib().ogs_sort_host(self.h, kind, ctypes.c_void_p(a.ctypes.data), a.size, chunk))
        return a

    TUNE_PARAMS = ("plain_block", "plain_merge", "plain_corank", "seg_tiny_block", "seg_tiny_max",
                   "seg_small_block", "seg_small_max", "seg_large_block", "seg_merge", "seg_plain_min",
                   "partition_lanes", "plain_fanin", "quad_lanes", "kway_fanin", "kway_shape", "kway_sample",
                   "kway_scorank", "kway_swhole")

    def generate_ptr(self, kind, dist, seed, n, out, stream=0):
        """Counter-based input written on the device (ogs_generate)."""
        check(lib().ogs_generate(self.h, kind, dist, seed, n, ctypes.c_void_p(out), ctypes.c_void_p(stream)))

    def pass_count(self, kind, n):
        return check(lib().ogs_pass_count(self.h, kind, n))

    def tune_set(self, kind, param, value):
        check(lib().ogs_tune_set(self.h, kind, param.encode(), value))

    def tune_load(self, path):
        check(lib().ogs_tune_load(self.h, path.encode()))

    def tune_reset(self):
        """Every key class back to the compiled-in tile table."""
        check(lib().ogs_tune_reset(self.h))

    def tune_get(self, kind, param):
        v = ctypes.c_int()
        check(lib().ogs_tune_get(self.h, kind, param.encode(), ctypes.byref(v)))
        return v.value

    # -- torch-tensor helpers (device memory plumbing) ------------------------
    def sort_tensor(self, kind, data, scratch, n=None, stream=0):
        """Sort uint8/int tensors viewed as raw item bytes; returns the tensor
        (data or scratch) that holds the result."""
        n = data.numel() * data.element_size() // WIDTH[kind] if n is None else n
        p = self.sort_ptr(kind, data.data_ptr(), scratch.data_ptr(), n, stream)
        return data if p == data.data_ptr() else scratch


def to_device(arr, device="cuda"):
    """numpy structured/int array -> torch uint8 device tensor of its bytes."""
    import torch
    b = np.ascontiguousarray(arr).view(np.uint8)
    return torch.from_numpy(b.copy()).to(device)


def from_device(t, kind, n):
    return t[: n * WIDTH[kind]].cpu().numpy().view(NP_DTYPE[kind]).copy()


def live_contexts():
    """The contexts this process has open."""
    return [c for c in list(_live_contexts) if c.h]


def declared_symbols():
    """Function names declared in include/ogs.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*(ogs_\w+)\s*\(", txt, re.M)))
