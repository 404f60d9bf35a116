"""ORACLE -- test infrastructure only.

ctypes binding of ``liboracle.so``, the plain-C restatement of Origami's CPU
mergesort (arif-arman/origami-sort, Origami/sorter.h, merger.h, merge_tree.h,
Partition.h).  Only the tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only as the
checker or the CPU baseline; the product library never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# key kinds: the same numbering as ogs_key_t in include/ogs.h
U32, I32, U64, I64, F32, KV_U64U64, KV_I64I64, KV_U32U32 = range(8)
KINDS = {"u32": U32, "i32": I32, "u64": U64, "i64": I64, "f32": F32,
         "kv_u64u64": KV_U64U64, "kv_i64i64": KV_I64I64, "kv_u32u32": KV_U32U32}
KV_DTYPE = np.dtype([("key", "<u8"), ("value", "<u8")])
KV_I_DTYPE = np.dtype([("key", "<i8"), ("value", "<i8")])
KV32_DTYPE = np.dtype([("key", "<u4"), ("value", "<u4")])
DTYPES = {U32: np.dtype("<u4"), I32: np.dtype("<i4"), U64: np.dtype("<u8"),
          I64: np.dtype("<i8"), F32: np.dtype("<f4"), KV_U64U64: KV_DTYPE,
          KV_I64I64: KV_I_DTYPE, KV_U32U32: KV32_DTYPE}
# distributions of orc_generate
UNIFORM, ZIPF, SORTED, REVERSE, DISTINCT16, NORMAL = range(6)
SAME, ALMOST_SORTED, PARETO, PARETO_BURSTS, PARETO_SHUFFLED, FIBONACCI = range(6, 12)
GEN_COUNTER = 0x100

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.orc_sort.argtypes = [i, vp, vp, u64, u64, i, i]
        L.orc_segmented_sort.argtypes = [i, vp, vp, u64, i]
        L.orc_network_u32.argtypes = [i, vp, u64]
        L.orc_merge.argtypes = [i, i, vp, u64, vp, u64, vp]
        L.orc_kway_merge.argtypes = [i, vp, vp, i, vp, u64]
        L.orc_partition.argtypes = [i, vp, vp, i, u64, vp]
        L.orc_generate.argtypes = [i, i, u64, u64, vp]
        L.orc_mix64.argtypes = [u64]
        L.orc_mix64.restype = u64
        L.orc_mt19937.argtypes = [ctypes.c_uint32, vp, u64]
        L.orc_mt19937_64.argtypes = [u64, vp, u64]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def sort(kind, a, sort_n=0, threads=1, min_k=2):
    """Return a sorted copy of ``a`` (kind = one of the ints above)."""
    a = np.ascontiguousarray(a).copy()
    tmp = np.empty_like(a)
    rc = lib().orc_sort(kind, _p(a), _p(tmp), a.size, sort_n, threads, min_k)
    assert rc == 0
    return a


def segmented_sort(kind, a, offsets, threads=1):
    a = np.ascontiguousarray(a).copy()
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    rc = lib().orc_segmented_sort(kind, _p(a), _p(off), off.size - 1, threads)
    assert rc == 0
    return a


def generate(kind, dist, n, seed=42):
    out = np.empty(n, dtype=DTYPES[kind])
    rc = lib().orc_generate(kind, dist, seed, n, _p(out))
    assert rc == 0
    return out


def mix64(z):
    return lib().orc_mix64(z)


def mt19937(seed, n):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_mt19937(seed, _p(out), n)
    return out


def mt19937_64(seed, n):
    out = np.empty(n, dtype=np.uint64)
    lib().orc_mt19937_64(seed, _p(out), n)
    return out


def network_u32(which, groups):
    g = np.ascontiguousarray(groups, dtype=np.uint32).copy()
    width = 16 if which == 16 else 8
    assert g.size % width == 0
    rc = lib().orc_network_u32(which, _p(g), g.size // width)
    assert rc == 0
    return g


def merge(kind, nreg, A, B):
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    C = np.empty(A.size + B.size, dtype=A.dtype)
    rc = lib().orc_merge(kind, nreg, _p(A), A.size, _p(B), B.size, _p(C))
    assert rc == 0
    return C


def kway_merge(kind, runs, bufn=64):
    base = np.ascontiguousarray(np.concatenate(runs))
    lens = np.array([r.size for r in runs], dtype=np.uint64)
    out = np.empty_like(base)
    rc = lib().orc_kway_merge(kind, _p(base), _p(lens), len(runs), _p(out), bufn)
    assert rc == 0
    return out


def partition(kind, runs, rank):
    base = np.ascontiguousarray(np.concatenate(runs))
    lens = np.array([r.size for r in runs], dtype=np.uint64)
    splits = np.zeros(len(runs), dtype=np.uint64)
    rc = lib().orc_partition(kind, _p(base), _p(lens), len(runs), rank, _p(splits))
    assert rc == 0
    return splits
