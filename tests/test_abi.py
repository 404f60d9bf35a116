"""Host-side checks of the C-ABI library (no GPU needed): it loads, exports every
function include/ogs.h declares, rejects bad arguments with status codes, and
its seeded generator reproduces the oracle's bytes."""
import ctypes

import numpy as np
import pytest

import ogs
import oracle


def test_library_exports_every_declared_symbol():
    lib = ogs.lib()
    names = ogs.declared_symbols()
    assert len(names) >= 13
    for name in names:
        assert hasattr(lib, name), name


def test_item_sizes():
    L = ogs.lib()
    assert [L.ogs_item_size(k) for k in range(8)] == [4, 4, 8, 8, 4, 16, 16, 8]
    assert L.ogs_item_size(99) == 0


def test_bad_arguments_return_status_not_exit():
    L = ogs.lib()
    out = ctypes.c_void_p()
    assert L.ogs_sort(None, 0, None, None, 10, ctypes.byref(out), None) == -1
    assert b"ctx" in L.ogs_last_error()
    buf = np.zeros(4, np.uint32)
    assert L.ogs_generate_host(0, 17, 1, 4, ctypes.c_void_p(buf.ctypes.data)) == -1
    assert L.ogs_generate_host(42, 0, 1, 4, ctypes.c_void_p(buf.ctypes.data)) == -1
    assert L.ogs_tune_set(None, 0, b"plain_block", 0) == -1


@pytest.mark.parametrize("kind", [oracle.U32, oracle.I32, oracle.F32, oracle.U64, oracle.I64,
                                  oracle.KV_U64U64, oracle.KV_I64I64, oracle.KV_U32U32])
@pytest.mark.parametrize("dist", [0, 1, 2, 3, 4, 5])
def test_generator_matches_oracle(kind, dist):
    if dist == 5 and kind != oracle.F32:
        pytest.skip("normal draws are float32 only")
    n = 20011
    a = ogs.generate(kind, dist, n, seed=42)
    b = oracle.generate(kind, dist, n, seed=42)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_generator_u32_is_raw_mt19937():
    a = ogs.generate(ogs.U32, ogs.UNIFORM, 10000, seed=5489)
    assert a[-1] == 4123659995  # [rand.predef]


def test_shape_registry():
    for kind in (ogs.U32, ogs.U64, ogs.KV_U64U64):
        bs = ogs.shapes(kind, 0)
        ms = ogs.shapes(kind, 1)
        assert bs and ms
        for b, ipt, w in bs:
            tile = b * ipt
            assert tile & (tile - 1) == 0 and b % 64 == 0 and 0 <= w <= 6
        for b, ipt, w in ms:
            tile = b * ipt
            assert tile & (tile - 1) == 0 and w == -1


def test_kway_shape_registry():
    # K-way merge shapes for every class; every block is whole LDS swizzle
    # blocks (256 items for 4-byte keys, 128 for 8-byte)
    for kind, blk in ((ogs.U32, 256), (ogs.U64, 128), (ogs.KV_U64U64, 128)):
        ks = ogs.shapes(kind, 2)
        assert ks
        for b, ipt, w in ks:
            assert b % blk == 0 and ipt & (ipt - 1) == 0 and w == -1


def test_kway_params_checked():
    try:
        ctx = ogs.Context(0)
    except ogs.OgsError:
        pytest.skip("creating a context needs a device")
    for bad in (("kway_fanin", 3), ("kway_fanin", 64), ("kway_sample", 2), ("kway_sample", 11)):
        with pytest.raises(ogs.OgsError):
            ctx.tune_set(ogs.U32, *bad)
    ctx.tune_set(ogs.U32, "kway_fanin", 32)
    assert ctx.tune_get(ogs.U32, "kway_fanin") == 32
    ctx.close()
