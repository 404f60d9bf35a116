"""include/origami_gpu_compat.h maps each item type the reference's host code
may pass (Origami/commons.h:47-72) to the ABI kind of the same width and
layout, and refuses every other type at compile time (a wrong mapping would
make ogs_sort_host read and write past the caller's buffer).  Compiled with
g++ on the CPU; nothing runs on a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")

MAPPED = [("ui", "OGS_U32", 4), ("int32_t", "OGS_I32", 4), ("float", "OGS_F32", 4),
          ("ui64", "OGS_U64", 8), ("i64", "OGS_I64", 8), ("KeyValue<i64, i64>", "OGS_KV_I64I64", 16),
          ("KeyValue<ui64, ui64>", "OGS_KV_U64U64", 16), ("KeyValue<ui, ui>", "OGS_KV_U32U32", 8)]
REFUSED = ["unsigned long long", "double", "KeyValue<int32_t, int32_t>", "KeyValue<ui, ui64>", "short"]


def compile_snippet(tmp_path, body):
    src = tmp_path / "t.cpp"
    src.write_text('#include "origami_gpu_compat.h"\n' + body + "\nint main() { return 0; }\n")
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", INC, str(src)], capture_output=True,
                          text=True, timeout=120)


def test_mapped_types(tmp_path):
    body = "\n".join(
        f"static_assert(origami_sorter::gpu_detail::key_type<{t}>() == {k}, \"{t}\");\n"
        f"static_assert(sizeof({t}) == {w}, \"width of {t}\");"
        for t, k, w in MAPPED)
    r = compile_snippet(tmp_path, body)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("t", REFUSED)
def test_refused_types(tmp_path, t):
    body = (f"void f({t} *d) {{ origami_sorter::sort_single_thread<{t}, {t}>(d, d, d + 1, 1); }}")
    r = compile_snippet(tmp_path, body)
    assert r.returncode != 0
    assert "unsupported item type" in r.stderr


def test_kv32_call_compiles(tmp_path):
    body = ("void f(KeyValue<ui, ui> *d) {\n"
            "  origami_sorter::sort_multi_thread<KeyValue<ui, ui>, KeyValue<ui, ui>>(d, d, 8, 8, 8, 4, nullptr);\n"
            "}")
    r = compile_snippet(tmp_path, body)
    assert r.returncode == 0, r.stderr
