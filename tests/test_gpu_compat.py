"""The reference-shaped C++ host program (tests/compat_bench_sort.cpp), built
against include/origami_gpu_compat.h + libogs.so, sorts correctly on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "origami-sort_amd", "compat_bench_sort")


def test_compat_program_is_built():
    assert os.path.exists(BIN), "run origami-sort_amd/build.sh"


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1000, 1 << 20, 3000017])
def test_compat_program_runs(n):
    print(f"case compat n={n}", flush=True)
    r = subprocess.run([BIN, str(n)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
