"""CPU check of the serving loop's HBM arena allocator (host bookkeeping,
era-boojum-cuda_amd/csrc/arena.hpp): compiled with g++ and driven by random
allocations and piecewise releases (tests/native/arena_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_arena_random_alloc_release(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-o", str(exe),
                    os.path.join(HERE, "native", "arena_check.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout + out.stderr
