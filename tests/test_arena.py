"""CPU checks of the serving loop's host bookkeeping: the HBM arena allocator
(era-boojum-cuda_amd/csrc/arena.hpp), driven by random allocations and
piecewise releases (tests/native/arena_check.cpp), and the admission policy
(era-boojum-cuda_amd/csrc/admission.hpp) on a simulated server
(tests/native/admission_check.cpp).  Both are compiled with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def build_and_run(tmp_path, name, timeout=240):
    exe = tmp_path / name
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-Wall", "-Werror"] + SANITIZE +
                   ["-o", str(exe), os.path.join(HERE, "native", name + ".cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout + out.stderr
    return out.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_arena_random_alloc_release(tmp_path):
    build_and_run(tmp_path, "arena_check")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_admission_never_stalls_and_keeps_its_reservation(tmp_path):
    out = build_and_run(tmp_path, "admission_check")
    # fragmentation failures are the rare exception: most blocked parts are split instead
    words = out.split()
    requests, failed = int(words[1]), int(words[3])
    assert failed * 1000 < requests, out
