"""CPU check of the GPU kernels' Goldilocks arithmetic (csrc/gl_field.hpp is
host-compilable with clang): every operation, including multiplication by 2^e for all
192 exponents, against 128-bit integer arithmetic (tests/native/field_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
# the header uses clang's carry builtins (__builtin_addc / __builtin_subc)
CLANG = shutil.which("clang++") or shutil.which("amdclang++") or "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(CLANG) and shutil.which(CLANG) is None, reason="clang++ not available")
def test_field_ops_match_128bit_reference(tmp_path):
    exe = tmp_path / "field_check"
    subprocess.run([CLANG, "-O2", "-std=c++17", "-o", str(exe), os.path.join(HERE, "native", "field_check.cpp")],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout + out.stderr
