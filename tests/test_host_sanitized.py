"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the
CPU: the pass and ragged-batch planner (era-boojum-cuda_amd/csrc/ntt_plan.cpp,
with the tuning table of ntt_tune.cpp) on a 1000-job plan drawn like the
ragged benchmark (tests/native/ragged_plan_check.cpp: level ranges, queue
positions, chunk index, column tables), and the oracle's C routines
(tests/native/oracle_check.c: round trips, DFT, Horner, threaded batches,
barycentric evaluation, fused Poseidon).  The planner is built with ROCm's
clang (the field header uses clang's carry builtins), the oracle with gcc."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "era-boojum-cuda_amd", "csrc")
CLANGXX = "/opt/rocm/llvm/bin/clang++"
SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def run(exe, timeout=240):
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout + out.stderr
    return out.stdout


@pytest.mark.skipif(not os.path.exists(CLANGXX), reason="ROCm clang++ not available")
def test_ragged_planner_invariants_sanitized(tmp_path):
    exe = tmp_path / "ragged_plan_check"
    subprocess.run([CLANGXX, "-O1", "-g", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include"] + SANITIZE +
                   ["-o", str(exe), os.path.join(HERE, "native", "ragged_plan_check.cpp"),
                    os.path.join(CSRC, "ntt_plan.cpp"), os.path.join(CSRC, "ntt_tune.cpp"), "-ldl"], check=True)
    words = run(exe).split()
    segments, levels, items, table_bytes = int(words[1]), int(words[2]), int(words[3]), int(words[4])
    # 1000 jobs of 14 sizes and 4 operations: several segments per level, at
    # most 3 passes per inverse stage plus 3 per coset stage
    assert segments > 100 and 2 <= levels <= 6 and items > segments
    # the packed table fits the 8 MiB static array of bc_ragged_run_no_alloc
    # with a wide margin
    assert table_bytes < (1 << 20)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_routines_sanitized(tmp_path):
    exe = tmp_path / "oracle_check"
    subprocess.run(["gcc", "-O1", "-g", "-std=c11", "-Wall", "-Werror"] + SANITIZE +
                   ["-o", str(exe), os.path.join(HERE, "native", "oracle_check.c"),
                    os.path.join(ROOT, "oracle", "gl_oracle.c"), os.path.join(ROOT, "oracle", "poseidon_oracle.c"),
                    "-lpthread"], check=True)
    run(exe)
