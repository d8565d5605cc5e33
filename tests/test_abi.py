"""The C ABI library loads, exports every symbol include/*.h declares, and
rejects bad arguments on the host exactly where the reference asserts
(src/ntt.rs:192-200,309-320,350-352).  CPU only: no call here gets past
argument validation, so nothing touches a GPU."""
import ctypes
import glob
import os
import re

import pytest

from boojum_cuda_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INVALID = _lib.BC_ERROR_INVALID_VALUE


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(bc_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert "bc_batch_ntt_out_of_place" in names and "bc_context_create" in names
    for n in sorted(names):
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"


def test_omega_log_order():
    assert _lib.lib().bc_omega_log_order() == 32


FAKE_IN = 0x10000
FAKE_OUT = 0x80000


def internal(**kw):
    a = dict(i=FAKE_IN, o=FAKE_OUT, log_n=4, num=1, si=16, so=16, br=False, inv=False, ext=0, coset=0)
    a.update(kw)
    return _lib.lib().bc_batch_ntt_internal(a["i"], a["o"], a["log_n"], a["num"], a["si"], a["so"], a["br"],
                                            a["inv"], a["ext"], a["coset"], None)


@pytest.mark.parametrize("kw", [
    dict(log_n=0),                              # assert!(log_n >= 1)
    dict(log_n=30, ext=3, si=2**30, so=2**30),  # log_n + log_ext <= OMEGA_LOG_ORDER
    dict(si=15),                                # n <= stride_in
    dict(so=8),                                 # n <= stride_out
    dict(coset=1, ext=0),                       # coset needs an extension
    dict(coset=4, ext=2),                       # coset < 2^log_ext
])
def test_internal_rejects_reference_asserts(kw):
    assert internal(**kw) == INVALID


def test_internal_zero_columns_is_a_noop():
    assert internal(num=0) == 0


def oop(**kw):
    a = dict(i=FAKE_IN, il=64, o=FAKE_OUT, ol=64, log_n=4, num=2, ioff=8, ooff=8, si=32, so=32)
    a.update(kw)
    return _lib.lib().bc_batch_ntt_out_of_place(a["i"], a["il"], a["o"], a["ol"], a["log_n"], a["num"], a["ioff"],
                                                a["ooff"], a["si"], a["so"], False, False, 0, 0, None)


@pytest.mark.parametrize("kw", [
    dict(ioff=17),                  # inputs_offset + n <= stride
    dict(ooff=17),                  # outputs_offset + n <= stride
    dict(il=63),                    # len >= num_ntts * stride
    dict(ol=63),
    dict(o=FAKE_IN),                # in != out
    dict(o=FAKE_IN - 8 * 8, ooff=16, ioff=8),  # in != out after offsets
    dict(i=0),
])
def test_out_of_place_rejects_reference_asserts(kw):
    assert oop(**kw) == INVALID


def test_in_place_rejects_reference_asserts():
    L = _lib.lib()
    assert L.bc_batch_ntt_in_place(FAKE_IN, 64, 4, 2, 17, 32, False, False, 0, 0, None) == INVALID
    assert L.bc_batch_ntt_in_place(FAKE_IN, 63, 4, 2, 8, 32, False, False, 0, 0, None) == INVALID


def test_context_create_rejects_out_of_range_split():
    h = ctypes.c_void_p()
    assert _lib.lib().bc_context_create(33, 12, ctypes.byref(h)) == INVALID
    assert _lib.lib().bc_context_create(12, 40, ctypes.byref(h)) == INVALID
    assert _lib.lib().bc_context_destroy(None) == INVALID


def test_python_mirror_raises_assertion_on_bad_args():
    from boojum_cuda_amd import ntt
    with pytest.raises(AssertionError):
        ntt.batch_ntt_internal(FAKE_IN, FAKE_OUT, 0, 1, 16, 16, False, False, 0, 0, None)


def test_missing_library_fails_loudly():
    """No fallback: with the HIP library absent, the first call into the
    package raises instead of computing anything on the CPU."""
    import subprocess
    import sys
    code = ("import boojum_cuda_amd._lib as L; L.LIB_PATH = '/nonexistent/libboojum_ntt.so'\n"
            "from boojum_cuda_amd import ntt\n"
            "try:\n    ntt.batch_ntt_internal(0, 0, 4, 1, 16, 16, False, False, 0, 0, 0)\n"
            "except ImportError as e:\n    print('raised', e)\n")
    pkg = os.path.join(ROOT, "era-boojum-cuda_amd")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                         env={**os.environ, "PYTHONPATH": pkg})
    assert out.returncode == 0 and out.stdout.startswith("raised"), out.stdout + out.stderr


def test_plan_of_a_batch_beyond_the_grid_limit_widens_pair_groups():
    # 2^20 columns of 2^13 (64 GiB): 2^21 (column, block) pairs in the strided
    # pass, more than 65535 groups at the default pairs per workgroup
    from boojum_cuda_amd import ntt
    for bitrev in (False, True):
        plan = ntt.describe_plan(13, 1 << 20, bitrev_inputs=bitrev, inverse=bitrev)
        assert sum(p["log_h"] for p in plan) == 13
        strided = [p for p in plan if not p["contiguous"]]
        assert strided
        for p in strided:
            assert 1 <= p["groups"] <= 65535
            assert p["groups"] * p["pairs_per_workgroup"] >= p["work_items"]
            assert (p["groups"] - 1) * p["pairs_per_workgroup"] < p["work_items"]


@pytest.mark.parametrize("log_n", [8, 12, 13, 16, 20, 21, 24, 26])
def test_plan_covers_every_bit_once(log_n):
    from boojum_cuda_amd import ntt
    for br, inv in [(False, False), (True, True), (True, False), (False, True)]:
        plan = ntt.describe_plan(log_n, 3, br, inv)
        assert sum(p["log_h"] for p in plan) == log_n
        # the contiguous pass works on the lowest bits: last in n2b, first in b2n
        assert plan[0 if br else -1]["contiguous"] == 1
        assert sum(p["contiguous"] for p in plan) == 1


def test_plan_rejects_reference_asserts():
    from boojum_cuda_amd import ntt
    with pytest.raises(AssertionError):
        ntt.describe_plan(0, 1)
    with pytest.raises(AssertionError):
        ntt.describe_plan(20, 1, log_extension_degree=2, coset_index=4)
    assert ntt.describe_plan(7, 4) == []


def test_plan_of_the_latency_path():
    """1..4 columns of 2^16 without a coset take two one-wave 8-bit passes
    (the top one strided, with n^-1 for inverses); 5 columns, or a coset,
    take the 4096-point pass kernels."""
    from boojum_cuda_amd import ntt
    for cols in (1, 4):
        for br, inv in [(False, False), (True, True)]:
            plan = ntt.describe_plan(16, cols, br, inv)
            assert [p["log_h"] for p in plan] == [8, 8]
            assert all(p["workgroup_threads"] == 64 for p in plan)
            top = plan[1 if br else 0]
            assert top["contiguous"] == 0 and top["positions"] == 256 and top["groups"] == cols
            assert top["mode"] == (1 if inv else 0)
            assert plan[0 if br else 1]["contiguous"] == 1
    assert all(p["workgroup_threads"] == 256 for p in ntt.describe_plan(16, 5))
    assert all(p["workgroup_threads"] == 256 for p in ntt.describe_plan(16, 1, log_extension_degree=1))


def test_ragged_wait_calls_check_arguments_without_a_device():
    """The ragged wait-status and wait-limit calls are host-side: a null plan
    is an argument error, and the limit setter accepts any value (0 restores
    the default)."""
    from boojum_cuda_amd._lib import lib
    assert lib().bc_ragged_wait_status(None) == 1  # BC_ERROR_INVALID_VALUE
    assert lib().bc_ragged_set_wait_limit(5) == 0
    assert lib().bc_ragged_set_wait_limit(0) == 0
