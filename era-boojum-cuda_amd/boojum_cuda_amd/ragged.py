"""Ragged batches of NTT / iNTT / coset-LDE jobs run by one persistent kernel
fed from a device work queue (C ABI: bc_ragged_*).  No reference
counterpart: the reference launches every job separately, which is what
`run_plain` does for comparison."""
import ctypes
from dataclasses import dataclass
from typing import Optional

from ._lib import BC_ERROR_INVALID_VALUE, HipError, lib

OP_NTT, OP_INTT, OP_LDE = 0, 1, 2


class _Job(ctypes.Structure):
    _fields_ = [("inputs", ctypes.c_void_p), ("outputs", ctypes.c_void_p), ("coeffs", ctypes.c_void_p),
                ("log_n", ctypes.c_uint32), ("num_cols", ctypes.c_uint32), ("stride_in", ctypes.c_uint32),
                ("stride_out", ctypes.c_uint32), ("stride_coeffs", ctypes.c_uint32), ("op", ctypes.c_uint32),
                ("log_blowup", ctypes.c_uint32)]


@dataclass
class Job:
    """One job: num_cols columns of 2^log_n points.

    OP_NTT: forward, natural -> bit-reversed.  OP_INTT: inverse,
    bit-reversed -> natural.  OP_LDE: bit-reversed evaluations ->
    coefficients (coeffs) -> bit-reversed LDE of blowup 2^log_blowup."""
    op: int
    log_n: int
    inputs: object                 # torch tensor (int64/uint64) on the device
    outputs: object
    num_cols: int = 1
    stride_in: Optional[int] = None
    stride_out: Optional[int] = None
    coeffs: object = None
    stride_coeffs: Optional[int] = None
    log_blowup: int = 0

    def _c(self):
        n = 1 << self.log_n
        out_len = n << (self.log_blowup if self.op == OP_LDE else 0)
        return _Job(self.inputs.data_ptr(), self.outputs.data_ptr(),
                    self.coeffs.data_ptr() if self.coeffs is not None else None, self.log_n, self.num_cols,
                    self.stride_in or n, self.stride_out or out_len, self.stride_coeffs or n, self.op,
                    self.log_blowup)


def _check(rc, what):
    if rc == 0:
        return
    if rc == BC_ERROR_INVALID_VALUE:
        raise AssertionError(f"{what}: argument assertion failed")
    raise HipError(rc, what)


def _array(jobs):
    arr = (_Job * max(1, len(jobs)))()
    for i, j in enumerate(jobs):
        arr[i] = j._c()
    return arr


class RaggedPlan:
    """Planned ragged batch; run() issues one persistent launch."""

    def __init__(self, jobs):
        self.jobs = list(jobs)
        h = ctypes.c_void_p()
        _check(lib().bc_ragged_plan_create(_array(self.jobs), len(self.jobs), ctypes.byref(h)), "ragged_plan_create")
        self._h = h

    def run(self, stream=None):
        import torch
        s = (stream or torch.cuda.current_stream()).cuda_stream
        _check(lib().bc_ragged_run(self._h, s), "ragged_run")

    def wait_status(self):
        """After the last run's stream has been synchronized: raises HipError
        (BC_ERROR_WAIT_TIMEOUT) if a wait between levels of that run gave up."""
        _check(lib().bc_ragged_wait_status(self._h), "ragged_run (wait between levels)")

    def info(self):
        u, lv, wg = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().bc_ragged_plan_info(self._h, ctypes.byref(u), ctypes.byref(lv), ctypes.byref(wg)), "plan_info")
        return {"units": u.value, "levels": lv.value, "workgroups": wg.value}

    def destroy(self):
        if self._h is not None:
            _check(lib().bc_ragged_plan_destroy(self._h), "ragged_plan_destroy")
            self._h = None


def run_plain(jobs, stream=None):
    """The same jobs as separate launch sequences (the per-job path)."""
    import torch
    s = (stream or torch.cuda.current_stream()).cuda_stream
    _check(lib().bc_ragged_run_plain(_array(jobs), len(jobs), s), "ragged_run_plain")


def run_no_alloc(jobs, stream=None):
    """Plans and runs the jobs with the plan in the library's static device
    array (no device allocation; results as RaggedPlan(jobs).run())."""
    import torch
    s = (stream or torch.cuda.current_stream()).cuda_stream
    _check(lib().bc_ragged_run_no_alloc(_array(jobs), len(jobs), s), "ragged_run_no_alloc")


def set_wait_limit(spins):
    """Polls of one wait between levels before it gives up (0: default); for
    tests of the timeout path."""
    _check(lib().bc_ragged_set_wait_limit(int(spins)), "ragged_set_wait_limit")
