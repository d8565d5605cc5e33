"""Serving loop over the C ABI (bc_server_*): requests of NTT / iNTT /
coset-LDE work arrive at any time, are admitted against an HBM budget and
run in continuous batches (see include/boojum_ntt.h).  Also the seeded input
generator (bc_fill_seeded).  No reference counterpart."""
import ctypes

from ._lib import BC_ERROR_INVALID_VALUE, HipError, lib
from .ragged import OP_INTT, OP_LDE, OP_NTT  # noqa: F401  (re-exported op codes)


class Config(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_part_elems", ctypes.c_uint64),
                ("max_batches_in_flight", ctypes.c_uint32), ("max_parts_per_batch", ctypes.c_uint32)]


class _Request(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint32), ("log_n", ctypes.c_uint32), ("num_cols", ctypes.c_uint32),
                ("log_blowup", ctypes.c_uint32), ("inputs", ctypes.c_void_p), ("stride_in", ctypes.c_uint64),
                ("seed", ctypes.c_uint64), ("keep_outputs", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("t_submit_ns", ctypes.c_int64), ("t_first_launch_ns", ctypes.c_int64), ("t_done_ns", ctypes.c_int64),
                ("parts", ctypes.c_uint32), ("parts_done", ctypes.c_uint32), ("status", ctypes.c_int32),
                ("reserved", ctypes.c_uint32)]


class PartInfo(ctypes.Structure):
    _fields_ = [("outputs", ctypes.c_void_p), ("stride_out", ctypes.c_uint64), ("first_col", ctypes.c_uint32),
                ("num_cols", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("requests_done", ctypes.c_uint64), ("parts_done", ctypes.c_uint64),
                ("batches_launched", ctypes.c_uint64), ("max_batches_in_flight_seen", ctypes.c_uint32),
                ("pending_parts", ctypes.c_uint32), ("arena_high_water_bytes", ctypes.c_uint64),
                ("arena_bytes_in_use", ctypes.c_uint64)]


HIP_ERROR_OUT_OF_MEMORY = 2


class OutOfBudget(RuntimeError):
    """The request cannot fit in the server's HBM budget."""


def _check(rc, what):
    if rc == 0:
        return
    if rc == BC_ERROR_INVALID_VALUE:
        raise AssertionError(f"{what}: argument assertion failed")
    if rc == HIP_ERROR_OUT_OF_MEMORY:
        raise OutOfBudget(f"{what}: request does not fit in the HBM budget")
    raise HipError(rc, what)


def fill_seeded(out, log_n, num_cols, stride, seed, first_col=0, stream=None):
    """Seeded raw u64 inputs on the device (the oracle's fill_raw stream)."""
    import torch
    s = (stream or torch.cuda.current_stream()).cuda_stream
    _check(lib().bc_fill_seeded(out.data_ptr(), log_n, num_cols, stride, seed, first_col, s), "fill_seeded")


def _dict(st):
    return {name: getattr(st, name) for name, _ in st._fields_ if name != "reserved"}


class Server:
    """A serving loop with its own HBM arena of `budget_bytes`."""

    def __init__(self, budget_bytes, max_part_elems=0, max_batches_in_flight=0, max_parts_per_batch=0):
        cfg = Config(budget_bytes, max_part_elems, max_batches_in_flight, max_parts_per_batch)
        h = ctypes.c_void_p()
        _check(lib().bc_server_create(ctypes.byref(cfg), ctypes.byref(h)), "server_create")
        self._h = h

    def submit(self, op, log_n, num_cols, log_blowup=0, seed=0, inputs=None, stride_in=None, keep_outputs=False):
        """Queues a request and returns its id.  inputs: device tensor of
        columns (stride_in apart) that must stay alive until the request
        completes, or None for seeded inputs."""
        req = _Request(op, log_n, num_cols, log_blowup, inputs.data_ptr() if inputs is not None else None,
                       stride_in or (1 << log_n), seed, 1 if keep_outputs else 0, 0)
        rid = ctypes.c_uint64()
        _check(lib().bc_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)), "server_submit")
        return rid.value

    def wait(self, rid):
        _check(lib().bc_server_wait(self._h, rid), "server_wait")

    def info(self, rid):
        st = RequestInfo()
        _check(lib().bc_server_request_info(self._h, rid, ctypes.byref(st)), "server_request_info")
        return _dict(st)

    def parts(self, rid):
        """Output locations of a completed keep_outputs request."""
        out = []
        for k in range(self.info(rid)["parts"]):
            p = PartInfo()
            _check(lib().bc_server_part(self._h, rid, k, ctypes.byref(p)), "server_part")
            out.append(_dict(p))
        return out

    def release(self, rid):
        _check(lib().bc_server_release(self._h, rid), "server_release")

    def drain(self):
        _check(lib().bc_server_drain(self._h), "server_drain")

    def stats(self):
        st = Stats()
        _check(lib().bc_server_stats_get(self._h, ctypes.byref(st)), "server_stats")
        return _dict(st)

    def destroy(self):
        if self._h is not None:
            _check(lib().bc_server_destroy(self._h), "server_destroy")
            self._h = None
