"""Coset low-degree extension over the batched NTT (C ABI: bc_batch_lde).

Not a reference function: it chains the reference's inverse transform and
per-coset forward transforms (src/ntt.rs:176-177 states that intent) so a
whole LDE is one call, and offers a captured hipGraph for fixed shapes.
"""
import ctypes

from ._lib import BC_ERROR_INVALID_VALUE, HipError, lib
from .ntt import _slice, _stream


def _check(rc, what):
    if rc == 0:
        return
    if rc == BC_ERROR_INVALID_VALUE:
        raise AssertionError(f"{what}: argument assertion failed")
    raise HipError(rc, what)


def batch_lde(inputs, coeffs, outputs, log_n, num_cols, stride_in, stride_coeffs, stride_out, log_blowup,
              bitrev_inputs, stream=None):
    """Evaluations on H -> coefficients (into `coeffs`, may be `inputs`) ->
    2^log_blowup coset blocks per output column (see include/boojum_ntt.h)."""
    ip, il = _slice(inputs)
    cp, cl = _slice(coeffs)
    op, ol = _slice(outputs)
    n = 1 << log_n
    assert il >= (num_cols - 1) * stride_in + n if num_cols else True
    assert cl >= (num_cols - 1) * stride_coeffs + n if num_cols else True
    assert ol >= (num_cols - 1) * stride_out + (n << log_blowup) if num_cols else True
    rc = lib().bc_batch_lde(ip, cp, op, log_n, num_cols, stride_in, stride_coeffs, stride_out, log_blowup,
                            bool(bitrev_inputs), _stream(inputs, stream))
    _check(rc, "batch_lde")


class LdeGraph:
    """A captured bc_batch_lde for fixed buffers and shape."""

    def __init__(self, inputs, coeffs, outputs, log_n, num_cols, stride_in, stride_coeffs, stride_out, log_blowup,
                 bitrev_inputs):
        ip, _ = _slice(inputs)
        cp, _ = _slice(coeffs)
        op, _ = _slice(outputs)
        self._keep = (inputs, coeffs, outputs)  # buffers must outlive the graph
        h = ctypes.c_void_p()
        rc = lib().bc_lde_graph_create(ip, cp, op, log_n, num_cols, stride_in, stride_coeffs, stride_out,
                                       log_blowup, bool(bitrev_inputs), ctypes.byref(h))
        _check(rc, "lde_graph_create")
        self._h = h

    def launch(self, stream=None):
        _check(lib().bc_graph_launch(self._h, _stream(self._keep[0], stream)), "graph_launch")

    def destroy(self):
        if self._h is not None:
            _check(lib().bc_graph_destroy(self._h), "graph_destroy")
            self._h = None
