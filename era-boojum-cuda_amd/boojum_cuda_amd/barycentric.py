"""Barycentric evaluation on the coset domain {coset * omega_N^i}, mirroring
the reference's `barycentric` module (src/barycentric.rs:83-280) over the C
ABI.  Tensors are HIP torch tensors of int64/uint64.  Extension values
(F_p[u]/(u^2 - 7)) are u64 pairs: one element as (c0, c1); a vector or a
matrix column of N elements as a double vector (N c0 values, then N c1)."""
import ctypes

from ._lib import BC_ERROR_INVALID_VALUE, HipError, lib

# src/barycentric.rs:265-267
EVAL_BASE_AT_BASE = 0
EVAL_BASE_AT_EXT = 1
EVAL_EXT_AT_EXT = 2

# PrecomputeImpl selectors (src/barycentric.rs:140-141)
PRECOMPUTE_AT_BASE = False
PRECOMPUTE_AT_EXT = True


def _check(rc, what):
    if rc == 0:
        return
    if rc == BC_ERROR_INVALID_VALUE:
        raise AssertionError(f"{what}: argument assertion failed")
    raise HipError(rc, what)


def _stream(stream):
    import torch
    return (stream or torch.cuda.current_stream()).cuda_stream


def precompute_lagrange_coeffs(at_ext, x, common_factor_storage, coset, lagrange_coeffs, bit_reverse, stream=None):
    """src/barycentric.rs:83-127.  x / common_factor_storage: 1 (base) or 2
    (extension) elements; lagrange_coeffs: N (base) or 2N (extension) elements,
    N a power of two."""
    width = 2 if at_ext else 1
    assert x.numel() >= width and common_factor_storage.numel() >= width
    assert lagrange_coeffs.numel() % width == 0
    count = lagrange_coeffs.numel() // width
    _check(lib().bc_barycentric_precompute(bool(at_ext), x.data_ptr(), common_factor_storage.data_ptr(),
                                           coset % (1 << 64), lagrange_coeffs.data_ptr(), count, bool(bit_reverse),
                                           _stream(stream)), "precompute_lagrange_coeffs")


def get_batch_eval_temp_storage_sizes(kind, stride, num_polys):
    """src/barycentric.rs:196-214: (partial-sum elements of the coefficient
    type, bytes of the final-reduce scratch)."""
    partial, final = ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib().bc_batch_eval_temp_sizes(kind, stride, num_polys, ctypes.byref(partial), ctypes.byref(final)),
           "get_batch_eval_temp_storage_sizes")
    return partial.value, final.value


def batch_eval(kind, batch_ys, stride, num_polys, lagrange_coeffs, temp_storage_partial_reduce,
               temp_storage_final_reduce, evals, stream=None):
    """src/barycentric.rs:216-261.  temp_storage_partial_reduce holds the
    partial elements reported by get_batch_eval_temp_storage_sizes (two u64
    each for extension kinds); evals: num_polys u64 (base) or 2 * num_polys."""
    coef_width = 1 if kind == EVAL_BASE_AT_BASE else 2
    partial_elems, _ = get_batch_eval_temp_storage_sizes(kind, stride, num_polys)
    assert temp_storage_partial_reduce.numel() >= partial_elems * coef_width
    assert evals.numel() == num_polys * coef_width, "batch_eval: evals length"  # src/barycentric.rs:235
    tf = temp_storage_final_reduce
    _check(lib().bc_batch_eval(kind, batch_ys.data_ptr(), stride, num_polys, lagrange_coeffs.data_ptr(),
                               temp_storage_partial_reduce.data_ptr(), partial_elems,
                               tf.data_ptr() if tf is not None else None, tf.numel() if tf is not None else 0,
                               evals.data_ptr(), _stream(stream)), "batch_eval")
