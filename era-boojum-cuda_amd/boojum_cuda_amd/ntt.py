"""Mirror of boojum-cuda's `ntt` module (reference: src/ntt.rs).

Device slices are 1-D torch tensors of dtype int64 or uint64 on a HIP device
(the bits are the Goldilocks u64 words), standing in for era_cudart's
`DeviceSlice<GoldilocksField>`.  `stream` is a `torch.cuda.Stream` (default:
the current stream), standing in for `CudaStream`.  Argument violations raise
AssertionError before anything is launched, as the reference's `assert!`s
panic; launch failures raise HipError (the reference's `CudaResult::Err`).
"""
import ctypes

from ._lib import BC_ERROR_INVALID_VALUE, BC_ERROR_NOT_INITIALIZED, HipError, lib


def _slice(t):
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor on a HIP device")
    if t.dtype not in (torch.int64, torch.uint64) or t.dim() != 1 or not t.is_contiguous():
        raise TypeError("expected a contiguous 1-D int64/uint64 tensor")
    if not t.is_cuda:
        raise TypeError("expected a device tensor")
    return t.data_ptr(), t.numel()


def _stream(t, stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(t.device)
    return stream.cuda_stream


def _check(rc, what):
    if rc == 0:
        return
    if rc == BC_ERROR_INVALID_VALUE:
        raise AssertionError(f"{what}: argument assertion failed")
    if rc == BC_ERROR_NOT_INITIALIZED:
        raise HipError(rc, f"{what}: no Context on this device")
    raise HipError(rc, what)


def batch_ntt_internal(inputs_ptr, outputs_ptr, log_n, num_ntts, stride_between_input_arrays,
                       stride_between_output_arrays, bitrev_inputs, inverse, log_extension_degree, coset_index,
                       stream_handle):
    """src/ntt.rs:179-291 over raw device addresses (ints) and a raw stream handle."""
    rc = lib().bc_batch_ntt_internal(inputs_ptr, outputs_ptr, log_n, num_ntts, stride_between_input_arrays,
                                     stride_between_output_arrays, bool(bitrev_inputs), bool(inverse),
                                     log_extension_degree, coset_index, stream_handle)
    _check(rc, "batch_ntt_internal")


def batch_ntt_out_of_place(inputs_matrix, outputs_matrix, log_n, num_ntts, inputs_offset, outputs_offset,
                           stride_between_input_arrays, stride_between_output_arrays, bitrev_inputs, inverse,
                           log_extension_degree, coset_index, stream=None):
    """src/ntt.rs:293-335."""
    ip, il = _slice(inputs_matrix)
    op, ol = _slice(outputs_matrix)
    rc = lib().bc_batch_ntt_out_of_place(ip, il, op, ol, log_n, num_ntts, inputs_offset, outputs_offset,
                                         stride_between_input_arrays, stride_between_output_arrays,
                                         bool(bitrev_inputs), bool(inverse), log_extension_degree, coset_index,
                                         _stream(inputs_matrix, stream))
    _check(rc, "batch_ntt_out_of_place")


def batch_ntt_in_place(inputs_matrix, log_n, num_ntts, inputs_offset, stride_between_input_arrays, bitrev_inputs,
                       inverse, log_extension_degree, coset_index, stream=None):
    """src/ntt.rs:337-369."""
    ip, il = _slice(inputs_matrix)
    rc = lib().bc_batch_ntt_in_place(ip, il, log_n, num_ntts, inputs_offset, stride_between_input_arrays,
                                     bool(bitrev_inputs), bool(inverse), log_extension_degree, coset_index,
                                     _stream(inputs_matrix, stream))
    _check(rc, "batch_ntt_in_place")


class _PassInfo(ctypes.Structure):
    _fields_ = [("contiguous", ctypes.c_uint32), ("log_h", ctypes.c_uint32), ("log_block", ctypes.c_uint32),
                ("mode", ctypes.c_uint32), ("work_items", ctypes.c_uint64), ("positions", ctypes.c_uint64),
                ("groups", ctypes.c_uint64), ("pairs_per_workgroup", ctypes.c_uint32),
                ("workgroup_threads", ctypes.c_uint32)]


def describe_plan(log_n, num_ntts, bitrev_inputs=False, inverse=False, log_extension_degree=0, coset_index=0):
    """The HBM passes batch_ntt_internal launches for this shape, as dicts
    (host only, no device needed).  Empty for log_n < 8 (one generic pass)."""
    arr = (_PassInfo * 8)()
    n = ctypes.c_uint32()
    rc = lib().bc_ntt_describe_plan(log_n, num_ntts, bool(bitrev_inputs), bool(inverse), log_extension_degree,
                                    coset_index, arr, 8, ctypes.byref(n))
    _check(rc, "ntt_describe_plan")
    return [{name: getattr(arr[i], name) for name, _ in _PassInfo._fields_}
            for i in range(n.value)]
