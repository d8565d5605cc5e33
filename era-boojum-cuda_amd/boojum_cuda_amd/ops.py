"""Order conversion and domain powers around the NTT, mirroring the
reference's `ops_complex` functions of the same names
(src/ops_complex.rs:29-86, 207-330) over the C ABI.  Tensors are HIP torch
tensors of int64/uint64; a slice's length is the element count, as a
DeviceSlice's is in the reference."""
from ._lib import BC_ERROR_INVALID_VALUE, HipError, lib


def _check(rc, what):
    if rc == 0:
        return
    if rc == BC_ERROR_INVALID_VALUE:
        raise AssertionError(f"{what}: argument assertion failed")
    raise HipError(rc, what)


def _stream(stream):
    import torch
    return (stream or torch.cuda.current_stream()).cuda_stream


def bit_reverse(src, dst, rows, cols, src_stride=None, dst_stride=None, stream=None):
    """dst[brev(i)] = src[i] for each of `cols` columns of `rows` points."""
    _check(lib().bc_bit_reverse(src.data_ptr(), src_stride or rows, dst.data_ptr(), dst_stride or rows, rows, cols,
                                _stream(stream)), "bit_reverse")


def bit_reverse_in_place(values, rows, cols, stride=None, stream=None):
    _check(lib().bc_bit_reverse_in_place(values.data_ptr(), stride or rows, rows, cols, _stream(stream)),
           "bit_reverse_in_place")


def get_powers_of_w(log_degree, offset, inverse, bit_reverse, result, stream=None):
    _check(lib().bc_get_powers_of_w(log_degree, offset, inverse, bit_reverse, result.data_ptr(), result.numel(),
                                    _stream(stream)), "get_powers_of_w")


def get_powers_of_g(log_degree, offset, inverse, bit_reverse, result, stream=None):
    _check(lib().bc_get_powers_of_g(log_degree, offset, inverse, bit_reverse, result.data_ptr(), result.numel(),
                                    _stream(stream)), "get_powers_of_g")


def omega_shift(src, log_degree, offset, inverse, shift, dst, stream=None):
    assert src.numel() == dst.numel(), "omega_shift: src and dst lengths differ"  # src/ops_complex.rs:233
    _check(lib().bc_omega_shift(src.data_ptr(), log_degree, offset, inverse, shift, dst.data_ptr(), dst.numel(),
                                _stream(stream)), "omega_shift")


def omega_shift_in_place(values, log_degree, offset, inverse, shift, stream=None):
    _check(lib().bc_omega_shift_in_place(values.data_ptr(), log_degree, offset, inverse, shift, values.numel(),
                                         _stream(stream)), "omega_shift_in_place")
