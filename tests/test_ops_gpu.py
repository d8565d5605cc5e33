"""GPU parity of the order conversion and domain powers around the NTT
(bit_reverse, get_powers_of_w/g, omega_shift), written like the reference's
own tests (src/ops_complex.rs:1366-1431, 1522-1672): closed forms checked
element by element, bit-exact."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

P = 0xFFFFFFFF00000001
G = 7


def brev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).to("cuda")


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def raw_matrix(cols, rows, stride, seed):
    m = np.zeros(cols * stride, dtype=np.uint64)
    for c in range(cols):
        m[c * stride:c * stride + rows] = oracle.fill_raw(rows, seed, c)
    return m


@pytest.mark.parametrize("in_place", [False, True])
def test_bit_reverse_reference_shape(gpu_ctx, in_place):
    """src/ops_complex.rs:1620-1672: 16 columns x 2^16, dst[brev(i)] = src[i]."""
    from boojum_cuda_amd import ops
    log_rows, cols = 16, 16
    rows = 1 << log_rows
    h = raw_matrix(cols, rows, rows, 5)
    perm = np.array([brev(i, log_rows) for i in range(rows)], dtype=np.int64)
    if in_place:
        d = to_dev(h)
        ops.bit_reverse_in_place(d, rows, cols)
        got = to_host(d)
    else:
        src, dst = to_dev(h), to_dev(np.zeros_like(h))
        ops.bit_reverse(src, dst, rows, cols)
        got = to_host(dst)
    for c in range(cols):
        want = np.empty(rows, dtype=np.uint64)
        want[perm] = h[c * rows:(c + 1) * rows]
        assert np.array_equal(got[c * rows:(c + 1) * rows], want), c


@pytest.mark.parametrize("log_rows", [0, 1, 5, 9, 10, 11, 12, 13, 21])
@pytest.mark.parametrize("in_place", [False, True])
def test_bit_reverse_sizes_strides(gpu_ctx, log_rows, in_place):
    """Both kernels (direct below 2^10, tiled from 2^10), padded strides,
    untouched padding, raw (non-canonical) values copied unchanged."""
    from boojum_cuda_amd import ops
    rows = 1 << log_rows
    cols = 3 if log_rows < 21 else 2
    sin, sout = rows + 7, rows + 3
    h = raw_matrix(cols, rows, sin, 9)
    h[h.size - 1] = 0xFFFFFFFFFFFFFFFF  # padding sentinel / raw value
    want_cols = []
    perm = np.array([brev(i, log_rows) for i in range(rows)], dtype=np.int64)
    for c in range(cols):
        w = np.empty(rows, dtype=np.uint64)
        w[perm] = h[c * sin:c * sin + rows]
        want_cols.append(w)
    if in_place:
        d = to_dev(h)
        ops.bit_reverse_in_place(d, rows, cols, stride=sin)
        got, stride = to_host(d), sin
        pad_ref = h
    else:
        sentinel = np.full(cols * sout, 0x1234, dtype=np.uint64)
        src, dst = to_dev(h), to_dev(sentinel)
        ops.bit_reverse(src, dst, rows, cols, src_stride=sin, dst_stride=sout)
        got, stride = to_host(dst), sout
        pad_ref = sentinel
    for c in range(cols):
        assert np.array_equal(got[c * stride:c * stride + rows], want_cols[c]), c
        assert np.array_equal(got[c * stride + rows:(c + 1) * stride], pad_ref[c * stride + rows:(c + 1) * stride])


def test_bit_reverse_same_buffer_is_in_place(gpu_ctx):
    from boojum_cuda_amd import ops
    rows, cols = 1 << 14, 2
    h = raw_matrix(cols, rows, rows, 3)
    d = to_dev(h)
    ops.bit_reverse(d, d, rows, cols)
    got = to_host(d)
    perm = np.array([brev(i, 14) for i in range(rows)], dtype=np.int64)
    for c in range(cols):
        want = np.empty(rows, dtype=np.uint64)
        want[perm] = h[c * rows:(c + 1) * rows]
        assert np.array_equal(got[c * rows:(c + 1) * rows], want)


def test_bit_reverse_rejects_bad_rows(gpu_ctx):
    import torch
    from boojum_cuda_amd import ops
    x = torch.zeros(64, dtype=torch.int64, device="cuda")
    with pytest.raises(AssertionError):
        ops.bit_reverse(x, x.clone(), 48, 1)  # not a power of two
    with pytest.raises(AssertionError):
        ops.bit_reverse(x, x.clone(), 32, 1, src_stride=16)  # stride below rows


def powers(base, count, start=0):
    out, v = [], pow(base, start, P)
    for _ in range(count):
        out.append(v)
        v = v * base % P
    return np.array(out, dtype=np.uint64)


@pytest.mark.parametrize("inverse", [False, True])
def test_get_powers_of_w(gpu_ctx, inverse):
    """src/ops_complex.rs:1366-1414: two halves, offsets 0 and n/2."""
    import torch
    from boojum_cuda_amd import ops
    log_degree = 16
    n = 1 << log_degree
    res = torch.zeros(n, dtype=torch.int64, device="cuda")
    ops.get_powers_of_w(log_degree, 0, inverse, False, res[:n // 2])
    ops.get_powers_of_w(log_degree, n // 2, inverse, False, res[n // 2:])
    w = oracle.domain_generator(log_degree)
    if inverse:
        w = pow(w, P - 2, P)
    assert np.array_equal(to_host(res), powers(w, n))


@pytest.mark.parametrize("inverse", [False, True])
def test_get_powers_of_g(gpu_ctx, inverse):
    """src/ops_complex.rs:1423-1431: generator g^(2^(24 - log_degree))."""
    import torch
    from boojum_cuda_amd import ops
    log_degree = 16
    n = 1 << log_degree
    res = torch.zeros(n, dtype=torch.int64, device="cuda")
    ops.get_powers_of_g(log_degree, 0, inverse, False, res[:n // 2])
    ops.get_powers_of_g(log_degree, n // 2, inverse, False, res[n // 2:])
    gen = pow(G, 1 << (24 - log_degree), P)
    if inverse:
        gen = pow(gen, P - 2, P)
    assert np.array_equal(to_host(res), powers(gen, n))


@pytest.mark.parametrize("which", ["w", "g"])
def test_get_powers_bit_reversed_and_edges(gpu_ctx, which):
    import torch
    from boojum_cuda_amd import ops
    fn = ops.get_powers_of_w if which == "w" else ops.get_powers_of_g
    log_degree = 10
    n = 1 << log_degree
    base = oracle.domain_generator(log_degree) if which == "w" else pow(G, 1 << (24 - log_degree), P)
    res = torch.zeros(n - 3, dtype=torch.int64, device="cuda")
    fn(log_degree, 3, False, True, res)
    want = np.array([pow(base, brev(i + 3, log_degree), P) for i in range(n - 3)], dtype=np.uint64)
    assert np.array_equal(to_host(res), want)
    one = torch.zeros(1, dtype=torch.int64, device="cuda")
    fn(0, 0, False, False, one)  # log_degree 0: the single power 1
    assert to_host(one)[0] == 1
    fn(log_degree, 0, False, False, torch.zeros(0, dtype=torch.int64, device="cuda"))  # empty: no-op
    if which == "g":
        with pytest.raises(AssertionError):
            fn(25, 0, False, False, one)  # beyond the reference's 24-bit g tables


@pytest.mark.parametrize("in_place", [False, True])
@pytest.mark.parametrize("inverse", [False, True])
def test_omega_shift(gpu_ctx, in_place, inverse):
    """src/ops_complex.rs:1522-1618: SHIFT = 42, two halves; raw inputs."""
    import torch
    from boojum_cuda_amd import ops
    log_degree, shift = 16, 42
    n = 1 << log_degree
    h = oracle.fill_raw(n, 17, 0)
    w = oracle.domain_generator(log_degree)
    if inverse:
        w = pow(w, P - 2, P)
    step = pow(w, shift, P)
    want = np.array([int(x) % P * int(f) % P for x, f in zip(h, powers(step, n))], dtype=np.uint64)
    if in_place:
        d = to_dev(h)
        ops.omega_shift_in_place(d[:n // 2], log_degree, 0, inverse, shift)
        ops.omega_shift_in_place(d[n // 2:], log_degree, n // 2, inverse, shift)
        got = to_host(d)
    else:
        src, dst = to_dev(h), torch.zeros(n, dtype=torch.int64, device="cuda")
        ops.omega_shift(src[:n // 2], log_degree, 0, inverse, shift, dst[:n // 2])
        ops.omega_shift(src[n // 2:], log_degree, n // 2, inverse, shift, dst[n // 2:])
        got = to_host(dst)
    assert np.array_equal(got, want)
