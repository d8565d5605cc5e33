"""GPU parity of barycentric evaluation (src/barycentric.rs:83-280), written
like the reference's own tests (src/barycentric.rs:304-680): Lagrange
coefficients for log_count 0..19 against the oracle, and batch_eval of coset
evaluations (monomial coefficients through the forward coset NTT) against
Horner's rule at x, for num_polys 1, 3, 32, 33 and log_count 0..16, all three
kinds, bit-exact."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

P = oracle.P
COSET = 7  # BF::multiplicative_generator(), as the reference's tests use


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).to("cuda")


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def rand_x(rng, ext):
    return (int(rng.integers(0, P, dtype=np.uint64)), int(rng.integers(0, P, dtype=np.uint64))) if ext else int(rng.integers(0, P, dtype=np.uint64))


def x_dev(x):
    return to_dev(np.array(list(x) if isinstance(x, tuple) else [x], dtype=np.uint64))


def precompute(x, log_count, bit_reverse):
    import torch
    from boojum_cuda_amd import barycentric as bary
    ext = isinstance(x, tuple)
    w = 2 if ext else 1
    coeffs = torch.zeros(w << log_count, dtype=torch.int64, device="cuda")
    common = torch.zeros(w, dtype=torch.int64, device="cuda")
    bary.precompute_lagrange_coeffs(ext, x_dev(x), common, COSET, coeffs, bit_reverse)
    return coeffs, common


@pytest.mark.parametrize("ext", [False, True])
def test_precompute_lagrange_coeffs(gpu_ctx, ext):
    """src/barycentric.rs:304-392: log_count 0..19, bit-reversed, random x."""
    rng = np.random.default_rng(11 + ext)
    for log_count in range(20):
        x = rand_x(rng, ext)
        coeffs, common = precompute(x, log_count, True)
        want = oracle.bary_lagrange_coeffs(x, COSET, log_count, True)
        got = to_host(coeffs)
        assert np.array_equal(got, want), log_count
        cf = oracle.bary_common_factor(x, COSET, 1 << log_count)
        assert tuple(int(v) for v in to_host(common)) == (cf if ext else (cf,))


@pytest.mark.parametrize("ext", [False, True])
@pytest.mark.parametrize("log_count", [0, 3, 12, 13])
def test_precompute_natural_order_and_raw_inputs(gpu_ctx, ext, log_count):
    """Natural order; x and coset given as raw u64 representatives >= p."""
    import torch
    from boojum_cuda_amd import barycentric as bary
    rng = np.random.default_rng(log_count)
    x = rand_x(rng, ext)
    raw = tuple(v + P if v + P < 2**64 else v for v in (x if ext else (x,)))
    w = 2 if ext else 1
    coeffs = torch.zeros(w << log_count, dtype=torch.int64, device="cuda")
    common = torch.zeros(w, dtype=torch.int64, device="cuda")
    bary.precompute_lagrange_coeffs(ext, to_dev(np.array(raw, dtype=np.uint64)), common, COSET + P, coeffs, False)
    assert np.array_equal(to_host(coeffs), oracle.bary_lagrange_coeffs(x, COSET, log_count, False))


def coset_evaluations(coeffs_host, log_count, cols):
    """ys = forward NTT on the coset 7 omega_N^i (log_extension_degree 1,
    coset index 0), outputs bit-reversed, on the GPU as the reference's tests
    make them (src/barycentric.rs:419-432)."""
    from boojum_cuda_amd import ntt
    d = to_dev(coeffs_host)
    if log_count > 0:
        ntt.batch_ntt_in_place(d, log_count, cols, 0, 1 << log_count, False, False, 1, 0)
    return d


KINDS = {"base_at_base": 0, "base_at_ext": 1, "ext_at_ext": 2}


def eval_gpu(kind, ys, log_count, num_polys, coeffs, misalign=False):
    import torch
    from boojum_cuda_amd import barycentric as bary
    n = 1 << log_count
    out_w = 1 if kind == 0 else 2
    partial, final = bary.get_batch_eval_temp_storage_sizes(kind, n, num_polys)
    assert final == 0 and partial >= num_polys
    temp = torch.full((partial * out_w,), -1, dtype=torch.int64, device="cuda")
    evals = torch.zeros(num_polys * out_w, dtype=torch.int64, device="cuda")
    if misalign:  # 8-byte aligned views: the two-load path
        ys2 = torch.zeros(ys.numel() + 1, dtype=torch.int64, device="cuda")
        ys2[1:] = ys
        c2 = torch.zeros(coeffs.numel() + 1, dtype=torch.int64, device="cuda")
        c2[1:] = coeffs
        ys, coeffs = ys2[1:], c2[1:]
    bary.batch_eval(kind, ys, n, num_polys, coeffs, temp, None, evals)
    return to_host(evals)


def horner_all(kind, coeffs_host, n, num_polys, x):
    out = []
    for p in range(num_polys):
        if kind == 2:
            a0, a1 = coeffs_host[2 * p * n:(2 * p + 1) * n], coeffs_host[(2 * p + 1) * n:(2 * p + 2) * n]
        else:
            a0, a1 = coeffs_host[p * n:(p + 1) * n], None
        h = oracle.e2_horner(a0, a1, x)
        out.extend(h if kind else h[:1])
    return np.array(out, dtype=np.uint64)


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("num_polys", [1, 3, 32, 33])
def test_batch_eval_matches_horner(gpu_ctx, kind, num_polys):
    """src/barycentric.rs:393-680, log_count 0..16."""
    k = KINDS[kind]
    for log_count in range(17):
        n = 1 << log_count
        cols = num_polys * (2 if k == 2 else 1)
        seed = 1000 * k + 10 * num_polys + log_count
        coeffs_host = oracle.matrix_uniform(cols, n, seed)
        ys = coset_evaluations(coeffs_host, log_count, cols)
        x = rand_x(np.random.default_rng(seed), k != 0)
        L, _ = precompute(x, log_count, True)
        got = eval_gpu(k, ys, log_count, num_polys, L)
        want = horner_all(k, coeffs_host, n, num_polys, x)
        assert np.array_equal(got, want), (kind, num_polys, log_count)


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("log_count", [6, 7, 8, 10, 17, 18])
def test_batch_eval_raw_misaligned(gpu_ctx, kind, log_count):
    """Raw u64 column values (>= p allowed) and 8-byte-aligned buffers, against
    the oracle's sum over the same inputs."""
    k = KINDS[kind]
    n = 1 << log_count
    num_polys = 9
    cols = num_polys * (2 if k == 2 else 1)
    ys_host = oracle.matrix_uniform(cols, n, 77 + log_count, raw=True)
    x = rand_x(np.random.default_rng(log_count), k != 0)
    L, _ = precompute(x, log_count, True)
    L_host = to_host(L)
    want = oracle.bary_batch_eval(ys_host, k == 2, L_host, k != 0, n, num_polys)
    for misalign in (False, True):
        got = eval_gpu(k, to_dev(ys_host), log_count, num_polys, L, misalign=misalign)
        assert np.array_equal(got, want), misalign


@pytest.mark.parametrize("kind", list(KINDS))
def test_batch_eval_full_size_against_oracle(gpu_ctx, kind):
    """2^20 rows x 24 columns (48 planes for extension columns) against the
    oracle's sum, and linearity: eval(a + b) = eval(a) + eval(b)."""
    import torch
    k = KINDS[kind]
    log_count, num_polys = 20, 24
    n = 1 << log_count
    cols = num_polys * (2 if k == 2 else 1)
    ys_host = oracle.matrix_uniform(cols, n, 5)
    x = rand_x(np.random.default_rng(5), k != 0)
    L, _ = precompute(x, log_count, True)
    ys = to_dev(ys_host)
    got = eval_gpu(k, ys, log_count, num_polys, L)
    want = oracle.bary_batch_eval(ys_host, k == 2, to_host(L), k != 0, n, num_polys)
    assert np.array_equal(got, want)
    # linearity over a shifted copy (b = a rolled by one column)
    shift = n * (2 if k == 2 else 1)
    b_host = np.roll(ys_host, shift)
    e_b = eval_gpu(k, to_dev(b_host), log_count, num_polys, L)
    e_ab = eval_gpu(k, to_dev(add_mod(ys_host, b_host)), log_count, num_polys, L)
    assert np.array_equal(e_ab, add_mod(got, e_b))


def add_mod(a, b):
    """Elementwise a + b mod p for canonical uint64 arrays."""
    s = a + b  # wraps mod 2^64
    wrapped = s < a
    return np.where(wrapped, s + np.uint64(0xFFFFFFFF), np.where(s >= np.uint64(P), s - np.uint64(P), s))


def test_batch_eval_argument_checks(gpu_ctx):
    import torch
    from boojum_cuda_amd import barycentric as bary
    with pytest.raises(AssertionError):
        bary.get_batch_eval_temp_storage_sizes(0, 48, 1)  # stride not a power of two
    with pytest.raises(AssertionError):
        bary.get_batch_eval_temp_storage_sizes(3, 64, 1)  # unknown kind
    ys = torch.zeros(1024, dtype=torch.int64, device="cuda")
    L = torch.zeros(1024, dtype=torch.int64, device="cuda")
    evals = torch.zeros(1, dtype=torch.int64, device="cuda")
    small = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(AssertionError):
        bary.batch_eval(0, ys, 1024, 1, L, small, None, evals)  # partial scratch too small
    with pytest.raises(AssertionError):
        bary.precompute_lagrange_coeffs(False, small, small, COSET, torch.zeros(48, dtype=torch.int64,
                                                                               device="cuda"), True)
    # the C ABI refuses a short scratch by itself (1 = BC_ERROR_INVALID_VALUE)
    from boojum_cuda_amd._lib import lib
    rc = lib().bc_batch_eval(0, ys.data_ptr(), 1024, 1, L.data_ptr(), small.data_ptr(), 1, None, 0,
                             evals.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert rc == 1
