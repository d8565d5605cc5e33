"""GPU parity of the coset-LDE pipeline (bc_batch_lde) and its hipGraph.

Expected values come from the oracle: coefficients = inverse transform of the
evaluations; coset c = forward coset transform (log_ext = blowup, coset c) of
the coefficients, i.e. the reference's own call sequence (src/ntt.rs:527-583
defines each step)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to("cuda")


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def expected_lde(evals, log_n, k, bitrev_inputs):
    n = 1 << log_n
    coeffs = oracle.ntt_column(evals, log_n, bitrev_inputs, True)
    out = np.zeros(n << k, dtype=np.uint64)
    for c in range(1 << k):
        if bitrev_inputs:
            blk = oracle.bitrev(c, k)
            out[blk * n:(blk + 1) * n] = oracle.ntt_column(coeffs, log_n, False, False, k, c)
        else:
            out[c * n:(c + 1) * n] = oracle.ntt_column(coeffs, log_n, True, False, k, c)
    return coeffs, out


@pytest.mark.parametrize("log_n,k,bitrev", [(3, 1, True), (8, 2, True), (10, 2, False), (13, 3, True),
                                             (16, 1, False), (17, 2, True), (14, 1, True), (15, 4, True),
                                             (21, 2, True)])
def test_batch_lde_matches_oracle(gpu_ctx, log_n, k, bitrev):
    from boojum_cuda_amd import lde
    n, cols = 1 << log_n, 3
    stride_in, stride_out = n + 8, (n << k) + 16
    evals = np.zeros(cols * stride_in, dtype=np.uint64)
    for c in range(cols):
        evals[c * stride_in: c * stride_in + n] = oracle.fill_raw(n, 31 + log_n, c)
    d_in = to_dev(evals)
    d_co = to_dev(np.zeros(cols * n, dtype=np.uint64))
    d_out = to_dev(np.zeros(cols * stride_out, dtype=np.uint64))
    lde.batch_lde(d_in, d_co, d_out, log_n, cols, stride_in, n, stride_out, k, bitrev)
    co, out = to_host(d_co), to_host(d_out)
    for c in range(cols):
        want_co, want = expected_lde(evals[c * stride_in: c * stride_in + n], log_n, k, bitrev)
        assert np.array_equal(co[c * n:(c + 1) * n], want_co)
        assert np.array_equal(out[c * stride_out: c * stride_out + (n << k)], want), (log_n, k, bitrev, c)


def test_bitreversed_lde_is_the_full_extension(gpu_ctx):
    # With bit-reversed input, the output column is bitreverse(f(g * w^m)),
    # m < n 2^k: the padded size-n 2^k transform of the coefficients on g H.
    from boojum_cuda_amd import lde
    log_n, k = 9, 2
    n = 1 << log_n
    evals = oracle.fill_uniform(n, 5)
    d_in, d_co, d_out = to_dev(evals), to_dev(np.zeros(n, np.uint64)), to_dev(np.zeros(n << k, np.uint64))
    lde.batch_lde(d_in, d_co, d_out, log_n, 1, n, n, n << k, k, True)
    co, out = to_host(d_co), to_host(d_out)
    g = 7
    P = oracle.P
    w = oracle.domain_generator(log_n + k)
    natural = oracle.bitreverse(out)
    for m in [0, 1, 2, 3, 5, 777, (n << k) - 1]:
        assert int(natural[m]) == oracle.horner(co, g * pow(w, m, P) % P)


def test_lde_in_place_coefficients(gpu_ctx):
    from boojum_cuda_amd import lde
    log_n, k, cols = 12, 2, 2
    n = 1 << log_n
    evals = np.concatenate([oracle.fill_raw(n, 3, c) for c in range(cols)])
    d = to_dev(evals)
    d_out = to_dev(np.zeros(cols * (n << k), dtype=np.uint64))
    lde.batch_lde(d, d, d_out, log_n, cols, n, n, n << k, k, True)
    out = to_host(d_out)
    for c in range(cols):
        _, want = expected_lde(evals[c * n:(c + 1) * n], log_n, k, True)
        assert np.array_equal(out[c * (n << k):(c + 1) * (n << k)], want)


def test_lde_graph_replays_match_eager(gpu_ctx):
    import torch
    from boojum_cuda_amd import lde, ntt
    log_n, k, cols = 14, 2, 4
    n = 1 << log_n
    evals = np.concatenate([oracle.fill_uniform(n, 9, c) for c in range(cols)])
    d_in = to_dev(evals)
    d_co = to_dev(np.zeros(cols * n, dtype=np.uint64))
    d_out = to_dev(np.zeros(cols * (n << k), dtype=np.uint64))
    lde.batch_lde(d_in, d_co, d_out, log_n, cols, n, n, n << k, k, True)
    eager = to_host(d_out).copy()
    g = lde.LdeGraph(d_in, d_co, d_out, log_n, cols, n, n, n << k, k, True)
    scratch = to_dev(np.zeros(1 << 16, dtype=np.uint64))
    try:
        for rep in range(3):
            d_out.zero_()
            # unrelated eager work between replays
            ntt.batch_ntt_in_place(scratch, 16, 1, 0, 1 << 16, rep % 2 == 1, False, 0, 0)
            torch.cuda.synchronize()
            g.launch()
            assert np.array_equal(to_host(d_out), eager), f"replay {rep}"
    finally:
        g.destroy()
