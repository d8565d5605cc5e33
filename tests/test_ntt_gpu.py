"""GPU parity of the HIP batched NTT against the CPU oracle, through the C ABI.

Mirrors the reference's `correctness()` (src/ntt.rs:391-611): for every log_n
it runs the four variants (n2b/b2n x in-place/out-of-place) on columns that
sit at offset n inside a stride of 2 * n_max, and compares every element,
bit-exact, with the oracle.  It also checks that nothing outside the
[offset, offset + n) window of each column is written.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SENTINEL = np.uint64(0xDEADBEEFCAFEF00D)


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to("cuda")


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def bitrev_chunks(buf, n):
    out = buf.copy()
    for s in range(0, buf.size, n):
        out[s:s + n] = oracle.bitreverse(buf[s:s + n])
    return out


def correctness(log_ns, inverse, log_ext, coset, num_ntts, seed=1):
    from boojum_cuda_amd import ntt
    n_max = 1 << max(log_ns)
    stride = 2 * n_max
    size = stride * num_ntts
    inputs = oracle.fill_raw(size, seed, 0)
    for log_n in log_ns:
        n = 1 << log_n
        off = n
        inputs_br = bitrev_chunks(inputs, n)
        for bitrev in (False, True):
            src = inputs_br if bitrev else inputs
            want = np.empty((num_ntts, n), dtype=np.uint64)
            for c in range(num_ntts):
                want[c] = oracle.ntt_column(src[c * stride + off: c * stride + off + n], log_n, bitrev, inverse,
                                            log_ext, coset)
            # in place
            d = to_dev(src)
            ntt.batch_ntt_in_place(d, log_n, num_ntts, off, stride, bitrev, inverse, log_ext, coset)
            got_ip = to_host(d)
            # out of place
            d_in = to_dev(src)
            d_out = to_dev(np.full(size, SENTINEL, dtype=np.uint64))
            ntt.batch_ntt_out_of_place(d_in, d_out, log_n, num_ntts, off, off, stride, stride, bitrev, inverse,
                                       log_ext, coset)
            got_oop = to_host(d_out)
            assert np.array_equal(to_host(d_in), src), "input modified by out-of-place call"
            for c in range(num_ntts):
                lo, hi = c * stride + off, c * stride + off + n
                for name, got in (("in-place", got_ip), ("out-of-place", got_oop)):
                    col = got[lo:hi]
                    if not np.array_equal(col, want[c]):
                        bad = np.nonzero(col != want[c])[0]
                        pytest.fail(f"{name} {'b2n' if bitrev else 'n2b'} inv={inverse} ext={log_ext} "
                                    f"coset={coset} log_n={log_n} col={c}: {bad.size} mismatches, first at "
                                    f"{bad[0]}: got {col[bad[0]]:#x} want {want[c][bad[0]]:#x}")
                # untouched outside the window
                assert np.array_equal(got_ip[c * stride: lo], src[c * stride: lo])
                assert np.array_equal(got_ip[hi: (c + 1) * stride], src[hi: (c + 1) * stride])
                assert np.all(got_oop[c * stride: lo] == SENTINEL)
                assert np.all(got_oop[hi: (c + 1) * stride] == SENTINEL)


# --- the reference's test matrix (src/ntt.rs:613-663) ------------------------

@pytest.mark.parametrize("chunk", [range(1, 9), range(9, 13), range(13, 17)])
def test_correctness_batch_ntt_fwd(gpu_ctx, chunk):
    correctness(list(chunk), False, 0, 0, 2 * 8 + 3)


@pytest.mark.parametrize("chunk", [range(1, 9), range(9, 17)])
def test_correctness_batch_ntt_inv(gpu_ctx, chunk):
    correctness(list(chunk), True, 0, 0, 1)


@pytest.mark.parametrize("inverse", [False, True])
def test_correctness_batch_lde(gpu_ctx, inverse):
    correctness(list(range(1, 17)), inverse, 3, 3, 3)


@pytest.mark.parametrize("log_n", [17, 18, 19, 20, 21, 22])
def test_correctness_batch_ntt_fwd_large(gpu_ctx, log_n):
    correctness([log_n], False, 0, 0, 3)


@pytest.mark.parametrize("log_n", [17, 20, 21, 23])
def test_correctness_batch_ntt_inv_large(gpu_ctx, log_n):
    correctness([log_n], True, 0, 0, 1)


@pytest.mark.parametrize("log_n", [17, 18, 19, 20, 21])
@pytest.mark.parametrize("inverse", [False, True])
def test_correctness_batch_lde_large(gpu_ctx, log_n, inverse):
    correctness([log_n], inverse, 3, 3, 3)


@pytest.mark.parametrize("log_n,inverse,log_ext,coset", [(24, False, 0, 0), (24, True, 0, 0), (25, False, 0, 0),
                                                          (26, True, 0, 0), (24, False, 2, 3), (23, True, 3, 5)])
def test_correctness_max_sizes(gpu_ctx, log_n, inverse, log_ext, coset):
    """log_n 24 is the reference's limit (OMEGA_LOG_ORDER, src/context.rs:12,
    src/ntt.rs:193); single columns and coset extensions up to 2^26 points are
    the sizes the ragged and serving configurations need."""
    correctness([log_n], inverse, log_ext, coset, 1)


@pytest.mark.parametrize("k,c", [(1, 0), (1, 1), (2, 0), (2, 3), (4, 9)])
def test_lde_cosets(gpu_ctx, k, c):
    correctness([5, 12, 14], False, k, c, 2)
    correctness([5, 12, 14], True, k, c, 2)


# --- committed fixtures -------------------------------------------------------

def test_golden_fixtures(gpu_ctx):
    from boojum_cuda_amd import ntt
    import gen_golden_keys as gk
    data = np.load(os.path.join(HERE, "golden", "ntt_golden.npz"))
    for log_n in gk.LOGS:
        inp = data[f"in_L{log_n}"]
        cols, n = inp.shape
        for bitrev in (False, True):
            for inverse in (False, True):
                for k, c in gk.LDE:
                    d_in = to_dev(inp.reshape(-1))
                    d_out = to_dev(np.zeros(cols * n, dtype=np.uint64))
                    ntt.batch_ntt_out_of_place(d_in, d_out, log_n, cols, 0, 0, n, n, bitrev, inverse, k, c)
                    got = to_host(d_out).reshape(cols, n)
                    assert np.array_equal(got, data[gk.case_key(log_n, bitrev, inverse, k, c)]), \
                        (log_n, bitrev, inverse, k, c)


def test_golden_checksums(gpu_ctx):
    from boojum_cuda_amd import ntt
    with open(os.path.join(HERE, "golden", "ntt_checksums.json")) as f:
        cases = json.load(f)
    for cs in cases:
        n = 1 << cs["log_n"]
        inp = np.concatenate([oracle.fill_uniform(n, cs["seed"], c) for c in range(cs["num_cols"])])
        d = to_dev(inp)
        ntt.batch_ntt_in_place(d, cs["log_n"], cs["num_cols"], 0, n, cs["bitrev_inputs"], cs["inverse"], 0, 0)
        got = to_host(d)
        assert hashlib.sha256(got.astype("<u8").tobytes()).hexdigest() == cs["sha256"], cs


# --- full-size properties (256 x 2^20, the batched-NTT benchmark shape) -------

def test_full_size_round_trip_and_sampled_parity(gpu_ctx):
    import torch
    from boojum_cuda_amd import ntt
    log_n, cols = 20, 256
    n = 1 << log_n
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    x = torch.randint(-2**63, 2**63 - 1, (cols * n,), dtype=torch.int64, device="cuda", generator=g)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    xh = x.cpu().numpy().view(np.uint64)
    for bitrev in (False, True):
        ntt.batch_ntt_out_of_place(x, y, log_n, cols, 0, 0, n, n, bitrev, False, 0, 0)
        ntt.batch_ntt_out_of_place(y, z, log_n, cols, 0, 0, n, n, not bitrev, True, 0, 0)
        zh = to_host(z)
        assert np.array_equal(zh, oracle.canon(xh)), "forward then inverse is not the identity"
        yh = to_host(y)
        for c in (0, 77, 255):
            want = oracle.ntt_column(xh[c * n:(c + 1) * n], log_n, bitrev, False)
            assert np.array_equal(yh[c * n:(c + 1) * n], want), (bitrev, c)


# --- latency path: 1..4 columns of 2^16 (one-wave workgroups) ----------------

@pytest.mark.parametrize("num_ntts", [1, 2, 4])
@pytest.mark.parametrize("inverse", [False, True])
def test_latency_path_2e16_matches_oracle(gpu_ctx, num_ntts, inverse):
    correctness([16], inverse, 0, 0, num_ntts, seed=11 + num_ntts)


def test_latency_path_equals_batched_path(gpu_ctx):
    """4 columns take the latency path, 5 columns the 4096-point pass kernels:
    the first four columns must come out identical, in all four variants."""
    import torch
    from boojum_cuda_amd import ntt
    n = 1 << 16
    x = to_dev(oracle.fill_raw(5 * n, 21, 0))
    for bitrev in (False, True):
        for inverse in (False, True):
            y4 = torch.empty(4 * n, dtype=torch.int64, device="cuda")
            y5 = torch.empty(5 * n, dtype=torch.int64, device="cuda")
            ntt.batch_ntt_out_of_place(x, y4, 16, 4, 0, 0, n, n, bitrev, inverse, 0, 0)
            ntt.batch_ntt_out_of_place(x, y5, 16, 5, 0, 0, n, n, bitrev, inverse, 0, 0)
            assert np.array_equal(to_host(y4), to_host(y5)[:4 * n]), (bitrev, inverse)


def test_latency_path_graph_replays_match_eager(gpu_ctx):
    """The latency path captured in a graph (forward n2b x -> y, then b2n
    inverse y -> z) gives the eager results on every replay, with eager work
    of other sizes between the replays."""
    import torch
    from boojum_cuda_amd import ntt
    n = 1 << 16
    host = oracle.fill_raw(n, 31, 0)
    x = to_dev(host)
    y, z = torch.empty_like(x), torch.empty_like(x)
    ntt.batch_ntt_out_of_place(x, y, 16, 1, 0, 0, n, n, False, False, 0, 0)
    ntt.batch_ntt_out_of_place(y, z, 16, 1, 0, 0, n, n, True, True, 0, 0)
    want_y, want_z = to_host(y).copy(), to_host(z).copy()
    assert np.array_equal(want_y, oracle.ntt_column(host, 16, False, False))
    assert np.array_equal(want_z, oracle.canon(host))
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            ntt.batch_ntt_out_of_place(x, y, 16, 1, 0, 0, n, n, False, False, 0, 0)
            ntt.batch_ntt_out_of_place(y, z, 16, 1, 0, 0, n, n, True, True, 0, 0)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    other = to_dev(oracle.fill_raw(4 << 12, 5, 0))
    other_out = torch.empty_like(other)
    for k in range(4):
        y.zero_()
        z.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(to_host(y), want_y), k
        assert np.array_equal(to_host(z), want_z), k
        ntt.batch_ntt_out_of_place(other, other_out, 12, 4, 0, 0, 1 << 12, 1 << 12, False, k % 2 == 1, 0, 0)
        torch.cuda.synchronize()
