"""Pins the CPU oracle (oracle/gl_oracle.c) against mathematics and the
committed fixtures.  CPU only."""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import oracle
import pyref

P = pyref.P
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
EDGE = [0, 1, 2, P - 1, P, P + 1, 2**64 - 1, 2**32 - 1, 2**32, 2**63, P - 2**32, 0xFFFFFFFF00000000]


def test_field_ops_edge_and_random():
    rng = random.Random(1)
    vals = EDGE + [rng.getrandbits(64) for _ in range(300)]
    for a in vals:
        for b in vals[:40]:
            assert oracle.add(a, b) == (a + b) % P
            assert oracle.sub(a, b) == (a - b) % P
            assert oracle.mul(a, b) == (a * b) % P
    for a in vals[:60]:
        if a % P:
            assert oracle.mul(oracle.inv(a), a) == 1


def test_generator_and_root_constants():
    # g = 7 generates F_p^*: p - 1 = 2^32 * 3 * 5 * 17 * 257 * 65537.
    assert (P - 1) == 2**32 * 3 * 5 * 17 * 257 * 65537
    for f in (2, 3, 5, 17, 257, 65537):
        assert pow(7, (P - 1) // f, P) != 1
    assert oracle.generator() == 7
    w = oracle.omega_2_32()
    assert w == 0x185629DCDA58878C == pyref.W32
    assert pow(w, 2**32, P) == 1 and pow(w, 2**31, P) == P - 1  # exact order 2^32
    for k in range(33):
        assert oracle.domain_generator(k) == pyref.omega(k)
    # every omega_{2^k}, k <= 6, is a power of two (omega_64 = 2^39)
    assert pyref.omega(6) == pow(2, 39, P)
    assert pow(2, 96, P) == P - 1


def test_twiddles_are_bitreversed_powers_with_prefix_property():
    # precompute_twiddles_for_fft(n)[r] = omega_n^brev(r); a size-n/2 table is a
    # prefix of the size-n one (src/ntt.rs:401-403).
    for log_n in range(1, 11):
        n = 1 << log_n
        for inverse in (False, True):
            tw = oracle.twiddles(n, inverse)
            w = pyref.omega(log_n)
            if inverse:
                w = pow(w, P - 2, P)
            assert [int(t) for t in tw] == [pow(w, pyref.brev(r, log_n - 1), P) for r in range(n // 2)]
            big = oracle.twiddles(2 * n, inverse)
            assert np.array_equal(big[: n // 2], tw)


@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("bitrev", [False, True])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("k,c", [(0, 0), (1, 0), (1, 1), (2, 2), (3, 5)])
def test_oracle_matches_direct_dft(log_n, bitrev, inverse, k, c):
    x = oracle.fill_raw(1 << log_n, 99, log_n)
    got = oracle.ntt_column(x, log_n, bitrev, inverse, k, c)
    assert [int(v) for v in got] == pyref.transform(x.tolist(), log_n, bitrev, inverse, k, c)


def test_oracle_dft_naive_matches_pyref():
    x = oracle.fill_uniform(16, 5)
    w = pyref.omega(4)
    got = oracle.dft_naive(x, w)
    assert [int(v) for v in got] == [sum(int(x[i]) * pow(w, i * kk, P) for i in range(16)) % P for kk in range(16)]


@pytest.mark.parametrize("log_n", [8, 10, 12])
def test_oracle_forward_is_polynomial_evaluation(log_n):
    # n2b forward with coset s: out[brev(j)] = poly(s * w^j) (Horner, sampled j)
    n = 1 << log_n
    x = oracle.fill_uniform(n, 7, log_n)
    for k, c in [(0, 0), (2, 1)]:
        out = oracle.ntt_column(x, log_n, False, False, k, c)
        s = pyref.coset_shift(log_n, k, c)
        w = pyref.omega(log_n)
        for j in [0, 1, 5, n // 2 + 3, n - 1]:
            pt = s * pow(w, j, P) % P
            assert int(out[pyref.brev(j, log_n)]) == oracle.horner(x, pt)


@pytest.mark.parametrize("log_n", [1, 7, 13])
@pytest.mark.parametrize("k,c", [(0, 0), (2, 3)])
def test_oracle_round_trip(log_n, k, c):
    n = 1 << log_n
    x = oracle.fill_raw(n, 3, log_n)
    for bitrev in (False, True):
        y = oracle.ntt_column(x, log_n, bitrev, False, k, c)
        # forward n2b output is bit-reversed -> b2n inverse brings it back (and vice versa)
        z = oracle.ntt_column(y, log_n, not bitrev, True, k, c)
        assert np.array_equal(z, oracle.canon(x))


def test_oracle_linearity_and_delta():
    log_n = 9
    n = 1 << log_n
    a = oracle.fill_uniform(n, 1)
    b = oracle.fill_uniform(n, 2)
    s = np.array([oracle.add(int(u), int(v)) for u, v in zip(a, b)], dtype=np.uint64)
    fa = oracle.ntt_column(a, log_n, False, False)
    fb = oracle.ntt_column(b, log_n, False, False)
    fs = oracle.ntt_column(s, log_n, False, False)
    assert all(int(fs[i]) == (int(fa[i]) + int(fb[i])) % P for i in range(n))
    delta = np.zeros(n, dtype=np.uint64)
    delta[0] = 1
    assert np.all(oracle.ntt_column(delta, log_n, False, False) == 1)
    assert np.all(oracle.ntt_column(delta, log_n, True, False) == 1)


def test_lde_interleave_identity():
    # coset c of a blowup-2^k LDE at position j equals the size-n*2^k transform
    # of the zero-padded coefficients on the coset g*H at index c + 2^k * j.
    log_n, k = 6, 2
    n = 1 << log_n
    coeffs = oracle.fill_uniform(n, 17)
    big = np.zeros(n << k, dtype=np.uint64)
    big[:n] = coeffs
    # natural-order coset-g evaluations of the padded polynomial
    full = oracle.bitreverse(oracle.ntt_column(big, log_n + k, False, False, 0, 0))
    g = 7
    full_g = [oracle.horner(coeffs, g * pow(pyref.omega(log_n + k), t, P) % P) for t in range(n << k)]
    assert [int(v) for v in full] != full_g  # plain transform is on H, not gH
    for c in range(1 << k):
        out = oracle.bitreverse(oracle.ntt_column(coeffs, log_n, False, False, k, c))
        assert all(int(out[j]) == full_g[c + (j << k)] for j in range(n))


def test_golden_fixtures_match_oracle_and_pyref():
    data = np.load(os.path.join(GOLDEN, "ntt_golden.npz"))
    import gen_golden_keys as gk
    for log_n in gk.LOGS:
        inp = data[f"in_L{log_n}"]
        for bitrev in (False, True):
            for inverse in (False, True):
                for k, c in gk.LDE:
                    want = data[gk.case_key(log_n, bitrev, inverse, k, c)]
                    for col in range(inp.shape[0]):
                        got = oracle.ntt_column(inp[col], log_n, bitrev, inverse, k, c)
                        assert np.array_equal(got, want[col])
                        if log_n <= 4:
                            ref = pyref.transform(inp[col].tolist(), log_n, bitrev, inverse, k, c)
                            assert [int(v) for v in want[col]] == ref


def test_golden_checksums_match_oracle():
    with open(os.path.join(GOLDEN, "ntt_checksums.json")) as f:
        cases = json.load(f)
    for cs in cases:
        if cs["log_n"] > 16:
            continue  # the GPU tests cover the 2^20 checksums
        n = 1 << cs["log_n"]
        inp = np.concatenate([oracle.fill_uniform(n, cs["seed"], c) for c in range(cs["num_cols"])])
        out = oracle.batch_ntt(inp, cs["log_n"], cs["num_cols"], n, n, cs["bitrev_inputs"], cs["inverse"],
                               threads=4)
        assert hashlib.sha256(oracle.canon(out).astype("<u8").tobytes()).hexdigest() == cs["sha256"]


def test_mt_matches_serial():
    log_n, cols = 11, 7
    n = 1 << log_n
    inp = oracle.matrix_uniform(cols, n, 4)
    a = oracle.batch_ntt(inp, log_n, cols, n, n, True, True, 2, 1, threads=1)
    b = oracle.batch_ntt(inp, log_n, cols, n, n, True, True, 2, 1, threads=4)
    assert np.array_equal(a, b)
