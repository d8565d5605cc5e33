"""Pins the oracle's quadratic extension and barycentric evaluation
(oracle/gl_oracle.c, restating native/barycentric.cu and
src/barycentric.rs:83-280) against mathematics written independently here:
big-int polynomial arithmetic mod u^2 - 7, the textbook Lagrange basis
prod_{j != i} (x - x_j) / (x_i - x_j), polynomial reproduction, and Horner's
rule on monomial coefficients whose coset evaluations feed batch_eval (the
reference's own test recipe, src/barycentric.rs:393-680).  CPU only."""
import random

import numpy as np
import pytest

import oracle

P = oracle.P
NR = 7  # u^2 = 7


def e2(a, b):
    """Product in F_p[u]/(u^2 - 7) by schoolbook polynomial multiplication."""
    c0 = a[0] * b[0] + NR * a[1] * b[1]
    c1 = a[0] * b[1] + a[1] * b[0]
    return c0 % P, c1 % P


def e2_add(a, b):
    return (a[0] + b[0]) % P, (a[1] + b[1]) % P


def e2_sub(a, b):
    return (a[0] - b[0]) % P, (a[1] - b[1]) % P


def e2_pow(a, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = e2(r, a)
        a = e2(a, a)
        e >>= 1
    return r


def e2_inv_fermat(a):
    """a^(p^2 - 2): the multiplicative group of F_{p^2} has order p^2 - 1."""
    return e2_pow(a, P * P - 2)


def brev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def test_seven_is_a_non_residue():
    # Euler's criterion: 7^((p-1)/2) = -1, so u^2 - 7 is irreducible.
    assert pow(7, (P - 1) // 2, P) == P - 1


def test_e2_ops_match_big_int():
    rng = random.Random(5)
    edge = [0, 1, P - 1, 2**32, 2**63]
    vals = [(a, b) for a in edge for b in edge] + [(rng.randrange(P), rng.randrange(P)) for _ in range(200)]
    for a in vals:
        for b in vals[:20]:
            assert oracle.e2_mul(a, b) == e2(a, b)
        if a != (0, 0):
            ia = oracle.e2_inv(a)
            assert ia == e2_inv_fermat(a)
            assert e2(ia, a) == (1, 0)


@pytest.mark.parametrize("ext", [False, True])
@pytest.mark.parametrize("log_count", [0, 1, 2, 3])
def test_lagrange_coeffs_match_textbook_basis(ext, log_count):
    """L_i(x) = prod_{j != i} (x - x_j) / (x_i - x_j) over the coset
    x_j = coset * omega_N^j, computed with no shared formula."""
    rng = random.Random(100 + log_count)
    n = 1 << log_count
    coset = 7
    w = pow(0x185629DCDA58878C, 1 << (32 - log_count), P)  # omega_N: order n
    assert pow(w, n, P) == 1 and (n == 1 or pow(w, n // 2, P) != 1)
    for _ in range(3):
        x = (rng.randrange(P), rng.randrange(P)) if ext else (rng.randrange(P), 0)
        pts = [coset * pow(w, j, P) % P for j in range(n)]
        want = []
        for i in range(n):
            num, den = (1, 0), 1
            for j in range(n):
                if j != i:
                    num = e2(num, e2_sub(x, (pts[j], 0)))
                    den = den * (pts[i] - pts[j]) % P
            want.append(e2(num, (pow(den, P - 2, P), 0)))
        for bit_reverse in (False, True):
            got = oracle.bary_lagrange_coeffs(x if ext else x[0], coset, log_count, bit_reverse)
            for i in range(n):
                pos = brev(i, log_count) if bit_reverse else i
                g = (int(got[pos]), int(got[n + pos]) if ext else 0)
                assert g == want[i], (i, bit_reverse)


@pytest.mark.parametrize("ext", [False, True])
def test_lagrange_coeffs_reproduce_polynomials(ext):
    """sum_i L_i(x) x_i^k = x^k for k < N (the basis interpolates exactly),
    checked at N = 64 for every k, and sum_i L_i = 1 up to N = 2^12."""
    rng = random.Random(9)
    coset = 7
    x = (rng.randrange(P), rng.randrange(P)) if ext else (rng.randrange(P), 0)
    log_count = 6
    n = 1 << log_count
    w = pow(0x185629DCDA58878C, 1 << (32 - log_count), P)
    c = oracle.bary_lagrange_coeffs(x if ext else x[0], coset, log_count, False)
    L = [(int(c[i]), int(c[n + i]) if ext else 0) for i in range(n)]
    pts = [coset * pow(w, i, P) % P for i in range(n)]
    for k in range(n):
        acc = (0, 0)
        for i in range(n):
            acc = e2_add(acc, e2(L[i], (pow(pts[i], k, P), 0)))
        assert acc == e2_pow(x, k), k
    for log_count in (8, 12):
        n = 1 << log_count
        c = oracle.bary_lagrange_coeffs(x if ext else x[0], coset, log_count, True)
        s0 = int(np.sum(c[:n] % np.uint64(1 << 32))) + (int(np.sum(c[:n] >> np.uint64(32))) << 32)
        assert s0 % P == 1
        if ext:
            s1 = int(np.sum(c[n:] % np.uint64(1 << 32))) + (int(np.sum(c[n:] >> np.uint64(32))) << 32)
            assert s1 % P == 0


def test_common_factor_closed_form():
    rng = random.Random(3)
    for log_count in (0, 1, 5, 20):
        n = 1 << log_count
        coset = rng.randrange(1, P)
        xb = rng.randrange(P)
        cn = pow(coset, n, P)
        f = coset * pow(n * cn % P, P - 2, P) % P
        assert oracle.bary_common_factor(xb, coset, n) == (pow(xb, n, P) - cn) * f % P
        xe = (rng.randrange(P), rng.randrange(P))
        xn = e2_pow(xe, n)
        assert oracle.bary_common_factor(xe, coset, n) == ((xn[0] - cn) * f % P, xn[1] * f % P)


def coset_evaluations(coeffs, log_count, num_cols):
    """Forward coset NTT (log_extension_degree 1, coset index 0: points
    7 omega_N^i, outputs bit-reversed), as the reference's tests make ys."""
    n = 1 << log_count
    if log_count == 0:
        return coeffs.copy()
    return oracle.batch_ntt(coeffs, log_count, num_cols, n, n, False, False, 1, 0)[: num_cols * n]


@pytest.mark.parametrize("kind", ["base_at_base", "base_at_ext", "ext_at_ext"])
@pytest.mark.parametrize("log_count", [0, 1, 4, 9])
def test_batch_eval_matches_horner(kind, log_count):
    """src/barycentric.rs:393-680: monomial coefficients -> coset NTT -> ys;
    batch_eval(ys, bit-reversed L(x)) equals Horner's rule at x."""
    rng = random.Random(1000 + log_count)
    n = 1 << log_count
    num_polys = 3
    ys_ext = kind == "ext_at_ext"
    x_ext = kind != "base_at_base"
    cols = num_polys * (2 if ys_ext else 1)
    coeffs = np.array([rng.randrange(P) for _ in range(cols * n)], dtype=np.uint64)
    ys = coset_evaluations(coeffs, log_count, cols)
    x = (rng.randrange(P), rng.randrange(P)) if x_ext else rng.randrange(P)
    L = oracle.bary_lagrange_coeffs(x, 7, log_count, True)
    got = oracle.bary_batch_eval(ys, ys_ext, L, x_ext, n, num_polys)
    for p in range(num_polys):
        if ys_ext:
            a0, a1 = coeffs[2 * p * n:(2 * p + 1) * n], coeffs[(2 * p + 1) * n:(2 * p + 2) * n]
        else:
            a0, a1 = coeffs[p * n:(p + 1) * n], None
        want = oracle.e2_horner(a0, a1, x)
        # Horner in Python big ints as the second, independent evaluation
        xx = x if x_ext else (x, 0)
        acc = (0, 0)
        for i in reversed(range(n)):
            acc = e2_add(e2(acc, xx), (int(a0[i]), int(a1[i]) if a1 is not None else 0))
        assert want == acc
        g = (int(got[2 * p]), int(got[2 * p + 1])) if x_ext else (int(got[p]), 0)
        assert g == want, p
