"""Direct statement of the transform semantics in pure Python big-int
arithmetic, independent of the oracle's algorithm.  Used to pin the oracle
(and its fixtures) on small sizes.

Semantics (reference src/ntt.rs:527-583):
  forward  X_k = sum_i x_i s^i w^(ik)
  inverse  y_i = n^-1 s^-i sum_k x_k w^(-ik)
  s = g * omega_{n 2^ext}^coset if ext > 0 else 1, g = 7,
  omega_{2^k} = (7^((p-1)/2^32))^(2^(32-k)).
  n2b: natural in, out[brev(j)] = X_j;  b2n: in[brev(i)] = x_i, natural out.
"""
P = 2**64 - 2**32 + 1
G = 7
W32 = pow(G, (P - 1) >> 32, P)


def omega(log_size):
    return pow(W32, 1 << (32 - log_size), P)


def brev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def coset_shift(log_n, log_ext, coset_index):
    if log_ext == 0:
        return 1
    return G * pow(omega(log_n + log_ext), coset_index, P) % P


def transform(x, log_n, bitrev_inputs, inverse, log_ext=0, coset_index=0):
    n = 1 << log_n
    xs = [int(v) % P for v in x]
    if bitrev_inputs:
        xs = [xs[brev(i, log_n)] for i in range(n)]
    s = coset_shift(log_n, log_ext, coset_index)
    w = omega(log_n)
    if not inverse:
        sp = [pow(s, i, P) for i in range(n)]
        out = [sum(xs[i] * sp[i] * pow(w, i * k, P) for i in range(n)) % P for k in range(n)]
    else:
        ninv = pow(n, P - 2, P)
        sinv = pow(s, P - 2, P)
        winv = pow(w, P - 2, P)
        out = [ninv * pow(sinv, i, P) * sum(xs[k] * pow(winv, i * k, P) for k in range(n)) % P for i in range(n)]
    if not bitrev_inputs:
        out = [out[brev(j, log_n)] for j in range(n)]
    return out
