"""ctypes wrapper around the CPU oracle (oracle/gl_oracle.c and
oracle/poseidon_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  The C file's
header lists the reference call sites (file:line) each function restates.
"""
import ctypes
import os
import subprocess

import numpy as np

P = 0xFFFFFFFF00000001
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_gl.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)


def build():
    """Compile the oracle with gcc (seconds)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle_gl.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, u32, i32 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        sig = {
            "gl_add": (u64, [u64, u64]),
            "gl_sub": (u64, [u64, u64]),
            "gl_mul": (u64, [u64, u64]),
            "gl_pow": (u64, [u64, u64]),
            "gl_inv": (u64, [u64]),
            "gl_generator": (u64, []),
            "gl_omega_2_32": (u64, []),
            "gl_domain_generator": (u64, [u32]),
            "gl_bitrev": (u64, [u64, u32]),
            "gl_bitreverse_inplace": (None, [_u64p, u64]),
            "gl_distribute_powers": (None, [_u64p, u64, u64]),
            "gl_precompute_twiddles": (None, [u64, i32, _u64p]),
            "gl_serial_ct_n2b": (None, [_u64p, u64, _u64p]),
            "gl_fft_natural_to_bitreversed": (None, [_u64p, u64, u64, _u64p]),
            "gl_ifft_natural_to_natural": (None, [_u64p, u64, u64, _u64p]),
            "gl_coset_shift": (u64, [u32, u32, u32]),
            "gl_ntt_column": (None, [_u64p, _u64p, u32, i32, i32, u32, u32, _u64p]),
            "gl_batch_ntt": (None, [_u64p, _u64p, u32, u32, u64, u64, i32, i32, u32, u32]),
            "gl_batch_ntt_mt": (i32, [_u64p, _u64p, u32, u32, u64, u64, i32, i32, u32, u32, u32]),
            "gl_dft_naive": (None, [_u64p, _u64p, u64, u64]),
            "gl_horner": (u64, [_u64p, u64, u64]),
            "gl_fill_uniform": (None, [_u64p, u64, u64, u64]),
            "gl_fill_raw": (None, [_u64p, u64, u64, u64]),
            "gl_e2_mul": (None, [_u64p, _u64p, _u64p]),
            "gl_e2_inv": (None, [_u64p, _u64p]),
            "gl_bary_common_factor": (None, [_u64p, i32, u64, u64, _u64p]),
            "gl_bary_lagrange_coeffs": (None, [_u64p, i32, u64, u32, i32, _u64p, _u64p]),
            "gl_bary_batch_eval": (None, [_u64p, i32, _u64p, i32, u64, u32, _u64p]),
            "gl_e2_horner": (None, [_u64p, _u64p, u64, _u64p, _u64p]),
            "pos_permute_textbook": (None, [_u64p, _u64p]),
            "pos_fused_words": (ctypes.c_size_t, []),
            "pos_derive": (i32, [_u64p, _u64p]),
            "pos_permute_fused": (None, [_u64p, _u64p, _u64p]),
            "pos_leaves": (None, [_u64p, _u64p, u32, u32, u32, i32, i32, _u64p, _u64p]),
            "pos_nodes": (None, [_u64p, _u64p, u32, _u64p, _u64p]),
            "pos_grain_constants": (None, [_u64p]),
            "pos2_matrices": (None, [_u64p, _u64p]),
            "pos2_permute": (None, [_u64p, _u64p]),
            "pos2_leaves": (None, [_u64p, _u64p, u32, u32, u32, i32, i32, _u64p]),
            "pos2_nodes": (None, [_u64p, _u64p, u32, _u64p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


# ---- scalars -------------------------------------------------------------

def add(a, b):
    return lib().gl_add(a, b)


def sub(a, b):
    return lib().gl_sub(a, b)


def mul(a, b):
    return lib().gl_mul(a, b)


def pow_(b, e):
    return lib().gl_pow(b, e)


def inv(a):
    return lib().gl_inv(a)


def generator():
    return lib().gl_generator()


def omega_2_32():
    return lib().gl_omega_2_32()


def domain_generator(log_size):
    return lib().gl_domain_generator(log_size)


def coset_shift(log_n, log_ext, coset_index):
    return lib().gl_coset_shift(log_n, log_ext, coset_index)


def bitrev(x, bits):
    return lib().gl_bitrev(x, bits)


# ---- arrays ----------------------------------------------------------------

def fill_uniform(n, seed, column=0):
    out = np.empty(n, dtype=np.uint64)
    lib().gl_fill_uniform(_ptr(out), n, seed, column)
    return out


def fill_raw(n, seed, column=0):
    out = np.empty(n, dtype=np.uint64)
    lib().gl_fill_raw(_ptr(out), n, seed, column)
    return out


def matrix_uniform(num_cols, n, seed, stride=None, offset=0, raw=False):
    """Column-major matrix: column c occupies [c*stride + offset, +n)."""
    stride = n if stride is None else stride
    m = np.zeros(num_cols * stride, dtype=np.uint64)
    for c in range(num_cols):
        col = fill_raw(n, seed, c) if raw else fill_uniform(n, seed, c)
        m[c * stride + offset: c * stride + offset + n] = col
    return m


def bitreverse(a):
    a = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().gl_bitreverse_inplace(_ptr(a), a.size)
    return a


def twiddles(n, inverse):
    out = np.zeros(max(n // 2, 1), dtype=np.uint64)
    lib().gl_precompute_twiddles(n, int(inverse), _ptr(out))
    return out[: n // 2]


def ntt_column(x, log_n, bitrev_inputs, inverse, log_ext=0, coset_index=0):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    assert x.size == 1 << log_n
    out = np.empty_like(x)
    lib().gl_ntt_column(_ptr(x), _ptr(out), log_n, int(bitrev_inputs), int(inverse), log_ext, coset_index, None)
    return out


def batch_ntt(inp, log_n, num_ntts, stride_in, stride_out, bitrev_inputs, inverse, log_ext=0, coset_index=0,
              out=None, threads=1):
    """Batched oracle over flat column-major buffers (offsets pre-applied by slicing)."""
    inp = np.ascontiguousarray(inp, dtype=np.uint64)
    if out is None:
        out = np.zeros(max(inp.size, (num_ntts - 1) * stride_out + (1 << log_n)), dtype=np.uint64)
    if threads > 1:
        lib().gl_batch_ntt_mt(_ptr(inp), _ptr(out), log_n, num_ntts, stride_in, stride_out, int(bitrev_inputs),
                              int(inverse), log_ext, coset_index, threads)
    else:
        lib().gl_batch_ntt(_ptr(inp), _ptr(out), log_n, num_ntts, stride_in, stride_out, int(bitrev_inputs),
                           int(inverse), log_ext, coset_index)
    return out


def dft_naive(a, root):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.empty_like(a)
    lib().gl_dft_naive(_ptr(a), _ptr(out), a.size, root)
    return out


def horner(a, x):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return lib().gl_horner(_ptr(a), a.size, x)


def canon(a):
    """Reduce a uint64 array to canonical representatives in [0, p)."""
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= np.uint64(P), a - np.uint64(P), a)


# ---- quadratic extension and barycentric evaluation -------------------------

def _pair(x):
    if isinstance(x, (tuple, list)):
        return np.array([int(x[0]), int(x[1]) if len(x) > 1 else 0], dtype=np.uint64)
    return np.array([int(x), 0], dtype=np.uint64)


def e2_mul(a, b):
    out = np.zeros(2, dtype=np.uint64)
    lib().gl_e2_mul(_ptr(_pair(a)), _ptr(_pair(b)), _ptr(out))
    return int(out[0]), int(out[1])


def e2_inv(a):
    out = np.zeros(2, dtype=np.uint64)
    lib().gl_e2_inv(_ptr(_pair(a)), _ptr(out))
    return int(out[0]), int(out[1])


def bary_common_factor(x, coset, count):
    """x: int (base) or (c0, c1) (extension).  Returns int or (c0, c1)."""
    ext = isinstance(x, (tuple, list))
    out = np.zeros(2, dtype=np.uint64)
    lib().gl_bary_common_factor(_ptr(_pair(x)), int(ext), coset, count, _ptr(out))
    return (int(out[0]), int(out[1])) if ext else int(out[0])


def bary_lagrange_coeffs(x, coset, log_count, bit_reverse):
    """Base x: a uint64 vector of 2^log_count coefficients.  Extension x
    ((c0, c1)): the double-vector layout, c0 plane then c1 plane."""
    ext = isinstance(x, (tuple, list))
    n = 1 << log_count
    out = np.zeros(2 * n if ext else n, dtype=np.uint64)
    c1 = out[n:].ctypes.data_as(_u64p) if ext else None
    lib().gl_bary_lagrange_coeffs(_ptr(_pair(x)), int(ext), coset, log_count, int(bit_reverse), _ptr(out), c1)
    return out


def bary_batch_eval(ys, ys_ext, coeffs, coef_ext, count, num_polys):
    """Evaluations: num_polys values (base) or 2 * num_polys (c0, c1) pairs."""
    ys = np.ascontiguousarray(ys, dtype=np.uint64)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64)
    assert ys.size >= num_polys * count * (2 if ys_ext else 1)
    assert coeffs.size >= count * (2 if coef_ext else 1)
    out = np.zeros(num_polys * (2 if (ys_ext or coef_ext) else 1), dtype=np.uint64)
    lib().gl_bary_batch_eval(_ptr(ys), int(ys_ext), _ptr(coeffs), int(coef_ext), count, num_polys, _ptr(out))
    return out


def e2_horner(a0, a1, x):
    """sum_i (a0[i] + a1[i] u) x^i; a1 None for base coefficients; x int or pair."""
    a0 = np.ascontiguousarray(a0, dtype=np.uint64)
    out = np.zeros(2, dtype=np.uint64)
    a1p = _ptr(np.ascontiguousarray(a1, dtype=np.uint64)) if a1 is not None else None
    lib().gl_e2_horner(_ptr(a0), a1p, a0.size, _ptr(_pair(x)), _ptr(out))
    return int(out[0]), int(out[1])


# ---- Poseidon (oracle/poseidon_oracle.c) ------------------------------------

POS_T, POS_RATE, POS_CAP, POS_PARTIAL = 12, 8, 4, 22


def pos_grain_constants():
    out = np.zeros(360, dtype=np.uint64)
    lib().pos_grain_constants(_ptr(out))
    return out


def pos_derive(rc):
    """Fused form of 360 textbook constants: dict of rcf (12), dense (12 x 12),
    sbox_c (22), vs (22 x 11), w_hats (22 x 11), plus the packed words."""
    rc = np.ascontiguousarray(rc, dtype=np.uint64)
    words = np.zeros(lib().pos_fused_words(), dtype=np.uint64)
    assert lib().pos_derive(_ptr(rc), _ptr(words)) == 1
    o = 0
    out = {"words": words}
    for name, shape in (("rcf", (12,)), ("dense", (12, 12)), ("sbox_c", (22,)), ("vs", (22, 11)),
                        ("w_hats", (22, 11))):
        k = int(np.prod(shape))
        out[name] = words[o:o + k].reshape(shape)
        o += k
    return out


def pos_permute_textbook(state, rc):
    s = np.ascontiguousarray(state, dtype=np.uint64).copy()
    lib().pos_permute_textbook(_ptr(s), _ptr(np.ascontiguousarray(rc, dtype=np.uint64)))
    return s


def pos_permute_fused(state, rc, fused_words):
    s = np.ascontiguousarray(state, dtype=np.uint64).copy()
    lib().pos_permute_fused(_ptr(s), _ptr(np.ascontiguousarray(rc, dtype=np.uint64)), _ptr(fused_words))
    return s


def pos_leaves(values, rows, cols, count, rc, fused_words, results=None, load_intermediate=False,
               store_intermediate=False):
    """Leaf digests, 4 columns of `count` (results updated in place when given)."""
    values = np.ascontiguousarray(values, dtype=np.uint64)
    assert values.size >= rows * cols * count
    if results is None:
        results = np.zeros(4 * count, dtype=np.uint64)
    lib().pos_leaves(_ptr(values), _ptr(results), rows, cols, count, int(load_intermediate),
                     int(store_intermediate), _ptr(np.ascontiguousarray(rc, dtype=np.uint64)), _ptr(fused_words))
    return results


def pos_nodes(values, count, rc, fused_words):
    values = np.ascontiguousarray(values, dtype=np.uint64)
    assert values.size >= 8 * count
    out = np.zeros(4 * count, dtype=np.uint64)
    lib().pos_nodes(_ptr(values), _ptr(out), count, _ptr(np.ascontiguousarray(rc, dtype=np.uint64)),
                    _ptr(fused_words))
    return out


def pos2_matrices():
    m_e = np.zeros((12, 12), dtype=np.uint64)
    m_i = np.zeros((12, 12), dtype=np.uint64)
    lib().pos2_matrices(_ptr(m_e), _ptr(m_i))
    return m_e, m_i


def pos2_permute(state, rc):
    s = np.ascontiguousarray(state, dtype=np.uint64).copy()
    lib().pos2_permute(_ptr(s), _ptr(np.ascontiguousarray(rc, dtype=np.uint64)))
    return s


def pos2_leaves(values, rows, cols, count, rc, results=None, load_intermediate=False, store_intermediate=False):
    values = np.ascontiguousarray(values, dtype=np.uint64)
    assert values.size >= rows * cols * count
    if results is None:
        results = np.zeros(4 * count, dtype=np.uint64)
    lib().pos2_leaves(_ptr(values), _ptr(results), rows, cols, count, int(load_intermediate),
                      int(store_intermediate), _ptr(np.ascontiguousarray(rc, dtype=np.uint64)))
    return results


def pos2_nodes(values, count, rc):
    values = np.ascontiguousarray(values, dtype=np.uint64)
    assert values.size >= 8 * count
    out = np.zeros(4 * count, dtype=np.uint64)
    lib().pos2_nodes(_ptr(values), _ptr(out), count, _ptr(np.ascontiguousarray(rc, dtype=np.uint64)))
    return out
