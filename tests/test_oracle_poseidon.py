"""Pins the Poseidon oracle (oracle/poseidon_oracle.c) on the CPU.

- The fused matrices it derives (dense, VS, W_HATS) depend only on the MDS
  matrix; they must equal the reference's tables
  (native/poseidon_constants_template.cuh:43-124), compared through the
  checksums in tests/golden/poseidon_fused.json.
- The fused permutation (the form native/poseidon_single_thread.cu:14-32 runs)
  must equal the textbook permutation, here restated independently in Python
  big ints, for any round constants.
- Leaf and node hashing (native/poseidon_utils.cuh:8-61) against a Python
  sponge built on that textbook permutation.

Parity unpinned: boojum's round constant values (absent from the reference),
so the two constant-dependent tables (fused constant layer, fused S-box
constants) cannot be checked; their checksums stay in the fixture for the
day the constants are supplied.
"""
import hashlib
import json
import os
import struct
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import oracle  # noqa: E402

P = oracle.P
GOLDEN = json.load(open(os.path.join(HERE, "golden", "poseidon_fused.json")))


def sha(values):
    return hashlib.sha256(b"".join(struct.pack("<Q", int(v)) for v in np.ravel(values))).hexdigest()


# ---- independent textbook statement (Python big ints) -----------------------

EXPS = GOLDEN["mds_matrix_exps"]


def py_permute(state, rc):
    s = [int(x) % P for x in state]
    for r in range(30):
        s = [(s[i] + int(rc[r * 12 + i])) % P for i in range(12)]
        full = r < 4 or r >= 26
        s = [pow(x, 7, P) if (full or i == 0) else x for i, x in enumerate(s)]
        s = [sum(s[(i + row) % 12] << EXPS[i] for i in range(12)) % P for row in range(12)]
    return s


def py_leaves(values, rows, cols, count, rc, results=None, load=False, store=False):
    out = [0] * (4 * count) if results is None else [int(x) for x in results]
    for g in range(count):
        s = [0] * 12
        if load:
            s[8:12] = [out[g + i * count] for i in range(4)]
        flat = [int(values[g * rows + r + c * rows * count]) for c in range(cols) for r in range(rows)]
        for k in range(0, len(flat), 8):
            chunk = flat[k:k + 8]
            s[:8] = chunk + [0] * (8 - len(chunk))
            s = py_permute(s, rc)
        for i in range(4):
            out[g + i * count] = s[8 + i] if store else s[i]
    return np.array(out, dtype=np.uint64)


@pytest.fixture(scope="module")
def grain():
    rc = oracle.pos_grain_constants()
    return rc, oracle.pos_derive(rc)


# ---- derived tables vs the reference ------------------------------------------

@pytest.mark.parametrize("key,name", [
    ("dense", "FUSED_DENSE_MATRIX_LAST_FULL_AND_FIRST_PARTIAL"),
    ("vs", "VS_FOR_PARTIAL_ROUNDS"),
    ("w_hats", "W_HATS_FOR_PARTIAL_ROUNDS"),
])
def test_fused_matrices_equal_reference_tables(grain, key, name):
    got = grain[1][key]
    ref = GOLDEN["tables"][name]
    assert got.size == ref["count"]
    assert hex(int(got.ravel()[0])) == ref["first"] and hex(int(got.ravel()[-1])) == ref["last"]
    assert sha(got) == ref["sha256"]


def test_fused_matrices_do_not_depend_on_constants(grain):
    rng = np.random.default_rng(5)
    other = oracle.pos_derive(rng.integers(0, P, 360, dtype=np.uint64))
    for key in ("dense", "vs", "w_hats"):
        assert np.array_equal(other[key], grain[1][key])
    assert not np.array_equal(other["rcf"], grain[1]["rcf"])


def test_dense_row0_is_mds_row0(grain):
    assert [int(x) for x in grain[1]["dense"][0]] == [1 << e for e in EXPS]


@pytest.mark.skip(reason="parity unpinned: boojum's 360 Poseidon round constants are not in the reference "
                         "tree, and no Grain LFSR seeding tried (field 0/1, S-box 0/1, field size 63/64 bits, "
                         "big- or little-endian 64-bit draws) reproduces the reference's constant-dependent "
                         "fused tables (native/poseidon_constants_template.cuh:48,65)")
def test_constant_dependent_tables_match_reference(grain):
    """The two constant-dependent fused tables derived from the candidate
    constants equal the reference's (checked through the fixture's SHA-256).
    It would pin the digests once boojum's constants are available."""
    for key, name in (("rcf", "ROUND_CONSTANTS_FUSED_LAST_FULL_AND_FIRST_PARTIAL"),
                      ("sbox_c", "ROUND_CONSTANTS_FOR_FUSED_S_BOXES")):
        assert sha(grain[1][key]) == GOLDEN["tables"][name]["sha256"]


# ---- permutation ------------------------------------------------------------

@pytest.mark.parametrize("const_seed", [None, 11, 12])
def test_fused_equals_textbook(grain, const_seed):
    rng = np.random.default_rng(7)
    if const_seed is None:
        rc, fused = grain[0], grain[1]["words"]
    else:
        rc = np.random.default_rng(const_seed).integers(0, P, 360, dtype=np.uint64)
        fused = oracle.pos_derive(rc)["words"]
    for _ in range(4):
        x = rng.integers(0, P, 12, dtype=np.uint64)
        want = py_permute(x, rc)
        assert [int(v) for v in oracle.pos_permute_textbook(x, rc)] == want
        assert [int(v) for v in oracle.pos_permute_fused(x, rc, fused)] == want


def test_permutation_edge_states(grain):
    rc, fused = grain[0], grain[1]["words"]
    for x in (np.zeros(12, dtype=np.uint64), np.full(12, P - 1, dtype=np.uint64),
              np.arange(12, dtype=np.uint64)):
        assert [int(v) for v in oracle.pos_permute_fused(x, rc, fused)] == py_permute(x, rc)


def test_grain_stream_shape(grain):
    rc = grain[0]
    assert rc.size == 360 and int(rc.max()) < P and len(set(rc.tolist())) == 360


# ---- leaves and nodes --------------------------------------------------------

@pytest.mark.parametrize("rows,cols", [(1, 1), (1, 8), (1, 13), (2, 3), (2, 4), (4, 5), (8, 2)])
def test_leaves_match_python_sponge(grain, rows, cols):
    rc, fused = grain[0], grain[1]["words"]
    count = 3
    vals = np.random.default_rng(rows * 100 + cols).integers(0, 2**64, rows * cols * count, dtype=np.uint64)
    got = oracle.pos_leaves(vals, rows, cols, count, rc, fused)
    assert np.array_equal(got, py_leaves(vals, rows, cols, count, rc))


def test_leaves_checkpoint_split_equals_one_shot(grain):
    """Two launches over column ranges, the first storing the capacity and the
    second loading it, hash like one launch (src/poseidon.rs:150-155 requires
    rows * cols of the first part to be a multiple of the rate)."""
    rc, fused = grain[0], grain[1]["words"]
    rows, count, c1, c2 = 2, 5, 4, 3
    vals = np.random.default_rng(3).integers(0, P, rows * (c1 + c2) * count, dtype=np.uint64)
    one = oracle.pos_leaves(vals, rows, c1 + c2, count, rc, fused)
    part = oracle.pos_leaves(vals[:rows * c1 * count], rows, c1, count, rc, fused, store_intermediate=True)
    assert np.array_equal(part, py_leaves(vals[:rows * c1 * count], rows, c1, count, rc, store=True))
    two = oracle.pos_leaves(vals[rows * c1 * count:], rows, c2, count, rc, fused, results=part,
                            load_intermediate=True)
    assert np.array_equal(two, one)


def test_nodes_match_python(grain):
    rc, fused = grain[0], grain[1]["words"]
    count = 4
    vals = np.random.default_rng(9).integers(0, P, 8 * count, dtype=np.uint64)
    got = oracle.pos_nodes(vals, count, rc, fused)
    for g in range(count):
        # children 2g, 2g + 1 of a layer of 2 * count digests, 4 columns
        left = [int(vals[i * 2 * count + 2 * g]) for i in range(4)]
        right = [int(vals[i * 2 * count + 2 * g + 1]) for i in range(4)]
        want = py_permute(left + right + [0] * 4, rc)[:4]
        assert [int(got[g + i * count]) for i in range(4)] == want


# ---- Poseidon2 ----------------------------------------------------------------

M4_PAPER = [[5, 7, 1, 3], [4, 6, 1, 1], [1, 3, 5, 7], [1, 1, 4, 6]]  # eprint 2023/323, Appendix B
MU = GOLDEN["poseidon2_log_mu_minus_one"]


def py2_permute(state, rc):
    """Poseidon2 built tile by tile: M_E = M4 on each 4-tile, then every tile
    plus the sum of all tiles; M_I = the sum of all elements plus 2^mu_i x_i."""
    def m_e(s):
        t = [sum(M4_PAPER[r][c] * s[4 * k + c] for c in range(4)) % P for k in range(3) for r in range(4)]
        tot = [sum(t[4 * k + j] for k in range(3)) for j in range(4)]
        return [(t[i] + tot[i % 4]) % P for i in range(12)]

    def m_i(s):
        tot = sum(s)
        return [(tot + (x << MU[i])) % P for i, x in enumerate(s)]

    s = m_e([int(x) % P for x in state])
    for r in range(30):
        full = r < 4 or r >= 26
        s = [pow((x + int(rc[r * 12 + i])) % P, 7, P) if (full or i == 0) else x for i, x in enumerate(s)]
        s = m_e(s) if full else m_i(s)
    return s


def det_mod_p(m):
    a = [[int(x) % P for x in row] for row in m]
    n, d = len(a), 1
    for c in range(n):
        piv = next((r for r in range(c, n) if a[r][c]), None)
        if piv is None:
            return 0
        if piv != c:
            a[c], a[piv] = a[piv], a[c]
            d = -d
        d = d * a[c][c] % P
        inv = pow(a[c][c], P - 2, P)
        for r in range(c + 1, n):
            f = a[r][c] * inv % P
            a[r] = [(x - f * y) % P for x, y in zip(a[r], a[c])]
    return d % P


def test_poseidon2_matrices():
    m_e, m_i = oracle.pos2_matrices()
    for k in range(3):
        for l in range(3):
            block = m_e[4 * k:4 * k + 4, 4 * l:4 * l + 4].astype(object)
            assert (block == np.array(M4_PAPER, dtype=object) * (2 if k == l else 1)).all()
    assert [int(m_i[i, i]) - 1 for i in range(12)] == [1 << m for m in MU]
    assert all(int(m_i[i, j]) == 1 for i in range(12) for j in range(12) if i != j)
    assert det_mod_p(m_e) != 0 and det_mod_p(m_i) != 0


@pytest.mark.parametrize("const_seed", [None, 21])
def test_poseidon2_permutation(grain, const_seed):
    rc = grain[0] if const_seed is None else np.random.default_rng(const_seed).integers(0, P, 360, dtype=np.uint64)
    rng = np.random.default_rng(8)
    for x in [np.zeros(12, dtype=np.uint64), np.full(12, P - 1, dtype=np.uint64)] + \
             [rng.integers(0, 2**64, 12, dtype=np.uint64) for _ in range(4)]:
        assert [int(v) for v in oracle.pos2_permute(x, rc)] == py2_permute(x, rc)


def test_poseidon2_differs_from_poseidon(grain):
    x = np.arange(12, dtype=np.uint64)
    assert not np.array_equal(oracle.pos2_permute(x, grain[0]), oracle.pos_permute_fused(x, grain[0], grain[1]["words"]))


@pytest.mark.parametrize("rows,cols", [(1, 1), (1, 13), (2, 4), (4, 3)])
def test_poseidon2_leaves_and_nodes(grain, rows, cols):
    rc = grain[0]
    count = 3
    vals = np.random.default_rng(rows * 10 + cols).integers(0, 2**64, rows * cols * count, dtype=np.uint64)
    got = oracle.pos2_leaves(vals, rows, cols, count, rc)
    for g in range(count):
        s = [0] * 12
        flat = [int(vals[g * rows + r + c * rows * count]) for c in range(cols) for r in range(rows)]
        for k in range(0, len(flat), 8):
            chunk = flat[k:k + 8]
            s[:8] = chunk + [0] * (8 - len(chunk))
            s = py2_permute(s, rc)
        assert [int(got[g + i * count]) for i in range(4)] == s[:4]
    nvals = np.random.default_rng(cols).integers(0, P, 8 * count, dtype=np.uint64)
    nodes = oracle.pos2_nodes(nvals, count, rc)
    for g in range(count):
        left = [int(nvals[i * 2 * count + 2 * g]) for i in range(4)]
        right = [int(nvals[i * 2 * count + 2 * g + 1]) for i in range(4)]
        assert [int(nodes[g + i * count]) for i in range(4)] == py2_permute(left + right + [0] * 4, rc)[:4]
