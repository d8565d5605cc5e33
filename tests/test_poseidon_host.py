"""The library's own host-side Poseidon / Poseidon2 table (bc_poseidon_fused_table,
no device call) against the oracle and against the reference's fused tables
(native/poseidon_constants_template.cuh:43-124, through the checksums in
tests/golden/poseidon_fused.json), plus the argument checks of the hashing
entry points (src/poseidon.rs:139-155, 224-232, 262-328), which reject bad
sizes before any device call."""
import hashlib
import json
import os
import struct
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "era-boojum-cuda_amd"))
import oracle  # noqa: E402
from boojum_cuda_amd import poseidon  # noqa: E402
from boojum_cuda_amd._lib import BC_ERROR_INVALID_VALUE, lib  # noqa: E402

GOLDEN = json.load(open(os.path.join(HERE, "golden", "poseidon_fused.json")))


def sha(values):
    return hashlib.sha256(b"".join(struct.pack("<Q", int(v)) for v in np.ravel(values))).hexdigest()


@pytest.mark.parametrize("seed", [None, 1, 2])
def test_library_derivation_equals_oracle(seed):
    rc = oracle.pos_grain_constants() if seed is None else \
        np.random.default_rng(seed).integers(0, 2**64, 360, dtype=np.uint64)  # raw: read mod p
    got = poseidon.fused_table(rc)
    want = oracle.pos_derive(rc)
    for key in ("rcf", "dense", "sbox_c", "vs", "w_hats"):
        assert np.array_equal(got[key], want[key]), key
    canon = rc % np.uint64(oracle.P) if seed is not None else rc
    assert np.array_equal(got["rc_full"][:4], canon[:48].reshape(4, 12))
    assert np.array_equal(got["rc_full"][4:], canon[27 * 12:].reshape(3, 12))
    # Poseidon2 reads the textbook constants directly
    assert np.array_equal(got["p2_full"][:4], canon[:48].reshape(4, 12))
    assert np.array_equal(got["p2_full"][4:], canon[26 * 12:].reshape(4, 12))
    assert np.array_equal(got["p2_partial"], canon[4 * 12:26 * 12:12])


@pytest.mark.parametrize("key,name", [
    ("dense", "FUSED_DENSE_MATRIX_LAST_FULL_AND_FIRST_PARTIAL"),
    ("vs", "VS_FOR_PARTIAL_ROUNDS"),
    ("w_hats", "W_HATS_FOR_PARTIAL_ROUNDS"),
])
def test_library_matrices_equal_reference_tables(key, name):
    got = poseidon.fused_table(oracle.pos_grain_constants())[key]
    assert sha(got) == GOLDEN["tables"][name]["sha256"]


def test_fused_table_rejects_short_output():
    rc = np.zeros(360, dtype=np.uint64)
    out = np.zeros(100, dtype=np.uint64)
    assert lib().bc_poseidon_fused_table(rc.ctypes.data, out.ctypes.data, out.size) == BC_ERROR_INVALID_VALUE


@pytest.mark.parametrize("values_len,results_len,log_rows,store", [
    (64, 6, 0, False),    # results not a multiple of the capacity
    (63, 8, 0, False),    # values not a multiple of count * rows
    (10, 8, 1, False),    # 2 leaves x 2 rows: 10 is not a multiple of 4
    (6, 8, 0, True),      # store_intermediate with rows * cols = 3, not a multiple of 8
    (8, 8, 32, False),    # log_rows_per_hash out of range
])
def test_leaves_rejects_reference_asserts(values_len, results_len, log_rows, store):
    assert lib().bc_poseidon_leaves(1, values_len, 1, results_len, log_rows, False, store, None) == \
        BC_ERROR_INVALID_VALUE


@pytest.mark.parametrize("values_len,results_len", [(12, 6), (16, 4), (16, 10)])
def test_nodes_rejects_reference_asserts(values_len, results_len):
    assert lib().bc_poseidon_nodes(1, values_len, 1, results_len, None) == BC_ERROR_INVALID_VALUE


@pytest.mark.parametrize("values_len,results_len,layers", [(24, 24, 1), (16, 8, 1), (16, 16, 3)])
def test_tree_nodes_rejects_reference_asserts(values_len, results_len, layers):
    assert lib().bc_poseidon_build_merkle_tree_nodes(1, values_len, 1, results_len, layers, None) == \
        BC_ERROR_INVALID_VALUE


@pytest.mark.parametrize("values_len,results_len,layers", [(8, 8, 0), (8, 12, 1), (8, 16, 3), (6, 32, 1)])
def test_tree_rejects_reference_asserts(values_len, results_len, layers):
    assert lib().bc_poseidon_build_merkle_tree(1, values_len, 1, results_len, 0, layers, None) == \
        BC_ERROR_INVALID_VALUE


@pytest.mark.parametrize("count,log_rpi,results_rows", [(4, 32, 4), (4, 1, 4), (3, 0, 4)])
def test_gather_rows_rejects_reference_asserts(count, log_rpi, results_rows):
    assert lib().bc_gather_rows(1, count, log_rpi, 1, 16, 16, 1, results_rows, results_rows, 2, None) == \
        BC_ERROR_INVALID_VALUE


@pytest.mark.parametrize("values_len,results_len,layers", [
    (30, 16, 1),    # not a multiple of the capacity
    (24, 16, 1),    # 6 digests: not a power of two
    (4, 0, 0),      # one digest: no leaf layer
    (64, 16 * 3, 3),  # 8 leaves: layers must be below 3
    (64, 20, 1),    # results != indexes * layers * 4
])
def test_gather_merkle_paths_rejects_reference_asserts(values_len, results_len, layers):
    assert lib().bc_gather_merkle_paths(1, 4, 1, values_len, 1, results_len, layers, None) == BC_ERROR_INVALID_VALUE
