"""Poseidon and Poseidon2 Merkle-tree hashing, mirroring the reference's
`poseidon` module (src/poseidon.rs:129-328) over the C ABI.  Where the
reference is generic over `P: PoseidonImpl`, these functions take
`p=POSEIDON` or `p=POSEIDON2`.
Tensors are HIP torch tensors of int64/uint64; digests are 4 elements, stored
as 4 columns of `count` (element i of digest g at results[g + i * count]).

The permutation's 360 round constants are boojum's
(poseidon_goldilocks_params::ALL_ROUND_CONSTANTS), which the reference
compiles in at build time; here they are handed to `set_round_constants`
once per device before any hashing call; both hashes read them."""
import ctypes

import numpy as np

from ._lib import BC_ERROR_INVALID_VALUE, BC_ERROR_NOT_INITIALIZED, HipError, lib

RATE = 8
CAPACITY = 4
STATE_WIDTH = 12
NUM_ROUND_CONSTANTS = 360
FUSED_TABLE_WORDS = 864

POSEIDON = "poseidon"    # struct Poseidon (src/poseidon.rs)
POSEIDON2 = "poseidon2"  # struct Poseidon2


def _check(rc, what):
    if rc == 0:
        return
    if rc == BC_ERROR_INVALID_VALUE:
        raise AssertionError(f"{what}: argument assertion failed")
    if rc == BC_ERROR_NOT_INITIALIZED:
        raise RuntimeError(f"{what}: no round constants on this device (call set_round_constants)")
    raise HipError(rc, what)


def _stream(stream):
    import torch
    return (stream or torch.cuda.current_stream()).cuda_stream


def _u64(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint64))
    return a, a.ctypes.data_as(ctypes.c_void_p)


def set_round_constants(round_constants):
    """360 textbook round constants (30 rounds x 12, round-major), any
    sequence of ints; derives the fused form and stores it on the current
    device."""
    rc, ptr = _u64(round_constants)
    assert rc.size == NUM_ROUND_CONSTANTS
    _check(lib().bc_poseidon_set_round_constants(ptr), "set_round_constants")


def fused_table(round_constants):
    """Host-side derivation only: dict of the fused tables the kernels use."""
    rc, ptr = _u64(round_constants)
    assert rc.size == NUM_ROUND_CONSTANTS
    out = np.zeros(FUSED_TABLE_WORDS, dtype=np.uint64)
    _check(lib().bc_poseidon_fused_table(ptr, out.ctypes.data_as(ctypes.c_void_p), out.size), "fused_table")
    res, o = {"words": out}, 0
    for name, shape in (("rc_full", (7, 12)), ("rcf", (12,)), ("dense", (12, 12)), ("sbox_c", (22,)),
                        ("vs", (22, 11)), ("w_hats", (22, 11)), ("p2_full", (8, 12)), ("p2_partial", (22,))):
        k = int(np.prod(shape))
        res[name] = out[o:o + k].reshape(shape)
        o += k
    return res


def _fn(p, name):
    assert p in (POSEIDON, POSEIDON2)
    return getattr(lib(), f"bc_{p}_{name}")


def launch_single_thread_leaves_kernel(values, results, log_rows_per_hash, load_intermediate=False,
                                       store_intermediate=False, stream=None, p=POSEIDON):
    """src/poseidon.rs:172-192 (argument checks of launch_leaves_kernel,
    src/poseidon.rs:129-170)."""
    _check(_fn(p, "leaves")(values.data_ptr(), values.numel(), results.data_ptr(), results.numel(),
                            log_rows_per_hash, bool(load_intermediate), bool(store_intermediate),
                            _stream(stream)), "launch_leaves_kernel")


def launch_single_thread_nodes_kernel(values, results, stream=None, p=POSEIDON):
    """src/poseidon.rs:218-238."""
    _check(_fn(p, "nodes")(values.data_ptr(), values.numel(), results.data_ptr(), results.numel(),
                           _stream(stream)), "launch_nodes_kernel")


def launch_cooperative_leaves_kernel(values, results, log_rows_per_hash, load_intermediate=False,
                                     store_intermediate=False, stream=None, p=POSEIDON2):
    """src/poseidon.rs:194-216: the same leaves with 16 lanes per permutation
    (the reference has it for Poseidon2 only; both hashes here)."""
    _check(_fn(p, "leaves_cooperative")(values.data_ptr(), values.numel(), results.data_ptr(), results.numel(),
                                        log_rows_per_hash, bool(load_intermediate), bool(store_intermediate),
                                        _stream(stream)), "launch_leaves_kernel")


def launch_cooperative_nodes_kernel(values, results, stream=None, p=POSEIDON2):
    """src/poseidon.rs:240-260."""
    _check(_fn(p, "nodes_cooperative")(values.data_ptr(), values.numel(), results.data_ptr(), results.numel(),
                                       _stream(stream)), "launch_nodes_kernel")


def build_merkle_tree_leaves(values, results, log_rows_per_hash, load_intermediate=False,
                             store_intermediate=False, stream=None, p=POSEIDON):
    """src/poseidon.rs:288-309."""
    assert results.numel() % CAPACITY == 0
    leaves = results.numel() // CAPACITY
    assert leaves and values.numel() % leaves == 0
    launch_single_thread_leaves_kernel(values, results, log_rows_per_hash, load_intermediate, store_intermediate,
                                       stream, p)


def build_merkle_tree_nodes(values, results, layers_count, stream=None, p=POSEIDON):
    """src/poseidon.rs:262-286."""
    _check(_fn(p, "build_merkle_tree_nodes")(values.data_ptr(), values.numel(), results.data_ptr(),
                                             results.numel(), layers_count, _stream(stream)),
           "build_merkle_tree_nodes")


def build_merkle_tree(values, results, log_rows_per_hash, stream=None, layers_count=1, p=POSEIDON):
    """src/poseidon.rs:311-328: results (8 leaves elements) holds the leaf
    digests in its first half and the layers_count - 1 node layers after."""
    _check(_fn(p, "build_merkle_tree")(values.data_ptr(), values.numel(), results.data_ptr(),
                                       results.numel(), log_rows_per_hash, layers_count, _stream(stream)),
           "build_merkle_tree")


def gather_rows(indexes, log_rows_per_index, values, values_rows, result, result_rows, stream=None):
    """src/poseidon.rs:330-358.  indexes: int32/uint32 tensor; values / result:
    column-major matrices (flat tensors, column stride = rows) with the same
    number of columns."""
    assert values_rows and result_rows and values.numel() % values_rows == 0 and result.numel() % result_rows == 0
    cols = values.numel() // values_rows
    assert result.numel() // result_rows == cols
    _check(lib().bc_gather_rows(indexes.data_ptr(), indexes.numel(), log_rows_per_index, values.data_ptr(),
                                values_rows, values_rows, result.data_ptr(), result_rows, result_rows, cols,
                                _stream(stream)), "gather_rows")


def gather_merkle_paths(indexes, values, results, layers_count, stream=None):
    """src/poseidon.rs:360-390."""
    _check(lib().bc_gather_merkle_paths(indexes.data_ptr(), indexes.numel(), values.data_ptr(), values.numel(),
                                        results.data_ptr(), results.numel(), layers_count, _stream(stream)),
           "gather_merkle_paths")
