"""Loader for libboojum_ntt.so, the HIP library behind this package.

There is no fallback: if the library is missing, or no GPU is present when a
transform is issued, the call raises.
"""
import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "libboojum_ntt.so")
_lib = None

c_u32, c_u64, c_size, c_bool, c_vp, c_int = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_bool,
                                             ctypes.c_void_p, ctypes.c_int)

# name -> (restype, argtypes); must match include/boojum_ntt.h
SIGNATURES = {
    "bc_omega_log_order": (c_u32, []),
    "bc_context_create": (c_int, [c_u32, c_u32, ctypes.POINTER(c_vp)]),
    "bc_context_destroy": (c_int, [c_vp]),
    "bc_batch_ntt_internal": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_bool, c_bool, c_u32, c_u32, c_vp]),
    "bc_batch_ntt_out_of_place": (c_int, [c_vp, c_size, c_vp, c_size, c_u32, c_u32, c_u32, c_u32, c_u32, c_u32,
                                          c_bool, c_bool, c_u32, c_u32, c_vp]),
    "bc_batch_ntt_in_place": (c_int, [c_vp, c_size, c_u32, c_u32, c_u32, c_u32, c_bool, c_bool, c_u32, c_u32,
                                      c_vp]),
    "bc_batch_lde": (c_int, [c_vp, c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_u64, c_u32, c_bool, c_vp]),
    "bc_lde_graph_create": (c_int, [c_vp, c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_u64, c_u32, c_bool,
                                    ctypes.POINTER(c_vp)]),
    "bc_graph_launch": (c_int, [c_vp, c_vp]),
    "bc_graph_destroy": (c_int, [c_vp]),
    "bc_ragged_plan_create": (c_int, [c_vp, c_u32, ctypes.POINTER(c_vp)]),
    "bc_ragged_run": (c_int, [c_vp, c_vp]),
    "bc_ragged_plan_destroy": (c_int, [c_vp]),
    "bc_ragged_plan_info": (c_int, [c_vp, ctypes.POINTER(c_u64), ctypes.POINTER(c_u32), ctypes.POINTER(c_u32)]),
    "bc_ragged_run_plain": (c_int, [c_vp, c_u32, c_vp]),
    "bc_ragged_run_no_alloc": (c_int, [c_vp, c_u32, c_vp]),
    "bc_ragged_wait_status": (c_int, [c_vp]),
    "bc_ragged_set_wait_limit": (c_int, [c_u32]),
    "bc_ntt_describe_plan": (c_int, [c_u32, c_u32, c_bool, c_bool, c_u32, c_u32, c_vp, c_u32,
                                     ctypes.POINTER(c_u32)]),
    "bc_tune_set": (c_int, [c_u32, c_u32, c_vp]),
    "bc_tune_get": (c_int, [c_u32, c_u32, c_vp]),
    "bc_tune_clear": (c_int, []),
    "bc_tune_load": (c_int, [ctypes.c_char_p]),
    "bc_tune_save": (c_int, [ctypes.c_char_p]),
    "bc_bit_reverse": (c_int, [c_vp, c_u64, c_vp, c_u64, c_u64, c_u32, c_vp]),
    "bc_bit_reverse_in_place": (c_int, [c_vp, c_u64, c_u64, c_u32, c_vp]),
    "bc_get_powers_of_w": (c_int, [c_u32, c_u32, c_bool, c_bool, c_vp, c_u32, c_vp]),
    "bc_get_powers_of_g": (c_int, [c_u32, c_u32, c_bool, c_bool, c_vp, c_u32, c_vp]),
    "bc_omega_shift": (c_int, [c_vp, c_u32, c_u32, c_bool, c_u32, c_vp, c_u32, c_vp]),
    "bc_omega_shift_in_place": (c_int, [c_vp, c_u32, c_u32, c_bool, c_u32, c_u32, c_vp]),
    "bc_barycentric_precompute": (c_int, [c_bool, c_vp, c_vp, c_u64, c_vp, c_u64, c_bool, c_vp]),
    "bc_batch_eval_temp_sizes": (c_int, [c_u32, c_u64, c_u32, ctypes.POINTER(c_size), ctypes.POINTER(c_size)]),
    "bc_batch_eval": (c_int, [c_u32, c_vp, c_u64, c_u32, c_vp, c_vp, c_size, c_vp, c_size, c_vp, c_vp]),
    "bc_poseidon_set_round_constants": (c_int, [c_vp]),
    "bc_poseidon_fused_table": (c_int, [c_vp, c_vp, c_u64]),
    "bc_poseidon_leaves": (c_int, [c_vp, c_u64, c_vp, c_u64, c_u32, c_bool, c_bool, c_vp]),
    "bc_poseidon_nodes": (c_int, [c_vp, c_u64, c_vp, c_u64, c_vp]),
    "bc_poseidon_build_merkle_tree_nodes": (c_int, [c_vp, c_u64, c_vp, c_u64, c_u32, c_vp]),
    "bc_poseidon_build_merkle_tree": (c_int, [c_vp, c_u64, c_vp, c_u64, c_u32, c_u32, c_vp]),
    "bc_poseidon2_leaves": (c_int, [c_vp, c_u64, c_vp, c_u64, c_u32, c_bool, c_bool, c_vp]),
    "bc_poseidon2_nodes": (c_int, [c_vp, c_u64, c_vp, c_u64, c_vp]),
    "bc_poseidon2_build_merkle_tree_nodes": (c_int, [c_vp, c_u64, c_vp, c_u64, c_u32, c_vp]),
    "bc_poseidon2_build_merkle_tree": (c_int, [c_vp, c_u64, c_vp, c_u64, c_u32, c_u32, c_vp]),
    "bc_poseidon_leaves_cooperative": (c_int, [c_vp, c_u64, c_vp, c_u64, c_u32, c_bool, c_bool, c_vp]),
    "bc_poseidon2_leaves_cooperative": (c_int, [c_vp, c_u64, c_vp, c_u64, c_u32, c_bool, c_bool, c_vp]),
    "bc_poseidon_nodes_cooperative": (c_int, [c_vp, c_u64, c_vp, c_u64, c_vp]),
    "bc_poseidon2_nodes_cooperative": (c_int, [c_vp, c_u64, c_vp, c_u64, c_vp]),
    "bc_gather_rows": (c_int, [c_vp, c_u32, c_u32, c_vp, c_u64, c_u64, c_vp, c_u64, c_u64, c_u32, c_vp]),
    "bc_gather_merkle_paths": (c_int, [c_vp, c_u32, c_vp, c_u64, c_vp, c_u64, c_u32, c_vp]),
    "bc_fill_seeded": (c_int, [c_vp, c_u32, c_u32, c_u64, c_u64, c_u32, c_vp]),
    "bc_server_create": (c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "bc_server_submit": (c_int, [c_vp, c_vp, ctypes.POINTER(c_u64)]),
    "bc_server_wait": (c_int, [c_vp, c_u64]),
    "bc_server_request_info": (c_int, [c_vp, c_u64, c_vp]),
    "bc_server_part": (c_int, [c_vp, c_u64, c_u32, c_vp]),
    "bc_server_release": (c_int, [c_vp, c_u64]),
    "bc_server_drain": (c_int, [c_vp]),
    "bc_server_stats_get": (c_int, [c_vp, c_vp]),
    "bc_server_destroy": (c_int, [c_vp]),
}

BC_SUCCESS = 0
BC_ERROR_INVALID_VALUE = 1
BC_ERROR_NOT_INITIALIZED = 3


class HipError(RuntimeError):
    """A HIP runtime error code returned by the library (CudaResult::Err)."""

    def __init__(self, code, what=""):
        super().__init__(f"{what} failed with HIP error {code}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C era-boojum-cuda_amd")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
