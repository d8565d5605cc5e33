"""Mirror of boojum-cuda's `context` module (reference: src/context.rs)."""
import ctypes

from ._lib import HipError, lib, BC_ERROR_INVALID_VALUE

OMEGA_LOG_ORDER = None  # filled from the library on first use


def omega_log_order():
    global OMEGA_LOG_ORDER
    if OMEGA_LOG_ORDER is None:
        OMEGA_LOG_ORDER = int(lib().bc_omega_log_order())
    return OMEGA_LOG_ORDER


class Context:
    """Power tables for the NTT (reference: src/context.rs:157-297).

    `Context.create(w_coarse_log, g_coarse_log)` builds the tables on the
    current device; they stay the process-wide tables of that device until
    `destroy()`.
    """

    def __init__(self, handle):
        self._handle = handle

    @classmethod
    def create(cls, powers_of_w_coarse_log_count, powers_of_g_coarse_log_count):
        # src/context.rs:175-176
        order = omega_log_order()
        assert powers_of_w_coarse_log_count <= order
        assert powers_of_g_coarse_log_count <= order
        h = ctypes.c_void_p()
        rc = lib().bc_context_create(powers_of_w_coarse_log_count, powers_of_g_coarse_log_count, ctypes.byref(h))
        if rc != 0:
            if rc == BC_ERROR_INVALID_VALUE:
                raise AssertionError("Context::create arguments out of range")
            raise HipError(rc, "Context::create")
        return cls(h)

    def destroy(self):
        if self._handle is None:
            return
        rc = lib().bc_context_destroy(self._handle)
        self._handle = None
        if rc != 0:
            raise HipError(rc, "Context::destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
        return False
