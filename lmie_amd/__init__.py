"""lmie_amd -- MI355X (gfx950) size-integrated Lorenz-Mie solver.

The product is the C library ``lmie_amd/lib/liblmie_amd.so``: the drop-in C ABI
of include/lmie_interface.h (same structs and ``lmie_solution`` call as the
reference, src/lmie_interface.h:20-157) over hand-written HIP kernels, plus
the extensions of include/lmie_hip.h.  This package only loads it and wraps
it with ctypes; there is no Python or CPU implementation of the solver here,
and loading fails loudly when the library has not been built.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liblmie_amd.so")

_lib = None


def load():
    """Load (once) and return the product library with its ABI declared."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "liblmie_amd.so is not built (%s); run __graft_entry__.build() or "
                "make -C lmie_amd/csrc" % LIB_PATH)
        from . import lmie
        _lib = lmie.bind_all(C.CDLL(LIB_PATH))
    return _lib


EXPORTED = (
    # drop-in ABI, include/lmie_interface.h
    "lmie_calc_max_coef", "lmie_in_alloc", "lmie_in_free", "lmie_out_alloc", "lmie_out_free",
    "lmie_in_zero_derivs", "lmie_solution", "lmie_solution2", "lmie_solution2_l",
    "lmie_solution_slave",
    # include/lmie_hip.h
    "lmie_core_solution", "lmie_core_solution_ex", "lmie_solution_ex", "lmie_solution_mix",
    "lmie_solution_adjoint", "lmie_last_error", "lmie_set_timing", "lmie_get_timing",
    # include/lmie_gpu.h
    "lmie_gpu_forward", "lmie_gpu_adjoint", "lmie_gpu_phase_func", "lmie_gpu_phase_func_static",
    "lmie_gpu_last_error",
)
