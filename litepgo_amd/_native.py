"""ctypes binding of liblitepgo_fem.so (the C ABI declared in include/litepgo_fem.h and litepgo_fem_ext.h).

There is no CPU fallback: if the library is missing or no gfx950 device is usable, the calls fail loudly.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liblitepgo_fem.so"

# Every exported symbol, with (restype, argtypes).  The C-ABI test checks this list against include/*.h.
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
SIGNATURES = {
    "lpg_last_error": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
    "lpg_tet_energy_create": (ctypes.c_int, [_I32, _I32, _P, _P, _P, _P, _I32, ctypes.POINTER(_P)]),
    "lpg_tet_energy_destroy": (None, [_P]),
    "lpg_tet_energy_num_dofs": (_I32, [_P]),
    "lpg_tet_energy_num_tets": (_I32, [_P]),
    "lpg_tet_energy_nnz": (_I32, [_P]),
    "lpg_tet_energy_pattern": (ctypes.c_int, [_P, _P, _P]),
    "lpg_tet_energy_slot_map": (ctypes.c_int, [_P, _P]),
    "lpg_tet_energy_eval": (ctypes.c_int, [_P, _P, _I32, _I32, _P, _P, _P]),
    "lpg_tet_energy_eval_device": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _P, _P, _P, _P]),
    "lpg_tet_energy_last_indefinite_count": (_I32, [_P]),
    "lpg_tet_energy_max_feasible_step": (ctypes.c_int, [_P, _P, _P, _P]),
    "lpg_tet_energy_max_feasible_step_device": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "lpg_tet_energy_dE_dtheta_device": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "lpg_tet_energy_force_vjp_theta_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "lpg_tet_energy_force_jvp_theta_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "lpg_tet_energy_force_fp_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "lpg_tet_energy_force_vjp_fp_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "lpg_tet_energy_force_jvp_fp_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "lpg_tet_energy_hvp_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "lpg_tet_energy_vjp_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "lpg_tet_energy_set_launch": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32]),
    "lpg_tet_energy_get_launch": (ctypes.c_int, [_P, _P]),
    "lpg_tet_energy_max_degree": (_I32, [_P]),
    "lpg_tet_energy_set_timing": (ctypes.c_int, [_P, _I32]),
    "lpg_tet_energy_kernel_times": (ctypes.c_int, [_P, _P, _P]),
    "lpg_lame_device": (ctypes.c_int, [_P, _P, ctypes.c_int64, _P, _P]),
    "lpg_lame_vjp_device": (ctypes.c_int, [_P, _P, _P, ctypes.c_int64, _P, _P, _P]),
    "lpg_tet_batch_create": (ctypes.c_int, [_I32, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "lpg_tet_batch_num_meshes": (_I32, [_P]),
    "lpg_tet_batch_offsets": (ctypes.c_int, [_P, _P, _P, _P]),
    "lpg_tet_batch_eval_device": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _P, _P, _P, _P]),
    "lpg_tet_batch_max_feasible_step_device": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "lpg_tet_energy_eval_enqueue": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _P, _P, _P, _P]),
    "lpg_tet_energy_eval_wait": (ctypes.c_int, [_P, _P]),
    "lpg_tet_energy_eval_device_noalloc": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _P, _P, _P, _P]),
    "lpg_tet_energy_set_graph": (ctypes.c_int, [_P, _I32]),
    "lpg_tet_energy_autotune": (ctypes.c_int, [_P, _P, _I32, _I32, _P, _P]),
    "lpg_newton_default_options": (None, [_P]),
    "lpg_tet_newton_solve_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "lpg_tet_equilibrium_vjp_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "lpg_tet_equilibrium_jvp_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
}

LPG_OK = 0
LPG_INVALID_ARGUMENT = 1
LPG_DOMAIN_ERROR = 2
LPG_LOGIC_ERROR = 3
LPG_BUSY = 4
LPG_DEVICE_ERROR = 5


class DomainError(ValueError):
    """std::domain_error of the reference (nanobind maps it to ValueError)."""


class BusyError(RuntimeError):
    """An overlapping call on one energy object (the reference's BusyScope logic_error)."""


class DeviceError(RuntimeError):
    """HIP runtime failure or no usable gfx950 device."""


_EXC = {LPG_INVALID_ARGUMENT: ValueError, LPG_DOMAIN_ERROR: DomainError, LPG_LOGIC_ERROR: RuntimeError,
        LPG_BUSY: BusyError, LPG_DEVICE_ERROR: DeviceError}

_lib = None


def lib():
    """The loaded library; raises if it has not been built (run __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                               "g.build()'` (there is no CPU fallback)")
        # One HIP runtime per process: when PyTorch is present, its runtime is loaded first so that the library's
        # libamdhip64 dependency binds to it (by soname) and device pointers of torch tensors are the library's.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().lpg_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code: int, what: str = "") -> None:
    if code != LPG_OK:
        msg = last_error()
        raise _EXC.get(code, RuntimeError)(f"{what}: {msg}" if what else msg)
