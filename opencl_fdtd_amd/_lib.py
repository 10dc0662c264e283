"""ctypes binding of ``libfdtd_hip.so`` (declared in ``include/fdtd.h``).

The structures below mirror the C declarations field for field.  Loading
fails loudly: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfdtd_hip.so")

ABI_VERSION = 7
MAX_FACES = 12

F32, F64 = 0, 1
POYNTING_OFF, POYNTING_FUSED, POYNTING_TWO_PASS = 0, 1, 2
BLOCH_PHASE, BLOCH_DECOUPLED = 0, 1
CHAIN_H_PENDING, CHAIN_DEFER_H, MOVED_H_PENDING = 1, 2, 8

# Every symbol include/fdtd.h declares.
EXPORTS = (
    "fdtd_abi_version", "fdtd_last_error", "fdtd_plan_create", "fdtd_plan_destroy", "fdtd_plan_describe",
    "fdtd_psi_elems", "fdtd_coef_elems", "fdtd_update_e", "fdtd_update_h",
    "fdtd_update_s", "fdtd_update_j", "fdtd_add_point", "fdtd_step_n", "fdtd_step_n_moved", "fdtd_step_n_observed",
    "fdtd_step_n_chain",
    "fdtd_update_he", "fdtd_update_he_supported", "fdtd_update_he2", "fdtd_update_he2_supported",
    "fdtd_update_he2_preferred",
    "fdtd_adjoint_step", "fdtd_plane_sum", "fdtd_reduce_partials", "fdtd_energy", "fdtd_energy_noalloc",
    "fdtd_shell_weight_elems", "fdtd_shell_hidden_elems", "fdtd_shell_partials_elems",
    "fdtd_shell_apply", "fdtd_shell_step_n", "fdtd_shell_backward",
)


class Face(ctypes.Structure):
    _fields_ = [("axis", ctypes.c_int32), ("polarity", ctypes.c_int32),
                ("thickness", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Desc(ctypes.Structure):
    _fields_ = [
        ("shape", ctypes.c_int64 * 3),
        ("dtype", ctypes.c_int32), ("n_face", ctypes.c_int32),
        ("face", Face * MAX_FACES),
        ("dt", ctypes.c_double), ("inv_dx", ctypes.c_double),
        ("nonuniform", ctypes.c_int32), ("poynting", ctypes.c_int32),
        ("bloch_mask", ctypes.c_int32), ("bloch_mode", ctypes.c_int32),
        ("bloch_re", ctypes.c_double * 3), ("bloch_im", ctypes.c_double * 3),
    ]


class Fields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "E", "H", "H_out", "eps", "S", "oS", "F", "G", "G_out",
        "psi_e", "psi_h", "psi_h_out", "psi_f", "psi_g", "psi_g_out", "coef", "inv_d", "E_out", "psi_e_out", "F_out")]


class AdjointBufs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "E_prev", "E_next", "H_prev", "H_next", "eps", "Eb", "Hb",
        "psib_e", "psib_e_out", "psib_h", "psib_h_out", "coef", "eps_grad")] + [
        ("src_index", ctypes.c_int64), ("src_amp", ctypes.c_double),
        ("obj_axis", ctypes.c_int32), ("obj_index", ctypes.c_int32),
        ("parts", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("grad_lo", ctypes.c_int32 * 3), ("grad_hi", ctypes.c_int32 * 3)]


class Observe(ctypes.Structure):
    _fields_ = [("flux_acc", ctypes.c_void_p), ("flux_comp", ctypes.c_int32), ("flux_axis", ctypes.c_int32),
                ("flux_index", ctypes.c_int64), ("flux_partials", ctypes.c_void_p),
                ("E_keep", ctypes.c_void_p), ("H_keep", ctypes.c_void_p)]


class FdtdError(RuntimeError):
    pass


_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FdtdError(
            "libfdtd_hip.so is not built (expected at %s); run __graft_entry__.build() "
            "or python -m opencl_fdtd_amd.build" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    F = ctypes.POINTER(Fields)
    lib.fdtd_abi_version.restype = ctypes.c_int
    lib.fdtd_abi_version.argtypes = []
    lib.fdtd_last_error.restype = ctypes.c_char_p
    lib.fdtd_last_error.argtypes = []
    lib.fdtd_plan_create.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(ctypes.c_void_p)]
    lib.fdtd_plan_destroy.argtypes = [P]
    lib.fdtd_plan_describe.argtypes = [P, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int64]
    lib.fdtd_plan_describe.restype = ctypes.c_int64
    lib.fdtd_psi_elems.argtypes = [P]
    lib.fdtd_psi_elems.restype = ctypes.c_int64
    lib.fdtd_coef_elems.argtypes = [P]
    lib.fdtd_coef_elems.restype = ctypes.c_int64
    for name in ("fdtd_update_e", "fdtd_update_h", "fdtd_update_s"):
        getattr(lib, name).argtypes = [P, F, P]
    lib.fdtd_update_j.argtypes = [P, F, P, P, ctypes.c_double, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_int64), P]
    lib.fdtd_add_point.argtypes = [P, P, ctypes.c_int64, P, P]
    lib.fdtd_step_n.argtypes = [P, F, ctypes.c_int32, ctypes.c_int64, P, P]
    lib.fdtd_step_n_moved.argtypes = [P, F, ctypes.c_int32, ctypes.c_int64, P, P, ctypes.POINTER(ctypes.c_int32)]
    lib.fdtd_step_n_chain.argtypes = [P, F, ctypes.c_int32, ctypes.c_int64, P, ctypes.c_int32, P,
                                      ctypes.POINTER(ctypes.c_int32)]
    lib.fdtd_step_n_observed.argtypes = [P, F, ctypes.c_int32, ctypes.c_int64, P, ctypes.POINTER(Observe), P,
                                         ctypes.POINTER(ctypes.c_int32)]
    lib.fdtd_update_he.argtypes = [P, F, ctypes.c_int64, P, P]
    lib.fdtd_update_he_supported.argtypes = [P]
    lib.fdtd_update_he2.argtypes = [P, F, ctypes.c_int64, P, P]
    lib.fdtd_update_he2_supported.argtypes = [P]
    lib.fdtd_update_he2_preferred.argtypes = [P]
    lib.fdtd_adjoint_step.argtypes = [P, ctypes.POINTER(AdjointBufs), P]
    lib.fdtd_plane_sum.argtypes = [P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, P, P]
    lib.fdtd_reduce_partials.argtypes = [P]
    lib.fdtd_reduce_partials.restype = ctypes.c_int64
    lib.fdtd_energy.argtypes = [P, P, P, P, P, P, P, P]
    lib.fdtd_energy_noalloc.argtypes = [P, P, P, P, P, P, P]
    lib.fdtd_shell_weight_elems.argtypes = []
    lib.fdtd_shell_hidden_elems.argtypes = [P, ctypes.c_int32]
    lib.fdtd_shell_partials_elems.argtypes = []
    for name in ("fdtd_shell_weight_elems", "fdtd_shell_hidden_elems", "fdtd_shell_partials_elems"):
        getattr(lib, name).restype = ctypes.c_int64
    lib.fdtd_shell_apply.argtypes = [P, ctypes.c_int32, P, P, P, P, P]
    lib.fdtd_shell_backward.argtypes = [P, ctypes.c_int32] + [P] * 10 + [P]
    lib.fdtd_shell_step_n.argtypes = [P, F, ctypes.c_int32, P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, P, P,
                                      ctypes.POINTER(ctypes.c_int32)]
    sized = ("fdtd_last_error", "fdtd_plan_describe", "fdtd_psi_elems", "fdtd_coef_elems", "fdtd_reduce_partials",
             "fdtd_shell_weight_elems", "fdtd_shell_hidden_elems", "fdtd_shell_partials_elems")
    for name in EXPORTS:
        if name not in sized:
            getattr(lib, name).restype = ctypes.c_int
    v = lib.fdtd_abi_version()
    if v != ABI_VERSION:
        raise FdtdError("libfdtd_hip.so ABI version %d, expected %d" % (v, ABI_VERSION))
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().fdtd_last_error().decode(errors="replace")
        raise FdtdError("%s failed (status %d): %s" % (what, rc, msg))
