"""ctypes binding of libdsv.so (declared in include/dsv.h).

The library is loaded from this package directory.  There is no fallback: if libdsv.so is
missing or fails to load, every entry point raises ``DsvLibraryError``.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdsv.so")

DSV_OK = 0
DSV_E_ARG = -1
DSV_E_MESH = -2
DSV_E_HIP = -3
DSV_E_NOMEM = -4
DSV_E_BREAKDOWN = -5

METHOD_CSR64 = 0
METHOD_MF32_IR64 = 1
METHOD_MF64 = 2
METHODS = {"csr64": METHOD_CSR64, "mf32_ir64": METHOD_MF32_IR64, "mf64": METHOD_MF64}

# every symbol include/dsv.h declares
EXPORTS = (
    "dsv_last_error", "dsv_device_count", "dsv_mesh_create", "dsv_mesh_destroy", "dsv_mesh_info",
    "dsv_mesh_view", "dsv_mesh_copy_view", "dsv_mesh_set_partition", "dsv_assemble", "dsv_rhs",
    "dsv_rhs_batch",
    "dsv_apply", "dsv_solve", "dsv_set_solver_options", "dsv_set_operator_form",
    "dsv_set_species_narrowing",
    "dsv_set_operator_reuse", "dsv_time_apply",
    "dsv_residual_norms", "dsv_warm_start", "dsv_solve_warm", "dsv_solve_warm_history",
    "dsv_step_warm_history", "dsv_build_low_modes",
    "dsv_mlp_features",
    "dsv_mlp_forward", "dsv_mlp_directions",
    "dsv_mlp_grad_workspace", "dsv_mlp_grad", "dsv_mlp_grad_noalloc", "dsv_mlp_adam", "dsv_nodal_gradient",
    "dsv_transport_coefficients", "dsv_effective_diffusion", "dsv_zinc_coefficients",
    "dsv_film_update", "dsv_under_relax",
    "dsv_interface_velocity", "dsv_levelset_advance", "dsv_mesh_h_avg", "dsv_redistance",
    "dsv_scaffold_volume",
)


class DsvLibraryError(RuntimeError):
    pass


class DsvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libdsv error {code}: {msg}")
        self.code = code


class Species(ctypes.Structure):
    _fields_ = [
        ("D", ctypes.c_void_p), ("D_len", ctypes.c_int32), ("D_value", ctypes.c_double),
        ("alpha", ctypes.c_void_p), ("alpha_len", ctypes.c_int32), ("alpha_value", ctypes.c_double),
        ("bc_dofs", ctypes.c_void_p), ("n_bc", ctypes.c_int32), ("bc_value", ctypes.c_double),
        ("b", ctypes.c_void_p), ("x", ctypes.c_void_p),
    ]


class CoefParams(ctypes.Structure):
    """dsv_coef_params (include/dsv.h)."""
    _fields_ = [("eps", ctypes.c_double), ("k_orr", ctypes.c_double), ("kd", ctypes.c_double),
                ("Fmax", ctypes.c_double), ("film_floor", ctypes.c_double)]


class FilmParams(ctypes.Structure):
    """dsv_film_params (include/dsv.h)."""
    _fields_ = [("kf", ctypes.c_double), ("kd", ctypes.c_double), ("Fmax", ctypes.c_double)]


class VelocityParams(ctypes.Structure):
    """dsv_velocity_params (include/dsv.h)."""
    _fields_ = [("h", ctypes.c_double), ("k_orr", ctypes.c_double),
                ("zn_solid", ctypes.c_double), ("zn_sat", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32 * 3), ("outer", ctypes.c_int32 * 3),
        ("converged", ctypes.c_int32 * 3), ("rel_residual", ctypes.c_double * 3),
        ("initial_rel_residual", ctypes.c_double * 3), ("solve_ms", ctypes.c_double),
        ("cg_launches", ctypes.c_int32), ("narrowings", ctypes.c_int32),
        ("dirs_reused", ctypes.c_int32), ("gram_reused", ctypes.c_int32),
        ("low_modes", ctypes.c_int32),
    ]

    def as_dict(self, n):
        return {
            "iterations": list(self.iterations[:n]), "outer": list(self.outer[:n]),
            "converged": [bool(c) for c in self.converged[:n]],
            "rel_residual": list(self.rel_residual[:n]),
            "initial_rel_residual": list(self.initial_rel_residual[:n]),
            "solve_ms": self.solve_ms, "cg_launches": self.cg_launches,
            "narrowings": self.narrowings, "dirs_reused": self.dirs_reused,
            "gram_reused": self.gram_reused, "low_modes": self.low_modes,
        }


_lib = None


def load():
    """Load libdsv.so once and declare the signatures; raises DsvLibraryError if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DsvLibraryError(f"{LIB_PATH} not built (run __graft_entry__.build() or make -C csrc)")
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise DsvLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    P = ctypes.c_void_p
    I32 = ctypes.c_int32
    D = ctypes.c_double
    sig = {
        "dsv_last_error": (ctypes.c_char_p, []),
        "dsv_device_count": (ctypes.c_int, [ctypes.POINTER(I32)]),
        "dsv_mesh_create": (ctypes.c_int, [ctypes.POINTER(P), I32, I32, P, P, I32]),
        "dsv_mesh_destroy": (ctypes.c_int, [P]),
        "dsv_mesh_info": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int64)]),
        "dsv_mesh_view": (ctypes.c_int, [P, I32, ctypes.POINTER(P)]),
        "dsv_mesh_copy_view": (ctypes.c_int, [P, I32, P, P]),
        "dsv_mesh_set_partition": (ctypes.c_int, [P, I32, I32, I32]),
        "dsv_assemble": (ctypes.c_int, [P, ctypes.POINTER(Species), D, I32, P, P]),
        "dsv_rhs": (ctypes.c_int, [P, P, P, I32, D, I32, P, I32, D, P, P]),
        "dsv_rhs_batch": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), ctypes.POINTER(P),
                                         ctypes.POINTER(P), D, I32, ctypes.POINTER(P), P]),
        "dsv_apply": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), D, I32, ctypes.POINTER(P),
                                     ctypes.POINTER(P), I32, P]),
        "dsv_solve": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), D, I32, I32, D, I32, P,
                                     ctypes.POINTER(Stats)]),
        "dsv_set_solver_options": (ctypes.c_int, [P, D, I32, I32, I32]),
        "dsv_set_operator_form": (ctypes.c_int, [P, I32]),
        "dsv_set_species_narrowing": (ctypes.c_int, [P, I32]),
        "dsv_set_operator_reuse": (ctypes.c_int, [P, I32]),
        "dsv_time_apply": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), D, I32, I32, I32, I32,
                                          P, ctypes.POINTER(D)]),
        "dsv_residual_norms": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), D, I32,
                                              ctypes.POINTER(D), ctypes.POINTER(D), P]),
        "dsv_warm_start": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), D, I32, I32,
                                          ctypes.POINTER(P), ctypes.POINTER(D), ctypes.POINTER(D),
                                          ctypes.POINTER(D), P]),
        "dsv_solve_warm": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), D, I32, I32, D, I32,
                                          I32, ctypes.POINTER(P), I32, ctypes.POINTER(P), P,
                                          ctypes.POINTER(Stats), ctypes.POINTER(D),
                                          ctypes.POINTER(D), ctypes.POINTER(D)]),
        "dsv_solve_warm_history": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), D, I32, I32, D,
                                                  I32, I32, ctypes.POINTER(P), I32,
                                                  ctypes.POINTER(P), P, ctypes.POINTER(Stats),
                                                  ctypes.POINTER(D), ctypes.POINTER(D),
                                                  ctypes.POINTER(D)]),
        "dsv_step_warm_history": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), ctypes.POINTER(P),
                                                 D, I32, I32, D, I32, I32, ctypes.POINTER(P), I32,
                                                 ctypes.POINTER(P), P, ctypes.POINTER(Stats),
                                                 ctypes.POINTER(D), ctypes.POINTER(D),
                                                 ctypes.POINTER(D)]),
        "dsv_build_low_modes": (ctypes.c_int, [P, I32, ctypes.POINTER(Species), D, I32, I32, I32,
                                               I32, ctypes.c_uint64, ctypes.POINTER(D), P]),
        "dsv_nodal_gradient": (ctypes.c_int, [P, P, P, P, P]),
        "dsv_transport_coefficients": (ctypes.c_int, [P, P, P, P, P, P, ctypes.POINTER(CoefParams),
                                                      P, P, P, P, P]),
        "dsv_effective_diffusion": (ctypes.c_int, [P, P, P, ctypes.POINTER(CoefParams),
                                                   ctypes.POINTER(D), P, P, P, P, P]),
        "dsv_zinc_coefficients": (ctypes.c_int, [P, P, P, P, ctypes.POINTER(FilmParams), P, P, P,
                                                 ctypes.POINTER(I32), P]),
        "dsv_film_update": (ctypes.c_int, [P, P, P, P, P, ctypes.POINTER(FilmParams), D, P, P]),
        "dsv_under_relax": (ctypes.c_int, [I32, P, P, D, P, P]),
        "dsv_interface_velocity": (ctypes.c_int, [P, P, P, P, P, P, ctypes.POINTER(VelocityParams),
                                                  P, P]),
        "dsv_levelset_advance": (ctypes.c_int, [P, P, P, D, P, P]),
        "dsv_mesh_h_avg": (ctypes.c_int, [P, ctypes.POINTER(D)]),
        "dsv_redistance": (ctypes.c_int, [P, P, I32, D, P]),
        "dsv_scaffold_volume": (ctypes.c_int, [P, P, I32, D, P, ctypes.POINTER(D), P]),
        "dsv_mlp_features": (ctypes.c_int, [I32, P, P, P, P, P, D, D, D, D, P, P, P, P]),
        "dsv_mlp_forward": (ctypes.c_int, [P, P, I32, P, P]),
        "dsv_mlp_directions": (ctypes.c_int, [I32, I32, P, P, D, ctypes.POINTER(P), P]),
        "dsv_mlp_grad_workspace": (ctypes.c_size_t, [I32]),
        "dsv_mlp_grad": (ctypes.c_int, [P, P, P, I32, P, P, P, ctypes.c_size_t, P]),
        "dsv_mlp_grad_noalloc": (ctypes.c_int, [P, P, P, I32, P, P, P]),
        "dsv_mlp_adam": (ctypes.c_int, [P, P, P, P, I32, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_float, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != DSV_OK:
        msg = load().dsv_last_error().decode(errors="replace")
        if rc == DSV_E_ARG:
            raise ValueError(msg)
        raise DsvError(rc, msg)


def call(name, *args):
    check(getattr(load(), name)(*args))
