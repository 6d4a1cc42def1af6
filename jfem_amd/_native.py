"""ctypes binding of libjfem_hip.so (the C ABI declared in include/jfem_hip.h).

The library is loaded from this package directory.  There is no fallback:
if the shared object is missing or fails to load, NativeLibraryError is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjfem_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "jfem_hip.h")


class NativeLibraryError(RuntimeError):
    pass


class JfemError(RuntimeError):
    """A C-ABI call returned a non-zero jfem_status."""

    def __init__(self, code, msg):
        super().__init__(f"[jfem status {code}] {msg}")
        self.code = code


class IllegalArgumentException(JfemError, ValueError):
    pass


class IllegalStateException(JfemError):
    pass


class UnsupportedOperationException(JfemError, NotImplementedError):
    pass


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("fell_back", C.c_int), ("resnorm", C.c_double), ("seconds", C.c_double)]


class TrainOpts(C.Structure):
    _fields_ = [("rhs_per_field", C.c_int), ("cycles_per_rhs", C.c_int), ("lr", C.c_double), ("seed", C.c_ulonglong),
                ("mode", C.c_int)]


class StokesOpts(C.Structure):
    _fields_ = [("restart", C.c_int), ("maxit", C.c_long), ("pressure_mass", C.c_void_p), ("mass_its", C.c_int),
                ("mass_lmin", C.c_double), ("mass_lmax", C.c_double), ("component_mg", C.c_void_p * 2)]


class MgOpts(C.Structure):
    _fields_ = [("smoother", C.c_int), ("runs", C.c_int), ("omega", C.c_double), ("degree", C.c_int),
                ("cheb_ratio", C.c_double), ("cheb_safety", C.c_double), ("power_its", C.c_int),
                ("cycles", C.c_int), ("fmg", C.c_int), ("precision", C.c_int), ("galerkin", C.c_int),
                ("raw_transfers", C.c_int), ("matrix_free", C.c_int)]


_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_fp = C.POINTER(C.c_float)

_SIGS = {
    "jfem_last_error": (C.c_char_p, []),
    "jfem_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "jfem_ctx_destroy": (C.c_int, [_vp]),
    "jfem_ctx_sync": (C.c_int, [_vp]),
    "jfem_timer_start": (C.c_int, [_vp]),
    "jfem_timer_stop": (C.c_int, [_vp, _dp]),
    "jfem_copy_bench": (C.c_int, [_vp, C.c_long, C.c_int, _dp]),
    "jfem_read_bench": (C.c_int, [_vp, C.c_long, C.c_int, _dp]),
    "jfem_space_create": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, C.c_int, C.POINTER(_vp)]),
    "jfem_space_destroy": (C.c_int, [_vp]),
    "jfem_space_ndofs": (C.c_long, [_vp]),
    "jfem_space_cell_dofs": (C.c_int, [_vp, _ip]),
    "jfem_space_boundary_nodes": (C.c_int, [_vp, _ip, _lp]),
    "jfem_space_node_coords": (C.c_int, [_vp, _dp]),
    "jfem_space_evaluate": (C.c_int, [_vp, _vp, C.c_long, _dp, _dp]),
    "jfem_fe_l2_difference": (C.c_int, [_vp, _vp, _vp, _vp, C.c_long, _dp, _dp]),
    "jfem_vec_create": (C.c_int, [_vp, C.c_long, C.POINTER(_vp)]),
    "jfem_vec_destroy": (C.c_int, [_vp]),
    "jfem_vec_upload": (C.c_int, [_vp, _dp]),
    "jfem_vec_download": (C.c_int, [_vp, _dp]),
    "jfem_vec_fill": (C.c_int, [_vp, C.c_double]),
    "jfem_vec_dot": (C.c_int, [_vp, _vp, _dp]),
    "jfem_vec_axpby": (C.c_int, [C.c_double, _vp, C.c_double, _vp]),
    "jfem_vec_copy": (C.c_int, [_vp, _vp]),
    "jfem_assemble": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, C.c_int, C.POINTER(_vp)]),
    "jfem_assemble_rhs_sines": (C.c_int, [_vp, C.c_int, _ip, _dp, C.c_int, _vp]),
    "jfem_assemble_rhs_trig": (C.c_int, [_vp, C.c_int, _ip, _ip, _dp, C.c_int, _vp]),
    "jfem_space_lex_rank": (C.c_int, [_vp, _ip]),
    "jfem_th_assemble": (C.c_int, [_vp, _vp, C.c_int, _ip, _dp, C.c_int, C.POINTER(_vp)]),
    "jfem_th_assemble_rhs_trig": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _ip, _ip, _dp, C.c_int, _vp]),
    "jfem_mat_from_csr": (C.c_int, [_vp, C.c_long, C.c_long, _lp, _ip, _dp, C.POINTER(_vp)]),
    "jfem_mat_destroy": (C.c_int, [_vp]),
    "jfem_mat_info": (C.c_int, [_vp, _lp, _lp, _lp]),
    "jfem_mat_export_csr": (C.c_int, [_vp, _lp, _ip, _dp]),
    "jfem_apply_dirichlet": (C.c_int, [_vp, _vp, C.c_long, _ip, _dp]),
    "jfem_spmv_t": (C.c_int, [_vp, _vp, _vp]),
    "jfem_spmv": (C.c_int, [_vp, _vp, _vp]),
    "jfem_spmv_f32": (C.c_int, [_vp, _fp, _fp]),
    "jfem_spmv_bench": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _dp]),
    "jfem_solve_cg": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_long, C.POINTER(SolveInfo), _dp]),
    "jfem_solve_bicgstab": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_long, C.POINTER(SolveInfo)]),
    "jfem_solve_gmres": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, C.POINTER(SolveInfo)]),
    "jfem_smooth": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, _vp, _vp]),
    "jfem_smooth_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, _vp,
                                      _vp]),
    "jfem_solve_gmres2": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_long, C.POINTER(SolveInfo)]),
    "jfem_solve_fcg": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_long, C.POINTER(SolveInfo), _dp]),
    "jfem_solve_stokes_fgmres": (C.c_int, [_vp, _vp, _vp, C.POINTER(StokesOpts), _vp, _vp, C.c_double,
                                            C.POINTER(SolveInfo), _dp]),
    "jfem_mg_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, _ip, _dp, _dp, C.c_int, C.POINTER(MgOpts),
                                 C.POINTER(_vp)]),
    "jfem_mg_destroy": (C.c_int, [_vp]),
    "jfem_mg_level_matrix": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "jfem_mg_level_apply": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "jfem_mg_level_bench": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _dp]),
    "jfem_mg_level_matrix_free": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "jfem_mg_apply": (C.c_int, [_vp, _vp, _vp]),
    "jfem_mg_vcycle": (C.c_int, [_vp, _vp, _vp]),
    "jfem_mg_wcycle": (C.c_int, [_vp, _vp, _vp]),
    "jfem_mg_lmax": (C.c_int, [_vp, _dp]),
    "jfem_mg_prolongate": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "jfem_mg_restrict": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "jfem_nn_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(_vp)]),
    "jfem_nn_destroy": (C.c_int, [_vp]),
    "jfem_nn_num_params": (C.c_long, [_vp]),
    "jfem_nn_get_params": (C.c_int, [_vp, _fp]),
    "jfem_nn_set_params": (C.c_int, [_vp, _fp]),
    "jfem_nn_forward_host": (C.c_int, [_vp, _ip, _fp, _fp]),
    "jfem_nn_gradient_host": (C.c_int, [_vp, _ip, _fp, _fp, _fp]),
    "jfem_nn_adam_step": (C.c_int, [_vp, _fp, C.c_float]),
    "jfem_nn_save": (C.c_int, [_vp, C.c_char_p]),
    "jfem_nn_load": (C.c_int, [_vp, C.c_char_p, C.POINTER(_vp)]),
    "jfem_mg_set_learned": (C.c_int, [_vp, _vp, _dp, C.c_int, C.c_double, C.c_int]),
    "jfem_mg_set_learned_active": (C.c_int, [_vp, C.c_int]),
    "jfem_mg_set_learned_sides": (C.c_int, [_vp, C.c_int]),
    "jfem_mg_learned_status": (C.c_int, [_vp, _ip, _ip]),
    "jfem_nn_train_mg": (C.c_int, [_vp, C.POINTER(TrainOpts), _dp, _ip]),
    "jfem_mg_solve": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int, C.POINTER(SolveInfo), _dp]),
}

_lib = None


def declared_symbols():
    """Function names declared in include/jfem_hip.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(jfem_\w+)\s*\(", txt, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryError(f"{LIB_PATH} is missing: build it with `python jfem_amd/build.py`")
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:
            raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


_EXC = {1: IllegalArgumentException, 2: IllegalStateException, 3: UnsupportedOperationException}


def check(rc):
    if rc != 0:
        msg = lib().jfem_last_error().decode(errors="replace")
        raise _EXC.get(rc, JfemError)(rc, msg)


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def lptr(a):
    return a.ctypes.data_as(_lp)


def fptr(a):
    return a.ctypes.data_as(_fp)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
