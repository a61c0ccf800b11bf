"""ctypes binding of libddcx.so (the C ABI declared in include/ddcx.h).

The shared library is built in-tree by ``ddc_amd/csrc/Makefile`` (hipcc,
gfx950).  There is no fallback: if the library is missing or fails to load,
every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libddcx.so")

DDCX_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "GRID", 3: "SINGULAR", 4: "HIP", 5: "UNSUPPORTED"}


class ddcx_layout(C.Structure):
    _fields_ = [("row_stride", C.c_int64), ("inner_count", C.c_int64), ("inner_stride", C.c_int64),
                ("outer_stride", C.c_int64)]


class ddcx_bsplines_desc(C.Structure):
    _fields_ = [("degree", C.c_int), ("uniform", C.c_int), ("periodic", C.c_int), ("ncells", C.c_int),
                ("rmin", C.c_double), ("rmax", C.c_double), ("breaks", C.c_void_p)]


class ddcx_builder_desc(C.Structure):
    _fields_ = [("lower", C.c_int), ("upper", C.c_int), ("solver", C.c_int), ("precision", C.c_int),
                ("points", C.c_void_p), ("npoints", C.c_int), ("refine_steps", C.c_int)]


class ddcx_builder_info(C.Structure):
    _fields_ = [("npoints", C.c_int), ("nbasis", C.c_int), ("size", C.c_int), ("offset", C.c_int),
                ("nbv_lower", C.c_int), ("nbv_upper", C.c_int), ("matrix_kind", C.c_int), ("blocked", C.c_int),
                ("main_size", C.c_int), ("border", C.c_int), ("dx", C.c_double), ("kernel", C.c_int),
                ("pivoted", C.c_int), ("refine_steps", C.c_int), ("refine_contraction", C.c_double)]


class ddcx_extrap_rule(C.Structure):
    _fields_ = [("kind", C.c_int), ("eval_pos", C.c_double), ("eval_pos_not_interest_min", C.c_double),
                ("eval_pos_not_interest_max", C.c_double)]


class ddcx_problem_desc(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("kl", C.c_int), ("ku", C.c_int), ("pds", C.c_int),
                ("bottom_size", C.c_int), ("top_size", C.c_int)]


class ddcx_problem_info(C.Structure):
    _fields_ = [("size", C.c_int), ("rhs_rows", C.c_int), ("main_kind", C.c_int), ("blocked", C.c_int),
                ("main_size", C.c_int), ("pivoted", C.c_int)]


# name -> (argtypes) ; all return ddcx_status (int) unless listed in _RESTYPES
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_L = ddcx_layout
_I64 = C.c_int64
SIGNATURES = {
    "ddcx_bsplines_create": [C.POINTER(ddcx_bsplines_desc), _PP],
    "ddcx_bsplines_destroy": [_P],
    "ddcx_bsplines_knots": [_P, _P],
    "ddcx_bsplines_integrals": [_P, _P],
    "ddcx_builder_create": [_P, C.POINTER(ddcx_builder_desc), _PP],
    "ddcx_builder_destroy": [_P],
    "ddcx_builder_get_info": [_P, C.POINTER(ddcx_builder_info)],
    "ddcx_builder_points": [_P, _P],
    "ddcx_builder_build": [_P, _P, _L, _P, _L, _P, _P, _L, _I64, _P],
    "ddcx_builder_build_vjp": [_P, _P, _L, _P, _L, _P, _P, _L, _I64, _P],
    "ddcx_builder_quadrature": [_P, _P, _P, _P],
    "ddcx_evaluator_create": [_P, C.c_int, C.c_int, _PP],
    "ddcx_evaluator_create_with_rules": [_P, C.POINTER(ddcx_extrap_rule), C.POINTER(ddcx_extrap_rule), _PP],
    "ddcx_evaluator_destroy": [_P],
    "ddcx_evaluator_eval": [_P, _P, _P, _L, _I64, _P, _L, _I64, _P],
    "ddcx_evaluator_deriv": [_P, C.c_int, _P, _P, _L, _I64, _P, _L, _I64, _P],
    "ddcx_evaluator_eval_at": [_P, C.c_int, _P, _L, _P, _I64, _P, _L, _I64, _P],
    "ddcx_evaluator_device_view": [_P, _P],
    "ddcx_evaluator_integrate": [_P, _P, _I64, _P, _L, _I64, _P],
    "ddcx_evaluator_eval_vjp": [_P, _P, _P, _P, _L, _I64, _P, _P, _L, _I64, _P],
    "ddcx_evaluator2d_eval": [_P, _P, C.c_int, C.c_int, _P, _P, _I64, _P, _I64, _I64, _I64, _I64, _P],
    "ddcx_evaluator2d_eval_vjp": [_P, _P, _P, _P, _P, _I64, _P, _P, _I64, _I64, _I64, _I64, _P],
    "ddcx_evaluator2d_eval_workspace": [_P, _P, _I64, _P],
    "ddcx_evaluator2d_eval_packed": [_P, _P, C.c_int, C.c_int, _P, _P, _I64, _P, _I64, _I64, _I64, _I64, _P, _I64, _P],
    "ddcx_build_eval": [_P, _P, _P, _L, _P, _L, _P, _P, _L, _P, _P, _L, _I64, _I64, _P],
    "ddcx_advect": [_P, _P, _P, _L, _P, _L, _P, _P, _L, _P, _P, _L, _I64, _P],
    "ddcx_evaluatornd_eval": [_P, C.c_int, _P, _P, _P, _I64, _P, _I64, _P, _I64, _P],
    "ddcx_evaluatornd_eval_workspace": [_P, C.c_int, _I64, _I64, _P],
    "ddcx_evaluatornd_eval_binned": [_P, C.c_int, _P, _P, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P],
    "ddcx_evaluator2d_eval_vjp_workspace": [_P, _P, _I64, _I64, _P],
    "ddcx_evaluator2d_eval_vjp_binned": [_P, _P, _P, _P, _P, _I64, _P, _P, _I64, _I64, _I64, _I64, _P, _I64, _P],
    "ddcx_evaluatornd_integrate": [_P, C.c_int, _P, _I64, _P, _I64, _P, _I64, _P],
    "ddcx_evaluatornd_integrate_workspace": [_P, C.c_int, _I64, _P],
    "ddcx_evaluatornd_integrate_ws": [_P, C.c_int, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P],
    "ddcx_evaluator2d_eval_vjp_binned_static": [_P, _P, _P, _P, _P, _I64, _P, _P, _I64, _I64, _I64, _I64, _P],
    "ddcx_problem_create": [C.POINTER(ddcx_problem_desc), _PP],
    "ddcx_problem_destroy": [_P],
    "ddcx_problem_set_element": [_P, C.c_int, C.c_int, C.c_double],
    "ddcx_problem_get_element": [_P, C.c_int, C.c_int, C.POINTER(C.c_double)],
    "ddcx_problem_setup_solver": [_P],
    "ddcx_problem_get_info": [_P, C.POINTER(ddcx_problem_info)],
    "ddcx_problem_solve": [_P, _P, _I64, _I64, C.c_int, _P],
    "ddcx_last_error": [],
    "ddcx_version": [],
}
_RESTYPES = {"ddcx_last_error": C.c_char_p, "ddcx_version": C.c_char_p}

_lib = None


class DDCXError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"ddcx {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def lib():
    """Load libddcx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C ddc_amd/csrc` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, C.c_int)
        _lib = L
    return _lib


def check(status):
    if status != DDCX_OK:
        raise DDCXError(status, lib().ddcx_last_error().decode())


def call(name, *args):
    check(getattr(lib(), name)(*args))
