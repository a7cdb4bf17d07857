"""ctypes binding of libsgw_hip.so (the C ABI declared in include/sgw.h).

The library is built in-tree (``sternheimergw_amd/libsgw_hip.so``).  Loading
fails loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgw_hip.so")

c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p

SGW_E = {-1: "SGW_E_ARG", -2: "SGW_E_SHAPE", -3: "SGW_E_HIP", -4: "SGW_E_ALLOC", -5: "SGW_E_STATE"}


class SgwError(RuntimeError):
    """A negative status from the C ABI (argument, shape or HIP runtime error)."""

    def __init__(self, code, msg):
        super().__init__(f"{SGW_E.get(code, code)}: {msg}")
        self.code = code


class sgw_complex(ctypes.Structure):
    _fields_ = [("re", c_double), ("im", c_double)]


OP_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(sgw_complex), ctypes.POINTER(sgw_complex),
                         ctypes.POINTER(sgw_complex), c_int64, c_void_p)


class sgw_select_solver_config(ctypes.Structure):
    _fields_ = [("priority", ctypes.POINTER(c_int32)), ("npriority", c_int32),
                ("max_iter", c_int32), ("threshold", c_double), ("bicg_lmax", c_int32)]


class sgw_bicgstab_config(ctypes.Structure):
    _fields_ = [("lmax", c_int32), ("threshold", c_double), ("max_iter", c_int32)]


class sgw_linear_solver_config(ctypes.Structure):
    _fields_ = [("recover_subspace", c_int32), ("max_iter", c_int32), ("threshold", c_double),
                ("abs_threshold", c_double)]


class sgw_pw_desc(ctypes.Structure):
    _fields_ = [("n1", c_int32), ("n2", c_int32), ("n3", c_int32), ("vloc", c_void_p),
                ("ngroup", c_int32), ("npw", c_void_p), ("ld", c_int64), ("mill", c_void_p),
                ("kin", c_void_p), ("nproj", c_int32), ("proj", c_void_p), ("dcoef", c_void_p)]


class sgw_batch_config(ctypes.Structure):
    _fields_ = [("nsys", c_int32), ("nshift", c_int32), ("ld", c_int64), ("lmax", c_int32),
                ("max_iter", c_int32), ("threshold", c_double), ("shift_test", c_int32),
                ("relax_threshold", c_double)]


class sgw_sigma_conv_desc(ctypes.Structure):
    _fields_ = [("n", c_int32), ("ld", c_int64), ("nout", c_int32), ("nin", c_int32),
                ("g_stride_i", c_int64), ("g_stride_j", c_int64), ("w_stride", c_int64),
                ("s_stride", c_int64), ("precision", c_int32), ("accumulate", c_int32)]


SIGMA_FP64 = 0
SIGMA_FP32_SPLIT = 1
GODBY_NEEDS = 1
PADE_APPROX = 2
PADE_ROBUST = 3
AAA_APPROX = 4
AAA_POLE = 5

# (name, restype, argtypes) of every exported entry point of include/sgw.h
SIGNATURES = [
    ("sgw_select_solver", c_int32, [ctypes.POINTER(sgw_select_solver_config), c_void_p, c_void_p,
                                    c_void_p, c_int64, c_void_p, c_int32, c_void_p]),
    ("sgw_bicgstab", c_int32, [ctypes.POINTER(sgw_bicgstab_config), c_void_p, c_void_p, c_void_p,
                               c_int64, c_void_p, c_int32, c_void_p, c_void_p]),
    ("sgw_linear_solver", c_int32, [ctypes.POINTER(sgw_linear_solver_config), c_void_p, c_void_p,
                                    c_void_p, c_int64, c_void_p, c_int32, c_void_p]),
    ("sgw_operator_dense", c_void_p, [c_void_p, c_int64]),
    ("sgw_operator_callback", c_void_p, [c_void_p, c_void_p, c_int64]),
    ("sgw_operator_pw", c_void_p, [ctypes.POINTER(sgw_pw_desc)]),
    ("sgw_operator_destroy", None, [c_void_p]),
    ("sgw_operator_set_fft_precision", c_int32, [c_void_p, c_int32]),
    ("sgw_operator_apply", c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_int64,
                                     c_void_p]),
    ("sgw_operator_apply_device", c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                                            c_int64, c_void_p, c_void_p]),
    ("sgw_operator_apply_device_noalloc", c_int32, [c_void_p, c_int32, c_void_p, c_void_p,
                                                    c_void_p, c_int64, c_void_p, c_void_p]),
    ("sgw_select_solver_op", c_int32, [ctypes.POINTER(sgw_select_solver_config), c_void_p, c_int32,
                                       c_void_p, c_int64, c_void_p, c_int32, c_void_p]),
    ("sgw_batch_solver_create", c_void_p, [c_void_p, ctypes.POINTER(sgw_batch_config)]),
    ("sgw_batch_solver_destroy", None, [c_void_p]),
    ("sgw_batch_solver_run", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p]),
    ("sgw_batch_solver_run_stream", c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("sgw_sternheimer_solve_batched", c_int32, [c_void_p, ctypes.POINTER(sgw_batch_config),
                                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                c_void_p]),
    ("sgw_residual_norms_device", c_int32, [c_void_p, c_int32, c_int32, c_int64, c_void_p,
                                            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("sgw_sternheimer_adjoint", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p]),
    ("sgw_drho_accumulate", c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_int64,
                                      c_void_p, c_double, c_void_p, c_int32, c_void_p]),
    ("sgw_box_fft", c_int32, [c_int32, c_int32, c_int32, c_void_p, c_void_p]),
    ("sgw_wave_to_box", c_int32, [c_int32, c_int32, c_int32, c_void_p, c_void_p, c_int64,
                                  c_void_p, c_void_p]),
    ("sgw_box_to_wave", c_int32, [c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p,
                                  c_int64, c_int32, c_void_p]),
    ("sgw_dvscf_psi", c_int32, [c_int32, c_int32, c_int32, c_void_p, c_void_p, c_int32,
                                c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    ("sgw_orthogonalize", c_int32, [c_int32, c_int32, c_int32, c_void_p, c_void_p, c_int64,
                                    c_void_p]),
    ("sgw_dv_of_drho", c_int32, [c_int32, c_int32, c_void_p, c_void_p, c_double, c_int32,
                                 c_void_p, c_void_p]),
    ("sgw_average_pm", c_int32, [c_int32, c_int32, c_int32, c_int32, c_void_p, c_int64, c_double,
                                 c_void_p, c_void_p]),
    ("sgw_invfft6", c_int32, [c_int32, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_double,
                              c_int32, c_void_p, c_void_p]),
    ("sgw_fwfft6", c_int32, [c_int32, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_double,
                             c_int32, c_void_p, c_void_p]),
    ("sgw_green_assemble", c_int32, [c_int32, c_int32, c_int32, c_int32, c_void_p, c_int32,
                                     c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_void_p]),
    ("sgw_analytic_coeff", c_int32, [c_int32, c_double, c_int32, c_int32, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    ("sgw_analytic_eval", c_int32, [c_int32, c_int32, c_int32, c_void_p, c_void_p, sgw_complex,
                                    c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p,
                                    c_void_p, c_void_p]),
    ("sgw_sigma_prod", c_int32, [c_int32, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_double,
                                 sgw_complex, c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    ("sgw_broyden_create", c_void_p, [c_int64, c_int32]),
    ("sgw_broyden_destroy", None, [c_void_p]),
    ("sgw_mix_potential_c", c_int32, [c_void_p, c_int64, c_void_p, c_void_p, c_double,
                                      ctypes.POINTER(c_double), c_double, c_int32, c_int32,
                                      ctypes.POINTER(c_int32), c_void_p]),
    ("sgw_sigma_conv_device", c_int32, [ctypes.POINTER(sgw_sigma_conv_desc), c_void_p, c_void_p,
                                        c_void_p, sgw_complex, c_void_p, c_void_p]),
    ("sgw_sigma_synthetic_device", c_int32, [c_int32, c_int64, c_int32, c_int32, c_int32, c_int32,
                                             ctypes.c_uint64, c_void_p, c_void_p, c_int64, c_int64,
                                             c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                                             c_int64, c_void_p]),
    ("sgw_kernel_timing", c_int32, [c_int32]),
    ("sgw_kernel_timing_query", c_int32, [ctypes.c_char_p, ctypes.POINTER(c_double),
                                          ctypes.POINTER(c_int64)]),
    ("sgw_autotune_choice", c_int32, [ctypes.c_char_p]),
    ("sgw_sternheimer_solve_refined", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                c_void_p, c_double, c_int32, c_int32, c_void_p,
                                                c_void_p, c_void_p, c_void_p]),
    ("sgw_last_error", ctypes.c_char_p, []),
    ("sgw_version", ctypes.c_char_p, []),
]

_lib = None


def lib():
    """Load libsgw_hip.so (once).  Raises if the library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m sternheimergw_amd.build` "
                "(there is no CPU fallback)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> int:
    """Raise SgwError on a negative status, else return it (the solver ierr)."""
    if status < 0:
        raise SgwError(status, lib().sgw_last_error().decode())
    return status


def ptr(a: np.ndarray):
    return c_void_p(a.ctypes.data) if a is not None else None


def as_c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def make_callback(AA):
    """Wrap a Python operator AA(sigma, x) -> Ax as the C callback type."""

    def _cb(sig_p, x_p, ax_p, n, _ctx):
        n = int(n)
        sig = complex(sig_p[0].re, sig_p[0].im)
        x = np.ctypeslib.as_array(ctypes.cast(x_p, ctypes.POINTER(c_double)), shape=(2 * n,))
        x = x.view(np.complex128).copy()
        ax = np.ctypeslib.as_array(ctypes.cast(ax_p, ctypes.POINTER(c_double)), shape=(2 * n,))
        ax.view(np.complex128)[:] = np.asarray(AA(sig, x), dtype=np.complex128)

    return OP_FN(_cb)
