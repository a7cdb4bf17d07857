"""Python mirror of SternheimerGW's solver modules, running on the GPU.

Names, argument meaning and error codes follow the reference's Fortran
modules:

* ``select_solver_module``  algo/linear_solver/src/select_solver.f90:32-204
* ``bicgstab_module``       algo/linear_solver/src/bicgstab.f90:33-146
* ``linear_solver_module``  algo/linear_solver/src/linear_solver.f90:29-78

The operator argument ``AA`` is either a Python callable ``AA(sigma, x) -> Ax``
(the procedure argument of the reference, run on the host with the solver's
vector work on the GPU) or a device operator from
:mod:`sternheimergw_amd.operators`.  Results return ``(xx, ierr)`` with
``xx`` shaped (SIZE(bb), SIZE(sigma)) in Fortran order, i.e. ``xx[:, k]``.
"""
from __future__ import annotations

import dataclasses
from typing import Optional, Sequence

import numpy as np

from . import _lib

# select_solver.f90:39-45
bicgstab_multi = 1
bicgstab_no_multi = 2
sgw_linear_solver = 3


@dataclasses.dataclass
class SelectSolverType:
    """TYPE(select_solver_type), select_solver.f90:48-62."""
    priority: Optional[Sequence[int]] = None
    max_iter: int = 10000
    threshold: float = 1e-4
    bicg_lmax: int = 4


@dataclasses.dataclass
class BicgstabType:
    """TYPE(bicgstab_type), bicgstab.f90:44-55."""
    lmax: int = 4
    threshold: float = 1e-4
    max_iter: int = 10000


@dataclasses.dataclass
class LinearSolverConfig:
    """TYPE(linear_solver_config), linear_solver.f90:42-56."""
    recover_subspace: int = 1
    max_iter: int = 10000
    threshold: float = 1e-4
    abs_threshold: float = 0.0


def _prep(bb, sigma):
    bb = _lib.as_c128(bb)
    sigma = _lib.as_c128(sigma)
    if bb.ndim != 1 or sigma.ndim != 1:
        raise ValueError("bb and sigma must be vectors")
    xx = np.zeros((sigma.size, bb.size), dtype=np.complex128)  # xx[k] = xx(:, k)
    return bb, sigma, xx


def _is_device_op(AA):
    from .operators import Operator
    return isinstance(AA, Operator)


def select_solver(config: SelectSolverType, AA, bb, sigma, group: int = 0):
    """select_solver(config, AA, bb, sigma, xx, ierr) — select_solver.f90:67-161."""
    L = _lib.lib()
    bb, sigma, xx = _prep(bb, sigma)
    if config.priority is None or len(config.priority) == 0:
        raise ValueError("priority of the solvers not specified")
    prio = np.ascontiguousarray(config.priority, dtype=np.int32)
    c = _lib.sgw_select_solver_config(prio.ctypes.data_as(_lib.ctypes.POINTER(_lib.c_int32)),
                                      prio.size, config.max_iter, config.threshold,
                                      config.bicg_lmax)
    if _is_device_op(AA):
        st = L.sgw_select_solver_op(_lib.ctypes.byref(c), AA.handle, group, _lib.ptr(bb),
                                    bb.size, _lib.ptr(sigma), sigma.size, _lib.ptr(xx))
    else:
        cb = _lib.make_callback(AA)
        st = L.sgw_select_solver(_lib.ctypes.byref(c), _lib.ctypes.cast(cb, _lib.c_void_p), None,
                                 _lib.ptr(bb), bb.size, _lib.ptr(sigma), sigma.size,
                                 _lib.ptr(xx))
    ierr = _lib.check(st)
    return xx.T, ierr


def bicgstab(config: BicgstabType, AA, bb, sigma, return_iters: bool = False):
    """bicgstab(config, AA, bb, sigma, xx, ierr) — bicgstab.f90:154-275."""
    L = _lib.lib()
    bb, sigma, xx = _prep(bb, sigma)
    c = _lib.sgw_bicgstab_config(config.lmax, config.threshold, config.max_iter)
    it = _lib.c_int32(0)
    if _is_device_op(AA):
        # device operator: same engine through the batched entry (one system)
        x3, ierr, iters = AA.solve_batched(bb[None, :], sigma[None, :], lmax=config.lmax,
                                          threshold=config.threshold,
                                          max_iter=config.max_iter)
        xx = x3[0][:, : bb.size]
        out = (xx.T, int(ierr[0]))
        return out + (int(iters[0]),) if return_iters else out
    cb = _lib.make_callback(AA)
    st = L.sgw_bicgstab(_lib.ctypes.byref(c), _lib.ctypes.cast(cb, _lib.c_void_p), None,
                        _lib.ptr(bb), bb.size, _lib.ptr(sigma), sigma.size, _lib.ptr(xx),
                        _lib.ctypes.byref(it))
    ierr = _lib.check(st)
    if return_iters:
        return xx.T, ierr, it.value
    return xx.T, ierr


def linear_solver(config: LinearSolverConfig, AA, bb, sigma):
    """linear_solver(config, AA, bb, sigma, xx, ierr) — linear_solver.f90:82-194.
    ``config.abs_threshold`` is set on return (INTENT(INOUT))."""
    L = _lib.lib()
    bb, sigma, xx = _prep(bb, sigma)
    c = _lib.sgw_linear_solver_config(config.recover_subspace, config.max_iter, config.threshold,
                                      config.abs_threshold)
    cb = _lib.make_callback(AA)
    st = L.sgw_linear_solver(_lib.ctypes.byref(c), _lib.ctypes.cast(cb, _lib.c_void_p), None,
                             _lib.ptr(bb), bb.size, _lib.ptr(sigma), sigma.size, _lib.ptr(xx))
    ierr = _lib.check(st)
    config.abs_threshold = c.abs_threshold
    return xx.T, ierr
