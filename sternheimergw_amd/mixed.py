"""Mixed-precision batched solve with FP64 iterative refinement
(sgw_sternheimer_solve_refined of include/sgw.h) on torch device tensors.

FP32 local-potential FFTs in a multishift solve to an inner threshold, then
FP64 residuals of every shift and FP32-FFT single-shift corrections until
every shift meets the solver's threshold.  The reference's solver is FP64
throughout (algo/linear_solver/src/bicgstab.f90:154-275); this is an
extension of its path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .batch import BatchSolver, _p


def solve_refined(solver: BatchSolver, d_group, d_rhs, d_sigma, h_sigma, d_x,
                  inner_threshold: float = 1e-5, max_refine: int = 3,
                  correction_slots: int = 1024):
    """Solves into d_x (nsys, nshift, ld).  Returns (ierr per system, stats,
    relative residuals (nsys, nshift)); stats: FP32 H-applies of the
    multishift stage and of the corrections, FP64 residual applications,
    refinement rounds, multishift iterations, correction iterations."""
    import torch
    ierr = np.zeros(solver.nsys, dtype=np.int32)
    stats = np.zeros(6, dtype=np.int64)
    res = np.zeros((solver.nsys, solver.nshift), dtype=np.float64)
    h_sigma = np.ascontiguousarray(h_sigma, dtype=np.complex128)
    rc = _lib.lib().sgw_sternheimer_solve_refined(
        solver.handle, _p(d_group), _p(d_rhs), _p(d_sigma), ctypes.c_void_p(h_sigma.ctypes.data),
        _p(d_x), inner_threshold, max_refine, correction_slots,
        ctypes.c_void_p(ierr.ctypes.data), ctypes.c_void_p(stats.ctypes.data),
        ctypes.c_void_p(res.ctypes.data),
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc)
    torch.cuda.synchronize()
    keys = ("fp32_happly_multishift", "fp32_happly_corrections", "fp64_residual_happly",
            "refinements", "multishift_iterations", "correction_iterations")
    return ierr, dict(zip(keys, (int(v) for v in stats))), res
