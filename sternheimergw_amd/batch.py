"""Device-resident batched solves and their adjoint (sgw_batch_solver_* and
sgw_sternheimer_adjoint of include/sgw.h) on torch tensors.

The arrays stay on the GPU: right-hand sides (nsys, ld), shifts (nsys, nshift),
solutions (nsys, nshift, ld), all complex128; torch supplies the memory and
the stream, the library does the arithmetic.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class BatchSolver:
    """A batched multishift BiCGStab(l) solver bound to one operator."""

    def __init__(self, op, nsys, nshift, ld, lmax=4, max_iter=10000, threshold=1e-8,
                 shift_test=False, relax_threshold=0.0):
        self.op = op
        self.cfg = _lib.sgw_batch_config(nsys, nshift, ld, lmax, max_iter, threshold,
                                         int(shift_test), relax_threshold)
        self.handle = _lib.lib().sgw_batch_solver_create(op.handle, ctypes.byref(self.cfg))
        if not self.handle:
            raise _lib.SgwError(-1, _lib.lib().sgw_last_error().decode())
        self.nsys, self.nshift, self.ld = nsys, nshift, ld

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            _lib.lib().sgw_batch_solver_destroy(h)
            self.handle = None

    def _stream(self):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def solve(self, d_group, d_rhs, d_sigma, h_sigma, d_x):
        """Solves in place into d_x; returns (ierr, iters) per system."""
        ierr = np.zeros(self.nsys, dtype=np.int32)
        iters = np.zeros(self.nsys, dtype=np.int32)
        h_sigma = np.ascontiguousarray(h_sigma, dtype=np.complex128)
        _lib.check(_lib.lib().sgw_batch_solver_run(
            self.handle, _p(d_group), _p(d_rhs), _p(d_sigma), ctypes.c_void_p(h_sigma.ctypes.data),
            _p(d_x), ctypes.c_void_p(ierr.ctypes.data), ctypes.c_void_p(iters.ctypes.data),
            self._stream()))
        return ierr, iters

    def solve_stream(self, h_group, d_rhs, d_sigma, h_sigma, d_x):
        """A queue of len(h_group) systems through the solver's nsys slots
        (sgw_batch_solver_run_stream); d_rhs (nq, ld), d_sigma (nq, nshift),
        d_x (nq, nshift, ld).  Returns (ierr, iters) per queued system."""
        h_group = np.ascontiguousarray(h_group, dtype=np.int32)
        nq = h_group.size
        ierr = np.zeros(max(nq, 1), dtype=np.int32)
        iters = np.zeros(max(nq, 1), dtype=np.int32)
        h_sigma = np.ascontiguousarray(h_sigma, dtype=np.complex128)
        _lib.check(_lib.lib().sgw_batch_solver_run_stream(
            self.handle, nq, ctypes.c_void_p(h_group.ctypes.data), _p(d_rhs), _p(d_sigma),
            ctypes.c_void_p(h_sigma.ctypes.data), _p(d_x), ctypes.c_void_p(ierr.ctypes.data),
            ctypes.c_void_p(iters.ctypes.data), self._stream()))
        return ierr[:nq], iters[:nq]

    def adjoint(self, d_group, d_c, d_sigma, h_sigma, d_weight, d_x, d_y, d_grad=None):
        """Backward solve at conj(sigma) into d_y and, with d_grad (a real
        tensor of the operator's box, (n3, n2, n1)), dL/dV_loc for
        L = sum w Re <c_s | x_sk>.  Returns (L, ierr)."""
        L = ctypes.c_double(0.0)
        ierr = np.zeros(self.nsys, dtype=np.int32)
        h_sigma = np.ascontiguousarray(h_sigma, dtype=np.complex128)
        _lib.check(_lib.lib().sgw_sternheimer_adjoint(
            self.handle, _p(d_group), _p(d_c), _p(d_sigma), ctypes.c_void_p(h_sigma.ctypes.data),
            _p(d_weight), _p(d_x), _p(d_y), _p(d_grad), ctypes.byref(L),
            ctypes.c_void_p(ierr.ctypes.data), self._stream()))
        return L.value, ierr
