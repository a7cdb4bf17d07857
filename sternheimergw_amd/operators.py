"""Device-resident operators (A + sigma) for the GPU solvers.

``PwOperator`` is the plane-wave shifted Hamiltonian of SternheimerGW's
``linear_op`` (algo/linear_solver/src/linear_op.f90:46-206) composed with the
kinetic + local + nonlocal ``h_psi`` of Quantum ESPRESSO, one group per k+q
point.  ``DenseOperator`` is the reference's unit-test operator
``linear_problem_apply`` (algo/linear_solver/test/linear_solver.pf:100-113).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class Operator:
    handle = None
    ld = 0
    ngroup = 1

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                _lib.lib().sgw_operator_destroy(h)
            except Exception:
                pass
            self.handle = None

    def apply(self, x, sigma, group=None):
        """ax[v] = (A_{group[v]} + sigma[v]) x[v] on the GPU (host arrays)."""
        x = _lib.as_c128(np.atleast_2d(x))
        nvec = x.shape[0]
        if x.shape[1] < self.ld:
            xp = np.zeros((nvec, self.ld), dtype=np.complex128)
            xp[:, : x.shape[1]] = x
            x = xp
        sigma = _lib.as_c128(np.broadcast_to(np.asarray(sigma, dtype=np.complex128), (nvec,)))
        g = None
        if group is not None:
            g = np.ascontiguousarray(np.broadcast_to(np.asarray(group, dtype=np.int32), (nvec,)))
        ax = np.zeros_like(x)
        _lib.check(_lib.lib().sgw_operator_apply(self.handle, nvec, _lib.ptr(g), _lib.ptr(sigma),
                                                 _lib.ptr(x), x.shape[1], _lib.ptr(ax)))
        return ax

    def solve_batched(self, rhs, sigma, group=None, lmax=4, threshold=1e-4, max_iter=10000,
                      shift_test=False, relax_threshold=0.0):
        """Batched multishift BiCGStab(l): rhs (nsys, n), sigma (nsys, nshift).
        shift_test=False stops a system on the reference's seed-residual test;
        True also requires every shifted residual estimate below threshold.
        Returns x (nsys, nshift, ld), ierr (nsys,), iters (nsys,)."""
        rhs = _lib.as_c128(np.atleast_2d(rhs))
        nsys = rhs.shape[0]
        if rhs.shape[1] < self.ld:
            r = np.zeros((nsys, self.ld), dtype=np.complex128)
            r[:, : rhs.shape[1]] = rhs
            rhs = r
        ld = rhs.shape[1]
        sigma = _lib.as_c128(np.atleast_2d(sigma))
        nshift = sigma.shape[1]
        g = None if group is None else np.ascontiguousarray(group, dtype=np.int32)
        x = np.zeros((nsys, nshift, ld), dtype=np.complex128)
        ierr = np.zeros(nsys, dtype=np.int32)
        iters = np.zeros(nsys, dtype=np.int32)
        cfg = _lib.sgw_batch_config(nsys, nshift, ld, lmax, max_iter, threshold, int(shift_test),
                                    relax_threshold)
        _lib.check(_lib.lib().sgw_sternheimer_solve_batched(
            self.handle, ctypes.byref(cfg), _lib.ptr(g), _lib.ptr(rhs), _lib.ptr(sigma),
            _lib.ptr(x), _lib.ptr(ierr), _lib.ptr(iters)))
        return x, ierr, iters


class DenseOperator(Operator):
    def __init__(self, a):
        a = np.asfortranarray(a, dtype=np.complex128)
        n = a.shape[0]
        self._a = a
        self.handle = _lib.lib().sgw_operator_dense(ctypes.c_void_p(a.ctypes.data), n)
        if not self.handle:
            raise _lib.SgwError(-1, _lib.lib().sgw_last_error().decode())
        self.ld = n
        self.n = n


class CallbackOperator(Operator):
    def __init__(self, AA, n):
        self._cb = _lib.make_callback(AA)
        self.handle = _lib.lib().sgw_operator_callback(ctypes.cast(self._cb, ctypes.c_void_p),
                                                       None, n)
        if not self.handle:
            raise _lib.SgwError(-1, _lib.lib().sgw_last_error().decode())
        self.ld = n
        self.n = n


class PwOperator(Operator):
    """Plane-wave operator over the k+q groups of a :class:`cell.Workload`."""

    def __init__(self, *, box, vloc, npw, ld, mill, kin, proj, dcoef):
        n1, n2, n3 = box
        self._keep = dict(
            vloc=np.ascontiguousarray(vloc, dtype=np.float64),
            npw=np.ascontiguousarray(npw, dtype=np.int32),
            mill=np.ascontiguousarray(mill, dtype=np.int32),
            kin=np.ascontiguousarray(kin, dtype=np.float64),
            proj=np.ascontiguousarray(proj, dtype=np.complex128),
            dcoef=np.ascontiguousarray(dcoef, dtype=np.float64),
        )
        k = self._keep
        ngroup = k["npw"].size
        nproj = k["proj"].shape[1] if k["proj"].ndim == 3 else 0
        desc = _lib.sgw_pw_desc(n1, n2, n3, k["vloc"].ctypes.data, ngroup, k["npw"].ctypes.data,
                                ld, k["mill"].ctypes.data, k["kin"].ctypes.data, nproj,
                                k["proj"].ctypes.data if nproj else None,
                                k["dcoef"].ctypes.data if nproj else None)
        self.handle = _lib.lib().sgw_operator_pw(ctypes.byref(desc))
        if not self.handle:
            raise _lib.SgwError(-1, _lib.lib().sgw_last_error().decode())
        self.ld = ld
        self.ngroup = ngroup
        self.npw = k["npw"]

    @classmethod
    def from_workload(cls, w, with_valence: bool = True):
        """Coulomb operator (alpha_pv P_v included) or, with_valence=False, the
        Green's-function operator (alpha_pv = 0, green.f90:282-283)."""
        proj, dcoef = w.proj, w.dcoef
        if not with_valence:
            proj, dcoef = proj[:, : w.nproj_beta], dcoef[:, : w.nproj_beta]
        return cls(box=w.box, vloc=w.vloc, npw=w.npw, ld=w.ld, mill=w.mill_padded(),
                   kin=w.kin_padded(), proj=proj, dcoef=dcoef)
