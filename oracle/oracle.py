"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
CPU-baseline leg of bench.py — never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

V = ctypes.c_void_p
I32 = ctypes.c_int
I64 = ctypes.c_int64
D = ctypes.c_double

_lib = None


class DenseCtx(ctypes.Structure):
    _fields_ = [("a", V), ("n", I64)]


class PwCtx(ctypes.Structure):
    _fields_ = [("n1", I32), ("n2", I32), ("n3", I32), ("vloc", V), ("npw", I64), ("nl", V),
                ("kin", V), ("nproj", I32), ("beta", V), ("dcoef", V), ("nocc", I32),
                ("evq", V), ("alpha_pv", D)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_norm.restype = D
        L.oracle_norm.argtypes = [V, I64]
        L.oracle_gram_schmidt.argtypes = [I32, V, I64, I32, V]
        L.oracle_bicgstab.restype = I32
        L.oracle_bicgstab.argtypes = [I32, D, I32, V, V, V, I64, V, I32, V, V]
        L.oracle_bicgstab_ex.restype = I32
        L.oracle_bicgstab_ex.argtypes = [I32, D, I32, V, V, V, I64, V, I32, V, V, I32]
        L.oracle_linear_solver.restype = I32
        L.oracle_linear_solver.argtypes = [I32, D, V, V, V, I64, V, I32, V, V]
        L.oracle_select_solver.restype = I32
        L.oracle_select_solver.argtypes = [V, I32, I32, D, I32, V, V, V, I64, V, I32, V]
        L.oracle_fft3d.argtypes = [V, I32, I32, I32, I32]
        L.oracle_op_calls.restype = I64
        _lib = L
    return _lib


def _p(a):
    return V(a.ctypes.data)


class DenseOp:
    """(A + sigma) with A dense column-major (linear_solver.pf:100-113)."""

    def __init__(self, a):
        self.a = np.asfortranarray(a, dtype=np.complex128)
        self.n = self.a.shape[0]
        self.ctx = DenseCtx(self.a.ctypes.data, self.n)
        self.fn = ctypes.cast(lib().oracle_dense_op, V)

    @property
    def ctx_ptr(self):
        return ctypes.cast(ctypes.byref(self.ctx), V)


class PwOp:
    """Plane-wave operator of one k+q group of a Workload (Ry)."""

    def __init__(self, w, iq: int, with_valence: bool = True):
        s = w.spheres[iq]
        self.n = s.npw
        n1, n2, n3 = w.box
        self.vloc = np.ascontiguousarray(w.vloc, dtype=np.float64)
        self.nl = s.box_index(w.box)
        self.kin = np.ascontiguousarray(s.kin)
        nb = w.nproj_beta
        self.beta = np.ascontiguousarray(w.proj[iq, :nb, : s.npw])
        self.dcoef = np.ascontiguousarray(w.dcoef[iq, :nb])
        self.evq = np.ascontiguousarray(w.evq[iq])
        nocc = self.evq.shape[0] if with_valence else 0
        self.ctx = PwCtx(n1, n2, n3, self.vloc.ctypes.data, self.n, self.nl.ctypes.data,
                         self.kin.ctypes.data, nb, self.beta.ctypes.data, self.dcoef.ctypes.data,
                         nocc, self.evq.ctypes.data, w.alpha_pv if with_valence else 0.0)
        self.fn = ctypes.cast(lib().oracle_pw_op, V)

    @property
    def ctx_ptr(self):
        return ctypes.cast(ctypes.byref(self.ctx), V)

    def apply(self, x, sigma):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        ax = np.zeros_like(x)
        sg = np.array([sigma], dtype=np.complex128)
        lib().oracle_pw_op(_p(sg), _p(x), _p(ax), I64(self.n), self.ctx_ptr)
        return ax


def bicgstab(op, b, sigma, lmax=4, threshold=1e-4, max_iter=10000, shift_test=False):
    """Returns x (nshift, n), ierr, iterations.  shift_test=True also waits
    for every shifted residual estimate (an extension of the reference's
    test); a shifted system whose estimate is below half the threshold at the
    end of an iteration keeps its solution from then on."""
    b = np.ascontiguousarray(b, dtype=np.complex128)
    sigma = np.ascontiguousarray(np.atleast_1d(sigma), dtype=np.complex128)
    x = np.zeros((sigma.size, b.size), dtype=np.complex128)
    it = ctypes.c_int(0)
    ierr = lib().oracle_bicgstab_ex(lmax, threshold, max_iter, op.fn, op.ctx_ptr, _p(b), b.size,
                                    _p(sigma), sigma.size, _p(x), ctypes.byref(it),
                                    int(shift_test))
    return x, ierr, it.value


def linear_solver(op, b, sigma, threshold=1e-4, max_iter=10000):
    b = np.ascontiguousarray(b, dtype=np.complex128)
    sigma = np.ascontiguousarray(np.atleast_1d(sigma), dtype=np.complex128)
    x = np.zeros((sigma.size, b.size), dtype=np.complex128)
    nb = ctypes.c_int(0)
    ierr = lib().oracle_linear_solver(max_iter, threshold, op.fn, op.ctx_ptr, _p(b), b.size,
                                      _p(sigma), sigma.size, _p(x), ctypes.byref(nb))
    return x, ierr, nb.value


def select_solver(op, b, sigma, priority, threshold=1e-4, max_iter=10000, bicg_lmax=4):
    b = np.ascontiguousarray(b, dtype=np.complex128)
    sigma = np.ascontiguousarray(np.atleast_1d(sigma), dtype=np.complex128)
    pr = np.ascontiguousarray(priority, dtype=np.int32)
    x = np.zeros((sigma.size, b.size), dtype=np.complex128)
    ierr = lib().oracle_select_solver(_p(pr), pr.size, max_iter, threshold, bicg_lmax, op.fn,
                                      op.ctx_ptr, _p(b), b.size, _p(sigma), sigma.size, _p(x))
    return x, ierr


def norm(v):
    v = np.ascontiguousarray(v, dtype=np.complex128)
    return lib().oracle_norm(_p(v), v.size)


def gram_schmidt(first, basis, vector=None):
    """basis, vector: (nbasis, n) row = column of the reference; modified in place."""
    nb, n = basis.shape
    lib().oracle_gram_schmidt(first, _p(basis), n, nb, _p(vector) if vector is not None else None)


def fft3d(box_arr, sign):
    """box_arr: complex (n3, n2, n1) array modified in place."""
    n3, n2, n1 = box_arr.shape
    lib().oracle_fft3d(_p(box_arr), n1, n2, n3, sign)
    return box_arr


def op_calls():
    return lib().oracle_op_calls()


def reset_counters():
    lib().oracle_reset_counters()
