"""sgw_operator_apply_device_noalloc (no counterpart in the reference): the
plane-wave application with its work buffers in the library's fixed
per-device scratch array gives the same results as sgw_operator_apply_device
(same kernels, same inputs: bitwise equal), on two calls in a row, the
second with a smaller input, so that leftovers of the first call in the
shared array would show.  A batch beyond the array is refused before any
work is queued."""
import ctypes

import numpy as np
import pytest
import torch

from sternheimergw_amd import _lib, cell, operators

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
SGW_E_SHAPE = -2


def _inputs(w, nvec, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    grp = (np.arange(nvec) % len(w.spheres)).astype(np.int32)
    x = np.zeros((nvec, w.ld), dtype=np.complex128)
    for v in range(nvec):
        n = w.npw[grp[v]]
        x[v, :n] = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    sig = rng.standard_normal(nvec) + 1j * rng.standard_normal(nvec)
    return (torch.from_numpy(grp).to(DEV), torch.from_numpy(sig).to(DEV),
            torch.from_numpy(x).to(DEV))


def _apply(fn, op, w, d_g, d_s, d_x):
    d_ax = torch.full_like(d_x, complex(np.nan, np.nan))
    rc = fn(op.handle, d_x.shape[0], ctypes.c_void_p(d_g.data_ptr()),
            ctypes.c_void_p(d_s.data_ptr()), ctypes.c_void_p(d_x.data_ptr()), w.ld,
            ctypes.c_void_p(d_ax.data_ptr()),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, d_ax.cpu().numpy()


@pytest.mark.parametrize("name", ["c1-24", "c2-48"])
def test_noalloc_apply_matches_device_apply(name):
    if name == "c1-24":
        q, scale = cell.c2_qpoints(3)
        w = cell.build_workload("multi", box=(24, 24, 24), ecut_ha=6.0, nbnd=3, lchan=[0, 1],
                                qpoints=q, ecut_scale=scale,
                                omega_ha=1j * np.linspace(0, 2, 5), seed=11)
        sizes = (150, 37)
    else:
        w = cell.workload_c2(nq=2, nbnd=2)
        sizes = (96, 21)
    op = operators.PwOperator.from_workload(w)
    L = _lib.lib()
    for call, nvec in enumerate(sizes):
        print(f"no-alloc apply {name}: call {call}, {nvec} vectors", flush=True)
        d_g, d_s, d_x = _inputs(w, nvec, 100 + call)
        rc0, ref = _apply(L.sgw_operator_apply_device, op, w, d_g, d_s, d_x)
        rc1, got = _apply(L.sgw_operator_apply_device_noalloc, op, w, d_g, d_s, d_x)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1)
        assert np.array_equal(got, ref), np.abs(got - ref).max()


def test_noalloc_refuses_batches_beyond_the_array():
    w = cell.workload_c2(nq=1, nbnd=2)
    op = operators.PwOperator.from_workload(w)
    d_g, d_s, d_x = _inputs(w, 400, 7)
    rc, out = _apply(_lib.lib().sgw_operator_apply_device_noalloc, op, w, d_g, d_s, d_x)
    assert rc == SGW_E_SHAPE
    assert b"scratch" in _lib.lib().sgw_last_error()
    assert np.all(np.isnan(out.real))
