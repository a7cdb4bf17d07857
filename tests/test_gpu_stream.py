"""Streaming batched solve (sgw_batch_solver_run_stream): a queue of systems
through fewer slots gives the same solutions and iteration counts as the
lock-step batched solve of all of them."""
import numpy as np
import pytest
import torch

from sternheimergw_amd import cell, operators
from sternheimergw_amd.batch import BatchSolver

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _both(w, op, order, slots, threshold, lmax=4, shift_test=True, max_iter=10000):
    rhs, sigma, group = w.rhs[order], np.ascontiguousarray(w.sigma[order]), w.group[order]
    nq, ns, ld = len(order), w.nshift, w.ld
    ref = BatchSolver(op, nq, ns, ld, lmax=lmax, threshold=threshold, shift_test=shift_test,
                      max_iter=max_iter)
    x_ref = torch.zeros((nq, ns, ld), dtype=torch.complex128, device=DEV)
    e_ref, it_ref = ref.solve(_t(group), _t(rhs), _t(sigma), sigma, x_ref)
    st = BatchSolver(op, slots, ns, ld, lmax=lmax, threshold=threshold, shift_test=shift_test,
                     max_iter=max_iter)
    x = torch.zeros((nq, ns, ld), dtype=torch.complex128, device=DEV)
    e, it = st.solve_stream(group, _t(rhs), _t(sigma), sigma, x)
    return (x_ref.cpu().numpy(), e_ref, it_ref), (x.cpu().numpy(), e, it)


def test_stream_equals_lockstep_48cubed():
    w = cell.workload_c2(nq=2, nbnd=16)
    op = operators.PwOperator.from_workload(w)
    # uneven groups: 20 systems of q 0, 12 of q 1, through 8 slots
    order = np.concatenate([np.arange(16), np.arange(16, 28), np.arange(0, 4)])
    print("stream vs lock-step: 32 systems through 8 slots, 48^3", flush=True)
    (xr, er, itr), (x, e, it) = _both(w, op, order, 8, 1e-8)
    assert np.all(er == 0) and np.all(e == 0)
    assert np.array_equal(it, itr)
    assert np.abs(x - xr).max() <= 1e-13 * np.abs(xr).max()


def test_stream_more_slots_than_systems(c1):
    w = c1
    op = operators.PwOperator.from_workload(w)
    order = np.arange(w.nsys)
    (xr, er, itr), (x, e, it) = _both(w, op, order, 16, 1e-10)
    assert np.array_equal(it, itr) and np.all(e == 0)
    assert np.abs(x - xr).max() <= 1e-13 * np.abs(xr).max()


def test_stream_max_iter_reports_ierr(c1):
    w = c1
    op = operators.PwOperator.from_workload(w)
    order = np.arange(w.nsys)
    (xr, er, itr), (x, e, it) = _both(w, op, order, 2, 1e-14, max_iter=3)
    assert np.all(er == 1) and np.all(e == 1)
    assert np.all(it == 4)


@pytest.mark.parametrize("chunk", ["1", "3"])
def test_graph_replay_equals_eager(c1, monkeypatch, chunk):
    """Iterations replayed from a captured hipGraph (SGW_SOLVER_GRAPH=1,
    `chunk` iterations per launch) give the eager launches' solutions and
    iteration counts bit for bit, lock-step and streamed (the captured
    sequence does not depend on the iteration number)."""
    w = c1
    op = operators.PwOperator.from_workload(w)
    order = np.arange(w.nsys)
    eager = _both(w, op, order, 2, 1e-10)
    monkeypatch.setenv("SGW_SOLVER_GRAPH", "1")
    monkeypatch.setenv("SGW_GRAPH_CHUNK", chunk)
    graph = _both(w, op, order, 2, 1e-10)
    for (xe, ee, ie), (xg, eg, ig) in zip(eager, graph):
        assert np.all(ee == 0) and np.array_equal(ee, eg)
        assert np.array_equal(ie, ig)
        assert np.array_equal(xe, xg)


@pytest.mark.parametrize("graph", ["0", "1"])
def test_lockstep_compaction_equals_full_slots(monkeypatch, graph):
    """Lock-step solves move their running systems into the leading slots as
    others finish (fewer blocks per launch late in the solve).  The moved
    systems' arithmetic does not change: solutions, iteration counts and
    ierr equal those of a solve that keeps every slot (SGW_COMPACT=0), bit
    for bit, on uneven groups of the 48^3 cell; with graph = 1 the compacted
    solve replays captured iterations (recaptured after each compaction)."""
    w = cell.workload_c2(nq=2, nbnd=16)
    op = operators.PwOperator.from_workload(w)
    order = np.concatenate([np.arange(16), np.arange(16, 28), np.arange(0, 4)])
    rhs, sigma, group = w.rhs[order], np.ascontiguousarray(w.sigma[order]), w.group[order]
    nq, ns, ld = len(order), w.nshift, w.ld
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("SGW_COMPACT", mode)
        monkeypatch.setenv("SGW_SOLVER_GRAPH", graph if mode == "1" else "0")
        print(f"lock-step 48^3, 32 systems, SGW_COMPACT={mode}", flush=True)
        sol = BatchSolver(op, nq, ns, ld, lmax=4, threshold=1e-8, shift_test=True)
        x = torch.zeros((nq, ns, ld), dtype=torch.complex128, device=DEV)
        e, it = sol.solve(_t(group), _t(rhs), _t(sigma), sigma, x)
        out[mode] = (x.cpu().numpy(), e, it)
    (xc, ec, itc), (xf, ef, itf) = out["1"], out["0"]
    assert np.all(ec == 0) and np.array_equal(ec, ef)
    assert np.array_equal(itc, itf)
    assert np.array_equal(xc, xf)
