"""The screened Coulomb interaction file of the reference: a Fortran
direct-access, unformatted file with one record per q point
(phys/coul/src/do_stern.f90:236 writes it with davcio; sigma.f90:303-331
opens it with RECL = 8 * lrcoul bytes and reads record iq).

Record iq holds scrcoul_g(num_g_corr, num_g_corr, nfs) as complex128 in
Fortran order (first index fastest), lrcoul = 2 num_g_corr^2 nfs reals
(algo/io/src/opengwfil.f90:55), with no record markers.  The device layout
of the library's matrices ([f][c][r]) is the same byte order, so a device
tensor is written as it is.
"""
from __future__ import annotations

import os

import numpy as np


def record_bytes(num_g_corr: int, nfs: int) -> int:
    """RECL in bytes: 8 * lrcoul, lrcoul = 2 num_g_corr^2 nfs."""
    return 8 * 2 * num_g_corr * num_g_corr * nfs


def _as_records(w, num_g_corr, nfs):
    """Accepts (nfs, num_g, num_g) in the device layout [f][c][r] (a numpy
    array or a torch tensor); returns the record's bytes."""
    try:
        import torch
        if isinstance(w, torch.Tensor):
            w = w.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    w = np.ascontiguousarray(w, dtype=np.complex128)
    if w.shape != (nfs, num_g_corr, num_g_corr):
        raise ValueError(f"W must be (nfs, num_g_corr, num_g_corr) = {(nfs, num_g_corr, num_g_corr)}")
    return w.tobytes()


def write_coulomb(path: str, iq: int, w, num_g_corr: int, nfs: int) -> None:
    """davcio(scrcoul_g, lrcoul, iuncoul, iq, +1): write record iq (1-based)
    of the file, creating it if needed; other records are kept."""
    if iq < 1:
        raise ValueError("records are numbered from 1")
    data = _as_records(w, num_g_corr, nfs)
    rec = record_bytes(num_g_corr, nfs)
    mode = "r+b" if os.path.exists(path) else "w+b"
    with open(path, mode) as f:
        f.seek((iq - 1) * rec)
        f.write(data)


def read_coulomb(path: str, iq: int, num_g_corr: int, nfs: int) -> np.ndarray:
    """davcio(..., iq, -1): record iq as (nfs, num_g_corr, num_g_corr) in the
    device layout [f][c][r] (Fortran scrcoul_g(r, c, f)).  Raises if the file
    has no such record, as davcio's errore does."""
    rec = record_bytes(num_g_corr, nfs)
    size = os.path.getsize(path)
    if iq < 1 or iq * rec > size:
        raise IOError(f"record {iq} is not in {path} ({size} bytes, {rec} per record)")
    with open(path, "rb") as f:
        f.seek((iq - 1) * rec)
        buf = f.read(rec)
    return np.frombuffer(buf, dtype=np.complex128).reshape(nfs, num_g_corr, num_g_corr).copy()
