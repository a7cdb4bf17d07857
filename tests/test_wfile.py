"""The direct-access W file (sternheimergw_amd/wfile.py) against a file that
a Fortran compiler writes with the reference's OPEN/WRITE statements
(direct access, unformatted, RECL = 8 * lrcoul bytes, record iq)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from sternheimergw_amd import wfile

FORTRAN = """
program w
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  integer :: n, nfs, lrcoul, ig, igp, iw, iq
  complex(dp), allocatable :: a(:,:,:)
  n = 3; nfs = 2; lrcoul = 2*n*n*nfs
  allocate(a(n,n,nfs))
  open(28, file='w.dat', access='direct', form='unformatted', recl=8*lrcoul, status='replace')
  do iq = 1, 2
    do iw = 1, nfs
      do igp = 1, n
        do ig = 1, n
          a(ig,igp,iw) = cmplx(ig + 10*igp + 100*iw + 1000*iq, -ig, kind=dp)
        end do
      end do
    end do
    write(28, rec=iq) a
  end do
  close(28)
end program
"""


def expected(iq, n=3, nfs=2):
    w = np.zeros((nfs, n, n), dtype=np.complex128)       # [f][c][r]
    for iw in range(nfs):
        for igp in range(n):
            for ig in range(n):
                w[iw, igp, ig] = complex(ig + 1 + 10 * (igp + 1) + 100 * (iw + 1) + 1000 * iq,
                                         -(ig + 1))
    return w


def test_round_trip_and_layout(tmp_path):
    p = str(tmp_path / "w.dat")
    wfile.write_coulomb(p, 2, expected(2), 3, 2)
    wfile.write_coulomb(p, 1, expected(1), 3, 2)
    assert os.path.getsize(p) == 2 * wfile.record_bytes(3, 2) == 576
    for iq in (1, 2):
        assert np.array_equal(wfile.read_coulomb(p, iq, 3, 2), expected(iq))
    raw = np.fromfile(p, dtype=np.complex128)
    assert raw[1] == complex(1112, -2) and raw[3] == complex(1121, -1)  # Fortran order
    with pytest.raises(IOError):
        wfile.read_coulomb(p, 3, 3, 2)


@pytest.mark.skipif(shutil.which("amdflang") is None and not os.path.exists("/opt/rocm/bin/amdflang"),
                    reason="no Fortran compiler")
def test_matches_fortran_direct_access_file(tmp_path):
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    src = tmp_path / "w.f90"
    src.write_text(FORTRAN)
    subprocess.run([fc, str(src), "-o", str(tmp_path / "w")], check=True, capture_output=True,
                   timeout=300)
    subprocess.run([str(tmp_path / "w")], check=True, cwd=tmp_path, timeout=60)
    ref = (tmp_path / "w.dat").read_bytes()
    mine = str(tmp_path / "mine.dat")
    for iq in (1, 2):
        wfile.write_coulomb(mine, iq, expected(iq), 3, 2)
        assert np.array_equal(wfile.read_coulomb(str(tmp_path / "w.dat"), iq, 3, 2), expected(iq))
    assert open(mine, "rb").read() == ref
