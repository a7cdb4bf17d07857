import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def lin_prob():
    """The reference's solver fixture (algo/linear_solver/test/lin_prob.xml.bz2)."""
    d = np.load(os.path.join(GOLDEN, "lin_prob.npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


@pytest.fixture(scope="session")
def c1():
    from sternheimergw_amd import cell
    return cell.workload_c1()
