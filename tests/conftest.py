import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


def golden(name):
    """A golden-vector case written by tests/golden/make_golden.py from the
    reference's own compiled Fortran."""
    path = os.path.join(GOLDEN, f"{name}.npz")
    return dict(np.load(path, allow_pickle=False))


def problem_from_golden(g):
    from bem3d_elastostatic_amd.mesh import BemProblem
    return BemProblem("golden", g["x"], g["y"], g["z"], g["nconec"].astype(np.int32),
                      g["kconec"].astype(np.int32), g["kode"].astype(np.int32), g["bc"],
                      float(g["mu"]), float(g["nu"]), g["points"])


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import Oracle
    return Oracle()


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def dfi_mask(p):
    """(ne, 27) mask of the element-node dofs that exist (3*nconec per element)."""
    return np.arange(27)[None, :] < 3 * np.asarray(p.nconec)[:, None]


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
