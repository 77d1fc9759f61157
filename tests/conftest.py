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


def hip_runtime():
    """The HIP runtime this process already uses (the library and torch are
    linked to it): test-side device copies go through the same one."""
    import ctypes as C
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                lib = C.CDLL(line.split()[-1])
                lib.hipMemcpy.restype = C.c_int
                lib.hipMemcpy2D.restype = C.c_int
                lib.hipMemcpy2D.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t,
                                            C.c_size_t, C.c_int]
                return lib
    raise RuntimeError("HIP runtime not loaded")


def device_rows(q, nodes):
    """Rows 3i, 3i+1, 3i+2 of the handle's device-resident A for each node i
    (0-based), gathered column by column from the column-major A: (3k, N)."""
    import ctypes as C
    hip = hip_runtime()
    pa, lda, _ = q.system_device()
    N = q.N
    out = np.zeros((len(nodes), N, 3))
    for j, i in enumerate(nodes):
        rc = hip.hipMemcpy2D(out[j].ctypes.data, 24, pa + 24 * int(i), 8 * lda, 24, N, 2)
        assert rc == 0, rc
    return out.transpose(0, 2, 1).reshape(3 * len(nodes), N)
