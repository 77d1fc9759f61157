"""qd_solve_device_noalloc: the allocation-free solve (no counterpart in the
reference; for callers that cannot allocate device memory).

Two calls in a row on one GPU, the second with a smaller system: the
18,438-DOF cube with two right-hand sides, then the 3,078-DOF cantilever with
one.  Each call's solutions must equal those of qd_solve_device (the existing
entry point, which factors the system first) to 1e-10 relative, the first
right-hand side must equal the compiled reference's solution to 1e-10, and
the library's device memory must not change across the call.  The existing
entry point keeps working on the same handle afterwards.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import golden, problem_from_golden, rel_err

pytestmark = pytest.mark.gpu


def _solve(q, fn, precision, B):
    import torch
    from bem3d_elastostatic_amd import _lib
    b = torch.tensor(np.ascontiguousarray(B.T), dtype=torch.float64, device="cuda")  # column-major (N, k)
    rep = _lib.SolveReport()
    rc = fn(q.h, _lib.PRECISIONS[precision], B.shape[1], C.c_void_p(b.data_ptr()), q.N, 1e-14, 30, C.byref(rep))
    return rc, b.cpu().numpy().T.copy(), rep.as_dict()


def _mem(lib):
    cur, peak = C.c_longlong(), C.c_longlong()
    lib.qd_memory_usage(C.byref(cur), C.byref(peak), 0)
    return cur.value


@pytest.mark.parametrize("precision", ["fp16", "fp64"])
def test_noalloc_solve_matches_solve_device(precision):
    from bem3d_elastostatic_amd import _lib
    from bem3d_elastostatic_amd.quadesth import Quadesth
    for name, k in (("c1", 2), ("cantilever_h05", 1)):
        g = golden(name)
        q = Quadesth(problem_from_golden(g))
        q.ghmat8()
        _, FI = q.system(with_matrix=False)
        rng = np.random.default_rng(17)
        B = np.stack([FI] + [np.abs(FI).max() * rng.normal(size=q.N) for _ in range(k - 1)], axis=1)
        print(f"[noalloc] {name} N={q.N} rhs={k} {precision}", flush=True)
        rc, X0, _ = _solve(q, q.lib.qd_solve_device_noalloc, precision, B)
        assert rc == -3, rc        # no factors held yet: QD_ERR_STATE
        rc, X_ref, rep_ref = _solve(q, q.lib.qd_solve_device, precision, B)
        _lib.check(q.h, rc, "solve_device")
        before = _mem(q.lib)
        rc, X, rep = _solve(q, q.lib.qd_solve_device_noalloc, precision, B)
        _lib.check(q.h, rc, "solve_device_noalloc")
        assert _mem(q.lib) == before
        print(f"[noalloc] {name}: {rep}", flush=True)
        assert rep["converged"] == 1 and rep["residual"] <= 1e-14, rep
        for v in range(k):
            assert rel_err(X[:, v], X_ref[:, v]) < 1e-10
        assert rel_err(X[:, 0], g["X"]) < 1e-10
        rc, X2, _ = _solve(q, q.lib.qd_solve_device, precision, B)
        _lib.check(q.h, rc, "solve_device after noalloc")
        assert rel_err(X2[:, 0], g["X"]) < 1e-10
        q.close()
