"""The GMRES fallback of the mixed-precision refinement: restarted GMRES in
FP64, left-preconditioned by the same FP16/FP32 factors and started from the
refinement's iterate.  It takes over when a sweep fails to halve the
residual, which the collocation systems here never do; QUADESTH_IR_SWEEPS=1
hands over after the first plain sweep, so these tests run the fallback on
ordinary systems.  Its solutions must meet the bars of the plain refinement:
backward error <= 1e-14 and the compiled reference's solution to 1e-10, with
several right-hand sides (GMRES runs per right-hand side) and through the
allocation-free entry point (whose workspace holds the Krylov basis).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import golden, problem_from_golden, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
@pytest.mark.parametrize("name", ["c0", "cantilever_h05", "cube8"])
def test_gmres_fallback_solution(name, precision, monkeypatch):
    from bem3d_elastostatic_amd.quadesth import Quadesth
    monkeypatch.setenv("QUADESTH_IR_SWEEPS", "1")
    g = golden(name)
    q = Quadesth(problem_from_golden(g))
    q.ghmat8()
    _, FI = q.system(with_matrix=False)
    print(f"[gmres] {name} {precision}", flush=True)
    B = np.stack([FI, -3.0 * FI], axis=1)
    X, rep = q.solve_linear_system(B, precision=precision, tol=1e-14)
    print(f"[gmres] {name} {precision}: {rep}", flush=True)
    assert rep["gmres"] == 1 and rep["converged"] == 1 and rep["residual"] <= 1e-14, rep
    assert rep["iterations"] > 1
    assert rel_err(X[:, 0], g["X"]) < 1e-10
    assert rel_err(X[:, 1], -3.0 * g["X"]) < 1e-10
    q.close()


def test_gmres_fallback_without_allocation(monkeypatch):
    import torch
    from bem3d_elastostatic_amd import _lib
    from bem3d_elastostatic_amd.quadesth import Quadesth
    g = golden("cantilever_h05")
    q = Quadesth(problem_from_golden(g))
    q.ghmat8()
    _, FI = q.system(with_matrix=False)
    print("[gmres] noalloc cantilever_h05", flush=True)
    q.solve_linear_system(FI, precision="fp16")   # the factors the no-allocation call needs
    monkeypatch.setenv("QUADESTH_IR_SWEEPS", "1")
    b = torch.tensor(FI, dtype=torch.float64, device="cuda")
    rep = _lib.SolveReport()
    cur, peak = C.c_longlong(), C.c_longlong()
    q.lib.qd_memory_usage(C.byref(cur), C.byref(peak), 0)
    before = cur.value
    rc = q.lib.qd_solve_device_noalloc(q.h, _lib.PRECISIONS["fp16"], 1, C.c_void_p(b.data_ptr()), q.N, 1e-14, 30,
                                       C.byref(rep))
    _lib.check(q.h, rc, "solve_device_noalloc")
    q.lib.qd_memory_usage(C.byref(cur), C.byref(peak), 0)
    assert cur.value == before
    r = rep.as_dict()
    print(f"[gmres] noalloc: {r}", flush=True)
    assert r["gmres"] == 1 and r["converged"] == 1 and r["residual"] <= 1e-14, r
    assert rel_err(b.cpu().numpy(), g["X"]) < 1e-10
    q.close()
