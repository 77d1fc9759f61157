"""Pivoting of the mixed-precision factorisations.

They first keep their diagonal pivots and accept the factors only when every
multiplier satisfies |l| <= 1 (then each pivot is the one partial pivoting
would choose); otherwise they refactor with tournament pivoting.  The BEM
systems of the golden cases pass the check.  The same system with its rows
reversed (and its right-hand side with them) has the same solution but fails
it, so it exercises the fallback."""
import ctypes as C

import numpy as np
import pytest

from conftest import golden, problem_from_golden, rel_err

pytestmark = pytest.mark.gpu


def _hip():
    from conftest import hip_runtime
    return hip_runtime()


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
def test_diagonal_pivots_kept(precision):
    from bem3d_elastostatic_amd.quadesth import Quadesth
    g = golden("cube8")
    q = Quadesth(problem_from_golden(g))
    q.ghmat8()
    print(f"[pivoting] cube8 {precision}: diagonal", flush=True)
    X, rep = q.solve_linear_system(precision=precision, tol=1e-14)
    assert rep["pivoting"] == 1, rep
    assert rep["converged"] == 1 and rep["residual"] <= 1e-14
    assert rel_err(X, g["X"]) < 1e-10


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
def test_reversed_rows_fall_back_to_tournament(precision):
    from bem3d_elastostatic_amd.quadesth import Quadesth
    g = golden("cube8")
    q = Quadesth(problem_from_golden(g))
    q.ghmat8()
    A, FI = q.system()
    N = A.shape[0]
    pa, lda, _ = q.system_device()
    # column-major with leading dimension lda, rows reversed
    Ar = np.zeros((N, lda))
    Ar[:, :N] = A[::-1, :].T
    hip = _hip()
    hip.hipMemcpy.restype = C.c_int
    rc = hip.hipMemcpy(C.c_void_p(pa), Ar.ctypes.data_as(C.c_void_p), C.c_size_t(Ar.nbytes), C.c_int(1))
    assert rc == 0
    print(f"[pivoting] cube8 reversed rows {precision}: fallback", flush=True)
    X, rep = q.solve_linear_system(FI[::-1].copy(), precision=precision, tol=1e-14)
    assert rep["pivoting"] == 0, rep
    assert rep["converged"] == 1 and rep["residual"] <= 1e-14
    assert rel_err(X, g["X"]) < 1e-10


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
@pytest.mark.parametrize("row", [10, 300, 2000])
def test_nan_multiplier_falls_back(precision, row):
    """A NaN entry of A in column 5 gives a NaN multiplier in the first
    panel (1024 columns of N = 2598): row 10 inside its first 64 x 64 tile
    (the one-wave elimination), row 300 further down its diagonal block (the
    64-column steps' L tiles, or the recursive panel's L21 rows in FP32), row
    2000 below the diagonal block (with FP16 updates: L21 = A21 U11^-1 by the
    FP16 product, whose epilogue takes the growth).  The largest-multiplier
    word must rank the NaN above every bound, so the factors are redone with
    tournament pivoting (pivoting == 0); the refinement then reports
    QD_ERR_NOCONV instead of a result."""
    import torch
    from bem3d_elastostatic_amd import _lib
    from bem3d_elastostatic_amd.quadesth import Quadesth
    g = golden("cube8")
    q = Quadesth(problem_from_golden(g))
    q.ghmat8()
    _, FI = q.system(with_matrix=False)
    N = FI.shape[0]
    pa, lda, _ = q.system_device()
    nan = np.array([np.nan])
    rc = _hip().hipMemcpy(C.c_void_p(pa + 8 * (5 * lda + row)), nan.ctypes.data_as(C.c_void_p), C.c_size_t(8),
                          C.c_int(1))
    assert rc == 0
    b = torch.tensor(FI, dtype=torch.float64, device="cuda")
    rep = _lib.SolveReport()
    print(f"[pivoting] cube8 NaN at A[{row}, 5] {precision}", flush=True)
    rc = q.lib.qd_solve_device(q.h, _lib.PRECISIONS[precision], 1, C.c_void_p(b.data_ptr()), N, 1e-14, 30,
                               C.byref(rep))
    print(f"[pivoting] rc {rc} {rep.as_dict()}", flush=True)
    assert rep.pivoting == 0, rep.as_dict()
    assert rc == -6 and rep.converged == 0
