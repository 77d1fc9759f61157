"""The 76,806-DOF thick cantilever (config 3) at its full size, against the
oracle.

At this size the system matrix is 47 GB (lda 76,864), so every column from
27,940 on lies beyond 2^31 elements of the matrix's start; the FP16
factorisation runs its default 4096-then-1024 panel schedule and the
look-back solves their largest grids.  The oracle cannot hold the whole
matrix, so the checks sample it:
* rows of A and FI of 32 collocation nodes (each row spans every column)
  against ``ghmat8_rows`` of the oracle (assembly_mod.f90:53-223), 1e-12
  relative to the largest sampled entry;
* the FP64 solve's backward error ||b - Ax|| / (||A|| ||x|| + ||b||) (inf-norms)
  against the FP64 A, <= 1e-14, reported by the library over every row and
  recomputed here on the sampled rows;
* the FP16-IR solution against the FP64 solution, 1e-10 relative;
* internal u and sigma at all 65,536 points plus 64 points 0.02-0.06 from the
  faces (near-field rule), from the FP64 solution: 256 of them against the
  oracle's Somigliana (stress_mod.f90:511-781) on the same FI and DFI, 1e-10;
  all of them against the FP16-IR solution's, 1e-10.
"""
import numpy as np
import pytest

from conftest import device_rows, rel_err

pytestmark = pytest.mark.gpu


def near_face_points(k, seed=7):
    """k points inside the box [0,10]x[0,2]x[0,1], each 0.02-0.06 from one face."""
    rng = np.random.default_rng(seed)
    lo, hi = np.array([0.0, 0.0, 0.0]), np.array([10.0, 2.0, 1.0])
    pts = rng.uniform(lo + 0.1, hi - 0.1, (k, 3))
    for i in range(k):
        ax, side = divmod(i % 6, 2)
        d = rng.uniform(0.02, 0.06)
        pts[i, ax] = lo[ax] + d if side == 0 else hi[ax] - d
    return pts


@pytest.fixture(scope="module")
def c3():
    from bem3d_elastostatic_amd import mesh
    from bem3d_elastostatic_amd.quadesth import Quadesth
    p = mesh.config3()
    assert (p.n, p.ne, 3 * p.n, p.l) == (25602, 10400, 76806, 65536)
    print("[c3] ghmat8 n=25602 ne=10400 N=76806", flush=True)
    q = Quadesth(p)
    q.precompute_element_geometry()
    q.ghmat8()
    _, lda, _ = q.system_device()
    assert lda * 27940 >= 2 ** 31
    state = {"p": p, "q": q}
    yield state
    q.close()


def test_c3_rows_match_oracle(c3, oracle):
    p, q = c3["p"], c3["q"]
    rng = np.random.default_rng(5)
    x, y, z = p.x, p.y, p.z
    extremes = [0, p.n - 1, int(np.argmin(x + y + z)), int(np.argmax(x + y + z)),
                int(np.argmax(x - y)), int(np.argmin(x - z)),
                int(np.argmin(np.abs(x - 5.0) + np.abs(z - 1.0))),   # top face (tri6)
                int(np.argmin(np.abs(x - 10.0) + np.abs(y - 1.0)))]  # loaded end (traction)
    nodes = np.unique(np.concatenate([extremes, rng.choice(p.n, 24, replace=False)])).astype(np.int32)
    print(f"[c3] {len(nodes)} rows of A: nodes {nodes.tolist()}", flush=True)
    rows, A_ref, FI_ref = oracle.ghmat8_rows(p, nodes)
    A_gpu = device_rows(q, nodes)
    _, FI = q.system(with_matrix=False)
    assert rel_err(A_gpu, A_ref) < 1e-12
    assert rel_err(FI[rows], FI_ref) < 1e-12
    # the far columns (offsets beyond 2^31 elements) on their own
    assert rel_err(A_gpu[:, 27940:], A_ref[:, 27940:]) < 1e-12
    c3["rows"], c3["A_rows"], c3["FI"] = rows, A_gpu, FI


def test_c3_solves(c3):
    q = c3["q"]
    FI = c3["FI"]
    print("[c3] fp16-IR solve", flush=True)
    X16, rep16 = q.solve_linear_system(FI, precision="fp16", tol=1e-14)
    print(f"[c3] fp16-IR: {rep16}", flush=True)
    assert rep16["converged"] == 1 and rep16["residual"] <= 1e-14, rep16
    print("[c3] fp64 solve", flush=True)
    X64, rep64 = q.solve_linear_system(FI, precision="fp64")
    print(f"[c3] fp64: {rep64}", flush=True)
    assert rep64["info"] == 0
    assert 0.0 < rep64["residual"] <= 1e-14, rep64
    A_rows, rows = c3["A_rows"], c3["rows"]
    r = A_rows @ X64 - FI[rows]
    anorm_lb = np.abs(A_rows).sum(axis=1).max()   # a lower bound of ||A||_inf
    berr = np.abs(r).max() / (anorm_lb * np.abs(X64).max() + np.abs(FI).max())
    print(f"[c3] fp64 backward error on the sampled rows {berr:.2e}", flush=True)
    assert berr <= 1e-14
    assert rel_err(X16, X64) < 1e-10
    c3["X16"], c3["X64"] = X16, X64


def test_c3_refinement_sweeps(c3):
    """How the FP16-IR solution approaches the FP64 solution sweep by sweep:
    the refinement stops at the first sweep whose backward error is within
    tol, so tol = 1e-8 ... 1e-14 on the kept factors give the sweep counts and
    the forward errors against the FP64 solution (printed).  The library's
    default (tol <= 0) is LAPACK DSGESV's bound 2^-53 sqrt(N) = 3.08e-14 here,
    between 1e-12 and 1e-14; the forward error it reaches is below 1e-10."""
    q, FI, X64 = c3["q"], c3["FI"], c3["X64"]
    prev_sweeps, prev_err = 0, np.inf
    dflt = 2.0**-53 * np.sqrt(q.N)
    for tol in (1e-8, 1e-10, 1e-12, 0.0, 1e-14):
        X, rep = q.solve_linear_system(FI, precision="fp16", tol=tol)
        err = rel_err(X, X64)
        print(f"[c3] tol {tol if tol > 0 else dflt:.2e}: {rep['iterations']} sweeps, backward {rep['residual']:.2e}, "
              f"forward vs fp64 {err:.2e}", flush=True)
        assert rep["converged"] == 1 and rep["residual"] <= (tol if tol > 0 else dflt)
        assert rep["iterations"] >= prev_sweeps and err <= 2.0 * prev_err + 1e-15
        if tol == 0.0:
            assert err < 1e-10
        prev_sweeps, prev_err = rep["iterations"], err
    assert prev_err < 1e-10


def test_c3_internal_points(c3, oracle):
    p, q = c3["p"], c3["q"]
    extra = near_face_points(64)
    pts = np.concatenate([p.points, extra])
    FI64, DFI64 = q.reorganize_solution(c3["X64"])
    FI16, DFI16 = q.reorganize_solution(c3["X16"])
    print(f"[c3] internal points l={len(pts)}", flush=True)
    r64 = q.compute_internal_points(FI64, DFI64, points=pts)
    r16 = q.compute_internal_points(FI16, DFI16, points=pts)
    assert rel_err(r16["u"], r64["u"]) < 1e-10
    assert rel_err(r16["sigma"], r64["sigma"]) < 1e-10
    sel = np.concatenate([np.random.default_rng(9).choice(p.l, 192, replace=False), p.l + np.arange(64)])
    print("[c3] oracle at 256 points", flush=True)
    ref = oracle.internal_points(p, oracle.geometry(p), FI64, DFI64, points=pts[sel])
    assert ref["counts"][1] > 0   # the near-field rule ran
    for k in ("u", "sigma", "vm"):
        assert rel_err(r64[k][sel], ref[k]) < 1e-10, k
    # the 64 near-face points on their own (their fields are the largest)
    near = np.arange(192, 256)
    assert rel_err(r64["sigma"][sel][near], ref["sigma"][near]) < 1e-10
