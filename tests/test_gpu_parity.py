"""Parity of the HIP library (through its C ABI) with the golden vectors of
the reference's compiled Fortran and with the CPU oracle.

Tolerances: the reference path is FP64; the bar is 1e-10 relative for the
solution, internal u and sigma.  The assembled matrix is compared at 1e-12
relative to its largest entry: the library sums the element contributions
chunk by chunk in colour order, the reference element by element, and the
kernels are regrouped algebraically, so the last bits differ.  The measured
errors are printed."""
import numpy as np
import pytest

from conftest import dfi_mask, golden, problem_from_golden, rel_err

pytestmark = pytest.mark.gpu

CASES = ["c0", "mixed_small", "cantilever_h05", "cube8"]


def _q(p):
    from bem3d_elastostatic_amd.quadesth import Quadesth
    return Quadesth(p)


@pytest.mark.parametrize("name", CASES)
def test_geometry_tables(oracle, name):
    g = golden(name)
    p = problem_from_golden(g)
    print(f"case {name}: geometry", flush=True)
    q = _q(p)
    t = q.precompute_element_geometry(tables=True)
    ref = oracle.geometry(p)
    # unit-normal components that vanish carry ~1e-13 of rounding noise
    # (FMA contraction in the dc/dct sums); positions and Jacobians are tighter
    for k in ("eta1", "eta2", "eta3", "xja", "xg1", "xg2", "xg3"):
        assert rel_err(t[k].reshape(-1), ref[k]) < 2e-12, k
    assert rel_err(t["ff"].reshape(-1), ref["ff"]) < 1e-13


@pytest.mark.parametrize("name", CASES)
def test_assembly_matches_reference(name):
    g = golden(name)
    p = problem_from_golden(g)
    print(f"case {name}: assembly n={p.n} ne={p.ne}", flush=True)
    q = _q(p)
    q.ghmat8()
    A, FI0 = q.system()
    if "A" in g:
        ea = rel_err(A, g["A"])
    else:
        rows = g["A_rows"]
        ea = rel_err(A[rows], g["A_sampled"])
        assert rel_err(np.abs(A).sum(axis=1), g["A_rowsum_abs"]) < 1e-12
        assert rel_err(A.sum(axis=0), g["A_colsum"]) < 1e-10
    ef = rel_err(FI0, g["FI0"])
    print(f"[parity] {name}: A {ea:.1e} FI {ef:.1e}", flush=True)
    assert ea < 1e-12
    assert ef < 1e-12


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_solve_matches_reference(name, precision):
    g = golden(name)
    p = problem_from_golden(g)
    print(f"case {name}: solve {precision}", flush=True)
    q = _q(p)
    q.ghmat8()
    X, rep = q.solve_linear_system(precision=precision, tol=1e-14)
    assert rep["info"] == 0
    assert rel_err(X, g["X"]) < 1e-10, rep
    if precision != "fp64":
        assert rep["converged"] == 1 and rep["residual"] <= 1e-14


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("variant", [0, 1])
def test_internal_points_match(oracle, name, variant):
    g = golden(name)
    p = problem_from_golden(g)
    if p.l == 0:
        pytest.skip("no internal points")
    print(f"case {name}: internal points variant {variant}", flush=True)
    q = _q(p)
    r = q.compute_internal_points(g["FI"], g["DFI"], variant=variant)
    if variant == 0:
        ref = {k: g[k] for k in ("u", "sigma", "vm", "prin")}
    else:
        ref = oracle.internal_points(p, oracle.geometry(p), g["FI"], g["DFI"], variant=1)
    e = {k: rel_err(r[k], ref[k]) for k in ("u", "sigma", "vm", "prin")}
    print(f"[parity] {name} variant {variant}: " + " ".join(f"{k} {v:.1e}" for k, v in e.items()), flush=True)
    assert e["u"] < 1e-10
    assert e["sigma"] < 1e-10
    assert e["vm"] < 1e-10
    assert e["prin"] < 1e-7
    oc = oracle.internal_points(p, oracle.geometry(p), g["FI"], g["DFI"])["counts"]
    assert np.array_equal(r["counts"], oc)


@pytest.mark.parametrize("name", CASES)
def test_end_to_end(name):
    g = golden(name)
    p = problem_from_golden(g)
    print(f"case {name}: end to end", flush=True)
    q = _q(p)
    out = q.run(precision="fp64")
    assert rel_err(out["FI"], g["FI"]) < 1e-10
    valid = dfi_mask(p)
    assert rel_err(out["DFI"][valid], g["DFI"][valid]) < 1e-10
    if p.l > 0:
        assert rel_err(out["u"], g["u"]) < 1e-10
        assert rel_err(out["sigma"], g["sigma"]) < 1e-10
    b = q.compute_boundary_stresses(out["FI"], out["DFI"])
    assert rel_err(b["sigma"], g["bsigma"]) < 1e-10


def test_c1_full_size():
    """c1: 6146 nodes, 18,438 DOF, 4096 internal points."""
    g = golden("c1")
    p = problem_from_golden(g)
    print("case c1: end to end fp64 + fp16-IR", flush=True)
    q = _q(p)
    q.ghmat8()
    A, FI0 = q.system()
    rows = g["A_rows"]
    assert rel_err(A[rows], g["A_sampled"]) < 1e-12
    assert rel_err(np.abs(A).sum(axis=1), g["A_rowsum_abs"]) < 1e-12
    assert rel_err(FI0, g["FI0"]) < 1e-12
    del A
    for prec in ("fp64", "fp32", "fp16"):
        X, rep = q.solve_linear_system(precision=prec)
        assert rel_err(X, g["X"]) < 1e-10, (prec, rep)
    FI, DFI = q.reorganize_solution(X)
    r = q.compute_internal_points(FI, DFI)
    assert rel_err(r["u"], g["u"]) < 1e-10
    assert rel_err(r["sigma"], g["sigma"]) < 1e-10


def test_ghmat8_rhs_matches_full_assembly():
    from bem3d_elastostatic_amd import mesh
    p = mesh.config1()
    p2 = mesh.unit_cube(4)
    rhs = mesh.config2_rhs(mesh.unit_cube(4))
    q = _q(p2)
    q.ghmat8()
    for r in range(rhs.shape[0]):
        fi = q.ghmat8_rhs(rhs[r])
        p3 = mesh.unit_cube(4)
        p3.bc = rhs[r]
        q3 = _q(p3)
        q3.ghmat8()
        _, fi_full = q3.system(with_matrix=False)
        assert rel_err(fi, fi_full) < 1e-14
    del p


def test_dgesv_dropin_matches_lapack():
    import scipy.linalg.lapack as lapack
    from bem3d_elastostatic_amd.quadesth import dgesv
    rng = np.random.default_rng(11)
    for n in (1, 5, 64, 100, 517, 1100):
        print(f"dgesv n={n}", flush=True)
        A = rng.normal(size=(n, n))
        b = rng.normal(size=(n, 3))
        lu, ipiv, x, info = dgesv(A, b)
        lu_s, piv_s, x_s, info_s = lapack.dgesv(A, b)
        assert info == 0
        assert np.array_equal(ipiv, piv_s + 1)
        assert np.allclose(lu, lu_s, rtol=1e-9, atol=1e-11)
        assert np.allclose(x, x_s, rtol=1e-9, atol=1e-10)
    A = np.zeros((4, 4))
    A[0, 0] = 2.0
    A[1, 1] = 3.0
    _, _, _, info = dgesv(A, np.ones(4))
    assert info == 3


def test_transpose_solve():
    import ctypes as C
    import torch
    from bem3d_elastostatic_amd import _lib
    g = golden("cantilever_h05")
    p = problem_from_golden(g)
    q = _q(p)
    q.ghmat8()
    A, _ = q.system()
    c = np.random.default_rng(2).normal(size=A.shape[0])
    y_ref = np.linalg.solve(A.T, c)
    for prec in (0, 1, 2):
        cd = torch.tensor(c, dtype=torch.float64, device="cuda")
        rep = _lib.SolveReport()
        rc = q.lib.qd_solve_transpose_device(q.h, prec, 1, C.c_void_p(cd.data_ptr()), cd.numel(), 1e-14, 30,
                                             C.byref(rep))
        _lib.check(q.h, rc)
        assert rel_err(cd.cpu().numpy(), y_ref) < 1e-10, prec


def test_bad_mesh_is_rejected():
    from bem3d_elastostatic_amd import mesh
    from bem3d_elastostatic_amd._lib import QuadesthError
    p = mesh.config0()
    p.kconec = p.kconec.copy()
    p.kconec[0, 0] = p.n + 5
    with pytest.raises(QuadesthError):
        _q(p)
