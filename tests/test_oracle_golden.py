"""The CPU oracle against the golden vectors of the reference's own compiled
Fortran (tests/golden/*.npz) and against the output file the reference ships
for its cube8.dat deck."""
import numpy as np
import pytest

from conftest import dfi_mask, golden, problem_from_golden, rel_err

SMALL = ["c0", "mixed_small", "cantilever_h05", "cube8"]


@pytest.mark.parametrize("name", SMALL)
def test_oracle_assembly_and_solution(oracle, name):
    g = golden(name)
    p = problem_from_golden(g)
    t = oracle.geometry(p)
    A, FI0 = oracle.ghmat8(p, t)
    # same statements in the same order as the reference: bit-identical here
    if "A" in g:
        assert np.array_equal(A, g["A"])
    else:
        rows = g["A_rows"]
        assert np.array_equal(A[rows], g["A_sampled"])
        assert rel_err(np.abs(A).sum(axis=1), g["A_rowsum_abs"]) < 1e-14
    assert np.array_equal(FI0, g["FI0"])
    _, _, X, info = oracle.dgesv(A, FI0)
    assert info == 0
    assert rel_err(X, g["X"]) < 1e-11
    FI, DFI = oracle.reorganize(p, g["X"])
    # the reference leaves DFI slots beyond 3*nconec unset (tri6 elements)
    valid = dfi_mask(p)
    assert np.array_equal(FI, g["FI"]) and np.array_equal(DFI[valid], g["DFI"][valid])
    if p.l > 0:
        r = oracle.internal_points(p, t, g["FI"], g["DFI"])
        assert rel_err(r["u"], g["u"]) < 1e-13
        assert rel_err(r["sigma"], g["sigma"]) < 1e-13
        assert rel_err(r["vm"], g["vm"]) < 1e-12
        # principal stresses of nearly equal eigenvalues: acos is ill-conditioned
        assert rel_err(r["prin"], g["prin"]) < 1e-8
    b = oracle.boundary_stresses(p, g["FI"], g["DFI"])
    assert rel_err(b["sigma"], g["bsigma"]) < 1e-12
    assert rel_err(b["vm"], g["bvm"]) < 1e-11


def test_oracle_reproduces_shipped_internal_points(oracle):
    """cube8.dat -> internal_points.dat as shipped by the reference (7
    significant digits, E15.7)."""
    g = golden("cube8")
    p = problem_from_golden(g)
    ship = g["shipped_internal_points"]
    assert np.array_equal(ship[:, 1:4].round(4), p.points.round(4))
    r = oracle.run(p, solver="numpy")
    for cols, key, tol in (((4, 7), "u", 6e-7), ((7, 13), "sigma", 1e-6)):
        got = r[key]
        ref = ship[:, cols[0]:cols[1]]
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() / scale < tol
    assert rel_err(r["vm"], ship[:, 13]) < 1e-6


def test_oracle_c1_rows(oracle):
    """c1 (18,438 DOF): sampled rows of A and FI against the reference."""
    g = golden("c1")
    p = problem_from_golden(g)
    rows = g["A_rows"]
    nodes = np.unique(rows // 3)
    rr, Ar, Fr = oracle.ghmat8_rows(p, nodes)
    sel = np.isin(rr, rows)
    order = np.searchsorted(rows, rr[sel])
    assert np.array_equal(Ar[sel], g["A_sampled"][order])
    assert np.array_equal(Fr, g["FI0"][rr])
