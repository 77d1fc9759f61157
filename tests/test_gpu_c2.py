"""Eight right-hand sides on the 18,438-DOF cube (config 2).

The matrix depends on kode only (assembly_mod.f90:126-190), so one ghmat8
serves all eight load cases of ``mesh.config2_rhs``: uniaxial x, y, z; simple
shear; four smooth random top-face tractions.  qd_ghmat8_rhs forms each FI.
Each mixed-precision solve (FP16 and FP32 factors, FP64 refinement, all
eight in one call) is compared per right-hand side against the FP64 LU of
the same GPU and against the compiled reference (tests/golden/c2.npz: the
reference's ghmat8 for each load case, solved with one LU of its A).
* Every FI from qd_ghmat8_rhs equals the reference's FI to 1e-12, and every
  solution (FP64, FP32-IR, FP16-IR) the reference's to 1e-10.
* Uniaxial z is config 1's own load case (tests/golden/c1.npz).
* Every FP64 solution has a small residual against the host copy of A.
"""
import numpy as np
import pytest

from conftest import golden, rel_err

pytestmark = pytest.mark.gpu


def test_eight_rhs_mixed_vs_fp64():
    from bem3d_elastostatic_amd import mesh
    from bem3d_elastostatic_amd.quadesth import Quadesth
    p = mesh.config1()
    rhs_bc = mesh.config2_rhs(p)
    print("[c2] ghmat8 n=6146 + 8 right-hand sides", flush=True)
    q = Quadesth(p)
    q.precompute_element_geometry()
    q.ghmat8()
    A, FI_assembled = q.system(with_matrix=True)
    B = np.stack([q.ghmat8_rhs(rhs_bc[r]) for r in range(8)], axis=1)
    # uniaxial z is the assembled load case
    assert np.array_equal(B[:, 2], FI_assembled)

    X64, rep64 = q.solve_linear_system(B, precision="fp64")
    res = np.abs(A @ X64 - B).max(axis=0) / (np.abs(A).sum(axis=1).max() * np.abs(X64).max(axis=0)
                                            + np.abs(B).max(axis=0))
    assert res.max() < 1e-14, res
    g = golden("c1")
    assert rel_err(X64[:, 2], g["X"]) < 1e-10
    ref = golden("c2")
    assert np.array_equal(ref["bc"], rhs_bc)
    for r in range(8):
        assert rel_err(B[:, r], ref["FI0"][:, r]) < 1e-12, r
    e64 = [rel_err(X64[:, r], ref["X"][:, r]) for r in range(8)]
    print(f"[c2] fp64 vs reference per rhs {['%.1e' % e for e in e64]}", flush=True)
    assert max(e64) < 1e-10

    for prec in ("fp16", "fp32"):
        print(f"[c2] solve 8 rhs {prec}", flush=True)
        X, rep = q.solve_linear_system(B, precision=prec)
        assert rep["converged"], rep
        errs = [rel_err(X[:, r], X64[:, r]) for r in range(8)]
        eref = [rel_err(X[:, r], ref["X"][:, r]) for r in range(8)]
        print(f"[c2] {prec}: iterations {rep['iterations']} gmres {rep['gmres']} "
              f"error vs FP64 per rhs {['%.1e' % e for e in errs]} vs reference {['%.1e' % e for e in eref]}",
              flush=True)
        assert max(errs) < 1e-10, errs
        assert max(eref) < 1e-10, eref
    q.close()
