"""Panel-width schedules of the mixed-precision factorisation on c1 (18,438 DOF).

With the look-ahead, the aux stream's U12 solve can still run while the main
stream starts the next step's solve for the panel after next.  Panels wider
than the strip solve takes (2048) give that main-stream solve inner FP16
updates, so the two solves need separate scratch: shared scratch corrupts the
factors (the refinement then needs tens of sweeps instead of a handful).  The
schedules below cover 1024 throughout, 2048 throughout, 2048 then 1024, 4096
then 1024 (the default schedule's widths) and 768 throughout, and the default
widths with the diagonal-block panels or the U12 solves through L11^-1 turned
off."""
import os

import pytest

from conftest import golden, problem_from_golden, rel_err

pytestmark = pytest.mark.gpu

SCHEDULES = {
    "narrow": {"QUADESTH_NB": "1024", "QUADESTH_NB_WIDE": "0"},
    "wide": {"QUADESTH_NB": "2048", "QUADESTH_NB_WIDE": "0"},
    "wide_then_narrow": {"QUADESTH_NB": "1024", "QUADESTH_NB_WIDE": "2048", "QUADESTH_NB_WIDE_REM": "8192"},
    "wider_then_narrow": {"QUADESTH_NB": "1024", "QUADESTH_NB_WIDE": "4096", "QUADESTH_NB_WIDE_REM": "8192"},
    # a first panel that is not whole 512-column slices: A's conversion into
    # the factors' matrix rounds its first launch up to whole slices (and
    # U11^-1 of a 768-column block is padded to 1024)
    "odd_width": {"QUADESTH_NB": "768", "QUADESTH_NB_WIDE": "0"},
    # the recursive panels over all rows and the strip U12 solves, which the
    # FP32 path and the fallback keep
    "recursive_panels": {"QUADESTH_NB": "1024", "QUADESTH_NB_WIDE": "4096", "QUADESTH_NB_WIDE_REM": "8192",
                         "QUADESTH_BLOCK_PANEL": "0"},
    "strip_u12": {"QUADESTH_NB": "1024", "QUADESTH_NB_WIDE": "4096", "QUADESTH_NB_WIDE_REM": "8192",
                  "QUADESTH_INV_U12": "0"},
    # the two-launch steps of the diagonal blocks (from 8 trailing tiles on)
    "split_steps": {"QUADESTH_NB": "1024", "QUADESTH_NB_WIDE": "4096", "QUADESTH_NB_WIDE_REM": "8192",
                    "QUADESTH_BLK_SPLIT": "8"},
}


@pytest.mark.parametrize("name", list(SCHEDULES))
def test_fp16_schedule_matches_golden(name, monkeypatch):
    from bem3d_elastostatic_amd.quadesth import Quadesth
    for k, v in SCHEDULES[name].items():
        monkeypatch.setenv(k, v)
    g = golden("c1")
    q = Quadesth(problem_from_golden(g))
    q.ghmat8()
    print(f"[schedule] c1 fp16 {name}: {SCHEDULES[name]}", flush=True)
    X, rep = q.solve_linear_system(precision="fp16", tol=1e-14)
    print(f"[schedule] {name}: {rep}", flush=True)
    assert rep["converged"] == 1 and rep["gmres"] == 0, rep
    assert rep["iterations"] <= 10, rep
    assert rel_err(X, g["X"]) < 1e-10
