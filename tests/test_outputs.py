"""The text outputs (io_mod.f90:44-140, 191-255; stress_mod.f90:412-468,
892-936) against the files the reference's own writers produced, compiled
from its sources, for the same numbers (tests/golden/outputs.npz).

The library's writers (csrc/qd_output.cpp) are built here for the host
alone, with tests/output_harness.cpp, and fed the reference's results of
each golden case (its FI, DFI, internal and boundary results; the
element-nodal stresses, which the golden omits, from the oracle, whose nodal
averages equal the reference's bit for bit).  Every file must equal the
reference's byte for byte: cube8 (quad9, 1965 points) and the mixed
tri6/quad9 box with prescribed displacements."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden, problem_from_golden

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FILES = ("cube8_out.dat", "stresses.dat", "internal_points.dat", "disp_trac.dat")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("harness") / "output_harness")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--cuda-host-only",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "bem3d_elastostatic_amd", "csrc"),
                    os.path.join(ROOT, "tests", "output_harness.cpp"),
                    os.path.join(ROOT, "bem3d_elastostatic_amd", "csrc", "qd_output.cpp"), "-o", exe],
                   check=True, capture_output=True)
    return exe


def _write_inputs(d, p, g, sigma_elem):
    arrs = dict(x=p.x, y=p.y, z=p.z, bc=p.bc.reshape(-1), FI=g["FI"], DFI=g["DFI"].reshape(-1),
                u=g["u"].reshape(-1), s=g["sigma"].reshape(-1), pr=g["prin"].reshape(-1), vm=g["vm"],
                bs=g["bsigma"].reshape(-1), bp=g["bprin"].reshape(-1), bv=g["bvm"],
                se=sigma_elem.reshape(-1), pts=p.points.reshape(-1))
    for k, v in arrs.items():
        np.ascontiguousarray(v, dtype=np.float64).tofile(os.path.join(d, f"{k}.bin"))
    for k in ("nconec", "kconec", "kode"):
        np.ascontiguousarray(getattr(p, k), dtype=np.int32).tofile(os.path.join(d, f"{k}.bin"))
    with open(os.path.join(d, "meta.txt"), "w") as fh:
        fh.write(f"{p.n} {p.ne} {p.l} {p.ntrack} {float(g['mu'])!r} {float(g['nu'])!r}\n{p.title}\n")


@pytest.mark.parametrize("name,title,ntrack", [
    ("cube8", "CUBIC DOMAIN UNDER UNIFORM TRACTION - 3D ELASTOSTATICS (QUADESTH)", 770),
    ("mixed_small", "MIXED SMALL BOX", 1)])
def test_writers_match_reference_files(harness, oracle, tmp_path, name, title, ntrack):
    g = golden(name)
    p = problem_from_golden(g)
    p.title, p.ntrack = title, ntrack
    b = oracle.boundary_stresses(p, g["FI"], g["DFI"])
    assert np.array_equal(b["sigma"], g["bsigma"])   # the element-nodal stresses average to the reference's
    d = str(tmp_path)
    _write_inputs(d, p, g, b["sigma_elem"])
    subprocess.run([harness, d], check=True)
    ref = np.load(os.path.join(ROOT, "tests", "golden", "outputs.npz"))
    for f in FILES:
        want = ref[f"{name}/{f}"].tobytes()
        with open(os.path.join(d, f), "rb") as fh:
            got = fh.read()
        assert got == want, f
