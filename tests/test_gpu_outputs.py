"""The reference's text outputs written from the library's own results
(Quadesth.write_outputs: the deck echo, output_data, the boundary- and
internal-stress tables, stresses.dat, internal_points.dat, disp_trac.dat),
against the files the reference's compiled writers produced for the same
decks (tests/golden/outputs.npz), and internal_points.dat against the file
the reference ships for cube8.dat.

Text outside the numbers must be identical, line for line.  Numbers are
printed to 7 or 8 significant digits from FP64 results that differ from the
reference's by about 1e-13 relative, so a printed value may move by one unit
in its last digit; values that are cancellation noise (shear stresses whose
exact value is 0) may move by more, which the bound 1e-10 x the largest
magnitude of the same table (records of one layout) covers."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden, problem_from_golden

pytestmark = pytest.mark.gpu

FILES = ("cube8_out.dat", "stresses.dat", "internal_points.dat", "disp_trac.dat")
NUM = re.compile(r"[-+]?\d*\.\d+(?:[EeDd][-+]?\d+)?")
# a number with the blanks before it: a sign takes one of its field's blanks
FIELD = re.compile(r"\s*[-+]?\d*\.\d+(?:[EeDd][-+]?\d+)?")


def _unit(b):
    """One unit in the last printed digit of b (E format, 7 or 8 digits, taken from the token)."""
    mant = b.split("E")[0].lstrip("-+")
    digits = len(mant.split(".")[1]) if "." in mant else 0
    e = int(b.split("E")[1]) if "E" in b else 0
    return 10.0 ** (e - digits)


INT = re.compile(r"[-+]?\d+")


def compare_text(got, want):
    gl, wl = got.splitlines(), want.splitlines()
    assert len(gl) == len(wl)
    # a table: the records of one layout (numbers and integers masked)
    scales = {}
    for b in wl:
        key = INT.sub("#", FIELD.sub(" #", b))
        vals = [abs(float(t)) for t in NUM.findall(b)]
        if vals:
            scales[key] = max(scales.get(key, 0.0), max(vals))
    worst = 0.0
    for a, b in zip(gl, wl):
        if a == b:
            continue
        assert FIELD.sub(" #", a) == FIELD.sub(" #", b) and len(a) == len(b), (a, b)
        ta, tb = NUM.findall(a), NUM.findall(b)
        scale = scales[INT.sub("#", FIELD.sub(" #", b))]
        for x, y in zip(ta, tb):
            dx = abs(float(x) - float(y))
            if dx == 0.0:
                continue
            assert dx <= 1.01 * _unit(y) or dx <= 1e-10 * scale, (a, b)
            worst = max(worst, dx / max(scale, 1e-300))
    return worst


@pytest.mark.parametrize("name,title,ntrack", [
    ("cube8", "CUBIC DOMAIN UNDER UNIFORM TRACTION - 3D ELASTOSTATICS (QUADESTH)", 770),
    ("mixed_small", "MIXED SMALL BOX", 1)])
def test_outputs_match_reference_writers(tmp_path, name, title, ntrack):
    from bem3d_elastostatic_amd.quadesth import Quadesth
    g = golden(name)
    p = problem_from_golden(g)
    p.title, p.ntrack = title, ntrack
    print(f"[outputs] {name}", flush=True)
    q = Quadesth(p)
    out = q.run(precision="fp64")
    q.write_outputs(str(tmp_path), out["FI"], out["DFI"], internal=out)
    ref = np.load(os.path.join(ROOT, "tests", "golden", "outputs.npz"))
    for f in FILES:
        with open(tmp_path / f) as fh:
            got = fh.read()
        w = compare_text(got, ref[f"{name}/{f}"].tobytes().decode())
        print(f"[outputs] {name} {f}: largest change {w:.1e} of its table's scale", flush=True)
    if name == "cube8":
        mine = np.loadtxt(tmp_path / "internal_points.dat", comments="#")
        ship = g["shipped_internal_points"]
        assert mine.shape == ship.shape
        assert np.array_equal(mine[:, :4], ship[:, :4])   # point numbers and coordinates
        d = np.abs(mine[:, 4:] - ship[:, 4:])
        unit = np.where(ship[:, 4:] == 0, 0.0, 10.0 ** (np.floor(np.log10(np.abs(ship[:, 4:]) + 1e-300)) + 1 - 7))
        sig_scale = np.abs(ship[:, 7:13]).max(axis=1, keepdims=True)
        ok = (d <= 1.01 * unit) | (d <= 1e-10 * np.maximum(sig_scale, np.abs(ship[:, 4:7]).max()))
        assert ok.all()
    q.close()
