"""The assembly's write plan (bem3d_elastostatic_amd/csrc/qd_plan.cpp) on the CPU.

The H and G passes write each column block of A by the plan: a chain of one
node's element updates inside a chunk starts from zero where no earlier
update wrote the column (no memset of A), is carried between elements in an
LDS slot, and is stored once.  tests/plan_harness.cpp replays the passes for
a sample of collocation rows both ways, plain read-modify-write of a zeroed
matrix and by the plan, and requires bit-identical results; the order of the
floating-point additions per entry is the same in both.  The cases cover
quad9-only and mixed quad9/tri6 meshes, prescribed displacements on traction
faces (the kode0/vp columns), nodes in more than eight elements, and the
76,806-DOF cantilever.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden, problem_from_golden

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("plan") / "plan_harness")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--cuda-host-only",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "bem3d_elastostatic_amd", "csrc"),
                    os.path.join(ROOT, "tests", "plan_harness.cpp"),
                    os.path.join(ROOT, "bem3d_elastostatic_amd", "csrc", "qd_plan.cpp"), "-o", exe],
                   check=True, capture_output=True)
    return exe


def _case(name):
    from bem3d_elastostatic_amd import deck, mesh
    if name == "cantilever_76806":
        return mesh.config3(npoints=16)
    if name == "fan_top":
        return mesh.unit_cube_fan_top(4)
    if name == "legacy_mixed":
        return deck.read_deck(os.path.join(ROOT, "tests", "golden", "legacy_mixed.dat"))
    return problem_from_golden(golden(name))


def _run(harness, p, d, nrows, seed=0):
    n, ne = int(p.n), int(p.ne)
    kconec = np.zeros((ne, 9), dtype=np.int32)
    kc = np.asarray(p.kconec, dtype=np.int32).reshape(ne, -1)
    kconec[:, :kc.shape[1]] = kc
    nconec = np.asarray(p.nconec, dtype=np.int32)
    for e in range(ne):
        kconec[e, nconec[e]:] = 0
    rng = np.random.default_rng(seed)
    if nrows >= n:
        rows = np.arange(n, dtype=np.int32)
    else:
        # random rows plus the nodes of a few elements (their own blocks are
        # the singular ones)
        own = kconec[rng.choice(ne, 4, replace=False)].reshape(-1) - 1
        rows = np.unique(np.concatenate([rng.choice(n, nrows, replace=False), own[own >= 0]])).astype(np.int32)
    os.makedirs(d, exist_ok=True)
    nconec.tofile(os.path.join(d, "nconec.bin"))
    kconec.tofile(os.path.join(d, "kconec.bin"))
    np.ascontiguousarray(p.kode, dtype=np.int32).reshape(-1).tofile(os.path.join(d, "kode.bin"))
    np.ascontiguousarray(p.bc, dtype=np.float64).reshape(-1).tofile(os.path.join(d, "bc.bin"))
    rows.tofile(os.path.join(d, "rows.bin"))
    with open(os.path.join(d, "meta.txt"), "w") as fh:
        fh.write(f"{n} {ne}\n")
    r = subprocess.run([harness, d], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    stats = dict(zip(r.stdout.split()[0::2], (int(v) for v in r.stdout.split()[1::2])))
    return stats


@pytest.mark.parametrize("name,nrows", [("c0", 10 ** 9), ("mixed_small", 10 ** 9), ("cantilever_h05", 10 ** 9),
                                        ("cube8", 10 ** 9), ("legacy_mixed", 10 ** 9), ("fan_top", 10 ** 9),
                                        ("c1", 96), ("cantilever_76806", 40)])
def test_plan_matches_read_modify_write(harness, tmp_path, name, nrows):
    p = _case(name)
    s = _run(harness, p, str(tmp_path / name), nrows)
    assert s["mismatches"] == 0
    # every column of a consistent deck is written by some pass
    assert s["unwritten"] == 0
    assert s["chains"] + s["carried"] == s["incidences"]
    if name == "cantilever_76806":
        # the carry removes about a third of the chains on the cantilever's
        # strip-ordered elements (69,600 element-node updates, 25,602 nodes)
        assert s["incidences"] == 69600
        assert s["chains"] < 0.75 * s["incidences"]
