"""The oracle on perturbed meshes: the objective of the gradient
configuration (J = sum of internal von Mises stress squared) and its central
differences reproduce the reference's (tests/golden/grad_*.npz, written by
the compiled reference), so the oracle can check the GPU gradient kernels."""
import dataclasses

import numpy as np
import pytest

from conftest import golden, problem_from_golden


def _J(oracle, p):
    r = oracle.run(p, solver="numpy")
    return float(np.sum(r["vm"] ** 2))


@pytest.mark.parametrize("name", ["grad_c0", "grad_mixed"])
def test_oracle_fd_matches_reference_fd(name, oracle):
    g = golden(name)
    p = problem_from_golden(g)
    assert abs(_J(oracle, p) - float(g["J"])) <= 1e-12 * abs(float(g["J"]))
    h = float(g["h"])
    xyz = p.coords()
    for i in range(2):
        vals = []
        for s in (1.0, -1.0):
            c = xyz + s * h * g["directions"][i]
            q = dataclasses.replace(p, x=c[:, 0].copy(), y=c[:, 1].copy(), z=c[:, 2].copy(),
                                    nu=p.nu + s * h * float(g["dnu"][i]))
            vals.append(_J(oracle, q))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g["fd"][i]) <= 1e-6 * abs(g["fd"][i]), (i, fd, g["fd"][i])
