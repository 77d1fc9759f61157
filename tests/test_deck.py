"""Input-deck hardening (io_mod.f90:8-146): the legacy deck layout and nodes
shared by more than 8 elements.

* tests/golden/legacy_mixed.dat holds the mixed tri6/quad9 box in the legacy
  layout: an 8-value header ``n, ne, l, nfr, mu, dam, ro, nu`` plus a
  frequency line (io_mod.f90:51-61), and prescribed values as complex
  ``(re, im)`` pairs, some with D exponents (io_mod.f90:18-27).  The
  reference's own reader, compiled from its sources, read that file for the
  golden legacy_mixed.npz; deck.read_deck must yield the same
  problem, so that the oracle assembles the reference's A and FI from it bit
  for bit.
* unit_cube_fan_top(4) has a node in 16 elements.  The reference keeps 8 in
  ke(8, n) (io_mod.f90:131) and its kode0 search (assembly_mod.f90:149-161)
  and reorganize_solution (io_mod.f90:168-180) then read past them; the
  library and the oracle search every element of the node in ascending
  order.  Pinned by brute force (the first displacement-prescribing element)
  and by the exact uniaxial solution, which the 8-entry search misses by 14x.
"""
import os

import numpy as np

from conftest import GOLDEN, golden, rel_err


def test_legacy_deck_reads_as_the_reference_reads_it(oracle):
    from bem3d_elastostatic_amd import deck
    p = deck.read_deck(os.path.join(GOLDEN, "legacy_mixed.dat"))
    g = golden("legacy_mixed")
    base = golden("mixed_small")   # the same problem, written as a modern deck
    for k in ("x", "y", "z", "nconec", "kconec", "kode", "bc", "points"):
        assert np.array_equal(getattr(p, k if k != "points" else "points"), base[k]), k
    assert p.mu == float(g["mu"]) == float(base["mu"]) and p.nu == float(g["nu"]) == float(base["nu"])
    assert p.ntrack == 7
    A, FI0 = oracle.ghmat8(p)
    assert np.array_equal(A, g["A"]) and np.array_equal(FI0, g["FI0"])
    # and the reference's outputs for the two layouts are the same numbers
    for k in ("A", "FI0", "X", "u", "sigma", "bsigma"):
        assert np.array_equal(g[k], base[k]), k


def test_modern_deck_round_trip(tmp_path):
    from bem3d_elastostatic_amd import deck, mesh
    p = mesh.config3(npoints=16, h=0.5)
    path = str(tmp_path / "c.dat")
    deck.write_deck(p, path)
    q = deck.read_deck(path)
    for k in ("x", "y", "z", "nconec", "kconec", "kode", "bc", "points"):
        assert np.array_equal(getattr(p, k), getattr(q, k)), k
    assert (q.mu, q.nu) == (p.mu, p.nu)


def _first_kode0(p):
    """(n, 3) prescribed displacement of the first element (ascending number)
    that contains the node and prescribes direction ld, NaN if none."""
    out = np.full((p.n, 3), np.nan)
    for e in range(p.ne):
        for k in range(int(p.nconec[e])):
            i = int(p.kconec[e, k]) - 1
            for ld in range(3):
                if p.kode[e, ld] == 0 and np.isnan(out[i, ld]):
                    out[i, ld] = p.bc[e, 3 * k + ld]
    return out


def test_node_in_more_than_eight_elements(oracle):
    from bem3d_elastostatic_amd import mesh
    p = mesh.unit_cube_fan_top(4)
    nesim, _ = oracle.adjacency(p)
    centre = int(np.argmin((p.x - 0.5) ** 2 + (p.y - 0.5) ** 2 + (p.z - 1.0) ** 2))
    assert nesim[centre] == 16 and nesim.max() == 16
    r = oracle.run(p)
    # reorganisation: every displacement-prescribed dof takes the first such
    # element's value (brute force over all elements)
    vp = _first_kode0(p)
    FI = r["FI"].reshape(p.n, 3)
    has = ~np.isnan(vp)
    assert np.array_equal(FI[has], vp[has])
    assert FI[centre, 2] == 1.0 / p.young       # found at the centre's 9th element
    E, nu = p.young, p.nu
    exact = np.stack([-nu * p.x / E, -nu * p.y / E, p.z / E], axis=1)
    err = np.abs(FI - exact).max() / np.abs(exact).max()
    # the fan's long triangles with the fixed 4x4 rule on neighbours cost
    # ~3e-3 (the same fan with tractions only: 3.8e-3); a search stopping at
    # ke's 8 entries mixes the centre's displacement and traction unknowns in
    # one column and misses by 4.7e-2
    assert err < 5e-3, err


def test_eight_entry_search_would_miss(oracle):
    """The same system assembled and reorganised with ke cut to the
    reference's 8 entries: the centre's u_z is then an unknown beside the
    traction of its displacement-prescribing fan elements, and the solution
    misses the exact field by more than 10x the full search's error."""
    import ctypes as C
    from bem3d_elastostatic_amd import mesh
    p = mesh.unit_cube_fan_top(4)
    t = oracle.geometry(p)
    lib = oracle.lib
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)   # noqa: E731
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    lib.qo_set_ke_cap(8)
    nesim = np.zeros(p.n, dtype=np.int32)
    ke = np.zeros(8 * p.n, dtype=np.int32)
    lib.qo_adjacency(p.n, p.ne, i32(p.nconec), i32(p.kconec), nesim, ke)
    N = 3 * p.n
    A, FI0 = np.zeros(N * N), np.zeros(N)
    m = oracle.material(p.mu, p.nu)
    lib.qo_ghmat8(p.n, p.ne, f64(p.x), f64(p.y), f64(p.z), i32(p.nconec), i32(p.kconec), i32(p.kode), f64(p.bc),
                  nesim, ke, C.byref(m), *oracle._tabs(t), A, FI0)
    X = np.linalg.solve(A.reshape(N, N).T, FI0)
    DFI = np.zeros(27 * p.ne)
    lib.qo_reorganize(p.n, p.ne, i32(p.nconec), i32(p.kconec), i32(p.kode), f64(p.bc), nesim, ke, p.mu, X, DFI)
    E, nu = p.young, p.nu
    exact = np.stack([-nu * p.x / E, -nu * p.y / E, p.z / E], axis=1)
    err8 = np.abs(X.reshape(p.n, 3) - exact).max() / np.abs(exact).max()
    full = oracle.run(p)["FI"].reshape(p.n, 3)
    err_full = np.abs(full - exact).max() / np.abs(exact).max()
    assert err8 > 10 * err_full, (err8, err_full)
