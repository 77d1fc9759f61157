"""Structured boundary meshes, boundary conditions and internal points for the
QUADESTH hot path.

The solver consumes the same arrays the reference keeps as module globals
(data_mod.f90:38-52): node coordinates ``x, y, z(n)``, element connectivity
``nconec(ne)`` / ``kconec(9, ne)`` (1-based node numbers, Fortran column-major,
so element ``e`` occupies ``kconec[e, :]`` in a C-ordered ``(ne, 9)`` array),
per-element boundary-condition codes ``kode(3*ne)`` (0 = displacement
prescribed, 1 = traction prescribed, io_mod.f90:97-113) and prescribed values
``bc(27, ne)`` stored as ``bc[e, 3*k + ld]``.

Node orderings follow the reference's element definitions:

* quad9 (stress_mod.f90:47-50): 1 (-1,-1), 2 (0,-1), 3 (1,-1), 4 (1,0),
  5 (1,1), 6 (0,1), 7 (-1,1), 8 (-1,0), 9 (0,0);
* tri6 (stress_mod.f90:84-87): vertices 1, 2, 3 then mid-side 4 (1-2),
  5 (2-3), 6 (3-1).

The unit normal is ``x_,xi1 x x_,xi2`` (assembly_mod.f90:318-334); every
generator here orients it out of the solid, which is what the rigid-body
diagonal of ghmat8 (assembly_mod.f90:100-123) assumes.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

# quad9 local node -> (xi1, xi2) in half-element grid units (0, 1, 2)
_Q9_LOCAL = [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (1, 2), (0, 2), (0, 1), (1, 1)]


@dataclass
class BemProblem:
    """A complete QUADESTH input deck held as arrays (io_mod.f90:36-146)."""

    title: str
    x: np.ndarray            # (n,) float64
    y: np.ndarray
    z: np.ndarray
    nconec: np.ndarray       # (ne,) int32, 9 (quad9) or 6 (tri6)
    kconec: np.ndarray       # (ne, 9) int32, 1-based, unused slots 0
    kode: np.ndarray         # (ne, 3) int32
    bc: np.ndarray           # (ne, 27) float64, bc[e, 3*k + ld]
    mu: float
    nu: float
    points: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))  # (l, 3)
    ntrack: int = 1
    extra_rhs: Optional[np.ndarray] = None   # (nrhs-1, ne, 27) additional bc sets

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def ne(self) -> int:
        return int(self.nconec.shape[0])

    @property
    def l(self) -> int:
        return int(self.points.shape[0])

    @property
    def young(self) -> float:
        return 2.0 * self.mu * (1.0 + self.nu)

    def coords(self) -> np.ndarray:
        return np.stack([self.x, self.y, self.z], axis=1)


class _NodeTable:
    """Assigns 1-based node numbers to half-grid lattice points in order of
    first appearance, so that element order and node order stay spatially
    coherent."""

    def __init__(self, origin, step):
        self.origin = np.asarray(origin, dtype=np.float64)
        self.step = np.asarray(step, dtype=np.float64)
        self.ids = {}
        self.pts = []

    def node(self, ijk):
        key = tuple(int(v) for v in ijk)
        nid = self.ids.get(key)
        if nid is None:
            nid = len(self.pts) + 1
            self.ids[key] = nid
            self.pts.append(key)
        return nid

    def coords(self):
        p = np.asarray(self.pts, dtype=np.float64)
        return self.origin[None, :] + p * self.step[None, :]


def _face_frames(nx, ny, nz):
    """Per box face: (name, fixed axis, fixed half-grid value, u axis, v axis,
    u count, v count) with u x v pointing out of the box."""
    m = {0: nx, 1: ny, 2: nz}
    faces = [
        ("x0", 0, 0, 2, 1),
        ("x1", 0, 2 * nx, 1, 2),
        ("y0", 1, 0, 0, 2),
        ("y1", 1, 2 * ny, 2, 0),
        ("z0", 2, 0, 1, 0),
        ("z1", 2, 2 * nz, 0, 1),
    ]
    return [(nm, ax, val, ua, va, m[ua], m[va]) for (nm, ax, val, ua, va) in faces]


def box_mesh(lo, hi, nx, ny, nz, tri_faces=()):
    """Surface mesh of the box ``[lo, hi]`` with ``nx x ny x nz`` element
    divisions. Faces named in ``tri_faces`` (any of x0 x1 y0 y1 z0 z1) are
    split into tri6 pairs along the diagonal from the (u, v) origin corner;
    all other faces use quad9.

    Returns ``(coords (n,3), nconec (ne,), kconec (ne,9), face_of (ne,) str)``.
    """
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    step = (hi - lo) / (2.0 * np.array([nx, ny, nz], dtype=np.float64))
    nt = _NodeTable(lo, step)
    nconec, kconec, face_of = [], [], []
    for (nm, ax, val, ua, va, nu_, nv_) in _face_frames(nx, ny, nz):
        def lattice(u, v):
            p = [0, 0, 0]
            p[ax] = val
            p[ua] = u
            p[va] = v
            return p
        for b in range(nv_):
            for a in range(nu_):
                if nm in tri_faces:
                    c00 = (2 * a, 2 * b)
                    c10 = (2 * a + 2, 2 * b)
                    c11 = (2 * a + 2, 2 * b + 2)
                    c01 = (2 * a, 2 * b + 2)
                    for (p1, p2, p3) in ((c00, c10, c11), (c00, c11, c01)):
                        mids = [((p1[0] + p2[0]) // 2, (p1[1] + p2[1]) // 2),
                                ((p2[0] + p3[0]) // 2, (p2[1] + p3[1]) // 2),
                                ((p3[0] + p1[0]) // 2, (p3[1] + p1[1]) // 2)]
                        loc = [p1, p2, p3] + mids
                        ids = [nt.node(lattice(*q)) for q in loc]
                        nconec.append(6)
                        kconec.append(ids + [0, 0, 0])
                        face_of.append(nm)
                else:
                    ids = [nt.node(lattice(2 * a + du, 2 * b + dv)) for (du, dv) in _Q9_LOCAL]
                    nconec.append(9)
                    kconec.append(ids)
                    face_of.append(nm)
    return (nt.coords(), np.asarray(nconec, dtype=np.int32),
            np.asarray(kconec, dtype=np.int32), np.asarray(face_of))


def _bc_arrays(nconec, face_of, face_bcs):
    """``face_bcs[name] = ((kx, ky, kz), (vx, vy, vz))`` -> kode (ne,3), bc (ne,27)."""
    ne = nconec.shape[0]
    kode = np.ones((ne, 3), dtype=np.int32)
    bc = np.zeros((ne, 27), dtype=np.float64)
    for e in range(ne):
        codes, vals = face_bcs.get(face_of[e], ((1, 1, 1), (0.0, 0.0, 0.0)))
        kode[e] = codes
        for k in range(int(nconec[e])):
            for ld in range(3):
                bc[e, 3 * k + ld] = vals[ld]
    return kode, bc


def _mu_from_e(young, nu):
    return young / (2.0 * (1.0 + nu))


UNIAXIAL_CUBE_BCS = {
    "z0": ((1, 1, 0), (0.0, 0.0, 0.0)),
    "x0": ((0, 1, 1), (0.0, 0.0, 0.0)),
    "y0": ((1, 0, 1), (0.0, 0.0, 0.0)),
    "z1": ((1, 1, 1), (0.0, 0.0, 1.0)),
}


def unit_cube(m: int, young: float = 1000.0, nu: float = 0.3, points=None,
              bcs=None, title=None) -> BemProblem:
    """Unit cube, ``m x m`` quad9 per face, symmetry-plane supports and a unit
    top traction (the uniaxial-stress box)."""
    coords, nconec, kconec, face_of = box_mesh((0, 0, 0), (1, 1, 1), m, m, m)
    kode, bc = _bc_arrays(nconec, face_of, bcs if bcs is not None else UNIAXIAL_CUBE_BCS)
    pts = np.zeros((0, 3)) if points is None else np.asarray(points, dtype=np.float64)
    p = BemProblem(title or f"UNIT CUBE {m}x{m} QUAD9 PER FACE, UNIAXIAL TRACTION",
                   coords[:, 0].copy(), coords[:, 1].copy(), coords[:, 2].copy(),
                   nconec, kconec, kode, bc, _mu_from_e(young, nu), nu, pts)
    p.face_of = face_of
    return p


def config0() -> BemProblem:
    """Unit cube, 2x2 quad9 per face (24 elements, 98 nodes, 294 DOF), E=1000,
    nu=0.3, 27 internal points on {0.25, 0.5, 0.75}^3 (x slowest, z fastest)."""
    g = np.array([0.25, 0.5, 0.75])
    pts = np.array([[a, b, c] for a in g for b in g for c in g])
    return unit_cube(2, points=pts, title="UNIT CUBE 2x2 QUAD9, UNIAXIAL STRESS (E=1000, NU=0.3)")


def config1(npoints: int = 4096) -> BemProblem:
    """Unit cube, 16x16 quad9 per face (1536 elements, 6146 nodes, 18438 DOF),
    4096 internal points uniform in [0.05, 0.95]^3 from default_rng(0)."""
    pts = np.random.default_rng(0).uniform(0.05, 0.95, (4096, 3))[:npoints]
    return unit_cube(16, points=pts, title="UNIT CUBE 16x16 QUAD9, UNIAXIAL STRESS (E=1000, NU=0.3)")


def config2_rhs(p: BemProblem) -> np.ndarray:
    """The 8 traction right-hand sides on the config-1 mesh and KODE pattern:
    uniaxial x, y, z; simple shear (t_x on the top face); and 4 smooth top-face
    fields t_z = sum_{k,l<=3} a_kl sin(k pi x) sin(l pi y), a ~ N(0,1) from
    default_rng(1), 9 coefficients per field drawn in order.

    Returned as ``(8, ne, 27)`` bc arrays. The KODE pattern of ``p`` is kept,
    so one system matrix serves all eight (the matrix depends only on KODE,
    assembly_mod.f90:126-190).
    """
    face_of = p.face_of
    ne = p.ne
    out = np.zeros((8, ne, 27))
    rng = np.random.default_rng(1)
    amps = [rng.standard_normal(9) for _ in range(4)]
    xyz = p.coords()
    for e in range(ne):
        f = face_of[e]
        for k in range(int(p.nconec[e])):
            node = p.kconec[e, k] - 1
            xx, yy = xyz[node, 0], xyz[node, 1]
            if f == "x1":
                out[0, e, 3 * k + 0] = 1.0
            if f == "y1":
                out[1, e, 3 * k + 1] = 1.0
            if f == "z1":
                out[2, e, 3 * k + 2] = 1.0
                out[3, e, 3 * k + 0] = 1.0
                for r in range(4):
                    a = amps[r]
                    s = 0.0
                    i = 0
                    for kk in range(1, 4):
                        for ll in range(1, 4):
                            s += a[i] * np.sin(kk * np.pi * xx) * np.sin(ll * np.pi * yy)
                            i += 1
                    out[4 + r, e, 3 * k + 2] = s
    # displacement-prescribed components keep their (zero) prescribed values
    return out


def config3(npoints: int = 65536, h: float = 0.1) -> BemProblem:
    """Thick cantilever box [0,10]x[0,2]x[0,1] with element size ``h``: tri6 on
    z=0 and z=1 (each square split on its diagonal), quad9 elsewhere. x=0
    clamped, x=10 uniform traction t_z=-1, other faces traction-free. E=210e3,
    nu=0.3. Points uniform in [0.1,9.9]x[0.1,1.9]x[0.1,0.9] from default_rng(2).
    At h=0.1: 25602 nodes, 2400 quad9 + 8000 tri6, 76806 DOF."""
    nx, ny, nz = int(round(10 / h)), int(round(2 / h)), int(round(1 / h))
    coords, nconec, kconec, face_of = box_mesh((0, 0, 0), (10, 2, 1), nx, ny, nz,
                                               tri_faces=("z0", "z1"))
    bcs = {"x0": ((0, 0, 0), (0.0, 0.0, 0.0)), "x1": ((1, 1, 1), (0.0, 0.0, -1.0))}
    kode, bc = _bc_arrays(nconec, face_of, bcs)
    pts = np.random.default_rng(2).uniform([0.1, 0.1, 0.1], [9.9, 1.9, 0.9], (65536, 3))[:npoints]
    p = BemProblem("CANTILEVER BOX 10x2x1, TRI6 ON Z FACES, QUAD9 ELSEWHERE",
                   coords[:, 0].copy(), coords[:, 1].copy(), coords[:, 2].copy(),
                   nconec, kconec, kode, bc, _mu_from_e(210e3, 0.3), 0.3, pts)
    p.face_of = face_of
    return p


def radial_bump(p: BemProblem, amplitude: float, seed: int) -> BemProblem:
    """Moves the nodes of ``p`` (in place) by a seeded radial bump field:
    displacement amplitude * exp(-|x - c|^2 / w^2) along x - c, with the
    centre c and width w drawn from ``seed`` (c in [0.3, 0.7]^3 scaled to the
    bounding box, w in [0.25, 0.45] times its size).  A node on one of the
    bounding box's minimum planes (the supported faces of the configs) keeps
    that coordinate, so those faces stay planar."""
    rng = np.random.default_rng(seed)
    xyz = p.coords()
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    size = float(np.max(hi - lo))
    centre = lo + rng.uniform(0.3, 0.7, 3) * (hi - lo)
    width = rng.uniform(0.25, 0.45) * size
    d = xyz - centre[None, :]
    r2 = np.sum(d * d, axis=1)
    bump = amplitude * np.exp(-r2 / (width * width))
    rn = np.sqrt(np.maximum(r2, 1e-30))
    disp = (bump / rn)[:, None] * d
    for ax in range(3):
        disp[np.isclose(xyz[:, ax], lo[ax]), ax] = 0.0
    xyz = xyz + disp
    p.x, p.y, p.z = xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()
    return p


def config4(amplitude: float = 0.02, seed: int = 3, npoints: int = 4096) -> BemProblem:
    """Config-1 cube with node coordinates perturbed by a seeded radial bump
    field that keeps the supported faces (x=0, y=0, z=0) planar."""
    p = radial_bump(config1(npoints), amplitude, seed)
    p.title = "PERTURBED UNIT CUBE 16x16 QUAD9 (RADIAL BUMP 0.02)"
    return p


def unit_cube_fan_top(m: int = 4, young: float = 1000.0, nu: float = 0.3, points=None) -> BemProblem:
    """Unit cube with m x m quad9 per face, except the top face z=1, which is a
    fan of 4m tri6 elements around its centre: the centre node belongs to 4m
    elements (16 for m = 4), more than the reference's ke(8, n) holds
    (io_mod.f90:131).  Boundary conditions are config 0's uniaxial case
    (exact solution u = (-nu x, -nu y, z)/E) except that the second half of
    the fan, in element order, prescribes u_z = 1/E instead of t_z = 1: the
    centre's first displacement-prescribing element (assembly_mod.f90:149-161)
    is then its (2m+1)-th."""
    coords, nconec, kconec, face_of = box_mesh((0, 0, 0), (1, 1, 1), m, m, m)
    keep = np.array([f != "z1" for f in face_of])
    nconec, kconec = nconec[keep], kconec[keep]
    face_of = [f for f, k in zip(face_of, keep) if k]
    used = np.unique(kconec[kconec > 0])
    # perimeter of the top face, counter-clockwise seen from +z: (vertex, mid) pairs
    h = 1.0 / m
    ring = []
    for i in range(m):                      # y = 0, x increasing
        ring.append(((i * h, 0.0), ((i + 0.5) * h, 0.0)))
    for i in range(m):                      # x = 1, y increasing
        ring.append(((1.0, i * h), (1.0, (i + 0.5) * h)))
    for i in range(m):                      # y = 1, x decreasing
        ring.append(((1.0 - i * h, 1.0), (1.0 - (i + 0.5) * h, 1.0)))
    for i in range(m):                      # x = 0, y decreasing
        ring.append(((0.0, 1.0 - i * h), (0.0, 1.0 - (i + 0.5) * h)))
    pts = [tuple(coords[j - 1]) for j in used]
    index = {}
    for j, c in zip(used, pts):
        index[(round(c[0], 12), round(c[1], 12), round(c[2], 12))] = len(index) + 1
    old_to_new = {int(j): k + 1 for k, j in enumerate(used)}
    kconec = np.vectorize(lambda v: old_to_new.get(int(v), 0))(kconec).astype(np.int32)
    xyz = list(pts)

    def node(x, y, z=1.0):
        key = (round(x, 12), round(y, 12), round(z, 12))
        if key not in index:
            xyz.append((x, y, z))
            index[key] = len(xyz)
        return index[key]

    centre = node(0.5, 0.5)
    nk = len(ring)
    fan_k = np.zeros((nk, 9), dtype=np.int32)
    for k in range(nk):
        (v0, mid), (v1, _) = ring[k], ring[(k + 1) % nk]
        a, b = node(*v0), node(*v1)
        sa = node(0.5 * (0.5 + v0[0]), 0.5 * (0.5 + v0[1]))
        sb = node(0.5 * (0.5 + v1[0]), 0.5 * (0.5 + v1[1]))
        # vertices 1-2-3 counter-clockwise from +z, then mid-sides 1-2, 2-3, 3-1
        fan_k[k, :6] = (centre, a, b, sa, node(*mid), sb)
    nconec = np.concatenate([nconec, np.full(nk, 6, dtype=np.int32)])
    kconec = np.concatenate([kconec, fan_k])
    face_of = face_of + ["z1"] * nk
    coords = np.array(xyz, dtype=np.float64)
    bcs = {"z0": ((1, 1, 0), (0.0, 0.0, 0.0)), "x0": ((0, 1, 1), (0.0, 0.0, 0.0)),
           "y0": ((1, 0, 1), (0.0, 0.0, 0.0)), "z1": ((1, 1, 1), (0.0, 0.0, 1.0))}
    kode, bc = _bc_arrays(nconec, face_of, bcs)
    ne = nconec.shape[0]
    for e in range(ne - nk // 2, ne):
        kode[e] = (1, 1, 0)
        bc[e, 2::3] = 0.0
        bc[e, 2:3 * 6:3] = 1.0 / young
    if points is None:
        g = np.array([0.25, 0.5, 0.75])
        points = np.array([[a, b, c] for a in g for b in g for c in g])
    return BemProblem(f"UNIT CUBE {m}x{m} QUAD9, TOP FACE A FAN OF {nk} TRI6", coords[:, 0].copy(),
                      coords[:, 1].copy(), coords[:, 2].copy(), nconec, kconec, kode, bc,
                      _mu_from_e(young, nu), nu, points)


def adjacency(p: BemProblem, cap: int = 8):
    """``nesim(n)`` and ``ke(n, cap)`` (1-based element numbers, ascending),
    exactly as io_mod.f90:124-135 fills them (entries beyond ``cap`` dropped)."""
    nesim = np.zeros(p.n, dtype=np.int32)
    ke = np.zeros((p.n, cap), dtype=np.int32)
    for e in range(p.ne):
        for j in range(int(p.nconec[e])):
            node = int(p.kconec[e, j]) - 1
            nesim[node] += 1
            if nesim[node] <= cap:
                ke[node, nesim[node] - 1] = e + 1
    return nesim, ke
