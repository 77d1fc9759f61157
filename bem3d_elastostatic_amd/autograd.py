"""Adjoint gradients of the hot path with respect to node coordinates and nu.

Two ``torch.autograd.Function`` s wrap the C ABI:

* ``BoundarySolve``: (coords (n, 3), nu) -> (FI (3n,), DFI (ne, 27)).
  Forward runs ghmat8 (assembly_mod.f90:53-223), the solve
  (solver_mod.f90:17-36) and reorganize_solution (io_mod.f90:148-183) on the
  GPU. Backward:
  1. transposes the reorganisation: g_X;
  2. solves A^T lambda = g_X with the forward's LU factors
     (``qd_solve_transpose``);
  3. returns -d/d(coords, nu) of lambda^T (A X - FI) from the hand-written
     kernels of ``qd_ghmat8_vjp``, which recompute the Gauss-point geometry.
* ``InternalPoints``: (coords, nu, FI, DFI) -> (u (l, 3), sigma (l, 6)).
  Forward runs compute_internal_points (stress_mod.f90:511-781); backward
  runs ``qd_internal_points_vjp``.

``objective_vm2`` composes them into config 4's J = sum over the internal
points of von Mises stress squared.  Tensors are float64 on the host: the
arrays crossing the boundary are O(n) and every O(n^2) and O(n^3) step stays
on the GPU.  One ``GradSession`` holds one library handle. A backward must
run while that handle still holds the forward's coordinates and factors;
otherwise it raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import SolveReport, check
from .mesh import BemProblem
from .quadesth import PRECISIONS, Quadesth, _f64, _ptr


class GradSession:
    """A handle configured for one mesh, boundary-value set and point set; the
    coordinates and nu change per evaluation."""

    def __init__(self, problem: BemProblem, points=None, variant: int = 0, precision: str = "fp64",
                 device: int = 0):
        self.p = problem
        self.q = Quadesth(problem, device)
        self.lib = self.q.lib
        self.points = _f64(problem.points if points is None else points).reshape(-1, 3)
        self.variant = int(variant)
        self.precision = precision
        self.version = 0
        self._state = None
        n, ne = problem.n, problem.ne
        kode = np.asarray(problem.kode, dtype=np.int32).reshape(ne, 3)
        kc = np.asarray(problem.kconec, dtype=np.int64).reshape(ne, 9) - 1
        nc = np.asarray(problem.nconec, dtype=np.int64)
        # reorganize_solution's maps (io_mod.f90:148-183): the dofs whose
        # displacement is prescribed somewhere, and the DFI slots equal to mu*X.
        prescribed = np.zeros(3 * n, dtype=bool)
        slots, dofs = [], []
        for e in range(ne):
            for ld in range(3):
                if kode[e, ld] == 0:
                    for k in range(nc[e]):
                        prescribed[3 * kc[e, k] + ld] = True
                        slots.append(e * 27 + 3 * k + ld)
                        dofs.append(3 * kc[e, k] + ld)
        self._free = ~prescribed
        self._slots = np.asarray(slots, dtype=np.int64)
        self._dofs = np.asarray(dofs, dtype=np.int64)

    # -- state ---------------------------------------------------------------
    def set_state(self, coords: np.ndarray, nu: float) -> None:
        xyz = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).reshape(self.p.n, 3))
        key = (xyz.tobytes(), float(nu))
        if self._state == key:
            return
        x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
        h = self.q.h
        check(h, self.lib.qd_set_coordinates(h, _ptr(x), _ptr(y), _ptr(z)), "set_coordinates")
        check(h, self.lib.qd_set_material(h, float(self.p.mu), float(nu)), "set_material")
        self._state = key
        self.version += 1
        self._solved = None

    def solve(self):
        """Assembly, solve and reorganisation at the current state."""
        q = self.q
        q.precompute_element_geometry()
        q.ghmat8()
        X, rep = q.solve_linear_system(precision=self.precision)
        FI, DFI = q.reorganize_solution(X)
        self._solved = self.version
        return X, FI, DFI, rep

    def internal_points(self, FI, DFI):
        r = self.q.compute_internal_points(FI, DFI, points=self.points, variant=self.variant)
        return r["u"], r["sigma"]

    # -- adjoints ------------------------------------------------------------
    def reorganize_transpose(self, g_fi: np.ndarray, g_dfi: np.ndarray) -> np.ndarray:
        """g_X from (g_FI, g_DFI): the transpose of reorganize_solution."""
        gx = np.where(self._free, np.asarray(g_fi, dtype=np.float64).reshape(-1), 0.0)
        np.add.at(gx, self._dofs, float(self.p.mu) * np.asarray(g_dfi, dtype=np.float64).reshape(-1)[self._slots])
        return gx

    def adjoint_solve(self, g: np.ndarray) -> np.ndarray:
        if self._solved != self.version:
            raise RuntimeError("the handle's factors do not belong to the current state")
        c = np.ascontiguousarray(np.asarray(g, dtype=np.float64).copy())
        rep = SolveReport()
        N = 3 * self.p.n
        rc = self.lib.qd_solve_transpose(self.q.h, PRECISIONS[self.precision], 1, _ptr(c), N, 1e-14, 30,
                                         C.byref(rep))
        check(self.q.h, rc, "solve_transpose")
        if rc > 0:
            raise _lib.QuadesthError(rc, "singular factor in the adjoint solve")
        return c

    def ghmat8_vjp(self, lam, FI, DFI):
        n = self.p.n
        gx, gy, gz = np.zeros(n), np.zeros(n), np.zeros(n)
        gnu = C.c_double()
        check(self.q.h, self.lib.qd_ghmat8_vjp(self.q.h, _ptr(_f64(lam)), _ptr(_f64(FI)),
                                               _ptr(_f64(DFI).reshape(-1)), _ptr(gx), _ptr(gy), _ptr(gz),
                                               C.byref(gnu)), "ghmat8_vjp")
        return np.stack([gx, gy, gz], axis=1), gnu.value

    def internal_points_vjp(self, FI, DFI, g_u=None, g_sigma=None):
        n, ne = self.p.n, self.p.ne
        pts = self.points
        l = pts.shape[0]
        cx, cy, cz = (np.ascontiguousarray(pts[:, i]) for i in range(3))
        gfi, gdfi = np.zeros(3 * n), np.zeros(27 * ne)
        gx, gy, gz = np.zeros(n), np.zeros(n), np.zeros(n)
        gnu = C.c_double()
        gu = None if g_u is None else _f64(g_u).reshape(-1)
        gs = None if g_sigma is None else _f64(g_sigma).reshape(-1)
        check(self.q.h, self.lib.qd_internal_points_vjp(
            self.q.h, l, _ptr(cx), _ptr(cy), _ptr(cz), _ptr(_f64(FI)), _ptr(_f64(DFI).reshape(-1)),
            self.variant, None if gu is None else _ptr(gu), None if gs is None else _ptr(gs), _ptr(gfi),
            _ptr(gdfi), _ptr(gx), _ptr(gy), _ptr(gz), C.byref(gnu)), "internal_points_vjp")
        return gfi, gdfi.reshape(ne, 27), np.stack([gx, gy, gz], axis=1), gnu.value

    def close(self):
        self.q.close()


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).numpy()


class BoundarySolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, nu, session: GradSession):
        session.set_state(_np(coords), float(nu))
        X, FI, DFI, rep = session.solve()
        ctx.session = session
        ctx.version = session.version
        ctx.fi, ctx.dfi = FI, DFI
        ctx.report = rep
        out = (torch.from_numpy(FI.copy()), torch.from_numpy(DFI.copy()))
        return tuple(t.to(coords.device) for t in out)

    @staticmethod
    def backward(ctx, g_fi, g_dfi):
        s: GradSession = ctx.session
        if s.version != ctx.version:
            raise RuntimeError("GradSession state changed between forward and backward")
        gfi = np.zeros(3 * s.p.n) if g_fi is None else _np(g_fi)
        gdfi = np.zeros((s.p.ne, 27)) if g_dfi is None else _np(g_dfi)
        lam = s.adjoint_solve(s.reorganize_transpose(gfi, gdfi))
        gc, gnu = s.ghmat8_vjp(lam, ctx.fi, ctx.dfi)
        dev = g_fi.device if g_fi is not None else g_dfi.device
        return (torch.from_numpy(-gc).to(dev), torch.tensor(-gnu, dtype=torch.float64, device=dev), None)


class InternalPoints(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, nu, FI, DFI, session: GradSession):
        session.set_state(_np(coords), float(nu))
        fi, dfi = _np(FI), _np(DFI)
        u, sigma = session.internal_points(fi, dfi)
        ctx.session = session
        ctx.version = session.version
        ctx.fi, ctx.dfi = fi, dfi
        return torch.from_numpy(u.copy()).to(coords.device), torch.from_numpy(sigma.copy()).to(coords.device)

    @staticmethod
    def backward(ctx, g_u, g_sigma):
        s: GradSession = ctx.session
        if s.version != ctx.version:
            raise RuntimeError("GradSession state changed between forward and backward")
        gfi, gdfi, gc, gnu = s.internal_points_vjp(ctx.fi, ctx.dfi, None if g_u is None else _np(g_u),
                                                   None if g_sigma is None else _np(g_sigma))
        dev = (g_sigma if g_sigma is not None else g_u).device
        return (torch.from_numpy(gc).to(dev), torch.tensor(gnu, dtype=torch.float64, device=dev),
                torch.from_numpy(gfi).to(dev), torch.from_numpy(gdfi).to(dev), None)


def von_mises_sq(sigma: torch.Tensor) -> torch.Tensor:
    """vm^2 of calc_principal_and_vm (stress_mod.f90:299-359) per row of
    sigma (xx, yy, zz, xy, yz, zx)."""
    sxx, syy, szz, sxy, syz, szx = sigma.unbind(-1)
    return 0.5 * ((sxx - syy) ** 2 + (syy - szz) ** 2 + (szz - sxx) ** 2) + 3.0 * (sxy ** 2 + syz ** 2 + szx ** 2)


def objective_vm2(session: GradSession, coords: torch.Tensor, nu: torch.Tensor) -> torch.Tensor:
    """J = sum over the internal points of (von Mises stress)^2."""
    FI, DFI = BoundarySolve.apply(coords, nu, session)
    _, sigma = InternalPoints.apply(coords, nu, FI, DFI, session)
    return von_mises_sq(sigma).sum()


def gradient_vm2(session: GradSession, coords: np.ndarray, nu: float):
    """(J, dJ/dcoords (n, 3), dJ/dnu) by one forward and one backward."""
    c = torch.tensor(np.asarray(coords, dtype=np.float64), requires_grad=True)
    v = torch.tensor(float(nu), dtype=torch.float64, requires_grad=True)
    J = objective_vm2(session, c, v)
    J.backward()
    return float(J.detach()), c.grad.numpy().copy(), float(v.grad)


def value_vm2(session: GradSession, coords: np.ndarray, nu: float) -> float:
    """J alone (no graph)."""
    session.set_state(coords, nu)
    _, FI, DFI, _ = session.solve()
    _, sigma = session.internal_points(FI, DFI)
    s = torch.from_numpy(sigma)
    return float(von_mises_sq(s).sum())
