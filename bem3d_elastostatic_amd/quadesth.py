"""Host mirror of QUADESTH's hot-path routines on the MI355X library.

``Quadesth`` holds one device problem (a libquadesth handle) and exposes the
reference's phases under the reference's names, in the order main.f90:38-68
calls them:

    q = Quadesth(problem)                 # input_data's arrays (io_mod.f90:36-146)
    q.precompute_element_geometry()       # assembly_mod.f90:28-51
    q.ghmat8()                            # assembly_mod.f90:53-223
    X = q.solve_linear_system()           # solver_mod.f90:17-36
    FI, DFI = q.reorganize_solution(X)    # io_mod.f90:148-183
    res = q.compute_internal_points(FI, DFI)   # stress_mod.f90:511-781

Arrays use the reference layouts (see mesh.BemProblem).  Every compute call
goes through libquadesth.so; there is no host fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PRECISIONS, SolveReport, check
from .mesh import BemProblem


def _ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Quadesth:
    def __init__(self, problem: BemProblem, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.qd_create(C.byref(h), device)
        if rc != 0:
            msg = self.lib.qd_last_error(None)
            raise _lib.QuadesthError(rc, msg.decode() if msg else f"qd_create(device={device}) failed")
        self.h = h
        self.p = problem
        self.n, self.ne = problem.n, problem.ne
        self.N = 3 * self.n
        self._keep = [_f64(problem.x), _f64(problem.y), _f64(problem.z),
                      _i32(problem.nconec), _i32(problem.kconec)]
        x, y, z, nc, kc = self._keep
        check(h, self.lib.qd_set_mesh(h, self.n, self.ne, _ptr(x), _ptr(y), _ptr(z),
                                      _ptr(nc, C.c_int), _ptr(kc, C.c_int)), "set_mesh")
        check(h, self.lib.qd_set_material(h, float(problem.mu), float(problem.nu)), "set_material")
        self.set_bc(problem.kode, problem.bc)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.qd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle: int):
        check(self.h, self.lib.qd_set_stream(self.h, C.c_void_p(stream_handle)))

    def synchronize(self):
        check(self.h, self.lib.qd_synchronize(self.h))

    def profile_enable(self, kinds=()):
        mask = 0
        for k in kinds:
            mask |= 1 << _lib.PROF_KINDS[k]
        check(self.h, self.lib.qd_profile_enable(self.h, mask))

    def profile_read(self, kind):
        """(launches, device ms, algorithmic work) accumulated for a kernel
        class since the last read."""
        n = C.c_longlong()
        ms, work = C.c_double(), C.c_double()
        check(self.h, self.lib.qd_profile_read(self.h, _lib.PROF_KINDS[kind], C.byref(n), C.byref(ms),
                                               C.byref(work)))
        return n.value, ms.value, work.value

    def set_bc(self, kode, bc):
        self._kode = _i32(kode).reshape(-1)
        self._bc = _f64(bc).reshape(-1)
        check(self.h, self.lib.qd_set_bc(self.h, _ptr(self._kode, C.c_int), _ptr(self._bc)), "set_bc")

    # -- reference phases --------------------------------------------------
    def precompute_element_geometry(self, tables: bool = False):
        """assembly_mod.f90:28-51. With ``tables`` returns the all_* arrays
        as (ne, 36) (and ff as (ne, 36, 9))."""
        if not tables:
            check(self.h, self.lib.qd_precompute_element_geometry(self.h, *([None] * 8)))
            return None
        ne = self.ne
        out = {k: np.zeros(ne * 36) for k in ("eta1", "eta2", "eta3", "xja", "xg1", "xg2", "xg3")}
        out["ff"] = np.zeros(ne * 36 * 9)
        order = ("eta1", "eta2", "eta3", "xja", "xg1", "xg2", "xg3", "ff")
        check(self.h, self.lib.qd_precompute_element_geometry(self.h, *[_ptr(out[k]) for k in order]))
        return {k: (v.reshape(ne, 36, 9) if k == "ff" else v.reshape(ne, 36)) for k, v in out.items()}

    def set_assembly_precision(self, mode: str = "fp64"):
        """Arithmetic of the regular element integrals for the next ghmat8:
        "fp64" (the reference's), "far_fp32" (pairs beyond two element
        diameters in FP32, A itself) or "far_fp32_factor" (A in FP64; only the
        mixed-precision factorisations factor the FP32 far-field matrix)."""
        check(self.h, self.lib.qd_set_assembly_precision(self.h, _lib.ASSEMBLY[mode]), "set_assembly_precision")

    def ghmat8(self):
        """assembly_mod.f90:53-223 into device memory."""
        check(self.h, self.lib.qd_ghmat8(self.h), "ghmat8")

    def ghmat8_rhs(self, bc):
        """FI for another set of prescribed values with the same kode."""
        b = _f64(bc).reshape(-1)
        fi = np.zeros(self.N)
        check(self.h, self.lib.qd_ghmat8_rhs(self.h, _ptr(b), _ptr(fi)), "ghmat8_rhs")
        return fi

    def system(self, with_matrix: bool = True):
        """Host copies: A as an (N, N) array indexed A[row, col], and FI."""
        FI = np.zeros(self.N)
        if with_matrix:
            Acm = np.zeros((self.N, self.N))  # rows of this buffer are columns of A
            check(self.h, self.lib.qd_get_system(self.h, _ptr(Acm), self.N, _ptr(FI)))
            return Acm.T, FI
        check(self.h, self.lib.qd_get_system(self.h, None, self.N, _ptr(FI)))
        return None, FI

    def system_device(self):
        """(A device pointer, leading dimension, FI device pointer)."""
        pa, pf = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        lda = C.c_longlong()
        check(self.h, self.lib.qd_system_device(self.h, C.byref(pa), C.byref(lda), C.byref(pf)))
        return C.cast(pa, C.c_void_p).value, lda.value, C.cast(pf, C.c_void_p).value

    def solve_linear_system(self, rhs=None, precision: str = "fp64", tol: float = 0.0,
                            max_iter: int = 30):
        """solver_mod.f90:17-36. ``rhs`` None solves for the assembled FI;
        otherwise (N,) or (N, k).  Returns (X, report dict).  ``tol`` bounds
        the mixed precisions' refinement (normwise backward error); 0 selects
        the library's default, max(1e-14, 2^-53 sqrt(N)) (LAPACK DSGESV's)."""
        rep = SolveReport()
        prec = PRECISIONS[precision]
        if rhs is None:
            _, FI = self.system(with_matrix=False)
            rhs = FI
        b = np.asarray(rhs, dtype=np.float64)
        k = 1 if b.ndim == 1 else b.shape[1]
        bb = np.ascontiguousarray(b.reshape(self.N, k).T).copy()   # column-major (N, k)
        rc = self.lib.qd_solve(self.h, prec, k, _ptr(bb), self.N, tol, max_iter, C.byref(rep))
        check(self.h, rc, "solve")
        if rc > 0:
            raise _lib.QuadesthError(rc, f"singular matrix: U({rc},{rc}) = 0")
        X = bb.reshape(k, self.N).T
        return (X[:, 0].copy() if b.ndim == 1 else X.copy()), rep.as_dict()

    def reorganize_solution(self, X):
        """io_mod.f90:148-183 -> (FI (3n,), DFI (ne, 27))."""
        x = _f64(X)
        fi = np.zeros(self.N)
        dfi = np.zeros(27 * self.ne)
        check(self.h, self.lib.qd_reorganize_solution(self.h, _ptr(x), _ptr(fi), _ptr(dfi)))
        return fi, dfi.reshape(self.ne, 27)

    def compute_internal_points(self, FI, DFI, points=None, variant: int = 0):
        """stress_mod.f90:511-781 -> dict(u (l,3), sigma (l,6), prin (l,3), vm (l,),
        counts (far, near Gauss-point evaluations))."""
        pts = _f64(self.p.points if points is None else points)
        l = pts.shape[0]
        cx, cy, cz = _f64(pts[:, 0]), _f64(pts[:, 1]), _f64(pts[:, 2])
        u, s, pr, vm = np.zeros(3 * l), np.zeros(6 * l), np.zeros(3 * l), np.zeros(l)
        counts = np.zeros(2, dtype=np.int64)
        fi, dfi = _f64(FI), _f64(DFI).reshape(-1)
        check(self.h, self.lib.qd_internal_points(
            self.h, l, _ptr(cx), _ptr(cy), _ptr(cz), _ptr(fi), _ptr(dfi), variant, _ptr(u), _ptr(s),
            _ptr(pr), _ptr(vm), _ptr(counts, C.c_longlong)), "internal_points")
        return dict(u=u.reshape(l, 3), sigma=s.reshape(l, 6), prin=pr.reshape(l, 3), vm=vm,
                    counts=counts)

    def compute_boundary_stresses(self, FI, DFI):
        """stress_mod.f90:366-406 -> dict(sigma_elem (ne,9,6), sigma (n,6), prin (n,3), vm (n,))."""
        fi, dfi = _f64(FI), _f64(DFI).reshape(-1)
        se, sn = np.zeros(54 * self.ne), np.zeros(6 * self.n)
        pr, vm = np.zeros(3 * self.n), np.zeros(self.n)
        check(self.h, self.lib.qd_boundary_stresses(self.h, _ptr(fi), _ptr(dfi), _ptr(se), _ptr(sn),
                                                    _ptr(pr), _ptr(vm)), "boundary_stresses")
        return dict(sigma_elem=se.reshape(self.ne, 9, 6), sigma=sn.reshape(self.n, 6),
                    prin=pr.reshape(self.n, 3), vm=vm)

    def write_outputs(self, directory, FI, DFI, internal=None, boundary=None, ntrack=None,
                      fileout: str = "cube8_out.dat"):
        """The reference's result files in ``directory``, in main.f90:75-80's
        order: ``fileout`` (the deck echo of io_mod.f90:44-140, output_data,
        the boundary- and internal-stress tables), stresses.dat,
        internal_points.dat and disp_trac.dat.  ``internal`` is
        compute_internal_points' dict (u, sigma, prin, vm) for the problem's
        points, ``boundary`` compute_boundary_stresses' dict (computed here
        when None).  Returns the paths written."""
        import os
        os.makedirs(directory, exist_ok=True)
        p = self.p
        out = os.path.join(directory, fileout).encode()
        spath = os.path.join(directory, "stresses.dat").encode()
        ipath = os.path.join(directory, "internal_points.dat").encode()
        tpath = os.path.join(directory, "disp_trac.dat").encode()
        l = p.l if internal is not None else 0
        pts = _f64(p.points if l else np.zeros((1, 3)))
        cx, cy, cz = _f64(pts[:, 0]), _f64(pts[:, 1]), _f64(pts[:, 2])
        nt = int(p.ntrack if ntrack is None else ntrack)
        check(self.h, self.lib.qd_write_input_echo(self.h, out, p.title.encode(), l, _ptr(cx), _ptr(cy), _ptr(cz), nt),
              "write_input_echo")
        fi, dfi = _f64(FI), _f64(DFI).reshape(-1)
        u = _f64(internal["u"]).reshape(-1) if l else np.zeros(3)
        check(self.h, self.lib.qd_output_data(self.h, out, _ptr(fi), _ptr(dfi), l, _ptr(u)), "output_data")
        b = boundary if boundary is not None else self.compute_boundary_stresses(FI, DFI)
        check(self.h, self.lib.qd_output_boundary_stresses(
            self.h, out, spath, _ptr(_f64(b["sigma_elem"]).reshape(-1)), _ptr(_f64(b["sigma"]).reshape(-1)),
            _ptr(_f64(b["prin"]).reshape(-1)), _ptr(_f64(b["vm"]))), "output_boundary_stresses")
        paths = [out.decode(), spath.decode()]
        if l:
            check(self.h, self.lib.qd_output_internal_points(
                self.h, out, ipath, l, _ptr(cx), _ptr(cy), _ptr(cz), _ptr(u),
                _ptr(_f64(internal["sigma"]).reshape(-1)), _ptr(_f64(internal["prin"]).reshape(-1)),
                _ptr(_f64(internal["vm"]))), "output_internal_points")
            paths.append(ipath.decode())
        check(self.h, self.lib.qd_save_disp_trac(self.h, tpath, 0, nt, _ptr(fi), _ptr(dfi)), "save_disp_trac")
        paths.append(tpath.decode())
        return paths

    def run(self, precision: str = "fp64", variant: int = 0, points=None, tol: float = 0.0):
        """The whole hot path in main.f90's order; returns every result."""
        self.precompute_element_geometry()
        self.ghmat8()
        X, rep = self.solve_linear_system(precision=precision, tol=tol)
        FI, DFI = self.reorganize_solution(X)
        out = dict(X=X, FI=FI, DFI=DFI, report=rep)
        if (points is not None and len(points)) or self.p.l > 0:
            out.update(self.compute_internal_points(FI, DFI, points=points, variant=variant))
        return out


def dgesv(a, b):
    """LAPACK DGESV through the library's drop-in ``dgesv_`` (host arrays).
    a: (n, n) indexed a[row, col]; b: (n,) or (n, k).  Returns (LU in
    LAPACK column-major order as an (n, n) array [row, col], ipiv (1-based), x, info)."""
    lib = _lib.load()
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    acm = np.ascontiguousarray(a.T).copy()    # column-major buffer
    bb = np.asarray(b, dtype=np.float64)
    k = 1 if bb.ndim == 1 else bb.shape[1]
    bcm = np.ascontiguousarray(bb.reshape(n, k).T).copy()
    ipiv = np.zeros(n, dtype=np.int32)
    info = C.c_int()
    nn, kk = C.c_int(n), C.c_int(k)
    lda, ldb = C.c_int(n), C.c_int(n)
    lib.dgesv_(C.byref(nn), C.byref(kk), _ptr(acm), C.byref(lda), _ptr(ipiv, C.c_int), _ptr(bcm),
               C.byref(ldb), C.byref(info))
    x = bcm.reshape(k, n).T
    return acm.T, ipiv, (x[:, 0] if bb.ndim == 1 else x), info.value
