"""ctypes front end of the CPU checkers.  TEST INFRASTRUCTURE ONLY.

* ``Oracle``: the C restatement (quadesth_oracle.c), FP64, OpenMP.
* ``Reference``: the reference's own Fortran modules compiled by
  ``make -C oracle ref`` into oracle/_ref/libqref.so and driven through
  ref_driver.f90; the dense solve it would hand to LAPACK dgesv
  (solver_mod.f90:28) is done here with numpy's LAPACK.

Only tests, the smoke check and the benchmark's CPU-baseline leg import this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libquadesth_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libqref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class _Material(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("mu", "nu", "young", "c3", "c4", "c5", "c6")]


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def build_reference():
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    """FP64 CPU restatement; every array in the reference's own layout."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build_oracle()
        lib = C.CDLL(path)
        self.lib = lib
        lib.qo_material.argtypes = [C.c_double, C.c_double, C.POINTER(_Material)]
        lib.qo_precompute.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp] + [_dp] * 8
        lib.qo_adjacency.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip, _ip]
        lib.qo_set_ke_cap.argtypes = [C.c_int]
        lib.qo_ke_cap.restype = C.c_int
        lib.qo_ghmat8.argtypes = ([C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip, _ip, _dp, _ip, _ip,
                                   C.POINTER(_Material)] + [_dp] * 8 + [_dp, _dp])
        lib.qo_ghmat8_rows.argtypes = ([C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip, _ip, _dp, _ip, _ip,
                                        C.POINTER(_Material)] + [_dp] * 8 + [C.c_int, _ip, _dp, _dp])
        lib.qo_element_pair.argtypes = ([C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip,
                                         C.POINTER(_Material)] + [_dp] * 8 + [_dp, _dp])
        lib.qo_dgesv.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, _dp, C.c_int]
        lib.qo_dgesv.restype = C.c_int
        lib.qo_reorganize.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip, _dp, _ip, _ip, C.c_double,
                                      _dp, _dp]
        lib.qo_gauss_legendre.argtypes = [C.c_int, _dp, _dp]
        lib.qo_principal_vm.argtypes = [_dp, _dp, _dp]
        lib.qo_funkelvin.argtypes = [C.POINTER(_Material), C.c_double, C.c_double, C.c_double, _dp,
                                     C.c_double, C.c_double, C.c_double, _dp, _dp]
        lib.qo_shape.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        lib.qo_internal_points.argtypes = ([C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip,
                                            C.POINTER(_Material)] + [_dp] * 8 +
                                           [_dp, _dp, C.c_int, _dp, _dp, _dp, C.c_int,
                                            _dp, _dp, _dp, _dp,
                                            np.ctypeslib.ndpointer(dtype=np.int64)])
        lib.qo_boundary_stresses.argtypes = ([C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip, _ip, _ip,
                                              C.POINTER(_Material), _dp, _dp, _dp, _dp, _dp, _dp])

    # -- small pieces -----------------------------------------------------
    def material(self, mu, nu):
        m = _Material()
        self.lib.qo_material(mu, nu, C.byref(m))
        return m

    def shape(self, ncone, g1, g2):
        ff, fd1, fd2 = np.zeros(9), np.zeros(9), np.zeros(9)
        self.lib.qo_shape(ncone, g1, g2, ff, fd1, fd2)
        return ff, fd1, fd2

    def funkelvin(self, mu, nu, xp, eta, xg):
        m = self.material(mu, nu)
        u, t = np.zeros(9), np.zeros(9)
        self.lib.qo_funkelvin(C.byref(m), xp[0], xp[1], xp[2], _f64(eta), xg[0], xg[1], xg[2], u, t)
        return u.reshape(3, 3).T.copy(), t.reshape(3, 3).T.copy()   # [jj, ii]

    def gauss_legendre(self, npts):
        x, w = np.zeros(npts), np.zeros(npts)
        self.lib.qo_gauss_legendre(npts, x, w)
        return x, w

    def principal_vm(self, sig):
        prin, vm = np.zeros(3), np.zeros(1)
        self.lib.qo_principal_vm(_f64(sig), prin, vm)
        return prin, vm[0]

    # -- phases -----------------------------------------------------------
    def geometry(self, p):
        ne = p.ne
        t = {k: np.zeros(ne * 36) for k in ("eta1", "eta2", "eta3", "xja", "xg1", "xg2", "xg3")}
        t["ff"] = np.zeros(ne * 36 * 9)
        self.lib.qo_precompute(ne, _i32(p.nconec), _i32(p.kconec), _f64(p.x), _f64(p.y), _f64(p.z),
                               t["eta1"], t["eta2"], t["eta3"], t["xja"], t["xg1"], t["xg2"],
                               t["xg3"], t["ff"])
        return t

    def _tabs(self, t):
        return [t[k] for k in ("eta1", "eta2", "eta3", "xja", "xg1", "xg2", "xg3", "ff")]

    def adjacency(self, p):
        """nesim and ke.  ke is as wide as the most-shared node needs (at least
        the reference's 8), so every element of every node is listed; the
        width is set in the C oracle for the calls that follow."""
        counts = np.bincount(np.concatenate([p.kconec[e, :int(p.nconec[e])] for e in range(p.ne)]) - 1,
                             minlength=p.n)
        cap = max(8, int(counts.max()))
        self.lib.qo_set_ke_cap(cap)
        nesim = np.zeros(p.n, dtype=np.int32)
        ke = np.zeros(p.n * cap, dtype=np.int32)
        self.lib.qo_adjacency(p.n, p.ne, _i32(p.nconec), _i32(p.kconec), nesim, ke)
        return nesim, ke

    def element_pair(self, p, t, node, elem):
        h, g = np.zeros(81), np.zeros(81)
        m = self.material(p.mu, p.nu)
        self.lib.qo_element_pair(node, elem, _f64(p.x), _f64(p.y), _f64(p.z), _i32(p.nconec),
                                 _i32(p.kconec), C.byref(m), *self._tabs(t), h, g)
        return h.reshape(27, 3).T.copy(), g.reshape(27, 3).T.copy()

    def ghmat8(self, p, t=None, bc=None):
        """Returns (A column-major as an (N, N) array indexed A[row, col], FI)."""
        if t is None:
            t = self.geometry(p)
        nesim, ke = self.adjacency(p)
        N = 3 * p.n
        A = np.zeros(N * N)
        FI = np.zeros(N)
        m = self.material(p.mu, p.nu)
        self.lib.qo_ghmat8(p.n, p.ne, _f64(p.x), _f64(p.y), _f64(p.z), _i32(p.nconec),
                           _i32(p.kconec), _i32(p.kode), _f64(p.bc if bc is None else bc),
                           nesim, ke, C.byref(m), *self._tabs(t), A, FI)
        return A.reshape(N, N).T, FI

    def ghmat8_rows(self, p, nodes, t=None):
        """Rows of A (as A[row, :]) and FI for the listed collocation nodes
        (0-based): returns (rows index array, A_rows (3k, N), FI_rows (3k,))."""
        if t is None:
            t = self.geometry(p)
        nesim, ke = self.adjacency(p)
        nodes = _i32(nodes)
        k = nodes.shape[0]
        N = 3 * p.n
        Ar = np.zeros(3 * k * N)
        Fr = np.zeros(3 * k)
        m = self.material(p.mu, p.nu)
        self.lib.qo_ghmat8_rows(p.n, p.ne, _f64(p.x), _f64(p.y), _f64(p.z), _i32(p.nconec),
                                _i32(p.kconec), _i32(p.kode), _f64(p.bc), nesim, ke, C.byref(m),
                                *self._tabs(t), k, nodes, Ar, Fr)
        rows = (3 * nodes[:, None] + np.arange(3)[None, :]).reshape(-1)
        return rows, Ar.reshape(3 * k, N), Fr

    def dgesv(self, A, b):
        """A[row, col] (any layout), b (N,) or (N, k). Returns (LU col-major, ipiv, x, info)."""
        N = A.shape[0]
        a = np.ascontiguousarray(np.asarray(A, dtype=np.float64).T).reshape(-1).copy()
        bb = np.asarray(b, dtype=np.float64)
        nrhs = 1 if bb.ndim == 1 else bb.shape[1]
        bcol = np.ascontiguousarray(bb.reshape(N, nrhs).T).reshape(-1).copy()
        ipiv = np.zeros(N, dtype=np.int32)
        info = self.lib.qo_dgesv(N, nrhs, a, N, ipiv, bcol, N)
        x = bcol.reshape(nrhs, N).T
        return a.reshape(N, N).T, ipiv, (x[:, 0] if bb.ndim == 1 else x), info

    def reorganize(self, p, X, bc=None):
        nesim, ke = self.adjacency(p)
        FI = _f64(X).copy()
        DFI = np.zeros(p.ne * 27)
        self.lib.qo_reorganize(p.n, p.ne, _i32(p.nconec), _i32(p.kconec), _i32(p.kode),
                               _f64(p.bc if bc is None else bc), nesim, ke, p.mu, FI, DFI)
        return FI, DFI.reshape(p.ne, 27)

    def internal_points(self, p, t, FI, DFI, points=None, variant=0):
        pts = p.points if points is None else points
        pts = _f64(pts)
        l = pts.shape[0]
        u = np.zeros(3 * l)
        s = np.zeros(6 * l)
        pr = np.zeros(3 * l)
        vm = np.zeros(l)
        counts = np.zeros(2, dtype=np.int64)
        m = self.material(p.mu, p.nu)
        self.lib.qo_internal_points(p.n, p.ne, _f64(p.x), _f64(p.y), _f64(p.z), _i32(p.nconec),
                                    _i32(p.kconec), C.byref(m), *self._tabs(t), _f64(FI),
                                    _f64(DFI).reshape(-1), l, _f64(pts[:, 0]), _f64(pts[:, 1]),
                                    _f64(pts[:, 2]), variant, u, s, pr, vm, counts)
        return dict(u=u.reshape(l, 3), sigma=s.reshape(l, 6), prin=pr.reshape(l, 3), vm=vm,
                    counts=counts)

    def boundary_stresses(self, p, FI, DFI):
        nesim, ke = self.adjacency(p)
        se = np.zeros(p.ne * 54)
        sn = np.zeros(p.n * 6)
        pr = np.zeros(p.n * 3)
        vm = np.zeros(p.n)
        m = self.material(p.mu, p.nu)
        self.lib.qo_boundary_stresses(p.n, p.ne, _f64(p.x), _f64(p.y), _f64(p.z), _i32(p.nconec),
                                      _i32(p.kconec), nesim, ke, C.byref(m), _f64(FI),
                                      _f64(DFI).reshape(-1), se, sn, pr, vm)
        return dict(sigma_elem=se.reshape(p.ne, 9, 6), sigma=sn.reshape(p.n, 6),
                    prin=pr.reshape(p.n, 3), vm=vm)

    def run(self, p, variant=0, solver="oracle"):
        """Whole FP64 pipeline, main.f90:38-68 order."""
        t = self.geometry(p)
        A, FI0 = self.ghmat8(p, t)
        if solver == "oracle":
            _, _, X, info = self.dgesv(A, FI0)
            assert info == 0, info
        else:
            X = np.linalg.solve(A, FI0)
        FI, DFI = self.reorganize(p, X)
        out = dict(A=A, FI0=FI0, X=X, FI=FI, DFI=DFI, tables=t)
        if p.l > 0:
            out.update(self.internal_points(p, t, FI, DFI, variant=variant))
        return out


class Reference:
    """The reference's compiled modules (oracle/_ref/libqref.so)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        self.lib = lib
        lib.qref_load.argtypes = [C.c_char_p, C.c_char_p]
        lib.qref_load.restype = C.c_int
        ip = C.POINTER(C.c_int)
        dpp = C.POINTER(C.c_double)
        lib.qref_sizes.argtypes = [ip, ip, ip, dpp, dpp]
        lib.qref_assemble.argtypes = [_dp, _dp]
        lib.qref_reorganize.argtypes = [_dp, _dp, _dp]
        lib.qref_internal.argtypes = [_dp, _dp, _dp, _dp]
        lib.qref_boundary.argtypes = [_dp, _dp, _dp]
        lib.qref_outputs.argtypes = [C.c_char_p]

    def run(self, deck_path: str, solve=None, boundary=False, outputs=False):
        """Run the reference on a deck file.  ``solve(A, b) -> x`` defaults to
        numpy's LAPACK gesv.  With ``outputs`` (implies ``boundary``) the
        reference's writers run too and their files come back as
        ``out["files"]``: name -> bytes."""
        boundary = boundary or outputs
        td_out = tempfile.mkdtemp(dir="/tmp") if outputs else None
        with tempfile.TemporaryDirectory(dir="/tmp") as td:
            short = os.path.join(td, "d.dat")
            with open(deck_path, "rb") as src, open(short, "wb") as dst:
                dst.write(src.read())
            outp = os.path.join(td_out or td, "cube8_out.dat")
            self.lib.qref_load(short.encode(), outp.encode())
        n, ne, l = C.c_int(), C.c_int(), C.c_int()
        mu, nu = C.c_double(), C.c_double()
        self.lib.qref_sizes(C.byref(n), C.byref(ne), C.byref(l), C.byref(mu), C.byref(nu))
        n, ne, l = n.value, ne.value, l.value
        N = 3 * n
        A = np.zeros(N * N)
        FI0 = np.zeros(N)
        self.lib.qref_assemble(A, FI0)
        A = A.reshape(N, N).T            # A[row, col]
        X = (solve or np.linalg.solve)(A, FI0)
        FI = np.zeros(N)
        DFI = np.zeros(27 * ne)
        self.lib.qref_reorganize(_f64(X), FI, DFI)
        out = dict(A=A, FI0=FI0, X=X, FI=FI, DFI=DFI.reshape(ne, 27), n=n, ne=ne, l=l,
                   mu=mu.value, nu=nu.value)
        if l > 0:
            u, s, pr, vm = np.zeros(3 * l), np.zeros(6 * l), np.zeros(3 * l), np.zeros(l)
            self.lib.qref_internal(u, s, pr, vm)
            out.update(u=u.reshape(l, 3), sigma=s.reshape(l, 6), prin=pr.reshape(l, 3), vm=vm)
        if boundary:
            s, pr, vm = np.zeros(6 * n), np.zeros(3 * n), np.zeros(n)
            self.lib.qref_boundary(s, pr, vm)
            out.update(bsigma=s.reshape(n, 6), bprin=pr.reshape(n, 3), bvm=vm)
        if outputs:
            import shutil
            self.lib.qref_outputs(td_out.encode())
            files = {}
            for name in ("cube8_out.dat", "stresses.dat", "internal_points.dat", "disp_trac.dat"):
                path = os.path.join(td_out, name)
                if os.path.exists(path):
                    with open(path, "rb") as fh:
                        files[name] = fh.read()
            shutil.rmtree(td_out)
            out["files"] = files
        return out
