"""Benchmark of the QUADESTH hot path on one MI355X.

One step = the whole hot path over the synthetic thick-cantilever box
(10 x 2 x 1, h = 0.1: 25,602 nodes, 2,400 quad9 + 8,000 tri6, 76,806 DOF,
65,536 internal points): ghmat8 assembly, the mixed-precision solve (FP16
MFMA trailing updates, FP64 refinement to LAPACK DSGESV's backward-error bound
2^-53 sqrt(N)), reorganisation, and the
Somigliana internal-point integrals.  Mesh, codes and points are resident in
HBM before the timed region.  After the timed steps the same GPU's FP64 LU
path is timed once on the same system, for the speed-up and for the error of
the mixed path against it.

Prints ONE JSON line (rank 0).  Usage:
    python bench.py [--steps K] [--warmup W] [--precision fp16|fp32|fp64]
                    [--config c3|c1|c4] [--no-fp64] [--no-cpu] [--fd-dirs D]
--config c4 measures the adjoint gradient on the perturbed cube instead
(forward and backward times, peak device memory, finite-difference check).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s/solve @76.8k DOF: mixed-IR vs FP64 LU @1e-10 rel; H,G entries/s; LU MFMA %"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "fp16": 2500.0, "fp64_vector": 78.6}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def hg_entries(p):
    """H,G entries of the reference's assembly: every (collocation node,
    element) pair yields a 3 x 3*nconec block of H and of G."""
    return float(p.n) * float(np.sum(18 * p.nconec.astype(np.float64)))


class Pipeline:
    def __init__(self, p, precision, torch):
        from bem3d_elastostatic_amd.quadesth import Quadesth
        self.torch = torch
        self.p = p
        self.q = Quadesth(p)
        self.prec = precision
        dev = torch.device("cuda:0")
        self.pts = torch.tensor(np.ascontiguousarray(p.points), dtype=torch.float64, device=dev)
        l = p.l
        self.u = torch.empty(3 * l, dtype=torch.float64, device=dev)
        self.s = torch.empty(6 * l, dtype=torch.float64, device=dev)
        self.pr = torch.empty(3 * l, dtype=torch.float64, device=dev)
        self.vm = torch.empty(l, dtype=torch.float64, device=dev)
        self.b = torch.empty(3 * p.n, dtype=torch.float64, device=dev)
        self.fi = torch.empty(3 * p.n, dtype=torch.float64, device=dev)
        self.dfi = torch.empty(27 * p.ne, dtype=torch.float64, device=dev)
        self.q.precompute_element_geometry()
        torch.cuda.synchronize()

    def step(self, precision=None, phases=None):
        from bem3d_elastostatic_amd import _lib
        torch, q = self.torch, self.q
        prec = precision or self.prec
        t0 = time.perf_counter()
        q.ghmat8()
        q.synchronize()
        t1 = time.perf_counter()
        _, _, fi_ptr = q.system_device()
        N = 3 * self.p.n
        rc = _hip().hipMemcpy(C.c_void_p(self.b.data_ptr()), C.c_void_p(fi_ptr), C.c_size_t(8 * N), 3)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy (device to device) failed: {rc}")
        rep = _lib.SolveReport()
        # tol 0: the library's default bound, max(1e-14, 2^-53 sqrt(N)) (LAPACK DSGESV's)
        rc = q.lib.qd_solve_device(q.h, _lib.PRECISIONS[prec], 1, C.c_void_p(self.b.data_ptr()), N, 0.0, 30,
                                   C.byref(rep))
        _lib.check(q.h, rc, "solve")
        t2 = time.perf_counter()
        X = self.b.cpu().numpy()
        FI, DFI = q.reorganize_solution(X)
        self.fi.copy_(torch.from_numpy(FI))
        self.dfi.copy_(torch.from_numpy(DFI.reshape(-1)))
        t3 = time.perf_counter()
        counts = np.zeros(2, dtype=np.int64)
        rc = q.lib.qd_internal_points_device(
            q.h, self.p.l, C.c_void_p(self.pts.data_ptr()), C.c_void_p(self.fi.data_ptr()),
            C.c_void_p(self.dfi.data_ptr()), 0, C.c_void_p(self.u.data_ptr()), C.c_void_p(self.s.data_ptr()),
            C.c_void_p(self.pr.data_ptr()), C.c_void_p(self.vm.data_ptr()),
            counts.ctypes.data_as(C.POINTER(C.c_longlong)))
        _lib.check(q.h, rc, "internal points")
        q.synchronize()
        t4 = time.perf_counter()
        if phases is not None:
            phases.append(dict(assembly=t1 - t0, solve=t2 - t1, reorganize=t3 - t2, internal=t4 - t3,
                               factor_ms=rep.t_factor_ms, refine_ms=rep.t_solve_ms, iterations=rep.iterations,
                               gmres=rep.gmres, residual=rep.residual, pivoting=rep.pivoting,
                               counts=counts.tolist()))
        return FI, DFI

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipMemcpy.restype = C.c_int
    return _HIP


def host_description():
    """CPU model, logical CPUs this process may run on, physical cores among
    them, and the BLAS/LAPACK numpy calls."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        try:
            pkg = open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id").read().strip()
            core = open(f"/sys/devices/system/cpu/cpu{c}/topology/core_id").read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    blas = "unknown"
    try:
        from threadpoolctl import threadpool_info
        info = [d for d in threadpool_info() if d.get("user_api") == "blas"]
        if info:
            blas = "%s %s (%s threads)" % (info[0].get("internal_api"), info[0].get("version"),
                                            info[0].get("num_threads"))
    except Exception:
        pass
    return dict(cpu_model=model, logical_cpus=len(cpus), physical_cores=len(cores), blas=blas)


def cpu_baseline(p, FI, DFI, budget_s=25.0):
    """The oracle (a C/OpenMP restatement of the reference's loops) and numpy
    LAPACK timed on bounded samples of the same workload, extrapolated to one
    whole solve."""
    from oracle.pyoracle import Oracle
    o = Oracle()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    host = host_description()
    t = o.geometry(p)
    rng = np.random.default_rng(0)
    # assembly: sampled collocation nodes, all elements
    k = min(p.n, 1024)
    nodes = np.sort(rng.choice(p.n, k, replace=False))
    t0 = time.perf_counter()
    o.ghmat8_rows(p, nodes, t)
    ta = time.perf_counter() - t0
    per_node = ta / k
    t_asm = per_node * p.n
    # LU: numpy (LAPACK gesv) at a size that runs in seconds, scaled by N^3
    ns = 8000
    A = rng.normal(size=(ns, ns))
    bb = rng.normal(size=ns)
    t0 = time.perf_counter()
    np.linalg.solve(A, bb)
    tl = time.perf_counter() - t0
    N = 3 * p.n
    t_lu = tl * (N / ns) ** 3
    # Somigliana: two sample sizes, slope per point (the near table is built per call)
    s1, s2 = 128, 1024
    pts = p.points
    t0 = time.perf_counter()
    o.internal_points(p, t, FI, DFI, points=pts[:s1])
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.internal_points(p, t, FI, DFI, points=pts[:s2])
    t2 = time.perf_counter() - t0
    slope = max(t2 - t1, 1e-9) / (s2 - s1)
    fixed = max(t1 - slope * s1, 0.0)
    t_som = fixed + slope * p.l
    total = t_asm + t_lu + t_som
    return dict(value=round(total, 3), unit="s/solve", cores=threads, kind="port", host=host,
                sample=(f"oracle ghmat8 rows of {k} of {p.n} collocation nodes ({ta:.2f} s), numpy LAPACK "
                        f"gesv N={ns} ({tl:.2f} s) scaled by (N/{ns})^3, oracle Somigliana at {s1} and {s2} of "
                        f"{p.l} points ({t1:.2f}+{t2:.2f} s); extrapolated phases s: assembly {t_asm:.1f}, "
                        f"LU {t_lu:.1f}, internal points {t_som:.1f}"),
                phases=dict(assembly=t_asm, lu=t_lu, internal=t_som))


HBM_PEAK_GBS = 8000.0


def trailing_shapes(N, precision, diagonal=True):
    """(M, Nn, K) of every outer trailing-update launch of one factorisation:
    the library's panel widths (FP16: 4096 while more than 30,000 columns
    remain, then 1024; else 512; QUADESTH_NB, QUADESTH_NB_WIDE and
    QUADESTH_NB_WIDE_REM override) and its look-ahead split (the next panel's columns, then the
    rest).  With diagonal pivots the FP16 update of the rest is two launches,
    the next panel's rows first."""
    nb = int(os.environ.get("QUADESTH_NB", "1024" if precision == "fp16" else "512"))
    nb = max(64, nb // 32 * 32)
    wide = rem = 0
    if precision == "fp16":
        wide = max(0, int(os.environ.get("QUADESTH_NB_WIDE", "4096")) // 32 * 32)
        rem = max(0, int(os.environ.get("QUADESTH_NB_WIDE_REM", "30000")))
    tail = tail_rem = 0
    if precision == "fp16":
        tail = max(0, int(os.environ.get("QUADESTH_NB_TAIL", "0")) // 32 * 32)
        tail_rem = max(0, int(os.environ.get("QUADESTH_NB_TAIL_REM", "0")))

    def width(j):
        left = N - j
        if wide > nb and left > rem:
            return min(wide, left)
        return min(tail if tail > 0 and left <= tail_rem else nb, left)

    look = os.environ.get("QUADESTH_LOOKAHEAD", "1")[:1] != "0"
    split = look and diagonal and precision == "fp16"
    out = []
    j = 0
    while j < N:
        jb = width(j)
        nj = j + jb
        j = nj
        if nj >= N:
            break
        njb = width(nj)
        if look and nj + njb < N:
            out.append((N - nj, njb, jb))
            if split:
                out.append((njb, N - nj - njb, jb))
                out.append((N - nj - njb, N - nj - njb, jb))
            else:
                out.append((N - nj, N - nj - njb, jb))
        else:
            out.append((N - nj, N - nj, jb))
    return out


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "lu_trailing_gemm16_traffic.txt")


def measured_traffic(N, precision):
    """HBM bytes per trailing-update launch from the committed counter passes
    (FETCH_SIZE and WRITE_SIZE over the replayed launches, each scaled by the
    same counter's reading for a copy of known size), for the workload they
    were taken on (FP16 update, 76,806 DOF); None otherwise."""
    if precision != "fp16" or N != 76806 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) == 2 and f[0] == "traffic_bytes_per_launch":
            return float(f[1])
    return None


def trailing_roofline(N, precision, d, diagonal=True):
    """Roofline of the outer trailing update from its algorithmic work: 2 M Nn K
    flops and, in bytes, the C read and write (8 per FP32 element) plus one
    pass over the operands in the precision the kernel reads them.  The bound
    is whichever floor (peak MFMA rate or 8 TB/s) is higher for these shapes."""
    shapes = trailing_shapes(N, precision, diagonal)
    ob = {"fp16": 2, "fp32": 4, "fp64": 8}[precision]   # operand bytes
    cb = {"fp16": 8, "fp32": 8, "fp64": 16}[precision]  # C read + write bytes per element
    flops = sum(2.0 * m * n * k for m, n, k in shapes)
    nbytes = sum(cb * m * n + ob * k * (m + n) for m, n, k in shapes)
    runs = max(1, round(d["launches"] / max(len(shapes), 1)))
    if abs(d["work"] - runs * flops) > 0.01 * d["work"]:
        log(f"trailing-update shapes do not match the profiled work: {runs * flops:.4e} vs {d['work']:.4e}")
    t = d["ms"] * 1e-3 / runs   # device seconds of one factorisation's launches
    peak_f = PEAK_TFLOPS[precision]
    t_mfma = flops / (peak_f * 1e12)
    t_hbm = nbytes / (HBM_PEAK_GBS * 1e9)
    common = dict(kernel="lu_trailing_gemm", traffic=measured_traffic(N, precision), launches=d["launches"],
                  avg_launch_ms=d["ms"] / max(d["launches"], 1), flops_per_launch=flops / len(shapes),
                  bytes_per_launch=nbytes / len(shapes), floor_s=dict(mfma=t_mfma, hbm=t_hbm))
    if t_hbm >= t_mfma:
        ach = nbytes / t / 1e9
        return dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), mfma_frac=round(flops / t / 1e12 / peak_f, 4), **common)
    ach = flops / t / 1e12
    return dict(bound="mfma", achieved=round(ach, 3), peak=peak_f, unit="TFLOP/s", frac=round(ach / peak_f, 4),
                hbm_frac=round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4), **common)


def bench_trailing_update():
    """The outer trailing updates of one FP16 factorisation at 76,806 DOF, in
    its launch order and shapes, through qd_fp16_update on synthetic operands
    (one FP32 C of the largest shape, 23 GB).  Used under rocprofv3 counter
    passes, whose launch interception does not survive the full solve."""
    import torch
    from bem3d_elastostatic_amd import _lib
    N = 76806
    shapes = trailing_shapes(N, "fp16")
    K = max(k for _, _, k in shapes)
    mx = max(m for m, _, _ in shapes)
    nx = max(n for _, n, _ in shapes)
    dev = torch.device("cuda:0")
    a16 = (0.1 * torch.randn(mx, K, device=dev)).half()
    b16 = (0.1 * torch.randn(nx, K, device=dev)).half()
    sc = torch.ones((nx + 63) // 64, device=dev)
    ldc = (nx + 63) // 64 * 64    # the factorisation's row pitch rule
    c = torch.zeros(mx, ldc, device=dev)
    lib = _lib.load()
    h = C.c_void_p()
    _lib.check(None, lib.qd_create(C.byref(h), 0), "create")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m, n, k in shapes:
        _lib.check(h, lib.qd_fp16_update(h, m, n, k, C.c_void_p(a16.data_ptr()), C.c_void_p(b16.data_ptr()),
                                         C.c_void_p(sc.data_ptr()), C.c_void_p(c.data_ptr()), ldc), "fp16_update")
    _lib.check(h, lib.qd_synchronize(h), "sync")
    t = time.perf_counter() - t0
    flops = sum(2.0 * m * n * k for m, n, k in shapes)
    print(json.dumps({"mode": "trailing-update-only", "launches": len(shapes), "seconds": round(t, 4),
                      "tflops": round(flops / t / 1e12, 1)}), flush=True)
    lib.qd_destroy(h)


def bench_gradient(args):
    """Adjoint gradient of J = sum of internal von Mises stress squared on the
    perturbed 16x16-per-face cube (18,438 DOF, 4096 points), FP64 forward:
    forward and backward times, peak device memory of the library, and the
    agreement with central differences (h = 1e-6) along seeded directions."""
    import torch
    from bem3d_elastostatic_amd import _lib, mesh
    from bem3d_elastostatic_amd.autograd import BoundarySolve, GradSession, InternalPoints, von_mises_sq, value_vm2

    p = mesh.config4()
    s = GradSession(p)
    lib = _lib.load()
    xyz = p.coords()
    log(f"gradient workload: n={p.n} ne={p.ne} N={3 * p.n} l={p.l}")

    def run_once():
        s._state = None                       # force the forward to recompute everything
        c = torch.tensor(xyz, requires_grad=True)
        v = torch.tensor(float(p.nu), dtype=torch.float64, requires_grad=True)
        t0 = time.perf_counter()
        FI, DFI = BoundarySolve.apply(c, v, s)
        _, sigma = InternalPoints.apply(c, v, FI, DFI, s)
        J = von_mises_sq(sigma).sum()
        t1 = time.perf_counter()
        J.backward()
        t2 = time.perf_counter()
        return float(J.detach()), c.grad.numpy().copy(), float(v.grad), t1 - t0, t2 - t1

    for _ in range(max(args.warmup, 1)):
        run_once()
    cur, peak = C.c_longlong(), C.c_longlong()
    lib.qd_memory_usage(C.byref(cur), C.byref(peak), 1)
    fw, bw = [], []
    for _ in range(args.steps):
        J, gc, gnu, tf, tb = run_once()
        fw.append(tf)
        bw.append(tb)
    lib.qd_memory_usage(C.byref(cur), C.byref(peak), 0)
    log(f"J={J:.12e} forward {np.mean(fw):.3f} s backward {np.mean(bw):.3f} s peak {peak.value / 2**30:.2f} GiB")

    rng = np.random.default_rng(64)
    h = 1e-6

    def central(d, dnu, hh):
        return (value_vm2(s, xyz + hh * d, p.nu + hh * dnu) - value_vm2(s, xyz - hh * d, p.nu - hh * dnu)) / (2 * hh)

    errs, abs_errs, dirs = [], [], []
    for i in range(args.fd_dirs):
        d = rng.standard_normal((p.n, 3))
        dnu = 0.1 * rng.standard_normal()
        an = float(np.sum(gc * d) + gnu * dnu)
        fd = central(d, dnu, h)
        errs.append(abs(an - fd) / max(abs(an), abs(fd), 1e-300))
        abs_errs.append(abs(an - fd))
        dirs.append((d, dnu, an))
        if i % 8 == 0:
            log(f"fd direction {i}: adjoint {an:.10e} fd {fd:.10e} rel {errs[-1]:.2e}")
    # The central difference carries J's rounding noise / h and a truncation
    # term ~h^2.  For the worst direction, larger steps and Richardson's
    # combination show which of the two the gap is; a gradient error would
    # survive both.
    iw = int(np.argmax(errs))
    d, dnu, an = dirs[iw]
    worst = {"direction": iw, "adjoint": an, "rel_err_h": {}, "richardson_rel_err_h": {}}
    cd = {}
    for hh in (1e-6, 1e-5, 2e-5, 1e-4, 2e-4):
        cd[hh] = central(d, dnu, hh)
    for hh in (1e-6, 1e-5, 1e-4):
        worst["rel_err_h"][f"{hh:g}"] = abs(an - cd[hh]) / max(abs(an), abs(cd[hh]), 1e-300)
    # Richardson: (4 D(h) - D(2h)) / 3 removes the h^2 term of the central difference
    for hh in (1e-5, 1e-4):
        r = (4.0 * cd[hh] - cd[2 * hh]) / 3.0
        worst["richardson_rel_err_h"][f"{hh:g}"] = abs(an - r) / max(abs(an), abs(r), 1e-300)
    log(f"worst direction {iw}: {worst}")
    out = {
        "metric": "adjoint gradient of sum(vm^2) w.r.t. 18438 coordinates + nu: backward/forward time",
        "value": round(float(np.mean(bw)) / float(np.mean(fw)), 3), "unit": "ratio",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
        "forward_s": round(float(np.mean(fw)), 4), "backward_s": round(float(np.mean(bw)), 4),
        "peak_library_hbm_bytes": int(peak.value), "J": J,
        "fd_check": {"directions": args.fd_dirs, "h": h, "max_rel_err": float(np.max(errs)),
                     "median_rel_err": float(np.median(errs)), "max_abs_err": float(np.max(abs_errs)),
                     "grad_norm": float(np.sqrt(np.sum(gc * gc) + gnu * gnu)), "worst": worst},
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "unit cube 16x16 quad9 per face, radial bump 0.02 (seed 3): 6146 nodes, "
                               "18438 DOF, 4096 internal points; J = sum vm^2", "precision": "fp64"},
    }
    print(json.dumps(out), flush=True)
    s.close()


def _c3_handle(args):
    import torch
    from bem3d_elastostatic_amd import mesh
    from bem3d_elastostatic_amd.quadesth import Quadesth
    p = mesh.config3() if args.config == "c3" else mesh.config1()
    q = Quadesth(p)
    if args.assembly != "fp64":
        q.set_assembly_precision(args.assembly)
    q.precompute_element_geometry()
    torch.cuda.synchronize()
    return p, q


def bench_assembly_only(args):
    """ghmat8 alone, --steps times after --warmup (no solve): the launches a
    counter pass of the assembly kernels needs."""
    p, q = _c3_handle(args)
    for _ in range(args.warmup):
        q.ghmat8()
        q.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        q.ghmat8()
        q.synchronize()
        log(f"assembly {s} done")
    t = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"mode": "assembly-only", "config": args.config, "assembly": args.assembly,
                      "s_per_assembly": round(t, 5), "hg_entries_per_s": hg_entries(p) / t}), flush=True)


def bench_internal_only(args):
    """compute_internal_points alone at the configuration's points with a
    seeded smooth synthetic boundary field (the work does not depend on the
    values), --steps times after --warmup."""
    import torch
    from bem3d_elastostatic_amd import _lib
    p, q = _c3_handle(args)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(4)
    fi = torch.tensor(1e-3 * np.sin(np.outer(np.arange(3 * p.n) % 97, [1.0]).ravel() + rng.normal(size=3 * p.n)),
                      dtype=torch.float64, device=dev)
    dfi = torch.tensor(rng.normal(size=27 * p.ne), dtype=torch.float64, device=dev)
    pts = torch.tensor(np.ascontiguousarray(p.points), dtype=torch.float64, device=dev)
    out = torch.empty(13 * p.l, dtype=torch.float64, device=dev)
    base = out.data_ptr()
    counts = np.zeros(2, dtype=np.int64)

    def once():
        rc = q.lib.qd_internal_points_device(q.h, p.l, C.c_void_p(pts.data_ptr()), C.c_void_p(fi.data_ptr()),
                                             C.c_void_p(dfi.data_ptr()), 0, C.c_void_p(base),
                                             C.c_void_p(base + 8 * 3 * p.l), C.c_void_p(base + 8 * 9 * p.l),
                                             C.c_void_p(base + 8 * 12 * p.l),
                                             counts.ctypes.data_as(C.POINTER(C.c_longlong)))
        _lib.check(q.h, rc, "internal points")
        q.synchronize()
    for _ in range(args.warmup):
        once()
    t0 = time.perf_counter()
    for s in range(args.steps):
        once()
        log(f"internal points {s} done")
    t = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"mode": "internal-only", "config": args.config, "s_per_call": round(t, 5),
                      "gp_evals": {"far": int(counts[0]), "near": int(counts[1])}}), flush=True)


def bench_cpu_full(args):
    """The oracle's assembly (every collocation row, in chunks) and Somigliana
    (every point, in chunks) of the 76,806-DOF cantilever on the host cores,
    timed in full; the LU at N=8000 with numpy's LAPACK, scaled by N^3."""
    from bem3d_elastostatic_amd import mesh
    from oracle.pyoracle import Oracle
    p = mesh.config3()
    o = Oracle()
    host = host_description()
    log(f"host {host}")
    t = o.geometry(p)
    t0 = time.perf_counter()
    for c0 in range(0, p.n, 2048):
        o.ghmat8_rows(p, np.arange(c0, min(p.n, c0 + 2048)), t)
        log(f"assembly rows of nodes {c0}..{min(p.n, c0 + 2048)} done, {time.perf_counter() - t0:.1f} s")
    t_asm = time.perf_counter() - t0
    rng = np.random.default_rng(4)
    FI = 1e-3 * rng.normal(size=3 * p.n)
    DFI = rng.normal(size=(p.ne, 27))
    t0 = time.perf_counter()
    for c0 in range(0, p.l, 8192):
        o.internal_points(p, t, FI, DFI, points=p.points[c0:c0 + 8192])
        log(f"internal points {c0}..{c0 + 8192} done, {time.perf_counter() - t0:.1f} s")
    t_som = time.perf_counter() - t0
    ns = 8000
    A = rng.normal(size=(ns, ns))
    t0 = time.perf_counter()
    np.linalg.solve(A, rng.normal(size=ns))
    tl = time.perf_counter() - t0
    N = 3 * p.n
    print(json.dumps({"mode": "cpu-full", "workload": "76806-DOF cantilever, 65536 points", "host": host,
                      "omp_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                      "assembly_s": round(t_asm, 2), "internal_points_s": round(t_som, 2),
                      "lu_s_extrapolated": round(tl * (N / ns) ** 3, 1), "lu_sample": f"numpy gesv N={ns}: {tl:.2f} s",
                      "note": "assembly and internal points timed in full (the near table is built per chunk of "
                              "8192 points); the LU extrapolated by N^3"}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp16", choices=["fp16", "fp32", "fp64"])
    ap.add_argument("--config", default="c3", choices=["c3", "c1", "c4"],
                    help="c3: the 76,806-DOF cantilever step (default); c1: the 18,438-DOF cube; "
                         "c4: adjoint gradient on the perturbed cube")
    ap.add_argument("--fd-dirs", type=int, default=64, help="c4: directions of the finite-difference check")
    ap.add_argument("--trailing-update-only", action="store_true",
                    help="only the FP16 trailing-update launches (154) of one 76,806-DOF factorisation "
                         "(qd_fp16_update on synthetic operands), for hardware counter passes")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (reference for the HBM byte counters)")
    ap.add_argument("--assembly-only", action="store_true", help="only ghmat8, --steps times (no solve)")
    ap.add_argument("--internal-only", action="store_true",
                    help="only the internal points, --steps times, on a synthetic boundary field (no solve)")
    ap.add_argument("--assembly", default="fp64", choices=["fp64", "far_fp32", "far_fp32_factor"],
                    help="assembly precision of the --assembly-only / --internal-only modes")
    ap.add_argument("--cpu-full", action="store_true",
                    help="only the CPU baseline of the cantilever with assembly and internal points in full")
    args = ap.parse_args()
    if args.cpu_full:
        bench_cpu_full(args)
        return
    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")   # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the dense LU of this path runs on a single device in this "
              "build; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from bem3d_elastostatic_amd import mesh

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.config == "c4":
        bench_gradient(args)
        return
    if args.trailing_update_only:
        bench_trailing_update()
        return
    if args.assembly_only:
        bench_assembly_only(args)
        return
    if args.internal_only:
        bench_internal_only(args)
        return
    if args.config == "c3":
        p = mesh.config3()
        workload = ("thick cantilever box [0,10]x[0,2]x[0,1], h=0.1: 25602 nodes, 2400 quad9 + 8000 tri6, "
                    "76806 DOF, 65536 internal points; E=210e3 nu=0.3, x=0 clamped, t_z=-1 on x=10")
    else:
        p = mesh.config1()
        workload = "unit cube 16x16 quad9 per face: 6146 nodes, 18438 DOF, 4096 internal points"
    N = 3 * p.n
    log(f"workload {args.config}: n={p.n} ne={p.ne} N={N} l={p.l}")
    pipe = Pipeline(p, args.precision, torch)
    q = pipe.q

    for w in range(args.warmup):
        ph = []
        pipe.step(phases=ph)
        log(f"warmup {w}: {ph[-1]}")
    q.profile_enable(["lu_trailing_gemm", "lu_panel", "assembly_h", "assembly_g", "somigliana_far",
                      "somigliana_near"])
    for k in ("lu_trailing_gemm", "lu_panel", "assembly_h", "assembly_g", "somigliana_far", "somigliana_near"):
        q.profile_read(k)
    phases = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        FI, DFI = pipe.step(phases=phases)
        log(f"step {s}: {phases[-1]}")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = {k: q.profile_read(k) for k in ("lu_trailing_gemm", "lu_panel", "assembly_h", "assembly_g",
                                          "somigliana_far", "somigliana_near")}
    q.profile_enable([])
    u_mixed, s_mixed = pipe.u.cpu().numpy(), pipe.s.cpu().numpy()
    ms_per_step = 1000.0 * dt / args.steps

    def mean(key):
        return float(np.mean([ph[key] for ph in phases]))

    t_asm = mean("assembly")
    t_solve = mean("solve")
    entries = hg_entries(p)
    lu_flops = 2.0 / 3.0 * N ** 3
    t_factor = mean("factor_ms") / 1000.0
    extra = dict(
        phases_s=dict(assembly=t_asm, solve=t_solve, factor=t_factor, refine=mean("refine_ms") / 1000.0,
                      reorganize=mean("reorganize"), internal_points=mean("internal")),
        hg_entries_per_s=entries / t_asm,
        lu_mfma_pct=100.0 * lu_flops / t_factor / (PEAK_TFLOPS[args.precision] * 1e12),
        ir_iterations=phases[-1]["iterations"], ir_gmres=phases[-1]["gmres"],
        ir_residual=phases[-1]["residual"],
        somigliana_gp_evals=dict(far=phases[-1]["counts"][0], near=phases[-1]["counts"][1]),
    )
    prof_out = {}
    for k, (n_l, ms, work) in prof.items():
        prof_out[k] = dict(launches=n_l, ms=ms, work=work,
                           tflops=(work / (ms * 1e-3) / 1e12) if ms > 0 else None)
    extra["kernels"] = prof_out
    # Dominant single kernel by device time.  lu_panel is excluded: it is the
    # panel factorisation's whole sequence (about 130 small dependent
    # launches per panel), whose time is reported in "kernels" and "phases_s".
    dom = max((k for k in prof_out if k != "lu_panel"), key=lambda k: prof_out[k]["ms"])
    d = prof_out[dom]
    if dom == "lu_trailing_gemm":
        roofline = trailing_roofline(N, args.precision, d, diagonal=bool(phases[-1]["pivoting"]))
    else:
        peak = PEAK_TFLOPS["fp64_vector"]
        achieved = d["tflops"] or 0.0
        roofline = dict(bound="valu", kernel=dom, achieved=round(achieved, 3), peak=peak, unit="TFLOP/s",
                        frac=round(achieved / peak, 4), traffic=None, launches=d["launches"],
                        avg_launch_ms=d["ms"] / max(d["launches"], 1),
                        work_per_launch=d["work"] / max(d["launches"], 1))

    if not args.no_fp64 and args.precision != "fp64":
        log("FP64 LU path, one step")
        ph64 = []
        FI64, DFI64 = pipe.step(precision="fp64", phases=ph64)
        log(f"fp64 step: {ph64[-1]}")
        u64, s64 = pipe.u.cpu().numpy(), pipe.s.cpu().numpy()
        t64 = ph64[-1]["assembly"] + ph64[-1]["solve"] + ph64[-1]["reorganize"] + ph64[-1]["internal"]

        def rel(a, b):
            return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
        mask = np.arange(27)[None, :] < 3 * p.nconec[:, None]
        extra["fp64_path_s_per_solve"] = t64
        extra["fp64_factor_s"] = ph64[-1]["factor_ms"] / 1000.0
        extra["fp64_backward_error"] = ph64[-1]["residual"]
        extra["fp64_lu_mfma_pct"] = 100.0 * lu_flops / (ph64[-1]["factor_ms"] / 1000.0) / (PEAK_TFLOPS["fp64"] * 1e12)
        extra["speedup_vs_fp64"] = t64 / (ms_per_step / 1000.0)
        extra["errors_vs_fp64"] = dict(u_boundary=rel(FI, FI64), t_boundary=rel(DFI[mask], DFI64[mask]),
                                       u_internal=rel(u_mixed, u64), sigma_internal=rel(s_mixed, s64))
    cpu = None
    if not args.no_cpu:
        log("CPU baseline sample")
        cpu = cpu_baseline(p, FI, DFI)
        log(f"cpu baseline {cpu}")
    out = {
        "metric": METRIC, "value": round(ms_per_step / 1000.0, 4), "unit": "s/solve", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp16": "f16 MFMA LU (f32 acc) + f64 refine/assembly/Somigliana",
                  "fp32": "f32 MFMA LU + f64 refine/assembly/Somigliana", "fp64": "f64"}[args.precision],
        "data": "synthetic", "config": {"workload": workload, "precision": args.precision, "dof": N,
                                        "internal_points": p.l},
        "roofline": roofline, "cpu_baseline": cpu, **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
