"""Builds libquadesth.so (HIP, gfx950) in-tree with hipcc.

    python -m bem3d_elastostatic_amd.build

The shared library carries the C ABI of include/quadesth.h; Python reaches it
with ctypes (bem3d_elastostatic_amd/_lib.py), Fortran with ISO_C_BINDING.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libquadesth.so")
SOURCES = ["qd_capi.cpp", "qd_assembly.hip", "qd_somigliana.hip", "qd_lu.hip", "qd_lapack.cpp", "qd_grad.hip",
           "qd_output.cpp", "qd_plan.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QUADESTH_ARCH", "gfx950")

# Packed FP32 VALU forms are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def compile_commands(out=LIB):
    objs, cmds = [], []
    bdir = os.path.join(PKG, "_build")
    for src in SOURCES:
        obj = os.path.join(bdir, src + ".o")
        objs.append(obj)
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-c",
               "-x", "hip", os.path.join(CSRC, src), "-I", INCLUDE, "-I", CSRC,
               "-o", obj] + DEVICE_FLAGS
        cmds.append(cmd)
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out] + objs
    return bdir, cmds, link


def build(verbose: bool = False, jobs: int = 5) -> str:
    bdir, cmds, link = compile_commands()
    os.makedirs(bdir, exist_ok=True)
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # the host pass of hipcc does not know the device feature flag
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln or "error" in ln]
        if p.returncode != 0:
            failed = True
            sys.stderr.write("\n".join(lines) + "\n")
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed")
    subprocess.run(link, check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
