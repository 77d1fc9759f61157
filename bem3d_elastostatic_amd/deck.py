"""Reader and writer for the QUADESTH input deck (io_mod.f90:36-146).

Layout (list-directed Fortran reads, so commas and blanks both separate):

1. title line;
2. header, modern ``n, ne, l, mu, nu`` or legacy ``n, ne, l, nfr, mu, dam, ro, nu``
   followed by one frequency line (io_mod.f90:51-61);
3. ``n`` lines ``node x y z``;
4. ``ne`` lines ``elem nconec k1 .. k_nconec``;
5. ``3*ne`` lines ``elem kode v1 .. v_nconec`` (direction x, y, z per element,
   io_mod.f90:97-113); legacy lines hold complex pairs ``(re, im)`` of which
   the real part is kept (io_mod.f90:18-27);
6. ``3*l`` reals: the internal points (read as one list, io_mod.f90:115-117);
7. optionally the tracked node (io_mod.f90:137-140, default 1).

The writer emits full precision (``%.17g``): the reference's own preprocessor
rounds coordinates to 4 decimals and mu/nu to 1/2 decimals, which is a
different problem, not a different rounding of the same one.
"""
from __future__ import annotations

import re

import numpy as np

from .mesh import BemProblem

_SPLIT = re.compile(r"[,\s]+")


def _tokens(line: str):
    return [t for t in _SPLIT.split(line.strip()) if t]


def _num(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def read_deck(path: str) -> BemProblem:
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    pos = 0
    title = lines[pos].rstrip()
    pos += 1
    head = _tokens(lines[pos])
    pos += 1
    if len(head) >= 8:
        n, ne, l = int(head[0]), int(head[1]), int(head[2])
        mu, nu = _num(head[4]), _num(head[7])
        pos += 1  # legacy frequency line
    else:
        n, ne, l = int(head[0]), int(head[1]), int(head[2])
        mu, nu = _num(head[3]), _num(head[4])
    x = np.zeros(n)
    y = np.zeros(n)
    z = np.zeros(n)
    for _ in range(n):
        t = _tokens(lines[pos])
        pos += 1
        i = int(t[0]) - 1
        x[i], y[i], z[i] = _num(t[1]), _num(t[2]), _num(t[3])
    nconec = np.zeros(ne, dtype=np.int32)
    kconec = np.zeros((ne, 9), dtype=np.int32)
    for _ in range(ne):
        t = _tokens(lines[pos])
        pos += 1
        e = int(t[0]) - 1
        nconec[e] = int(t[1])
        kconec[e, : nconec[e]] = [int(v) for v in t[2: 2 + nconec[e]]]
    kode = np.ones((ne, 3), dtype=np.int32)
    bc = np.zeros((ne, 27))
    for _ in range(ne):
        for k in range(3):
            raw = lines[pos]
            pos += 1
            legacy = "(" in raw
            t = _tokens(raw.replace("(", " ").replace(")", " "))
            e = int(t[0]) - 1
            kode[e, k] = int(t[1])
            nc = int(nconec[e])
            vals = [_num(v) for v in t[2:]]
            if legacy:
                vals = vals[0::2]
            for j in range(nc):
                bc[e, 3 * j + k] = vals[j]
    rest = []
    while pos < len(lines):
        rest.extend(_tokens(lines[pos]))
        pos += 1
    pts = np.array([_num(v) for v in rest[: 3 * l]]).reshape(l, 3) if l > 0 else np.zeros((0, 3))
    ntrack = int(_num(rest[3 * l])) if len(rest) > 3 * l else 1
    return BemProblem(title, x, y, z, nconec, kconec, kode, bc, mu, nu, pts, ntrack)


def write_deck(p: BemProblem, path: str) -> None:
    f17 = "%.17g"
    with open(path, "w") as fh:
        fh.write(p.title + "\n")
        fh.write(f"{p.n},{p.ne},{p.l},{f17 % p.mu},{f17 % p.nu}\n")
        for i in range(p.n):
            fh.write(f"{i + 1} {f17 % p.x[i]} {f17 % p.y[i]} {f17 % p.z[i]}\n")
        for e in range(p.ne):
            nc = int(p.nconec[e])
            fh.write(f"{e + 1} {nc} " + " ".join(str(int(v)) for v in p.kconec[e, :nc]) + "\n")
        for e in range(p.ne):
            nc = int(p.nconec[e])
            for k in range(3):
                vals = " ".join(f17 % p.bc[e, 3 * j + k] for j in range(nc))
                fh.write(f"{e + 1} {int(p.kode[e, k])} {vals}\n")
        for i in range(p.l):
            fh.write(f"{f17 % p.points[i, 0]} {f17 % p.points[i, 1]} {f17 % p.points[i, 2]}\n")
        fh.write(f"{int(p.ntrack)}\n")
