"""Tabulated (numerical) radial basis functions: CHAMP's basis.<ict> files
and the spline data the engine evaluates.

A basis file (qmc/vmc/read_bas_num.f90:81-114) holds, after '#' comment
lines, one header record `nrbas igrid nr exp_h_bas r0_bas icusp`, for
all-electron systems (nloc = 0) a record of the l of each function, and nr
rows `x R_1(x) ... R_nrbas(x)`.  The grid is

    igrid 1: x_i = r0 + (i - 1) h                 (h = exp_h_bas)
    igrid 2: x_i = r0 h^(i - 1)
    igrid 3: x_i = r0' (h^(i - 1) - 1), r0' = r0 / (h^(nr - 1) - 1)

(for igrid 3 the file's r0_bas is the last point; read_bas_num.f90:225
turns it into r0').  For each function the setup follows
read_bas_num.f90:238-381: a 5-term polynomial through the first five
points (the Vandermonde system dgelg solves; the cusp-constrained fit of
all-electron s functions with icusp = 1 is not provided), an exponential
a0 exp(-ak r) matched at the last two points, and the cubic spline of
spline2.f90 with the polynomial's slope at x_1 and the exponential's slope
at x_nr as end conditions.  splfit.f90:1-79 evaluates the three pieces.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

NCOEF = 5


@dataclass
class BasisFile:
    igrid: int
    nr: int
    exp_h: float
    r0: float             # as written in the file
    icusp: int
    x: np.ndarray         # [nr]
    rwf: np.ndarray       # [nrbas][nr]
    l: list | None        # per function (all-electron files only)


def read_basis_file(text: str, nloc: int) -> BasisFile:
    """Parse one basis.<ict> file (read_bas_num.f90:87-114)."""
    lines = [ln for ln in text.splitlines()]
    i = 0
    while lines[i].lstrip().startswith("#") or not lines[i].strip():
        i += 1
    hdr = lines[i].split()
    nrbas, igrid, nr = int(hdr[0]), int(hdr[1]), int(hdr[2])
    exp_h, r0, icusp = float(hdr[3]), float(hdr[4]), int(hdr[5])
    i += 1
    if igrid not in (1, 2, 3):
        raise ValueError("grid not implemented (igrid 1, 2 or 3)")
    if igrid == 3 and 0.0 < r0 < 1.0:
        raise ValueError("for igrid = 3 r0_bas is the last point of the grid, not the first")
    if igrid == 2 and r0 > 1.0:
        raise ValueError("for igrid = 2 r0_bas is the first point of the grid")
    ls = None
    if nloc == 0:
        ls = [int(v) for v in lines[i].split()[:nrbas]]
        i += 1
    vals = []
    while len(vals) < nr:
        ln = lines[i]
        i += 1
        if not ln.strip():
            continue
        vals.append([float(t.lower().replace("d", "e")) for t in ln.split()[:nrbas + 1]])
    a = np.array(vals)
    return BasisFile(igrid=igrid, nr=nr, exp_h=exp_h, r0=r0, icusp=icusp, x=a[:, 0].copy(),
                     rwf=a[:, 1:].T.copy(), l=ls)


def spline2(x, y, yp1, ypn):
    """Second derivatives of the cubic spline through (x, y) with end slopes
    yp1 and ypn, > 1e30 meaning a natural end (spline2.f90:1-79)."""
    n = len(x)
    y2 = np.zeros(n)
    u = np.zeros(n)
    if yp1 > 1e30:
        y2[0] = 0.0
        u[0] = 0.0
    else:
        y2[0] = -0.5
        u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1)
    for i in range(1, n - 1):
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
        p = sig * y2[i - 1] + 2.0
        y2[i] = (sig - 1.0) / p
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1])
                       / (x[i] - x[i - 1])) / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p
    if ypn > 1e30:
        qn = un = 0.0
    else:
        qn = 0.5
        un = (3.0 / (x[n - 1] - x[n - 2])) * (ypn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]))
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0)
    for k in range(n - 2, -1, -1):
        y2[k] = y2[k] * y2[k + 1] + u[k]
    return y2


@dataclass
class RadialTable:
    """One tabulated function in the form qmcw_system takes it."""
    igrid: int
    nr: int
    h: float
    r0: float             # r0 as splfit_bas uses it
    rwf: np.ndarray
    d2: np.ndarray
    ce: np.ndarray        # [5]
    ae: np.ndarray        # [2]


def grid_params(bf: BasisFile):
    """(h, r0) of the evaluation (read_bas_num.f90:224-233)."""
    h, r0 = bf.exp_h, bf.r0
    if bf.igrid == 2 and h <= 1.0:
        h = bf.x[1] / bf.x[0]
    if bf.igrid == 3:
        r0 = r0 / (h ** (bf.nr - 1) - 1.0)
    if bf.igrid == 2 and (r0 <= 0.0 or h <= 0.0):
        r0 = bf.x[0]
        h = bf.x[1] / bf.x[0]
    return h, r0


def radial_tables(bf: BasisFile, nloc: int) -> list:
    """Spline data of every function of a basis file
    (read_bas_num.f90:238-381, ndim = 3)."""
    h, r0 = grid_params(bf)
    x = bf.x
    nr = bf.nr
    out = []
    for k, y in enumerate(bf.rwf):
        if nloc == 0 and bf.l is not None and bf.l[k] == 0 and bf.icusp == 1:
            raise ValueError("the cusp-constrained small-r fit (icusp = 1) is not provided")
        V = np.vander(x[:NCOEF], NCOEF, increasing=True)
        ce = np.linalg.solve(V, y[:NCOEF])
        dwf1 = sum((i - 1) * ce[i - 1] * x[0] ** (i - 2) for i in range(2, NCOEF + 1))
        wfm = 0.5 * (y[nr - 1] + y[nr - 2])
        dwfm = (y[nr - 1] - y[nr - 2]) / (x[nr - 1] - x[nr - 2])
        if abs(wfm) > 1e-99:
            ak = -dwfm / wfm if dwfm / wfm < 0 else 0.1
            a0 = y[nr - 1] * np.exp(ak * x[nr - 1])
            dwfn = -ak * y[nr - 1]
        else:
            a0 = ak = dwfn = 0.0
        out.append(RadialTable(igrid=bf.igrid, nr=nr, h=h, r0=r0, rwf=y.copy(),
                               d2=spline2(x, y, dwf1, dwfn), ce=ce, ae=np.array([a0, ak])))
    return out

