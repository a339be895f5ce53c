"""System descriptions for the wavefunction engine.

A :class:`System` holds, as numpy arrays, the data CHAMP keeps in its common
modules for one wavefunction (nuclei, basis, orbital coefficients,
determinants/CSFs, ijas=4 Jastrow, Gaussian pseudopotential and the lists of
parameters to differentiate).  ``System.struct()`` exposes it as the
``qmcw_system`` C struct of ``include/champ_amd.h``.

Builders:
  * :func:`from_champ_input` -- CHAMP's keyword-menu input (the subset used by
    finite molecules with analytic Slater/Gaussian bases), e.g.
    ``examples/be_vb1_jcas24_vmc_j_opt_lin.inp``;
  * :func:`from_champ_old_input` -- CHAMP's old line-oriented input (finite
    all-electron systems, Slater s/p basis), e.g.
    ``examples/c2_ktt_hfj_vmc_emin_j_opt3_lin.inp``;
  * :func:`be_atom`, :func:`c2_dimer`, :func:`hchain` and
    :func:`random_molecule` -- the seeded synthetic workloads of the benchmark
    configurations.
"""
from __future__ import annotations

import ctypes as C
import math
import re
from dataclasses import dataclass, field

import numpy as np

from .abi import (QMCW_BASIS_GAUSSIAN, QMCW_BASIS_SLATER, QMCW_MAX_L, QMCW_MAX_LPOT,
                  QmcwSystem, dptr, iptr)


def nterms4(nord: int) -> int:
    """Number of e-e-n terms of order ``nord`` (qmc/vmc/jastrow4.f90:495-516)."""
    n_terms = 0
    for n in range(2, nord + 1):
        for k in range(n - 1, -1, -1):
            l_hi = n - k - 2 if k == 0 else n - k
            for l in range(l_hi, -1, -1):
                if (n - k - l) % 2 == 0:
                    n_terms += 1
    return n_terms


def een_terms(nord: int):
    """(k, l, m) of each e-e-n term in CHAMP's order."""
    out = []
    for n in range(2, nord + 1):
        for k in range(n - 1, -1, -1):
            l_hi = n - k - 2 if k == 0 else n - k
            for l in range(l_hi, -1, -1):
                if (n - k - l) % 2 == 0:
                    out.append((k, l, (n - k - l) // 2))
    return out


def cusp_matrix(nordc: int):
    """Cusp-condition matrix d and dependent-term indices iwc4 (1-based):
    qmc/fit/cuspinit4.f90:1-60."""
    nt = nterms4(nordc)
    neqs = 2 * (nordc - 1) if nordc > 1 else 0
    d = np.zeros((neqs, nt))
    iwc4 = [0] * neqs
    for i, (k, l, m) in enumerate(een_terms(nordc), start=1):
        n = k + l + 2 * m
        if k == 1 and iwc4[n - 2] == 0:
            iwc4[n - 2] = i
        if k == 0 and iwc4[n + nordc - 3] == 0:
            iwc4[n + nordc - 3] = i
        d[l + 2 * m - 1, i - 1] -= 2 * k
        if l + m > 0:
            d[nordc - 1 + k + m - 1, i - 1] -= (l + m)
        if m > 0:
            d[nordc - 1 + k + l + m - 1, i - 1] -= m
    return d, iwc4


def cusp_exact(c: np.ndarray, nordc: int) -> np.ndarray:
    """Set the dependent e-e-n coefficients from the independent ones
    (qmc/fit/cuspexact4.f90:1-74); c is [nctype][nterms]."""
    c = np.array(c, dtype=np.float64, copy=True)
    d, iwc4 = cusp_matrix(nordc)
    order = list(range(1, nordc)) + list(range(nordc + 1, 2 * (nordc - 1) + 1))
    if nordc > 1:
        order.append(nordc)
    for ct in c:
        for i in order:
            dep = iwc4[i - 1] - 1
            s = sum(d[i - 1, j] * ct[j] for j in range(len(ct)) if j != dep)
            ct[dep] = -s / d[i - 1, dep]
    return c


def radial_norm(basis_type: int, n: int, l: int, zetas, coefs) -> float:
    """Normalisation of r^(n-1) sum_p d_p exp(-zeta_p r^q) times a real solid
    harmonic with <Y^2> = 4 pi/(2l+1): for one primitive this is
    qmc/vmc/basis_norm.f90:28-37 (Slater (2n)!, Gaussian Gamma(n+1/2))."""
    s = 0.0
    for zp, dp in zip(zetas, coefs):
        for zq, dq in zip(zetas, coefs):
            if basis_type == QMCW_BASIS_SLATER:
                s += dp * dq * math.factorial(2 * n) / (zp + zq) ** (2 * n + 1)
            else:
                s += dp * dq * math.gamma(n + 0.5) / (2.0 * (zp + zq) ** (n + 0.5))
    return math.sqrt((2 * l + 1) / (4 * math.pi) / s)


_LABEL_LM = {"s": (0, [0]), "px": (1, [1]), "py": (1, [-1]), "pz": (1, [0])}


def parse_label(lab: str):
    """CHAMP basis label (qmc/basis_mod.f90:400-470) -> (n, l, m)."""
    lab = lab.lower()
    mt = re.fullmatch(r"(\d)([spdfgh])(x|y|z|[+-]?\d)?", lab)
    if not mt:
        raise ValueError(f"unknown basis function label {lab!r}")
    n = int(mt.group(1))
    l = "spdfgh".index(mt.group(2))
    sub = mt.group(3)
    if l == 0:
        return n, 0, 0
    if l == 1 and sub in ("x", "y", "z"):
        return n, 1, {"x": 1, "y": -1, "z": 0}[sub]
    return n, l, int(sub)


@dataclass
class System:
    nelec: int
    nup: int
    cent: np.ndarray
    iwctype: np.ndarray
    znuc: np.ndarray
    basis_type: int
    rad_n: np.ndarray
    rad_prim_off: np.ndarray
    prim_zeta: np.ndarray
    prim_coef: np.ndarray
    ctype_bas_off: np.ndarray
    bas_rad: np.ndarray
    bas_l: np.ndarray
    bas_m: np.ndarray
    coef: np.ndarray            # [norb][nbasis], normalisation folded in
    det_up: np.ndarray
    det_dn: np.ndarray
    csf_coef: np.ndarray
    csf_det_off: np.ndarray
    csf_det: np.ndarray
    csf_det_coef: np.ndarray
    isc: int = 4
    nspin2b: int = 1
    nocuspb: int = 0
    norda: int = 5
    nordb: int = 5
    nordc: int = 5
    scalek: float = 0.8
    a4: np.ndarray = None
    b: np.ndarray = None
    c: np.ndarray = None
    nloc: int = 0
    npotd: np.ndarray = None
    lpotp1: np.ndarray = None
    ecp_off: np.ndarray = None
    ecp_coef: np.ndarray = None
    ecp_power: np.ndarray = None
    ecp_expo: np.ndarray = None
    nquad: int = 12
    jas_a: list = field(default_factory=list)   # [(type, idx1)]
    jas_b: list = field(default_factory=list)   # [(spin, idx1)]
    jas_c: list = field(default_factory=list)   # [(type, idx1)] independent
    csf_parm: list = field(default_factory=list)
    orb_parm: list = field(default_factory=list)  # [(i, j, rev)]
    hb: float = 0.5
    # tabulated radial functions (numerical basis): rad_num[r] >= 0 selects
    # table k for radial function r; the tables as in qmcw_system
    rad_num: np.ndarray = None
    num_igrid: np.ndarray = None
    num_nr: np.ndarray = None
    num_h: np.ndarray = None
    num_r0: np.ndarray = None
    num_off: np.ndarray = None
    num_rwf: np.ndarray = None
    num_d2: np.ndarray = None
    num_ce: np.ndarray = None
    num_ae: np.ndarray = None
    name: str = ""
    _keep: list = field(default_factory=list, repr=False)

    # ---- derived sizes -------------------------------------------------
    @property
    def ncent(self):
        return len(self.iwctype)

    @property
    def nctype(self):
        return len(self.znuc)

    @property
    def nbasis(self):
        return int(sum(self.ctype_bas_off[t + 1] - self.ctype_bas_off[t]
                       for t in self.iwctype))

    @property
    def norb(self):
        return self.coef.shape[0]

    @property
    def ndn(self):
        return self.nelec - self.nup

    @property
    def nterms(self):
        return nterms4(self.nordc)

    @property
    def nparm(self):
        return (len(self.csf_parm) + len(self.jas_a) + len(self.jas_b)
                + len(self.jas_c) + len(self.orb_parm))

    def copy(self):
        import copy as _c
        s = _c.copy(self)
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                setattr(s, k, v.copy())
            elif isinstance(v, list):
                setattr(s, k, list(v))
        s._keep = []
        return s

    # ---- C struct --------------------------------------------------------
    def struct(self) -> QmcwSystem:
        f64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
        i32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int32))
        arrs = {}

        def D(name, a):
            arrs[name] = f64(a)
            return dptr(arrs[name])

        def I(name, a):
            arrs[name] = i32(a) if len(np.atleast_1d(a)) else np.zeros(1, np.int32)
            return iptr(arrs[name])

        s = QmcwSystem()
        s.nelec, s.nup = self.nelec, self.nup
        s.ncent, s.nctype = self.ncent, self.nctype
        s.cent = D("cent", self.cent.reshape(-1))
        s.iwctype = I("iwctype", self.iwctype)
        s.znuc = D("znuc", self.znuc)
        s.basis_type = self.basis_type
        s.nrad = len(self.rad_n)
        s.rad_n = I("rad_n", self.rad_n)
        s.rad_prim_off = I("rad_prim_off", self.rad_prim_off)
        s.prim_zeta = D("prim_zeta", self.prim_zeta)
        s.prim_coef = D("prim_coef", self.prim_coef)
        s.ctype_bas_off = I("ctype_bas_off", self.ctype_bas_off)
        s.bas_rad = I("bas_rad", self.bas_rad)
        s.bas_l = I("bas_l", self.bas_l)
        s.bas_m = I("bas_m", self.bas_m)
        s.nbasis = self.nbasis
        s.norb = self.norb
        s.coef = D("coef", self.coef.reshape(-1))
        s.ndet = self.det_up.shape[0]
        s.det_up = I("det_up", self.det_up.reshape(-1))
        s.det_dn = I("det_dn", self.det_dn.reshape(-1) if self.ndn else np.zeros(1))
        s.ncsf = len(self.csf_coef)
        s.csf_coef = D("csf_coef", self.csf_coef)
        s.csf_det_off = I("csf_det_off", self.csf_det_off)
        s.csf_det = I("csf_det", self.csf_det)
        s.csf_det_coef = D("csf_det_coef", self.csf_det_coef)
        s.isc, s.nspin2b, s.nocuspb = self.isc, self.nspin2b, self.nocuspb
        s.norda, s.nordb, s.nordc = self.norda, self.nordb, self.nordc
        s.scalek = self.scalek
        s.a4 = D("a4", self.a4.reshape(-1))
        s.b = D("b", self.b.reshape(-1))
        s.c = D("c", self.c.reshape(-1) if self.c.size else np.zeros(1))
        s.nloc = self.nloc
        if self.nloc:
            s.npotd = I("npotd", self.npotd)
            s.lpotp1 = I("lpotp1", self.lpotp1)
            s.ecp_off = I("ecp_off", self.ecp_off)
            s.ecp_coef = D("ecp_coef", self.ecp_coef if len(self.ecp_coef) else [0.0])
            s.ecp_power = I("ecp_power", self.ecp_power if len(self.ecp_power) else [0])
            s.ecp_expo = D("ecp_expo", self.ecp_expo if len(self.ecp_expo) else [0.0])
        s.nquad = self.nquad
        ja = np.array(self.jas_a, dtype=np.int32).reshape(-1, 2)
        jb = np.array(self.jas_b, dtype=np.int32).reshape(-1, 2)
        jc = np.array(self.jas_c, dtype=np.int32).reshape(-1, 2)
        s.nparma_tot = len(ja)
        s.jas_a_type = I("ja_t", ja[:, 0])
        s.jas_a_idx = I("ja_i", ja[:, 1])
        s.nparmb_tot = len(jb)
        s.jas_b_spin = I("jb_s", jb[:, 0])
        s.jas_b_idx = I("jb_i", jb[:, 1])
        s.nparmc_tot = len(jc)
        s.jas_c_type = I("jc_t", jc[:, 0])
        s.jas_c_idx = I("jc_i", jc[:, 1])
        s.nparmcsf = len(self.csf_parm)
        s.csf_parm = I("csf_parm", self.csf_parm)
        op = np.array(self.orb_parm, dtype=np.int32).reshape(-1, 3)
        s.nparmorb = len(op)
        s.orb_parm_i = I("op_i", op[:, 0])
        s.orb_parm_j = I("op_j", op[:, 1])
        s.orb_parm_rev = I("op_r", op[:, 2])
        s.hb = self.hb
        if self.rad_num is not None and len(self.num_nr):
            s.rad_num = I("rad_num", self.rad_num)
            s.nnum = len(self.num_nr)
            s.num_igrid = I("num_igrid", self.num_igrid)
            s.num_nr = I("num_nr", self.num_nr)
            s.num_h = D("num_h", self.num_h)
            s.num_r0 = D("num_r0", self.num_r0)
            s.num_off = I("num_off", self.num_off)
            s.num_rwf = D("num_rwf", self.num_rwf)
            s.num_d2 = D("num_d2", self.num_d2)
            s.num_ce = D("num_ce", np.asarray(self.num_ce).reshape(-1))
            s.num_ae = D("num_ae", np.asarray(self.num_ae).reshape(-1))
        s._arrays = arrs          # keep the buffers alive with the struct
        self._keep.append(arrs)
        return s

    def struct_ptr(self):
        s = self.struct()
        self._keep.append(s)
        return C.pointer(s)


# ---------------------------------------------------------------------------
# common construction helpers
# ---------------------------------------------------------------------------
class _BasisBuilder:
    """Collect radial functions and basis functions per centre type."""

    def __init__(self, basis_type, nctype):
        self.basis_type = basis_type
        self.rad_n, self.prim_zeta, self.prim_coef = [], [], []
        self.rad_off = [0]
        self.rad_num = []                              # table index or -1
        self.tables = []                               # numbas.RadialTable
        self.per_type = [[] for _ in range(nctype)]   # (rad, l, m, norm)

    def add_radial(self, n, zetas, coefs):
        self.rad_n.append(n)
        self.prim_zeta.extend(zetas)
        self.prim_coef.extend(coefs)
        self.rad_off.append(len(self.prim_zeta))
        self.rad_num.append(-1)
        return len(self.rad_n) - 1

    def add_table(self, table):
        """Register a tabulated radial function; returns its table index."""
        self.tables.append(table)
        return len(self.tables) - 1

    def add_numeric_radial(self, k):
        """A radial function that is table k (no primitives)."""
        self.rad_n.append(0)
        self.rad_off.append(len(self.prim_zeta))
        self.rad_num.append(k)
        return len(self.rad_n) - 1

    def add_function(self, ictype, rad, l, m):
        if self.rad_num[rad] >= 0:
            # tabulated R(r) Y_lm: qmc/vmc/basis_norm.f90:38-39
            self.per_type[ictype].append((rad, l, m, math.sqrt((2 * l + 1) / (4 * math.pi))))
            return
        n = self.rad_n[rad]
        z = self.prim_zeta[self.rad_off[rad]:self.rad_off[rad + 1]]
        d = self.prim_coef[self.rad_off[rad]:self.rad_off[rad + 1]]
        self.per_type[ictype].append((rad, l, m, radial_norm(self.basis_type, n, l, z, d)))

    def arrays(self, iwctype):
        off = [0]
        bas_rad, bas_l, bas_m = [], [], []
        for t in self.per_type:
            for (r, l, m, _) in t:
                bas_rad.append(r)
                bas_l.append(l)
                bas_m.append(m)
            off.append(len(bas_rad))
        anorm = []
        for t in iwctype:
            anorm.extend(a for (_, _, _, a) in self.per_type[t])
        num = {}
        if self.tables:
            T = self.tables
            num = dict(rad_num=np.array(self.rad_num, np.int32),
                       num_igrid=np.array([t.igrid for t in T], np.int32),
                       num_nr=np.array([t.nr for t in T], np.int32),
                       num_h=np.array([t.h for t in T]), num_r0=np.array([t.r0 for t in T]),
                       num_off=np.cumsum([0] + [t.nr for t in T]).astype(np.int32),
                       num_rwf=np.concatenate([t.rwf for t in T]),
                       num_d2=np.concatenate([t.d2 for t in T]),
                       num_ce=np.array([t.ce for t in T]), num_ae=np.array([t.ae for t in T]))
        return dict(basis_type=self.basis_type, **num,
                    rad_n=np.array(self.rad_n, np.int32),
                    rad_prim_off=np.array(self.rad_off, np.int32),
                    prim_zeta=np.array(self.prim_zeta, np.float64),
                    prim_coef=np.array(self.prim_coef, np.float64),
                    ctype_bas_off=np.array(off, np.int32),
                    bas_rad=np.array(bas_rad, np.int32),
                    bas_l=np.array(bas_l, np.int32),
                    bas_m=np.array(bas_m, np.int32)), np.array(anorm)


def _csf_arrays(csfs, normalize):
    """csfs: list of (coef, [(det, cdet), ...])."""
    off, dets, cd = [0], [], []
    for _, lst in csfs:
        vals = [c for _, c in lst]
        nrm = math.sqrt(sum(v * v for v in vals)) if normalize else 1.0
        for d, c in lst:
            dets.append(d)
            cd.append(c / nrm)
        off.append(len(dets))
    return (np.array([c for c, _ in csfs], np.float64), np.array(off, np.int32),
            np.array(dets, np.int32), np.array(cd, np.float64))


def sort_dets(det_up, det_dn, csf_det, csf_det_coef):
    """Order orbital labels within each determinant and flip the CSF
    coefficients by the permutation parity (qmc/vmc/read_input.f90:2235-2273)."""
    det_up = np.array(det_up, np.int32)
    det_dn = np.array(det_dn, np.int32)
    sign = np.ones(len(det_up))
    for arr in (det_up, det_dn):
        for i, row in enumerate(arr):
            r = list(row)
            for j in range(len(r)):
                for k in range(j + 1, len(r)):
                    if r[k] < r[j]:
                        r[j], r[k] = r[k], r[j]
                        sign[i] = -sign[i]
            arr[i] = r
    cdc = np.array([c * sign[d] for d, c in zip(csf_det, csf_det_coef)])
    return det_up, det_dn, cdc


def _random_orthonormal(rng, n, m=None):
    """n x m matrix with orthonormal columns, Gram-Schmidt of a seeded
    Gaussian matrix."""
    m = m or n
    a = rng.standard_normal((n, m))
    q = np.zeros_like(a)
    for j in range(m):
        v = a[:, j].copy()
        for i in range(j):
            v -= (q[:, i] @ a[:, j]) * q[:, i]
        for i in range(j):          # second pass for stability
            v -= (q[:, i] @ v) * q[:, i]
        q[:, j] = v / np.linalg.norm(v)
    return q


# ---------------------------------------------------------------------------
# CHAMP input (keyword menus, qmc/main_menu_mod.f90 and the *_mod menus)
# ---------------------------------------------------------------------------
def _blocks(text):
    """Split CHAMP's menu input into {menu: [lines]} (nested 'end' aware)."""
    lines = [ln.split("#")[0].rstrip() for ln in text.splitlines()]
    menus, i = {}, 0
    while i < len(lines):
        tok = lines[i].strip()
        if not tok:
            i += 1
            continue
        name = tok.split()[0].lower()
        body, depth = [], 1
        i += 1
        while i < len(lines) and depth > 0:
            t = lines[i].strip()
            if t.lower() == "end":
                depth -= 1
                if depth == 0:
                    break
            elif t.lower() in ("geometry", "basis_functions", "coefficients",
                               "energies", "symmetry", "determinants",
                               "dets_in_csfs", "parameters"):
                depth += 1
            body.append(lines[i])
            i += 1
        menus[name] = body
        i += 1
    return menus


def _sub(body, key):
    """Lines of a sub-block 'key ... end' of a menu body."""
    out, on = [], False
    for ln in body:
        t = ln.strip()
        if not on and t.lower().startswith(key):
            on = True
            rest = t[len(key):].strip()
            if rest:
                out.append(rest)
            continue
        if on:
            if t.lower() == "end" or t.lower().endswith(" end") and key != "csf_coef":
                if t.lower() != "end":
                    out.append(t[:-3])
                break
            out.append(t)
    return out


def _nums(line, conv=float):
    out = []
    for tok in line.replace(",", " ").split():
        try:
            out.append(conv(tok))
        except ValueError:
            break
    return out


def _kv(body):
    kv = {}
    for ln in body:
        for m in re.finditer(r"(\w+)\s*=\s*([-+.\w]+)", ln):
            kv[m.group(1).lower()] = m.group(2)
    return kv


def from_champ_input(text: str, name: str = "champ") -> System:
    """Build a System from CHAMP's keyword-menu input.

    Supported: finite systems, analytic Slater (all-electron default) or
    Gaussian (``which_analytical_basis = gaussian``) bases, orbitals given by
    coefficients, determinants/CSFs, ijas=4 Jastrow.  The dependent e-e-n
    coefficients are re-imposed with cuspexact4.
    """
    menus = _blocks(text)
    nuc = menus["nuclei"]
    geo = _sub(nuc, "geometry")
    types, cent = [], []
    for ln in geo:
        v = ln.split()
        if len(v) < 5:
            continue
        types.append((int(v[0]) - 1, float(v[1])))
        cent.append([float(x) for x in v[2:5]])
    nctype = max(t for t, _ in types) + 1
    znuc = np.zeros(nctype)
    for t, z in types:
        znuc[t] = z
    iwctype = np.array([t for t, _ in types], np.int32)
    cent = np.array(cent)
    kvn = _kv(nuc)
    nloc = int(kvn.get("nloc", 0))
    if nloc not in (0,):
        raise ValueError("pseudopotential inputs are built with the synthetic builders")
    wf = _kv(menus["wavefunction"])
    nelec, nup = int(wf["nelec"]), int(wf["nup"])
    bas = menus["basis"]
    kvb = _kv(bas)
    btype = QMCW_BASIS_GAUSSIAN if kvb.get("which_analytical_basis", "slater").lower() == "gaussian" \
        else QMCW_BASIS_SLATER
    bb = _BasisBuilder(btype, nctype)
    ict = 0
    for ln in _sub(bas, "basis_functions"):
        t = ln.strip()
        if not t:
            continue
        if re.fullmatch(r"\d+", t):
            ict = int(t) - 1
            continue
        lab, z = t.split()[:2]
        n, l, m = parse_label(lab)
        r = bb.add_radial(n, [float(z)], [1.0])
        bb.add_function(ict, r, l, m)
    barr, anorm = bb.arrays(iwctype)
    rows = [_nums(ln) for ln in _sub(menus["orbitals"], "coefficients")]
    rows = [r for r in rows if r]
    coef = np.array(rows) * anorm[None, :]
    csfm = menus["csfs"]
    dets = [_nums(ln, int) for ln in _sub(csfm, "determinants")]
    dets = [d for d in dets if d]
    det_up = [[x - 1 for x in d[:nup]] for d in dets]
    det_dn = [[x - 1 for x in d[nup:nelec]] for d in dets]
    ccl = " ".join(ln for ln in csfm if "csf_coef" in ln.lower())
    ccl = ccl.lower().split("csf_coef", 1)[1].split("end")[0]
    csf_coef = _nums(ccl)
    dic = _sub(csfm, "dets_in_csfs")
    dic = [ln for ln in dic if ln.strip()]
    normalize = not any("." in ln for ln in dic[1::2])
    csfs = []
    for k in range(len(csf_coef)):
        ds = [x - 1 for x in _nums(dic[2 * k], int)]
        cs = _nums(dic[2 * k + 1])
        csfs.append((csf_coef[k], list(zip(ds, cs))))
    cc, off, cdet, cdc = _csf_arrays(csfs, normalize)
    det_up, det_dn, cdc = sort_dets(det_up, det_dn, cdet, cdc)
    jm = menus["jastrow"]
    kvj = _kv(jm)
    if int(kvj.get("ijas", 4)) != 4:
        raise ValueError("only ijas=4 is supported")
    nord = int(kvj.get("nord", 5))
    norda = int(kvj.get("norda", nord))
    nordb = int(kvj.get("nordb", nord))
    nordc = int(kvj.get("nordc", nord))
    nspin2b = int(kvj.get("nspin2b", 1))
    prows = [_nums(ln) for ln in _sub(jm, "parameters")]
    prows = [r for r in prows if r]
    a4 = np.array(prows[:nctype])[:, :norda + 1]
    b = np.array(prows[nctype:nctype + nspin2b])[:, :nordb + 1]
    c = np.array(prows[nctype + nspin2b:2 * nctype + nspin2b])[:, :nterms4(nordc)]
    c = cusp_exact(c, nordc)
    return System(nelec=nelec, nup=nup, cent=cent, iwctype=iwctype, znuc=znuc,
                  coef=coef, det_up=det_up, det_dn=det_dn, csf_coef=cc,
                  csf_det_off=off, csf_det=cdet, csf_det_coef=cdc,
                  isc=int(kvj.get("isc", 4)), nspin2b=nspin2b,
                  nocuspb=int(kvj.get("nocuspb", 0)), norda=norda, nordb=nordb,
                  nordc=nordc, scalek=float(kvj.get("scalek", 0.5)), a4=a4, b=b,
                  c=c, nloc=0, name=name, **barr)


class _OldReader:
    """Fortran list-directed records of CHAMP's old line-oriented input: each
    read takes the numeric prefix of lines (the trailing text is the
    variable-name comment) and continues on the next line while values are
    missing."""

    def __init__(self, lines, i=0):
        self.lines, self.i = lines, i

    def seek(self, key):
        while key not in self.lines[self.i]:
            self.i += 1
        self.i += 1

    def _numeric(self, ln):
        vals = []
        for tok in ln.split():
            try:
                vals.append(float(tok.lower().replace("d", "e")))
            except ValueError:
                break
        return vals

    def read(self, n=None):
        """One record: n values (None: the numbers on the next non-empty line)."""
        vals = []
        while True:
            ln = self.lines[self.i]
            self.i += 1
            if not ln.strip():
                continue
            vals.extend(self._numeric(ln))
            if n is None or len(vals) >= n:
                break
        if n is not None:
            if len(vals) < n:
                raise ValueError(f"record of {n} values expected, got {len(vals)}: {ln!r}")
            vals = vals[:n]          # a record's surplus values are skipped
        return vals

    def ints(self, n=None):
        return [int(v) for v in self.read(n)]


# m of each value of the count record of numerical-basis inputs
# (qmc/vmc/read_orb_loc.f90:241-245): s; p 1, -1, 0; d 0, 2, -2, 1, -1;
# f 3, -3, 2, -2, 1, -1, 0; g 4, -4, ..., 0; h 5, -5, ..., 0.  The basis
# functions are listed in the same order (orb_loc_ana_original_order,
# read_orb_loc.f90:449-840).
_NUMR1_LM = ([(0, 0)] + [(1, m) for m in (1, -1, 0)] + [(2, m) for m in (0, 2, -2, 1, -1)] +
             [(l, m) for l in (3, 4, 5) for m in
              [v for mm in range(l, 0, -1) for v in (mm, -mm)] + [0]])


def read_gauss_pseudopotential(text: str, lpotp1: int):
    """One pseudopot<ict> file in the format of qmc/vmc/readps_gauss.f90:
    a label line, npotd, then npotd blocks `nterms` + nterms rows `coef power
    exponent`; the first block is the local component (stored at lpotp1),
    block l >= 2 is component l - 1.  Returns (npotd, lpotp1, components),
    components[i] the (coef, power, exponent) terms of CHAMP's component
    i + 1 (an l = i nonlocal channel unless i + 1 == lpotp1)."""
    rows = [ln.split() for ln in text.splitlines()[1:] if ln.strip()]
    npotd = int(rows[0][0])
    if lpotp1 < 0:
        lpotp1 = npotd
    if lpotp1 != npotd:
        # the file lists the local component first and the l = 0, 1, ...
        # channels after it, which fills components 1..npotd only when the
        # local one is the last (readps_gauss.f90:46-53)
        raise ValueError("the local component must be the last one (lpotp1 = npotd)")
    if npotd > QMCW_MAX_LPOT:
        raise ValueError(f"more than {QMCW_MAX_LPOT} pseudopotential components")
    comps = [[] for _ in range(npotd)]
    i = 1
    for l in range(1, npotd + 1):
        idx = lpotp1 if l == 1 else l - 1
        nt = int(rows[i][0])
        i += 1
        for _ in range(nt):
            c, pw, ex = rows[i][:3]
            comps[idx - 1].append((float(c.lower().replace("d", "e")), int(pw),
                                   float(ex.lower().replace("d", "e"))))
            i += 1
    return npotd, lpotp1, comps


def _gauss_ecp_arrays(per_type):
    """per_type: (npotd, lpotp1, components) of each centre type."""
    npotd, lpotp1, off, coef, power, expo = [], [], [0], [], [], []
    for (nd, lp, comps) in per_type:
        npotd.append(nd)
        lpotp1.append(lp)
        for i in range(QMCW_MAX_LPOT):
            for (cf, pw, ex) in (comps[i] if i < nd else []):
                coef.append(cf)
                power.append(pw)
                expo.append(ex)
            off.append(len(coef))
    return (np.array(npotd, np.int32), np.array(lpotp1, np.int32),
            np.array(off, np.int32), np.array(coef), np.array(power, np.int32),
            np.array(expo))


def from_champ_old_input(text: str, name: str = "champ", files: dict | None = None) -> System:
    """Build a System from CHAMP's old line-oriented input (read_input.f90;
    e.g. examples/c2_ktt_hfj_vmc_emin_j_opt3_lin.inp and
    examples/c2_bfd_vtz_rhf_vmc.inp).

    Supported: finite systems, all-electron (nloc = 0) or with Gaussian
    pseudopotentials (nloc = 6, readps_gauss files pseudopot<ict>, nquad 6
    or 12); the analytic Slater basis of s and p functions in the order of
    orb_loc_ana_gamess_order (numr = -1: all s by n, then the px, py and pz
    blocks by n; qmc/vmc/read_orb_loc.f90:843-1000) or numerical radial
    functions (numr = 1: files basis.<ict>, count record and function order
    of orb_loc_ana_original_order, functions up to h, the radial function of
    each from the iwrwf record); orbitals given by coefficients,
    determinants/CSFs, ijas = 4 Jastrow (isc = 2 or 4) and, for one centre
    type, the Jastrow parameters the optimisation section varies (iwjasa,
    iwjasb, iwjasc).  The dependent e-e-n coefficients are re-imposed with
    cuspexact4.  ``files`` maps file names (basis.1, pseudopot1, ...) to
    their text.
    """
    from . import numbas
    files = files or {}
    R = _OldReader(text.splitlines())
    R.seek("iperiodic,ibasis")
    R.i -= 1
    iperiodic, ibasis = R.ints(2)
    if iperiodic != 0 or ibasis != 1:
        raise ValueError("finite systems with a localized basis (iperiodic 0, ibasis 1) only")
    R.seek("nfprod,tau")
    nloc, numr = R.ints(4)[:2]
    if nloc not in (0, 6):
        raise ValueError("nloc = 0 (all-electron) or 6 (Gaussian pseudopotential) only")
    if numr not in (-1, 1):
        raise ValueError("numr = -1 (analytic basis in the GAMESS order) or 1 (numerical) only")
    nquad = 12
    if nloc == 6:
        nquad = R.ints(1)[0]
        if nquad not in (6, 12):
            raise ValueError("nquad = 6 or 12 only")
    nelec, nup = R.ints(2)
    R.seek("Geometry section")
    if R.ints(1)[0] != 3:
        raise ValueError("ndim = 3 only")
    nctype, ncent = R.ints(2)
    iwctype = np.array(R.ints(ncent), np.int32) - 1
    znuc = np.array(R.read(nctype))
    lpotp1 = R.ints(nctype) if nloc > 0 else None
    cent = np.array([R.read(3) for _ in range(ncent)])
    R.seek("Determinantal section")
    R.read()                                          # inum_orb
    ndet, nbasis, norb = R.ints(3)
    funcs, iwrwf = [], []
    for _ in range(nctype):
        v = R.ints()
        if numr == 1:
            if len(v) < len(_NUMR1_LM):
                raise ValueError("numerical-basis count record too short")
            f = []
            for (l, m), cnt in zip(_NUMR1_LM, v):
                if cnt and l > QMCW_MAX_L:
                    raise ValueError(f"basis functions with l > {QMCW_MAX_L} are not supported")
                f += [(0, l, m)] * abs(cnt)
            funcs.append(f)
            iwrwf.append([k - 1 for k in R.ints(len(f))])
            continue
        if len(v) < 18:
            raise ValueError("basis count record too short")
        if any(c < 0 for c in v):
            raise ValueError("gaussian (negative) basis counts are not supported")
        if any(v[9:14]) or any(v[18:]):
            raise ValueError("only s and p basis functions are supported in the old format")
        # n1s, n2s, n2p(x,y,z), n3s, n3p(x,y,z), n3d(5), n4s, n4p(x,y,z)
        cn = dict(s={1: v[0], 2: v[1], 3: v[5], 4: v[14]},
                  p={m: {2: v[2 + k], 3: v[6 + k], 4: v[15 + k]} for k, m in enumerate((1, -1, 0))})
        f = []
        for n in (1, 2, 3, 4):
            f += [(n, 0, 0)] * cn["s"][n]
        for m in (1, -1, 0):
            for n in (2, 3, 4):
                f += [(n, 1, m)] * cn["p"][m][n]
        funcs.append(f)
    coef = np.array([R.read(nbasis) for _ in range(norb)])
    zex = R.read(nbasis)
    if sum(len(funcs[t]) for t in iwctype) != nbasis:
        raise ValueError("basis counts do not add up to nbasis")
    bb = _BasisBuilder(QMCW_BASIS_SLATER, nctype)
    first = {}
    ib = 0
    for ic, t in enumerate(iwctype):
        z = zex[ib:ib + len(funcs[t])]
        ib += len(funcs[t])
        if t in first:
            if not np.allclose(first[t], z, rtol=0, atol=0):
                raise ValueError("centres of one type must share their exponents")
            continue
        first[t] = z
        if numr == 1:
            # numerical radial functions: zex 0 (read_orb_loc.f90:428-436),
            # or 1 in inputs written before that convention
            if any(zz not in (0.0, 1.0) for zz in z):
                raise ValueError("mixed analytic and numerical radial functions are not supported")
            fname = f"basis.{t + 1}"
            if fname not in files:
                raise ValueError(f"numerical basis file {fname} not given")
            bf = numbas.read_basis_file(files[fname], nloc)
            tabs = numbas.radial_tables(bf, nloc)
            rad = {}
            for (_, l, m), k in zip(funcs[t], iwrwf[t]):
                if not 0 <= k < len(tabs):
                    raise ValueError(f"{fname}: radial function {k + 1} not in the file")
                if k not in rad:
                    rad[k] = bb.add_numeric_radial(bb.add_table(tabs[k]))
                bb.add_function(t, rad[k], l, m)
            continue
        for (n, l, m), zz in zip(funcs[t], z):
            bb.add_function(t, bb.add_radial(n, [float(zz)], [1.0]), l, m)
    barr, anorm = bb.arrays(iwctype)
    coef = coef * anorm[None, :]
    dets = [[x - 1 for x in R.ints(nelec)] for _ in range(ndet)]
    det_up = [d[:nup] for d in dets]
    det_dn = [d[nup:] for d in dets]
    ncsf = R.ints(1)[0]
    csf_coef = R.read(ncsf)
    ndic = R.ints(ncsf)
    csfs = []
    for k in range(ncsf):
        ds = [x - 1 for x in R.ints(ndic[k])]
        cs = R.read(ndic[k])
        csfs.append((csf_coef[k], list(zip(ds, cs))))
    cc, off, cdet, cdc = _csf_arrays(csfs, False)
    det_up, det_dn, cdc = sort_dets(det_up, det_dn, cdet, cdc)
    R.seek("Jastrow section")
    R.read()                                          # ianalyt_lap
    ijas, isc, nspin1, nspin2 = R.ints(6)[:4]
    if ijas != 4:
        raise ValueError("only ijas=4 is supported")
    norda, nordb, nordc = R.ints(3)
    scalek = R.read(2)[0]
    nspin2b = nspin2 - nspin1 + 1
    a4 = np.array([R.read(norda + 1) for _ in range(nctype)])
    b = np.array([R.read(nordb + 1) for _ in range(nspin2b)])
    c = np.array([R.read(nterms4(nordc)) for _ in range(nctype)])
    c = cusp_exact(c, nordc)
    # the Jastrow parameters the optimisation section varies (iwjasa, iwjasb,
    # iwjasc; one centre type)
    jas = {}
    for key in ("iwjasa", "iwjasb", "iwjasc"):
        for ln in R.lines[R.i:]:
            if "(" + key in ln:
                jas[key] = [int(v) for v in R._numeric(ln)]
                break
    ja = [(0, i) for i in jas.get("iwjasa", [])] if nctype == 1 else []
    jb = [(0, i) for i in jas.get("iwjasb", [])] if nspin2b == 1 else []
    jc = [(0, i) for i in jas.get("iwjasc", [])] if nctype == 1 else []
    psp = {}
    if nloc == 6:
        per = []
        for t in range(nctype):
            fname = f"pseudopot{t + 1}"
            if fname not in files:
                raise ValueError(f"pseudopotential file {fname} not given")
            per.append(read_gauss_pseudopotential(files[fname], lpotp1[t]))
        ecp = _gauss_ecp_arrays(per)
        psp = dict(npotd=ecp[0], lpotp1=ecp[1], ecp_off=ecp[2], ecp_coef=ecp[3],
                   ecp_power=ecp[4], ecp_expo=ecp[5], nquad=nquad)
    return System(nelec=nelec, nup=nup, cent=cent, iwctype=iwctype, znuc=znuc,
                  coef=coef, det_up=det_up, det_dn=det_dn, csf_coef=cc,
                  csf_det_off=off, csf_det=cdet, csf_det_coef=cdc, isc=isc,
                  nspin2b=nspin2b, norda=norda, nordb=nordb, nordc=nordc,
                  scalek=scalek, a4=a4, b=b, c=c, nloc=nloc, jas_a=ja, jas_b=jb, jas_c=jc,
                  name=name, **psp, **barr)


def jastrow_opt_list(sysm: System, vary_a1: bool = False):
    """Jastrow parameters CHAMP varies by default for ijas=4 (a3.., b2..,
    independent c's), qmc/optimization_mod.f90:520-570."""
    ja = [(t, i) for t in range(sysm.nctype)
          for i in range(1 if vary_a1 else 3, sysm.norda + 2)]
    jb = [(s, i) for s in range(sysm.nspin2b) for i in range(2, sysm.nordb + 2)]
    _, iwc4 = cusp_matrix(sysm.nordc)
    dep = set(iwc4)
    jc = [(t, i) for t in range(sysm.nctype)
          for i in range(1, sysm.nterms + 1) if i not in dep]
    return ja, jb, jc


# ---------------------------------------------------------------------------
# synthetic workloads
# ---------------------------------------------------------------------------
def _jastrow_random(rng, nctype, norda, nordb, nordc, sd, a1=0.0, nspin2b=1):
    a4 = rng.normal(0, sd, (nctype, norda + 1))
    a4[:, 0] = a1
    a4[:, 1] = np.abs(a4[:, 1]) + 0.5     # keep 1 + a2 r > 0
    b = rng.normal(0, sd, (nspin2b, nordb + 1))
    b[:, 0] = 0.5                          # e-e cusp
    b[:, 1] = np.abs(b[:, 1]) + 1.0
    c = rng.normal(0, sd, (nctype, nterms4(nordc)))
    c = cusp_exact(c, nordc) if nordc > 1 else c
    return a4, b, c


def be_atom(seed: int = 1) -> System:
    """All-electron Be-like atom: Z=4 at the origin, 2+2 electrons, one
    determinant of 2 orbitals per spin over 10 STOs (1s x2, 2s x2, 2p x6) with
    exponents U[0.8, 6.0]; MOs a seeded Gram-Schmidt-orthonormal matrix;
    Jastrow e-n, e-e (Pade, kappa = 1) and 4th-order e-e-n with coefficients
    N(0, 0.05)."""
    rng = np.random.default_rng(seed)
    bb = _BasisBuilder(QMCW_BASIS_SLATER, 1)
    for n in (1, 1, 2, 2):
        r = bb.add_radial(n, [rng.uniform(0.8, 6.0)], [1.0])
        bb.add_function(0, r, 0, 0)
    for _ in range(2):
        r = bb.add_radial(2, [rng.uniform(0.8, 6.0)], [1.0])
        for m in (1, -1, 0):
            bb.add_function(0, r, 1, m)
    iwctype = np.array([0], np.int32)
    barr, anorm = bb.arrays(iwctype)
    q = _random_orthonormal(rng, 10)
    coef = q.T * anorm[None, :]
    a4, b, c = _jastrow_random(rng, 1, 4, 4, 4, 0.05)
    cc, off, cdet, cdc = _csf_arrays([(1.0, [(0, 1.0)])], False)
    return System(nelec=4, nup=2, cent=np.zeros((1, 3)), iwctype=iwctype,
                  znuc=np.array([4.0]), coef=coef, det_up=np.array([[0, 1]], np.int32),
                  det_dn=np.array([[0, 1]], np.int32), csf_coef=cc, csf_det_off=off,
                  csf_det=cdet, csf_det_coef=cdc, isc=4, scalek=1.0, norda=4,
                  nordb=4, nordc=4, a4=a4, b=b, c=c, name="be_atom", **barr)


# BFD-like Gaussian pseudopotential for Z_eff = 4 (local channel and s channel
# as in examples/pseudopot1; the p channel is a synthetic Gaussian term).
_ECP_C = {
    "local": [(4.0, 1, 8.35973821), (33.43895285, 3, 4.48361888), (-19.17537323, 2, 3.93831258)],
    "l0": [(22.55164191, 2, 5.02991637)],
    "l1": [(9.65, 2, 4.21)],
}


def _ecp_arrays(types):
    """types: list of None (no PP, local -Z/r only) or the _ECP_C dict.
    Component order per type: l=0, l=1, local (lpotp1 = 3)."""
    npotd, lpotp1, off, coef, power, expo = [], [], [0], [], [], []
    for t in types:
        comps = [[]] * QMCW_MAX_LPOT
        if t is None:
            npotd.append(1)
            lpotp1.append(1)
            comps = [[] for _ in range(QMCW_MAX_LPOT)]
        else:
            npotd.append(3)
            lpotp1.append(3)
            comps = [t["l0"], t["l1"], t["local"]] + [[] for _ in range(QMCW_MAX_LPOT - 3)]
        for lst in comps:
            for (cf, pw, ex) in lst:
                coef.append(cf)
                power.append(pw)
                expo.append(ex)
            off.append(len(coef))
    return (np.array(npotd, np.int32), np.array(lpotp1, np.int32),
            np.array(off, np.int32), np.array(coef), np.array(power, np.int32),
            np.array(expo))


def _gto_shells(bb, rng, ictype, shells, nprim=2):
    """Add contracted spherical GTO shells ((l, count) pairs, l <= 5) with
    exponents log-uniform in [0.1, 20]."""
    for l, count in shells:
        for _ in range(count):
            z = np.exp(rng.uniform(np.log(0.1), np.log(20.0), nprim))
            d = rng.uniform(0.3, 1.0, nprim)
            r = bb.add_radial(l + 1, list(z), list(d))
            ms = [0] + [s * mm for mm in range(1, l + 1) for s in (1, -1)]
            for m in ms:
                bb.add_function(ictype, r, l, m)


def c2_dimer(seed: int = 2, derivs: bool = False) -> System:
    """C2-like dimer: two Z_eff=4 centres at +-1.17 bohr on z, 4+4 valence
    electrons, 58 contracted spherical GTOs (3s3p2d1f per atom), Gaussian
    pseudopotential with l=0,1 nonlocal channels (12-point quadrature),
    24 CSFs over 48 determinants with CSF coefficients N(0, 0.1) (first = 1).
    With ``derivs`` the ~250 parameters of the linear method are listed:
    Jastrow, 23 CSFs and ~190 occupied-virtual orbital rotations."""
    rng = np.random.default_rng(seed)
    bb = _BasisBuilder(QMCW_BASIS_GAUSSIAN, 1)
    _gto_shells(bb, rng, 0, [(0, 3), (1, 3), (2, 2), (3, 1)])
    iwctype = np.array([0, 0], np.int32)
    barr, anorm = bb.arrays(iwctype)
    nb = len(anorm)
    q = _random_orthonormal(rng, nb)
    coef = q.T * anorm[None, :]
    ref = [0, 1, 2, 3]
    # 24 CSFs: reference, 22 singlet singles (2 dets), one 3-det double
    csfs, dets = [], []

    def det_index(up, dn):
        key = (tuple(sorted(up)), tuple(sorted(dn)))
        for i, d in enumerate(dets):
            if d == key:
                return i
        dets.append(key)
        return len(dets) - 1

    csfs.append((1.0, [(det_index(ref, ref), 1.0)]))
    singles = [(i, a) for i in (3, 2, 1, 0) for a in range(4, 10)][:22]
    for (i, a) in singles:
        exc = [a if x == i else x for x in ref]
        csfs.append((rng.normal(0, 0.1), [(det_index(exc, ref), 1.0), (det_index(ref, exc), 1.0)]))
    d1 = [4 if x == 3 else x for x in ref]
    d2 = [5 if x == 3 else x for x in ref]
    csfs.append((rng.normal(0, 0.1), [(det_index(d1, d1), 1.0), (det_index(d1, d2), -1.0),
                                       (det_index(d2, d1), -1.0)]))
    cc, off, cdet, cdc = _csf_arrays(csfs, True)
    det_up = [list(d[0]) for d in dets]
    det_dn = [list(d[1]) for d in dets]
    det_up, det_dn, cdc = sort_dets(det_up, det_dn, cdet, cdc)
    norda, nordb, nordc = 5, 5, 5
    sd = 0.02 if derivs else 0.05
    a4, b, c = _jastrow_random(rng, 1, norda, nordb, nordc, sd, a1=rng.normal(0, sd))
    ecp = _ecp_arrays([_ECP_C])
    s = System(nelec=8, nup=4, cent=np.array([[0, 0, -1.17], [0, 0, 1.17]]),
               iwctype=iwctype, znuc=np.array([4.0]), coef=coef,
               det_up=det_up, det_dn=det_dn, csf_coef=cc, csf_det_off=off,
               csf_det=cdet, csf_det_coef=cdc, isc=4, scalek=0.8,
               norda=norda, nordb=nordb, nordc=nordc, a4=a4, b=b, c=c, nloc=6,
               npotd=ecp[0], lpotp1=ecp[1], ecp_off=ecp[2], ecp_coef=ecp[3],
               ecp_power=ecp[4], ecp_expo=ecp[5], nquad=12, name="c2_dimer", **barr)
    if derivs:
        ja, jb, jc = jastrow_opt_list(s, vary_a1=True)
        # extend to 40 Jastrow parameters with the optional e-n a2 term
        ja = [(0, 2)] + ja
        s.jas_a, s.jas_b, s.jas_c = ja, jb, jc
        s.csf_parm = list(range(1, len(cc)))
        occ = sorted(set(x for d in dets for x in d[0]) | set(x for d in dets for x in d[1]))
        virt = [o for o in range(nb) if o not in occ]
        s.orb_parm = [(i, a, 0) for i in ref for a in virt]
        s.name = "c2_dimer_derivs"
    return s


def hchain(seed: int = 4, ncarbon: int = 16) -> System:
    """Chain of Z_eff=4 centres spaced 2.6 bohr along z with two Z=1 centres
    +-2.0 bohr off-axis per carbon gap (34 for 16 carbons); 98 electrons,
    424 contracted GTOs (C: 4s3p1d, H: 1s1p); pseudopotential on the Z_eff=4
    centres only; one determinant; e-n, e-e, e-e-n Jastrow."""
    rng = np.random.default_rng(seed)
    cent, types = [], []
    for k in range(ncarbon):
        cent.append([0.0, 0.0, 2.6 * k])
        types.append(0)
    nh = 34 if ncarbon == 16 else 2 * ncarbon + 2
    for k in range(nh):
        zc = 2.6 * (k // 2) * (ncarbon - 1) / max(1, (nh // 2 - 1)) if nh > 2 else 0.0
        s = 2.0 if k % 2 == 0 else -2.0
        cent.append([s, 0.0, zc])
        types.append(1)
    bb = _BasisBuilder(QMCW_BASIS_GAUSSIAN, 2)
    _gto_shells(bb, rng, 0, [(0, 4), (1, 3), (2, 1)])
    _gto_shells(bb, rng, 1, [(0, 1), (1, 1)])
    iwctype = np.array(types, np.int32)
    barr, anorm = bb.arrays(iwctype)
    nb = len(anorm)
    nelec = 4 * ncarbon + nh
    nup = nelec // 2
    q = _random_orthonormal(rng, nb, nup + 8)
    coef = q.T * anorm[None, :]
    cc, off, cdet, cdc = _csf_arrays([(1.0, [(0, 1.0)])], False)
    a4, b, c = _jastrow_random(rng, 2, 4, 4, 4, 0.03)
    ecp = _ecp_arrays([_ECP_C, None])
    return System(nelec=nelec, nup=nup, cent=np.array(cent), iwctype=iwctype,
                  znuc=np.array([4.0, 1.0]), coef=coef,
                  det_up=np.arange(nup, dtype=np.int32)[None, :],
                  det_dn=np.arange(nelec - nup, dtype=np.int32)[None, :],
                  csf_coef=cc, csf_det_off=off, csf_det=cdet, csf_det_coef=cdc,
                  isc=4, scalek=0.8, norda=4, nordb=4, nordc=4, a4=a4, b=b, c=c,
                  nloc=6, npotd=ecp[0], lpotp1=ecp[1], ecp_off=ecp[2],
                  ecp_coef=ecp[3], ecp_power=ecp[4], ecp_expo=ecp[5], nquad=12,
                  name="hchain", **barr)


def random_molecule(seed: int, nelec: int | None = None, ncent: int | None = None,
                    zeff: float | None = None, big_basis: bool = False,
                    gh_shells: bool = False) -> System:
    """Synthetic molecule of the multi-system batch: electron count uniform
    in [4, 128] (split evenly between spins, the extra one up), 1-24 nuclei
    with Z_eff in {1, 4, 6} placed by random packing at >= 2 bohr spacing,
    contracted GTOs (Z=1: 2s1p, Z_eff=4: 3s2p1d, Z_eff=6: 3s3p2d1f),
    pseudopotentials on the Z_eff > 1 centres, random orthonormal MOs, one
    determinant, e-n / e-e / e-e-n Jastrow.  Charged systems are allowed:
    the electron count is drawn independently of the nuclei, then capped so
    that the basis holds the occupied orbitals of each spin plus two.
    zeff fixes every nucleus' charge; big_basis gives the Z_eff=6 centres
    4s3p3d1f (35 functions), so that 23 of them reach ~800 functions;
    gh_shells adds one g and one h shell to every centre type."""
    rng = np.random.default_rng(seed)
    N = int(nelec if nelec is not None else rng.integers(4, 129))
    M = int(ncent if ncent is not None else rng.integers(1, 25))
    zs = rng.choice([1.0, 4.0, 6.0], size=M)
    if zeff is not None:
        zs[:] = float(zeff)
    if not (zs > 1).any():
        zs[0] = 4.0
    cent = []
    box = 2.0 * (M ** (1.0 / 3.0)) + 1.0
    while len(cent) < M:
        c = rng.uniform(-box, box, 3)
        if all(np.linalg.norm(c - q) >= 2.0 for q in cent):
            cent.append(c)
        else:
            box *= 1.01
    types = {1.0: 0, 4.0: 1, 6.0: 2}
    iwctype = np.array([types[z] for z in zs], np.int32)
    bb = _BasisBuilder(QMCW_BASIS_GAUSSIAN, 3)
    _gto_shells(bb, rng, 0, [(0, 2), (1, 1)])
    _gto_shells(bb, rng, 1, [(0, 3), (1, 2), (2, 1)])
    _gto_shells(bb, rng, 2, [(0, 4), (1, 3), (2, 3), (3, 1)] if big_basis else
                [(0, 3), (1, 3), (2, 2), (3, 1)])
    if gh_shells:
        for t in range(3):
            _gto_shells(bb, rng, t, [(4, 1), (5, 1)])
    barr, anorm = bb.arrays(iwctype)
    nb = len(anorm)
    N = max(2, min(N, 2 * (nb - 2)))
    nup = (N + 1) // 2
    norb = min(nb, nup + 4)
    q = _random_orthonormal(rng, nb, norb)
    coef = q.T * anorm[None, :]
    cc, off, cdet, cdc = _csf_arrays([(1.0, [(0, 1.0)])], False)
    a4, b, c = _jastrow_random(rng, 3, 4, 4, 4, 0.03)
    ecp6 = dict(local=[(c_ * 1.5, p_, e_) for c_, p_, e_ in _ECP_C["local"]],
                l0=[(c_ * 1.5, p_, e_ * 1.2) for c_, p_, e_ in _ECP_C["l0"]],
                l1=[(c_ * 1.5, p_, e_ * 1.2) for c_, p_, e_ in _ECP_C["l1"]])
    ecp = _ecp_arrays([None, _ECP_C, ecp6])
    return System(nelec=N, nup=nup, cent=np.array(cent), iwctype=iwctype,
                  znuc=np.array([1.0, 4.0, 6.0]), coef=coef,
                  det_up=np.arange(nup, dtype=np.int32)[None, :],
                  det_dn=np.arange(N - nup, dtype=np.int32)[None, :],
                  csf_coef=cc, csf_det_off=off, csf_det=cdet, csf_det_coef=cdc,
                  isc=4, scalek=0.8, norda=4, nordb=4, nordc=4, a4=a4, b=b, c=c,
                  nloc=6, npotd=ecp[0], lpotp1=ecp[1], ecp_off=ecp[2],
                  ecp_coef=ecp[3], ecp_power=ecp[4], ecp_expo=ecp[5], nquad=12,
                  name=f"mol{seed}", **barr)


def molecule_derivs(s: System) -> System:
    """The linear method's parameters of a single-determinant molecule: the
    Jastrow set CHAMP varies (jastrow_opt_list, a1 included) and the
    rotations from the four highest up-spin occupied orbitals into the
    virtual orbitals the MO set holds (four at most for random_molecule)."""
    ja, jb, jc = jastrow_opt_list(s, vary_a1=True)
    s.jas_a, s.jas_b, s.jas_c = ja, jb, jc
    nup, norb = int(s.nup), int(np.asarray(s.coef).shape[0])
    s.orb_parm = [(i, a, 0) for i in range(max(0, nup - 4), nup) for a in range(nup, norb)]
    s.name = f"{s.name}_derivs"
    return s


def random_walkers(sysm: System, nw: int, seed: int = 7, sigma: float = 1.0):
    """Electrons Gaussian (sigma bohr) around the nuclei, assigned round-robin."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, sigma, (nw, sysm.nelec, 3))
    for i in range(sysm.nelec):
        x[:, i, :] += sysm.cent[i % sysm.ncent][None, :]
    return x
