"""Walker configurations in CHAMP's forms.

* ``sites``: the starting configuration CHAMP generates when there is no
  configuration file (qmc/vmc/mc_configs.f90:64-89 distributes the electrons
  over the centres, qmc/vmc/sites.f90:63-159 places them), drawing from
  CHAMP's rannyu stream.
* ``mc_configs_start``: one configuration, 3N numbers in free format
  (mc_configs.f90:60-62 reads it, :122-125 writes it).
* configuration files for optimisation and DMC: one line per sampled
  configuration, the 3N coordinates followed by the sign of the determinant
  part, ln|Psi|, the local energy, Psi_D and ln J (qmc/vmc/vmc.f90:273-281,
  Fortran format ``(3N f13.8, i3, es12.4, f12.5, 2es12.4)``).
"""
from __future__ import annotations

import math

import numpy as np

from . import rng as _rng


def electron_sites(s):
    """(nsite, nsiteup): electrons per centre and up electrons per centre
    (mc_configs.f90:64-83): int(Z + 0.5) per centre, trimmed to nelec, the
    rest on the first centre; up electrons in proportion, the remainder
    handed out round-robin from the first centre."""
    M, nelec, nup = s.ncent, s.nelec, s.nup
    nsite = [0] * M
    l = 0
    for i in range(M):
        nsite[i] = int(s.znuc[s.iwctype[i]] + 0.5)
        l += nsite[i]
        if l > nelec:
            nsite[i] -= l - nelec
            l = nelec
    if l < nelec:
        nsite[0] += nelec - l
    nsiteup = [(nup * nsite[i]) // nelec for i in range(M)]
    count = sum(nsiteup)
    i = 0
    while count < nup:
        nsiteup[i] += 1
        count += 1
        i = (i + 1) % M
    return nsite, nsiteup


def sites(s, seed):
    """CHAMP's starting configuration of an atom or molecule (sites.f90,
    atoms and molecules) from rannyu state ``seed`` (four limbs, e.g.
    ``rng.seed_limbs("1837465927472523")``): up electrons centre by centre,
    then down electrons; the j-th electron of a centre with charge Z at
    scale 1/max(Z, 1) (j = 1), 2/max(Z-2, 1) (j <= 5), 3/max(Z-10, 1)
    (j <= 9), 4/max(Z-18, 1) (j <= 18), else 5/max(Z-36, 1), each
    coordinate offset by +-scale * (-ln u1) with the sign of u2 - 1/2.
    Returns (x [N][3], the rannyu state after the draws)."""
    nsite, nsiteup = electron_sites(s)
    r = _rng.limbs_to_int(list(seed))

    def rannyu():
        nonlocal r
        r = r * _rng.MULT % _rng.MOD
        return r / _rng.MOD

    x = np.zeros((s.nelec, 3))
    ie = 0
    for spin in (0, 1):
        for i in range(s.ncent):
            z = float(s.znuc[s.iwctype[i]])
            if z == 0.0:
                raise ValueError("sites needs nonzero nuclear charges")
            ju = nsiteup[i] if spin == 0 else nsite[i] - nsiteup[i]
            for j in range(1, ju + 1):
                if ie >= s.nelec:
                    return x, _rng.int_to_limbs(r)
                if j == 1:
                    sc = 1 / max(z, 1.0)
                elif j <= 5:
                    sc = 2 / max(z - 2, 1.0)
                elif j <= 9:
                    sc = 3 / max(z - 10, 1.0)
                elif j <= 18:
                    sc = 4 / max(z - 18, 1.0)
                else:
                    sc = 5 / max(z - 36, 1.0)
                for k in range(3):
                    site = -math.log(rannyu())
                    site = math.copysign(site, rannyu() - 0.5)
                    x[ie, k] = sc * site + s.cent[i][k]
                ie += 1
    if ie < s.nelec:
        raise ValueError("bad input to sites: fewer sites than electrons")
    return x, _rng.int_to_limbs(r)


def read_mc_configs_start(text: str, nelec: int) -> np.ndarray:
    """The configuration of an mc_configs_start file, [N][3]."""
    vals = [float(t.lower().replace("d", "e")) for t in text.split()]
    if len(vals) < 3 * nelec:
        raise ValueError(f"mc_configs_start holds {len(vals)} numbers, {3 * nelec} needed")
    return np.array(vals[:3 * nelec]).reshape(nelec, 3)


def write_mc_configs_start(x) -> str:
    """One configuration as mc_configs_start holds it (free format)."""
    return " ".join(f"{v: .16E}" for v in np.asarray(x, np.float64).reshape(-1)) + "\n"


def _es(v: float, w: int = 12, d: int = 4) -> str:
    """Fortran ESw.d for exponents of two digits."""
    return f"{v:{w}.{d}E}"


def write_mc_configs(x, sign, logpsi, eloc, psid, psij) -> str:
    """Configuration-file lines (vmc.f90:276-279): per configuration the 3N
    coordinates (f13.8), sign (i3), ln|Psi| (es12.4), E_L (f12.5), Psi_D and
    ln J (es12.4)."""
    x = np.asarray(x, np.float64)
    lines = []
    for c in range(x.shape[0]):
        coords = "".join(f"{v:13.8f}" for v in x[c].reshape(-1))
        lines.append(coords + f"{int(sign[c]):3d}" + _es(logpsi[c]) + f"{eloc[c]:12.5f}" +
                     _es(psid[c]) + _es(psij[c]))
    return "\n".join(lines) + ("\n" if lines else "")


def read_mc_configs(text: str, nelec: int):
    """The configurations of a configuration file, [nconf][N][3], and the
    columns after them when present (sign, ln|Psi|, E_L, Psi_D, ln J).  The
    fixed-width coordinates are read by their width (f13.8), so that values
    that fill the field still separate."""
    xs, extra = [], []
    w = 13
    for ln in text.splitlines():
        if not ln.strip():
            continue
        head = ln[:w * 3 * nelec]
        coords = None
        if len(head) == w * 3 * nelec:
            try:
                coords = [float(head[k * w:(k + 1) * w]) for k in range(3 * nelec)]
                rest = ln[w * 3 * nelec:].split()
            except ValueError:
                coords = None
        if coords is None:   # free format
            toks = ln.split()
            coords = [float(t.lower().replace("d", "e")) for t in toks[:3 * nelec]]
            rest = toks[3 * nelec:]
        if len(coords) < 3 * nelec:
            raise ValueError("configuration line with fewer than 3N coordinates")
        xs.append(np.array(coords).reshape(nelec, 3))
        extra.append([float(t.lower().replace("d", "e")) for t in rest])
    cols = None
    if extra and all(len(e) >= 5 for e in extra):
        e = np.array([r[:5] for r in extra])
        cols = dict(sign=e[:, 0].astype(int), logpsi=e[:, 1], eloc=e[:, 2], psid=e[:, 3], psij=e[:, 4])
    return np.array(xs), cols
