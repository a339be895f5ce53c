"""Per-walker streams of CHAMP's rannyu generator (qmc/vmc/rannyu.f90).

rannyu is the multiplicative congruential generator R <- 11^13 R mod 2^48
held in four 12-bit limbs.  Walker w gets the master stream advanced by
``(w + 1) * stride`` draws (jump-ahead by modular exponentiation), so that
the walkers' streams are disjoint segments of CHAMP's own sequence.
"""
from __future__ import annotations

import numpy as np

MULT = 34522712143931          # 11**13 = 502*8^12 + 1521*8^8 + 4071*8^4 + 2107
MOD = 1 << 48


def seed_limbs(seed16: str | int):
    """CHAMP's 16-digit seed read as four 4-digit limbs
    (qmc/control_mod.f90:328), last limb made odd (rannyu.f90:11-21)."""
    s = str(seed16).zfill(16)
    limbs = [int(s[0:4]), int(s[4:8]), int(s[8:12]), int(s[12:16])]
    limbs[3] = 2 * (limbs[3] // 2) + 1
    return limbs


def limbs_to_int(l):
    return ((l[0] * 4096 + l[1]) * 4096 + l[2]) * 4096 + l[3]


def int_to_limbs(r):
    return [(r >> 36) & 4095, (r >> 24) & 4095, (r >> 12) & 4095, r & 4095]


def jump(limbs, n: int):
    """State after n draws."""
    return int_to_limbs(limbs_to_int(limbs) * pow(MULT, n, MOD) % MOD)


def walker_seeds(seed16, nwalk: int, stride: int = 1 << 30) -> np.ndarray:
    """[nwalk][4] int32 limb states, one disjoint stream per walker."""
    base = seed_limbs(seed16)
    r0 = limbs_to_int(base)
    step = pow(MULT, stride, MOD)
    out = np.empty((nwalk, 4), np.int32)
    r = r0
    for w in range(nwalk):
        r = r * step % MOD
        out[w] = int_to_limbs(r)
    return out


def rannyu_sequence(limbs, n: int):
    """n draws and the final limbs (pure Python restatement, for checks)."""
    r = limbs_to_int(limbs)
    out = []
    for _ in range(n):
        r = r * MULT % MOD
        out.append(r / MOD)
    return out, int_to_limbs(r)
