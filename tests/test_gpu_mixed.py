"""Mixed-precision engines against the FP64 oracle on identical walker
configurations (sampled by a few FP64 VMC steps from random positions), for
the C2-like dimer, the 98-electron chain and CHAMP's pseudopotential C2
example (tabulated radial basis).

QMCW_MIXED evaluates the basis functions and the Jastrow in FP32 and keeps
the orbital accumulation, determinants, inverses, E_L and reductions in
FP64.  Stated bounds (measured maxima over 4,096 C2 and 256 chain walkers
were 5.7e-6, 2.6e-5, 2.3e-4 Ha, 6.5e-6, 1.4e-5, 2.7e-3, 2.9e-6 and 8.8e-6):
  |d ln|Psi_D||       <= 5e-5
  |d ln J|            <= 2e-4
  |d E_L|             <= 2e-3 Ha per walker, |mean d E_L| <= 0.1 mHa
  |d grad ln Psi|     <= 5e-5 x the walker's largest gradient component
  |d lap| (divv)      <= 1e-4 x the walker's largest |divv|
  one-electron ratio, new gradient and laplacian: relative error <= 1e-2
      (near nodes the ratio is small and its relative error largest), median <= 1e-5
  d ln Psi/dp, dE_L/dp: <= 3e-5 and 1e-4 x the parameter's largest |value|

QMCW_MIXED_BF16 rounds the pseudopotential projection (the value-only
quadrature rows) to bfloat16 on top of that; the Slater-matrix rows are not
rounded (bfloat16 there gave local-energy errors of tens of Ha near nodes).
Everything but the pseudopotential terms equals the mixed engine's, and
|d E_L| <= 0.5 Ha (measured maxima 0.14 Ha on C2, 0.09 Ha on the chain; the
mean over walkers cannot be shown within 0.1 mHa, so this mode is not the
reported one).
"""
import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu

def _c2_bfd():
    import os, sys
    sys.path.insert(0, os.path.dirname(__file__))
    from transcripts import c2_bfd
    return c2_bfd(3)[0]


CASES = {"c2": (lambda: system.c2_dimer(2), 512, 0.6),
         "chain": (lambda: system.hchain(4), 48, 0.3),
         "c2_bfd": (_c2_bfd, 512, 0.6)}


def _configs(s, W, delta):
    e = Engine(s, W, "fp64")
    e.set_walkers(system.random_walkers(s, W, seed=5))
    e.set_rng(rng.walker_seeds("1837465927472523", W))
    e.vmc_device(10, delta)
    e.sync()
    x = e.get_walkers()
    e.set_walkers(x)
    e.psi_at()
    return e, x


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_psi_at_within_stated_bounds(name):
    mk, W, delta = CASES[name]
    s = mk()
    e64, x = _configs(s, W, delta)
    r = oracle.Oracle(s).eval_batch(x, nthreads=8)
    o9 = r["out9"]
    res = {}
    for prec in ("mixed", "mixed_bf16"):
        e = Engine(s, W, prec)
        e.set_walkers(x)
        g = e.psi_at()
        res[prec] = g
        assert np.abs(g["logpsid"] - o9[:, 0]).max() <= 5e-5
        assert np.array_equal(g["sign"], o9[:, 1])
        assert np.abs(g["psij"] - o9[:, 2]).max() <= 2e-4
        gs = np.abs(r["grad"]).max(axis=(1, 2))[:, None, None]
        assert (np.abs(g["grad"] - r["grad"]) / gs).max() <= 5e-5
        ds = np.abs(r["divv"]).max(axis=1)[:, None]
        assert (np.abs(g["divv"] - r["divv"]) / ds).max() <= 1e-4
        de = g["eloc"] - o9[:, 3]
        if prec == "mixed":
            assert np.abs(de).max() <= 2e-3
            assert abs(de.mean()) <= 1e-4
        else:
            assert np.abs(de).max() <= 0.5
        e.close()
    for k in ("logpsid", "psij", "grad", "divv"):
        assert np.array_equal(res["mixed"][k], res["mixed_bf16"][k])


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_one_electron_moves_within_stated_bounds(name):
    mk, W, delta = CASES[name]
    s = mk()
    e64, x = _configs(s, W, delta)
    em = Engine(s, W, "mixed")
    em.set_walkers(x)
    em.psi_at()
    rs = np.random.default_rng(3)
    rel = {"ratio": [], "grad": [], "lap": []}
    for ie in range(0, s.nelec, max(1, s.nelec // 8)):
        xn = x[:, ie, :] + rs.normal(0, delta, (W, 3))
        a, b = em.move_1e(ie, xn), e64.move_1e(ie, xn)
        rel["ratio"].append(np.abs(a["ratio"] - b["ratio"]) / np.abs(b["ratio"]))
        rel["grad"].append(np.abs(a["grad"] - b["grad"]).max(axis=1) / np.abs(b["grad"]).max(axis=1))
        rel["lap"].append(np.abs(a["lap"] - b["lap"]) / np.maximum(np.abs(b["lap"]), 1.0))
        acc = (rs.random(W) < 0.5).astype(np.uint8)
        em.copy_psi_1e(ie, acc)
        e64.copy_psi_1e(ie, acc)
    for k, v in rel.items():
        v = np.concatenate(v)
        assert v.max() <= 1e-2, k
        assert np.median(v) <= 1e-5, k


def test_mixed_derivatives_within_stated_bounds():
    s = system.c2_dimer(2, derivs=True)
    W = 256
    e64, x = _configs(s, W, 0.6)
    d64, l64 = e64.derivs()
    for prec in ("mixed", "mixed_bf16"):
        e = Engine(s, W, prec)
        e.set_walkers(x)
        d, l = e.derivs()
        assert (np.abs(d - d64) / (np.abs(d64).max(axis=0) + 1e-12)).max() <= 3e-5
        bound = 1e-4 if prec == "mixed" else 5e-3
        assert (np.abs(l - l64) / (np.abs(l64).max(axis=0) + 1e-12)).max() <= bound
        e.close()


@pytest.mark.parametrize("name", ["c2", "c2_bfd"])
def test_mixed_walk_decisions_against_fp64(name):
    """The mixed engine decides each move from its own ratio, so its walk can
    leave the FP64 walk where u falls between the two ratios.  From identical
    configurations and random streams, one sweep (a decision per electron)
    of 2,048 walkers: the mixed and FP64 walks agree in at least 99.9% of
    the decisions, and where the walks still agree the positions equal to
    1e-12 bohr.  (Measured over ten sweeps of 4,096 walkers, 327,680
    decisions each: no walk left the FP64 walk on the C2-like system, two
    on the pseudopotential C2 example.)"""
    mk, _, delta = CASES[name]
    s = mk()
    W = 2048
    e64, x = _configs(s, W, delta)
    seeds = rng.walker_seeds("5555666677778889", W)
    out = {}
    for prec in ("fp64", "mixed"):
        e = Engine(s, W, prec)
        e.set_walkers(x)
        e.set_rng(seeds)
        v = e.vmc(1, delta=delta)
        out[prec] = (v["naccept"], e.get_walkers(), e.get_rng())
        e.close()
    n64, x64, r64 = out["fp64"]
    nmx, xmx, rmx = out["mixed"]
    same = (n64 == nmx) & np.all(np.abs(x64 - xmx) <= 1e-12, axis=(1, 2))
    ndiff = int((~same).sum())
    assert ndiff <= max(1, int(1e-3 * W * s.nelec)), ndiff
    assert np.array_equal(r64[same], rmx[same])
