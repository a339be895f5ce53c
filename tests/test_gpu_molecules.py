"""GPU parity on synthetic molecules of the multi-system batch (uneven
electron counts, H centres without and Z_eff = 4, 6 centres with
pseudopotentials, s-f shells, and g and h shells in the last case): full
evaluation and one VMC step against the oracle."""
import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _rel(a, b, floor):
    a, b = np.asarray(a), np.asarray(b)
    return float((np.abs(a - b) / (np.abs(b) + floor)).max())


@pytest.mark.parametrize("seed,nelec,ncent,gh", [(101, 5, 1, False), (102, 12, 3, False),
                                                 (103, 24, 6, False), (104, 37, 9, False),
                                                 (105, 14, 3, True)])
def test_molecule_parity(seed, nelec, ncent, gh):
    s = system.random_molecule(seed, nelec=nelec, ncent=ncent, gh_shells=gh)
    if gh:
        assert 5 in set(int(v) for v in s.bas_l)
    W = 9
    x = system.random_walkers(s, W, seed=seed)
    e = Engine(s, W)
    e.set_walkers(x)
    g = e.psi_at()
    r = oracle.Oracle(s).eval_batch(x)
    o9 = r["out9"]
    assert _rel(g["logpsid"], o9[:, 0], 1e-10) < 1e-10
    assert _rel(g["psij"], o9[:, 2], 1e-10) < 1e-10
    assert _rel(g["eloc"], o9[:, 3], 1e-8) < 1e-9
    assert _rel(g["grad"], r["grad"], 1e-10) < 1e-9
    seeds = rng.walker_seeds("1837465927472523", W)
    e.set_rng(seeds)
    v = e.vmc(2, delta=0.4)
    o = oracle.Oracle(s).vmc_run(x, seeds, 2, 0.4)
    assert np.array_equal(v["naccept"], o["naccept"])
    assert np.array_equal(e.get_walkers(), o["x"])
    assert _rel(v["eloc"], o["eloc"], 1e-8) < 1e-9


@pytest.mark.parametrize("which", ["c2", "mol5", "mol1"])
def test_projection_on_mfma_matches_valu(monkeypatch, which):
    """The pseudopotential projection gp = G^T coef_occ on FP64 MFMA
    (k_gproj_mfma: rows = (walker, electron) pairs of many walkers, 16 walkers
    of a 4-electron spin per 64-row block) against the VALU loop of
    k_full_det (QMCW_NO_GPROJ_MFMA=1), on walker counts that leave the last
    block partly empty, spins of 4 + 4, 3 + 2 and 1 + 1 electrons: the local
    energies (whose nonlocal part is linear in gp) agree to 1e-12 relative
    and equal the oracle's to 1e-9."""
    s = {"c2": lambda: system.c2_dimer(),
         "mol5": lambda: system.random_molecule(101, nelec=5, ncent=1),
         "mol1": lambda: system.random_molecule(106, nelec=1, ncent=2)}[which]()
    assert s.nloc > 0
    W = 37
    x = system.random_walkers(s, W, seed=44)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("QMCW_NO_GPROJ_MFMA", flag)
        e = Engine(s, W)
        e.set_walkers(x)
        out[flag] = e.psi_at()
        e.close()
    assert _rel(out["0"]["eloc"], out["1"]["eloc"], 1e-8) < 1e-12
    assert np.array_equal(out["0"]["logpsid"], out["1"]["logpsid"])
    r = oracle.Oracle(s).eval_batch(x)
    assert _rel(out["0"]["eloc"], r["out9"][:, 3], 1e-8) < 1e-9
