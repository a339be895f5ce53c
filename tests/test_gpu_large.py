"""GPU parity at the size of the 98-electron chain (16 Z_eff=4 centres with a
pseudopotential, 34 H, 424 GTOs, one determinant): full evaluation against
the oracle and one VMC step with bit-exact decisions, at two lane-group
widths.  Few walkers: the oracle needs ~0.4 s per walker evaluation."""
import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu

S = system.hchain(4)


def _rel(a, b, floor):
    a, b = np.asarray(a), np.asarray(b)
    return float((np.abs(a - b) / (np.abs(b) + floor)).max())


@pytest.mark.parametrize("G", [32, 64])
def test_chain98_psi_at_and_vmc_step(G):
    W = 5
    x = system.random_walkers(S, W, seed=41)
    e = Engine(S, W)
    e.set_group_width(G)
    e.set_walkers(x)
    g = e.psi_at()
    r = oracle.Oracle(S).eval_batch(x)
    o9 = r["out9"]
    assert _rel(g["logpsid"], o9[:, 0], 1e-10) < 1e-10
    assert _rel(g["psij"], o9[:, 2], 1e-10) < 1e-10
    assert _rel(g["eloc"], o9[:, 3], 1e-8) < 1e-9
    assert _rel(g["grad"], r["grad"], 1e-10) < 1e-9
    seeds = rng.walker_seeds("1837465927472523", W)
    e.set_rng(seeds)
    v = e.vmc(1, delta=0.3)
    o = oracle.Oracle(S).vmc_run(x, seeds, 1, 0.3)
    assert np.array_equal(v["naccept"], o["naccept"])
    assert np.array_equal(e.get_walkers(), o["x"])
    assert _rel(v["eloc"], o["eloc"], 1e-8) < 1e-9


def test_chain98_projection_on_mfma_matches_valu(monkeypatch):
    """The pseudopotential projection gp = G^T coef_occ of the chain's
    49 x 49 determinants on FP64 MFMA (k_gproj_mfma, the default; a block of
    64 (walker, electron) rows holds one walker and part of the next here)
    against the VALU loop of
    k_full_det (QMCW_NO_GPROJ_MFMA=1): the local energies (whose nonlocal
    part is linear in gp) agree to 1e-12 relative."""
    W = 5
    x = system.random_walkers(S, W, seed=42)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("QMCW_NO_GPROJ_MFMA", flag)
        e = Engine(S, W)
        e.set_walkers(x)
        out[flag] = e.psi_at()
        e.close()
    assert _rel(out["0"]["eloc"], out["1"]["eloc"], 1e-8) < 1e-12
    assert np.array_equal(out["0"]["logpsid"], out["1"]["logpsid"])


@pytest.mark.parametrize("mode", ["vmc", "dmc"])
def test_single_determinant_move_equals_the_general_path(monkeypatch, mode):
    """One-determinant systems take a shorter move (A^-1 only: T, the alpha
    blocks and Y stay as the full evaluation left them).  It leaves every
    observable unchanged: against the general path (QMCW_NO_SINGLE_DET=1) the
    walk, its random streams, weights and local energies are bit-identical."""
    from champ_amd.abi import DmcParams
    W = 24
    x = system.random_walkers(S, W, seed=43)
    seeds = rng.walker_seeds("1837465927472523", W)
    out = []
    for general in (False, True):
        if general:
            monkeypatch.setenv("QMCW_NO_SINGLE_DET", "1")
        e = Engine(S, W, "mixed")
        e.set_walkers(x)
        e.set_rng(seeds)
        if mode == "vmc":
            v = e.vmc(3, delta=0.3)
            out.append((v["naccept"], v["eloc"], e.get_walkers(), e.get_rng()))
        else:
            brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
            brng[3] |= 1
            e.set_branch_rng(brng)
            e0 = float(np.mean(e.psi_at()["eloc"][:16]))
            e.dmc_begin(16)
            p = DmcParams(0.01, e0, eest=e0, esigma=1.0)
            for _ in range(2):
                e.dmc_step(p)
            xg, wg, rg = e.dmc_population()
            out.append((xg, wg, rg, e.get_eloc()))
        e.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("tables", ["1", "0"])
def test_chain98_one_electron_ratios_match_oracle(monkeypatch, tables):
    """One-electron moves of the chain against the oracle's ratio_1e: the
    move's e-e-n terms of the 97 pairs x 50 centres through the moved
    electron's per-centre tables (the default) and summed term by term
    (QMCW_NO_EEN_TAB=1); ratios, gradients and laplacians to 1e-9."""
    monkeypatch.setenv("QMCW_NO_EEN_TAB", "0" if tables == "1" else "1")
    W = 4
    x = system.random_walkers(S, W, seed=45)
    e = Engine(S, W)
    e.set_group_width(64)
    e.set_walkers(x)
    e.psi_at()
    O = oracle.Oracle(S)
    ws = [O.L.orc_walker_create(O.h) for _ in range(W)]
    for w in range(W):
        O.eval_full(x[w], walker=ws[w])
    rs = np.random.default_rng(7)
    for ie in (0, 37, 61, 97):
        xn = x[:, ie, :] + rs.normal(0, 0.3, (W, 3))
        g = e.move_1e(ie, xn)
        ref = [O.ratio_1e(ie, xn[w], walker=ws[w]) for w in range(W)]
        assert _rel(g["ratio"], [r["ratio"] for r in ref], 1e-12) < 1e-9
        assert _rel(g["grad"], [r["grad"] for r in ref], 1e-8) < 1e-9
        assert _rel(g["lap"], [r["lap"] for r in ref], 1e-8) < 1e-8
    for w in ws:
        O.L.orc_walker_destroy(w)
