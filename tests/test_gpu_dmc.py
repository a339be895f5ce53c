"""GPU parity of DMC generations (qmcw_dmc_step) against the oracle's DMC on
the same population, random streams and branching stream.

Both movers of sample_1e_move are covered: drift-diffusion for the
pseudopotential system and drift-diffusion with the nucleus cusp for the
all-electron one.  Per generation: accepted-move and node-crossing counts and the population
after split_join are identical; population sums, weights and positions agree
to 1e-9 relative / 1e-11 absolute (the gaussian proposals go through cos and
log, whose last bits may differ between the GPU and the host library);
random streams are identical.
"""
import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.abi import DmcParams
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _setup(s, n0, cap, seed):
    x = system.random_walkers(s, cap, seed=seed)
    seeds = np.asarray(rng.walker_seeds("1837465927472523", cap), np.int32)
    return x, seeds


# T-moves act on pseudopotentials, so the all-electron atom runs without them.
_CASES = ([("c2_dimer", t, m) for t, m in [(0.02, 0), (0.1, 0), (0.1, 1), (0.1, 2), (0.3, 1)]]
          + [("be_atom", t, 0) for t in (0.02, 0.1)] + [("c2_bfd", 0.1, 0), ("c2_bfd", 0.1, 1)])


def _system(name):
    if name == "c2_bfd":
        # CHAMP's pseudopotential C2 example: tabulated radial basis, 6-point quadrature
        import os, sys
        sys.path.insert(0, os.path.dirname(__file__))
        from transcripts import c2_bfd
        return c2_bfd(3)[0]
    return system.c2_dimer(2) if name == "c2_dimer" else system.be_atom(1)


@pytest.mark.parametrize("name,tau,tmoves", _CASES)
def test_dmc_generations_match_oracle(name, tau, tmoves):
    """c2_dimer: pseudopotential mover, with T-moves (1: det_balance1, 2:
    casula1) before the moves; be_atom: all-electron mover with the nucleus
    cusp (gaussian or exponential branch per proposal); c2_bfd: CHAMP's
    pseudopotential example with and without T-moves."""
    s = _system(name)
    n0, cap = 20, 48
    x, seeds = _setup(s, n0, cap, seed=31)
    O = oracle.Oracle(s)
    e0 = float(O.eval_batch(x[:n0])["out9"][:, 3].mean())
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    brng[3] |= 1
    e = Engine(s, cap)
    e.set_walkers(x)
    e.set_rng(seeds)
    e.set_branch_rng(brng)
    e.dmc_begin(n0)
    d = oracle.OracleDMC(s, cap)
    assert d.set(x[:n0], seeds[:n0]) == 0
    bo = brng.copy()
    p = DmcParams(tau, e0, eest=e0, esigma=1.0, tmoves=tmoves)
    for gen in range(4):
        sg = e.dmc_step(p)
        r, so = d.generation(p)
        assert r == 0
        r, sb = d.branch(bo)
        assert r == 0
        assert sg.nacc == so.nacc and sg.ncross == so.ncross and sg.ntry == so.ntry
        assert abs(sg.wsum - so.wsum) <= 1e-9 * abs(so.wsum)
        assert abs(sg.wesum - so.wesum) <= 1e-9 * abs(so.wesum)
        assert sg.nwalk == sb.nwalk and sg.nbranch == sb.nbranch
        xg, wg, rg = e.dmc_population()
        go = d.get()
        assert np.array_equal(rg, go["rng"])
        assert np.allclose(wg, go["weight"], rtol=1e-9, atol=0)
        assert np.allclose(xg, go["x"], rtol=0, atol=1e-11)
        assert np.array_equal(e.get_branch_rng(), bo)


@pytest.mark.parametrize("name,tmoves", [("c2_dimer", 1), ("be_atom", 0)])
def test_dmc_five_launch_path_matches_oracle(monkeypatch, name, tmoves):
    """The DMC move as separate launches per electron (gradient, proposal,
    ratio, decision, commit; the path of the three-launch move) gives the
    generations of the fused k_ratio_1e mode."""
    monkeypatch.setenv("QMCW_DMC_UNFUSED", "1")
    test_dmc_generations_match_oracle(name, 0.1, tmoves)


class _OracleStepper:
    def __init__(self, d, brng):
        self.d, self.brng = d, brng

    def dmc_step(self, params):
        r, st = self.d.generation(params)
        assert r == 0
        r, sb = self.d.branch(self.brng)
        assert r == 0
        st.nwalk, st.nbranch = sb.nwalk, sb.nbranch
        return st


def test_dmc_driver_run_matches_oracle():
    """Six generations under the rundmc driver (population control, running
    eest/taueff) on the GPU and on the oracle: same populations, energies to
    1e-9."""
    from champ_amd.dmc import DmcDriver
    s = system.c2_dimer(2)
    n0, cap = 24, 64
    x, seeds = _setup(s, n0, cap, seed=33)
    e0 = float(oracle.Oracle(s).eval_batch(x[:n0])["out9"][:, 3].mean())
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    brng[3] |= 1
    e = Engine(s, cap)
    e.set_walkers(x)
    e.set_rng(seeds)
    e.set_branch_rng(brng)
    e.dmc_begin(n0)
    d = oracle.OracleDMC(s, cap)
    assert d.set(x[:n0], seeds[:n0]) == 0
    dg = DmcDriver(e, s.nelec, n0, tau=0.05, etrial=e0)
    do = DmcDriver(_OracleStepper(d, brng.copy()), s.nelec, n0, tau=0.05, etrial=e0)
    for _ in range(6):
        sg, so = dg.generation(), do.generation()
        assert sg.nwalk == so.nwalk and sg.nacc == so.nacc
        assert abs(sg.wesum - so.wesum) <= 1e-9 * abs(so.wesum)
    assert abs(dg.eest - do.eest) <= 1e-9 * abs(do.eest)


def test_gpu_dmc_recovers_the_exact_helium_energy():
    """The GPU engine's all-electron DMC of He (1 + 1 electrons, nucleus-cusp
    mover) under the rundmc driver: the exact -2.903724 Ha within 4 sigma
    (or 5 mHa) at tau = 0.01."""
    import os, sys
    sys.path.insert(0, os.path.dirname(__file__))
    from test_oracle_dmc import he_atom
    from champ_amd.dmc import DmcDriver
    s = he_atom()
    W0, cap = 1000, 2500
    e = Engine(s, cap)
    e.set_walkers(system.random_walkers(s, cap, seed=7, sigma=0.6))
    e.set_rng(np.asarray(rng.walker_seeds("1837465927472523", cap), np.int32))
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    brng[3] |= 1
    e.set_branch_rng(brng)
    e.dmc_begin(W0)
    drv = DmcDriver(e, 2, W0, tau=0.01, etrial=-2.9)
    drv.run(200)
    en = np.array([g["eave"] for g in drv.run(400)])
    e.close()
    blocks = en.reshape(20, 20).mean(axis=1)
    est, err = blocks.mean(), blocks.std() / np.sqrt(len(blocks))
    assert abs(est + 2.903724) < max(4 * err, 0.005), (est, err)


def test_dmc_walkers_that_move_nothing_match_oracle():
    """At tau = 3 many proposals fail, so some walkers move no electron in a
    generation (counted on the oracle) and, after branching, sit in scratch
    slots other walkers used before.  Their evaluation draws no grid
    rotations and keeps the local energy of the generation's start; counts,
    streams, weights and positions still match the oracle."""
    s = system.c2_dimer(2)
    n0, cap = 40, 96
    x, seeds = _setup(s, n0, cap, seed=31)
    e0 = float(oracle.Oracle(s).eval_batch(x[:n0])["out9"][:, 3].mean())
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    brng[3] |= 1
    e = Engine(s, cap)
    e.set_walkers(x)
    e.set_rng(seeds)
    e.set_branch_rng(brng)
    e.dmc_begin(n0)
    d = oracle.OracleDMC(s, cap)
    assert d.set(x[:n0], seeds[:n0]) == 0
    bo = brng.copy()
    p = DmcParams(3.0, e0, eest=e0, esigma=1.0, tmoves=0)
    unmoved = 0
    for gen in range(4):
        x0 = d.get()["x"]
        sg = e.dmc_step(p)
        r, so = d.generation(p)
        assert r == 0
        unmoved += int(np.all(x0 == d.get()["x"], axis=(1, 2)).sum())
        r, sb = d.branch(bo)
        assert r == 0
        assert sg.nacc == so.nacc and sg.ntry == so.ntry and sg.nwalk == sb.nwalk
        assert abs(sg.wesum - so.wesum) <= 1e-9 * abs(so.wesum)
        xg, wg, rg = e.dmc_population()
        go = d.get()
        assert np.array_equal(rg, go["rng"])
        assert np.allclose(wg, go["weight"], rtol=1e-9, atol=0)
        assert np.allclose(xg, go["x"], rtol=0, atol=1e-11)
    assert unmoved > 0


def test_tmove_record_overflow_is_reported(monkeypatch):
    """An electron in range of more pseudopotential centres than the T-move
    record holds is reported as an error, not dropped: QMCW_TM_KMAX=1 makes
    the record hold one centre while the dimer's electrons between the
    nuclei are in range of two."""
    from champ_amd.engine import QmcwError
    monkeypatch.setenv("QMCW_TM_KMAX", "1")
    s = system.c2_dimer(2)
    n0, cap = 20, 48
    x, seeds = _setup(s, n0, cap, seed=31)
    x[:, :, :] *= 0.2                      # every electron near the bond centre
    e = Engine(s, cap)
    e.set_walkers(x)
    e.set_rng(seeds)
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    brng[3] |= 1
    e.set_branch_rng(brng)
    with pytest.raises(QmcwError, match="T-move record"):
        e.dmc_begin(n0)                    # the first DMC-mode evaluation keeps the points
        e.dmc_step(DmcParams(0.1, -11.0, eest=-11.0, esigma=1.0, tmoves=1))


def test_chain98_dmc_generations_match_oracle():
    """The 98-electron chain (pseudopotential mover, 424 GTOs): a small
    population (6 walkers, capacity 12), 2 generations at tau = 0.01 on the
    GPU and in the oracle: counts, streams and population identical, weights
    to 1e-9 relative, positions to 1e-10 bohr."""
    s = system.hchain(4)
    n0, cap = 6, 12
    x, seeds = _setup(s, n0, cap, seed=43)
    e0 = float(oracle.Oracle(s).eval_batch(x[:n0], nthreads=8)["out9"][:, 3].mean())
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    brng[3] |= 1
    e = Engine(s, cap)
    e.set_walkers(x)
    e.set_rng(seeds)
    e.set_branch_rng(brng)
    e.dmc_begin(n0)
    d = oracle.OracleDMC(s, cap)
    assert d.set(x[:n0], seeds[:n0]) == 0
    bo = brng.copy()
    p = DmcParams(0.01, e0, eest=e0, esigma=1.0, tmoves=0)
    for gen in range(2):
        sg = e.dmc_step(p)
        r, so = d.generation(p, nthreads=8)
        assert r == 0
        r, sb = d.branch(bo)
        assert r == 0
        assert sg.nacc == so.nacc and sg.ncross == so.ncross and sg.ntry == so.ntry
        assert abs(sg.wsum - so.wsum) <= 1e-9 * abs(so.wsum)
        assert sg.nwalk == sb.nwalk and sg.nbranch == sb.nbranch
        xg, wg, rg = e.dmc_population()
        go = d.get()
        assert np.array_equal(rg, go["rng"])
        assert np.allclose(wg, go["weight"], rtol=1e-9, atol=0)
        assert np.allclose(xg, go["x"], rtol=0, atol=1e-10)
        assert np.array_equal(e.get_branch_rng(), bo)
