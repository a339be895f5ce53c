"""Uneven multi-system batch (the 4-128-electron, 10-800-function molecules
of the batch workload), run two ways: one handle and HIP stream per system
with every system's VMC steps queued before any is awaited, and as one
batch (qmcw_batch_*) whose VMC step runs each kernel once over the walkers
of all systems.  Each system must reproduce the oracle's walk bit for bit
(positions, streams) with local energies to 1e-9, and the largest (128
electrons, 23 centres, 805 GTOs) also its full evaluation."""
import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu


def _rel(a, b, floor):
    a, b = np.asarray(a), np.asarray(b)
    return float((np.abs(a - b) / (np.abs(b) + floor)).max())


def test_concurrent_uneven_systems_match_oracle():
    systems = [system.random_molecule(301, nelec=7, ncent=2),
               system.random_molecule(302, nelec=48, ncent=10),
               system.random_molecule(303, nelec=128, ncent=23, zeff=6.0, big_basis=True)]
    assert systems[2].nbasis >= 800
    W, nst, delta = 3, 2, 0.3
    runs = []
    for k, s in enumerate(systems):
        x = system.random_walkers(s, W, seed=50 + k)
        seeds = np.asarray(rng.walker_seeds("1837465927472523", W, stride=(1 << 20) + k), np.int32)
        e = Engine(s, W)
        e.set_walkers(x)
        e.set_rng(seeds)
        runs.append((s, x, seeds, e))
    big = runs[2]
    g = big[3].psi_at()
    r = oracle.Oracle(big[0]).eval_batch(big[1], nthreads=8)
    assert _rel(g["logpsid"], r["out9"][:, 0], 1e-10) < 1e-10
    assert _rel(g["eloc"], r["out9"][:, 3], 1e-8) < 1e-9
    assert _rel(g["grad"], r["grad"], 1e-10) < 1e-9
    for s, x, seeds, e in runs:             # queued on every stream before any wait
        e.vmc_device(nst, delta)
    for s, x, seeds, e in runs:
        e.sync()
    for s, x, seeds, e in runs:
        o = oracle.Oracle(s).vmc_run(x, seeds, nst, delta, nthreads=8)
        assert np.array_equal(e.get_walkers(), o["x"]), s.name
        assert np.array_equal(e.get_rng(), o["rng"]), s.name
        assert _rel(e.get_eloc(), o["eloc"][-1], 1e-8) < 1e-9, s.name
        e.close()


@pytest.mark.parametrize("sweep", ["0", "1"], ids=["launch_per_electron", "sweep_per_launch"])
def test_one_launch_batch_matches_oracle(monkeypatch, sweep):
    """Four systems of 4 to 128 electrons in one batch: Gaussian bases with
    pseudopotentials and the all-electron Be atom with its Slater basis and
    no pseudopotential; two VMC steps (each a sweep of up to 128
    one-electron launches over all systems, then the full evaluation)
    against the oracle's walk of each system, at W = 3, 5, 2 and 4; the
    moves either one launch per electron index over all systems, or each
    system's whole sweep in one launch (QMCW_BATCH_SWEEP=1)."""
    monkeypatch.setenv("QMCW_BATCH_SWEEP", sweep)
    systems = [system.random_molecule(311, nelec=7, ncent=2),
               system.random_molecule(312, nelec=48, ncent=10),
               system.be_atom(1),
               system.random_molecule(314, nelec=128, ncent=23, zeff=6.0, big_basis=True)]
    Ws = [3, 5, 2, 4]
    nst, delta = 2, 0.3
    runs = []
    for k, (s, W) in enumerate(zip(systems, Ws)):
        x = system.random_walkers(s, W, seed=70 + k)
        seeds = np.asarray(rng.walker_seeds("1837465927472523", W, stride=(1 << 20) + k), np.int32)
        e = Engine(s, W)
        e.set_walkers(x)
        e.set_rng(seeds)
        runs.append((s, x, seeds, e))
    b = Batch([r[3] for r in runs])
    b.vmc_device(nst, delta)
    for s, x, seeds, e in runs:
        e.sync()
    b.close()
    for s, x, seeds, e in runs:
        o = oracle.Oracle(s).vmc_run(x, seeds, nst, delta, nthreads=8)
        assert np.array_equal(e.get_walkers(), o["x"]), s.name
        assert np.array_equal(e.get_rng(), o["rng"]), s.name
        assert _rel(e.get_eloc(), o["eloc"][-1], 1e-8) < 1e-9, s.name
        e.close()


def test_concurrent_systems_with_derivatives_on_half():
    """The batch workload's mix: concurrent handles, the derivatives and
    the linear method's product sums on for every second system (its
    Jastrow set and orbital rotations).  Sampling leaves each walk as the
    oracle's, bit for bit; the derivative systems' last-step d ln Psi/dp
    match the oracle's at the final positions and their dE_L/dp a fresh
    evaluation of the same state."""
    systems = [system.random_molecule(311, nelec=9, ncent=3),
               system.molecule_derivs(system.random_molecule(1011)),
               system.random_molecule(313, nelec=40, ncent=6),
               system.molecule_derivs(system.random_molecule(314, nelec=30, ncent=5))]
    W, nst, delta = 3, 2, 0.3
    runs = []
    for k, s in enumerate(systems):
        x = system.random_walkers(s, W, seed=60 + k)
        seeds = np.asarray(rng.walker_seeds("1837465927472523", W, stride=(1 << 21) + k), np.int32)
        e = Engine(s, W)
        e.set_walkers(x)
        e.set_rng(seeds)
        if s.nparm:
            e.accum_reset()
            e.sample_derivs(2)
        runs.append((s, x, seeds, e))
    for s, x, seeds, e in runs:
        e.vmc_device(nst, delta)
    for s, x, seeds, e in runs:
        e.sync()
    for s, x, seeds, e in runs:
        o = oracle.Oracle(s).vmc_run(x, seeds, nst, delta, nthreads=8)
        assert np.array_equal(e.get_walkers(), o["x"]), s.name
        assert np.array_equal(e.get_rng(), o["rng"]), s.name
        assert _rel(e.get_eloc(), o["eloc"][-1], 1e-8) < 1e-9, s.name
        if s.nparm:
            dp, de = e.get_derivs()
            r = oracle.Oracle(s).eval_batch(o["x"], derivs=True)
            assert r["status"] == 0
            assert _rel(dp, r["dpsi"], 1e-9) < 1e-9, s.name
            # dE_L/dp was formed on the last step's rotated quadrature grid:
            # a fresh derivative evaluation of the same state gives it again
            dp2, de2 = e.derivs()
            assert np.array_equal(dp, dp2) and np.array_equal(de, de2), s.name
            assert e.accum_get()["n"] == W * nst
        e.close()
