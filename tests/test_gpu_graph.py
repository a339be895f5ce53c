"""The device VMC driver replays one captured HIP graph per step
(qmcw_vmc_run_device).  Replays, interleaved with eager work, give the same
walk bit for bit as the eager driver (qmcw_vmc_run) from the same state, and
a parameter change rebuilds the graph."""
import numpy as np
import pytest

from champ_amd import rng, system
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _pair(s, W, seed):
    x = system.random_walkers(s, W, seed=seed)
    seeds = rng.walker_seeds("1837465927472523", W)
    a, b = Engine(s, W), Engine(s, W)
    for e in (a, b):
        e.set_walkers(x)
        e.set_rng(seeds)
    return a, b


def test_graph_replays_match_eager_steps():
    s = system.c2_dimer(2)
    a, b = _pair(s, 24, 51)
    for rep in range(3):
        a.vmc_device(2, 0.6)          # graph: two replays
        b.vmc(2, 0.6)                 # eager, traced
        assert np.array_equal(a.get_walkers(), b.get_walkers())
        assert np.array_equal(a.get_rng(), b.get_rng())
        ea, eb = a.psi_at()["eloc"], b.psi_at()["eloc"]   # eager work between replays
        assert np.array_equal(ea, eb)


def test_graph_rebuilt_after_parameter_change():
    s = system.c2_dimer(2)
    a, b = _pair(s, 16, 52)
    a.vmc_device(1, 0.6)
    b.vmc(1, 0.6)
    b2 = s.b.copy()
    b2[0, 1] += 0.05              # Pade denominator: changes the Jastrow asymptote
    for e in (a, b):
        e.set_params(b=b2)
    a.vmc_device(2, 0.6)
    b.vmc(2, 0.6)
    assert np.array_equal(a.get_walkers(), b.get_walkers())
    assert np.array_equal(a.get_eloc(), b.get_eloc())
