"""The alternative code paths of the engine agree with the oracle as the
default ones do: global scratch instead of LDS staging (basis values, the
pseudopotential point scratch, eval_y's inputs), electron chunks in global
scratch in the full evaluation, the one-electron move as three launches, and
full evaluations that recompute the occupied orbitals instead of keeping the
rows the moves committed, and the move's e-e-n terms summed term by term
per pair instead of through the moved electron's per-centre tables.
Large systems take some of these paths by size; here they are forced on the
C2-like system through the engine's environment switches, read when a
handle is created."""
import os

import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu

S = system.c2_dimer(2)


def _rel(a, b, floor):
    a, b = np.asarray(a), np.asarray(b)
    return float((np.abs(a - b) / (np.abs(b) + floor)).max())


@pytest.fixture
def env(monkeypatch):
    def set_(**kv):
        for k, v in kv.items():
            monkeypatch.setenv(k, v)
    return set_


@pytest.mark.parametrize("switches", [
    {"QMCW_NO_LDS": "1"},
    {"QMCW_DET_LDS": "0", "QMCW_CHUNK": "3"},
    {"QMCW_NO_LDS_Y": "1"},
    {"QMCW_SPLIT_MOVE": "1"},
    {"QMCW_SPLIT_MOVE": "1", "QMCW_NO_LDS": "1"},
    {"QMCW_NO_KEEP_ORB": "1"},
    {"QMCW_NO_EEN_TAB": "1"},
], ids=["no_lds", "det_chunks", "no_lds_y", "split_move", "split_no_lds", "recompute_orbitals",
        "een_per_pair"])
def test_alternative_paths_match_oracle(env, switches):
    env(**switches)
    W = 17
    x = system.random_walkers(S, W, seed=61)
    e = Engine(S, W)
    e.set_walkers(x)
    g = e.psi_at()
    r = oracle.Oracle(S).eval_batch(x)
    o9 = r["out9"]
    assert _rel(g["logpsid"], o9[:, 0], 1e-10) < 1e-10
    assert _rel(g["eloc"], o9[:, 3], 1e-8) < 1e-9
    assert _rel(g["grad"], r["grad"], 1e-10) < 1e-9
    seeds = rng.walker_seeds("1837465927472523", W)
    e.set_rng(seeds)
    v = e.vmc(2, delta=0.6)
    o = oracle.Oracle(S).vmc_run(x, seeds, 2, 0.6)
    assert np.array_equal(v["naccept"], o["naccept"])
    assert np.array_equal(e.get_walkers(), o["x"])
    assert _rel(v["eloc"], o["eloc"], 1e-8) < 1e-9
    e.close()
