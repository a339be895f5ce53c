"""GPU sums of the linear method's derivative products (qmcw_accum_*, FP64
MFMA) against the oracle's sums (orc_accum) of the same derivative samples:
1e-12 relative (the sums are reassociated), accumulation over several calls,
reset, and the sampling mode that accumulates every VMC step."""
import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, float(np.abs(b).max()))
    return float(np.abs(a - b).max()) / scale < tol


@pytest.mark.parametrize("W", [13, 700])   # one partial chunk; two chunks, the second ragged
def test_accum_matches_oracle(W):
    s = system.c2_dimer(2, derivs=True)
    e = Engine(s, W)
    e.set_walkers(system.random_walkers(s, W, seed=71))
    dp, dl = e.derivs()
    el = e.get_eloc()
    e.accum_reset()
    e.accum_add()
    g = e.accum_get()
    o = oracle.accum(dp, dl, el)
    assert g["n"] == W
    for k in ("se", "se2", "sd", "sde", "sl", "dd", "dde", "dl"):
        assert _close(g[k], o[k], 1e-12), k
    # a second add doubles every sum; reset clears them
    e.accum_add()
    g2 = e.accum_get()
    for k in ("n", "sd", "dd", "dde", "dl"):
        assert _close(g2[k], 2 * np.asarray(g[k]), 1e-12), k
    e.accum_reset()
    g0 = e.accum_get()
    assert g0["n"] == 0 and not np.any(g0["dd"]) and not np.any(g0["dl"])
    e.close()


def test_sampling_mode_accumulates_every_step():
    s = system.c2_dimer(2, derivs=True)
    W = 64
    e = Engine(s, W)
    e.set_walkers(system.random_walkers(s, W, seed=72))
    e.set_rng(rng.walker_seeds("1837465927472523", W))
    e.accum_reset()
    e.sample_derivs(2)
    tot = None
    for step in range(3):
        e.vmc(1, delta=0.6)
        dp, dl = e.get_derivs()
        o = oracle.accum(dp, dl, e.get_eloc())
        tot = o if tot is None else {k: tot[k] + o[k] for k in tot}
    g = e.accum_get()
    assert g["n"] == 3 * W
    for k in ("se", "sd", "sde", "sl", "dd", "dde", "dl"):
        assert _close(g[k], tot[k], 1e-12), k
    e.close()
