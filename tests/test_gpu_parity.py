"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle on
identical walker configurations and identical rannyu streams.

Tolerances (FP64 engine): log|Psi_D|, ln J, E_L, gradients and laplacians
agree to 1e-10 relative (1e-9 absolute on E_L, whose terms cancel); VMC
accept/reject decisions are bit-exact, so acceptance counts and positions
match exactly.
"""
import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _systems():
    import json, os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "be_transcript.json")))
    be = system.from_champ_input(fx["input"], "be")
    p = [r for r in fx["runs"] if r["iteration"] == 4][0]["params"]
    be.a4 = np.array([p["a"]]); be.b = np.array([p["b"]])
    be.c = system.cusp_exact(np.array([p["c"]]), be.nordc)
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "c2_ktt_transcript.json")))
    c2k = system.from_champ_old_input(fx["input"], "c2_ktt")
    p = [r for r in fx["runs"] if r["iteration"] == 3][0]["params"]
    c2k.a4 = np.array([p["a"]]); c2k.b = np.array([p["b"]])
    c2k.c = system.cusp_exact(np.array([p["c"]]), c2k.nordc)
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from transcripts import c2_bfd
    return {"be": be, "be_atom": system.be_atom(1), "c2": system.c2_dimer(2),
            "chain4": system.hchain(4, ncarbon=4), "c2_ktt": c2k, "c2_bfd": c2_bfd(3)[0]}


SYS = _systems()


def _close(a, b, rtol=RTOL, atol=1e-10):
    a, b = np.asarray(a), np.asarray(b)
    err = np.abs(a - b) / (np.abs(b) + atol / rtol)
    return float(err.max())


@pytest.mark.parametrize("name", list(SYS))
def test_psi_at_matches_oracle(name):
    s = SYS[name]
    W = 37          # not a multiple of the walkers per wavefront
    x = system.random_walkers(s, W, seed=11)
    e = Engine(s, W)
    e.set_walkers(x)
    g = e.psi_at()
    r = oracle.Oracle(s).eval_batch(x)
    o9 = r["out9"]
    assert _close(g["logpsid"], o9[:, 0]) < RTOL
    assert np.array_equal(g["sign"], o9[:, 1])
    assert _close(g["psij"], o9[:, 2]) < RTOL
    assert _close(g["epot"], o9[:, 4]) < RTOL
    assert _close(g["eloc"], o9[:, 3], atol=1e-9) < 1e-9
    assert _close(g["grad"], r["grad"]) < RTOL
    assert _close(g["divv"], r["divv"], atol=1e-9) < 1e-9


@pytest.mark.parametrize("name", list(SYS))
def test_move_accept_sequence_matches_oracle(name):
    s = SYS[name]
    W = 19
    x = system.random_walkers(s, W, seed=12)
    e = Engine(s, W)
    e.set_walkers(x)
    e.psi_at()
    O = oracle.Oracle(s)
    ws = [O.L.orc_walker_create(O.h) for _ in range(W)]
    for w in range(W):
        O.eval_full(x[w], walker=ws[w])
    rs = np.random.default_rng(5)
    for step in range(2 * s.nelec):
        ie = step % s.nelec
        xn = x[:, ie, :] + rs.normal(0, 0.4, (W, 3))
        g = e.move_1e(ie, xn)
        ref = [O.ratio_1e(ie, xn[w], walker=ws[w]) for w in range(W)]
        assert _close(g["ratio"], [r["ratio"] for r in ref]) < 1e-9
        assert _close(g["grad"], [r["grad"] for r in ref]) < 1e-9
        assert _close(g["lap"], [r["lap"] for r in ref], atol=1e-9) < 1e-8
        acc = rs.random(W) < 0.6
        e.copy_psi_1e(ie, acc)
        for w in range(W):
            O.accept_1e(ie, bool(acc[w]), walker=ws[w])
            if acc[w]:
                x[w, ie] = xn[w]
        je = (ie + 1) % s.nelec
        gg, ll = e.calc_1e_derivatives(je)
        rr = [O.grad_1e(je, walker=ws[w]) for w in range(W)]
        assert _close(gg, [r[0] for r in rr]) < 1e-9
    assert np.allclose(e.get_walkers(), x, rtol=0, atol=0)
    for w in ws:
        O.L.orc_walker_destroy(w)


@pytest.mark.parametrize("name", ["be", "c2", "chain4", "c2_bfd"])
def test_vmc_bit_exact_decisions(name):
    s = SYS[name]
    W = 24
    x = system.random_walkers(s, W, seed=13)
    seeds = rng.walker_seeds("1837465927472523", W)
    e = Engine(s, W)
    e.set_walkers(x)
    e.set_rng(seeds)
    nstep = 4
    g = e.vmc(nstep, delta=0.6)
    o = oracle.Oracle(s).vmc_run(x, seeds, nstep, 0.6)
    assert np.array_equal(g["naccept"], o["naccept"])
    assert np.array_equal(e.get_walkers(), o["x"])
    assert np.array_equal(e.get_rng(), o["rng"])
    assert _close(g["eloc"], o["eloc"], atol=1e-9) < 1e-9


@pytest.mark.parametrize("name,path", [("be", "fused"), ("c2_ktt", "fused"), ("c2", "fused"),
                                       ("chain4", "fused"), ("c2", "split"), ("be", "device"),
                                       ("c2_bfd", "fused"), ("c2_bfd", "device")])
def test_accelerated_mover_bit_exact_decisions(monkeypatch, name, path):
    """CHAMP's default mover (imetro = 6, metrop_slat_mov1; deltar 5,
    deltat 1, fbias 1 as in the example inputs) on the GPU and in the
    oracle: identical decisions and streams (the rejection loops draw a
    varying number of rannyu), positions to 1e-11 bohr, E_L to 1e-9.  `split`: the
    three-launch move; `device`: the graph-captured device loop."""
    if path == "split":
        monkeypatch.setenv("QMCW_SPLIT_MOVE", "1")
    s = SYS[name]
    W = 24
    x = system.random_walkers(s, W, seed=13)
    seeds = rng.walker_seeds("1837465927472523", W)
    e = Engine(s, W)
    e.set_walkers(x)
    e.set_rng(seeds)
    e.set_vmc_mover(6, 5.0, 1.0)
    O = oracle.Oracle(s)
    O.set_mover(6, 5.0, 1.0)
    nstep = 4
    o = O.vmc_run(x, seeds, nstep, 1.0)
    if path == "device":
        e.vmc_device(nstep, 1.0)
        e.sync()
        assert _close(e.get_eloc(), o["eloc"][-1], atol=1e-9) < 1e-9
    else:
        g = e.vmc(nstep, delta=1.0)
        assert np.array_equal(g["naccept"], o["naccept"])
        assert _close(g["eloc"], o["eloc"], atol=1e-9) < 1e-9
    # the proposal goes through cos, exp and sqrt, whose last bits may differ
    # between the GPU and the host library
    assert np.allclose(e.get_walkers(), o["x"], rtol=0, atol=1e-11)
    assert np.array_equal(e.get_rng(), o["rng"])


@pytest.mark.parametrize("G", [8, 16, 32, 64])
def test_every_group_width_gives_identical_results(G):
    s = SYS["c2"]
    W = 21
    x = system.random_walkers(s, W, seed=14)
    seeds = rng.walker_seeds("1837465927472523", W)
    e = Engine(s, W)
    e.set_group_width(G)
    e.set_walkers(x)
    e.set_rng(seeds)
    g = e.vmc(3, delta=0.6)
    o = oracle.Oracle(s).vmc_run(x, seeds, 3, 0.6)
    assert np.array_equal(g["naccept"], o["naccept"])
    assert np.array_equal(e.get_walkers(), o["x"])
    assert _close(g["eloc"], o["eloc"], atol=1e-9) < 1e-9


def test_autotune_keeps_state_and_parity():
    s = SYS["c2"]
    W = 40
    x = system.random_walkers(s, W, seed=15)
    seeds = rng.walker_seeds("1837465927472523", W)
    e = Engine(s, W)
    e.set_walkers(x)
    e.set_rng(seeds)
    assert e.autotune() in (8, 16, 32, 64)
    assert np.array_equal(e.get_walkers(), x)
    assert np.array_equal(e.get_rng(), np.asarray(seeds, np.int32))
    g = e.vmc(2, delta=0.6)
    o = oracle.Oracle(s).vmc_run(x, seeds, 2, 0.6)
    assert np.array_equal(e.get_walkers(), o["x"])
    assert _close(g["eloc"], o["eloc"], atol=1e-9) < 1e-9


def test_new_parameters_are_used_at_unchanged_positions():
    """After moves, the full evaluation keeps the orbital rows and Jastrow
    pairs the moves committed; qmcw_set_params must make it recompute them:
    psi_at with new MO coefficients and Jastrow b at the walkers' current
    positions against the oracle with the same parameters (an all-electron
    system: the moves also rotated the walkers' pseudopotential grids, which
    the oracle's batch evaluation does not share)."""
    s = SYS["be_atom"]
    W = 19
    x = system.random_walkers(s, W, seed=71)
    e = Engine(s, W)
    e.set_walkers(x)
    e.set_rng(rng.walker_seeds("1837465927472523", W))
    e.vmc(2, delta=0.6)
    s2 = s.copy()
    r = np.random.default_rng(3)
    s2.coef = s.coef + 0.01 * r.normal(size=s.coef.shape)
    s2.b = s.b.copy()
    s2.b[0, 2] += 0.03
    e.set_params(b=s2.b, coef=s2.coef)
    g = e.psi_at()
    xw = e.get_walkers()
    o = oracle.Oracle(s2).eval_batch(xw)
    assert _close(g["logpsid"], o["out9"][:, 0]) < RTOL
    assert _close(g["eloc"], o["out9"][:, 3], atol=1e-9) < 1e-9
    e.close()


@pytest.mark.parametrize("name,every", [("c2", 3), ("chain4", 5)])
def test_periodic_reinversion_keeps_the_walk(name, every):
    """qmcw_set_reinversion: the Slater inverses rebuilt from the kept rows
    every few moves within the sweeps (host loop and graph-captured device
    loop) change them only by the Sherman-Morrison rounding, so decisions,
    positions and streams still equal the oracle's (which never re-inverts
    within a sweep) and E_L agrees to 1e-9."""
    s = SYS[name]
    W = 24
    x = system.random_walkers(s, W, seed=13)
    seeds = rng.walker_seeds("1837465927472523", W)
    o = oracle.Oracle(s).vmc_run(x, seeds, 3, 0.6)
    for device in (False, True):
        e = Engine(s, W)
        e.set_walkers(x)
        e.set_rng(seeds)
        e.set_reinversion(every)
        if device:
            e.psi_at()
            e.vmc_device(3, 0.6)
            e.sync()
            assert _close(e.get_eloc(), o["eloc"][-1], atol=1e-9) < 1e-9
        else:
            g = e.vmc(3, delta=0.6)
            assert np.array_equal(g["naccept"], o["naccept"])
            assert _close(g["eloc"], o["eloc"], atol=1e-9) < 1e-9
        assert np.array_equal(e.get_walkers(), o["x"])
        assert np.array_equal(e.get_rng(), o["rng"])
        e.reinvert()
        g2 = e.psi_at()
        assert _close(g2["eloc"], o["eloc"][-1], atol=1e-9) < 1e-9
        e.close()
