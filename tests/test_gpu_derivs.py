"""GPU parity of the parameter derivatives (qmcw_derivs) against the FP64
oracle on identical walker configurations.

d ln Psi/dp and dE_L/dp of every parameter (24 CSF, 40 Jastrow, the orbital
rotations) agree with the oracle to 1e-9 relative (absolute floor 1e-9:
several dE_L/dp are differences of O(10) terms).  The oracle's derivatives
are pinned by central finite differences in tests/test_oracle.py.
"""
import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _rel(a, b, floor):
    a, b = np.asarray(a), np.asarray(b)
    return float((np.abs(a - b) / (np.abs(b) + floor)).max())


def _derivs_system(name):
    if name == "c2":
        return system.c2_dimer(2, derivs=True)
    if name == "c2_ktt":
        # the all-electron C2 example: isc = 2 scaled distances, the 24
        # Jastrow parameters its optimisation section varies
        import json, os
        fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                         "c2_ktt_transcript.json")))
        s = system.from_champ_old_input(fx["input"], "c2_ktt")
        p = [r for r in fx["runs"] if r["iteration"] == 3][0]["params"]
        s.a4 = np.array([p["a"]]); s.b = np.array([p["b"]])
        s.c = system.cusp_exact(np.array([p["c"]]), s.nordc)
        return s
    if name == "mol":
        # a 39-electron random molecule of seven centres (Z_eff 1, 4, 6 with
        # pseudopotentials): its Jastrow set and 16 orbital rotations
        return system.molecule_derivs(system.random_molecule(1011))
    if name == "c2_bfd":
        # the pseudopotential C2 example (tabulated radial basis, 6-point
        # quadrature): the 24 Jastrow parameters its optimisation section varies
        import os, sys
        sys.path.insert(0, os.path.dirname(__file__))
        from transcripts import c2_bfd
        return c2_bfd(3)[0]
    s = system.be_atom(1)
    ja, jb, jc = system.jastrow_opt_list(s, vary_a1=False)
    s.jas_a, s.jas_b, s.jas_c = ja, jb, jc
    return s


@pytest.mark.parametrize("name,valu", [("c2", "0"), ("be_jastrow", "0"), ("c2_ktt", "0"),
                                       ("c2_bfd", "0"), ("mol", "0"), ("c2", "1"),
                                       ("c2_bfd", "1"), ("mol", "1")])
def test_derivs_match_oracle(monkeypatch, name, valu):
    """The derivatives against the oracle; valu = 1 keeps the contractions
    of every orbital and of psp_nonloc_orb on per-walker VALU loops
    (QMCW_NO_GPROJ_MFMA=1) instead of the batched MFMA GEMMs."""
    monkeypatch.setenv("QMCW_NO_GPROJ_MFMA", valu)
    s = _derivs_system(name)
    assert s.nparm > 0
    W = 13
    x = system.random_walkers(s, W, seed=21)
    e = Engine(s, W)
    assert e.P == s.nparm
    e.set_walkers(x)
    dp, de = e.derivs()
    r = oracle.Oracle(s).eval_batch(x, derivs=True)
    assert r["status"] == 0
    assert _rel(dp, r["dpsi"], 1e-9) < 1e-9
    assert _rel(de, r["deloc"], 1e-8) < 1e-8
    # the derivative evaluation refreshes the ordinary state as psi_at does
    g = e.psi_at()
    assert _rel(g["eloc"], r["out9"][:, 3], 1e-9) < 1e-9


def test_mixed_precision_derivs_close_to_oracle():
    """Mixed precision (FP32 basis values and Jastrow at the quadrature
    points, FP64 accumulation): d ln Psi/dp and dE_L/dp within 2e-3
    relative (absolute floor 1e-2).  Measured on the C2-like system: 2.6e-4
    and 4.7e-4, the FP32 rounding of the basis values carried through the
    orbital-rotation blocks; a wrong term gives errors of order one."""
    s = system.c2_dimer(2, derivs=True)
    W = 13
    x = system.random_walkers(s, W, seed=23)
    e = Engine(s, W, "mixed")
    e.set_walkers(x)
    dp, de = e.derivs()
    r = oracle.Oracle(s).eval_batch(x, derivs=True)
    edp, ede = _rel(dp, r["dpsi"], 1e-2), _rel(de, r["deloc"], 1e-2)
    print(f"mixed derivs: dpsi rel {edp:.3g}, deloc rel {ede:.3g}")
    assert edp < 2e-3
    assert ede < 2e-3


def test_sampling_mode_gives_last_step_derivatives(monkeypatch):
    s = system.c2_dimer(2, derivs=True)
    W = 11
    x = system.random_walkers(s, W, seed=22)
    seeds = rng.walker_seeds("1837465927472523", W)
    e = Engine(s, W)
    e.set_walkers(x)
    e.set_rng(seeds)
    e.sample_derivs(True)
    e.vmc(2, delta=0.6)
    dp, de = e.get_derivs()
    xs = e.get_walkers()
    # d ln Psi/dp does not depend on the quadrature grid: against the oracle
    r = oracle.Oracle(s).eval_batch(xs, derivs=True)
    assert _rel(dp, r["dpsi"], 1e-9) < 1e-9
    # dE_L/dp was formed on the last step's rotated grid: a fresh derivative
    # evaluation at the same state gives the same numbers
    dp2, de2 = e.derivs()
    assert np.array_equal(dp, dp2) and np.array_equal(de, de2)
    # those evaluations took the distances and Jastrow pairs the moves kept:
    # the same walk with every evaluation recomputing them gives the same
    # derivatives (the grid rotations come from the same streams)
    monkeypatch.setenv("QMCW_NO_KEEP_ORB", "1")
    e3 = Engine(s, W)
    e3.set_walkers(x)
    e3.set_rng(seeds)
    e3.sample_derivs(True)
    e3.vmc(2, delta=0.6)
    dp3, de3 = e3.get_derivs()
    assert np.array_equal(e3.get_walkers(), xs)
    assert _rel(dp3, dp, 1e-9) < 1e-9 and _rel(de3, de, 1e-8) < 1e-8
    # the walk itself is unchanged by the derivative evaluation
    o = oracle.Oracle(s).vmc_run(x, seeds, 2, 0.6)
    assert np.array_equal(xs, o["x"])


def test_derivs_without_parameters_is_an_error():
    s = system.c2_dimer(2)
    e = Engine(s, 8)
    e.set_walkers(system.random_walkers(s, 8, seed=1))
    with pytest.raises(Exception):
        e.derivs()
