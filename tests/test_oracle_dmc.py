"""Pins of the oracle's DMC (orc_dmc_*), independent of the GPU.

  reweighting      with the exact ground state every E_L is -Z^2/2, so each
                   generation multiplies every weight by exp(-(E0 - ET) tau)
                   pop_ctrl (ene_int_v9 collapses to eest when E_L = eest)
  drift-diffusion  the walk samples |Psi|^2 Green's-function-wise: for the
                   exact hydrogenic 1s state <r> = 3/(2Z) within 4 sigma
  split_join       total weight is conserved, no walker keeps weight zero,
                   population = n - kills - joins + splits, and every walker
                   after branching is a copy of one before
  cusp mover       (all-electron systems) the proposal density is the
                   density its sampler draws from (E_s[f/p] = 1 for
                   normalized f); far from the nucleus and moving away it is
                   the textbook drift-diffusion gaussian about x + v tau; when
                   the drift crosses the nucleus it is centred on the nucleus
                   (isotropic); it is invariant under rotations; and the
                   Metropolis walk built on it samples |Psi|^2 (<r> = 3/(2Z))
"""
import math

import numpy as np
import pytest

import oracle
from champ_amd import rng, system
from champ_amd.abi import DmcParams


def _hydrogenic_pp(Z):
    """1 electron, exact 1s orbital, no Jastrow; a pseudopotential table with
    only the local channel (which is empty), so that V = -Z/r and the
    pseudopotential mover of the reference applies."""
    bb = system._BasisBuilder(system.QMCW_BASIS_SLATER, 1)
    r = bb.add_radial(1, [float(Z)], [1.0])
    bb.add_function(0, r, 0, 0)
    barr, anorm = bb.arrays(np.array([0], np.int32))
    ecp = system._ecp_arrays([dict(local=[], l0=[], l1=[])])
    npotd = np.array([1], np.int32)
    lpotp1 = np.array([1], np.int32)
    return system.System(nelec=1, nup=1, cent=np.zeros((1, 3)), iwctype=np.array([0], np.int32),
                         znuc=np.array([float(Z)]), coef=np.array([[anorm[0]]]),
                         det_up=np.array([[0]], np.int32), det_dn=np.zeros((1, 0), np.int32),
                         csf_coef=np.array([1.0]), csf_det_off=np.array([0, 1], np.int32),
                         csf_det=np.array([0], np.int32), csf_det_coef=np.array([1.0]),
                         isc=4, scalek=1.0, norda=2, nordb=2, nordc=0,
                         a4=np.zeros((1, 3)), b=np.array([[0.5, 1.0, 0.0]]), c=np.zeros((1, 0)),
                         nloc=6, npotd=npotd, lpotp1=lpotp1, ecp_off=np.zeros_like(ecp[2]),
                         ecp_coef=ecp[3][:0], ecp_power=ecp[4][:0], ecp_expo=ecp[5][:0],
                         nquad=12, **barr)


def _hydrogenic_ae(Z):
    """The same exact 1s state as an all-electron system (V = -Z/r), which
    the reference moves with the nucleus-cusp mover."""
    s = _hydrogenic_pp(Z)
    s.nloc = 0
    return s


@pytest.mark.parametrize("Z", [1, 2])
@pytest.mark.parametrize("ae", [False, True])
def test_exact_wavefunction_weights(Z, ae):
    s = _hydrogenic_ae(Z) if ae else _hydrogenic_pp(Z)
    W = 64
    x = system.random_walkers(s, W, seed=3, sigma=1.0 / Z)
    seeds = rng.walker_seeds("1837465927472523", W)
    d = oracle.OracleDMC(s, W)
    assert d.set(x, seeds) == 0
    e0, et, tau = -0.5 * Z * Z, -0.5 * Z * Z + 0.3, 0.02
    p = DmcParams(tau, et, eest=e0, esigma=0.5, pop_ctrl=1.1, itau_eff=0)
    w = np.ones(W)
    for _ in range(3):
        st, stats = d.generation(p)
        assert st == 0
        w *= np.exp(-(e0 - et) * tau) * 1.1
        g = d.get()
        assert np.allclose(g["eloc"], e0, rtol=0, atol=1e-12)
        assert np.allclose(g["weight"], w, rtol=1e-12)
        assert stats.nacc > 0 and stats.ntry == W


@pytest.mark.parametrize("ae", [False, True])
def test_drift_diffusion_samples_the_ground_state(ae):
    Z = 1
    s = _hydrogenic_ae(Z) if ae else _hydrogenic_pp(Z)
    W = 2000
    x = system.random_walkers(s, W, seed=5, sigma=1.0)
    seeds = rng.walker_seeds("1122334455667789", W)
    d = oracle.OracleDMC(s, W)
    assert d.set(x, seeds) == 0
    p = DmcParams(0.01, -0.5, eest=-0.5, esigma=0.5, idmc=0)
    rs = []
    for gen in range(120):
        assert d.generation(p)[0] == 0
        if gen >= 40 and gen % 4 == 0:
            rs.append(np.linalg.norm(d.get()["x"][:, 0, :], axis=1))
    r = np.array(rs)
    # generations 4 apart are correlated; use block means per walker
    m = r.mean(axis=0)
    est, err = m.mean(), m.std() / np.sqrt(W)
    assert abs(est - 1.5 / Z) < 4 * err + 0.01, (est, err)


def test_split_join_invariants():
    s = system.c2_dimer(2)
    W, cap = 24, 48
    x = system.random_walkers(s, W, seed=9)
    seeds = rng.walker_seeds("1837465927472523", W)
    d = oracle.OracleDMC(s, cap)
    rs = np.random.default_rng(2)
    w = rs.choice([0.0, 0.2, 0.3, 0.45, 1.0, 1.3, 2.6, 5.5, 9.0], size=W)
    assert d.set(x, seeds, w) == 0
    before = d.get()
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    st, stats = d.branch(brng)
    assert st == 0
    after = d.get()
    assert abs(after["weight"].sum() - w.sum()) < 1e-12 * w.sum()
    assert (after["weight"] > 0).all()
    nkill = int((w == 0).sum())
    nsmall = int(((w > 0) & (w < 0.5)).sum())
    njoin = nsmall // 2
    nsplit = stats.nbranch - nkill - njoin
    assert nsplit >= int((w > 9.0 - 1e-12).sum())
    assert stats.nwalk == W - nkill - njoin + nsplit == len(after["weight"])
    # copies: each walker after branching has the positions of one before
    flat = before["x"].reshape(W, -1)
    for xa in after["x"].reshape(len(after["weight"]), -1):
        assert (np.abs(flat - xa).max(axis=1) == 0).any()
    # at most one unpaired walker below 1/2 remains
    assert int((after["weight"] < 0.5).sum()) <= 1


class OracleStepper:
    """dmc_step over the oracle's population: generation + split_join."""

    def __init__(self, d, brng):
        self.d, self.brng = d, brng

    def dmc_step(self, params):
        r, st = self.d.generation(params)
        assert r == 0
        r, sb = self.d.branch(self.brng)
        assert r == 0
        st.nwalk, st.nbranch = sb.nwalk, sb.nbranch
        return st


def _hydrogen_sto(zeta):
    s = _hydrogenic_pp(1)
    bb = system._BasisBuilder(system.QMCW_BASIS_SLATER, 1)
    r = bb.add_radial(1, [float(zeta)], [1.0])
    bb.add_function(0, r, 0, 0)
    barr, anorm = bb.arrays(np.array([0], np.int32))
    for k, v in barr.items():
        setattr(s, k, v)
    s.coef = np.array([[anorm[0]]])
    return s


def test_dmc_recovers_the_exact_energy_of_a_nodeless_state():
    """H atom with the wrong orbital exp(-1.3 r): VMC gives zeta^2/2 - zeta =
    -0.455 Ha; one electron has no nodes, so DMC must give -0.5 Ha (time step
    0.02, error ~ a few mHa)."""
    from champ_amd.dmc import DmcDriver
    s = _hydrogen_sto(1.3)
    W0, cap = 1000, 2500
    x = system.random_walkers(s, W0, seed=7, sigma=1.0)
    seeds = rng.walker_seeds("1837465927472523", W0)
    d = oracle.OracleDMC(s, cap)
    assert d.set(x, seeds) == 0
    evmc = float(d.get()["eloc"].mean())
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    brng[3] |= 1
    drv = DmcDriver(OracleStepper(d, brng), 1, W0, tau=0.02, etrial=-0.45)
    drv.run(150)                         # equilibration
    h = drv.run(300)
    e = np.array([g["eave"] for g in h])
    blocks = e.reshape(15, 20).mean(axis=1)
    est, err = blocks.mean(), blocks.std() / np.sqrt(len(blocks))
    assert abs(est + 0.5) < max(4 * err, 0.004), (est, err, evmc)
    assert 0.3 * W0 < h[-1]["nwalk"] < 2 * W0      # population control holds


# ---- the all-electron mover (qmc/green_functions_mod.f90:83-145, 268-323)

def _stream(seed="1837465927472523"):
    return np.asarray(rng.walker_seeds(seed, 1), np.int32)[0].copy()


def _gauss3(d, var):
    return np.exp(-np.dot(d, d) / (2 * var)) / (2 * np.pi * var) ** 1.5


def test_cusp_mover_density_matches_its_sampler():
    """Electron 0.1 bohr from a Z=2 nucleus with the drift pointing inward,
    drift length ~ distance, so both branches are drawn (pgaus near 1/2).  For densities f normalized in
    closed form, a gaussian about the electron narrower than sqrt(tau) and an
    exponential about the nucleus sharper than the mover's, f/p is bounded
    and E_s[f/p] = 1 when p is the density the sampler draws from."""
    Z, tau = 2.0, 0.05
    o = oracle.Oracle(_hydrogenic_ae(Z))
    xe = np.array([0.03, -0.05, 0.08])
    g = -Z * xe / np.linalg.norm(xe) + np.array([0.3, 0.1, -0.2])
    st = _stream()
    n, var, zf = 30000, 0.5 * tau, 8.0
    r1, r2 = np.empty(n), np.empty(n)
    for i in range(n):
        xn, _ = o.cusp_move(tau, g, xe, rng=st)
        _, p = o.cusp_move(tau, g, xe, pt=xn)
        r = np.linalg.norm(xn)
        r1[i] = _gauss3(xn - xe, var) / p
        r2[i] = zf ** 3 * np.exp(-2 * zf * r) / np.pi / p
    for est in (r1, r2):
        m, err = est.mean(), est.std() / np.sqrt(n)
        assert err < 0.01 and abs(m - 1) < 5 * err, (m, err)


def test_cusp_mover_far_from_the_nucleus_is_drift_diffusion():
    """Electron 2.4 bohr from the nucleus, moving away (g . xe > 0): zf > zi,
    pgaus = 1 to machine precision and the bent drift point is x + v tau, so
    the density is the plain drift-diffusion gaussian with the velocity of
    averaged_1e_velocity (green_functions_mod.f90:616-630) for this frame:
    adrift = (1 + 1e-10 + cos(g, xe)) / 2 + 0.1 h / (1 + h), h = (Z r / 2)^2,
    |v| = (sqrt(1 + 2 adrift |g|^2 tau) - 1) / (adrift |g| tau)."""
    Z, tau = 2.0, 0.01
    o = oracle.Oracle(_hydrogenic_ae(Z))
    xe = np.array([1.2, -0.7, 2.0])
    g = 0.8 * xe / np.linalg.norm(xe) + np.array([0.5, 0.9, -0.3])
    r, gn = np.linalg.norm(xe), np.linalg.norm(g)
    assert g @ xe > 0
    h = (Z * r / 2) ** 2
    a = (1 + 1e-10 + (g @ xe) / (gn * r)) / 2 + 0.1 * h / (1 + h)
    vav = (np.sqrt(1 + 2 * a * gn ** 2 * tau) - 1) / (a * gn * tau)
    xd = xe + vav * g / gn * tau
    rs = np.random.default_rng(4)
    for _ in range(10):
        pt = xd + rs.normal(size=3) * np.sqrt(tau)
        _, p = o.cusp_move(tau, g, xe, pt=pt)
        assert abs(p - _gauss3(pt - xd, tau)) <= 1e-12 * p


def test_cusp_mover_toward_the_nucleus_is_centred_on_it():
    """Drift past the nucleus (zf < 0): the electron cannot drift through
    it, so the drift point is the nucleus and both branches are isotropic
    about it: p(c + d u) is the same for every unit vector u."""
    Z, tau = 2.0, 0.05
    o = oracle.Oracle(_hydrogenic_ae(Z))
    xe = np.array([0.01, 0.02, -0.015])
    g = -3.0 * xe / np.linalg.norm(xe) + np.array([0.2, -0.4, 0.1])
    rs = np.random.default_rng(5)
    for d in (0.02, 0.1, 0.3):
        ps = []
        for _ in range(6):
            u = rs.normal(size=3)
            _, p = o.cusp_move(tau, g, xe, pt=d * u / np.linalg.norm(u))
            ps.append(p)
        assert np.ptp(ps) <= 1e-12 * max(ps), ps


def test_cusp_mover_rotation_invariance():
    """Rotating the electron, its gradient and the target point about the
    nucleus leaves the density unchanged."""
    Z, tau = 2.0, 0.05
    o = oracle.Oracle(_hydrogenic_ae(Z))
    rs = np.random.default_rng(6)
    for _ in range(5):
        q, _ = np.linalg.qr(rs.normal(size=(3, 3)))
        xe = rs.normal(size=3) * 0.3
        g = rs.normal(size=3) * 2
        pt = xe + rs.normal(size=3) * 0.2
        _, p0 = o.cusp_move(tau, g, xe, pt=pt)
        _, p1 = o.cusp_move(tau, q @ g, q @ xe, pt=q @ pt)
        assert abs(p1 - p0) <= 1e-12 * p0


# ---- DMC-mode pseudopotential grid (qmc/vmc/nonloc.f90:89, l_mode_dmc)

def _rotqua(st, q0):
    """rotqua from the stream st (advanced in place): the rotated grid."""
    L = oracle.lib()
    xq, yq, zq = (np.zeros(12) for _ in range(3))
    from oracle import _d, _i
    L.orc_rotqua(_i(st), _d(q0[0]), _d(q0[1]), _d(q0[2]), 12, _d(xq), _d(yq), _d(zq))
    return xq, yq, zq


@pytest.mark.parametrize("near", [0, 1, 2])
def test_dmc_evaluation_rotates_the_grid_per_pair_in_range(near):
    """A DMC evaluation draws one rotqua per (electron, centre) pair within
    the nonlocal range, in pair order, from the walker's stream: with `near`
    electrons close to the first centre of the C2-like dimer and the others
    far from both, the stream advances by exactly `near` rotations; with one,
    E_L equals the evaluation on the grid rotated by that draw."""
    s = system.c2_dimer(2)
    assert np.allclose(s.cent[:, :2], 0) and abs(s.cent[0, 2]) < 1.5
    N = s.nelec
    rs = np.random.default_rng(3)
    phi = 2 * np.pi * np.arange(N) / N                 # a ring 4 bohr from both centres
    x = np.stack([4 * np.cos(phi), 4 * np.sin(phi), rs.normal(size=N) * 0.3], axis=1)
    c0 = s.cent[0]
    for i in range(near):
        d = rs.normal(size=3)
        x[i] = c0 + 0.35 * d / np.linalg.norm(d)
    seed = np.asarray(rng.walker_seeds("1837465927472523", 1), np.int32)
    d = oracle.OracleDMC(s, 1)
    assert d.set(x[None], seed) == 0
    g = d.get()
    O = oracle.Oracle(s)
    q0 = O.quadrature()
    st = seed[0].copy()
    grids = [_rotqua(st, q0) for _ in range(near)]
    assert np.array_equal(g["rng"][0], st)
    if near == 1:
        O.set_quad(*grids[0])
        r = O.eval_full(x)
        assert r["status"] == 0
        assert abs(r["eloc"] - g["eloc"][0]) <= 1e-12 * abs(r["eloc"])
        O.set_quad(*q0[:3])
        assert abs(O.eval_full(x)["eloc"] - g["eloc"][0]) > 1e-9   # the grid matters


# ---- T-moves (qmc/dmc/dmc_elec/tmove.f90)

def _hydrogenic_nonlocal(Z, v0):
    """One electron in the exact 1s state of -Z/r plus a nonlocal s channel
    v_s(r) = v0 exp(-r^2): Psi is spherical, so every quadrature-point ratio
    at the electron's radius is 1."""
    s = _hydrogenic_pp(Z)
    ecp = system._ecp_arrays([dict(local=[], l0=[(v0, 2, 1.0)], l1=[])])
    s.npotd, s.lpotp1, s.ecp_off, s.ecp_coef, s.ecp_power, s.ecp_expo = ecp
    return s


@pytest.mark.parametrize("casula", [False, True])
def test_tmove_frequency_of_a_spherical_state(casula):
    """Spherical Psi and an s channel: each of the 12 points weighs
    w_q (exp(-tau v_s) - 1) with sum w_q = 1, the backward sum equals the
    forward one (det_balance1 accepts every proposal), so the electron moves
    with probability S / (1 + S), S = exp(-tau v_s(r)) - 1, onto the sphere
    of its radius."""
    v0, tau, r = -2.0, 0.5, 0.8
    s = _hydrogenic_nonlocal(1.0, v0)
    O = oracle.Oracle(s)
    S = math.exp(-tau * v0 * math.exp(-r * r)) - 1
    st = _stream()
    rs = np.random.default_rng(8)
    n, moved = 3000, 0
    for _ in range(n):
        u = rs.normal(size=3)
        x = (r * u / np.linalg.norm(u))[None]
        a, xo = O.tmove_run(x, st, tau, casula)
        assert a in (0, 1)
        if a:
            moved += 1
            assert abs(np.linalg.norm(xo[0]) - r) < 1e-12
            assert np.linalg.norm(xo[0] - x[0]) > 1e-3
        else:
            assert np.array_equal(xo, x)
    p = S / (1 + S)
    sd = math.sqrt(p * (1 - p) / n)
    assert abs(moved / n - p) < 4 * sd, (moved / n, p, sd)


def test_tmove_saved_ratios_are_the_one_electron_ratios():
    """The ratios a DMC-mode evaluation keeps for T-moves (G^T phi times the
    Jastrow factor at each rotated point) equal the one-electron move ratio
    Psi(x_i -> c + r q) / Psi of the same point, for every electron of the
    C2-like dimer in range of a pseudopotential."""
    s = system.c2_dimer(2)
    O = oracle.Oracle(s)
    x = system.random_walkers(s, 1, seed=31)[0]
    st = _stream()
    O.tmove_run(x, st, 1e-12)            # evaluation; tau ~ 0 makes no T-move
    npts = 0
    for i in range(s.nelec):
        q, rat = O.walker_tq(i)
        assert len(rat) % 12 == 0
        npts += len(rat)
        for k in range(len(rat)):
            # the point lies on the sphere of electron i about one of the centres
            errs = []
            for ic in range(s.ncent):
                xn = s.cent[ic] + np.linalg.norm(x[i] - s.cent[ic]) * q[k]
                errs.append(abs(O.ratio_1e(i, xn)["ratio"] - rat[k]) / max(1, abs(rat[k])))
            assert min(errs) <= 1e-10, (i, k, errs)
    assert npts > 0


# ---- the all-electron mover in a DMC run

def he_atom(zeta=1.6875):
    """He: Z = 2, one 1s Slater orbital per spin (exponent 27/16), e-e Jastrow
    with the antiparallel cusp (b1 = 1/2), no e-n or e-e-n terms.  The
    ground state is nodeless, so DMC gives the exact -2.903724 Ha."""
    bb = system._BasisBuilder(system.QMCW_BASIS_SLATER, 1)
    r = bb.add_radial(1, [zeta], [1.0])
    bb.add_function(0, r, 0, 0)
    barr, anorm = bb.arrays(np.array([0], np.int32))
    cc, off, cdet, cdc = system._csf_arrays([(1.0, [(0, 1.0)])], False)
    return system.System(nelec=2, nup=1, cent=np.zeros((1, 3)), iwctype=np.array([0], np.int32),
                         znuc=np.array([2.0]), coef=np.array([[anorm[0]]]),
                         det_up=np.array([[0]], np.int32), det_dn=np.array([[0]], np.int32),
                         csf_coef=cc, csf_det_off=off, csf_det=cdet, csf_det_coef=cdc,
                         isc=4, scalek=1.0, norda=4, nordb=4, nordc=4,
                         a4=np.zeros((1, 5)), b=np.array([[0.5, 0.4, 0.0, 0.0, 0.0]]),
                         c=np.zeros((1, system.nterms4(4))), name="he", **barr)


def test_dmc_recovers_the_exact_helium_energy():
    """All-electron DMC (the nucleus-cusp mover) of He from an approximate
    wavefunction: -2.903724 Ha within 4 sigma (or 5 mHa) at tau = 0.01."""
    from champ_amd.dmc import DmcDriver
    s = he_atom()
    W0, cap = 1000, 2500
    x = system.random_walkers(s, W0, seed=7, sigma=0.6)
    d = oracle.OracleDMC(s, cap)
    assert d.set(x, rng.walker_seeds("1837465927472523", W0)) == 0
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    brng[3] |= 1
    drv = DmcDriver(OracleStepper(d, brng), 2, W0, tau=0.01, etrial=-2.9)
    drv.run(200)
    e = np.array([g["eave"] for g in drv.run(400)])
    blocks = e.reshape(20, 20).mean(axis=1)
    est, err = blocks.mean(), blocks.std() / np.sqrt(len(blocks))
    assert abs(est + 2.903724) < max(4 * err, 0.005), (est, err)


def test_end_of_equilibration_resets_etrial_and_clears_the_averages():
    """CHAMP's end of an equilibration run (qmc/dmc_mod.f90:386-403) with
    l_reset_etrial: etrial takes eest, every population-control factor is
    multiplied by exp((eest - etrial) taueff), and the averages restart, so
    the next generation's eest is that generation's weighted E_L and its
    pop_ctrl has no energy factor left: (wtarg / wtot)^(1 / ngen_pop_ctrl)."""
    import math
    from champ_amd.dmc import DmcDriver
    s = he_atom()
    W0, cap = 200, 500
    x = system.random_walkers(s, W0, seed=8, sigma=0.6)
    d = oracle.OracleDMC(s, cap)
    assert d.set(x, rng.walker_seeds("1837465927472523", W0)) == 0
    brng = np.array(rng.seed_limbs("0000111122223333"), np.int32)
    brng[3] |= 1
    drv = DmcDriver(OracleStepper(d, brng), 2, W0, tau=0.01, etrial=-2.5)
    drv.run(30)
    eest, etrial, taueff = drv.eest, drv.etrial, drv.swt / drv.sw
    ff0, wtot = list(drv.ff), drv.wtot
    assert abs(eest - etrial) > 1e-3
    drv.end_equilibration(reset_etrial=True)
    dff = math.exp((eest - etrial) * taueff)
    assert drv.etrial == eest
    assert np.allclose(drv.ff, [f * dff for f in ff0], rtol=1e-15, atol=0)
    assert drv.sw == drv.swe == drv.swt == 0.0 and drv.is_ == 0
    h = drv.run(1)[-1]
    assert abs(h["pop_ctrl"] - (W0 / wtot) ** (1.0 / drv.ngen_pop_ctrl)) < 1e-14
    assert abs(drv.eest - h["eave"]) < 1e-12 * abs(h["eave"])
    # without the reset only the averages restart
    etr = drv.etrial
    drv.end_equilibration(reset_etrial=False)
    assert drv.etrial == etr and drv.sw == 0.0
