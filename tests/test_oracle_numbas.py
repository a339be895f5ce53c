"""Tabulated radial basis, Gaussian pseudopotential files and the 6-point
quadrature, on CHAMP's pseudopotential C2 example
(examples/c2_bfd_vtz_rhf_vmc.inp with basis.1 and pseudopot1).

  spline setup      spline2's second derivatives equal those of the clamped
                    cubic spline (scipy) with the same end slopes; the small-r
                    polynomial passes through the first five points; the tail
                    matches the last point (read_bas_num.f90:238-381)
  evaluation        the oracle's splfit_bas (value, R', R'') equals the scipy
                    spline inside the grid and the exponential beyond it, on
                    the example's shifted exponential grid (igrid 3) and on
                    linear and logarithmic grids (igrid 1, 2) of r e^-r, whose
                    spline also reproduces the function itself
  reader            sizes, the basis order CHAMP printed, the pseudopotential
                    terms it printed, the l of each tabulated function (its
                    log derivative at the origin, as CHAMP reports them)
  VMC               energy, sigma, potential and interaction energies and the
                    acceptance of the accelerated mover against the runs of
                    CHAMP's transcript (iteration 1: the input's Jastrow;
                    iteration 3: the optimised one it printed)
"""
import math
import os
import sys

import numpy as np
import pytest
import scipy.interpolate

from champ_amd import numbas, rng, system
import oracle

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from transcripts import c2_bfd, c2_bfd_files  # noqa: E402


def _basis_file():
    _, files = c2_bfd_files()
    return numbas.read_basis_file(files["basis.1"], nloc=6)


def test_spline_setup_matches_clamped_cubic_spline():
    bf = _basis_file()
    assert (bf.igrid, bf.nr, bf.exp_h, bf.r0) == (3, 2000, 1.003, 20.0)
    tabs = numbas.radial_tables(bf, nloc=6)
    assert len(tabs) == 9
    # igrid 3: the file's points are r0' (h^(i-1) - 1) with r0' from the last point
    h, r0 = numbas.grid_params(bf)
    assert np.allclose(bf.x, r0 * (h ** np.arange(bf.nr) - 1), rtol=1e-11, atol=1e-14)
    for t, y in zip(tabs, bf.rwf):
        # the polynomial through the first five points
        c = t.ce
        xs = bf.x[:5]
        assert np.allclose(sum(c[i] * xs ** i for i in range(5)), y[:5], rtol=1e-9, atol=1e-14)
        dwf1 = c[1]                          # slope of the polynomial at x_1 = 0
        dwfn = -t.ae[1] * y[-1]
        if t.ae[0] != 0:
            assert abs(t.ae[0] * math.exp(-t.ae[1] * bf.x[-1]) - y[-1]) <= 1e-12 * abs(y[-1])
        cs = scipy.interpolate.CubicSpline(bf.x, y, bc_type=((1, dwf1), (1, dwfn)))
        d2 = cs(bf.x, 2)
        scale = max(1.0, np.abs(d2).max())
        assert np.abs(t.d2 - d2).max() < 1e-7 * scale


def _one_table_system(tab):
    """A one-electron, one-centre system whose s function is table 0 (the
    oracle's struct needs a complete system)."""
    bb = system._BasisBuilder(system.QMCW_BASIS_SLATER, 1)
    rad = bb.add_numeric_radial(bb.add_table(tab))
    bb.add_function(0, rad, 0, 0)
    barr, anorm = bb.arrays(np.array([0], np.int32))
    return system.System(nelec=1, nup=1, cent=np.zeros((1, 3)), iwctype=np.array([0], np.int32),
                         znuc=np.array([1.0]), coef=np.array([[anorm[0]]]),
                         det_up=np.array([[0]], np.int32), det_dn=np.zeros((1, 0), np.int32), csf_coef=np.array([1.0]),
                         csf_det_off=np.array([0, 1], np.int32), csf_det=np.array([0], np.int32),
                         csf_det_coef=np.array([1.0]), isc=4, nspin2b=1, norda=1, nordb=1,
                         nordc=0, scalek=0.8, a4=np.zeros((1, 2)), b=np.array([[0.5, 0.0]]),
                         c=np.zeros((1, 0)), **barr)


def test_tabulated_evaluation_matches_spline_and_tail():
    bf = _basis_file()
    tabs = numbas.radial_tables(bf, nloc=6)
    rs = np.random.default_rng(3)
    r_in = np.concatenate([rs.uniform(0, 20, 400), rs.uniform(0, 0.01, 100)])
    r_out = rs.uniform(20.0, 35.0, 50)
    for k in (0, 3, 5, 8):
        O = oracle.Oracle(_one_table_system(tabs[k]))
        t = tabs[k]
        cs = scipy.interpolate.CubicSpline(bf.x, bf.rwf[k], bc_type=((1, t.ce[1]), (1, -t.ae[1] * bf.rwf[k][-1])))
        got = O.num_radial(0, r_in)
        scale = np.abs(bf.rwf[k]).max()
        for d in range(3):
            ref = cs(r_in, d)
            tol = 1e-9 * max(1.0, np.abs(ref).max()) if d else 1e-10 * scale
            assert np.abs(got[:, d] - ref).max() < tol, (k, d)
        tail = O.num_radial(0, r_out)
        v = t.ae[0] * np.exp(-t.ae[1] * r_out)
        assert np.allclose(tail[:, 0], v, rtol=1e-13, atol=0)
        assert np.allclose(tail[:, 1], -t.ae[1] * v, rtol=1e-13, atol=0)
        assert np.allclose(tail[:, 2], t.ae[1] ** 2 * v, rtol=1e-13, atol=0)


@pytest.mark.parametrize("igrid", [1, 2])
def test_linear_and_log_grids_reproduce_the_function(igrid):
    """R = r e^-r on a linear grid from 0 (h = 0.01) or a logarithmic grid
    from 1e-3 (h = 1.01): the spline reproduces R, R', R'' to the accuracy of
    a cubic spline, equals scipy's clamped spline, and below the first point
    of the logarithmic grid the polynomial takes over."""
    if igrid == 1:
        x = 0.01 * np.arange(2000)
        hdr = f"1 1 2000 0.01 0.0 0"
    else:
        x = 1e-3 * 1.01 ** np.arange(1000)
        hdr = f"1 2 1000 1.01 1.0e-3 0"
    y = x * np.exp(-x)
    text = "# test grid\n" + hdr + "\n" + "\n".join(f"{a:.17e} {b:.17e}" for a, b in zip(x, y))
    bf = numbas.read_basis_file(text, nloc=6)
    t = numbas.radial_tables(bf, nloc=6)[0]
    O = oracle.Oracle(_one_table_system(t))
    rs = np.random.default_rng(7)
    r = rs.uniform(x[0], 15.0, 300)
    got = O.num_radial(0, r)
    ex = [r * np.exp(-r), (1 - r) * np.exp(-r), (r - 2) * np.exp(-r)]
    tol = [1e-7, 1e-5, 1e-3] if igrid == 1 else [1e-6, 1e-4, 1e-2]
    for d in range(3):
        assert np.abs(got[:, d] - ex[d]).max() < tol[d], (igrid, d)
    cs = scipy.interpolate.CubicSpline(x, y, bc_type=((1, t.ce[1] + 2 * t.ce[2] * x[0] + 3 * t.ce[3] * x[0] ** 2
                                                        + 4 * t.ce[4] * x[0] ** 3),
                                                       (1, -t.ae[1] * y[-1])))
    for d in range(3):
        assert np.abs(got[:, d] - cs(r, d)).max() < 1e-9 * max(1.0, np.abs(cs(r, d)).max()), (igrid, d)
    if igrid == 2:
        small = np.array([1e-5, 3e-4, 9e-4])
        v = O.num_radial(0, small)
        c = t.ce
        assert np.allclose(v[:, 0], sum(c[i] * small ** i for i in range(5)), rtol=1e-14)
        assert np.allclose(v[:, 0], small * np.exp(-small), rtol=1e-6)


def test_reader_on_the_c2_bfd_example():
    s, _ = c2_bfd(1)
    assert (s.nelec, s.nup, s.ncent, s.nbasis, s.norb, s.nloc, s.nquad, s.isc) == \
        (8, 4, 2, 68, 68, 6, 6, 4)
    # CHAMP's listing: 5 s, 4 px, 4 py, 4 pz, 2 each of dzr, dx2, dxy, dxz,
    # dyz, then f 3, f -3, f 2, f -2, f 1, f -1, f 0
    assert s.bas_l.tolist() == [0] * 5 + [1] * 12 + [2] * 10 + [3] * 7
    assert s.bas_m.tolist() == ([0] * 5 + [1] * 4 + [-1] * 4 + [0] * 4 +
                                [0, 0, 2, 2, -2, -2, 1, 1, -1, -1] + [3, -3, 2, -2, 1, -1, 0])
    # radial functions by first use: iwrwf 1 2 3 7 8 | 4 5 6 9 | ... | 7 8 | 9
    assert s.rad_num.tolist() == list(range(9))
    assert s.bas_rad.tolist()[:9] == list(range(9))
    # the printed pseudopotential: local (component 2) 3 terms, l = 0 one term
    assert s.npotd.tolist() == [2] and s.lpotp1.tolist() == [2]
    off = s.ecp_off
    l0 = list(zip(s.ecp_coef[off[0]:off[1]], s.ecp_power[off[0]:off[1]], s.ecp_expo[off[0]:off[1]]))
    loc = list(zip(s.ecp_coef[off[1]:off[2]], s.ecp_power[off[1]:off[2]], s.ecp_expo[off[1]:off[2]]))
    assert l0 == [(22.55164191, 2, 5.02991637)]
    assert loc == [(4.0, 1, 8.35973821), (33.43895285, 3, 4.48361888), (-19.17537323, 2, 3.93831258)]
    # the l of each tabulated function from its log derivative at the origin
    # (CHAMP: radial 1-3 l = 0, 4-6 l = 1, 7-8 l = 2, 9 l = 3)
    O = oracle.Oracle(s)
    r = np.array([1e-3, 2e-3])
    ls = []
    for k in range(9):
        v = O.num_radial(k, r)[:, 0]
        ls.append(round(math.log(v[1] / v[0]) / math.log(2.0)))
    assert ls == [0, 0, 0, 2, 2, 1, 1, 1, 3]        # tables in first-use order: 1 2 3 7 8 4 5 6 9
    q = O.quadrature()
    assert len(q[0]) == 6 and np.allclose(q[3], 1 / 6)


def test_six_point_rule_is_exact_to_l3():
    """The octahedral rule integrates the Legendre polynomials P_1..P_3 of
    any axis to zero (their average over the sphere), and P_4 not."""
    s, _ = c2_bfd(1)
    O = oracle.Oracle(s)
    xq, yq, zq, wq = O.quadrature()
    assert abs(wq.sum() - 1) < 1e-15
    rs = np.random.default_rng(11)
    for _ in range(5):
        u = rs.standard_normal(3)
        u /= np.linalg.norm(u)
        c = xq * u[0] + yq * u[1] + zq * u[2]
        p = [c, 0.5 * (3 * c ** 2 - 1), 0.5 * (5 * c ** 3 - 3 * c)]
        for pl in p:
            assert abs((wq * pl).sum()) < 1e-14
    c = zq
    assert abs((wq * (35 * c ** 4 - 30 * c ** 2 + 3) / 8).sum()) > 0.1


def _vmc(s, nw, neq, nstep, nchunk):
    O = oracle.Oracle(s)
    O.set_mover(6, 5.0, 1.0)
    x = system.random_walkers(s, nw, seed=5)
    seeds = rng.walker_seeds("1837465927472523", nw)
    r = O.vmc_run(x, seeds, neq, 1.0, nthreads=8)
    xs, st = r["x"], r["rng"]
    e, pe, pee, nacc = [], [], [], 0
    for _ in range(nchunk):
        r = O.vmc_run(xs, st, nstep // nchunk, 1.0, nthreads=8)
        xs, st = r["x"], r["rng"]
        e.append(r["eloc"])
        nacc += int(r["naccept"].sum())
        for w in range(nw):
            q = O.eval_full(xs[w])
            pe.append(q["pe"])
            pee.append(q["pe_ee"])
    return np.concatenate(e), np.array(pe), np.array(pee), nacc / (nw * nstep * s.nelec)


@pytest.mark.parametrize("iteration", [1, 3])
def test_vmc_matches_champ_c2_bfd_transcript(iteration):
    s, run = c2_bfd(iteration)
    nw, nstep = 256, 400
    e, pe, pee, acc = _vmc(s, nw, 50, nstep, 8)
    blk = e.reshape(-1, 25, nw).mean(axis=1).mean(axis=1)
    err = blk.std(ddof=1) / math.sqrt(len(blk))
    assert abs(e.mean() - run["energy"]) < 3 * math.hypot(err, run["error"]), (e.mean(), err, run)
    assert abs(e.std() - run["sigma"]) < 0.05, (e.std(), run["sigma"])
    # potential and e-e interaction energies from 8 x 256 sampled configurations
    for v, ref, rerr in ((pe, run["potential"], run["potential_error"]),
                         (pee, run["interaction"], run["interaction_error"])):
        se = v.std(ddof=1) / math.sqrt(len(v))
        assert abs(v.mean() - ref) < 3 * math.hypot(se, rerr), (v.mean(), se, ref, rerr)
    nmove = nw * nstep * s.nelec
    pa = run["acceptance"]
    se = 2 * math.hypot(math.sqrt(pa * (1 - pa) / nmove),
                        math.sqrt(pa * (1 - pa) / (run["passes"] * s.nelec)))
    assert abs(acc - pa) < 3.5 * se, (acc, pa, se)


def test_c2_bfd_jastrow_derivatives_fd():
    """The 24 Jastrow parameters the example varies (isc = 4 scaled
    distances, a pseudopotential whose nonlocal part moves with the
    Jastrow): d ln Psi / dp and d E_L / dp against central finite
    differences (h = 1e-5), the dependent e-e-n coefficients re-imposed on
    each side."""
    s, _ = c2_bfd(3)
    assert (len(s.jas_a), len(s.jas_b), len(s.jas_c)) == (4, 5, 15)
    O = oracle.Oracle(s)
    x = system.random_walkers(s, 1, seed=19)[0]
    assert O.eval_full(x)["status"] == 0
    dpsi, deloc = O.derivs()
    assert len(dpsi) == 24
    h = 1e-5

    def lnpsi_el(s2):
        q = oracle.Oracle(s2).eval_full(x)
        return q["logpsid"] + q["psij"], q["eloc"]

    p = 0
    for blk, arr in (("a4", s.jas_a), ("b", s.jas_b), ("c", s.jas_c)):
        for (t, idx) in arr:
            sp, sm = s.copy(), s.copy()
            getattr(sp, blk)[t, idx - 1] += h
            getattr(sm, blk)[t, idx - 1] -= h
            if blk == "c":
                sp.c = system.cusp_exact(sp.c, s.nordc)
                sm.c = system.cusp_exact(sm.c, s.nordc)
            lp, ep = lnpsi_el(sp)
            lm, em = lnpsi_el(sm)
            assert abs((lp - lm) / (2 * h) - dpsi[p]) < 1e-6 * max(1, abs(dpsi[p])), (blk, idx)
            assert abs((ep - em) / (2 * h) - deloc[p]) < 2e-5 * max(1, abs(deloc[p])), (blk, idx)
            p += 1
