"""Pins of the FP64 oracle (oracle/oracle.c), independent of the GPU.

Each oracle function is checked against something that does not retype its
formula:
  rannyu, setrn, gauss     bit-exact vs the reference's rannyu.f90 / gauss.f90
                           compiled from its sources (oracle/_ref) and vs the
                           closed form R <- 11^13 R mod 2^48
  matinv                   bit-exact vs the reference's matinv.f90; numpy
                           inverse/determinant; singular -> det 0, zeros
  gesqua / rotqua          bit-exact vs the reference's rotqua.f90; exactness
                           of the 12-point rule for sum_q w P_l(a.q), l <= 5
  basis functions          normalisation integrals (product Gauss rule);
                           finite-difference gradient and laplacian
  Jastrow                  finite differences; e-e cusp 1/2 (anti-parallel)
                           and 1/4 (parallel); the cusp conditions of
                           CHAMP's printed e-e-n coefficients (Be transcript)
  determinants (AMF)       brute-force sum_d c_d det(up) det(dn) with numpy
  local energy             hydrogenic closed form -Z^2/2; E_L identity
                           against the finite-difference laplacian of Psi
  pseudopotential          12-point rule against a high-order quadrature of
                           the same angular integral (exact for s, p orbitals)
  parameter derivatives    central finite differences (h = 1e-5) of ln Psi
                           and E_L for CSF, Jastrow (incl. dependent c's) and
                           orbital-rotation parameters
  VMC                      energies of CHAMP's Be transcript within 3 sigma;
                           energies and sigmas of its all-electron C2
                           transcript (old-format input, isc = 2) likewise
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import oracle
from champ_amd import rng, system

HERE = os.path.dirname(os.path.abspath(__file__))


def _ref_lib():
    if not os.path.exists(oracle.REF_LIB):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = C.CDLL(oracle.REF_LIB)
    L.ref_rannyu.restype = C.c_double
    L.ref_gauss.restype = C.c_double
    L.ref_matinv.restype = C.c_double
    L.ref_matinv.argtypes = [C.POINTER(C.c_double), C.c_int32]
    return L


def _be(iteration=4):
    fx = json.load(open(os.path.join(HERE, "golden", "be_transcript.json")))
    s = system.from_champ_input(fx["input"], "be")
    run = [r for r in fx["runs"] if r["iteration"] == iteration][0]
    if run["params"]:
        p = run["params"]
        s.a4 = np.array([p["a"]])
        s.b = np.array([p["b"]])
        s.c = system.cusp_exact(np.array([p["c"]]), s.nordc)
    return s, run


# ---------------------------------------------------------------- RNG
def test_rannyu_matches_closed_form_and_reference():
    seed = rng.seed_limbs("1837465927472523")
    seq, last = oracle.rng_stream(seed, 2000)
    ref, lref = rng.rannyu_sequence(seed, 2000)
    assert np.array_equal(seq, np.array(ref))
    assert list(last) == lref
    L = _ref_lib()
    L.ref_setrn((C.c_int32 * 4)(*seed))
    r = np.array([L.ref_rannyu() for _ in range(2000)])
    assert np.array_equal(seq, r)


def test_walker_streams_are_jumps_of_the_master_stream():
    seeds = rng.walker_seeds("1837465927472523", 3, stride=1000)
    base = rng.seed_limbs("1837465927472523")
    _, l1000 = rng.rannyu_sequence(base, 1000)
    _, l2000 = rng.rannyu_sequence(base, 2000)
    assert list(seeds[0]) == l1000 and list(seeds[1]) == l2000


def test_gauss_matches_reference():
    L = _ref_lib()
    seed = rng.seed_limbs("0000123456789011")
    L.ref_setrn((C.c_int32 * 4)(*seed))
    st = np.zeros(4, np.int32)
    lib = oracle.lib()
    lib.orc_setrn(st.ctypes.data_as(C.POINTER(C.c_int32)),
                  np.array(seed, np.int32).ctypes.data_as(C.POINTER(C.c_int32)))
    for _ in range(200):
        a = lib.orc_gauss(st.ctypes.data_as(C.POINTER(C.c_int32)))
        b = L.ref_gauss()
        assert a == b


# ---------------------------------------------------------------- matinv
@pytest.mark.parametrize("n", [1, 2, 4, 7, 13])
def test_matinv_matches_numpy_and_reference(n):
    rs = np.random.default_rng(n)
    a = rs.normal(size=(n, n))
    inv, det = oracle.matinv(a)
    assert np.allclose(inv, np.linalg.inv(a), rtol=1e-12, atol=1e-12)
    assert math.isclose(det, np.linalg.det(a), rel_tol=1e-12)
    L = _ref_lib()
    flat = np.ascontiguousarray(a.T.reshape(-1))
    dref = L.ref_matinv(flat.ctypes.data_as(C.POINTER(C.c_double)), n)
    assert dref == det
    assert np.array_equal(flat.reshape(n, n).T, inv)


def test_matinv_singular():
    a = np.array([[1.0, 2.0], [2.0, 4.0]])
    inv, det = oracle.matinv(a)
    a[1] = 0.0
    inv, det = oracle.matinv(a)
    assert det == 0.0 and not inv.any()


# ---------------------------------------------------------------- quadrature
def _ref_driver(seed):
    exe = os.path.join(os.path.dirname(oracle.REF_LIB), "ref_driver")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import subprocess
    out = subprocess.run([exe, *map(str, seed)], capture_output=True, text=True, check=True).stdout
    q = np.array([[float(t) for t in ln.split()[1:]] for ln in out.splitlines() if ln.startswith("Q ")])
    r = np.array([[float(t) for t in ln.split()[1:]] for ln in out.splitlines() if ln.startswith("R ")])
    return q, r


def test_quadrature_matches_reference_and_is_exact():
    s = system.c2_dimer(2)
    O = oracle.Oracle(s)
    xq, yq, zq, wq = O.quadrature()
    seed = rng.seed_limbs("4444333322221111")
    q, rot = _ref_driver(seed)
    assert np.array_equal(np.stack([xq, yq, zq, wq], 1), q)
    rs = np.random.default_rng(3)
    for _ in range(5):
        a = rs.normal(size=3)
        a /= np.linalg.norm(a)
        c = xq * a[0] + yq * a[1] + zq * a[2]
        for l in range(6):
            pl = np.polynomial.legendre.Legendre.basis(l)(c)
            assert abs((wq * pl).sum() - (1.0 if l == 0 else 0.0)) < 1e-10
    # grid rotated with the same rannyu stream as the reference
    st = np.zeros(4, np.int32)
    lib = oracle.lib()
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    lib.orc_setrn(st.ctypes.data_as(C.POINTER(C.c_int32)),
                  np.array(seed, np.int32).ctypes.data_as(C.POINTER(C.c_int32)))
    xo, yo, zo = (np.zeros(12) for _ in range(3))
    lib.orc_rotqua(st.ctypes.data_as(C.POINTER(C.c_int32)), P(xq), P(yq), P(zq), 12,
                   P(xo), P(yo), P(zo))
    assert np.array_equal(np.stack([xo, yo, zo], 1), rot)


# ---------------------------------------------------------------- basis
def _radial_grid(n=200, rmax=25.0):
    x, w = np.polynomial.legendre.leggauss(n)
    r = 0.5 * rmax * (x + 1)
    return r, 0.5 * rmax * w


@pytest.mark.parametrize("which", ["be", "c2"])
def test_basis_normalisation(which):
    s = _be(1)[0] if which == "be" else system.c2_dimer(2)
    if which == "c2":  # one centre only
        s = s.copy()
        s.cent = s.cent[:1]
        s.iwctype = s.iwctype[:1]
        nb = s.ctype_bas_off[1]
        s.coef = np.zeros((s.coef.shape[0], nb))   # keep norb: det lists index it
    O = oracle.Oracle(s)
    nb = s.nbasis
    # unnormalised values times the stored normalisation -> unit norm
    from champ_amd.system import radial_norm
    r, wr = _radial_grid(120, 20.0 if which == "c2" else 15.0)
    ct, wt = np.polynomial.legendre.leggauss(24)
    ph = np.linspace(0, 2 * np.pi, 48, endpoint=False)
    acc = np.zeros(nb)
    for ri, wri in zip(r, wr):
        for c, w_c in zip(ct, wt):
            st_ = math.sqrt(1 - c * c)
            for p in ph:
                xe = ri * np.array([st_ * math.cos(p), st_ * math.sin(p), c]) + s.cent[0]
                phi, _, _ = O.basis(xe)
                acc += phi ** 2 * wri * ri * ri * w_c * (2 * np.pi / len(ph))
    for b in range(nb):
        t = s.iwctype[0]
        q = s.ctype_bas_off[t] + b
        rr = s.bas_rad[q]
        p0, p1 = s.rad_prim_off[rr], s.rad_prim_off[rr + 1]
        nrm = radial_norm(s.basis_type, int(s.rad_n[rr]), int(s.bas_l[q]),
                          s.prim_zeta[p0:p1], s.prim_coef[p0:p1])
        assert abs(acc[b] * nrm * nrm - 1.0) < 1e-6, (b, acc[b] * nrm * nrm)


def test_basis_derivatives_fd():
    s = system.c2_dimer(2)
    O = oracle.Oracle(s)
    rs = np.random.default_rng(1)
    h = 1e-4
    for _ in range(3):
        xe = rs.normal(size=3)
        phi, dphi, d2 = O.basis(xe)
        g = np.zeros_like(dphi)
        lap = np.zeros_like(d2)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            pp, pm = O.basis(xe + e)[0], O.basis(xe - e)[0]
            pp2, pm2 = O.basis(xe + 2 * e)[0], O.basis(xe - 2 * e)[0]
            g[:, k] = (-pp2 + 8 * pp - 8 * pm + pm2) / (12 * h)
            lap += (-pp2 + 16 * pp - 30 * phi + 16 * pm - pm2) / (12 * h * h)
        assert np.allclose(g, dphi, rtol=1e-7, atol=1e-8)
        assert np.allclose(lap, d2, rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------- Jastrow
def test_jastrow_fd_and_cusps():
    s, _ = _be(4)
    O = oracle.Oracle(s)
    x = system.random_walkers(s, 1, seed=3)[0]
    J = O.jastrow(x)
    h = 1e-4
    g = np.zeros_like(x)
    lap = 0.0
    for i in range(s.nelec):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fp, fm = O.jastrow(xp)["value"], O.jastrow(xm)["value"]
            g[i, k] = (fp - fm) / (2 * h)
            lap += (fp - 2 * J["value"] + fm) / h ** 2
    assert np.allclose(g, J["v"], atol=1e-7)
    assert abs(lap - J["d2"]) < 1e-4
    # e-e cusp: d J/d r_ij at coalescence, b1 = 1/2 with sspinn = 1/2 for
    # parallel spins (qmc/vmc/jastrow4.f90:79-93).  The outward slopes along
    # +d and -d share the cusp and carry the smooth gradient with opposite
    # signs, so their mean is the cusp alone.
    for (i, j, want) in [(0, 2, 0.5), (0, 1, 0.25)]:
        xx = x.copy()
        d = np.array([0.3, -0.2, 0.5]) / np.linalg.norm([0.3, -0.2, 0.5])
        slopes = []
        for sgn in (1.0, -1.0):
            vals = []
            for t in (1e-5, 2e-5):
                xx[i] = x[j] + sgn * t * d
                vals.append(O.jastrow(xx)["value"])
            slopes.append((vals[1] - vals[0]) / 1e-5)
        slope = 0.5 * (slopes[0] + slopes[1])
        assert abs(slope - want) < 1e-3, (i, j, slope)


def test_champ_een_coefficients_satisfy_the_cusp_matrix():
    fx = json.load(open(os.path.join(HERE, "golden", "be_transcript.json")))
    for run in fx["runs"]:
        if run["params"] is None:
            continue
        c = np.array(run["params"]["c"])
        d, iwc4 = system.cusp_matrix(5)
        assert np.abs(d @ c).max() < 1e-8      # printed with 12 digits
        # the line cuspinit4 prints for nordc=5
        # (examples/c2_ktt_hfj_vmc_emin_j_opt3_lin.out:344)
        assert iwc4 == [1, 4, 10, 19, 2, 6, 12, 22]


def test_jastrow_parameter_derivatives_fd():
    s = system.c2_dimer(2, derivs=True)
    O = oracle.Oracle(s)
    x = system.random_walkers(s, 1, seed=5)[0]
    J = O.jastrow(x, derivs=True)
    h = 1e-5
    p = 0
    def val_of(a4=None, b=None, c=None):
        s2 = s.copy()
        if a4 is not None: s2.a4 = a4
        if b is not None: s2.b = b
        if c is not None: s2.c = c
        return oracle.Oracle(s2).jastrow(x)["value"]
    for (t, idx) in s.jas_a:
        ap, am = s.a4.copy(), s.a4.copy()
        ap[t, idx - 1] += h
        am[t, idx - 1] -= h
        assert abs((val_of(a4=ap) - val_of(a4=am)) / (2 * h) - J["gvalue"][p]) < 1e-7
        p += 1
    for (sp, idx) in s.jas_b:
        bp, bm = s.b.copy(), s.b.copy()
        bp[sp, idx - 1] += h
        bm[sp, idx - 1] -= h
        assert abs((val_of(b=bp) - val_of(b=bm)) / (2 * h) - J["gvalue"][p]) < 1e-7
        p += 1
    for (t, idx) in s.jas_c:
        cp, cm = s.c.copy(), s.c.copy()
        cp[t, idx - 1] += h
        cm[t, idx - 1] -= h
        cp = system.cusp_exact(cp, s.nordc)
        cm = system.cusp_exact(cm, s.nordc)
        assert abs((val_of(c=cp) - val_of(c=cm)) / (2 * h) - J["gvalue"][p]) < 1e-7
        p += 1


# ---------------------------------------------------------------- determinants
def _brute_psid(s, O, x):
    orb = np.array([O.orbitals(x[i])[0] for i in range(s.nelec)])
    nup = s.nup
    cdet = np.zeros(len(s.det_up))
    for k in range(len(s.csf_coef)):
        for q in range(s.csf_det_off[k], s.csf_det_off[k + 1]):
            cdet[s.csf_det[q]] += s.csf_coef[k] * s.csf_det_coef[q]
    tot = 0.0
    for d in range(len(s.det_up)):
        du = np.linalg.det(orb[:nup][:, s.det_up[d]])
        dd = np.linalg.det(orb[nup:][:, s.det_dn[d]]) if s.ndn else 1.0
        tot += cdet[d] * du * dd
    return tot


@pytest.mark.parametrize("which", ["be", "c2"])
def test_multideterminant_matches_brute_force(which):
    s = _be(4)[0] if which == "be" else system.c2_dimer(2)
    O = oracle.Oracle(s)
    for seed in range(3):
        x = system.random_walkers(s, 1, seed=20 + seed)[0]
        d = O.determinant(x)
        ref = _brute_psid(s, O, x)
        assert math.isclose(d["logdet"], math.log(abs(ref)), rel_tol=1e-12)
        assert d["sign"] == np.sign(ref)


# ---------------------------------------------------------------- local energy
def _hydrogenic(Z):
    s, _ = _be(1)
    bb = system._BasisBuilder(system.QMCW_BASIS_SLATER, 1)
    r = bb.add_radial(1, [float(Z)], [1.0])
    bb.add_function(0, r, 0, 0)
    barr, anorm = bb.arrays(np.array([0], np.int32))
    return system.System(nelec=1, nup=1, cent=np.zeros((1, 3)), iwctype=np.array([0], np.int32),
                         znuc=np.array([float(Z)]), coef=np.array([[anorm[0]]]),
                         det_up=np.array([[0]], np.int32), det_dn=np.zeros((1, 0), np.int32),
                         csf_coef=np.array([1.0]), csf_det_off=np.array([0, 1], np.int32),
                         csf_det=np.array([0], np.int32), csf_det_coef=np.array([1.0]),
                         isc=4, scalek=1.0, norda=2, nordb=2, nordc=0,
                         a4=np.zeros((1, 3)), b=np.array([[0.5, 1.0, 0.0]]), c=np.zeros((1, 0)),
                         **barr)


@pytest.mark.parametrize("Z", [1, 2, 3])
def test_hydrogenic_local_energy(Z):
    s = _hydrogenic(Z)
    O = oracle.Oracle(s)
    rs = np.random.default_rng(Z)
    for _ in range(5):
        x = rs.normal(size=(1, 3))
        e = O.eval_full(x)["eloc"]
        assert abs(e + 0.5 * Z * Z) < 1e-12


@pytest.mark.parametrize("which", ["be", "c2_noPP"])
def test_local_energy_identity(which):
    s = _be(4)[0] if which == "be" else system.c2_dimer(2)
    if which == "c2_noPP":
        s = s.copy()
        s.nloc = 0
    O = oracle.Oracle(s)
    x = system.random_walkers(s, 1, seed=9)[0]
    r = O.eval_full(x)
    def psi(xx):
        q = O.eval_full(xx)
        return q["sign"] * math.exp(q["logpsid"] + q["psij"])
    h = 2e-4
    p0 = psi(x)
    lap = 0.0
    for i in range(s.nelec):
        for k in range(3):
            e = np.zeros_like(x)
            e[i, k] = h
            lap += (-psi(x + 2 * e) + 16 * psi(x + e) - 30 * p0 + 16 * psi(x - e) - psi(x - 2 * e)) / (12 * h * h)
    e_fd = r["pe"] - 0.5 * lap / p0
    assert abs(e_fd - r["eloc"]) < 1e-5 * max(1.0, abs(r["eloc"]))


# ---------------------------------------------------------------- pseudopotential
def test_nonlocal_quadrature_exact_for_sp_orbitals():
    s = system.c2_dimer(2)
    # one centre, one electron, s and p functions only, no Jastrow
    s = s.copy()
    s.cent = np.zeros((1, 3))      # nucleus at the origin: x below is relative to it
    s.iwctype = s.iwctype[:1]
    nbt = s.ctype_bas_off[1]
    keep = [q for q in range(nbt) if s.bas_l[q] <= 1]
    rs = np.random.default_rng(4)
    coef = np.zeros((1, nbt))
    coef[0, keep] = rs.normal(size=len(keep))
    s.coef = coef
    s.nelec, s.nup = 1, 1
    s.det_up = np.array([[0]], np.int32)
    s.det_dn = np.zeros((1, 0), np.int32)
    s.csf_coef = np.array([1.0]); s.csf_det_off = np.array([0, 1], np.int32)
    s.csf_det = np.array([0], np.int32); s.csf_det_coef = np.array([1.0])
    s.a4 = np.zeros_like(s.a4); s.b = np.zeros_like(s.b); s.c = np.zeros_like(s.c)
    O = oracle.Oracle(s)
    x = np.array([[0.31, -0.22, 0.41]])
    r = O.eval_full(x)
    # reference: sum_l v_l(r) (2l+1)/(4 pi) int dOmega' P_l(cos) psi(r')/psi(r)
    rr = np.linalg.norm(x[0])
    psi0 = O.orbitals(x[0])[0][0]
    ct, wt = np.polynomial.legendre.leggauss(40)
    ph = np.linspace(0, 2 * np.pi, 80, endpoint=False)
    ecp = system._ECP_C
    def v(l):
        lst = ecp["l0"] if l == 0 else ecp["l1"]
        return sum(cf * (rr ** (pw - 2) if pw != 2 else 1.0) * math.exp(-ex * rr * rr) for cf, pw, ex in lst)
    xhat = x[0] / rr
    acc = 0.0
    for c, w in zip(ct, wt):
        st_ = math.sqrt(1 - c * c)
        for p in ph:
            q = np.array([st_ * math.cos(p), st_ * math.sin(p), c])
            cos = q @ xhat
            ratio = O.orbitals(rr * q)[0][0] / psi0
            acc += w * (2 * np.pi / len(ph)) / (4 * np.pi) * ratio * (v(0) + 3 * cos * v(1))
    assert abs(r["vpsp"] - acc) < 1e-11 * max(1.0, abs(acc))


# ---------------------------------------------------------------- derivatives
@pytest.mark.parametrize("name", ["c2", "mol"])
def test_parameter_derivatives_fd(name):
    """name = mol: a 39-electron random molecule of seven centres (Z_eff 1,
    4 and 6, pseudopotentials) with its Jastrow set and 16 orbital rotations"""
    if name == "c2":
        s = system.c2_dimer(2, derivs=True)
    else:
        s = system.molecule_derivs(system.random_molecule(1011))
    O = oracle.Oracle(s)
    x = system.random_walkers(s, 1, seed=17)[0]
    r0 = O.eval_full(x)
    dpsi, deloc = O.derivs()
    h = 1e-5

    def lnpsi_el(s2):
        q = oracle.Oracle(s2).eval_full(x)
        return q["logpsid"] + q["psij"], q["eloc"]

    p = 0
    checks = []
    for icsf in s.csf_parm:
        sp, sm = s.copy(), s.copy()
        sp.csf_coef = s.csf_coef.copy(); sp.csf_coef[icsf] += h
        sm.csf_coef = s.csf_coef.copy(); sm.csf_coef[icsf] -= h
        checks.append((p, sp, sm)); p += 1
    for (t, idx) in s.jas_a:
        sp, sm = s.copy(), s.copy()
        sp.a4[t, idx - 1] += h; sm.a4[t, idx - 1] -= h
        checks.append((p, sp, sm)); p += 1
    for (sb, idx) in s.jas_b:
        sp, sm = s.copy(), s.copy()
        sp.b[sb, idx - 1] += h; sm.b[sb, idx - 1] -= h
        checks.append((p, sp, sm)); p += 1
    for (t, idx) in s.jas_c:
        sp, sm = s.copy(), s.copy()
        sp.c[t, idx - 1] += h; sm.c[t, idx - 1] -= h
        sp.c = system.cusp_exact(sp.c, s.nordc); sm.c = system.cusp_exact(sm.c, s.nordc)
        checks.append((p, sp, sm)); p += 1
    for (i, j, rev) in s.orb_parm:
        sp, sm = s.copy(), s.copy()
        for ss_, sg in ((sp, 1), (sm, -1)):
            ci, cj = s.coef[i].copy(), s.coef[j].copy()
            ss_.coef[i] = ci + sg * h * cj
            if rev:
                ss_.coef[j] = cj - sg * h * ci
        checks.append((p, sp, sm)); p += 1
    assert p == len(dpsi)
    rs = np.random.default_rng(0)
    sel = sorted(set(rs.choice(len(checks), min(60, len(checks)), replace=False).tolist()) |
                 set(range(len(s.csf_parm) + 3)))
    for k in sel:
        p, sp, sm = checks[k]
        lp, ep = lnpsi_el(sp)
        lm, em = lnpsi_el(sm)
        assert abs((lp - lm) / (2 * h) - dpsi[p]) < 1e-6 * max(1, abs(dpsi[p])), (p, dpsi[p])
        assert abs((ep - em) / (2 * h) - deloc[p]) < 2e-5 * max(1, abs(deloc[p])), (p, deloc[p])


# ---------------------------------------------------------------- VMC
@pytest.mark.parametrize("iteration", [5])
def test_vmc_energy_matches_champ_transcript(iteration):
    s, run = _be(iteration)
    O = oracle.Oracle(s)
    nw, nstep = 512, 700
    x = system.random_walkers(s, nw, seed=5)
    seeds = rng.walker_seeds("1837465927472523", nw)
    r = O.vmc_run(x, seeds, nstep, 1.0, nthreads=8)
    e = r["eloc"][100:]
    blk = e.reshape(-1, 20, nw).mean(axis=1).mean(axis=1)
    err = blk.std(ddof=1) / math.sqrt(len(blk))
    sig = math.hypot(err, run["error"])
    assert abs(e.mean() - run["energy"]) < 3 * sig, (e.mean(), err, run)
    assert abs(e.std() - run["sigma"]) < 0.01


def _c2_ktt(iteration):
    fx = json.load(open(os.path.join(HERE, "golden", "c2_ktt_transcript.json")))
    s = system.from_champ_old_input(fx["input"], "c2_ktt")
    run = [r for r in fx["runs"] if r["iteration"] == iteration][0]
    p = run["params"]
    if p:
        s.a4 = np.array([p["a"]])
        s.b = np.array([p["b"]])
        s.c = system.cusp_exact(np.array([p["c"]]), s.nordc)
    return s, run


def test_old_format_input_reads_the_c2_example():
    """examples/c2_ktt_hfj_vmc_emin_j_opt3_lin.inp through the old-format
    reader: sizes, the GAMESS-order basis (6 s, then 4 px, 4 py, 4 pz per
    centre, as CHAMP's printout lists them) and the exponents it printed as
    screening constants."""
    s, _ = _c2_ktt(1)
    assert (s.nelec, s.nup, s.ncent, s.nbasis, s.norb, s.isc) == (12, 6, 2, 36, 36, 2)
    assert s.bas_l.tolist() == [0] * 6 + [1] * 12
    assert s.bas_m.tolist() == [0] * 6 + [1] * 4 + [-1] * 4 + [0] * 4
    assert s.rad_n.tolist() == [1, 1, 2, 2, 2, 3] + [2] * 12
    assert np.allclose(s.prim_zeta[:10], [5.450236, 9.584061, 1.304429, 1.899079, 3.753271,
                                          3.753486, 1.058381, 1.648879, 2.808762, 6.781327])
    assert s.det_up.tolist() == [[0, 1, 2, 3, 4, 5]] and s.det_dn.tolist() == [[0, 1, 2, 3, 4, 5]]


@pytest.mark.parametrize("iteration", [1, 2, 3])
def test_vmc_energy_matches_champ_c2_transcript(iteration):
    """All-electron C2 (12 electrons, 36 Slater functions, isc = 2 Jastrow)
    from the old-format input: the oracle's VMC energy and sigma against the
    runs of CHAMP's transcript (iteration 1: the input's Jastrow; 2, 3: the
    optimised ones it printed), within 3 sigma of the combined errors."""
    s, run = _c2_ktt(iteration)
    O = oracle.Oracle(s)
    nw, nstep = 256, 400
    x = system.random_walkers(s, nw, seed=5)
    seeds = rng.walker_seeds("1837465927472523", nw)
    r = O.vmc_run(x, seeds, nstep, 1.0, nthreads=8)
    e = r["eloc"][100:]
    blk = e.reshape(-1, 20, nw).mean(axis=1).mean(axis=1)
    err = blk.std(ddof=1) / math.sqrt(len(blk))
    sig = math.hypot(err, run["error"])
    assert abs(e.mean() - run["energy"]) < 3 * sig, (e.mean(), err, run)
    assert abs(e.std() - run["sigma"]) < 0.1, (e.std(), run)


@pytest.mark.parametrize("case", [("be", 5), ("c2_ktt", 4)])
def test_accelerated_mover_matches_champ_transcript_acceptance(case):
    """CHAMP's default mover (imetro = 6, metrop_slat_mov1) with the
    transcripts' settings (deltar = 5, deltat = 1, fbias = 1): the oracle's
    acceptance ratio against the one each run printed (Be: 234,400 passes,
    C2: 108,400), within 3.5 combined standard errors (binomial, doubled for
    the walk's correlation), and the energy within 3 sigma."""
    name, it = case
    s, run = _be(it) if name == "be" else _c2_ktt(it)
    O = oracle.Oracle(s)
    O.set_mover(6, 5.0, 1.0)
    nw, nstep, neq = (512, 300, 50) if name == "be" else (256, 250, 50)
    x = system.random_walkers(s, nw, seed=5)
    seeds = rng.walker_seeds("1837465927472523", nw)
    r0 = O.vmc_run(x, seeds, neq, 1.0, nthreads=8)
    r = O.vmc_run(r0["x"], r0["rng"], nstep, 1.0, nthreads=8)
    nmove = nw * nstep * s.nelec
    acc = float(r["naccept"].sum()) / nmove
    pa = run["acceptance"]
    se = 2 * math.hypot(math.sqrt(pa * (1 - pa) / nmove),
                        math.sqrt(pa * (1 - pa) / (run["passes"] * s.nelec)))
    assert abs(acc - pa) < 3.5 * se, (acc, pa, se)
    e = r["eloc"]
    blk = e.reshape(-1, 25, nw).mean(axis=1).mean(axis=1)
    err = blk.std(ddof=1) / math.sqrt(len(blk))
    assert abs(e.mean() - run["energy"]) < 3 * math.hypot(err, run["error"]), (e.mean(), err, run)


def test_c2_example_jastrow_derivatives_fd():
    """The 24 Jastrow parameters the C2 example varies (isc = 2 scaled
    distances, read from its optimisation section): d ln Psi / dp and
    d E_L / dp against central finite differences (h = 1e-5), with the
    dependent e-e-n coefficients re-imposed on each side."""
    s, _ = _c2_ktt(3)
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
    for blk, arr in (("a", s.jas_a), ("b", s.jas_b), ("c", s.jas_c)):
        for (t, idx) in arr:
            sp, sm = s.copy(), s.copy()
            name = {"a": "a4", "b": "b", "c": "c"}[blk]
            getattr(sp, name)[t, idx - 1] += h
            getattr(sm, name)[t, idx - 1] -= h
            if blk == "c":
                sp.c = system.cusp_exact(sp.c, s.nordc)
                sm.c = system.cusp_exact(sm.c, s.nordc)
            lp, ep = lnpsi_el(sp)
            lm, em = lnpsi_el(sm)
            assert abs((lp - lm) / (2 * h) - dpsi[p]) < 1e-6 * max(1, abs(dpsi[p])), (blk, idx)
            assert abs((ep - em) / (2 * h) - deloc[p]) < 2e-5 * max(1, abs(deloc[p])), (blk, idx)
            p += 1


def test_accum_products_against_numpy():
    """orc_accum: sums over samples of dpsi dpsi^T, E dpsi dpsi^T, dpsi deloc^T
    and the vector sums, against numpy's matrix products."""
    r = np.random.default_rng(5)
    W, P = 53, 17
    dp, dl, e = r.normal(size=(W, P)), r.normal(size=(W, P)), r.normal(-5, 1, size=W)
    a = oracle.accum(dp, dl, e)
    assert a["n"] == W
    assert np.allclose(a["se"], e.sum(), rtol=1e-14) and np.allclose(a["se2"], (e * e).sum(), rtol=1e-14)
    assert np.allclose(a["sd"], dp.sum(0), rtol=1e-13, atol=1e-13)
    assert np.allclose(a["sde"], dp.T @ e, rtol=1e-13, atol=1e-13)
    assert np.allclose(a["sl"], dl.sum(0), rtol=1e-13, atol=1e-13)
    assert np.allclose(a["dd"], dp.T @ dp, rtol=1e-13, atol=1e-12)
    assert np.allclose(a["dde"], (dp * e[:, None]).T @ dp, rtol=1e-13, atol=1e-12)
    assert np.allclose(a["dl"], dp.T @ dl, rtol=1e-13, atol=1e-12)
    # the deloc product is not symmetric; dpsi dpsi is
    assert np.array_equal(a["dd"], a["dd"].T)
    assert not np.allclose(a["dl"], a["dl"].T)
