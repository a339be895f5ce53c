"""The linear method's update (champ_amd/linear_method.py, after
qmc/opt_lin_mod.f90) pinned by the strong zero-variance principle on the
1D harmonic oscillator H = -1/2 d2/dx2 + x^2/2 with samples drawn exactly
from |Psi|^2: when the exact ground state is reachable from Psi and its
first derivatives, the right eigenvector of the sampled H v = lambda S v is
that state for ANY finite sample.

* Nonlinear parameter, Psi = exp(-a x^2 / 2): at a = 1 the update is zero
  to rounding although the sample mean of dE_L/da (the H_0j entries a
  transposed H would use) is not; from a = 0.7 the iteration converges to 1.
* Linear (CSF-like) parameters, Psi = phi0 + c1 phi2 + c2 phi4 with phi_n
  the oscillator's eigenstates: one update lands exactly on c = 0, which
  checks the semiorthogonal first-coefficient formula for CSF parameters.
"""
import numpy as np
import pytest

from champ_amd import linear_method as lm


def _sums(dpsi, eloc, deloc):
    dpsi, deloc = np.atleast_2d(dpsi.T).T, np.atleast_2d(deloc.T).T
    return dict(n=len(eloc), se=eloc.sum(), se2=(eloc * eloc).sum(), sd=dpsi.sum(0),
                sde=dpsi.T @ eloc, sl=deloc.sum(0), dd=dpsi.T @ dpsi,
                dde=(dpsi * eloc[:, None]).T @ dpsi, dl=dpsi.T @ deloc)


def _gauss_samples(a, n, seed):
    return np.random.default_rng(seed).normal(0.0, np.sqrt(0.5 / a), n)   # |Psi|^2 = exp(-a x^2)


def _gauss_terms(a, x):
    return -0.5 * x * x, 0.5 * a + 0.5 * x * x * (1 - a * a), 0.5 - a * x * x


def test_exact_nonlinear_parameter_gives_zero_update():
    x = _gauss_samples(1.0, 4000, 1)
    dp, el, de = _gauss_terms(1.0, x)
    assert abs(de.mean()) > 1e-3                         # H_0j != 0 on this sample
    r = lm.step(_sums(dp, el, de), nparmcsf=0)
    assert abs(r.delta[0]) < 1e-12
    assert abs(r.eigval - 0.5) < 1e-7


def test_nonlinear_parameter_iteration_converges():
    a = 0.7
    for it in range(8):
        x = _gauss_samples(a, 20000, 10 + it)
        r = lm.step(_sums(*_gauss_terms(a, x)), nparmcsf=0)
        a += r.delta[0]
    assert abs(a - 1.0) < 0.02


def _phi(x):
    g = np.exp(-0.5 * x * x)
    return g, (2 * x * x - 1) * g, (4 * x ** 4 - 12 * x * x + 3) * g      # n = 0, 2, 4 (unnormalised)


@pytest.mark.parametrize("c", [(0.3,), (-0.45,), (0.2, -0.05)])
def test_linear_parameters_reach_the_exact_state_in_one_step(c):
    c = np.array(c)
    k = len(c)
    en = np.array([0.5, 2.5, 4.5])
    # exact sampling of |Psi|^2 by rejection from a wide gaussian
    rs = np.random.default_rng(7)
    xs = []
    while sum(len(v) for v in xs) < 6000:
        x = rs.normal(0, 1.6, 20000)
        ph = _phi(x)
        psi = ph[0] + sum(c[j] * ph[j + 1] for j in range(k))
        env = np.exp(-x * x / (2 * 1.6 ** 2))
        ratio = psi * psi / env
        xs.append(x[rs.random(len(x)) * ratio.max() * 1.01 < ratio])
    x = np.concatenate(xs)[:6000]
    ph = _phi(x)
    psi = ph[0] + sum(c[j] * ph[j + 1] for j in range(k))
    hpsi = en[0] * ph[0] + sum(c[j] * en[j + 1] * ph[j + 1] for j in range(k))
    el = hpsi / psi
    dp = np.stack([ph[j + 1] / psi for j in range(k)], axis=1)
    de = np.stack([(en[j + 1] * ph[j + 1] - el * ph[j + 1]) / psi for j in range(k)], axis=1)
    r = lm.step(_sums(dp, el, de), nparmcsf=k, diag_stab=0.0, add_diag_ovlp=0.0)
    assert np.allclose(c + r.delta, 0.0, atol=1e-9), c + r.delta
    assert abs(r.eigval - 0.5) < 1e-9


def test_update_solves_the_sampled_eigenproblem():
    """On the engine's parameter layout (the C2-like system's 242
    parameters, oracle samples): the chosen vector and eigenvalue satisfy
    H v = lambda S v to rounding, the eigenvalue lies in CHAMP's window, and
    the update is the vector divided by its (semiorthogonal) first entry."""
    import oracle
    from champ_amd import system
    s = system.c2_dimer(2, derivs=True)
    x = system.random_walkers(s, 400, seed=4)
    r0 = oracle.Oracle(s).eval_batch(x, derivs=True, nthreads=8)
    acc = oracle.accum(r0["dpsi"], r0["deloc"], r0["out9"][:, 3])
    r = lm.step(acc, nparmcsf=len(s.csf_parm))
    assert r.delta.shape == (s.nparm,) and np.all(np.isfinite(r.delta))
    v = r.vector
    res = r.ham @ v - r.eigval * (r.ovlp @ v)
    assert np.abs(res).max() <= 1e-9 * np.abs(r.ham).max() * np.abs(v).max()
    assert r.energy - r.sigma < r.eigval < r.energy + 0.01
    nc = len(s.csf_parm)
    first = 1.0 - (v[1:1 + nc] / v[0]) @ (acc["sd"][:nc] / acc["n"])
    assert np.allclose(r.delta, v[1:] / v[0] / first, rtol=1e-12, atol=1e-14)
