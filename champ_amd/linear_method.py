"""The linear method's parameter update from the engine's product sums.

The sums come from qmcw_accum_get (Engine.accum_get): the number of samples,
the sums of E_L and E_L^2, of d ln Psi/dp (dpsi), E_L dpsi and dE_L/dp
(deloc), and the P x P sums of dpsi_i dpsi_j, E_L dpsi_i dpsi_j and
dpsi_i deloc_j.  This module turns them into the parameter variations of
one linear-method iteration, as qmc/opt_lin_mod.f90 does with CHAMP's
defaults for energy minimisation:

* averages and covariances (qmc/deriv_mod.f90:756, 1447, 1515);
* the overlap matrix of Psi and its first derivatives, S (ovlp_lin_bld,
  qmc/opt_lin_mod.f90:249-330, add_diag_ovlp on its diagonal), optionally
  renormalised by sqrt(S_ii) (ovlp_lin_renorm_bld, :332-370);
* the Hamiltonian matrix H (ham_lin_energy_bld, :404-520): H_00 = <E_L>,
  H_i0 = cov(dpsi_i, E_L), H_0j = cov(dpsi_j, E_L) + <deloc_j>,
  H_ij = <dpsi_i dpsi_j E_L> - <dpsi_j><dpsi_i E_L> - <dpsi_i><dpsi_j E_L>
         + <dpsi_i><dpsi_j><E_L> + cov(dpsi_i, deloc_j);
* stabilisation by diag_stab on the parameter diagonal of H
  (ham_lin_renorm_stab_bld, 'identity', :590-630);
* the generalised eigenproblem H v = lambda S v solved by LAPACK's dggev
  (scipy.linalg.eig), as delta_lin_bld does (:885-911);
* CHAMP's eigenvector choice (:950-1125): the lowest eigenvalue within
  [E - sigma, E + 0.01] unless the norm of its wave-function change for the
  nonlinear parameters exceeds 100 times the smallest such norm, else the
  eigenvector of smallest norm;
* the 'semiorthogonal' update with xi (:1226-1250): delta_p = v_p / (v_0 -
  sum over CSF parameters of v_i <dpsi_i> + (1 - xi) n / ((1 - xi) + xi (1
  + n))), n the norm above.

The update is host work once per optimisation iteration (P ~ 250).  The
variance-minimisation mix (p_var > 0) and the left-eigenvector and
excited-state options of the reference are not provided.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.linalg


@dataclass
class LinearStep:
    delta: np.ndarray          # parameter variations [P]
    eigval: float              # eigenvalue of the chosen eigenvector
    eigvals: np.ndarray        # all eigenvalues (real parts)
    chosen: int                # index of the chosen eigenvector
    norm_change: float         # norm of the linear wave-function change (nonlinear parameters)
    energy: float              # <E_L> of the samples
    sigma: float               # sqrt(<E_L^2> - <E_L>^2)
    ham: np.ndarray            # H (augmented, stabilised, as solved)
    ovlp: np.ndarray           # S (augmented, as solved)
    vector: np.ndarray         # the chosen eigenvector of (ham, ovlp)


def matrices(acc: dict, add_diag_ovlp: float = 1e-8):
    """(H, S, averages) of the linear method, (P+1) x (P+1), from the sums."""
    n = float(acc["n"])
    if n <= 0:
        raise ValueError("no samples in the product sums")
    e = acc["se"] / n
    dp = np.asarray(acc["sd"]) / n
    dpe = np.asarray(acc["sde"]) / n
    dl = np.asarray(acc["sl"]) / n
    dd = np.asarray(acc["dd"]) / n
    dde = np.asarray(acc["dde"]) / n
    ddl = np.asarray(acc["dl"]) / n
    P = dp.size
    dd_cov = dd - np.outer(dp, dp)
    de_cov = dpe - dp * e
    ddl_cov = ddl - np.outer(dp, dl)
    S = np.zeros((P + 1, P + 1))
    S[0, 0] = 1.0
    S[1:, 1:] = dd_cov
    S[np.diag_indices(P + 1)] += add_diag_ovlp
    H = np.zeros((P + 1, P + 1))
    H[0, 0] = e
    H[1:, 0] = de_cov
    H[0, 1:] = de_cov + dl
    H[1:, 1:] = dde - np.outer(dpe, dp) - np.outer(dp, dpe) + np.outer(dp, dp) * e + ddl_cov
    avg = dict(eloc=e, eloc2=acc["se2"] / n, dpsi=dp)
    return H, S, avg


def step(acc: dict, nparmcsf: int, diag_stab: float = 1e-8, add_diag_ovlp: float = 1e-8,
         renormalize: bool = False, xi: float = 1.0) -> LinearStep:
    """One linear-method update from the product sums (parameters in
    CHAMP's order, the first nparmcsf of them CSF coefficients)."""
    H, S, avg = matrices(acc, add_diag_ovlp)
    P = H.shape[0] - 1
    ren = np.sqrt(np.diag(S)) if renormalize else np.ones(P + 1)
    Sr = S / np.outer(ren, ren)
    Hr = H / np.outer(ren, ren)
    Hr[np.arange(1, P + 1), np.arange(1, P + 1)] += diag_stab
    w, V = scipy.linalg.eig(Hr, Sr)               # dggev: right eigenvectors
    lam = np.real(w)
    V = np.real(V)
    e = avg["eloc"]
    sigma = float(np.sqrt(max(0.0, avg["eloc2"] - e * e)))
    nl = slice(1 + nparmcsf, P + 1)
    Snl = S[nl, nl] / np.outer(ren[nl], ren[nl])

    def norm_change(k):
        v = V[:, k]
        return float(v[nl] @ Snl @ v[nl]) / v[0] ** 2

    norms = np.array([norm_change(k) for k in range(P + 1)])
    # smallest non-negative norm, lowest eigenvalue among equal norms
    smallest, ismall = 9.0e99, 0
    for k in range(P + 1):
        if 0.0 <= norms[k] < smallest or (norms[k] == smallest and lam[k] < lam[ismall]):
            smallest, ismall = norms[k], k
    upper, lower = e + 0.01, e - sigma
    lowest, ilow = 9.0e99, -1
    for k in range(P + 1):
        if lam[k] < lowest and lower < lam[k] < upper:
            lowest, ilow = lam[k], k
        if ilow >= 0 and lam[k] == lowest and abs(V[0, k]) > abs(V[0, ilow]):
            ilow = k
    low_norm = norms[ilow] if ilow >= 0 and norms[ilow] >= 0.0 else 9.0e99
    k = ilow if ilow >= 0 and low_norm < 100 * smallest else ismall
    vec = V[:, k].copy()
    v = V[:, k] / ren
    v = v / v[0]
    nrm = float(v[nl] @ S[nl, nl] @ v[nl])
    first = v[0] - float(v[1:1 + nparmcsf] @ avg["dpsi"][:nparmcsf])
    first += (1.0 - xi) * nrm / ((1.0 - xi) + xi * (1.0 + nrm))
    return LinearStep(delta=v[1:] / first, eigval=float(lam[k]), eigvals=lam, chosen=int(k),
                      norm_change=nrm, energy=float(e), sigma=sigma, ham=Hr, ovlp=Sr, vector=vec)
