"""torch.autograd.Function over the engine's adjoint derivatives, for the
test harness only (the product never imports torch).

forward(d) sets the parameters to the system's own plus the displacement d
(in CHAMP's parameter order: CSF, Jastrow a, b, c, orbital rotations;
qmc/deriv_mod.f90:219-239), evaluates the walkers through qmcw_eval_full and
returns (ln|Psi|, E_L) per walker.  backward uses the engine's
d ln Psi/dp and dE_L/dp (qmcw_derivs) at the same parameters.  Orbital
rotations enter to first order (coef_i += k coef_j, and coef_j -= k coef_i
for a pair that CHAMP rotates both ways), which is exact for the derivative
at d = 0.  CPU tensors only: torch never touches the GPU here."""
import numpy as np
import torch

from champ_amd import system


def params_at(s, d):
    """(a4, b, c, csf_coef, coef) of system s displaced by d."""
    d = np.asarray(d, np.float64)
    a4, b, c = s.a4.copy(), s.b.copy(), s.c.copy()
    cc, coef = np.asarray(s.csf_coef, np.float64).copy(), s.coef.copy()
    p = 0
    for icsf in s.csf_parm:
        cc[icsf] += d[p]; p += 1
    for (t, idx) in s.jas_a:
        a4[t, idx - 1] += d[p]; p += 1
    for (sb, idx) in s.jas_b:
        b[sb, idx - 1] += d[p]; p += 1
    for (t, idx) in s.jas_c:
        c[t, idx - 1] += d[p]; p += 1
    if s.jas_c:
        c = system.cusp_exact(c, s.nordc)
    for (i, j, rev) in s.orb_parm:
        coef[i] += d[p] * s.coef[j]
        if rev:
            coef[j] -= d[p] * s.coef[i]
        p += 1
    assert p == len(d)
    return a4, b, c, cc, coef


class PsiEloc(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d, engine, s):
        engine.set_params(*params_at(s, d.detach().numpy()))
        g = engine.psi_at()
        ctx.engine, ctx.s = engine, s
        ctx.save_for_backward(d)
        return (torch.from_numpy(g["logpsid"] + g["psij"]), torch.from_numpy(g["eloc"]))

    @staticmethod
    def backward(ctx, g_lnpsi, g_eloc):
        (d,) = ctx.saved_tensors
        ctx.engine.set_params(*params_at(ctx.s, d.detach().numpy()))
        dpsi, deloc = ctx.engine.derivs()
        grad = torch.from_numpy(dpsi).T @ g_lnpsi + torch.from_numpy(deloc).T @ g_eloc
        return grad, None, None
