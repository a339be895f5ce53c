"""The adjoint derivatives through torch.autograd (tests/torch_adjoint.py):
torch.autograd.gradcheck compares the backward (qmcw_derivs) with central
finite differences of forward evaluations of the engine at displaced
parameters (qmcw_set_params + qmcw_eval_full), for all 242 parameters of the
C2-like system (24 CSF, 40 Jastrow, ~180 orbital rotations) and both
outputs, ln|Psi| and E_L."""
import numpy as np
import pytest

from champ_amd import system
from champ_amd.engine import Engine

pytestmark = pytest.mark.gpu


def test_gradcheck_of_psi_and_eloc_through_the_adjoint():
    import os, sys
    sys.path.insert(0, os.path.dirname(__file__))
    import torch
    from torch_adjoint import PsiEloc
    s = system.c2_dimer(2, derivs=True)
    W = 3
    e = Engine(s, W)
    e.set_walkers(system.random_walkers(s, W, seed=17))
    d = torch.zeros(s.nparm, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda dd: PsiEloc.apply(dd, e, s), (d,), eps=1e-6,
                                    atol=1e-5, rtol=1e-4)
    # and a loss through both outputs: the gradient torch assembles
    lnpsi, el = PsiEloc.apply(d, e, s)
    loss = (el * torch.exp(2 * lnpsi - 2 * lnpsi.detach())).mean()
    loss.backward()
    dpsi, deloc = e.derivs()
    el0 = el.detach().numpy()
    ref = (deloc + 2 * dpsi * el0[:, None]).mean(axis=0)
    assert np.allclose(d.grad.numpy(), ref, rtol=1e-10, atol=1e-12)
    e.close()
