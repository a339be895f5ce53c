"""Host driver of a DMC run over the engine's generations, following the
outer loop of CHAMP's rundmc (qmc/rundmc_mod.f90:97-330).

Per generation (rundmc_mod.f90:100-112):
    pop_ctrl = exp((eest - etrial) taueff) (wtarg / wtot)^(1 / ngen_pop_ctrl)
    ff(ipass mod nfprod) = 1 / pop_ctrl,  undo_pop_ctrl = prod(ff)
the engine moves, reweights and branches the population (qmcw_dmc_step);
then the weighted averages of E_L and taunow give eest and taueff
(rundmc_mod.f90:386-398).  esigma is 0.2 sqrt(nelec) for the first
max(10, 10/tau) generations and sqrt(wtarg) sigma(E_L) afterwards
(rundmc_mod.f90:235-242).

`stepper` is any object with dmc_step(params) -> stats (the Engine), so the
same driver can also run other implementations of a generation.
"""
from __future__ import annotations

import math

from .abi import DmcParams


class DmcDriver:
    def __init__(self, stepper, nelec: int, nwalk0: int, tau: float, etrial: float,
                 nfprod: int = 50, **params):
        self.stepper = stepper
        self.nelec = nelec
        self.tau = tau
        self.etrial = etrial
        self.wtarg = float(nwalk0)
        self.wtot = float(nwalk0)
        self.ngen_pop_ctrl = math.ceil(1.0 / tau)
        self.nfprod = nfprod
        self.ff = [1.0] * (nfprod + 1)
        self.ipass = 0
        self.eest = etrial
        self.taueff = tau
        self.params = params
        self.is_ = 0
        # running weighted sums (weights wg = w undo_pop_ctrl)
        self.sw = self.swe = self.swe2 = self.swt = 0.0
        self.history = []

    def generation(self):
        self.is_ += 1
        pop_ctrl = math.exp((self.eest - self.etrial) * self.taueff) * \
            (self.wtarg / self.wtot) ** (1.0 / self.ngen_pop_ctrl)
        self.ipass += 1
        self.ff[self.ipass % self.nfprod] = 1.0 / pop_ctrl
        undo = 1.0
        for f in self.ff[:self.nfprod]:
            undo *= f
        if self.is_ > max(10, round(10.0 / self.tau)) and self.sw > 0:
            var = max(self.swe2 / self.sw - (self.swe / self.sw) ** 2, 0.0)
            esigma = math.sqrt(self.wtarg) * math.sqrt(var / max(self.is_ - 1, 1))
        else:
            esigma = 0.2 * math.sqrt(self.nelec)
        p = DmcParams(self.tau, self.etrial, eest=self.eest, esigma=esigma, pop_ctrl=pop_ctrl,
                      taueff=self.taueff, **self.params)
        st = self.stepper.dmc_step(p)
        self.wtot = st.wsum
        wg = undo
        self.sw += wg * st.wsum
        self.swe += wg * st.wesum
        self.swe2 += wg * st.we2sum
        self.swt += wg * st.wtausum
        self.eest = self.swe / self.sw
        self.taueff = self.swt / self.sw
        self.history.append(dict(eave=st.wesum / st.wsum, wsum=st.wsum, nwalk=st.nwalk,
                                 acc=st.nacc / max(st.ntry, 1), pop_ctrl=pop_ctrl))
        return st

    def end_equilibration(self, reset_etrial: bool = True):
        """End of an equilibration run (qmc/dmc_mod.f90:380-403): with
        reset_etrial, etrial becomes eest and every population-control
        factor ff is multiplied by exp((eest - etrial) taueff), taueff the
        weighted average of the run; then the running averages are cleared
        (stats%clear), so that eest follows the generations after this
        call only, and the esigma schedule starts again as rundmc's step
        counter does.  eest keeps its value until the next generation."""
        if reset_etrial and self.sw > 0:
            dff = math.exp((self.eest - self.etrial) * (self.swt / self.sw))
            self.ff = [f * dff for f in self.ff]
            self.etrial = self.eest
        self.sw = self.swe = self.swe2 = self.swt = 0.0
        self.is_ = 0

    def run(self, ngen: int):
        for _ in range(ngen):
            self.generation()
        return self.history[-ngen:]
