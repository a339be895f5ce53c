"""Host-side mirror of CHAMP's per-walker psi call surface over the C ABI.

:class:`Engine` owns one ``qmcw_handle`` (W walkers resident in GPU memory).
Its methods keep the names of the reference routines they replace:

=====================  =====================================================
``psi_at``             qmc/psi_utils_mod.f90:13-42 (full evaluation, hpsi)
``move_1e``            qmc/psi_utils_mod.f90:45-90 (one-electron move, hpsie)
``copy_psi_1e``        qmc/psi_utils_mod.f90:299-387 (commit a move)
``calc_1e_derivatives`` qmc/psi_utils_mod.f90:457-483
``derivs``             qmc/deriv_mod.f90:248-556 (dpsi, deloc)
``vmc``                qmc/vmc/vmc.f90:250-331 with metrop_mov1's proposal
=====================  =====================================================

Errors from the library raise :class:`QmcwError` carrying the status code
and the library's message (the reference stops the program instead).
There is no CPU fallback: the HIP library must be built and a GPU present.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

_PREC = {"fp64": abi.QMCW_FP64, "mixed": abi.QMCW_MIXED, "mixed_bf16": abi.QMCW_MIXED_BF16}


class QmcwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qmcw error {code}: {msg}")
        self.code = code


def _d(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class Engine:
    def __init__(self, system, nwalk: int, precision: str = "fp64"):
        self.lib = abi.load_library()
        self.system = system
        self.N = system.nelec
        self.W = int(nwalk)
        self._struct = system.struct()
        h = C.c_void_p()
        st = self.lib.qmcw_create(C.byref(self._struct), self.W, _PREC[precision], C.byref(h))
        if st != 0:
            raise QmcwError(st, self.lib.qmcw_last_error(None).decode())
        self.h = h
        self.P = self.lib.qmcw_nparm(h)

    def _chk(self, st):
        if st != 0:
            raise QmcwError(st, self.lib.qmcw_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.qmcw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- walkers and streams
    def set_walkers(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(self.W, self.N, 3)
        self._chk(self.lib.qmcw_set_walkers(self.h, _d(x)))

    def get_walkers(self):
        x = np.zeros((self.W, self.N, 3))
        self._chk(self.lib.qmcw_get_walkers(self.h, _d(x)))
        return x

    def set_rng(self, seeds):
        s = np.ascontiguousarray(seeds, np.int32).reshape(self.W, 4)
        self._chk(self.lib.qmcw_set_rng(self.h, s.ctypes.data_as(C.POINTER(C.c_int32))))

    def get_rng(self):
        s = np.zeros((self.W, 4), np.int32)
        self._chk(self.lib.qmcw_get_rng(self.h, s.ctypes.data_as(C.POINTER(C.c_int32))))
        return s

    def set_params(self, a4=None, b=None, c=None, csf_coef=None, coef=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)
        arrs = [conv(a) for a in (a4, b, c, csf_coef, coef)]
        self._keep = arrs
        self._chk(self.lib.qmcw_set_params(self.h, *[_d(a) for a in arrs]))

    # ---- psi call surface
    def psi_at(self):
        W, N = self.W, self.N
        out = dict(logpsid=np.zeros(W), sign=np.zeros(W), psij=np.zeros(W),
                   grad=np.zeros((W, N, 3)), divv=np.zeros((W, N)), eloc=np.zeros(W),
                   epot=np.zeros(W))
        self._chk(self.lib.qmcw_eval_full(self.h, _d(out["logpsid"]), _d(out["sign"]),
                                          _d(out["psij"]), _d(out["grad"]), _d(out["divv"]),
                                          _d(out["eloc"]), _d(out["epot"])))
        return out

    eval_full = psi_at

    def move_1e(self, ie: int, xnew):
        W = self.W
        xn = np.ascontiguousarray(xnew, np.float64).reshape(W, 3)
        ratio, grad, lap = np.zeros(W), np.zeros((W, 3)), np.zeros(W)
        self._chk(self.lib.qmcw_ratio_1e(self.h, int(ie), _d(xn), _d(ratio), _d(grad), _d(lap)))
        return dict(ratio=ratio, grad=grad, lap=lap)

    ratio_1e = move_1e

    def copy_psi_1e(self, ie: int, accept):
        a = np.ascontiguousarray(np.asarray(accept).astype(np.uint8)).reshape(self.W)
        self._chk(self.lib.qmcw_accept_1e(self.h, int(ie),
                                          a.ctypes.data_as(C.POINTER(C.c_uint8))))

    accept_1e = copy_psi_1e

    def calc_1e_derivatives(self, ie: int):
        grad, lap = np.zeros((self.W, 3)), np.zeros(self.W)
        self._chk(self.lib.qmcw_grad_1e(self.h, int(ie), _d(grad), _d(lap)))
        return grad, lap

    grad_1e = calc_1e_derivatives

    def derivs(self):
        """Full evaluation with parameter derivatives: (dpsi, deloc), [W][P]."""
        dp, de = np.zeros((self.W, self.P)), np.zeros((self.W, self.P))
        self._chk(self.lib.qmcw_derivs(self.h, _d(dp), _d(de)))
        return dp, de

    def sample_derivs(self, enable=True):
        """Every VMC step's full evaluation also forms the derivatives;
        enable=2 also adds them to the product sums (accum_get)."""
        self._chk(self.lib.qmcw_sample_derivs(self.h, int(enable)))

    def get_derivs(self):
        """(dpsi, deloc) of the last derivative evaluation."""
        dp, de = np.zeros((self.W, self.P)), np.zeros((self.W, self.P))
        self._chk(self.lib.qmcw_get_derivs(self.h, _d(dp), _d(de)))
        return dp, de

    # ---- sums of the linear method's derivative products (qmcw_accum_*)
    def accum_reset(self):
        self._chk(self.lib.qmcw_accum_reset(self.h))

    def accum_add(self):
        """Add the last derivative sample of every walker."""
        self._chk(self.lib.qmcw_accum_add(self.h))

    def accum_add_static(self, acc_device_ptr: int):
        """qmcw_accum_add_static: add the last derivative sample of every
        walker to a caller-owned device accumulator (QMCW_ACCUM_DOUBLES(P)
        doubles at `acc_device_ptr`), with no device allocation."""
        self._chk(self.lib.qmcw_accum_add_static(self.h, C.c_void_p(int(acc_device_ptr))))

    def accum_get(self):
        """dict: n, se, se2 (sums of 1, E_L, E_L^2); sd, sde, sl [P] (sums of
        dpsi, E_L dpsi, deloc); dd, dde, dl [P][P] (sums of dpsi_i dpsi_j,
        E_L dpsi_i dpsi_j, dpsi_i deloc_j)."""
        P = self.P
        o = np.zeros(3 + 3 * P + 3 * P * P)
        self._chk(self.lib.qmcw_accum_get(self.h, _d(o)))
        v = o[3:3 + 3 * P].reshape(3, P)
        m = o[3 + 3 * P:].reshape(3, P, P)
        return dict(n=o[0], se=o[1], se2=o[2], sd=v[0], sde=v[1], sl=v[2], dd=m[0], dde=m[1], dl=m[2])

    # ---- drivers
    def set_reinversion(self, every: int):
        """Rebuild the Slater inverses after every `every` one-electron moves
        (0: only at each step's full evaluation)."""
        self._chk(self.lib.qmcw_set_reinversion(self.h, int(every)))

    def reinvert(self):
        self._chk(self.lib.qmcw_reinvert(self.h))

    def set_vmc_mover(self, imetro: int, deltar: float = 5.0, deltat: float = 1.0):
        """CHAMP's imetro: 1 metrop_mov1 (default), 6 metrop_slat_mov1 with
        radial step multiplier deltar and cos(theta) step deltat."""
        self._chk(self.lib.qmcw_set_vmc_mover(self.h, int(imetro), float(deltar), float(deltat)))

    def vmc(self, nstep: int, delta: float, fbias: float = 1.0, trace: bool = True):
        tr = np.zeros((nstep, self.W)) if trace else None
        nacc = np.zeros(self.W, np.int64)
        self._chk(self.lib.qmcw_vmc_run(self.h, int(nstep), float(delta), float(fbias), _d(tr),
                                        nacc.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(eloc=tr, naccept=nacc)

    def vmc_device(self, nstep: int, delta: float, fbias: float = 1.0):
        self._chk(self.lib.qmcw_vmc_run_device(self.h, int(nstep), float(delta), float(fbias)))

    # ---- DMC (qmc/rundmc_mod.f90)
    def dmc_begin(self, nwalk: int, weight=None):
        w = None if weight is None else np.ascontiguousarray(weight, np.float64)
        self._chk(self.lib.qmcw_dmc_begin(self.h, int(nwalk), _d(w)))

    def dmc_step(self, params):
        """One generation (moves, reweighting, split_join); returns the
        qmcw_dmc_stats of the generation."""
        st = abi.DmcStats()
        self._chk(self.lib.qmcw_dmc_step(self.h, C.byref(params), C.byref(st)))
        return st

    def dmc_population(self):
        """(x[nwalk][N][3], weight[nwalk], rng[nwalk][4]) of the population."""
        n = C.c_int32(0)
        self._chk(self.lib.qmcw_dmc_get(self.h, C.byref(n), None))
        w = np.zeros(max(n.value, 1))
        self._chk(self.lib.qmcw_dmc_get(self.h, C.byref(n), _d(w)))
        return self.get_walkers()[:n.value], w[:n.value], self.get_rng()[:n.value]

    def set_branch_rng(self, seed):
        s = np.ascontiguousarray(seed, np.int32).reshape(4)
        self._chk(self.lib.qmcw_set_branch_rng(self.h, s.ctypes.data_as(C.POINTER(C.c_int32))))

    def get_branch_rng(self):
        s = np.zeros(4, np.int32)
        self._chk(self.lib.qmcw_get_branch_rng(self.h, s.ctypes.data_as(C.POINTER(C.c_int32))))
        return s

    def get_eloc(self):
        e = np.zeros(self.W)
        self._chk(self.lib.qmcw_get_eloc(self.h, _d(e)))
        return e

    def autotune(self) -> int:
        g = C.c_int32(0)
        self._chk(self.lib.qmcw_autotune(self.h, C.byref(g)))
        return g.value

    def group_widths(self):
        """Lane-group widths per kernel class (Jastrow, determinants,
        pseudopotential, moves)."""
        g = np.zeros(4, np.int32)
        self._chk(self.lib.qmcw_get_group_widths(self.h, g.ctypes.data_as(C.POINTER(C.c_int32))))
        return [int(v) for v in g]

    def set_group_width(self, g: int):
        self._chk(self.lib.qmcw_set_group_width(self.h, int(g)))

    def profile(self, enable: bool = True):
        self._chk(self.lib.qmcw_profile(self.h, int(enable)))

    def profile_get(self):
        """{kernel name: (total ms, launches)} of the launches timed since
        profile(True)."""
        n = len(abi.PROF_KERNELS)
        ms, cnt = np.zeros(n), np.zeros(n, np.int64)
        self._chk(self.lib.qmcw_profile_get(self.h, _d(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64))))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(abi.PROF_KERNELS)}

    def sync(self):
        self._chk(self.lib.qmcw_sync(self.h))


class Batch:
    """Several engines (systems of uneven electron and basis counts) whose
    VMC steps run as one launch per kernel and lane-group width
    (qmcw_batch_*).  The engines must share the precision and the e-e-n
    order; each keeps its own widths.  Close the batch before the engines."""

    def __init__(self, engines):
        self.lib = abi.load_library()
        self.engines = list(engines)
        hs = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        b = C.c_void_p()
        st = self.lib.qmcw_batch_create(hs, len(self.engines), C.byref(b))
        if st != 0:
            raise QmcwError(st, self.lib.qmcw_batch_last_error(None).decode())
        self.b = b

    def vmc_device(self, nstep: int, delta: float, fbias: float = 1.0):
        """nstep VMC steps of every engine (a full evaluation first), as each
        engine's vmc_device would run them alone."""
        st = self.lib.qmcw_batch_vmc_run(self.b, int(nstep), float(delta), float(fbias))
        if st != 0:
            raise QmcwError(st, self.lib.qmcw_batch_last_error(self.b).decode())

    def close(self):
        if getattr(self, "b", None):
            self.lib.qmcw_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
