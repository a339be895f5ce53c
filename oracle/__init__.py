"""ctypes bindings of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
Parity status of each oracle function is listed in the module docstring of
tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_here, "liboracle.so")
REF_LIB = os.path.join(_here, "_ref", "libchamp_ref.so")

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int32)
_vp = C.c_void_p


def build(ref: bool = True):
    """Compile liboracle.so (and _ref/libchamp_ref.so when the reference
    sources are present)."""
    subprocess.run(["make", "-s", "-C", _here, "liboracle.so"], check=True)
    if ref and os.path.isdir("/root/reference/qmc"):
        subprocess.run(["make", "-s", "-C", _here, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = C.CDLL(LIB)
        sig = {
            "orc_create": (_vp, [_vp]),
            "orc_destroy": (None, [_vp]),
            "orc_error": (C.c_char_p, []),
            "orc_nterms": (C.c_int32, [_vp]),
            "orc_nparm": (C.c_int32, [_vp]),
            "orc_asymp": (None, [_vp, _pd, _pd]),
            "orc_cusp": (C.c_int32, [_vp, _pd, _pi]),
            "orc_cuspexact": (None, [_vp, _pd]),
            "orc_set_emulation": (None, [_vp, C.c_int32]),
            "orc_set_mover": (None, [_vp, C.c_int32, C.c_double, C.c_double]),
            "orc_slat_sample": (None, [_vp, _vp, C.c_int32, C.c_double, _pi, _pd, _pd]),
            "orc_slat_density": (C.c_double, [_vp, _vp, C.c_int32, C.c_double, _pd]),
            "orc_slat_area": (C.c_double, [C.c_double] * 5),
            "orc_set_params": (None, [_vp, _pd, _pd, _pd, _pd, _pd]),
            "orc_det_info": (None, [_vp, _pi]),
            "orc_setrn": (None, [_pi, _pi]),
            "orc_rannyu": (C.c_double, [_pi]),
            "orc_gauss": (C.c_double, [_pi]),
            "orc_quadrature": (None, [_vp, _pd, _pd, _pd, _pd]),
            "orc_num_radial": (None, [_vp, C.c_int32, C.c_int32, _pd, _pd]),
            "orc_solid_harmonic": (None, [C.c_int32, C.c_int32, _pd, _pd]),
            "orc_rotqua": (None, [_pi, _pd, _pd, _pd, C.c_int32, _pd, _pd, _pd]),
            "orc_matinv": (C.c_double, [_pd, C.c_int32]),
            "orc_basis": (None, [_vp, _pd, _pd, _pd, _pd]),
            "orc_orbitals": (None, [_vp, _pd, _pd, _pd, _pd]),
            "orc_walker_create": (_vp, [_vp]),
            "orc_walker_destroy": (None, [_vp]),
            "orc_walker_copy": (None, [_vp, _vp, _vp]),
            "orc_walker_set_quad": (None, [_vp, _pd, _pd, _pd, C.c_int]),
            "orc_eval_full": (C.c_int, [_vp, _vp, _pd, _pd, _pd, _pd]),
            "orc_jastrow_full": (None, [_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
            "orc_determinant_full": (C.c_int, [_vp, _pd, _pd, _pd, _pd, _pd]),
            "orc_ratio_1e": (C.c_int, [_vp, _vp, C.c_int32, _pd, _pd, _pd, _pd]),
            "orc_accept_1e": (None, [_vp, _vp, C.c_int32, C.c_int]),
            "orc_grad_1e": (None, [_vp, _vp, C.c_int32, _pd, _pd]),
            "orc_derivs": (C.c_int, [_vp, _vp, _pd, _pd]),
            "orc_accum": (None, [C.c_int32, C.c_int32, _pd, _pd, _pd]),
            "orc_walker_get": (None, [_vp, _vp, _pd, _pd]),
            "orc_vmc_run": (C.c_int, [_vp, C.c_int32, _pd, _pi, C.c_int32, C.c_double,
                                      C.c_double, _pd, C.POINTER(C.c_int64), C.c_int32]),
            "orc_eval_batch": (C.c_int, [_vp, C.c_int32, _pd, _pd, _pd, _pd, _pd, _pd,
                                         C.c_int32]),
            "orc_cusp_move": (C.c_double, [_vp, C.c_double, C.c_int32, _pd, _pd, _pi, _pd, _pd]),
            "orc_tmove_run": (C.c_int, [_vp, _vp, _pd, _pi, C.c_double, C.c_int32, _pd]),
            "orc_walker_tq": (C.c_int32, [_vp, _vp, C.c_int32, _pd, _pd]),
            "orc_dmc_create": (_vp, [_vp, C.c_int32]),
            "orc_dmc_destroy": (None, [_vp]),
            "orc_dmc_set": (C.c_int, [_vp, C.c_int32, _pd, _pi, _pd]),
            "orc_dmc_generation": (C.c_int, [_vp, _vp, _vp, C.c_int32]),
            "orc_dmc_branch": (C.c_int, [_vp, _pi, _vp]),
            "orc_dmc_get": (C.c_int32, [_vp, _pd, _pi, _pd, _pd]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_pd)


def _i(a):
    return None if a is None else a.ctypes.data_as(_pi)


class Oracle:
    """One wavefunction (an orc_sys) plus a scratch walker."""

    def __init__(self, system):
        self.system = system
        self._struct = system.struct()
        self.L = lib()
        self.h = self.L.orc_create(C.byref(self._struct))
        if not self.h:
            raise RuntimeError(self.L.orc_error().decode())
        self.N = system.nelec
        self.P = self.L.orc_nparm(self.h)
        self.w = self.L.orc_walker_create(self.h)

    def __del__(self):
        try:
            self.L.orc_walker_destroy(self.w)
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    # ---- setup info ----
    def asymp(self):
        aa = np.zeros(self.system.nctype)
        ab = np.zeros(2)
        self.L.orc_asymp(self.h, _d(aa), _d(ab))
        return aa, ab

    def cusp(self):
        nt = self.L.orc_nterms(self.h)
        d = np.zeros((64, max(nt, 1)))
        iw = np.zeros(64, np.int32)
        n = self.L.orc_cusp(self.h, _d(d), _i(iw))
        return d[:n, :nt].copy(), iw[:n].copy()

    def det_info(self):
        info = np.zeros(5, np.int32)
        self.L.orc_det_info(self.h, _i(info))
        return dict(ndetup=info[0], ndetdn=info[1], noccup=info[2], noccdn=info[3], kmax=info[4])

    def set_emulation(self, mode: int):
        """Error studies only: bit 0 rounds the contraction inputs of the
        full orbital rows to bfloat16, bit 1 those of the quadrature rows."""
        self.L.orc_set_emulation(self.h, int(mode))

    def set_mover(self, imetro: int, deltar: float = 5.0, deltat: float = 1.0):
        """VMC mover of vmc_run: 1 metrop_mov1, 6 metrop_slat_mov1."""
        self.L.orc_set_mover(self.h, int(imetro), float(deltar), float(deltat))

    def set_params(self, a4=None, b=None, c=None, csf_coef=None, coef=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)
        arrs = [conv(a) for a in (a4, b, c, csf_coef, coef)]
        self.L.orc_set_params(self.h, *[_d(a) for a in arrs])

    # ---- pieces ----
    def basis(self, xe):
        nb = self.system.nbasis
        xe = np.ascontiguousarray(xe, np.float64)
        phi, dphi, d2 = np.zeros(nb), np.zeros((nb, 3)), np.zeros(nb)
        self.L.orc_basis(self.h, _d(xe), _d(phi), _d(dphi), _d(d2))
        return phi, dphi, d2

    def orbitals(self, xe):
        no = self.system.norb
        xe = np.ascontiguousarray(xe, np.float64)
        o, do, d2 = np.zeros(no), np.zeros((no, 3)), np.zeros(no)
        self.L.orc_orbitals(self.h, _d(xe), _d(o), _d(do), _d(d2))
        return o, do, d2

    def jastrow(self, x, derivs=False):
        N = self.N
        x = np.ascontiguousarray(x, np.float64)
        val, d2 = np.zeros(1), np.zeros(1)
        v, dv = np.zeros((N, 3)), np.zeros(N)
        npj = len(self.system.jas_a) + len(self.system.jas_b) + len(self.system.jas_c)
        gv = g = d2g = None
        if derivs:
            gv, g, d2g = np.zeros(npj), np.zeros((npj, N, 3)), np.zeros(npj)
        self.L.orc_jastrow_full(self.h, _d(x), _d(val), _d(v), _d(d2), _d(dv),
                                _d(gv), _d(g), _d(d2g))
        out = dict(value=val[0], v=v, d2=d2[0], divvj=dv)
        if derivs:
            out.update(gvalue=gv, g=g, d2g=d2g)
        return out

    def determinant(self, x):
        N = self.N
        x = np.ascontiguousarray(x, np.float64)
        ld, sg = np.zeros(1), np.zeros(1)
        vd, dv = np.zeros((N, 3)), np.zeros(N)
        st = self.L.orc_determinant_full(self.h, _d(x), _d(ld), _d(sg), _d(vd), _d(dv))
        return dict(logdet=ld[0], sign=sg[0], vd=vd, divvd=dv, status=st)

    def eval_full(self, x, walker=None):
        N = self.N
        x = np.ascontiguousarray(x, np.float64)
        out9, grad, divv = np.zeros(9), np.zeros((N, 3)), np.zeros(N)
        st = self.L.orc_eval_full(self.h, walker or self.w, _d(x), _d(out9), _d(grad), _d(divv))
        return dict(logpsid=out9[0], sign=out9[1], psij=out9[2], eloc=out9[3],
                    pe=out9[4], d2lnpsi=out9[5], vpsp=out9[6], pe_ee=out9[7],
                    pe_loc=out9[8], grad=grad, divv=divv, status=st)

    def derivs(self, walker=None):
        dp, de = np.zeros(self.P), np.zeros(self.P)
        st = self.L.orc_derivs(self.h, walker or self.w, _d(dp), _d(de))
        return dp, de

    def ratio_1e(self, ie, xnew, walker=None):
        xnew = np.ascontiguousarray(xnew, np.float64)
        o4, v, lap = np.zeros(4), np.zeros(3), np.zeros(1)
        st = self.L.orc_ratio_1e(self.h, walker or self.w, ie, _d(xnew), _d(o4), _d(v), _d(lap))
        return dict(ratio=o4[0], logpsid=o4[1], sign=o4[2], psij=o4[3], grad=v,
                    lap=lap[0], status=st)

    def accept_1e(self, ie, accept, walker=None):
        self.L.orc_accept_1e(self.h, walker or self.w, ie, int(bool(accept)))

    def grad_1e(self, ie, walker=None):
        g, lap = np.zeros(3), np.zeros(1)
        self.L.orc_grad_1e(self.h, walker or self.w, ie, _d(g), _d(lap))
        return g, lap[0]

    def set_quad(self, xq, yq, zq, walker=None):
        xq, yq, zq = (np.ascontiguousarray(a, np.float64) for a in (xq, yq, zq))
        self.L.orc_walker_set_quad(walker or self.w, _d(xq), _d(yq), _d(zq), len(xq))

    def quadrature(self):
        """The unrotated grid (x, y, z, w) of the nonlocal quadrature."""
        q = [np.zeros(12) for _ in range(4)]
        self.L.orc_quadrature(self.h, *[_d(a) for a in q])
        n = int((q[3] > 0).sum())
        return [a[:n] for a in q]

    def num_radial(self, k, r):
        """(R, R', R'') of tabulated radial function k at the distances r, [n][3]."""
        r = np.ascontiguousarray(r, np.float64)
        out = np.zeros((len(r), 3))
        self.L.orc_num_radial(self.h, k, len(r), _d(r), _d(out))
        return out

    def cusp_move(self, tau, g, xe, rng=None, pt=None, n_drift=2):
        """The all-electron DMC mover at xe with gradient g: (xn or None,
        proposal density at pt or None); rng [4] int32 advances in place."""
        g, xe = (np.ascontiguousarray(a, np.float64) for a in (g, xe))
        xn = np.zeros(3) if rng is not None else None
        ptd = None if pt is None else np.ascontiguousarray(pt, np.float64)
        p = self.L.orc_cusp_move(self.h, tau, n_drift, _d(g), _d(xe), _i(rng), _d(xn), _d(ptd))
        return xn, (p if pt is not None else None)

    def tmove_run(self, x, rng, tau, casula=False, walker=None):
        """DMC-mode evaluation at x, then T-moves; (accepted, x after);
        rng [4] int32 advances in place."""
        x = np.ascontiguousarray(x, np.float64)
        xo = np.zeros_like(x)
        a = self.L.orc_tmove_run(self.h, walker or self.w, _d(x), _i(rng), tau, int(casula), _d(xo))
        return a, xo

    def walker_tq(self, i, walker=None):
        """(points [n][3], ratios [n]) kept for electron i by the last
        DMC-mode evaluation."""
        q, r = np.zeros((self.system.ncent * 12, 3)), np.zeros(self.system.ncent * 12)
        n = self.L.orc_walker_tq(self.h, walker or self.w, i, _d(q), _d(r))
        return q[:n], r[:n]

    # ---- batched ----
    def eval_batch(self, x, derivs=False, nthreads=0):
        x = np.ascontiguousarray(x, np.float64)
        nw, N = x.shape[0], self.N
        out9, grad, divv = np.zeros((nw, 9)), np.zeros((nw, N, 3)), np.zeros((nw, N))
        dp = de = None
        if derivs and self.P:
            dp, de = np.zeros((nw, self.P)), np.zeros((nw, self.P))
        st = self.L.orc_eval_batch(self.h, nw, _d(x), _d(out9), _d(grad), _d(divv),
                                   _d(dp), _d(de), nthreads)
        return dict(out9=out9, grad=grad, divv=divv, dpsi=dp, deloc=de, status=st)

    def vmc_run(self, x, rng, nstep, delta, fbias=1.0, trace=True, nthreads=0):
        x = np.ascontiguousarray(x, np.float64).copy()
        rng = np.ascontiguousarray(rng, np.int32).copy()
        nw = x.shape[0]
        tr = np.zeros((nstep, nw)) if trace else None
        nacc = np.zeros(nw, np.int64)
        st = self.L.orc_vmc_run(self.h, nw, _d(x), _i(rng), nstep, delta, fbias, _d(tr),
                                nacc.ctypes.data_as(C.POINTER(C.c_int64)), nthreads)
        return dict(x=x, rng=rng, eloc=tr, naccept=nacc, status=st)


def solid_harmonic(l, m, x):
    """(S_lm, grad S_lm) at x from the oracle's closed form."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(4)
    lib().orc_solid_harmonic(int(l), int(m), _d(x), _d(out))
    return out[0], out[1:]


def rng_stream(seed4, n):
    L = lib()
    st = np.zeros(4, np.int32)
    s = np.ascontiguousarray(seed4, np.int32)
    L.orc_setrn(_i(st), _i(s))
    return np.array([L.orc_rannyu(_i(st)) for _ in range(n)]), st


def accum(dpsi, deloc, eloc):
    """Sums over samples of the linear method's derivative products (orc_accum):
    dict with n, se, se2, sd, sde, sl [P] and dd, dde, dl [P][P]."""
    dpsi, deloc = np.asarray(dpsi, np.float64), np.asarray(deloc, np.float64)
    W, P = dpsi.shape
    der = np.ascontiguousarray(np.concatenate([dpsi, deloc], axis=1))
    e = np.ascontiguousarray(eloc, np.float64)
    out = np.zeros(3 + 3 * P + 3 * P * P)
    lib().orc_accum(P, W, _d(der), _d(e), _d(out))
    return split_accum(out, P)


def split_accum(out, P):
    """The accumulator layout of qmcw_accum_get / orc_accum as a dict."""
    o = np.asarray(out)
    PP = P * P
    v = o[3:3 + 3 * P].reshape(3, P)
    m = o[3 + 3 * P:].reshape(3, P, P)
    return dict(n=o[0], se=o[1], se2=o[2], sd=v[0], sde=v[1], sl=v[2], dd=m[0], dde=m[1], dl=m[2])


def matinv(a):
    a = np.asfortranarray(np.array(a, np.float64))
    n = a.shape[0]
    flat = np.ascontiguousarray(a.T.reshape(-1))   # column-major storage
    det = lib().orc_matinv(_d(flat), n)
    return flat.reshape(n, n).T.copy(), det


class OracleDMC:
    """A DMC population in the oracle (orc_dmc): per-walker streams, one
    generation of moves + reweighting, split_join with a branching stream."""

    def __init__(self, system, cap):
        self.o = Oracle(system)
        self.N = system.nelec
        self.cap = int(cap)
        self.h = self.o.L.orc_dmc_create(self.o.h, self.cap)

    def set(self, x, rng, weight=None):
        x = np.ascontiguousarray(x, np.float64)
        rng = np.ascontiguousarray(rng, np.int32)
        w = None if weight is None else np.ascontiguousarray(weight, np.float64)
        return self.o.L.orc_dmc_set(self.h, x.shape[0], _d(x), _i(rng), _d(w))

    def generation(self, params, nthreads=0):
        from champ_amd.abi import DmcStats
        st = DmcStats()
        r = self.o.L.orc_dmc_generation(self.h, C.byref(params), C.byref(st), nthreads)
        return r, st

    def branch(self, brng):
        from champ_amd.abi import DmcStats
        st = DmcStats()
        r = self.o.L.orc_dmc_branch(self.h, _i(brng), C.byref(st))
        return r, st

    def get(self):
        x = np.zeros((self.cap, self.N, 3))
        rng = np.zeros((self.cap, 4), np.int32)
        w, e = np.zeros(self.cap), np.zeros(self.cap)
        n = self.o.L.orc_dmc_get(self.h, _d(x), _i(rng), _d(w), _d(e))
        return dict(x=x[:n], rng=rng[:n], weight=w[:n], eloc=e[:n])

    def close(self):
        if self.h:
            self.o.L.orc_dmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
