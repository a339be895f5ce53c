"""ctypes mirror of ``include/champ_amd.h`` and the loader of libchamp_amd.so.

The layout of :class:`QmcwSystem` must match ``struct qmcw_system`` field by
field; ``tests/test_abi.py`` checks the exported symbols and the struct size
against the compiled library.
"""
from __future__ import annotations

import ctypes as C
import os

_here = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_here, "libchamp_amd.so")

QMCW_OK = 0
QMCW_ERR_ARG = 1
QMCW_ERR_HIP = 2
QMCW_ERR_SINGULAR = 3
QMCW_ERR_NAN = 4
QMCW_ERR_NOMEM = 5

QMCW_FP64 = 0
QMCW_MIXED = 1
QMCW_MIXED_BF16 = 2

# kernels timed by qmcw_profile, in the order of the QMCW_PROF_* indices
PROF_KERNELS = ("k_full_jas", "k_full_det", "k_full_nlpp", "k_full_fin",
                "k_ratio_1e", "k_accept_1e", "k_rotqua",
                "k_deriv_jas", "k_deriv_nlpp", "k_derivs", "k_accum")

QMCW_BASIS_SLATER = 1
QMCW_BASIS_GAUSSIAN = 2
QMCW_MAX_L = 5
QMCW_MAX_ORD = 8
QMCW_MAX_LPOT = 5
QMCW_MAX_NQUAD = 12

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int32)


class QmcwSystem(C.Structure):
    _fields_ = [
        ("nelec", C.c_int32), ("nup", C.c_int32),
        ("ncent", C.c_int32), ("nctype", C.c_int32),
        ("cent", _pd), ("iwctype", _pi), ("znuc", _pd),
        ("basis_type", C.c_int32), ("nrad", C.c_int32),
        ("rad_n", _pi), ("rad_prim_off", _pi),
        ("prim_zeta", _pd), ("prim_coef", _pd),
        ("ctype_bas_off", _pi), ("bas_rad", _pi), ("bas_l", _pi), ("bas_m", _pi),
        ("nbasis", C.c_int32),
        ("norb", C.c_int32), ("coef", _pd),
        ("ndet", C.c_int32), ("det_up", _pi), ("det_dn", _pi),
        ("ncsf", C.c_int32), ("csf_coef", _pd), ("csf_det_off", _pi),
        ("csf_det", _pi), ("csf_det_coef", _pd),
        ("isc", C.c_int32), ("nspin2b", C.c_int32), ("nocuspb", C.c_int32),
        ("norda", C.c_int32), ("nordb", C.c_int32), ("nordc", C.c_int32),
        ("scalek", C.c_double),
        ("a4", _pd), ("b", _pd), ("c", _pd),
        ("nloc", C.c_int32), ("npotd", _pi), ("lpotp1", _pi), ("ecp_off", _pi),
        ("ecp_coef", _pd), ("ecp_power", _pi), ("ecp_expo", _pd),
        ("nquad", C.c_int32),
        ("nparma_tot", C.c_int32), ("jas_a_type", _pi), ("jas_a_idx", _pi),
        ("nparmb_tot", C.c_int32), ("jas_b_spin", _pi), ("jas_b_idx", _pi),
        ("nparmc_tot", C.c_int32), ("jas_c_type", _pi), ("jas_c_idx", _pi),
        ("nparmcsf", C.c_int32), ("csf_parm", _pi),
        ("nparmorb", C.c_int32), ("orb_parm_i", _pi), ("orb_parm_j", _pi),
        ("orb_parm_rev", _pi),
        ("hb", C.c_double),
        ("rad_num", _pi), ("nnum", C.c_int32), ("num_igrid", _pi), ("num_nr", _pi),
        ("num_h", _pd), ("num_r0", _pd), ("num_off", _pi), ("num_rwf", _pd), ("num_d2", _pd),
        ("num_ce", _pd), ("num_ae", _pd),
    ]


class DmcParams(C.Structure):
    """struct qmcw_dmc_params; defaults are CHAMP's."""
    _fields_ = [("tau", C.c_double), ("etrial", C.c_double), ("eest", C.c_double),
                ("esigma", C.c_double), ("pop_ctrl", C.c_double), ("taueff", C.c_double),
                ("c_rewt", C.c_double), ("pow_rewt", C.c_double), ("limit_rewt", C.c_double),
                ("adrift", C.c_double), ("n_drift", C.c_int32), ("itau_eff", C.c_int32),
                ("itau_integ", C.c_int32), ("iacc_rej", C.c_int32), ("icross", C.c_int32),
                ("ene_int", C.c_int32), ("idmc", C.c_int32), ("tmoves", C.c_int32)]

    def __init__(self, tau, etrial, eest=None, esigma=None, pop_ctrl=1.0, taueff=None, **kw):
        super().__init__()
        self.tau, self.etrial = tau, etrial
        self.eest = etrial if eest is None else eest
        self.esigma = 1.0 if esigma is None else esigma
        self.pop_ctrl = pop_ctrl
        self.taueff = tau if taueff is None else taueff
        self.c_rewt, self.pow_rewt, self.limit_rewt, self.adrift = 3.5, 0.5, 10.0, 0.5
        self.n_drift, self.itau_eff, self.itau_integ = 2, 1, 1
        self.iacc_rej, self.icross, self.ene_int, self.idmc = 1, 1, 0, 2
        self.tmoves = 0
        for k, v in kw.items():
            setattr(self, k, v)


class DmcStats(C.Structure):
    """struct qmcw_dmc_stats"""
    _fields_ = [("wsum", C.c_double), ("wesum", C.c_double), ("we2sum", C.c_double),
                ("wtausum", C.c_double), ("pacc", C.c_double), ("nacc", C.c_double),
                ("ntry", C.c_double), ("ncross", C.c_double), ("nwalk", C.c_int32),
                ("nbranch", C.c_int32)]


def dptr(a):
    """Pointer to a contiguous float64 numpy array (or NULL)."""
    if a is None:
        return _pd()
    assert a.dtype.name == "float64" and a.flags["C_CONTIGUOUS"], a.dtype
    return a.ctypes.data_as(_pd)


def iptr(a):
    """Pointer to a contiguous int32 numpy array (or NULL)."""
    if a is None:
        return _pi()
    assert a.dtype.name == "int32" and a.flags["C_CONTIGUOUS"], a.dtype
    return a.ctypes.data_as(_pi)


# exported C-ABI entry points and their ctypes signatures
_vp = C.c_void_p
SIGNATURES = {
    "qmcw_nterms4": (C.c_int32, [C.c_int32]),
    "qmcw_sizeof_system": (C.c_int64, []),
    "qmcw_create": (C.c_int, [C.POINTER(QmcwSystem), C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "qmcw_destroy": (C.c_int, [_vp]),
    "qmcw_set_params": (C.c_int, [_vp, _pd, _pd, _pd, _pd, _pd]),
    "qmcw_set_walkers": (C.c_int, [_vp, _pd]),
    "qmcw_get_walkers": (C.c_int, [_vp, _pd]),
    "qmcw_eval_full": (C.c_int, [_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
    "qmcw_ratio_1e": (C.c_int, [_vp, C.c_int32, _pd, _pd, _pd, _pd]),
    "qmcw_accept_1e": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_uint8)]),
    "qmcw_grad_1e": (C.c_int, [_vp, C.c_int32, _pd, _pd]),
    "qmcw_nparm": (C.c_int32, [_vp]),
    "qmcw_derivs": (C.c_int, [_vp, _pd, _pd]),
    "qmcw_sample_derivs": (C.c_int, [_vp, C.c_int32]),
    "qmcw_get_derivs": (C.c_int, [_vp, _pd, _pd]),
    "qmcw_accum_reset": (C.c_int, [_vp]),
    "qmcw_accum_add": (C.c_int, [_vp]),
    "qmcw_accum_get": (C.c_int, [_vp, _pd]),
    "qmcw_accum_add_static": (C.c_int, [_vp, C.c_void_p]),
    "qmcw_split_join": (C.c_int, [C.c_int32, C.c_int32, _pd, _pi, _pi, _pi, _pi, _pi]),
    "qmcw_split_join_gpu": (C.c_int, [C.c_int32, C.c_int32, _pd, _pi, _pi, _pi, _pi, _pi]),
    "qmcw_set_rng": (C.c_int, [_vp, _pi]),
    "qmcw_get_rng": (C.c_int, [_vp, _pi]),
    "qmcw_vmc_run": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double, _pd,
                               C.POINTER(C.c_int64)]),
    "qmcw_vmc_run_device": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double]),
    "qmcw_batch_create": (C.c_int, [C.POINTER(_vp), C.c_int32, C.POINTER(_vp)]),
    "qmcw_batch_vmc_run": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double]),
    "qmcw_batch_destroy": (C.c_int, [_vp]),
    "qmcw_batch_last_error": (C.c_char_p, [_vp]),
    "qmcw_set_vmc_mover": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double]),
    "qmcw_set_reinversion": (C.c_int, [_vp, C.c_int32]),
    "qmcw_reinvert": (C.c_int, [_vp]),
    "qmcw_get_eloc": (C.c_int, [_vp, _pd]),
    "qmcw_sync": (C.c_int, [_vp]),
    "qmcw_autotune": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "qmcw_set_group_width": (C.c_int, [_vp, C.c_int32]),
    "qmcw_get_group_widths": (C.c_int, [_vp, _pi]),
    "qmcw_profile": (C.c_int, [_vp, C.c_int32]),
    "qmcw_profile_get": (C.c_int, [_vp, _pd, C.POINTER(C.c_int64)]),
    "qmcw_last_error": (C.c_char_p, [_vp]),
    "qmcw_dmc_begin": (C.c_int, [_vp, C.c_int32, _pd]),
    "qmcw_dmc_step": (C.c_int, [_vp, C.c_void_p, C.c_void_p]),
    "qmcw_dmc_get": (C.c_int, [_vp, C.POINTER(C.c_int32), _pd]),
    "qmcw_set_branch_rng": (C.c_int, [_vp, _pi]),
    "qmcw_get_branch_rng": (C.c_int, [_vp, _pi]),
}

_lib = None


def load_library(path: str | None = None):
    """Load libchamp_amd.so (built by __graft_entry__.build()).

    Raises ``RuntimeError`` when the library is missing: there is no CPU
    fallback for the product path.
    """
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
