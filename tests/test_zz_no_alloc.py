"""qmcw_accum_add_static (the product sums with no device allocation)
against qmcw_accum_add on the same handles: bit-identical accumulators on
two calls in a row, the second on a handle with fewer walkers; sizes beyond
the static array are rejected.  Device buffers for the caller-owned
accumulator come from the HIP runtime through ctypes."""
import ctypes as C

import numpy as np
import pytest

from champ_amd import system
from champ_amd.engine import Engine, QmcwError

pytestmark = pytest.mark.gpu


def _ref_sums(e):
    """The accumulator qmcw_accum_add leaves after one add from zero."""
    o = np.zeros(3 + 3 * e.P + 3 * e.P * e.P)
    e.accum_reset()
    e.accum_add()
    e._chk(e.lib.qmcw_accum_get(e.h, o.ctypes.data_as(C.POINTER(C.c_double))))
    return o


def _hip():
    """The HIP runtime the library itself uses: the libamdhip64 file mapped
    from ROCm's own directory (torch loads a second copy under the bare
    name, which cannot initialise the GPU in the same process)."""
    from champ_amd import abi
    abi.load_library()
    path = None
    for ln in open("/proc/self/maps"):
        f = ln.split()[-1]
        if "libamdhip64.so" in f and "torch" not in f:
            path = f
            break
    hip = C.CDLL(path or "/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _static_sums(e):
    """The caller's device accumulator (cleared by the caller) after one
    qmcw_accum_add_static."""
    hip = _hip()
    n = 3 + 3 * e.P + 3 * e.P * e.P
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), n * 8) == 0
    try:
        assert hip.hipMemset(p, 0, n * 8) == 0
        e.accum_add_static(p.value)
        e.sync()
        out = np.zeros(n)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, n * 8, 2) == 0   # device to host
    finally:
        hip.hipFree(p)
    return out


def test_static_scratch_sums_match_allocating_path():
    s = system.c2_dimer(seed=2, derivs=True)
    for W, seed in ((600, 11), (300, 12)):     # the second call has fewer walkers
        e = Engine(s, W)
        e.set_walkers(system.random_walkers(s, W, seed=seed))
        e.derivs()
        got = _static_sums(e)
        ref = _ref_sums(e)
        assert got[0] == W
        assert np.array_equal(got, ref)
        e.close()


def test_static_scratch_rejects_too_many_walkers():
    s = system.c2_dimer(seed=2, derivs=True)
    e = Engine(s, 6000)                        # 24 chunks of 256 walkers at P = 242
    assert e.P == 242
    with pytest.raises(QmcwError, match="static array"):
        e.accum_add_static(1 << 20)            # rejected before any launch
    e.close()
