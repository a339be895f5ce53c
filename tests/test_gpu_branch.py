"""The device split_join (k_split_join, the branching qmcw_dmc_step runs)
against the host routine qmcw_split_join -- itself bit-exact against the
reference's split_join (tests/test_branch_ref.py) -- on the same 70 weight
vectors: identical weights, copies, streams, populations, branching counts
and branching stream.  The DMC generations of tests/test_gpu_dmc.py run the
device path against the oracle; one of them is repeated on the host path."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from champ_amd import abi

sys.path.insert(0, os.path.dirname(__file__))
from test_branch_ref import _cases  # noqa: E402

pytestmark = pytest.mark.gpu
_pd, _pi = C.POINTER(C.c_double), C.POINTER(C.c_int32)


def _run(fn, w, cap, brng, rng):
    wt = np.zeros(cap)
    wt[:len(w)] = w
    origin = np.zeros(cap, np.int32)
    r = np.array(rng, np.int32).copy()
    b = np.array(brng, np.int32).copy()
    n, nb = C.c_int32(), C.c_int32()
    st = fn(len(w), cap, wt.ctypes.data_as(_pd), origin.ctypes.data_as(_pi), r.ctypes.data_as(_pi),
            b.ctypes.data_as(_pi), C.byref(n), C.byref(nb))
    k = n.value
    return st, (wt[:k], origin[:k], r[:k], b, k, nb.value)


def test_device_split_join_matches_host_routine():
    lib = abi.load_library()
    nsplit = 0
    for k, (w, cap) in enumerate(_cases()):
        b = [(17 * k + 5) % 4096, (311 * k + 9) % 4096, (1001 * k + 3) % 4096, (2 * k + 1) % 4096]
        rng = np.arange(4 * cap, dtype=np.int32).reshape(cap, 4) % 4096
        sh, h = _run(lib.qmcw_split_join, w, cap, b, rng)
        sg, g = _run(lib.qmcw_split_join_gpu, w, cap, b, rng)
        assert sh == 0 and sg == 0, (w, cap)
        for a, c in zip(h, g):
            assert np.array_equal(np.asarray(a), np.asarray(c)), (w, cap)
        nsplit += int(h[4] > len(w) - sum(1 for v in w if v == 0))
    assert nsplit > 5


def test_dmc_generations_on_the_host_split_join(monkeypatch):
    from test_gpu_dmc import test_dmc_generations_match_oracle
    monkeypatch.setenv("QMCW_HOST_SPLIT_JOIN", "1")
    test_dmc_generations_match_oracle("c2_dimer", 0.3, 1)
