"""The engine driven from Fortran: tests/fortran/qmcw_driver.f90 (built by
__graft_entry__.build() with flang against include/champ_amd_mod.f90) fills
a qmcw_system member by member, creates a handle, sets walkers, evaluates
them and runs VMC steps with CHAMP's default mover.  Its results equal those
of the same calls made through ctypes bit for bit (same library, same
inputs).  The system is CHAMP's pseudopotential C2 example, whose tabulated
radial basis, pseudopotential and Jastrow parameter lists exercise most
members of the struct."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from champ_amd import abi, rng
from champ_amd.engine import Engine
from champ_amd import system

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from transcripts import c2_bfd  # noqa: E402

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "qmcw_driver")


def dump_system(s, path):
    """The members of struct qmcw_system in declaration order: scalars as
    they are, arrays as an int64 count and the elements (-1: NULL)."""
    st = s.struct()
    arrays = {a.ctypes.data: a for a in st._arrays.values()}
    with open(path, "wb") as f:
        for name, typ in abi.QmcwSystem._fields_:
            v = getattr(st, name)
            if typ is C.c_int32:
                f.write(np.int32(v).tobytes())
            elif typ is C.c_double:
                f.write(np.float64(v).tobytes())
            else:
                addr = C.cast(v, C.c_void_p).value
                if addr is None:
                    f.write(np.int64(-1).tobytes())
                    continue
                a = arrays[addr]
                assert a.dtype == (np.int32 if typ is abi._pi else np.float64), name
                f.write(np.int64(a.size).tobytes())
                f.write(a.tobytes())


def test_fortran_driver_matches_the_ctypes_path(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("the Fortran driver is not built (flang absent when building)")
    s, _ = c2_bfd(3)
    W, nstep, imetro, delta, fbias = 24, 3, 6, 1.0, 1.0
    x = system.random_walkers(s, W, seed=17)
    seeds = np.asarray(rng.walker_seeds("1837465927472523", W), np.int32)
    dump_system(s, tmp_path / "sys.bin")
    with open(tmp_path / "run.bin", "wb") as f:
        f.write(np.array([W, s.nelec, nstep, imetro], np.int32).tobytes())
        f.write(np.array([delta, fbias], np.float64).tobytes())
        f.write(np.ascontiguousarray(x, np.float64).tobytes())
        f.write(seeds.tobytes())
    print(f"[fortran driver] c2_bfd W={W} nstep={nstep}", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path / "sys.bin"), str(tmp_path / "run.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(tmp_path / "out.bin", "rb").read()
    N = s.nelec
    st = np.frombuffer(raw[:64], np.int32)
    assert (st == 0).all(), st
    o = 64
    def take(n, dt=np.float64):
        nonlocal o
        a = np.frombuffer(raw[o:o + n * np.dtype(dt).itemsize], dt)
        o += n * np.dtype(dt).itemsize
        return a
    f_logpsid, f_sgn, f_psij, f_eloc, f_epot = (take(W) for _ in range(5))
    f_grad, f_lap = take(3 * N * W).reshape(W, N, 3), take(N * W).reshape(W, N)
    f_trace = take(W * nstep).reshape(nstep, W)
    f_nacc = take(W, np.int64)
    f_x = take(3 * N * W).reshape(W, N, 3)
    f_seeds = take(4 * W, np.int32).reshape(W, 4)
    assert o == len(raw)
    # the same calls through ctypes
    e = Engine(s, W)
    e.set_walkers(x)
    g = e.psi_at()
    e.set_rng(seeds)
    e.set_vmc_mover(imetro, 5.0, 1.0)
    v = e.vmc(nstep, delta=delta, fbias=fbias)
    for a, b in ((f_logpsid, g["logpsid"]), (f_sgn, g["sign"]), (f_psij, g["psij"]),
                 (f_eloc, g["eloc"]), (f_epot, g["epot"]), (f_grad, g["grad"]), (f_lap, g["divv"]),
                 (f_trace, v["eloc"]), (f_nacc, v["naccept"]), (f_x, e.get_walkers()),
                 (f_seeds, e.get_rng())):
        assert np.array_equal(a, b)
    assert np.isfinite(f_eloc).all() and f_nacc.sum() > 0
    e.close()
