"""split_join against the reference's own split_join
(qmc/dmc/dmc_elec/split_join_mod.f90:3-219), compiled from its module
closure into oracle/_ref/libref_branch.so (oracle/ref_closure.py).

On fixed weight vectors -- zero weights, pairs to join, splits at 2x, 4x
and 8x the split threshold, populations near the capacity (where the
threshold rises) and seeded random populations -- the engine's host
split_join (qmcw_split_join, the branching of qmcw_dmc_step) and the
oracle's orc_dmc_branch must give the reference's weights bit for bit, the
same walker copied into each slot, the same population and branching count,
and the same join decisions (one rannyu per joined pair, in walker order).
The one documented difference is checked exactly: each split copy gets four
12-bit limbs drawn from the branching stream after the reference's last
draw, so the final stream is the reference's advanced by 4 draws per split.

The reference library is called in a child process, because the reference
stops the program when a population exceeds MWALK."""
import ctypes as C
import multiprocessing as mp
import os

import numpy as np
import pytest

import oracle
from champ_amd import abi, system

REF_SO = os.path.join(os.path.dirname(oracle.__file__), "_ref", "libref_branch.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF_SO),
                                reason="reference split_join not built (needs /root/reference)")

_pd, _pi = C.POINTER(C.c_double), C.POINTER(C.c_int32)


def _ref_run(cases):
    """Child process: the reference split_join on each (weights, cap, brng):
    weights, copied-walker tags, population, nbrnch, the stream after it, and
    4*nsplit further draws of that stream with the stream after those."""
    L = C.CDLL(REF_SO, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    L.ref_split_join.argtypes = [C.c_int32, C.c_int32, _pd, _pi, _pi, _pi]
    L.ref_branch_setrn.argtypes = [_pi]
    L.ref_branch_savern.argtypes = [_pi]
    L.ref_branch_rannyu.restype = C.c_double
    out = []
    for w, cap, brng, nsplit in cases:
        wt = np.zeros(cap)
        wt[:len(w)] = w
        tag = np.zeros(cap, np.int32)
        nout, nbr = C.c_int32(), C.c_int32()
        b = np.array(brng, np.int32)
        L.ref_branch_setrn(b.ctypes.data_as(_pi))
        L.ref_split_join(len(w), cap, wt.ctypes.data_as(_pd), tag.ctypes.data_as(_pi),
                         C.byref(nout), C.byref(nbr))
        after = np.zeros(4, np.int32)
        L.ref_branch_savern(after.ctypes.data_as(_pi))
        draws = [L.ref_branch_rannyu() for _ in range(4 * nsplit)]
        final = np.zeros(4, np.int32)
        L.ref_branch_savern(final.ctypes.data_as(_pi))
        out.append(dict(w=wt[:nout.value].copy(), tag=tag[:nout.value].copy(), n=nout.value,
                        nbr=nbr.value, after=after, draws=draws, final=final))
    return out


def _engine(w, cap, brng, rng):
    lib = abi.load_library()
    wt = np.zeros(cap)
    wt[:len(w)] = w
    origin = np.zeros(cap, np.int32)
    r = np.array(rng, np.int32).copy()
    b = np.array(brng, np.int32).copy()
    n, nb = C.c_int32(), C.c_int32()
    st = lib.qmcw_split_join(len(w), cap, wt.ctypes.data_as(_pd), origin.ctypes.data_as(_pi),
                             r.ctypes.data_as(_pi), b.ctypes.data_as(_pi), C.byref(n), C.byref(nb))
    return st, dict(w=wt[:n.value], origin=origin[:n.value], rng=r[:n.value], brng=b, n=n.value,
                    nbr=nb.value)


def _cases():
    rs = np.random.default_rng(7)
    c = [
        ([1.0, 0.0, 0.7, 0.0, 1.2], 8),                                  # kills only
        ([0.3, 1.0, 0.2, 0.45, 1.1, 0.1, 0.49, 0.05, 1.0], 12),           # joins (one left unpaired)
        ([1.0, 2.5, 0.4, 0.3, 1.0], 10),                                  # split at 2x, join
        ([4.5, 0.0, 9.0, 17.0, 1.0, 0.2, 0.25], 24),                      # splits at 2x, 4x, 8x threshold
        ([2.0, 2.0000001, 0.5, 0.4999999, 0.0], 8),                       # at the thresholds
        ([3.0] * 6 + [0.2] * 4, 12),                                      # near capacity: threshold rises
        ([1.5, 6.0, 2.2, 0.0, 0.3, 0.35, 3.3, 12.0], 12),
        ([0.7], 4),
        ([], 4),
        ([0.0, 0.0, 0.0, 2.5], 6),
    ]
    for k in range(60):                                                   # seeded random populations
        n = int(rs.integers(5, 60))
        w = rs.lognormal(0.0, rs.uniform(0.3, 1.2), n)
        w[rs.random(n) < 0.1] = 0.0
        c.append((list(w), int(n * rs.uniform(1.3, 3.0)) + 4))
    return c


def _nsplit(w, nbr):
    w = np.asarray(w)
    nlow = int(((w > 0) & (w < 0.5)).sum())
    return nbr - int((w == 0).sum()) - nlow // 2


def test_split_join_matches_reference_build():
    cases = _cases()
    brng = [[(17 * k + 5) % 4096, (311 * k + 9) % 4096, (1001 * k + 3) % 4096, (2 * k + 1) % 4096]
            for k in range(len(cases))]
    eng = []
    for (w, cap), b in zip(cases, brng):
        rng = np.arange(4 * cap, dtype=np.int32).reshape(cap, 4) % 4096
        st, r = _engine(w, cap, b, rng)
        assert st == 0, (w, cap)
        eng.append((r, rng))
    ctx = mp.get_context("spawn")
    with ctx.Pool(1) as pool:
        ref = pool.apply(_ref_run, ([(w, cap, b, _nsplit(w, r["nbr"]))
                                    for (w, cap), b, (r, _) in zip(cases, brng, eng)],))
    nsplit_total = 0
    for (w, cap), (r, rng0), f in zip(cases, eng, ref):
        assert r["n"] == f["n"] and r["nbr"] == f["nbr"], (w, cap)
        assert np.array_equal(r["w"], f["w"]), (w, cap)                  # bit for bit
        assert np.array_equal(r["origin"], f["tag"]), (w, cap)
        ns = _nsplit(w, r["nbr"])
        nsplit_total += ns
        # the branching stream: the reference's, then 4 draws per split copy
        assert np.array_equal(r["brng"], f["final"]), (w, cap)
        fresh = sorted(tuple(int(d * 4096.0) & 4095 for d in f["draws"][4 * k:4 * k + 4])
                       for k in range(ns))
        fresh = sorted(t[:3] + (t[3] | 1,) for t in fresh)
        got = []
        for d in range(r["n"]):
            if not np.array_equal(r["rng"][d], rng0[r["origin"][d]]):
                got.append(tuple(int(v) for v in r["rng"][d]))
        assert sorted(got) == fresh, (w, cap)
    assert nsplit_total > 20


def test_oracle_branch_matches_engine_split_join():
    """orc_dmc_branch on the same weight vectors gives the engine's populations, copies, streams and
    branching stream; with the test above, the oracle's branching is the
    reference's."""
    s = system.be_atom(1)
    for k, (w, cap) in enumerate(_cases()):
        if not w:
            continue
        b = [(17 * k + 5) % 4096, (311 * k + 9) % 4096, (1001 * k + 3) % 4096, (2 * k + 1) % 4096]
        rng = np.arange(4 * cap, dtype=np.int32).reshape(cap, 4) % 4096
        st, r = _engine(w, cap, b, rng)
        assert st == 0
        d = oracle.OracleDMC(s, cap)
        x = system.random_walkers(s, len(w), seed=k)        # distinct rows identify the walkers
        assert d.set(x, rng[:len(w)], np.asarray(w, np.float64)) == 0
        bo = np.array(b, np.int32)
        rr, sb = d.branch(bo)
        assert rr == 0
        g = d.get()
        assert sb.nwalk == r["n"] and sb.nbranch == r["nbr"]
        assert np.array_equal(g["weight"], r["w"])
        assert np.array_equal(g["x"], x[r["origin"]])
        assert np.array_equal(g["rng"], r["rng"])
        assert np.array_equal(bo, r["brng"])
        d.close()
