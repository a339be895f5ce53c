"""CHAMP's walker configurations (champ_amd/configs.py).

  sites         the starting configuration from the input's seed equals the
                "sites:" line CHAMP printed, to its four significant digits,
                for the Be atom (one centre) and the pseudopotential C2
                (two centres): electron distribution, draw order, scales
  mc_configs    configuration files written in vmc.f90's format read back
                to the width of the format; mc_configs_start round trip
"""
import json
import math
import os
import sys

import numpy as np

from champ_amd import configs, rng, system

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from transcripts import c2_bfd  # noqa: E402


def _within_printed_digits(x, ref):
    """|x - ref| within half a unit of the fourth significant digit of ref
    (CHAMP prints 0.dddd D+ee)."""
    ref = np.asarray(ref)
    e = np.floor(np.log10(np.abs(ref))) + 1
    return np.all(np.abs(np.asarray(x) - ref) <= 0.505e-4 * 10.0 ** e)


def test_sites_reproduces_champ_starting_configurations():
    seed = rng.seed_limbs("1837465927472523")
    fx = json.load(open(os.path.join(HERE, "golden", "be_transcript.json")))
    be = system.from_champ_input(fx["input"], "be")
    x, _ = configs.sites(be, seed)
    assert x.shape == (4, 3)
    assert _within_printed_digits(x.reshape(-1), fx["sites"])
    s, _ = c2_bfd(1)
    fx = json.load(open(os.path.join(HERE, "golden", "c2_bfd_transcript.json")))
    x, after = configs.sites(s, seed)
    assert configs.electron_sites(s) == ([4, 4], [2, 2])
    assert _within_printed_digits(x.reshape(-1), fx["sites"])
    # six draws per electron
    assert after == rng.jump(seed, 6 * s.nelec)


def test_electron_sites_distribution():
    s = system.hchain(4)
    nsite, nup = configs.electron_sites(s)
    assert sum(nsite) == s.nelec and sum(nup) == s.nup
    assert all(0 <= u <= n for u, n in zip(nup, nsite))
    x, _ = configs.sites(s, rng.seed_limbs("1234567890123457"))
    assert x.shape == (s.nelec, 3) and np.isfinite(x).all()


def test_mc_configs_round_trip():
    rs = np.random.default_rng(3)
    N, nc = 5, 7
    x = rs.normal(0, 2, (nc, N, 3))
    sign = rs.choice([-1, 1], nc)
    logpsi, eloc = rs.normal(-3, 1, nc), rs.normal(-14, 0.5, nc)
    psid, psij = rs.normal(0, 1e-2, nc), rs.normal(-1, 0.1, nc)
    text = configs.write_mc_configs(x, sign, logpsi, eloc, psid, psij)
    lines = text.splitlines()
    assert len(lines) == nc and all(len(l) == 13 * 3 * N + 3 + 12 + 12 + 24 for l in lines)
    xr, cols = configs.read_mc_configs(text, N)
    assert np.allclose(xr, x, rtol=0, atol=5e-9)
    assert np.array_equal(cols["sign"], sign)
    assert np.allclose(cols["logpsi"], logpsi, rtol=1e-4)
    assert np.allclose(cols["eloc"], eloc, rtol=0, atol=5e-6)
    start = configs.write_mc_configs_start(x[0])
    assert np.array_equal(configs.read_mc_configs_start(start, N), x[0])
    # Fortran's D exponents read too
    assert np.allclose(configs.read_mc_configs_start("1.5D0 -2.D-1 3\n4 5 6", 2),
                       [[1.5, -0.2, 3], [4, 5, 6]])


def test_mc_configs_free_format_lines():
    """Configuration lines whose numbers do not sit in 13-character fields
    (another writer's free format) are read by their blanks."""
    N = 2
    line = " ".join(str(v) for v in [0.5, -1.25, 2.0, 3.5, 1e-3, -7.0]) + " 1 -2.5 -14.2 0.01 -0.9"
    line = line + " " * max(0, 13 * 3 * N - len(line))
    x, cols = configs.read_mc_configs(line + "\n", N)
    assert np.allclose(x[0], [[0.5, -1.25, 2.0], [3.5, 1e-3, -7.0]])
    assert cols["sign"][0] == 1 and cols["eloc"][0] == -14.2
