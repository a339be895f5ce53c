"""The C-ABI library loads on any host and exports every entry point that
include/champ_amd.h declares; the ctypes mirror of qmcw_system matches the
compiled struct.  No compute calls (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from champ_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "champ_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(qmcw_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("libchamp_amd.so not built")
    lib = C.CDLL(abi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_declared()) == set(abi.SIGNATURES)


def test_struct_layout_matches():
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("libchamp_amd.so not built")
    lib = abi.load_library()
    assert lib.qmcw_sizeof_system() == C.sizeof(abi.QmcwSystem)
    assert lib.qmcw_nterms4(5) == 23 and lib.qmcw_nterms4(4) == 13


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        abi.load_library(str(tmp_path / "nope.so"))
