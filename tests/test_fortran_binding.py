"""The Fortran binding (include/champ_amd_mod.f90) compiles with flang, links
against libchamp_amd.so, and its qmcw_system is layout-compatible with the C
struct (same sizeof).  Calls only entry points that need no GPU."""
import os
import shutil
import subprocess

import pytest

from champ_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"

PROGRAM = """
program t
  use champ_amd_mod
  implicit none
  type(qmcw_system) :: s
  type(qmcw_dmc_params) :: p
  type(qmcw_dmc_stats) :: d
  print '(5i8)', c_sizeof(s), qmcw_sizeof_system(), qmcw_nterms4(5_c_int32_t), c_sizeof(p), c_sizeof(d)
end program
"""


def test_fortran_module_matches_the_c_abi(tmp_path):
    if not os.path.exists(FLANG) or shutil.which("ld") is None:
        pytest.skip("flang not available")
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("libchamp_amd.so not built")
    libdir = os.path.dirname(abi.LIB_PATH)
    (tmp_path / "t.f90").write_text(PROGRAM)
    subprocess.run([FLANG, "-c", os.path.join(ROOT, "include", "champ_amd_mod.f90"), "-o", "mod.o"],
                   cwd=tmp_path, check=True, capture_output=True)
    subprocess.run([FLANG, "t.f90", "mod.o", f"-L{libdir}", "-lchamp_amd", f"-Wl,-rpath,{libdir}",
                    "-o", "t"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./t"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout
    f_size, c_size, nterms, p_size, d_size = map(int, out.split())
    assert f_size == c_size
    assert nterms == 23
    # the DMC structs have the sizes of their C layout (mirrored by abi.py)
    import ctypes
    assert p_size == ctypes.sizeof(abi.DmcParams) == 10 * 8 + 8 * 4
    assert d_size == ctypes.sizeof(abi.DmcStats) == 8 * 8 + 2 * 4


def test_fortran_module_binds_every_entry_point():
    """Every function include/champ_amd.h declares has a bind(C) interface
    of the same name in include/champ_amd_mod.f90."""
    import re
    hdr = open(os.path.join(ROOT, "include", "champ_amd.h")).read()
    mod = open(os.path.join(ROOT, "include", "champ_amd_mod.f90")).read()
    c_names = set(re.findall(r"^(?:int|int32_t|int64_t|double|const char \*|void)\s*\*?\s*(qmcw_\w+)\(",
                             hdr, re.M))
    f_names = set(re.findall(r'bind\(C, name="(qmcw_\w+)"\)', mod))
    assert len(c_names) > 40
    assert c_names - f_names == set()
