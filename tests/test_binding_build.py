"""CPU checks of the reference-side binding and of the library's exports.

  * reference_binding/hip_executor.cpp compiles as C++ against the
    reference's own public header (include/dsmvc/engine.hpp) -- the file a
    maintainer adds to the reference tree -- with nothing but that header and
    include/dsmvc_hip.h;
  * the built binding test library loads and exports its entry points;
  * libdsmvc_hip.so exports its C ABI and no C++ symbol, so it cannot
    interpose with the reference engine's identically named C++ symbols when
    both are loaded into one process.
"""
import ctypes
import shutil
import subprocess
from pathlib import Path

import pytest

from support import ROOT

REF_INCLUDE = Path("/root/reference/include")
PKG = ROOT / "vapoursynth-descalemvc_amd"
BINDING = PKG / "reference_binding" / "hip_executor.cpp"
BINDING_SO = ROOT / "oracle" / "_ref" / "libdsmvc_binding.so"
LIB = PKG / "dsmvc_hip" / "libdsmvc_hip.so"


@pytest.mark.skipif(not (REF_INCLUDE / "dsmvc" / "engine.hpp").exists(),
                    reason="reference headers absent")
@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_binding_compiles_against_the_reference_header():
    cmd = ["g++", "-std=c++2b", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
           f"-I{REF_INCLUDE}", f"-I{ROOT / 'include'}", f"-I{BINDING.parent}", str(BINDING)]
    done = subprocess.run(cmd, capture_output=True, text=True)
    assert done.returncode == 0, done.stderr


@pytest.mark.skipif(not BINDING_SO.exists(), reason="binding library not built")
def test_binding_library_loads_and_exports():
    lib = ctypes.CDLL(str(BINDING_SO))
    for name in ("bnd_create", "bnd_free", "bnd_prepare", "bnd_seal", "bnd_inverse_2d",
                 "bnd_inverse_rows", "bnd_inverse_columns", "bnd_available", "bnd_last_error"):
        assert hasattr(lib, name), name


@pytest.mark.skipif(shutil.which("nm") is None, reason="nm absent")
def test_library_exports_only_the_c_abi():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True,
                         text=True, check=True).stdout
    names = [line.split()[-1] for line in out.splitlines() if line.strip()]
    assert names, "no exported symbols"
    stray = [n for n in names if not n.startswith("dsmvc_hip_")]
    assert not stray, stray[:10]
