"""The ISO_C_BINDING module trekis-3_amd/fortran/trekis_gpu_c.f90 (what a
Fortran driver uses in place of `use Monte_Carlo`) compiles with the image's
Fortran compiler, its bind(C) types have the C layout, and calls through it
reach the library (host-side entry points only, no GPU needed)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from trekis_gpu import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"

PROG = """
program probe
   use trekis_gpu_c
   implicit none
   type(trk_material) :: m
   type(trk_ion) :: i
   type(trk_options) :: o
   type(trk_tallies) :: t
   real(c_double) :: A(1), E0(1), G(1), ks, fs
   integer(c_int) :: rc
   print '(4I8)', c_sizeof(m), c_sizeof(i), c_sizeof(o), c_sizeof(t)
   print '(ES25.17)', trk_rng_uniform(0_c_int64_t, 0_c_int64_t, 0_c_int32_t, 0_c_int32_t)
   A = 1.0d0; E0 = 15.0d0; G = 0.6d0
   rc = trk_sumrules(A, E0, G, 1_c_int32_t, 0.0d0, 1.0d30, ks, fs)
   print '(I4,2ES25.17)', rc, ks, fs
end program
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler in this image")
def test_fortran_module_layout_and_calls(tmp_path, lib):
    mod = os.path.join(ROOT, "trekis-3_amd", "fortran", "trekis_gpu_c.f90")
    prog = tmp_path / "probe.f90"
    prog.write_text(PROG)
    libdir = os.path.dirname(abi.LIB_PATH)
    subprocess.run([FLANG, "-c", mod, "-J", str(tmp_path), "-o", str(tmp_path / "m.o")], check=True)
    subprocess.run([FLANG, "-I", str(tmp_path), str(prog), str(tmp_path / "m.o"), "-L", libdir, "-ltrekis_gpu",
                    f"-Wl,-rpath,{libdir}", "-o", str(tmp_path / "probe")], check=True)
    out = subprocess.run([str(tmp_path / "probe")], check=True, capture_output=True, text=True).stdout.split("\n")
    sizes = [int(x) for x in out[0].split()]
    assert sizes == [ctypes.sizeof(s) for s in (abi.Material, abi.Ion, abi.Options, abi.Tallies)]
    assert float(out[1]) == lib.trk_rng_uniform(0, 0, 0, 0)
    rc, ks, fs = out[2].split()
    a = ctypes.c_double(); b = ctypes.c_double()
    import numpy as np
    from oracle import dp
    A, E0, G = np.array([1.0]), np.array([15.0]), np.array([0.6])
    lib.trk_sumrules(dp(A), dp(E0), dp(G), 1, 0.0, 1e30, ctypes.byref(a), ctypes.byref(b))
    assert int(rc) == 0 and float(ks) == a.value and float(fs) == b.value


REF_OBJECTS = "/root/reference/Source_files/Objects.f90"


@pytest.mark.skipif(not (os.path.exists(FLANG) and os.path.exists(REF_OBJECTS)),
                    reason="needs the Fortran compiler and the reference's Objects.f90")
def test_bridge_compiles_against_the_reference_types(lib):
    """trekis_gpu_bridge.f90 (do_Monte_Carlo with the reference's argument
    list, Monte_Carlo.f90:39-92) compiles with the reference's own
    Objects.f90 types, and a driver that calls it as MAIN does links."""
    d = os.path.join(ROOT, "trekis-3_amd", "fortran")
    subprocess.run(["make", "-C", d], check=True, capture_output=True)
    assert os.path.exists(os.path.join(d, "_build", "bridge_driver"))
    mod = open(os.path.join(d, "trekis_gpu_bridge.f90")).read()
    ref = open("/root/reference/Source_files/Monte_Carlo.f90").read()
    import re

    def args(text):
        i = text.index("subroutine do_Monte_Carlo(")
        j = text.index(")", i)
        return [a.strip().lower() for a in re.sub(r"&\s*\n\s*", "", text[i + 26:j]).split(",")]
    assert args(mod) == args(ref)
