"""The C-ABI library loads, exports every function include/trekis_gpu.h
declares, and the ctypes mirror agrees with the C layout of every struct.
No compute call is made here (no GPU needed)."""
import ctypes
import os
import subprocess
import tempfile

import pytest

from trekis_gpu import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {"trk_material": abi.Material, "trk_ion": abi.Ion, "trk_options": abi.Options,
           "trk_tallies": abi.Tallies}


def test_library_exports_header_symbols(lib):
    names = abi.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_c():
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "trekis_gpu.h"', 'int main(void){']
    for cname, py in STRUCTS.items():
        src.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, f, v = line.split()
        py = STRUCTS[cname]
        got = ctypes.sizeof(py) if f == "size" else getattr(py, f).offset
        assert got == int(v), (cname, f, got, v)


def test_product_has_no_oracle_dependency():
    """The shipped library must not link or name the oracle."""
    out = subprocess.run(["nm", "-D", abi.LIB_PATH], check=True, capture_output=True, text=True).stdout
    assert "or_" not in " ".join(l.split()[-1][:3] for l in out.splitlines())
    ldd = subprocess.run(["ldd", abi.LIB_PATH], capture_output=True, text=True).stdout
    assert "trk_oracle" not in ldd


def test_oracle_and_product_constants_bitwise(lib, oracle_lib):
    for w in range(8):
        a = lib.trk_constant(w)
        b = oracle_lib.or_constant(w)
        assert a == b and a != 0.0, (w, a, b)


def test_constants_defining_expressions(lib):
    # Universal_Constants.f90:28-60, evaluated in double as written there
    e, h, c, e0 = 1.602176487e-19, 1.05457162853e-34, 299792458.0, 8.854187817620e-12
    pi = 3.1415926535897932384626433832795
    assert lib.trk_constant(0) == pi
    assert lib.trk_constant(1) == e
    assert lib.trk_constant(2) == 9.1093821545e-31
    assert lib.trk_constant(3) == 1836.1526724780 * 9.1093821545e-31
    assert lib.trk_constant(4) == h
    assert lib.trk_constant(5) == e * e / (h * c * 4.0 * pi * e0)
    assert abs(1.0 / lib.trk_constant(5) - 137.036) < 1e-3
