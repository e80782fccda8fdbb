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


def _decay_material(lib):
    import trekis_gpu as tg
    mat, ion, opt, run = tg.config_material("test")
    mat.prepare(lib)
    m = abi.Material.from_buffer_copy(mat.to_struct())
    return m, ion, opt


def test_decay_time_rules_of_check_atomic_parameters(lib):
    """Dealing_with_EADL.f90:350-351, 366-370: no radiative decay of the valence
    band or with photons off, no Auger decay of the valence band; an Auger time
    the reference would read from EADL is refused."""
    m, ion, opt = _decay_material(lib)
    nsh = sum(m.n_shells[:m.n_atoms])
    vb = m.n_shells[0] - 1
    m.Radiat[0] = 5.0                # a deep shell with a finite radiative time
    m.Auger[vb] = 3.0
    m.Radiat[vb] = 3.0
    why = ctypes.create_string_buffer(256)
    # photons off: every radiative time becomes 1e23 fs
    a = abi.Material.from_buffer_copy(m)
    assert lib.trk_check_atomic_parameters(ctypes.byref(a), 0, why, 256) == 0
    assert all(a.Radiat[s] == 1.0e23 for s in range(nsh))
    assert a.Auger[vb] == 1.0e23 and a.Auger[0] == m.Auger[0]
    # photons on: the deep shell keeps its time, the valence band does not decay
    b = abi.Material.from_buffer_copy(m)
    assert lib.trk_check_atomic_parameters(ctypes.byref(b), 1, why, 256) == 0
    assert b.Radiat[0] == 5.0 and b.Radiat[vb] == 1.0e23 and b.Auger[vb] == 1.0e23
    # 'read it from EADL' Auger marks and non-positive radiative times are refused
    for field, v, photons in (("Auger", -3.6, 1), ("Auger", 2.0e30, 0), ("Auger", float("nan"), 0),
                              ("Radiat", 0.0, 1), ("Radiat", float("inf"), 1)):
        c = abi.Material.from_buffer_copy(m)
        getattr(c, field)[0] = v
        assert lib.trk_check_atomic_parameters(ctypes.byref(c), photons, why, 256) == abi.TRK_ERR_ARG, (field, v)
        assert why.value, (field, v)
    # trk_ctx_create refuses such a descriptor before it touches the device
    c = abi.Material.from_buffer_copy(m)
    c.Auger[0] = -1.0
    h = ctypes.c_void_p()
    rc = lib.trk_ctx_create(ctypes.byref(c), ctypes.byref(ion), ctypes.byref(opt), ctypes.byref(h))
    assert rc == abi.TRK_ERR_ARG
    buf = ctypes.create_string_buffer(512)
    lib.trk_last_error(h, buf, 512)
    assert b"EADL" in buf.value
    lib.trk_ctx_destroy(h)
