"""Table caches in the reference's file formats (trekis_gpu.cache).

* field writers: every double reads back bit for bit; the layouts are the
  Fortran edit descriptors' (0.ddd...E+ee, d.ddd...E+ee, exponents past 99
  without the letter), which parse_real reads back;
* shell names: the EADL designators the reference names its per-shell files
  by (define_PQN, Dealing_with_EADL.f90:1016-1214), at every group boundary;
* file names composed as Analytical_IMFPs.f90 composes them (electrons
  AI:259-290, holes 386-405, photons 505-520, elastic 917-950 / 1066-1127 /
  1195-1201 / 1307-1375, the ion AI:2316-2350, diff_CS rows 672-682);
* a full round trip on the oracle's tables: every table and every cumulative
  row written and read back equals the one written, bit for bit; a file one
  line short is refused.
"""
import math
import os

import numpy as np
import pytest


import trekis_gpu as tg  # noqa: E402
from trekis_gpu import abi, cache  # noqa: E402


def test_fields_round_trip_bitwise():
    g = np.random.default_rng(7)
    vals = list(np.exp(g.uniform(-700, 700, 4000))) + list(-np.exp(g.uniform(-50, 50, 200)))
    vals += [0.0, 1.0, 1e-300, 5e-324, 1.7976931348623157e308, 0.1, 1e30, 1e23, 123.456]
    for v in vals:
        for fmt in (cache.fortran_e, cache.fortran_es):
            s = fmt(float(v))
            assert len(s) == 25, s
            assert cache.parse_real(s) == float(v), (v, s)
    assert cache.fortran_e(123.456) == "  0.12345600000000000E+03"
    assert cache.fortran_es(123.456) == "   1.2345600000000000E+02"
    assert cache.fortran_e(1e-150).strip() == "0.10000000000000000-149"
    s = cache.fortran_es(-2.5e200).strip()
    assert s.startswith("-2.50000000000000") and s.endswith("+200") and "E" not in s
    assert cache.fortran_f(12.5) == "      12.5000000000000000"
    assert cache.parse_real("0.1234-100") == 0.1234e-100
    assert cache.parse_real("1.5D+02") == 150.0
    assert cache.parse_real("-0.25E-3") == -0.25e-3


def test_shell_names_at_every_group():
    pins = {0: "K-shell", 1: "K-shell", 2: "L-shell", 3: "L1-shell", 4: "L23-shell", 6: "L3-shell", 7: "M-shell",
            12: "M45-shell", 14: "M5-shell", 15: "N-shell", 23: "N67-shell", 25: "N7-shell", 26: "O-shell",
            37: "O89-shell", 39: "O9-shell", 40: "P-shell", 54: "P1011-shell", 56: "P11-shell", 57: "Q-shell",
            59: "Q23-shell", 61: "Q3-shell", 62: "Q...-shell", 63: "Valence", 70: "Valence"}
    for d, name in pins.items():
        assert cache.shell_name(d) == name, d
    assert len(cache.SHELL_NAMES) == 62 and len(set(cache.SHELL_NAMES)) == 62


def _names(T=300.0, m_eff=1.0, dr=1, emfp=1, zeff=0, ion_Z=54, kind_zeff=0):
    mat = abi.Material()
    mat.temp = T
    mat.El_eff_mass = m_eff
    opt = abi.Options()
    opt.kind_of_DR, opt.kind_of_EMFP, opt.CDF_elast_Zeff = dr, emfp, zeff
    ion = abi.Ion()
    ion.Zat, ion.Kind_Zeff, ion.Kind_ion = ion_Z, kind_zeff, 0
    return cache.file_names(mat, ion, opt)


def test_file_names_follow_the_reference_rules():
    n = _names()
    assert n["e_inel"] == "OUTPUT_Electron_IMFPs_Free_CDF_Me_300.00_K.dat"
    assert n["h_inel"] == "OUTPUT_Hole_IMFPs_CDF_CDF_300.00_K.dat"
    assert n["phot"] == "OUTPUT_Photon_IMFPs_CDF_300.00_K.dat"
    assert n["e_el"] == "OUTPUT_Electron_EMFPs_CDF_Zeff_300.00_K.dat"
    assert n["h_el"] == "OUTPUT_Hole_CDF_EMFPs_300.00_K.dat"
    assert n["shi"] == "OUTPUT_Xe_CDF_Barkas_P_IMFP.dat" and n["shi_dedx"] == "OUTPUT_Xe_CDF_Barkas_P_dEdx.dat"
    assert _names(m_eff=0.0, dr=3)["e_inel"] == "OUTPUT_Electron_IMFPs_Ritchie_CDF_DOS_300.00_K.dat"
    assert _names(m_eff=0.5, dr=2, T=0.0)["e_inel"] == "OUTPUT_Electron_IMFPs_Plasmon_pole_CDF_0.50_me_0.00_K.dat"
    assert _names(m_eff=-1.0)["e_inel"].endswith("_CDF_Me_300.00_K.dat")
    assert _names(zeff=1)["e_el"] == "OUTPUT_Electron_EMFPs_CDF_Z=1_300.00_K.dat"
    m = _names(emfp=0)
    assert m["e_el"] == "OUTPUT_Electron_EMFPs_Mott.dat" and m["h_el"] == "OUTPUT_Hole_Mott_EMFPs_300.00_K.dat"
    m = _names(emfp=-1)
    assert m["e_el"] == "OUTPUT_Electron_No_elas_EMFPs.dat" and m["h_el"] == "OUTPUT_Hole_No_elas_EMFPs.dat"
    m = _names(emfp=2)
    assert m["e_el"] == "OUTPUT_Electron_DSF_EMFPs_300.00_K.dat" and m["h_el"] == "OUTPUT_Hole_DSF_EMFPs_300.00_K.dat"
    assert _names(ion_Z=79, kind_zeff=4)["shi"] == "OUTPUT_Au_CDF_fixed_P_IMFP.dat"
    # diff_CS rows: the file name without 'OUTPUT_' and '.dat', atom, shell, f14.3 energy
    assert cache._stem(n["e_inel"]) == "Electron_IMFPs_Free_CDF_Me_300.00_K"
    assert os.path.basename(cache._row_file("x", "Electron_IMFPs_Free_CDF_Me_300.00_K_Al_Valence", 12.3456)) == \
        "Electron_IMFPs_Free_CDF_Me_300.00_K_Al_Valence_12.346.dat"


class _OracleTables:
    """The oracle's tables behind the context calls the cache module makes;
    set_table / set_row are recorded, not applied."""

    def __init__(self, o):
        self.o = o
        self.mat, self.ion, self.opt, self.desc = o.mat, o.ion, o.opt, o.desc
        self.tables, self.rows, self.finished = {}, {}, False

    def table(self, kind, shell=0):
        return self.o.table(kind, shell)

    def row(self, kind, shell, i_E):
        return self.o.row(kind, shell, i_E)

    def set_table(self, kind, shell, L=None, dEdx=None, E=None):
        self.tables[(kind, shell)] = (L, dEdx)

    def set_row(self, kind, shell, i_E, hw, lam):
        self.rows[(kind, shell, i_E)] = (np.array(hw), np.array(lam))

    def finish_tables(self):
        self.finished = True


@pytest.fixture(scope="module")
def oracle_cdf():
    from oracle import Oracle
    lib = tg.load_library()
    mat, ion, opt, _ = tg.config_material("test_cdf")
    mat.prepare(lib)
    o = Oracle(mat, ion, opt)
    o.build_tables(threads=8)
    return o


def test_round_trip_through_files_is_bitwise(oracle_cdf, tmp_path):
    o = oracle_cdf
    src = _OracleTables(o)
    names = cache.write_cache(src, str(tmp_path))
    for key in ("e_inel", "h_inel", "phot", "e_el", "h_el", "shi", "shi_dedx"):
        assert os.path.exists(tmp_path / names[key]), key
    dst = _OracleTables(o)
    cache.seed_from_cache(dst, str(tmp_path))
    assert dst.finished
    nsh = sum(o.desc.n_shells)
    kinds = [abi.TAB_E_INEL, abi.TAB_H_INEL, abi.TAB_PHOT, abi.TAB_SHI]
    for kind in kinds:
        for s in range(nsh):
            _, L, D = o.table(kind, s)
            got_L, got_D = dst.tables[(kind, s)]
            assert np.array_equal(got_L, L), (kind, s)
            if kind == abi.TAB_SHI:
                assert np.array_equal(got_D, D)
    for kind in (abi.TAB_E_EL, abi.TAB_H_EL):
        assert np.array_equal(dst.tables[(kind, 0)][0], o.table(kind, 0)[1])
    # every row the reference keeps in diff_CS/ (every electron shell, the
    # valence band of holes, both elastic sets)
    nE = len(o.table(abi.TAB_E_INEL, 0)[0])
    nH = len(o.table(abi.TAB_H_INEL, 0)[0])
    nEl = len(o.table(abi.TAB_E_EL, 0)[0]) + len(o.table(abi.TAB_H_EL, 0)[0])
    assert len(os.listdir(tmp_path / "diff_CS")) <= nE * nsh + nH + nEl
    n_rows = 0
    for (kind, s, i), (hw, lam) in dst.rows.items():
        h0, l0 = o.row(kind, s, i)
        assert np.array_equal(hw, h0) and np.array_equal(lam, l0), (kind, s, i)
        n_rows += 1
    assert n_rows > nE * nsh
    # the ion's energies are written in MeV: the reader checks them against
    # the grid to the conversion's rounding and keeps the grid
    E = o.table(abi.TAB_SHI, 0)[0]
    first = open(tmp_path / names["shi"]).readline()
    assert abs(cache.parse_real(first[:25]) * 1e6 / E[0] - 1.0) < 1e-15


def test_cache_of_another_grid_is_refused(oracle_cdf, tmp_path):
    o = oracle_cdf
    names = cache.write_cache(_OracleTables(o), str(tmp_path))
    p = tmp_path / names["e_inel"]
    lines = open(p).read().splitlines(True)
    open(p, "w").writelines(lines[:-1])
    with pytest.raises(cache.CacheMismatch, match="energy nodes"):
        cache.seed_from_cache(_OracleTables(o), str(tmp_path))
    # a shifted energy column is another grid too
    lines[3] = cache.fortran_f(float(lines[3][:25]) * (1 + 1e-9)) + lines[3][25:]
    open(p, "w").writelines(lines)
    with pytest.raises(cache.CacheMismatch, match="energy node 3"):
        cache.seed_from_cache(_OracleTables(o), str(tmp_path))
    assert math.isfinite(float(lines[0][:25]))


def test_dsf_elastic_files_round_trip(lib, tmp_path):
    # DSF elastic tables (AI:1392-1398): energy, total, emission and
    # absorption paths; no diff_CS rows
    from oracle import Oracle
    mat, ion, opt, _ = tg.config_material("opt_dsf")
    mat.prepare(lib)
    o = Oracle(mat, ion, opt)
    o.build_tables(threads=8, shi_window=(ion.E * 0.8, ion.E * 1.01))
    names = cache.write_cache(_OracleTables(o), str(tmp_path))
    for p, key in enumerate(("e_el", "h_el")):
        a = np.array(cache._rows(str(tmp_path / names[key])))
        t = mat.dsf[p]
        assert a.shape == (len(t.E), 4)
        assert np.array_equal(a[:, 2], t.dL_emit[:, -1]) and np.array_equal(a[:, 3], t.dL_absorb[:, -1])
    dst = _OracleTables(o)
    cache.seed_from_cache(dst, str(tmp_path))
    for kind in (abi.TAB_E_EL, abi.TAB_H_EL):
        assert np.array_equal(dst.tables[(kind, 0)][0], o.table(kind, 0)[1])
    assert not any(k[0] in (abi.TAB_E_EL, abi.TAB_H_EL) for k in dst.rows)
