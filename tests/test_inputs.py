"""Readers of the reference's material files (trekis_gpu.inputs): the
full-CDF material file (reading_material_parameters, Reading_files_and_
parameters.f90:1185-1644) and the DOS file (reading_material_DOS, RF:2340-
2450).  The format rules are pinned on small texts written here; the
shipped input folder of the reference, where present, is read as a whole."""
import glob
import math
import os

import numpy as np
import pytest

from trekis_gpu import configs, inputs

REF_CDF = "/root/reference/INPUT_CDF"

TWO_ELEMENTS = """Test oxide   ! name
2		! elements
13	2	! Z, contribution
8	3
3.95	6510.0	4.4	8.8	! density, sound, Fermi, gap
2		! shells of Al
2	2	80.0d0	8	27.27d0	! ncdf, designator, Ip, Nel, Auger
90.0	369.36	40.0
130.0	221.62	60.0
1	63	8.8D0	24	1.0d23
14.0	115.55	3.0
1
1	1	538.0e0	2	4.613e0
560.0	166.65	120.0
2		! phonon peaks
0.09	0.002	0.006
0.06	0.00005	0.002
"""


def test_full_cdf_file_fields():
    f = inputs.parse_cdf(TWO_ELEMENTS)
    assert f.name == "Test oxide"
    assert [(a["Z"], a["pers"], len(a["shells"])) for a in f.atoms] == [(13, 2.0, 2), (8, 3.0, 1)]
    assert (f.dens, f.vsound, f.E_F, f.egap) == (3.95, 6510.0, 4.4, 8.8)
    s = f.atoms[0]["shells"][0]
    assert (s["ncdf"], s["shl_num"], s["Ip"], s["Nel"], s["auger"]) == (2, 2, 80.0, 8.0, 27.27)
    assert s["osc"] == [(90.0, 369.36, 40.0), (130.0, 221.62, 60.0)]
    assert f.atoms[0]["shells"][1]["Ip"] == 8.8 and f.atoms[0]["shells"][1]["shl_num"] == 63
    assert f.phonon == [(0.09, 0.002, 0.006), (0.06, 0.00005, 0.002)]


def test_density_line_with_three_values_is_a_metal():
    """Four values are read, else three and gap 0 (RF:1326-1333): a '!'
    after three values is read as the fourth and fails, so the gap is the
    0.1 eV floor -- as for the reference's own files written that way."""
    t = TWO_ELEMENTS.replace("3.95	6510.0	4.4	8.8	! density, sound, Fermi, gap",
                             "3.95	6510.0	4.4	! density, sound, Fermi")
    assert inputs.parse_cdf(t).egap == 0.1
    t2 = TWO_ELEMENTS.replace("3.95	6510.0	4.4	8.8	!", "3.95	6510.0	4.4	0.05	!")
    assert inputs.parse_cdf(t2).egap == 0.1


def test_ionisation_potential_floor_and_no_phonon_block():
    t = TWO_ELEMENTS.replace("1	63	8.8D0	24	1.0d23", "1	63	0.0d0	24	1.0d23")
    t = t[:t.index("2		! phonon peaks")]
    f = inputs.parse_cdf(t)
    assert f.atoms[0]["shells"][1]["Ip"] == 0.1          # RF:1554
    assert f.phonon is None                               # end of file: single-pole phonon CDF (RF:1590-1593)


@pytest.mark.parametrize("edit,msg", [
    (("2		! elements\n13	2	! Z, contribution\n8	3\n", "Al2O3\n"), "EADL"),
    (("2	2	80.0d0	8	27.27d0", "0	2	80.0d0	8	27.27d0"), "BEB"),
    (("2	2	80.0d0	8	27.27d0", "2	-2	80.0d0	8	27.27d0"), "BEB"),
])
def test_forms_needing_other_inputs_are_rejected(edit, msg):
    with pytest.raises(ValueError, match=msg):
        inputs.parse_cdf(TWO_ELEMENTS.replace(*edit))


def test_records_span_blank_lines_and_bad_values_stop():
    """List-directed reads take values from the following records when a
    record runs short (blank lines hold none); a '!' where a value is
    expected stops the reference's reader and this one."""
    t = TWO_ELEMENTS.replace("130.0	221.62	60.0\n", "130.0	221.62	60.0\n\n\n")
    assert inputs.parse_cdf(t).atoms[0]["shells"][1]["osc"] == [(14.0, 115.55, 3.0)]
    t = TWO_ELEMENTS.replace("90.0	369.36	40.0", "90.0\n369.36	40.0")
    assert inputs.parse_cdf(t).atoms[0]["shells"][0]["osc"][0] == (90.0, 369.36, 40.0)
    with pytest.raises(ValueError, match="unreadable.*line 8"):
        inputs.parse_cdf(TWO_ELEMENTS.replace("90.0	369.36	40.0", "90.0	369.36	! no width"))
    with pytest.raises(ValueError, match="unreadable"):
        inputs.parse_cdf(TWO_ELEMENTS.replace("3.95	6510.0	4.4	8.8	!", "3.95	6510.0	!"))


def test_single_pole_marker_file_is_rejected():
    t = "Metal\n1\n13 1\n2.7 5000.0 11.17\nVALENCE\n1 63 0 0 0\n"
    with pytest.raises(ValueError, match="single-pole"):
        inputs.parse_cdf(t)


def test_read_cdf_builds_the_material():
    """Masses from the element table, the atomic density of RF:1338, the
    valence band as the first element's last shell, oscillators as given."""
    E, D = configs.dos_insulator()
    m = inputs.read_cdf(TWO_ELEMENTS, layer=100.0, dos=(E, D))
    sm = 2 * 26.9815386 + 3 * 15.9992
    assert m.At_Dens == 1.0e-3 * 3.95 / (configs.G_MP * sm / 5.0)
    assert m.N_VB_el == 24.0 and not m.gap_differs
    assert all(not s.normalize for a in m.atoms for s in a.shells)
    assert m.atoms[1].shells[0].osc == [(560.0, 166.65, 120.0)]
    assert m.phonon == [(0.09, 0.002, 0.006), (0.06, 0.00005, 0.002)] and not m.phonon_single_pole


def test_element_masses_cover_the_periodic_table():
    assert sorted(inputs.ELEMENT_MASS) == list(range(1, 119))
    assert inputs.ELEMENT_MASS[13] == 26.9815386 and inputs.ELEMENT_MASS[8] == 15.9992
    m = [inputs.ELEMENT_MASS[z] for z in range(1, 119)]
    assert all(b > a for a, b in zip(m[:17], m[1:18]))    # increasing through argon


def test_read_dos_two_columns():
    t = "  -0.74E+01   0.4E-02\n -7.3 0.007\n\n -7.2d0 3.7d-2\n"
    E, D = inputs.read_dos(t)
    assert np.array_equal(E, [-7.4, -7.3, -7.2]) and np.array_equal(D, [0.004, 0.007, 0.037])


def test_example_material_of_the_test_cdf_workload():
    mat, ion, opt, run = configs.config_material("test_cdf")
    assert mat.name == "Al2O3-like oxide" and len(mat.atoms) == 2 and not mat.gap_differs
    assert [len(a.shells) for a in mat.atoms] == [2, 1] and len(mat.phonon) == 2


@pytest.mark.skipif(not os.path.isdir(REF_CDF), reason="the reference's input folder is not present")
def test_reference_input_folder():
    """Every material file the reference ships parses, or is rejected for
    a stated reason (formula, single-pole markers, BEB shells, or values the
    reference's reader stops on as well)."""
    ok, rejected = 0, 0
    for p in sorted(glob.glob(os.path.join(REF_CDF, "*.cdf"))):
        try:
            f = inputs.parse_cdf(open(p, errors="replace").read())
        except ValueError as e:
            assert any(w in str(e) for w in ("EADL", "BEB", "single-pole", "unreadable")), (p, e)
            rejected += 1
            continue
        assert f.atoms and f.atoms[0]["shells"], p       # other elements may give all electrons to the VB
        assert all(s["Ip"] >= 0.1 and s["Nel"] > 0 for a in f.atoms for s in a["shells"]), p
        assert f.dens > 0 and math.isfinite(f.E_F), p
        ok += 1
    assert ok >= 10, (ok, rejected)
    f = inputs.parse_cdf(open(os.path.join(REF_CDF, "Al2O3.cdf")).read())
    assert [(a["Z"], a["pers"]) for a in f.atoms] == [(13, 2.0), (8, 3.0)]
    assert [s["Ip"] for a in f.atoms for s in a["shells"]] == [1559.1, 89.0, 8.8, 538.0]
    assert f.atoms[0]["shells"][2]["osc"][0] == (25.4, 275.0, 12.5)
    assert f.egap == 0.1 and len(f.phonon) == 2


def test_reference_radial_grid():
    """The 50 radii of Radius_for_distributions (SO:1395-1433) and Out_V:
    the cylinder shells' volumes add up to the outer cylinder's."""
    R, V = configs.reference_radial_grid(250.0)
    exp = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 15, 20, 30, 40, 50, 60, 70, 80, 90, 100,
           150, 200, 300, 400, 500, 600, 700, 800, 900, 1000, 1500, 2000, 3000, 4000, 5000, 6000, 7000, 8000,
           9000, 10000, 15000, 20000, 30000, 40000, 50000, 60000, 70000, 80000, 90000, 100000]
    assert R.tolist() == [float(x) for x in exp]
    assert V[0] == 1.0 / (1.0 * configs.G_PI * 250.0)
    assert np.isclose(np.sum(1.0 / V), configs.G_PI * 1.0e10 * 250.0, rtol=1e-12)


def _fake_tallies(Nt=5, NR=4, na=2, ms=2, seed=3):
    r = np.random.default_rng(seed)
    t = {k: r.uniform(0.0, 10.0, (Nt,)) for k in ["Out_tot_Ne", "Out_tot_Nphot", "Out_tot_E", "Out_E_e",
                                                   "Out_E_phot", "Out_E_at", "Out_diff_coeff"]}
    for k in ["Out_ne", "Out_Ee", "Out_nphot", "Out_Ephot", "Out_Ee_vs_E", "Out_Elat"]:
        t[k] = r.uniform(0.0, 10.0, (Nt, NR))
    t["Out_ne"][1, 2] = 0.0
    t["Out_Ee"][1, 2] = 0.0
    t["Out_Eh_vs_E"] = r.uniform(0.0, 1.0, (Nt, 7))
    for k in ["Out_nh", "Out_Eh", "Out_Ehkin"]:
        t[k] = r.uniform(0.0, 10.0, (Nt, NR, na, ms))
    t["Out_E_h"] = r.uniform(0.0, 1.0, (Nt, na, ms))
    t["Out_theta"] = r.uniform(0.0, 1.0, (Nt + 1, 180))
    t["Out_theta_h"] = r.uniform(0.0, 1.0, (Nt + 1, 180))
    return t


def test_output_tables_follow_save_output():
    """normalise() divides by NMC (MAIN:281-314); radial_tables() gives the
    files' columns: densities x 1e24, temperatures Ee/ne kB 2/3 (0 where no
    electron), lattice energy summed over the nodes so far (SO:931), the
    track energy adding the VB holes' energies (SO:960), the lattice
    temperature per atom (SO:988); rows are radii, columns time nodes."""
    from trekis_gpu import output
    t = _fake_tallies()
    n = output.normalise(t, 4)
    assert np.array_equal(n["Out_ne"], t["Out_ne"] / 4.0)
    E, D = configs.dos_insulator()
    mat = inputs.read_cdf(TWO_ELEMENTS, layer=100.0, dos=(E, D))
    tab = output.radial_tables(n, mat)
    assert tab["Radial_electron_density[1_cm^-3].txt"].shape == (4, 5)
    assert tab["Radial_electron_density[1_cm^-3].txt"][2, 3] == n["Out_ne"][3, 2] * 1.0e24
    Te = tab["Radial_electron_temperature[K].txt"]
    assert Te[2, 1] == 0.0 and Te[0, 0] == n["Out_Ee"][0, 0] / n["Out_ne"][0, 0] * 11604.0 * 2.0 / 3.0
    lat = tab["Radial_Lattice_energy[eV_A^-3].txt"]
    assert np.allclose(lat, np.cumsum(n["Out_Elat"], axis=0).T, rtol=1e-15, atol=0)
    assert np.allclose(np.diff(lat, axis=1), n["Out_Elat"][1:].T, rtol=1e-12, atol=1e-12)
    track = tab["Radial_Track_energy[eV_A^-3].txt"]
    assert np.allclose(track - lat, (n["Out_Eh"][:, :, 0, 1] + n["Out_Ehkin"][:, :, 0, 1]).T, rtol=1e-12, atol=1e-12)
    TL = tab["Radial_Lattice_temperature[K].txt"]
    assert np.allclose(TL / lat, 11604.0 * 2.0 / 3.0 / (mat.At_Dens * 1e-24), rtol=1e-13)
    assert tab[("holes_density", 1, 0)][3, 4] == n["Out_nh"][4, 3, 1, 0] * 1.0e24
