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
EXAMPLE_CDF = os.path.join(os.path.dirname(os.path.abspath(configs.__file__)), "examples", "INPUT_CDF", "al2o3_like.cdf")

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
])
def test_forms_needing_other_inputs_are_rejected(edit, msg):
    with pytest.raises(ValueError, match=msg):
        inputs.parse_cdf(TWO_ELEMENTS.replace(*edit))


def test_negative_designator_is_a_beb_shell():
    """A negative designator keeps the shell's oscillators for the ion and
    gives its electrons and holes the BEB cross section (RF:1556-1561): the
    shell becomes kocs = 2 with the given mean kinetic energy, and without
    one it is refused (the reference reads it from EADL)."""
    t = TWO_ELEMENTS.replace("2	2	80.0d0	8	27.27d0", "2	-2	80.0d0	8	27.27d0")
    f = inputs.parse_cdf(t)
    assert f.atoms[0]["shells"][0]["beb"] and f.atoms[0]["shells"][0]["shl_num"] == 2
    E, D = configs.dos_insulator()
    with pytest.raises(ValueError, match="EADL"):
        inputs.read_cdf(t, layer=100.0, dos=(E, D))
    mat = inputs.read_cdf(t, layer=100.0, dos=(E, D), ek=[[110.0, None], [None]])
    sh = mat.atoms[0].shells[0]
    assert (sh.kocs, sh.Ek) == (2, 110.0) and mat.atoms[0].shells[1].kocs == 1


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
        # Auger times the library can use: positive, at most 1e30 fs, 1e23 fs for the band
        assert all(0.0 < s["auger"] <= 1.0e30 for a in f.atoms for s in a["shells"]), p
        assert f.atoms[0]["shells"][-1]["auger"] == 1.0e23, p
        assert f.dens > 0 and math.isfinite(f.E_F), p
        ok += 1
    assert ok >= 10, (ok, rejected)
    f = inputs.parse_cdf(open(os.path.join(REF_CDF, "Al2O3.cdf")).read())
    assert [(a["Z"], a["pers"]) for a in f.atoms] == [(13, 2.0), (8, 3.0)]
    assert [s["Ip"] for a in f.atoms for s in a["shells"]] == [1559.1, 89.0, 8.8, 538.0]
    assert f.atoms[0]["shells"][2]["osc"][0] == (25.4, 275.0, 12.5)
    assert f.egap == 0.1 and len(f.phonon) == 2



@pytest.mark.skipif(not os.path.isdir(REF_CDF), reason="the reference's input folder is not present")
def test_reference_material_and_dos_prepare():
    """The reference's Al2O3 material and DOS files, read and prepared by the
    library's host functions (DOS regridded on 0.1 eV from the VB top, the
    integrated DOS holding the VB electrons)."""
    import trekis_gpu as tg
    dos = inputs.read_dos(os.path.join(os.path.dirname(REF_CDF), "INPUT_DOS", "Al2O3.dos"))
    assert len(dos[0]) > 50 and np.all(np.diff(dos[0]) > 0)
    mat = inputs.read_cdf(os.path.join(REF_CDF, "Al2O3.cdf"), layer=100.0, dos=dos)
    assert mat.gap_differs and mat.egap_file == 0.1          # the density line's quirk, see parse_cdf
    mat.prepare(tg.load_library())
    E = mat.dos["E"]
    assert mat.n_dos == int(np.floor((dos[0][-1] - dos[0][0]) / 0.1))
    assert E[0] == 0.0 or E[0] < 0.1
    assert np.all(mat.dos["int_DOS"] >= 0.0)


PARAMS = """Test    ! material name
79         ! SHI atomic number
2166.7     ! [MeV] total SHI energy
-1         ! [proton mass], negative: the element's
1000.0     ! [fs] total time
2000.0 0   ! [fs] timestep; kind of time grid
10.0d0     ! [eV] cut-off
100.0      ! [A] layer
0.0        ! [K] temperature
4   0.0    ! Zeff model; fixed value
0          ! point-like ion
0   0      ! elastic: Mott; target Zeff
7   -1     ! dispersion relation; m_eff
1          ! plasmon integration limit
1.0        ! hole mass
1          ! photons
-1.0 10.0d0 6.18d0  ! work function, barrier length, height
10000      ! iterations
-1         ! threads

!----------------
DOS Other.dos
!CDF Ignored.cdf
"""


def test_parameter_file_fields_and_corrections():
    """Read_input_file's order and corrections (RF:256-372)."""
    p = inputs.parse_parameters(PARAMS)
    assert p["material"] == "Test" and p["Z"] == 79 and p["E_MeV"] == 2166.7
    assert p["mass"] == inputs.ELEMENT_MASS[79]           # negative mass: the element's
    assert p["dt"] == 1000.0 and p["dt_flag"] == 0         # dt at most Tim on a linear grid
    assert p["fixed_Zeff"] == 79.0                         # fixed Zeff within (0, Z]
    assert p["kind_of_DR"] == 1 and p["El_eff_mass"] == 1.0
    assert p["plasmon_Emax"] == 1 and p["include_photons"] == 1 and p["work_function"] == 0.0
    assert p["NMC"] == 10000 and p["DOS_file"] == "Other.dos" and p["CDF_file"] is None
    q = inputs.parse_parameters(PARAMS.replace("2000.0 0   !", "1.5 1   !"))
    assert q["dt"] == 2.0 and q["dt_flag"] == 1           # at least 2 on a logarithmic grid
    with pytest.raises(ValueError, match="ends early"):
        inputs.parse_parameters("\n".join(PARAMS.splitlines()[:8]))


@pytest.mark.skipif(not os.path.isdir(REF_CDF), reason="the reference's input folder is not present")
def test_reference_input_tree_loads():
    """The reference's own INPUT_PARAMETERS.txt with its material and DOS
    files: material, ion and options, prepared by the library's host code."""
    import trekis_gpu as tg
    root = os.path.dirname(REF_CDF)
    p = inputs.parse_parameters(open(os.path.join(root, "INPUT_PARAMETERS.txt")).read())
    assert p["include_photons"] == 1
    # radiative decay is on there, and its times come from EADL in the reference
    with pytest.raises(ValueError, match="radiative decay times"):
        inputs.load_inputs(os.path.join(root, "INPUT_PARAMETERS.txt"))
    mat, ion, opt, nmc = inputs.load_inputs(os.path.join(root, "INPUT_PARAMETERS.txt"), radiat="off")
    assert all(s.radiat == 1.0e23 for a in mat.atoms for s in a.shells)
    assert mat.name and ion.Zat == p["Z"] and ion.E == p["E_MeV"] * 1.0e6 and nmc == p["NMC"]
    assert (opt.Tim, opt.dt, opt.dt_flag) == (p["Tim"], p["dt"], p["dt_flag"])
    mat.prepare(tg.load_library())
    assert mat.n_dos > 10


def test_auger_times_the_reference_reads_from_eadl_are_refused():
    """An Auger time <= 0 or > 1e30 fs in a material file makes the reference
    read it from EADL (Dealing_with_EADL.f90:368-370); this reader refuses it.
    The valence band never decays by Auger (1e23 fs, EADL:366-367)."""
    text = open(EXAMPLE_CDF).read()
    for bad in ("-3.6", "0.0", "2.0d31"):
        with pytest.raises(ValueError, match="EADL"):
            inputs.parse_cdf(text.replace("538.0e0\t2\t4.613e0", "538.0e0\t2\t" + bad))
    f = inputs.parse_cdf(text.replace("24\t1.0d23", "24\t-1.0"))
    assert f.atoms[0]["shells"][-1]["auger"] == 1.0e23


def test_dsf_files_of_the_input_tree(tmp_path):
    """kind_of_EMFP = 2 reads INPUT_DSF/<material>/<material>_{Electron,Hole}_
    DSF_Differential_EMFPs_<T>K.dat (RF:316-322); without them the run falls
    back to Mott scattering (RF:2543-2552)."""
    import shutil
    from trekis_gpu import dsf
    src = os.path.join(os.path.dirname(os.path.abspath(configs.__file__)), "examples")
    for d in ("INPUT_CDF", "INPUT_DOS"):
        shutil.copytree(os.path.join(src, d), tmp_path / d)
    lines = open(os.path.join(src, "INPUT_PARAMETERS.txt")).read().splitlines(True)
    lines[8] = "300.0      ! [K] temperature of the target\n"
    lines[11] = "2   0      ! elastic cross sections: 2=DSF\n"
    par = tmp_path / "INPUT_PARAMETERS.txt"
    par.write_text("".join(lines))
    mat, ion, opt, _ = inputs.load_inputs(str(par), radiat="off")
    assert opt.kind_of_EMFP == 0 and mat.dsf is None
    folder = tmp_path / "INPUT_DSF" / "al2o3_like"
    folder.mkdir(parents=True)
    blocks = {}
    for who, top in (("Electron", 1.0e4), ("Hole", 30.0)):
        blocks[who] = dsf.synthetic_dsf_blocks(0.01, top, 300.0, nE=6, nW=40)
        (folder / f"al2o3_like_{who}_DSF_Differential_EMFPs_300K.dat").write_text(dsf.dsf_file_text(*blocks[who]))
    mat, ion, opt, _ = inputs.load_inputs(str(par), radiat="off")
    assert opt.kind_of_EMFP == 2
    for t, who in zip(mat.dsf, ("Electron", "Hole")):
        want = dsf.process_dsf(*blocks[who])
        for x, y in zip(t.arrays(), want.arrays()):
            assert np.array_equal(x, y)
