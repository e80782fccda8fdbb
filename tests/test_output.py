"""Host-side numbers around a run: the reference's own radial grid
(Radius_for_distributions, Sorting_output_data.f90:1363-1435) and the values
MAIN and Save_output write after it (trekis_gpu.output)."""
import numpy as np
import pytest

from trekis_gpu import configs, inputs

from test_inputs import TWO_ELEMENTS


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


def _read_matrix(path):
    """(header tokens, row labels, values) of a radius x time file."""
    from trekis_gpu.cache import parse_real
    lines = open(path).read().splitlines()
    rows = [[parse_real(x) for x in l.split()] for l in lines[1:] if l.strip()]
    return lines[0], np.array([r[0] for r in rows]), np.array([r[1:] for r in rows])


def test_output_files_follow_save_output(tmp_path):
    """write_output() writes Save_output's files (SO:536-1138): names with
    element and shell designators ('Radial_Al_L-shell_holes_...',
    'Radial_Valence_holes_...'), time headers in f10.2 with '[fs]', rows
    labelled in f9.1 and values in E25.16 that read back as the normalised
    numbers to 16 digits; the VB temperature only for designator 63, the
    hole-energy files split into potential and kinetic for the VB."""
    from trekis_gpu import output
    t = _fake_tallies(Nt=5, NR=4)
    E, D = configs.dos_insulator()
    mat = inputs.read_cdf(TWO_ELEMENTS, layer=100.0, dos=(E, D))
    times = [10.0, 20.0, 30.0, 40.0, 50.0]
    R = [1.0, 2.5, 10.0, 100.0]
    dos_E = np.linspace(-10.0, 0.0, 7)
    names = output.write_output(str(tmp_path), t, 4, mat, times, R, include_photons=True, dos_E=dos_E,
                                hole_mass=1.0)
    n = output.normalise(t, 4)
    want = {"Electrons_theta_distribution.txt", "VB_holes_theta_distribution.txt", "Total_numbers.txt",
            "Hole_mean_diffusion_coefficient.txt", "Total_energies.txt", "Radial_electron_density[1_cm^-3].txt",
            "Radial_photon_energy[eV_A^-3].txt", "Electron_distribution_vs_E[1_eV].txt",
            "VB_holes_distribution_vs_E[1_eV].txt", "Radial_Lattice_temperature[K].txt",
            "Radial_Al_L-shell_holes_density[1_cm^-3].txt", "Radial_Al_L-shell_holes_energy[eV_A^-3].txt",
            "Radial_Valence_holes_density[1_cm^-3].txt", "Radial_Valence_holes_pot_energy[eV_A^-3].txt",
            "Radial_Valence_holes_kin_energy[eV_A^-3].txt", "Radial_Valence_holes_temperature[K].txt",
            "Radial_O_K-shell_holes_density[1_cm^-3].txt"}
    assert want <= set(names) and all((tmp_path / x).exists() for x in names)
    head, lab, val = _read_matrix(tmp_path / "Radial_electron_density[1_cm^-3].txt")
    assert head == "#Radius[A] " + "".join(f"{x:10.2f}[fs]   " for x in times) + " "
    assert lab.tolist() == R
    assert np.allclose(val, (n["Out_ne"] * 1e24).T, rtol=2e-16 * 8, atol=0)
    first = open(tmp_path / "Radial_electron_density[1_cm^-3].txt").read().splitlines()[1]
    assert first[:9] == "      1.0" and len(first) == 9 + 25 * len(times) + 1
    _, lab, val = _read_matrix(tmp_path / "Radial_Valence_holes_kin_energy[eV_A^-3].txt")
    assert np.allclose(val, n["Out_Ehkin"][:, :, 0, 1].T, rtol=1e-15, atol=0)
    _, lab, val = _read_matrix(tmp_path / "Radial_Al_L-shell_holes_energy[eV_A^-3].txt")
    assert np.allclose(val, n["Out_Eh"][:, :, 0, 0].T, rtol=1e-15, atol=0)
    # spectrum rows are labelled by Out_R (SO:867), its header has no unit
    lines = open(tmp_path / "Electron_distribution_vs_E[1_eV].txt").read().splitlines()
    assert lines[0] == "#Energy[eV]   Spectrum[arb.units]"
    assert lines[1] == "#time[fs]:  " + "".join(f"{x:10.2f}" for x in times) + " "
    # theta: 180 rows labelled 1..180
    _, lab, val = _read_matrix(tmp_path / "VB_holes_theta_distribution.txt")
    assert lab.tolist() == [float(i) for i in range(1, 181)] and np.allclose(val, n["Out_theta_h"][:5].T, rtol=1e-15)
    assert open(tmp_path / "VB_holes_theta_distribution.txt").readline().startswith("Angle[deg] ")
    # totals: six columns with photons, emitted quantities zero
    rows = np.loadtxt(tmp_path / "Total_numbers.txt", comments="#")
    assert rows.shape == (5, 6) and (rows[:, 2] == 0).all() and np.allclose(rows[:, 5], n["Out_tot_Nphot"])
    head = open(tmp_path / "Total_energies.txt").readline()
    assert "Al_L-shell[eV]" in head and "Valence[eV]" in head and "O_K-shell[eV]" in head
    rows = np.loadtxt(tmp_path / "Total_energies.txt", comments="#")
    assert rows.shape == (5, 5 + 3) and np.allclose(rows[:, 5], n["Out_E_h"][:, 0, 0])
    # no hole transport: the diffusion file exists and is empty
    output.write_output(str(tmp_path / "heavy"), t, 4, mat, times, R, include_photons=False, hole_mass=1e30)
    assert (tmp_path / "heavy" / "Hole_mean_diffusion_coefficient.txt").read_text() == ""
    assert not (tmp_path / "heavy" / "Radial_photon_density[1_cm^-3].txt").exists()


def test_emission_files_follow_save_output(tmp_path):
    """With emission on (work function > 0): the emitted electrons' spectrum
    file (SO:592-615) has the time header of the radial files, rows
    labelled by Out_R/10 in f9.1 and the normalised spectrum; the second and
    fourth columns of Total_numbers.txt (SO:630-640) are the normalised
    emitted number and energy.  Without emission the file is not written."""
    from trekis_gpu import output
    t = _fake_tallies(Nt=5, NR=4)
    r = np.random.default_rng(11)
    t["Out_Ne_Em"] = np.cumsum(r.integers(0, 5, 5)).astype(float)
    t["Out_E_Em"] = np.cumsum(r.uniform(0.0, 50.0, 5))
    t["Out_Ee_vs_E_Em"] = r.uniform(0.0, 1.0, (5, 4))
    E, D = configs.dos_insulator()
    mat = inputs.read_cdf(TWO_ELEMENTS, layer=100.0, dos=(E, D))
    mat.work_function, mat.bar_length, mat.bar_height = 4.5, 10.0, 6.18
    times = [10.0, 20.0, 30.0, 40.0, 50.0]
    R = [1.0, 3.0, 10.0, 100.0]
    names = output.write_output(str(tmp_path), t, 4, mat, times, R, include_photons=False)
    assert "Emitted_electron_distribution_vs_E[1_eV].txt" in names
    head, lab, val = _read_matrix(tmp_path / "Emitted_electron_distribution_vs_E[1_eV].txt")
    assert head == "#Energy[eV] " + "".join(f"{x:10.2f}[fs]   " for x in times) + " "
    assert lab.tolist() == [0.1, 0.3, 1.0, 10.0]
    assert np.allclose(val, (t["Out_Ee_vs_E_Em"] / 4.0).T, rtol=1e-15, atol=0)
    rows = np.loadtxt(tmp_path / "Total_numbers.txt", comments="#")
    assert rows.shape == (5, 5)
    assert np.allclose(rows[:, 2], t["Out_Ne_Em"] / 4.0, rtol=1e-15) and np.allclose(rows[:, 4], t["Out_E_Em"] / 4.0)
    mat.work_function = 0.0
    names = output.write_output(str(tmp_path / "none"), t, 4, mat, times, R, include_photons=False)
    assert "Emitted_electron_distribution_vs_E[1_eV].txt" not in names


def test_photon_paths_from_cross_sections():
    """EPDL branch of Tot_Phot_IMFP (CS:855-877): a core shell's photon path
    is 1/(n x_a sigma), the valence band's 1/(n sigma_VB); no absorption
    below a shell's edge gives 1e30."""
    mat = configs._sio2_like(layer=100.0)
    E = np.geomspace(1.0, 1.0e5, 50)
    sig = configs.synthetic_photoabsorption(mat)
    L = configs.photon_paths_from_cross_sections(mat, E, sig)
    n = mat.At_Dens * 1e-24
    shells = [(ia, js, sh) for ia, a in enumerate(mat.atoms) for js, sh in enumerate(a.shells)]
    assert L.shape == (len(shells), len(E))
    for s, (ia, js, sh) in enumerate(shells):
        x = 1.0 if (ia == 0 and js == len(mat.atoms[0].shells) - 1) else mat.atoms[ia].pers / mat.sum_pers
        above = E >= sh.Ip
        assert (L[s][~above] == 1e30).all()
        assert np.allclose(L[s][above] * n * x * sig(s, E[above]), 1.0, rtol=1e-14)
        # the E^-3 fall above the edge: the path grows as E^3
        if above.sum() > 2:
            e1, e2 = E[above][0], E[above][-1]
            assert L[s][above][-1] / L[s][above][0] == pytest.approx((e2 / e1) ** 3, rel=1e-12)


def test_parameter_echo_follows_print_parameters(lib, tmp_path):
    """'!Parameters.txt' (print_parameters, SO:42-337; Save_output SO:489-506):
    the option lines for the run's models, and the atomic table, whose k-sum
    of every core shell is its electron count (the oscillators are normalised
    by that sum rule with the same plasma frequency, CS:600-613)."""
    import datetime
    import trekis_gpu as tg
    from trekis_gpu import output
    mat, ion, opt, run = tg.config_material("opt_fixed_zeff")
    mat.prepare(lib)
    t = tg.TallySet(11, 4, mat)
    t["Out_tot_E"][:] = np.linspace(0.0, 5000.0, 11)
    name = output.write_parameters(str(tmp_path), mat, ion, opt, 4, lib, tallies=t, zeff=3.5, duration=3725.25,
                                   cdf_file="x.cdf", when=datetime.datetime(2026, 1, 2, 3, 4, 5, 6000))
    lines = open(tmp_path / name).read().splitlines()
    assert lines[0] == "*" * 56 and lines[-1] == "-" * 56
    assert " Ion fixed charge 3.500 [e] is used with 3.500 model for SHI." in lines
    assert " Inelastic scattering calculated with free electrons" in lines
    assert " Elastic scattering calculated with phonon CDF (Ritchie-Howie)" in lines
    assert " starting with 0.01 [fs] increasing by " + f"{opt.Tim:.3f}" + " in logarithmic time-scale" in lines
    assert " Temperature of the target: 300.000 [K]" in lines
    assert " CDF file used: x.cdf" in lines
    assert "Printed at 03:04:05:006 on 02/01/2026" in lines
    assert "MC calculated energy loss (Se): " + f"{5000.0 / 4 / mat.layer:.2f}" + " [eV/A]" in lines
    assert "Duration of calculation:        1 hour 2 mins 5.250 sec" in lines
    rows = [ln for ln in lines if ln[:11].rstrip().endswith("-shell")]
    core = [sh for ia, a in enumerate(mat.atoms) for js, sh in enumerate(a.shells)
            if not (ia == 0 and js == len(a.shells) - 1)]
    assert len(rows) == len(core)
    for ln, sh in zip(rows, core):
        assert float(ln[11 + 3 + 8 + 10 + 10 + 24:][:9]) == pytest.approx(sh.Nel, abs=0.006), ln
        assert float(ln[14:22]) == sh.Nel and float(ln[22:32]) == pytest.approx(sh.Ip, abs=0.006)
    assert sum(ln.startswith("Valence") for ln in lines) == 1
    assert output.duration_text(59.5) == "59.500 sec"
    assert output.duration_text(2 * 86400 + 7200 + 61.0) == "2 days 2 hours 1 min 1.000 sec"
