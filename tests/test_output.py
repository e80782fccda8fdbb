"""Host-side numbers around a run: the reference's own radial grid
(Radius_for_distributions, Sorting_output_data.f90:1363-1435) and the values
MAIN and Save_output write after it (trekis_gpu.output)."""
import numpy as np

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
