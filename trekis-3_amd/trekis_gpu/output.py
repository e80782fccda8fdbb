"""The numbers the reference writes after a run, from the Out_* arrays that
trk_run_mc fills (unnormalised sums).

normalise() divides every tally by the number of iterations, as MAIN does
before writing (Universal_MC_for_SHI_MAIN.f90:281-314); radial_tables()
forms the columns of the radial-distribution files of Save_output
(Sorting_output_data.f90:700-1060) with the same expressions: densities
in 1/A^3 scaled by 1e24 to 1/cm^3 (SO:728, 816, 1020), mean electron
energy as a temperature Ee/ne * kB * 2/3 (SO:782-786), lattice energy as the
sum over the time nodes so far (SO:931) and as a temperature per atom
(SO:988).  Each table has one row per radius (Out_R) and one column per time
node, as the files do; the text layout of the files (Intel list formats) is
not reproduced.
"""
import numpy as np

# Universal_Constants.f90:39, Boltzmann constant [K/eV]
G_KB = 11604.0

PER_ITERATION = ["Out_tot_Ne", "Out_tot_Nphot", "Out_tot_E", "Out_E_e", "Out_E_phot", "Out_E_at", "Out_E_h",
                 "Out_ne", "Out_Ee", "Out_nphot", "Out_Ephot", "Out_Ee_vs_E", "Out_Eh_vs_E", "Out_Elat",
                 "Out_theta", "Out_theta_h", "Out_diff_coeff", "Out_nh", "Out_Eh", "Out_Ehkin"]


def normalise(tallies, nmc):
    """Every Out_* array divided by dble(NMC) (MAIN:281-314)."""
    n = float(nmc)
    return {k: tallies[k] / n for k in PER_ITERATION}


def _cum_lattice(Elat):
    # SUM(Out_Elat(1:k, i)) for every node k: a running sum in node order
    out = np.empty_like(Elat)
    s = np.zeros(Elat.shape[1])
    for k in range(Elat.shape[0]):
        s = s + Elat[k]
        out[k] = s
    return out


def radial_tables(norm, mat, vb=None):
    """{file name: array [radius, time node]} of the radial files, from
    normalise()'s output.  mat: the MaterialDesc of the run (At_Dens, shells);
    vb: (atom, shell) whose hole energies the track-energy table adds to the
    lattice's (Lowest_Ip_At / Lowest_Ip_Shl, SO:960; default the first
    element's last shell, the valence band).  The per-shell hole tables
    (files 'Radial_<element>_<shell>_holes_density[1_cm^-3].txt' and so on,
    SO:1000-1060) are keyed ('holes_density', atom, shell),
    ('holes_pot_energy', ...), ('holes_kin_energy', ...)."""
    ne, Ee = norm["Out_ne"], norm["Out_Ee"]
    t = {}
    t["Radial_electron_density[1_cm^-3].txt"] = (ne * 1.0e24).T
    t["Radial_electron_energy[eV_A^-3].txt"] = Ee.T
    with np.errstate(divide="ignore", invalid="ignore"):
        temp = np.where(ne > 0.0, Ee / ne * G_KB * 2.0 / 3.0, 0.0)
    t["Radial_electron_temperature[K].txt"] = temp.T
    t["Radial_photon_density[1_cm^-3].txt"] = (norm["Out_nphot"] * 1.0e24).T
    t["Radial_photon_energy[eV_A^-3].txt"] = norm["Out_Ephot"].T
    t["Electron_distribution_vs_E[1_eV].txt"] = norm["Out_Ee_vs_E"].T
    t["VB_holes_distribution_vs_E[1_eV].txt"] = norm["Out_Eh_vs_E"].T
    lat = _cum_lattice(norm["Out_Elat"])
    t["Radial_Lattice_energy[eV_A^-3].txt"] = lat.T
    a, s = vb if vb is not None else (0, len(mat.atoms[0].shells) - 1)
    t["Radial_Track_energy[eV_A^-3].txt"] = (lat + norm["Out_Eh"][:, :, a, s] + norm["Out_Ehkin"][:, :, a, s]).T
    t["Radial_Lattice_temperature[K].txt"] = (lat / (mat.At_Dens * 1e-24) * G_KB * 2.0 / 3.0).T
    for ia, atom in enumerate(mat.atoms):
        for js in range(len(atom.shells)):
            t[("holes_density", ia, js)] = (norm["Out_nh"][:, :, ia, js] * 1.0e24).T
            t[("holes_pot_energy", ia, js)] = norm["Out_Eh"][:, :, ia, js].T
            t[("holes_kin_energy", ia, js)] = norm["Out_Ehkin"][:, :, ia, js].T
    return t
