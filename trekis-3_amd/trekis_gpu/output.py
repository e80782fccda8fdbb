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
node, as the files do.  write_output() writes the files themselves with
the reference's names, headers and Fortran fields.
"""
import math

import numpy as np

# Universal_Constants.f90:39, Boltzmann constant [K/eV]
G_KB = 11604.0

PER_ITERATION = ["Out_tot_Ne", "Out_tot_Nphot", "Out_tot_E", "Out_E_e", "Out_E_phot", "Out_E_at", "Out_E_h",
                 "Out_ne", "Out_Ee", "Out_nphot", "Out_Ephot", "Out_Ee_vs_E", "Out_Eh_vs_E", "Out_Elat",
                 "Out_theta", "Out_theta_h", "Out_diff_coeff", "Out_nh", "Out_Eh", "Out_Ehkin",
                 "Out_Ne_Em", "Out_E_Em", "Out_Ee_vs_E_Em"]


def normalise(tallies, nmc):
    """Every Out_* array divided by dble(NMC) (MAIN:281-314); the emitted
    electrons' arrays (MAIN:301-303) are zeros when the tallies have none."""
    n = float(nmc)
    out = {}
    for k in PER_ITERATION:
        try:
            out[k] = tallies[k] / n
        except KeyError:
            if not k.endswith("_Em"):
                raise
            shape = np.shape(tallies["Out_Ee_vs_E"] if k == "Out_Ee_vs_E_Em" else tallies["Out_tot_Ne"])
            out[k] = np.zeros(shape)
    return out


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
    t["Emitted_electron_distribution_vs_E[1_eV].txt"] = norm["Out_Ee_vs_E_Em"].T
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


# ---------------------------------------------------------------- files

def _e(x):
    # the '(e)' field of Intel Fortran for REAL(8): E25.16
    from .cache import fortran_e
    return fortran_e(float(x), 25, 16)


def _time_header(first, times, with_unit=True):
    return first + "".join(f"{t:10.2f}" + ("[fs]   " if with_unit else "") for t in times) + " \n"


def _matrix_file(path, first, times, labels, values, label_fmt="{:9.1f}"):
    """Header of time nodes, then a row per label: the label and one value
    per time node (values[row, node])."""
    with open(path, "w") as f:
        f.write(_time_header(first, times))
        for i, lab in enumerate(labels):
            f.write(label_fmt.format(lab) + "".join(_e(v) for v in values[i]) + " \n")


def _shell_tag(atom, shell):
    from .cache import ELEMENT_SYMBOL, shell_name
    if shell.shl_num < 63:
        return f"{ELEMENT_SYMBOL[atom.Z]}_{shell_name(shell.shl_num)}"
    return shell_name(shell.shl_num)


def write_output(folder, tallies, nmc, mat, time_grid, Out_R, include_photons, dos_E=None, hole_mass=None,
                 work_function=None):
    """Write the result files of Save_output (Sorting_output_data.f90:340-1140)
    into folder, from a run's unnormalised tallies (divided by nmc here, as
    MAIN:281-314 does): the angular distributions, Total_numbers.txt,
    Hole_mean_diffusion_coefficient.txt, Total_energies.txt, the radial
    electron / photon / lattice / track files, the energy spectra and every
    shell's hole files, with the reference's names, headers and Fortran
    fields ('(f10.2)' time headers, '(f9.1)' row labels, '(e)' values as
    E25.16); with emission on (work_function > 0, default the material's)
    also the emitted electrons' spectrum.  Quantities of the field model,
    which this library does not build, are written as zeros where the files
    hold them.  The parameter echo '!Parameters.txt' is write_parameters'.
    Returns the file names."""
    import os
    os.makedirs(folder, exist_ok=True)
    n = normalise(tallies, nmc)
    N = len(time_grid)
    times = np.asarray(time_grid, dtype=float)
    R = np.asarray(Out_R, dtype=float)
    hole_mass = mat.hole_mass if hole_mass is None else hole_mass
    work_function = getattr(mat, "work_function", 0.0) if work_function is None else work_function
    written = []

    def path(name):
        written.append(name)
        return os.path.join(folder, name)

    # angular distributions (SO:536-591)
    for name, key, first in (("Electrons_theta_distribution.txt", "Out_theta", "#Angle[deg] "),
                             ("VB_holes_theta_distribution.txt", "Out_theta_h", "Angle[deg] ")):
        th = n[key]
        with open(path(name), "w") as f:
            f.write(_time_header(first, times))
            for i in range(180):
                f.write(_e(float(i + 1)) + "".join(_e(th[k, i]) for k in range(N)) + " \n")
    # emitted electrons' spectrum (SO:592-615), rows labelled by Out_R/10
    if work_function > 0.0:
        _matrix_file(path("Emitted_electron_distribution_vs_E[1_eV].txt"), "#Energy[eV] ", times, R / 10.0,
                     radial_tables(n, mat)["Emitted_electron_distribution_vs_E[1_eV].txt"][:, :N])
    # totals (SO:618-640)
    Nem, Eem = n["Out_Ne_Em"], n["Out_E_Em"]
    with open(path("Total_numbers.txt"), "w") as f:
        if include_photons:
            f.write("#Time[fs]    Ne    Ne_Emitted    Energy[eV]     Energy_Emitted[eV] N_photons\n")
            for k in range(N):
                f.write("".join(_e(v) for v in (times[k], n["Out_tot_Ne"][k], Nem[k], n["Out_tot_E"][k], Eem[k],
                                                 n["Out_tot_Nphot"][k])) + "\n")
        else:
            f.write("#Time[fs]    Ne    Ne_Emitted    Energy[eV]     Energy_Emitted[eV]\n")
            for k in range(N):
                f.write("".join(_e(v) for v in (times[k], n["Out_tot_Ne"][k], Nem[k], n["Out_tot_E"][k], Eem[k]))
                        + "\n")
    # hole diffusion coefficient (SO:646-662): the file stays empty without hole transport
    with open(path("Hole_mean_diffusion_coefficient.txt"), "w") as f:
        if hole_mass < 1.0e3:
            f.write("#Time[fs]    Diffusion_coeff[cm^2/s]\n")
            for k in range(N):
                f.write(_e(times[k]) + _e(n["Out_diff_coeff"][k]) + "\n")
    # energies of electrons, atoms, field, photons and every shell's holes (SO:666-705)
    with open(path("Total_energies.txt"), "w") as f:
        f.write("#Time[fs]    Electrons[eV]   Atoms[eV]   Field[eV]  " + ("Photons[eV] " if include_photons else ""))
        for atom in mat.atoms:
            for sh in atom.shells:
                f.write(_shell_tag(atom, sh) + "[eV]    ")
        f.write(" \n")
        Eh = n["Out_E_h"]
        for k in range(N):
            vals = [times[k], n["Out_E_e"][k], n["Out_E_at"][k], 0.0] + ([n["Out_E_phot"][k]] if include_photons else [])
            for ia, atom in enumerate(mat.atoms):
                for js in range(len(atom.shells)):
                    vals.append(Eh[k, ia, js])
            f.write("".join(_e(v) for v in vals) + " \n")
    t = radial_tables(n, mat)
    radial = ["Radial_electron_density[1_cm^-3].txt", "Radial_electron_energy[eV_A^-3].txt",
              "Radial_electron_temperature[K].txt"]
    if include_photons:
        radial += ["Radial_photon_density[1_cm^-3].txt", "Radial_photon_energy[eV_A^-3].txt"]
    for name in radial:
        _matrix_file(path(name), "#Radius[A] ", times, R, t[name][:, :N])
    # spectra (SO:852-906): the electron spectrum's rows are labelled by Out_R
    ee = t["Electron_distribution_vs_E[1_eV].txt"]
    with open(path("Electron_distribution_vs_E[1_eV].txt"), "w") as f:
        f.write("#Energy[eV]   Spectrum[arb.units]\n")
        f.write(_time_header("#time[fs]:  ", times, with_unit=False))
        for i, r in enumerate(R):
            f.write(f"{r:9.1f}" + "".join(_e(v) for v in ee[i, :N]) + " \n")
    if dos_E is not None:
        _matrix_file(path("VB_holes_distribution_vs_E[1_eV].txt"), "#Energy[eV] ", times, np.asarray(dos_E),
                     t["VB_holes_distribution_vs_E[1_eV].txt"][:, :N])
    for name in ("Radial_Lattice_energy[eV_A^-3].txt", "Radial_Track_energy[eV_A^-3].txt",
                 "Radial_Lattice_temperature[K].txt"):
        _matrix_file(path(name), "#Radius[A] ", times, R, t[name][:, :N])
    # every shell's holes (SO:996-1138)
    for ia, atom in enumerate(mat.atoms):
        for js, sh in enumerate(atom.shells):
            _matrix_file(path(f"Radial_{_shell_tag(atom, sh)}_holes_density[1_cm^-3].txt"), "#Radius[A] ", times, R,
                         t[("holes_density", ia, js)][:, :N])
    for ia, atom in enumerate(mat.atoms):
        for js, sh in enumerate(atom.shells):
            tag = _shell_tag(atom, sh)
            if sh.shl_num < 63:
                _matrix_file(path(f"Radial_{tag}_holes_energy[eV_A^-3].txt"), "#Radius[A] ", times, R,
                             t[("holes_pot_energy", ia, js)][:, :N])
            else:
                _matrix_file(path(f"Radial_{tag}_holes_pot_energy[eV_A^-3].txt"), "#Radius[A] ", times, R,
                             t[("holes_pot_energy", ia, js)][:, :N])
                _matrix_file(path(f"Radial_{tag}_holes_kin_energy[eV_A^-3].txt"), "#Radius[A] ", times, R,
                             t[("holes_kin_energy", ia, js)][:, :N])
    for ia, atom in enumerate(mat.atoms):
        for js, sh in enumerate(atom.shells):
            if sh.shl_num == 63:
                nh = n["Out_nh"][:, :, ia, js]
                ek = n["Out_Ehkin"][:, :, ia, js]
                with np.errstate(divide="ignore", invalid="ignore"):
                    temp = np.where(nh > 0.0, ek / nh * G_KB * 2.0 / 3.0, 0.0)
                _matrix_file(path(f"Radial_{_shell_tag(atom, sh)}_holes_temperature[K].txt"), "#Radius[A] ", times,
                             R, temp.T[:, :N])
    return written


# Element names as Find_element_name gives them (Dealing_with_EADL.f90:
# 1217-1703), printed as the ion's full name in the parameter echo.
ELEMENT_NAME = (
    None, "Hydrogen", "Helium", "Lithium", "Beryllium", "Boron", "Carbon", "Nitrogen", "Oxygen", "Fluorine", "Neon",
    "Sodium", "Magnesium", "Aluminium", "Silicon", "Phosphorus", "Sulfur", "Chlorine", "Argon", "Potassium",
    "Calcium", "Scandium", "Titanium", "Vanadium", "Chromium", "Manganese", "Iron", "Cobalt", "Nickel", "Copper",
    "Zinc", "Gallium", "Germanium", "Arsenic", "Selenium", "Bromine", "Krypton", "Rubidium", "Strontium", "Yttrium",
    "Zirconium", "Niobium", "Molybdenum", "Technetium", "Ruthenium", "Rhodium", "Palladium", "Silver", "Cadmium",
    "Indium", "Tin", "Antimony", "Tellurium", "Iodine", "Xenon", "Caesium", "Barium", "Lanthanum", "Cerium",
    "Praseodymium", "Neodymium", "Promethium", "Samarium", "Europium", "Gadolinium", "Terbium", "Dysprosium",
    "Holmium", "Erbium", "Thulium", "Ytterbium", "Lutetium", "Hafnium", "Tantalum", "Tungsten", "Rhenium", "Osmium",
    "Iridium", "Platinum", "Gold", "Mercury", "Thallium", "Lead", "Bismuth", "Polonium", "Astatine", "Radon",
    "Francium", "Radium", "Actinium", "Thorium", "Protactinium", "Uranium", "Neptunium", "Plutonium", "Americium",
    "Curium", "Berkelium", "Californium", "Einsteinium", "Fermium", "Mendelevium", "Nobelium", "Lawrencium",
    "Rutherfordium", "Dubnium", "Seaborgium", "Bohrium", "Hassium", "Meitnerium", "Darmstadtium", "Roentgenium",
    "Copernicium", "Ununtrium", "Flerovium", "Ununpentium", "Livermorium", "Ununseptium", "Ununoctium",
)

_DASHES = "-" * 56
_STARS = "*" * 56


def duration_text(sec):
    """parse_time (SO:1439-1484): days, hours and minutes above their
    counts, then the seconds as f7.3."""
    days = hours = mins = 0.0
    if sec >= 60.0:
        mins = math.floor(sec / 60.0)
        sec = sec - mins * 60.0
        if mins > 60.0:
            hours = math.floor(mins / 60.0)
            mins = mins - hours * 60.0
            if hours > 24.0:
                days = math.floor(hours / 24.0)
                hours = hours - days * 24.0
    out = ""
    for v, one, many in ((days, " day", " days"), (hours, " hour", " hours"), (mins, " min", " mins")):
        if v > 1.0:
            out = (out + " " + str(int(v)) + many).strip()
        elif v > 0.0:
            out = (out + " " + str(int(v)) + one).strip()
    return (out + " " + f"{sec:7.3f}".strip() + " sec").strip()


def parameter_lines(mat, ion, opt, nmc, lib, material_name=None, target_name=None, dos_file="", cdf_file="",
                    when=None):
    """The lines print_parameters writes (Sorting_output_data.f90:42-337)
    for the run of material desc mat (prepared), ion and options: ion,
    material, time grid, models, options, and the atomic table with every
    shell's k- and f-sum (sumrules, with the plasma frequency of the shell's
    atoms).  Without the title block (print_title false); the thread line
    names the GPU instead of OpenMP threads.  when: the time stamp
    (datetime, default now)."""
    import ctypes
    import datetime
    from . import abi
    from .cache import ELEMENT_SYMBOL, fortran_e, fortran_es, shell_name
    from .configs import G_MP

    def f(x, w, d):
        return f"{x:{w}.{d}f}"

    def es(x, w, d):
        return fortran_es(float(x), w, d + 1)

    name = material_name or mat.name
    L = [_STARS]
    if ion.Zat > 0:
        L.append(f" Performing calculations for {ELEMENT_NAME[ion.Zat]} in {name}")
        L.append(f" Ion {ELEMENT_SYMBOL[ion.Zat]} (Element #{ion.Zat}, Mass {f(ion.Mass, 12, 3).strip()})")
        L.append(f" With energy {f(ion.E / 1.0e6, 9, 2).strip()} [MeV]")
    else:
        L.append(f" Performing calculations of the electron MFP in {name} only.")
    L.append(_DASHES)
    L.append(f" Material: {name} ({target_name or mat.name})")
    L.append(f" Material density: {f(mat.dens, 12, 3).strip()} [g/cm^3] or {fortran_e(mat.At_Dens, 16, 7).strip()}"
             " [1/cm^3]")
    L.append(f" Thickness of the analysed layer: {f(mat.layer, 12, 3).strip()} [A]")
    L.append(f" Temperature of the target: {f(mat.temp, 12, 3).strip()} [K]")
    if mat.el_eff_mass == 0:
        L.append(" Effective mass of valence electrons is calculated from DOS.")
    else:
        v = f(mat.el_eff_mass, 10, 2) if abs(mat.el_eff_mass) < 1e6 else fortran_e(mat.el_eff_mass, 16, 7)
        L.append(f" Effective mass of valence electrons: {v.strip()} [me]")
    if mat.hole_mass > 0:
        v = f(mat.hole_mass, 10, 2) if abs(mat.hole_mass) < 1e6 else fortran_e(mat.hole_mass, 16, 7)
        L.append(f" Effective mass of valence holes: {v.strip()} [me]")
    else:
        L.append(" Effective mass of valence holes is calculated from DOS")
    L.append(f" DOS file used: {dos_file}" if dos_file else " No DOS file was found, not even free-electron one!")
    L.append(f" CDF file used: {cdf_file}")
    L.append(_DASHES)
    if ion.Zat > 0:
        tim = f(opt.Tim, 12, 3).strip()
        L.append(f" Total time to be analysed: {tim} [fs]")
        if opt.dt_flag <= 0:
            L.append(f" with the timestep of {f(opt.dt, 12, 3).strip()} [fs], linear time-scale")
        else:
            # SO:141 prints the text last formatted, the total time, as the ratio
            L.append(f" starting with 0.01 [fs] increasing by {tim} in logarithmic time-scale")
    L.append(_DASHES)
    model = "Brandt-Kitagawa" if ion.Kind_ion == 1 else "point-like charge"
    if ion.Kind_Zeff == 4:
        # SO:157-158 prints the charge where the model's name would go
        z = f(ion.fixed_Zeff, 7, 3).strip()
        L.append(f" Ion fixed charge {z} [e] is used with {z} model for SHI.")
    else:
        formula = {1: "Bohr", 2: "Nikolaev-Dmitriev", 3: "Schiwietz-Grande"}.get(int(ion.Kind_Zeff), "Barkas")
        L.append(f" Ion equilibrium charge is used with {formula} formula, {model} model for SHI.")
    if max(sh.kocs for sh in mat.atoms[0].shells) == 1:
        dr = {3: "dispersion relation from Ritchie and Howie", 2: "plasmon-pole dispersion relation of scattering "
              "centers", 4: "Delta-function CDF"}.get(int(opt.kind_of_DR), "free electrons")
    else:
        dr = "BEB atomic cross sections"
    L.append(f" Inelastic scattering calculated with {dr}")
    if opt.plasmon_Emax:
        L.append(" Plasmon maximal energy is used as upper integration limit during cross-section calculation")
    L.append(" (using new integration energy grid)")
    el = {2: "calculated with DSF cross-sections", 1: "calculated with phonon CDF (Ritchie-Howie)",
          0: "calculated with Mott atomic cross-sections"}.get(int(opt.kind_of_EMFP), "is excluded")
    L.append(f" Elastic scattering {el}")
    if mat.cut_off > 0.0:
        L.append(f" Energy cut-off used is {f(mat.cut_off, 12, 3).strip()}")
    else:
        L.append(" No energy cut-off is used")
    if mat.work_function > 0:
        L.append(f" Electron emission included. Work function = {f(mat.work_function, 12, 3).strip()} [eV]")
        L.append(f" Potential barrier length = {f(mat.bar_length, 12, 3).strip()} [A] Barrier height = "
                 f"{f(mat.bar_height, 12, 3).strip()} [eV]")
    else:
        L.append(" Electron emission is excluded")
    L.append(" Radiative decays of deep-shell holes and photon transport are "
             + ("included" if opt.include_photons else "excluded"))
    L.append(" Transient electric fields are excluded")
    L.append(f" Number of MC iterations: {f'{int(nmc):5d}'.strip()}")
    L.append(" Transport on 1 GPU")
    # atomic parameters (SO:276-330)
    n_at_mol = sum(a.pers for a in mat.atoms)
    L.append(_DASHES)
    L.append(" The following atomic parameters of the target are used:")

    def sums(osc, Ip, Omega):
        if not osc:
            return 0.0, 0.0
        A = np.ascontiguousarray([o[1] for o in osc], dtype=np.float64)
        E0 = np.ascontiguousarray([o[0] for o in osc], dtype=np.float64)
        G = np.ascontiguousarray([o[2] for o in osc], dtype=np.float64)
        ks, fs = ctypes.c_double(), ctypes.c_double()
        lib.trk_sumrules(A.ctypes.data_as(abi.DP), E0.ctypes.data_as(abi.DP), G.ctypes.data_as(abi.DP), len(A), Ip,
                         Omega, ctypes.byref(ks), ctypes.byref(fs))
        return ks.value, fs.value

    for ia, atom in enumerate(mat.atoms):
        L.append(f"{ELEMENT_SYMBOL[atom.Z]} atom:")
        L.append(" Shell  Quantum_n  Ne    Ip[eV]  Ekin[eV]  t(Auger)[fs]  t(Rad)[fs]   k-sum    f-sum")
        for js, sh in enumerate(atom.shells):
            vb = ia == 0 and js == len(atom.shells) - 1
            dens = 1e6 * mat.At_Dens / n_at_mol if vb else 1e6 * mat.At_Dens * atom.pers / n_at_mol
            ks, fs = sums(sh.osc, sh.Ip, lib.trk_w_plasma(dens, 0.0))
            nm = ("Valence" if vb else shell_name(sh.shl_num))[:11].ljust(11)
            pqn = "   " if vb else f"{'KLMNOPQ'.index(nm[0]) + 1 if nm[0] in 'KLMNOPQ' else 0:3d}"
            L.append(nm + pqn + f(sh.Nel, 8, 2) + f(sh.Ip, 10, 2) + f(sh.Ek, 10, 2) + es(sh.auger, 12, 2)
                     + es(sh.radiat, 12, 2) + f(ks, 9, 2) + f(fs, 9, 2))
    mean_mass = sum(a.pers * a.mass for a in mat.atoms) * G_MP / n_at_mol
    ks, fs = sums([(e0, a, g) for e0, a, g in mat.phonon], 1.0e-8,
                  lib.trk_w_plasma(1e6 * mat.At_Dens / n_at_mol, mean_mass))
    if ks > 1.0e-2:
        L.append("Phonons" + "       " + f(n_at_mol, 8, 2) + f(0.0, 10, 2) + f(0.0, 10, 2) + es(0.0, 12, 2)
                 + es(0.0, 12, 2) + f(ks, 9, 2) + f(fs, 9, 2))
    else:
        L.append("Phonons" + "       " + f(n_at_mol, 8, 2) + f(0.0, 9, 2) + f(0.0, 9, 2) + es(0.0, 12, 2)
                 + es(0.0, 12, 2) + es(ks, 12, 2) + es(fs, 12, 2))
    L.append(_DASHES)
    t = when or datetime.datetime.now()
    L.append(f"Printed at {t.hour:02d}:{t.minute:02d}:{t.second:02d}:{t.microsecond // 1000:03d} on "
             f"{t.day:02d}/{t.month:02d}/{t.year:04d}")
    L.append(_STARS)
    return L


def write_parameters(folder, mat, ion, opt, nmc, lib, tallies=None, zeff=None, duration=None, **names):
    """'!Parameters.txt' as Save_output writes it (SO:489-506): the
    parameter lines, then the ion's equilibrium charge (zeff, when given),
    the energy loss Out_tot_E at the last node per layer thickness (from
    the run's unnormalised tallies and nmc) and the run's duration [s]."""
    import os
    os.makedirs(folder, exist_ok=True)
    L = parameter_lines(mat, ion, opt, nmc, lib, **names)
    if zeff is not None:
        L.append("Ion equilibrium charge:         " + f"{zeff:6.3f}".strip() + " [electron charge]")
    if tallies is not None:
        Se = float(tallies["Out_tot_E"][-1]) / float(nmc) / mat.layer
        L.append("MC calculated energy loss (Se): " + f"{Se:9.2f}".strip() + " [eV/A]")
    if duration is not None:
        L.append("Duration of calculation:        " + duration_text(float(duration)))
    L.append(_DASHES)
    with open(os.path.join(folder, "!Parameters.txt"), "w") as f:
        f.write("\n".join(L) + "\n")
    return "!Parameters.txt"
