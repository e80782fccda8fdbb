"""Material, ion and run descriptors for the benchmark configurations.

Input reading stays on the caller's side in the reference (Read_input_file,
Reading_files_and_parameters.f90:162-634); this module plays that role for
the Python host: it builds the flat trk_material from a readable description,
normalises CDF oscillator strengths with the f-sum rule the way
get_single_pole does (Cross_sections.f90:554-688), derives the VB density-of-
states arrays (reading_material_DOS, RF:2374-2450) through the library, and
builds the radial grid of Radius_for_distributions (Sorting_output_data.f90
:1363-1435) for a caller-chosen set of radii.

Every value a workload description leaves open (cut-off, time step,
oscillator widths, DOS shapes, decay times) is fixed in this file, next to
the workload that uses it.
"""
import ctypes
import math
import os
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from . import abi

# Universal_Constants.f90:28-60 (as folded by the library)
G_ME = 9.1093821545e-31
G_MP = 1836.1526724780 * G_ME
G_H = 1.05457162853e-34
G_E = 1.602176487e-19
G_PI = 3.1415926535897932384626433832795


@dataclass
class Shell:
    Ip: float
    Nel: float
    osc: List[Tuple[float, float, float]]       # (E0, A or weight, Gamma)
    auger: float = 1.0e23                       # [fs]
    radiat: float = 1.0e23                      # [fs]
    shl_num: int = 1
    normalize: bool = True                      # set A by the f-sum rule to Nel
    kocs: int = 1                               # electron/hole cross section: 1 CDF, 2 BEB (KOCS)
    Ek: float = 0.0                             # [eV] mean kinetic energy of the shell's electrons (BEB)


@dataclass
class Atom:
    Z: int
    mass: float                                 # [proton masses]
    pers: float
    shells: List[Shell]


@dataclass
class MaterialDesc:
    name: str
    atoms: List[Atom]
    dens: float                                 # [g/cm^3]
    layer: float                                # [A]
    vsound: float = 5000.0
    E_F: float = 10.0
    el_eff_mass: float = 1.0
    hole_mass: float = 1.0
    cut_off: float = 10.0
    temp: float = 0.0
    work_function: float = 0.0                  # [eV] > 0: emission through the surface Z = 0
    bar_length: float = 0.0                     # [A] surface barrier length
    bar_height: float = 0.0                     # [eV] surface barrier height
    phonon: List[Tuple[float, float, float]] = field(default_factory=list)
    phonon_single_pole: bool = False
    dos_raw_E: np.ndarray = None
    dos_raw: np.ndarray = None
    # (electron, hole) DSF tables (dsf.DsfTable) of DSF elastic scattering
    # (kind_of_EMFP = 2); handed to the library when a context is created
    dsf: tuple = None

    def __post_init__(self):
        self.n_atoms = len(self.atoms)
        self.n_shells = [len(a.shells) for a in self.atoms]
        sm = sum(a.mass * a.pers for a in self.atoms)
        sp = sum(a.pers for a in self.atoms)
        # At_Dens, RF:1335
        self.At_Dens = 1.0e-3 * self.dens / (G_MP * sm / sp)
        self.sum_pers = sp
        self.v_f = math.sqrt(2.0 * self.E_F / G_ME)
        vb = self.atoms[0].shells[-1]
        self.N_VB_el = vb.Nel
        self._lib = None
        self.dos = None

    # -------------------------------------------------------------- prep
    def prepare(self, lib):
        """Normalise oscillators and derive the DOS arrays (needs the library)."""
        self._lib = lib
        self._normalize(lib)
        self._prepare_dos(lib)
        return self

    def _normalize(self, lib):
        n_at_mol = self.sum_pers
        for ia, atom in enumerate(self.atoms):
            for js, sh in enumerate(atom.shells):
                if not sh.normalize:
                    continue
                A = np.array([o[1] for o in sh.osc], dtype=np.float64)
                E0 = np.array([o[0] for o in sh.osc], dtype=np.float64)
                G = np.array([o[2] for o in sh.osc], dtype=np.float64)
                is_vb = (ia == 0 and js == len(atom.shells) - 1)
                if is_vb:
                    target = sh.Nel / n_at_mol                       # CS:580
                    Omega = lib.trk_w_plasma(1e6 * self.At_Dens, 0.0)  # CS:590
                else:
                    target = sh.Nel                                  # CS:600
                    contrib = atom.pers / n_at_mol
                    Omega = lib.trk_w_plasma(1e6 * self.At_Dens * contrib, 0.0)  # CS:613
                ks = ctypes.c_double()
                fs = ctypes.c_double()
                rc = lib.trk_sumrules(A.ctypes.data_as(abi.DP), E0.ctypes.data_as(abi.DP), G.ctypes.data_as(abi.DP),
                                      len(A), sh.Ip, Omega, ctypes.byref(ks), ctypes.byref(fs))
                if rc:
                    raise ValueError("sumrules failed")
                A = A * (target / ks.value)
                sh.osc = [(float(e), float(a), float(g)) for e, a, g in zip(E0, A, G)]
                sh.normalize = False
        if self.phonon_single_pole:
            # get_single_pole phonon branch, CS:656-677
            qd = (6.0 * G_PI * G_PI * (self.At_Dens * 1e6)) ** 0.33333333
            Ed = G_H * self.vsound * qd / G_E
            Ee = Ed * (G_PI / 6.0) ** (1.0 / 3.0)
            E0 = 2.0 * Ee
            Gm = E0 * 0.5
            mean_mass = sum(a.pers * a.mass for a in self.atoms) * G_MP / n_at_mol
            Omega = lib.trk_w_plasma(1e6 * self.At_Dens / n_at_mol, mean_mass)
            A = np.array([1.0]); E = np.array([E0]); Gv = np.array([Gm])
            ks = ctypes.c_double(); fs = ctypes.c_double()
            lib.trk_sumrules(A.ctypes.data_as(abi.DP), E.ctypes.data_as(abi.DP), Gv.ctypes.data_as(abi.DP), 1, 1.0e-8,
                             Omega, ctypes.byref(ks), ctypes.byref(fs))
            self.phonon = [(E0, n_at_mol / ks.value, Gm)]
            self.phonon_single_pole = False

    def _prepare_dos(self, lib):
        rE = np.ascontiguousarray(self.dos_raw_E, dtype=np.float64)
        rD = np.ascontiguousarray(self.dos_raw, dtype=np.float64)
        M = ctypes.c_int32()
        if lib.trk_dos_size(rE.ctypes.data_as(abi.DP), len(rE), ctypes.byref(M)):
            raise ValueError("bad DOS grid")
        M = M.value
        arrs = {k: np.zeros(M) for k in ["E", "DOS", "int_DOS", "k", "Eff_m", "k_inv", "Eff_m_inv"]}
        rc = lib.trk_prepare_dos(rE.ctypes.data_as(abi.DP), rD.ctypes.data_as(abi.DP), len(rE), self.N_VB_el,
                                 self.At_Dens, self.sum_pers,
                                 *[arrs[k].ctypes.data_as(abi.DP) for k in ["E", "DOS", "int_DOS", "k", "Eff_m",
                                                                             "k_inv", "Eff_m_inv"]], M)
        if rc:
            raise ValueError("trk_prepare_dos failed")
        self.dos = arrs
        self.n_dos = M

    # -------------------------------------------------------------- struct
    def to_struct(self):
        if self.dos is None:
            raise RuntimeError("MaterialDesc.prepare(lib) must run first")
        m = abi.Material()
        m.n_atoms = self.n_atoms
        s = 0
        o = 0
        for ia, a in enumerate(self.atoms):
            m.Zat[ia] = a.Z
            m.Mass[ia] = a.mass
            m.Pers[ia] = a.pers
            m.n_shells[ia] = len(a.shells)
            for js, sh in enumerate(a.shells):
                m.Ip[s] = max(sh.Ip, 0.1)       # RF:1554
                m.Nel[s] = sh.Nel
                m.Auger[s] = sh.auger
                m.Radiat[s] = sh.radiat
                m.Shl_num[s] = 63 if (ia == 0 and js == len(a.shells) - 1) else sh.shl_num
                m.kocs[s] = sh.kocs
                m.Ek[s] = sh.Ek
                m.n_osc[s] = len(sh.osc)
                for (E0, A, G) in sh.osc:
                    m.osc_E0[o], m.osc_A[o], m.osc_Gamma[o] = E0, A, G
                    o += 1
                s += 1
        m.n_phonon = len(self.phonon)
        for i, (E0, A, G) in enumerate(self.phonon):
            m.ph_E0[i], m.ph_A[i], m.ph_Gamma[i] = E0, A, G
        m.Dens = self.dens
        m.At_Dens = self.At_Dens
        m.Vsound = self.vsound
        m.Layer = self.layer
        m.N_VB_el = self.N_VB_el
        m.El_eff_mass = self.el_eff_mass
        m.hole_mass = self.hole_mass
        m.E_F = self.E_F
        m.v_f = self.v_f
        m.cut_off = self.cut_off
        m.temp = self.temp
        m.work_function = self.work_function
        m.bar_length = self.bar_length
        m.bar_height = self.bar_height
        m.n_dos = self.n_dos
        for k, f in [("E", "dos_E"), ("DOS", "dos_DOS"), ("int_DOS", "dos_int_DOS"), ("k", "dos_k"),
                     ("Eff_m", "dos_Eff_m"), ("k_inv", "dos_k_inv"), ("Eff_m_inv", "dos_Eff_m_inv")]:
            setattr(m, f, self.dos[k].ctypes.data_as(abi.DP))
        self._struct = m
        return m


def radial_grid(NR, Rmin, Rmax, layer):
    """Log-spaced cylinder radii Out_R [A] and inverse volumes Out_V [1/A^3]
    (Out_V(j) = 1/((R_j^2 - R_{j-1}^2) pi L), Sorting_output_data.f90:1429)."""
    R = np.array([10.0 ** (math.log10(Rmin) + (math.log10(Rmax) - math.log10(Rmin)) * j / (NR - 1)) for j in range(NR)])
    V = np.zeros(NR)
    R0 = 0.0
    for j in range(NR):
        V[j] = 1.0 / ((R[j] * R[j] - R0 * R0) * G_PI * layer)
        R0 = R[j]
    return R, V


def reference_radial_grid(layer):
    """The reference's own 50 cylinder radii (Radius_for_distributions,
    Sorting_output_data.f90:1363-1435): 1..10 A in steps of 1, then per
    decade a first radius of 1.5 and steps of 1 times the decade (15, 20..100,
    150, 200..1000, ..., 15000, 20000..100000 A), and Out_V as radial_grid
    computes it."""
    R = []
    for i in range(1, 51):
        if i < 11:
            r = i
        elif i < 41:
            d = 10 ** ((i - 1) // 10)            # 10, 100, 1000
            k = i - 10 * ((i - 1) // 10)
            r = 15 * d // 10 if k == 1 else k * d
        else:
            r = 15000 if i == 41 else (i - 40) * 10000
        R.append(float(r))
    R = np.array(R)
    V = np.zeros(len(R))
    R0 = 0.0
    for j in range(len(R)):
        V[j] = 1.0 / ((R[j] * R[j] - R0 * R0) * G_PI * layer)
        R0 = R[j]
    return R, V


# -------------------------------------------------------------- DOS shapes
def dos_insulator(width=22.0, step=0.05):
    """Synthetic two-band oxide VB: an s-like band (Gaussian at -18 eV) under a
    p-like band (half ellipse over the top 9 eV).  Energies relative to the VB top."""
    E = np.arange(-width, 0.0 + step / 2, step)
    E[-1] = 0.0
    p = np.where(E > -9.0, np.sqrt(np.clip(1.0 - ((E + 4.5) / 4.5) ** 2, 0.0, None)), 0.0)
    s = 0.35 * np.exp(-0.5 * ((E + 18.0) / 1.2) ** 2)
    return E, p + s


def dos_free_electron(E_F=11.17, step=0.05):
    """Free-electron VB, DOS ~ sqrt(E - E_bottom), top at E_F."""
    E = np.arange(-E_F, 0.0 + step / 2, step)
    E[-1] = 0.0
    return E, np.sqrt(np.clip(E + E_F, 0.0, None))


# -------------------------------------------------------------- configurations
def _al2o3_like(cut_off=10.0, layer=100.0, decay_seed=0x5EED0002, photons=True):
    """Al2O3-like target of the 'insulator with hole kinetics' configuration."""
    rng = np.random.default_rng(decay_seed)
    br_L, br_K = rng.uniform(0.0, 1.0, size=2)
    tau_L, tau_K = 20.6, 3.98          # total decay times [fs] (INPUT_CDF/Al2O3.cdf values)
    def split(tau, br):
        # total rate 1/tau split into radiative fraction br and Auger 1-br
        return tau / (1.0 - br), (tau / br if photons else 1.0e23)
    aL, rL = split(tau_L, br_L)
    aK, rK = split(tau_K, br_K)
    al = Atom(13, 26.98, 2.0, [
        Shell(80.0, 8.0, [(90.0, 1.0, 40.0), (130.0, 0.6, 60.0)], auger=aL, radiat=rL, shl_num=2),
        Shell(8.8, 24.0, [(14.0, 0.15, 3.0), (24.0, 0.35, 8.0), (33.0, 0.35, 12.0), (55.0, 0.15, 20.0)]),
    ])
    o = Atom(8, 16.00, 3.0, [
        Shell(538.0, 2.0, [(560.0, 1.0, 120.0), (640.0, 0.5, 200.0)], auger=aK, radiat=rK, shl_num=1),
    ])
    E, D = dos_insulator()
    return MaterialDesc("Al2O3-like", [al, o], dens=3.95, layer=layer, vsound=6510.0, E_F=4.4,
                        el_eff_mass=1.0, hole_mass=1.0, cut_off=cut_off, dos_raw_E=E, dos_raw=D)


def _al_like(layer=100.0):
    """Al-like metal of the small CPU-runnable configuration."""
    al = Atom(13, 26.98, 1.0, [
        Shell(118.0, 2.0, [(130.0, 1.0, 60.0)], auger=0.47, shl_num=3),
        Shell(73.0, 6.0, [(80.0, 1.0, 30.0), (110.0, 1.0, 60.0)], auger=11.5, shl_num=5),
        Shell(0.1, 3.0, [(15.0, 1.0, 0.6)]),
    ])
    E, D = dos_free_electron(11.17)
    return MaterialDesc("Al-like", [al], dens=2.70, layer=layer, vsound=5000.0, E_F=11.17,
                        el_eff_mass=0.0, hole_mass=-1.0, cut_off=10.0, phonon_single_pole=True,
                        dos_raw_E=E, dos_raw=D)


def _sio2_like(layer=1.0e4, decay_seed=0x5EED0005):
    """SiO2-like target of the 'large single track' configuration: density
    2.32 g/cm^3, 9 eV gap, five valence oscillators, Si K / L1 / L2,3 and O K
    shells, three phonon oscillators with E0 in 0.06-0.15 eV (quasi-elastic
    scattering on the phonon CDF).  Oscillator positions, widths, ionisation
    potentials and Auger times are INPUT_CDF/SiO2.cdf's (the valence set
    without its weakest oscillator); the radiative times are drawn like the
    Al2O3-like target's, from a branching ratio per core shell."""
    rng = np.random.default_rng(decay_seed)
    br = rng.uniform(0.0, 1.0, size=4)
    def split(tau, b):
        return tau / (1.0 - b), tau / b
    aK, rK = split(1.6, br[0])
    aL1, rL1 = split(16.0, br[1])
    aL23, rL23 = split(16.0, br[2])
    aOK, rOK = split(8.0, br[3])
    si = Atom(14, 28.09, 1.0, [
        Shell(1844.1, 2.0, [(2060.0, 1.0, 946.5)], auger=aK, radiat=rK, shl_num=1),
        Shell(151.6, 2.0, [(232.0, 1.0, 124.4)], auger=aL1, radiat=rL1, shl_num=3),
        Shell(108.0, 6.0, [(149.75, 0.94, 120.0), (120.52, 0.06, 16.52)], auger=aL23, radiat=rL23, shl_num=4),
        Shell(9.0, 16.0, [(41.385, 0.26389, 0.1856), (23.63, 262.65, 11.79), (14.5927, 5.783, 3.0),
                          (53.904, 177.4, 32.89), (34.064, 55.788, 14.2)]),
    ])
    o = Atom(8, 16.00, 2.0, [
        Shell(538.25, 2.0, [(481.0, 1.0, 406.5)], auger=aOK, radiat=rOK, shl_num=1),
    ])
    E, D = dos_insulator()
    return MaterialDesc("SiO2-like", [si, o], dens=2.32, layer=layer, vsound=5170.0, E_F=10.0,
                        el_eff_mass=1.0, hole_mass=1.0, cut_off=10.0,
                        phonon=[(0.06, 0.0005, 0.009), (0.10, 0.0015, 0.013), (0.15, 0.0004, 0.009)],
                        dos_raw_E=E, dos_raw=D)


def config_material(name):
    """(MaterialDesc, Ion, Options, run dict) for a named configuration.

    config1  Al-like metal, Xe 1 MeV/u, 100 iterations, 0-100 fs (dt 0.1)
    config2  Al2O3-like, Au 11 MeV/u, 10^4 iterations, 0-1000 fs (dt 1)
    config5  SiO2-like with phonon quasi-elastic scattering and photons, U 20
             MeV/u over a 1 um track, 10^6 iterations, 0-1000 fs, 200 radial bins;
             its synthetic photon attenuation tables replace the built photon
             tables after the build (apply_photon_tables with
             synthetic_photoabsorption)
    test     config2's target with a reduced electron grid (2 keV), O ion at
             0.5 MeV/u, 0-100 fs: small enough for the CPU oracle
    test_sio2  config5 on a 100 A track segment, 2 iterations (the full
             tables; a cascade the CPU oracle replays in seconds)
    test_polymer  a material file whose hydrogen lists no shell
             (examples/INPUT_CDF/c2h4_like.cdf), O ion at 0.5 MeV/u, 0-100 fs, 4 iterations
    test_cdf  a material read from a full-CDF material file
             (examples/INPUT_CDF/al2o3_like.cdf), O ion at 0.5 MeV/u, 0-100 fs, phonon
             CDF elastic scattering and photons, 4 iterations
    """
    opt = abi.Options()
    ion = abi.Ion()
    if name == "config1":
        mat = _al_like()
        ion.Zat, ion.Mass, ion.E = 54, 131.29, 131.29e6
        opt.kind_of_EMFP, opt.CDF_elast_Zeff = 1, 1
        opt.Tim, opt.dt, opt.dt_flag = 100.0, 0.1, 0
        run = dict(nmc=100, NR=100, Rmin=1.0, Rmax=1000.0, seed=0x5EED0001)
    elif name == "config2":
        mat = _al2o3_like()
        ion.Zat, ion.Mass, ion.E = 79, 196.97, 11.0e6 * 196.97
        opt.kind_of_EMFP, opt.CDF_elast_Zeff = 0, 0
        opt.Tim, opt.dt, opt.dt_flag = 1000.0, 1.0, 0
        opt.include_photons = 1
        run = dict(nmc=10000, NR=100, Rmin=1.0, Rmax=1000.0, seed=0x5EED0002)
    elif name in ("config5", "test_sio2"):
        mat = _sio2_like(layer=1.0e4 if name == "config5" else 100.0)
        ion.Zat, ion.Mass, ion.E = 92, 238.03, 20.0e6 * 238.03
        opt.kind_of_EMFP, opt.CDF_elast_Zeff = 1, 1
        opt.Tim, opt.dt, opt.dt_flag = 1000.0, 1.0, 0
        opt.include_photons = 1
        run = dict(nmc=10 ** 6 if name == "config5" else 2, NR=200, Rmin=1.0, Rmax=1.0e4, seed=0x5EED0005)
    elif name == "test":
        mat = _al2o3_like()
        ion.Zat, ion.Mass, ion.E = 8, 16.00, 0.5e6 * 16.0
        opt.kind_of_EMFP, opt.CDF_elast_Zeff = 0, 0
        opt.Tim, opt.dt, opt.dt_flag = 100.0, 1.0, 0
        opt.include_photons = 1
        opt.e_grid_Emax = 2000.0
        run = dict(nmc=4, NR=40, Rmin=1.0, Rmax=1000.0, seed=0x5EED0003)
    elif name == "test_metal":
        mat = _al_like(layer=50.0)
        ion.Zat, ion.Mass, ion.E = 8, 16.00, 0.5e6 * 16.0
        opt.kind_of_EMFP, opt.CDF_elast_Zeff = 1, 1
        opt.Tim, opt.dt, opt.dt_flag = 50.0, 0.5, 0
        opt.e_grid_Emax = 2000.0
        run = dict(nmc=2, NR=40, Rmin=1.0, Rmax=1000.0, seed=0x5EED0005)
    elif name == "test_cdf":
        # the material read from a file in the reference's full-CDF format
        # (examples/INPUT_CDF/al2o3_like.cdf: the Al2O3-like target's shells with the
        # oscillator strengths written out, two phonon peaks)
        from .inputs import read_cdf
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples", "INPUT_CDF", "al2o3_like.cdf")
        mat = read_cdf(path, layer=100.0, dos=dos_insulator(), hole_mass=1.0, el_eff_mass=1.0)
        ion.Zat, ion.Mass, ion.E = 8, 16.00, 0.5e6 * 16.0
        opt.kind_of_EMFP, opt.CDF_elast_Zeff = 1, 1
        opt.Tim, opt.dt, opt.dt_flag = 100.0, 1.0, 0
        opt.include_photons = 1
        opt.e_grid_Emax = 2000.0
        run = dict(nmc=4, NR=40, Rmin=1.0, Rmax=1000.0, seed=0x5EED0006)
    elif name == "test_polymer":
        # a material whose second element lists no shell (examples/
        # INPUT_CDF/c2h4_like.cdf: hydrogen's electrons are in the valence band), Mott
        # elastic scattering, no phonon block (single-pole phonon CDF)
        from .inputs import read_cdf
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples", "INPUT_CDF", "c2h4_like.cdf")
        mat = read_cdf(path, layer=100.0, dos=dos_insulator(), hole_mass=1.0, el_eff_mass=1.0)
        ion.Zat, ion.Mass, ion.E = 8, 16.00, 0.5e6 * 16.0
        opt.kind_of_EMFP, opt.CDF_elast_Zeff = 0, 0
        opt.Tim, opt.dt, opt.dt_flag = 100.0, 1.0, 0
        opt.e_grid_Emax = 2000.0
        run = dict(nmc=4, NR=40, Rmin=1.0, Rmax=1000.0, seed=0x5EED0007)
    elif name in OPTION_WORKLOADS:
        # the small workloads with the reference's option branches that the
        # configurations above leave at their defaults switched on
        base, dr, kz, fz, temp, plasmon, dt_flag, dt, *rest = OPTION_WORKLOADS[name]
        mat, ion, opt, run = config_material(base)
        mat.temp = temp
        ion.Kind_Zeff, ion.fixed_Zeff = kz, fz
        ion.Kind_ion = rest[0] if rest else 0
        if len(rest) > 1 and rest[1]:
            mat.work_function, mat.bar_length, mat.bar_height = rest[1]
        if len(rest) > 2:
            shells = [sh for a in mat.atoms for sh in a.shells]
            for s, Ek in rest[2].items():
                shells[s].kocs, shells[s].Ek = 2, Ek
        if len(rest) > 3 and rest[3]:
            from .dsf import synthetic_dsf
            opt.kind_of_EMFP, opt.CDF_elast_Zeff = 2, 0
            mat.dsf = synthetic_dsf(temp)
        if dr == 4:
            # electrons slowed below the peaks scatter elastically for the
            # rest of the window: ten times the events per iteration
            run = dict(run, nmc=2)
        opt.kind_of_DR, opt.plasmon_Emax = dr, plasmon
        opt.dt_flag, opt.dt = dt_flag, dt
        return mat, ion, opt, run
    else:
        raise KeyError(name)
    ion.Kind_Zeff, ion.fixed_Zeff, ion.Kind_ion = 0, float(ion.Zat), 0
    opt.kind_of_DR = 1
    return mat, ion, opt, run


# Option branches (Flag / SHI fields, Objects.f90:44-53, 135-180) on the two
# small workloads: (base workload, kind_of_DR, Kind_Zeff, fixed Zeff, target
# temperature [K], plasmon_Emax, dt_flag, dt).
#   opt_plasmon_pole  plasmon-pole dispersion (CS:379-381), Bohr ion charge,
#                     plasmon integration limit (CS:930-936), logarithmic time
#                     grid of ratio 2 (MC:2132-2144)
#   opt_ritchie_warm  Ritchie's extended dispersion (CS:382-385),
#                     Nikolaev-Dmitriev charge, a 300 K target (Bose factor
#                     of the inelastic cross sections, CS:2274-2278)
#   opt_sg_warm       Schiwietz-Grande charge, 300 K
#   opt_fixed_zeff    a fixed ion charge of 3.5, a 300 K metal (Bose factor of
#                     the phonon CDF, CS:3290-3295), logarithmic time grid of
#                     ratio 10 (the reference's own INPUT_PARAMETERS.txt:6 setting)
#   opt_bk_ion        a Brandt-Kitagawa ion (Kind_ion = 1, a 9th entry): its
#                     mean free paths with the charge form factor inside the q
#                     integral (SHI_TotIMFP_BK, CS:2748-2889), Bohr charge
#   opt_beb           the BEB cross section (KOCS = 2) for the electrons and
#                     holes on the two core shells (an 11th entry: flat shell
#                     index -> mean kinetic energy of its electrons [eV]); the
#                     ion keeps their CDF
#   opt_delta_cdf     the Delta-CDF model of the electron and hole inelastic
#                     cross sections (kind_of_DR = 4, CS:952-956, 1894-1896)
#   opt_dsf           DSF elastic scattering (kind_of_EMFP = 2, a 12th entry
#                     True) on the synthetic DSF tables of dsf.synthetic_dsf,
#                     300 K
#   opt_combo         the branches together: Delta-CDF, a BEB shell, DSF
#                     elastic scattering, emission, a Brandt-Kitagawa ion with
#                     Bohr charge, 300 K, logarithmic time grid of ratio 2
#   opt_emission      emission through the surface Z = 0 (a 10th entry: work
#                     function [eV], barrier length [A], barrier height [eV];
#                     calculate_emission, MC:2477-2513): electrons above 15 eV
#                     leave, slower ones through the barrier's transmission or
#                     are turned back; the ion's track starts at the surface
OPTION_WORKLOADS = {
    "opt_plasmon_pole": ("test", 2, 1, 8.0, 0.0, 1, 1, 2.0),
    "opt_ritchie_warm": ("test", 3, 2, 8.0, 300.0, 0, 0, 1.0),
    "opt_sg_warm": ("test", 1, 3, 8.0, 300.0, 0, 0, 1.0),
    "opt_fixed_zeff": ("test_metal", 1, 4, 3.5, 300.0, 0, 1, 10.0),
    "opt_bk_ion": ("test", 1, 1, 8.0, 0.0, 0, 0, 1.0, 1),
    "opt_emission": ("test", 1, 0, 8.0, 0.0, 0, 0, 1.0, 0, (4.5, 10.0, 10.0)),
    "opt_delta_cdf": ("test", 4, 0, 8.0, 0.0, 0, 0, 1.0),
    "opt_beb": ("test", 1, 0, 8.0, 0.0, 0, 0, 1.0, 0, None, {0: 105.0, 2: 700.0}),
    "opt_dsf": ("test", 1, 0, 8.0, 300.0, 0, 0, 1.0, 0, None, {}, True),
    "opt_combo": ("test", 4, 1, 8.0, 300.0, 0, 1, 2.0, 1, (4.5, 10.0, 10.0), {0: 105.0}, True),
}


# -------------------------------------------------------------- the batch of impacts
CONFIG4_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples", "config4_impacts.json")


def generate_config4(n_impacts=256, n_materials=4, seed=0x5EED0004):
    """The batching workload: n_impacts ion impacts (Z uniform in 18-92,
    energies log-uniform in 1-20 MeV/u) on n_materials synthetic targets
    (one element of Z 6-16 each, density uniform in 1-10 g/cm^3, band gap
    uniform in 0-9 eV, 1-6 Drude-Lorentz valence oscillators with E0 in
    8-40 eV and widths in 2-15 eV, a K shell at the Moseley energy
    13.6 (Z-1)^2 eV with one oscillator and an Auger time of 2-10 fs).
    Drawn with numpy's Philox generator from `seed`; the result is what
    trekis_gpu/examples/config4_impacts.json holds."""
    from .inputs import ELEMENT_MASS
    g = np.random.Generator(np.random.Philox(seed))
    mats = []
    for m in range(n_materials):
        Zt = int(g.integers(6, 17))
        nosc = int(g.integers(1, 7))
        E0 = np.sort(g.uniform(8.0, 40.0, nosc))
        mats.append(dict(Z=Zt, mass=float(ELEMENT_MASS[Zt]), dens=float(g.uniform(1.0, 10.0)),
                         gap=float(g.uniform(0.0, 9.0)),
                         vb_osc=[(float(e), float(w), float(gm)) for e, w, gm in
                                 zip(E0, g.uniform(0.2, 1.0, nosc), g.uniform(2.0, 15.0, nosc))],
                         ip_k=float(round(13.6 * (Zt - 1) ** 2, 1)), auger_k=float(g.uniform(2.0, 10.0))))
    impacts = []
    for i in range(n_impacts):
        Z = int(g.integers(18, 93))
        Eu = float(np.exp(g.uniform(np.log(1.0), np.log(20.0))))
        impacts.append(dict(Z=Z, mass=float(ELEMENT_MASS[Z]), MeV_per_u=Eu, material=int(g.integers(0, n_materials))))
    return dict(seed=seed, materials=mats, impacts=impacts)


def config4_material(md, layer=100.0):
    """MaterialDesc of one synthetic target of the batching workload."""
    Zt, gap = md["Z"], max(md["gap"], 0.1)
    vb = Shell(gap, float(Zt - 2), [tuple(o) for o in md["vb_osc"]])
    k = Shell(md["ip_k"], 2.0, [(1.15 * md["ip_k"], 1.0, 0.6 * md["ip_k"])], auger=md["auger_k"], shl_num=1)
    atom = Atom(Zt, md["mass"], 1.0, [k, vb])
    E, D = dos_free_electron(10.0) if gap < 0.2 else dos_insulator()
    return MaterialDesc(f"synthetic Z={Zt}", [atom], dens=md["dens"], layer=layer, vsound=5000.0, E_F=10.0,
                        el_eff_mass=1.0, hole_mass=1.0, cut_off=10.0, dos_raw_E=E, dos_raw=D)


def config4_batch(n_impacts=None, precision=1):
    """(materials, ions, options, run) of the batching workload read from
    its fixture: MaterialDesc and Options per material (Mott elastic
    scattering, photons off, 0-1000 fs in 1 fs steps), one Ion per impact
    with the index of its material, and the radial grid (100 log bins,
    1-1000 A over a 100 A segment)."""
    import json
    with open(CONFIG4_FIXTURE) as fh:
        fx = json.load(fh)
    imps = fx["impacts"][: n_impacts] if n_impacts else fx["impacts"]
    mats, opts = [], []
    for md in fx["materials"]:
        mats.append(config4_material(md))
        opt = abi.Options()
        opt.kind_of_EMFP, opt.CDF_elast_Zeff, opt.kind_of_DR = 0, 0, 1
        opt.Tim, opt.dt, opt.dt_flag = 1000.0, 1.0, 0
        opt.include_photons = 0
        opt.precision = precision
        opts.append(opt)
    ions = []
    for im in imps:
        ion = abi.Ion()
        ion.Zat, ion.Mass, ion.E = im["Z"], im["mass"], im["MeV_per_u"] * 1.0e6 * im["mass"]
        ion.Kind_Zeff, ion.fixed_Zeff, ion.Kind_ion = 0, float(im["Z"]), 0
        ions.append((ion, im["material"]))
    run = dict(NR=100, Rmin=1.0, Rmax=1000.0, seed=0x5EED0004)
    return mats, opts, ions, run


# -------------------------------------------------------------- photon attenuation tables
def photon_paths_from_cross_sections(desc, E, sigma):
    """Per-shell photon mean free paths [A] on the photon grid E from
    per-shell photoabsorption cross sections [A^2 per atom], as the EPDL
    branch of Tot_Phot_IMFP converts them (Cross_sections.f90:855-877): a
    core shell of atom a, L = 1 / (n x_a sigma) with n the atomic density
    [1/A^3] (At_Dens 1e-24) and x_a = Pers_a / sum Pers; the valence band
    (the last shell of the first atom), L = 1 / (n sigma_VB) with sigma_VB
    the stoichiometry-weighted cross section of the shells that form it,
    given as it is.  sigma: callable (flat shell index, E array) -> array;
    a zero cross section gives 1e30 (no absorption)."""
    n = desc.At_Dens * 1.0e-24
    xs = np.array([a.pers for a in desc.atoms]) / desc.sum_pers
    out = []
    s = 0
    vb = len(desc.atoms[0].shells) - 1
    for ia, atom in enumerate(desc.atoms):
        for js in range(len(atom.shells)):
            sig = np.asarray(sigma(s, np.asarray(E)), dtype=float)
            w = n if (ia == 0 and js == vb) else n * xs[ia]
            with np.errstate(divide="ignore"):
                L = np.where(sig > 0.0, 1.0 / (w * sig), 1.0e30)
            out.append(L)
            s += 1
    return np.array(out)


def synthetic_photoabsorption(desc, sigma_edge=2.0e-3):
    """Synthetic photoabsorption cross sections of the 'large single track'
    configuration (no EPDL data here): per shell, zero below its ionisation
    potential and sigma_edge x (Nel / 2) x (Ip / E)^3 above it [A^2], the
    E^-3 fall of photoabsorption above an edge."""
    shells = [sh for a in desc.atoms for sh in a.shells]

    def sigma(s, E):
        sh = shells[s]
        return np.where(E >= sh.Ip, sigma_edge * (sh.Nel / 2.0) * (sh.Ip / np.maximum(E, 1e-30)) ** 3, 0.0)
    return sigma


def apply_photon_tables(ctx, sigma):
    """Replace a built context's photon tables (trk_table_set, kind PHOT) by
    those of the cross sections sigma and re-derive its fused totals."""
    E, _, _ = ctx.table(abi.TAB_PHOT, 0)
    L = photon_paths_from_cross_sections(ctx.desc, E, sigma)
    for s in range(L.shape[0]):
        ctx.set_table(abi.TAB_PHOT, s, L=L[s])
    ctx.finish_tables()
    return E, L
