"""Readers of the reference's material input files: the full-CDF material
file (INPUT_CDF/*.cdf, read by reading_material_parameters,
Reading_files_and_parameters.f90:1185-1644) and the valence-band DOS file
(INPUT_DOS/*.dos, reading_material_DOS, Reading_files_and_parameters.f90
:2340-2450).  They build the same MaterialDesc the workload descriptions in
configs.py build, so a material the reference reads from its input folder
runs here unchanged.

Supported is the file form the particle-history path needs: every shell
with its own CDF oscillators (kind_of_CDF = 0); a negative designator gives
the shell's electrons and holes the BEB cross section (KOCS = 2, RF:1556-1561),
whose mean kinetic energy of the shell's electrons the caller passes (the
reference reads it from EADL).  Forms that need the EADL database or cross
sections this library does not compute are rejected with the reason: a
chemical formula in place of the element list (Decompose_compound), the
single-pole file with VALENCE / PHONON markers (check_atomic_parameters for
the shells), shells without oscillators (the ion's BEB cross section, which
the reference marks as not working, MC:1772).
"""
import math
import re
from typing import List, Optional, Tuple

import numpy as np

from .configs import Atom, MaterialDesc, Shell

# Element masses [proton masses] by atomic number, as the reference's
# Find_element_name assigns them (Dealing_with_EADL.f90:1217-1703); the
# atomic density At_Dens is computed from these, so they are kept identical.
ELEMENT_MASS = {
    1: 1.0082, 2: 4.002602, 3: 6.942, 4: 9.012182, 5: 10.812, 6: 12.0112, 7: 14.0072, 8: 15.9992,
    9: 18.9984032, 10: 20.1797, 11: 22.98976928, 12: 24.3059, 13: 26.9815386, 14: 28.0854, 15: 30.973762, 16: 32.062,
    17: 35.452, 18: 39.948, 19: 39.0983, 20: 40.078, 21: 44.955912, 22: 47.867, 23: 50.9415, 24: 51.9961,
    25: 54.938045, 26: 55.845, 27: 58.933195, 28: 58.6934, 29: 63.546, 30: 65.38, 31: 69.723, 32: 72.63,
    33: 74.9216, 34: 78.96, 35: 79.9049, 36: 83.798, 37: 85.4678, 38: 87.62, 39: 88.90585, 40: 91.224,
    41: 92.90638, 42: 95.96, 43: 98.0, 44: 101.07, 45: 102.9055, 46: 106.42, 47: 107.8682, 48: 112.411,
    49: 114.818, 50: 118.71, 51: 121.76, 52: 127.6, 53: 126.90447, 54: 131.293, 55: 132.9054519, 56: 137.327,
    57: 138.90547, 58: 140.116, 59: 140.90765, 60: 144.242, 61: 145.0, 62: 150.36, 63: 151.964, 64: 157.25,
    65: 158.92535, 66: 162.5, 67: 164.93032, 68: 167.259, 69: 168.93421, 70: 173.054, 71: 174.9668, 72: 178.49,
    73: 180.94788, 74: 183.84, 75: 186.207, 76: 190.23, 77: 192.217, 78: 195.084, 79: 196.966569, 80: 200.592,
    81: 204.389, 82: 207.2, 83: 208.9804, 84: 209.0, 85: 210.0, 86: 222.0, 87: 223.0, 88: 226.0,
    89: 227.0, 90: 232.03806, 91: 231.03588, 92: 238.02891, 93: 237.0, 94: 244.0, 95: 243.0, 96: 247.0,
    97: 247.0, 98: 251.0, 99: 252.0, 100: 257.0, 101: 258.0, 102: 259.0, 103: 262.0, 104: 267.0,
    105: 268.0, 106: 269.0, 107: 270.0, 108: 269.0, 109: 278.0, 110: 281.0, 111: 281.0, 112: 285.0,
    113: 286.0, 114: 289.0, 115: 288.0, 116: 293.0, 117: 294.0, 118: 294.0,
}


def _fortran_real(tok: str) -> float:
    """One real of Fortran list-directed input (d/D/q exponents allowed)."""
    t = tok.strip().lower().replace("d", "e").replace("q", "e")
    if not re.fullmatch(r"[+-]?(\d+\.?\d*|\.\d+)(e[+-]?\d+)?", t):
        raise ValueError(f"not a real: {tok!r}")
    return float(t)


def _fortran_int(tok: str) -> int:
    t = tok.strip()
    if not re.fullmatch(r"[+-]?\d+", t):
        raise ValueError(f"not an integer: {tok!r}")
    return int(t)


def _items(line: str) -> List[str]:
    """Values of one record of list-directed input: separated by blanks,
    tabs or commas; a slash ends the record.  A '!' is not a comment there:
    it is read as a value (and fails as a number)."""
    line = line.split("/", 1)[0]
    return [t for t in re.split(r"[\s,]+", line.strip()) if t]


def _read(line: str, kinds: str):
    """Read len(kinds) values ('i' integer, 'r' real) from the start of a
    record, as READ(unit, *) does; the rest of the record is skipped."""
    it = _items(line)
    if len(it) < len(kinds):
        raise ValueError(f"record holds {len(it)} values, {len(kinds)} needed: {line!r}")
    return [(_fortran_int if k == "i" else _fortran_real)(t) for k, t in zip(kinds, it)]


class CDFFile:
    """Contents of a full-CDF material file, in the reference's terms."""

    def __init__(self):
        self.name = ""
        self.atoms: List[dict] = []        # Zat, Pers, shells: [dict(ncdf, shl_num, Ip, Nel, auger, osc)]
        self.dens = self.vsound = self.E_F = 0.0
        self.egap = 0.1
        self.phonon: Optional[List[Tuple[float, float, float]]] = None


def parse_cdf(text: str) -> CDFFile:
    """Parse a full-CDF material file (reading_material_parameters, RF:1255-
    1614, the kind_of_CDF = 0 branch).  A file the reference's reader stops
    on (too few or unreadable values) raises ValueError naming the line."""
    try:
        return _parse_cdf(text)
    except _Unreadable as e:
        raise ValueError(f"unreadable material file: {e}") from None


class _Unreadable(Exception):
    pass


def _parse_cdf(text: str) -> CDFFile:
    lines = text.splitlines()
    pos = 0

    def record():
        nonlocal pos
        if pos >= len(lines):
            raise EOFError("end of file")
        pos += 1
        return lines[pos - 1]

    def read(kinds):
        """READ(unit, *): values are taken from as many records as needed
        (blank records hold none); the rest of the last one is skipped."""
        start = pos + 1
        items = []
        while len(items) < len(kinds):
            items += _items(record())
        try:
            return [(_fortran_int if k == "i" else _fortran_real)(t) for k, t in zip(kinds, items)]
        except ValueError as e:
            raise _Unreadable(f"line {start}: {e}") from None

    f = CDFFile()
    # name: up to a '!' comment or a TAB (RF:1259-1276)
    name = record()
    for sep in ("!", "\t"):
        i = name.find(sep)
        if i > 0:
            name = name[:i]
        elif i == 0 and sep == "\t":
            name = name[1:]
    f.name = name.strip()
    rec = record()
    try:
        (n_el,) = _read(rec, "i")
    except ValueError:
        raise ValueError("a chemical formula in place of the element count needs the EADL database "
                         "(Decompose_compound); give the elements as 'Z  contribution' lines") from None
    for _ in range(n_el):
        Z, pers = read("ir")
        if Z not in ELEMENT_MASS:
            raise ValueError(f"no element with Z = {Z}")
        f.atoms.append(dict(Z=Z, pers=pers, shells=[]))
    # density, speed of sound, Fermi energy [, gap]: four values, else three
    # and a metal (RF:1326-1333); the gap is at least 0.1 eV
    rec = record()
    try:
        f.dens, f.vsound, f.E_F, egap = _read(rec, "rrrr")
    except ValueError:
        try:
            f.dens, f.vsound, f.E_F = _read(rec, "rrr")
        except ValueError as e:
            raise _Unreadable(f"line {pos}: {e}") from None
        egap = 0.0
    f.egap = egap if egap >= 0.1 else 0.1
    # the first shell count decides the form: a record that is not a number
    # (or the end of the file) is the single-pole file (RF:1344-1349)
    try:
        _read(lines[pos], "i")
    except (ValueError, IndexError):
        raise ValueError("a single-pole file (VALENCE / PHONON markers, no shell list) needs the EADL "
                         "database for its shells; give every shell with its oscillators") from None
    for j, at in enumerate(f.atoms):
        (nshl,) = read("i")
        for k in range(nshl):
            ncdf, shl_num, Ip, Nel, auger = read("iirrr")
            if ncdf <= 0:
                raise ValueError(f"element {at['Z']} shell {k + 1}: a shell without oscillators takes the BEB "
                                 "cross section for the ion too, which the reference marks as not working "
                                 "(MC:1772) and this library does not compute")
            osc = [tuple(read("rrr")) for _ in range(ncdf)]     # E0, A, Gamma
            at["shells"].append(dict(ncdf=ncdf, shl_num=abs(shl_num), Ip=max(Ip, 0.1), Nel=Nel, auger=auger,
                                     osc=osc, beb=shl_num < 0))
    # Auger times as check_atomic_parameters takes them (Dealing_with_EADL.f90
    # :366-370): none for the valence band (designator >= 63, or the last shell
    # of the first element, where this library keeps the band); a time <= 0 or
    # > 1e30 fs tells the reference to read it from EADL
    for j, at in enumerate(f.atoms):
        for k, s in enumerate(at["shells"]):
            vb = s["shl_num"] >= 63 or (j == 0 and k == len(at["shells"]) - 1)
            if vb:
                s["auger"] = 1.0e23
            elif not (0.0 < s["auger"] <= 1.0e30):
                raise ValueError(f"element {at['Z']} shell {k + 1}: Auger time {s['auger']} fs means 'take it from "
                                 "the EADL database', which this library does not read; give the time in fs")
    # phonon peaks: a count and as many oscillators, or the end of the file
    # (then the single-pole phonon CDF, RF:1584-1621)
    try:
        (nph,) = read("i")
    except EOFError:
        f.phonon = None
        return f
    f.phonon = [tuple(read("rrr")) for _ in range(nph)]
    return f


def read_dos(path_or_text: str) -> Tuple[np.ndarray, np.ndarray]:
    """Two columns, energy [eV] and DOS, one level per line (RF:2374-2380);
    MaterialDesc.prepare shifts and regrids them as the reference does."""
    text = _text(path_or_text)
    E, D = [], []
    for ln in text.splitlines():
        if not ln.strip():
            continue
        e, d = _read(ln, "rr")
        E.append(e)
        D.append(d)
    return np.array(E), np.array(D)


def _text(path_or_text: str) -> str:
    if "\n" in path_or_text:
        return path_or_text
    with open(path_or_text) as fh:
        return fh.read()


def read_cdf(path_or_text: str, layer: float, dos: Tuple[np.ndarray, np.ndarray], *, cut_off: float = 10.0,
             hole_mass: float = 1.0, el_eff_mass: float = 1.0, radiat=None, ek=None) -> MaterialDesc:
    """MaterialDesc of a full-CDF material file.  The oscillators are used as
    the file gives them (no sum-rule renormalisation in this form); the last
    shell of the first element is the valence band.  radiat: radiative decay
    times per (element, shell) [fs] (the reference takes them from EADL;
    default 1e23, no radiative decay).  ek: mean kinetic energies of the
    shells' electrons per (element, shell) [eV], needed for the shells with a
    negative designator (BEB, from EADL in the reference).  layer, dos,
    cut_off, hole_mass and el_eff_mass are the other input files' settings
    (INPUT_PARAMETERS.txt)."""
    f = parse_cdf(_text(path_or_text))
    atoms = []
    for j, at in enumerate(f.atoms):
        shells = []
        for k, s in enumerate(at["shells"]):
            rad = 1.0e23 if radiat is None else radiat[j][k]
            sh = Shell(s["Ip"], s["Nel"], list(s["osc"]), auger=s["auger"], radiat=rad, shl_num=s["shl_num"],
                       normalize=False)
            if s.get("beb"):
                if ek is None or ek[j][k] is None:
                    raise ValueError(f"element {at['Z']} shell {k + 1}: its BEB cross section needs the mean kinetic "
                                     "energy of the shell's electrons, which the reference reads from the EADL "
                                     "database; pass it as ek")
                sh.kocs, sh.Ek = 2, float(ek[j][k])
            shells.append(sh)
        atoms.append(Atom(at["Z"], ELEMENT_MASS[at["Z"]], at["pers"], shells))
    E, D = dos
    mat = MaterialDesc(f.name, atoms, dens=f.dens, layer=layer, vsound=f.vsound, E_F=f.E_F,
                       el_eff_mass=el_eff_mass, hole_mass=hole_mass, cut_off=cut_off,
                       phonon=list(f.phonon) if f.phonon else [], phonon_single_pole=f.phonon is None,
                       dos_raw_E=E, dos_raw=D)
    # The library takes the gap from the valence band's ionisation potential;
    # the file's own gap value is kept here for the caller to compare.
    mat.egap_file = f.egap
    mat.gap_differs = not math.isclose(f.egap, atoms[0].shells[-1].Ip, rel_tol=0.0, abs_tol=0.0)
    return mat


# -------------------------------------------------------------- run parameters
class Parameters(dict):
    """Fields of the reference's run-parameter file, after its own checks."""


def parse_parameters(text: str) -> Parameters:
    """The run-parameter file (INPUT_PARAMETERS.txt), read as Read_input_file
    reads it (Reading_files_and_parameters.f90:256-372): one list-directed
    READ per line in a fixed order, with the reference's corrections: the
    ion's mass from the element table unless a positive mass is given, dt
    at most Tim on a linear grid and at least 2 on a logarithmic one, a
    fixed Zeff within (0, Z], dispersion kind 1 and electron mass 1 where
    the values are out of range, no surface emission unless the work
    function, barrier length and height are all positive.  After the fixed
    lines, 'DOS <file>' and 'CDF <file>' name other material files."""
    lines = text.splitlines()
    pos = 0

    def read(kinds):
        nonlocal pos
        items = []
        while len(items) < len(kinds):
            if pos >= len(lines):
                raise ValueError("unreadable parameter file: it ends early")
            items += _items(lines[pos])
            pos += 1
        out = []
        for k, t in zip(kinds, items):
            if k == "s":
                out.append(t)
            else:
                try:
                    out.append((_fortran_int if k == "i" else _fortran_real)(t))
                except ValueError as e:
                    raise ValueError(f"unreadable parameter file: line {pos}: {e}") from None
        return out

    p = Parameters()
    (p["material"],) = read("s")
    (Z,) = read("i")
    p["Z"] = Z
    (p["E_MeV"],) = read("r")
    (M,) = read("r")
    p["mass"] = M if M > 0.0 else ELEMENT_MASS[abs(Z)]
    (p["Tim"],) = read("r")
    dt, flag = read("ri")
    if dt > p["Tim"] and flag <= 0:
        dt = p["Tim"]
    if dt < 2 and flag >= 1:
        dt = 2.0
    p["dt"], p["dt_flag"] = dt, flag
    (p["cut_off"],) = read("r")
    (p["layer"],) = read("r")
    (p["temp"],) = read("r")
    kz, fz = read("ir")
    if fz <= 0.0 or fz > Z:
        fz = float(Z)
    p["Kind_Zeff"], p["fixed_Zeff"] = kz, fz
    (p["Kind_ion"],) = read("i")
    p["kind_of_EMFP"], p["CDF_elast_Zeff"] = read("ii")
    dr, meff = read("ir")
    p["kind_of_DR"] = dr if 0 < dr <= 4 else 1
    p["El_eff_mass"] = meff if meff >= 0 else 1.0
    (pl,) = read("i")
    p["plasmon_Emax"] = int(pl == 1)
    (p["hole_mass"],) = read("r")
    (ph,) = read("i")
    p["include_photons"] = int(ph == 1)
    wf, bl, bh = read("rrr")
    # RF:378-381: emission is on only when all three are positive
    p["work_function"] = wf if (wf > 0.0 and bl > 0.0 and bh > 0.0) else 0.0
    p["bar_length"], p["bar_height"] = bl, bh
    (p["NMC"],) = read("i")
    (p["Num_th"],) = read("i")
    p["DOS_file"] = p["CDF_file"] = None
    for ln in lines[pos:]:
        it = _items(ln)
        if len(it) >= 2 and not it[0].startswith("!") and it[0] in ("DOS", "CDF"):
            p[it[0] + "_file"] = it[1]
    return p


def load_inputs(parameters_path: str, root: str = None, radiat=None):
    """(MaterialDesc, Ion, Options, NMC) from the reference's input tree: the
    parameter file and, under root (default: the parameter file's folder),
    INPUT_CDF/<material>.cdf and INPUT_DOS/<material>.dos (or the files the
    parameter file names).  With DSF elastic scattering (kind_of_EMFP = 2)
    the electron and hole DSF files INPUT_DSF/<material>/<material>_{Electron,
    Hole}_DSF_Differential_EMFPs_<T>K.dat (RF:316-322) are read into
    MaterialDesc.dsf; when one is missing the run falls back to Mott
    scattering, as the reference does (RF:2543-2552).

    radiat: radiative decay times [fs] per (element, shell), which the
    reference reads from the EADL database (Dealing_with_EADL.f90:356).  With
    radiative decay switched on in the parameter file they must be given;
    radiat="off" keeps photon transport but no shell decays radiatively."""
    import os
    from . import abi
    root = root or os.path.dirname(os.path.abspath(parameters_path))
    p = parse_parameters(_text(parameters_path))
    cdf = os.path.join(root, "INPUT_CDF", p["CDF_file"] or p["material"] + ".cdf")
    dos = os.path.join(root, "INPUT_DOS", p["DOS_file"] or p["material"] + ".dos")
    if p["include_photons"] and radiat is None:
        raise ValueError("radiative decay is on in the parameter file: give the radiative decay times per shell "
                         "(radiat=[[...], ...]; the reference reads them from the EADL database) or radiat='off'")
    mat = read_cdf(cdf, layer=p["layer"], dos=read_dos(dos), cut_off=p["cut_off"], hole_mass=p["hole_mass"],
                   el_eff_mass=p["El_eff_mass"], radiat=None if isinstance(radiat, str) else radiat)
    mat.temp = p["temp"]
    mat.work_function = p["work_function"]
    if mat.work_function > 0.0:
        mat.bar_length, mat.bar_height = p["bar_length"], p["bar_height"]
    ion = abi.Ion()
    ion.Zat, ion.Mass, ion.E = p["Z"], p["mass"], p["E_MeV"] * 1.0e6
    ion.Kind_Zeff, ion.fixed_Zeff, ion.Kind_ion = p["Kind_Zeff"], p["fixed_Zeff"], p["Kind_ion"]
    opt = abi.Options()
    opt.kind_of_EMFP, opt.CDF_elast_Zeff = p["kind_of_EMFP"], p["CDF_elast_Zeff"]
    opt.kind_of_DR, opt.plasmon_Emax, opt.include_photons = p["kind_of_DR"], p["plasmon_Emax"], p["include_photons"]
    opt.Tim, opt.dt, opt.dt_flag = p["Tim"], p["dt"], p["dt_flag"]
    if opt.kind_of_EMFP == 2:
        from .dsf import read_dsf
        name = p["material"]
        files = [os.path.join(root, "INPUT_DSF", name, f"{name}_{who}_DSF_Differential_EMFPs_{int(mat.temp)}K.dat")
                 for who in ("Electron", "Hole")]
        if all(os.path.exists(f) for f in files):
            mat.dsf = tuple(read_dsf(f) for f in files)
        else:
            opt.kind_of_EMFP = 0
    return mat, ion, opt, p["NMC"]
