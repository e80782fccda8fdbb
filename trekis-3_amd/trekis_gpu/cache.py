"""Table caches in the reference's on-disk formats.

The reference keeps every computed table in text files beside its output and
reads them back instead of integrating again (Analytical_IMFPs.f90):

* mean free paths, one file per particle: a row per energy node with the
  energy, the mean free path of every shell of every atom (atoms in order,
  shells in order) and the total, written '(f)', '(e)' ... '(e)' and read back
  with the same descriptors (read_write of AI:596-630; file names AI:259-290
  electrons, 386-405 holes, 505-520 photons);
* elastic mean free paths: a row per node, energy and total mean free path,
  '(f,e)' (AI:1394-1415; names AI:917-950 CDF, 1066-1070 Mott, 1124-1127 none
  for electrons, 1195-1201, 1307-1310, 1375 for holes);
* the ion: '<name>_IMFP.dat' and '<name>_dEdx.dat', energy in MeV then every
  shell then the total, '(e)' throughout (Analytical_SHI_dEdx, AI:2600-2660;
  read back by read_SHI_MFP, Reading_files_and_parameters.f90:2855-2903);
* cumulative transfer rows, folder diff_CS/, one file per (atom, shell,
  energy node): the transfer energy and the cumulative mean free path, one
  pair per line '(es,es)', read back list-directed (AI:663-711 electrons,
  745-790 holes (valence band only), 1450-1510 and 1540-1600 elastic).

Field layout: Intel Fortran's default fields for REAL(8), 25 characters wide
(F25.16, E25.16, ES25.16), which is what the reference's width-less '(f)',
'(e)' and '(es)' descriptors give under the compiler it is built with.
Mantissas are written with 17 significant digits, enough for every double
to read back bit for bit; the reader accepts any Fortran real (D or E
exponents, exponents of three digits without the letter).

write_cache() writes every table of a built context; seed_from_cache() reads
them back into a context whose tables were built on the same grids (the
reference reads caches only when the line count equals its grid, AI:318-327,
and otherwise computes again -- here a mismatch is an error) and re-derives
the fused totals.  Stopping powers of the inelastic kinds are not in the
reference's mean-free-path files; the built ones are kept.
"""
import math
import os
import re

import numpy as np

from . import abi

# Element symbols by atomic number (the names Find_element_name gives,
# Dealing_with_EADL.f90:1217-1703; they name the per-shell diff_CS files).
ELEMENT_SYMBOL = (
    None, "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar", "K",
    "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr",
    "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La",
    "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os",
    "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am",
    "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh", "Hs", "Mt", "Ds", "Rg", "Cn", "Nh", "Fl",
    "Mc", "Lv", "Ts", "Og")


def _subshells():
    """EADL subshell designators 1..62 and their names (define_PQN,
    Dealing_with_EADL.f90:1016-1214): each principal shell is named, then
    its subshells, grouped pairs (e.g. L23) before their members."""
    groups = {
        "K": [],
        "L": ["1", ("2", "3")],
        "M": ["1", ("2", "3"), ("4", "5")],
        "N": ["1", ("2", "3"), ("4", "5"), ("6", "7")],
        "O": ["1", ("2", "3"), ("4", "5"), ("6", "7"), ("8", "9")],
        "P": ["1", ("2", "3"), ("4", "5"), ("6", "7"), ("8", "9"), ("10", "11")],
        "Q": ["1", ("2", "3"), "..."],
    }
    names = []
    for shell, subs in groups.items():
        names.append(shell + "-shell")
        for s in subs:
            if isinstance(s, tuple):
                names.append(shell + s[0] + s[1] + "-shell")
                names += [shell + x + "-shell" for x in s]
            else:
                names.append(shell + s + "-shell")
    return tuple(names)


SHELL_NAMES = _subshells()      # SHELL_NAMES[d - 1] for designator d


def shell_name(designator):
    d = int(designator)
    if d >= 63:
        return "Valence"
    if d < 1:
        return "K-shell"
    return SHELL_NAMES[d - 1]


# ---------------------------------------------------------------- fields

def _fixed(x, w, d):
    s = f"{x:{w}.{d}f}"
    return "*" * w if len(s) > w else s


def _digits(x, nd):
    """(sign, nd significant digits, decimal exponent e) with x = 0.digits * 10^e."""
    if x == 0.0:
        return ("-" if math.copysign(1.0, x) < 0 else ""), "0" * nd, 0
    s = f"{abs(x):.{nd - 1}e}"
    mant, ex = s.split("e")
    return ("-" if x < 0 else ""), mant.replace(".", ""), int(ex) + 1


def _exponent(e):
    return f"E{e:+03d}" if abs(e) <= 99 else f"{e:+04d}"


def fortran_e(x, w=25, nd=17):
    """Ew.d: 0.ddd...E+ee, right-justified in w characters."""
    if not math.isfinite(x):
        return f"{'NaN' if math.isnan(x) else ('-Infinity' if x < 0 else 'Infinity'):>{w}}"
    sign, dig, e = _digits(x, nd)
    s = f"{sign}0.{dig}{_exponent(e)}"
    return "*" * w if len(s) > w else s.rjust(w)


def fortran_es(x, w=25, nd=17):
    """ESw.d: d.ddd...E+ee, right-justified in w characters."""
    if not math.isfinite(x):
        return f"{'NaN' if math.isnan(x) else ('-Infinity' if x < 0 else 'Infinity'):>{w}}"
    sign, dig, e = _digits(x, nd)
    s = f"{sign}{dig[0]}.{dig[1:]}{_exponent(e - 1 if x != 0.0 else 0)}"
    return "*" * w if len(s) > w else s.rjust(w)


def fortran_f(x, w=25, d=16):
    return _fixed(x, w, d)


_NOLETTER = re.compile(r"^([+-]?[0-9.]+)([+-][0-9]+)$")


def parse_real(tok):
    """A Fortran real: D/d/E/e exponents, or a three-digit exponent written
    without its letter (Ew.d output for |exponent| > 99)."""
    t = tok.strip().replace("D", "E").replace("d", "E")
    m = _NOLETTER.match(t)
    if m:
        t = m.group(1) + "E" + m.group(2)
    return float(t)


def _rows(path):
    with open(path) as f:
        return [[parse_real(t) for t in line.split()] for line in f if line.strip()]


# ---------------------------------------------------------------- file names

def _temp(T, w):
    return _fixed(T, w, 2).strip()


def _mass_tag(m_eff, T):
    t = _temp(T, 9)
    if m_eff == 0.0:
        return f"DOS_{t}_K"
    if m_eff < 0.0 or m_eff == 1.0:
        return f"Me_{t}_K"
    return f"{_fixed(m_eff, 5, 2)}_me_{t}_K".lstrip()


_DR_NAME = {2: "Plasmon_pole", 3: "Ritchie"}
_ZEFF_TAG = {1: "Z=1_", 2: "Z_CDFe_FF_", 3: "Z_CDFe_"}
_CHARGE_NAME = {1: "_Bohr", 2: "_ND", 3: "_SG", 4: "_fixed"}


def file_names(mat, ion, opt):
    """Cache file names of a material descriptor (abi.Material), ion and
    options, as the reference composes them (CDF cross sections of the
    Ritchie-Howie kind, KCS = '_CDF')."""
    T = float(mat.temp)
    kcs = "_CDF"
    e_tag = kcs[1:] + "_" + _mass_tag(float(mat.El_eff_mass), T)
    dr = _DR_NAME.get(int(opt.kind_of_DR), "Free")
    n = {"e_inel": f"OUTPUT_Electron_IMFPs_{dr}_{e_tag}.dat",
         "h_inel": f"OUTPUT_Hole_IMFPs_CDF{kcs}_{_temp(T, 7)}_K.dat",
         "phot": f"OUTPUT_Photon_IMFPs{kcs}_{_temp(T, 7)}_K.dat"}
    if opt.kind_of_EMFP == 1:
        n["e_el"] = f"OUTPUT_Electron_EMFPs{kcs}_{_ZEFF_TAG.get(int(opt.CDF_elast_Zeff), 'Zeff_')}{_temp(T, 7)}_K.dat"
        n["h_el"] = f"OUTPUT_Hole{kcs}_EMFPs_{_temp(T, 7)}_K.dat"
    elif opt.kind_of_EMFP == 2:
        # AI:809-811, 1142-1145
        n["e_el"] = f"OUTPUT_Electron_DSF_EMFPs_{_temp(T, 7)}_K.dat"
        n["h_el"] = f"OUTPUT_Hole_DSF_EMFPs_{_temp(T, 7)}_K.dat"
    elif opt.kind_of_EMFP == 0:
        n["e_el"] = "OUTPUT_Electron_EMFPs_Mott.dat"
        n["h_el"] = f"OUTPUT_Hole_Mott_EMFPs_{_temp(T, 7)}_K.dat"
    else:
        n["e_el"] = "OUTPUT_Electron_No_elas_EMFPs.dat"
        n["h_el"] = "OUTPUT_Hole_No_elas_EMFPs.dat"
    shi = (f"OUTPUT_{ELEMENT_SYMBOL[int(ion.Zat)]}{kcs}{_CHARGE_NAME.get(int(ion.Kind_Zeff), '_Barkas')}"
           f"{'_BK' if ion.Kind_ion == 1 else '_P'}")
    n["shi"] = shi + "_IMFP.dat"
    n["shi_dedx"] = shi + "_dEdx.dat"
    return n


def _shells(mat):
    """(atom symbol, shell name) of every shell, flat in the tables' order."""
    out = []
    s = 0
    for a in range(mat.n_atoms):
        for _ in range(mat.n_shells[a]):
            out.append((ELEMENT_SYMBOL[mat.Zat[a]], shell_name(mat.Shl_num[s])))
            s += 1
    return out


def _row_file(folder, stem, E):
    return os.path.join(folder, "diff_CS", f"{stem}_{_fixed(E, 14, 3).strip()}.dat")


def _stem(name):
    """file_name(8:len-4): the cache file's name without 'OUTPUT_' and '.dat'."""
    return name[7:-4]


def _row_files(ctx, names, folder):
    """(kind, shell, node, path) of every cumulative row the reference keeps
    in diff_CS/: every shell for electrons, the valence band for holes, the
    single elastic row set (phonon CDF only)."""
    m = ctx.mat
    shells = _shells(m)
    vb = m.n_shells[0] - 1
    out = []
    E, _, _ = ctx.table(abi.TAB_E_INEL, 0)
    for s, (sym, sh) in enumerate(shells):
        for i, e in enumerate(E):
            out.append((abi.TAB_E_INEL, s, i, _row_file(folder, f"{_stem(names['e_inel'])}_{sym}_{sh}", e)))
    E, _, _ = ctx.table(abi.TAB_H_INEL, 0)
    sym, sh = shells[vb]
    for i, e in enumerate(E):
        out.append((abi.TAB_H_INEL, vb, i, _row_file(folder, f"{_stem(names['h_inel'])}_{sym}_{sh}", e)))
    if ctx.opt.kind_of_EMFP == 1:
        for kind, key in ((abi.TAB_E_EL, "e_el"), (abi.TAB_H_EL, "h_el")):
            E, _, _ = ctx.table(kind, 0)
            for i, e in enumerate(E):
                out.append((kind, 0, i, _row_file(folder, _stem(names[key]), e)))
    return out


# ---------------------------------------------------------------- writing

def _write_mfp(path, E, L, scale=1.0, first=fortran_f):
    """E, every shell's column, then the total 1/sum(1/L)."""
    with np.errstate(divide="ignore"):
        tot = 1.0 / (1.0 / L).sum(axis=0)
    with open(path, "w") as f:
        for i in range(len(E)):
            f.write(first(E[i] * scale) + "".join(fortran_e(L[s, i]) for s in range(L.shape[0]))
                    + fortran_e(float(tot[i])) + "\n")


def _kind_tables(ctx, kind, nsh):
    E = None
    L, D = [], []
    for s in range(nsh):
        E, l_, d_ = ctx.table(kind, s)
        L.append(l_)
        D.append(d_)
    return E, np.array(L), np.array(D)


def write_cache(ctx, folder, ion_only=False):
    """Write the tables of a built context in the reference's cache layout
    under folder (and folder/diff_CS).  ion_only: the ion's files only (an
    ion context sharing its material's tables).  Returns the file names."""
    m = ctx.mat
    nsh = sum(m.n_shells[a] for a in range(m.n_atoms))
    names = file_names(m, ctx.ion, ctx.opt)
    os.makedirs(os.path.join(folder, "diff_CS"), exist_ok=True)
    E, L, D = _kind_tables(ctx, abi.TAB_SHI, nsh)
    _write_mfp(os.path.join(folder, names["shi"]), E, L, 1.0e-6, fortran_e)
    with open(os.path.join(folder, names["shi_dedx"]), "w") as f:
        for i in range(len(E)):
            f.write(fortran_e(E[i] * 1.0e-6) + "".join(fortran_e(D[s, i]) for s in range(nsh))
                    + fortran_e(float(D[:, i].sum())) + "\n")
    if ion_only:
        return names
    for kind, key in ((abi.TAB_E_INEL, "e_inel"), (abi.TAB_H_INEL, "h_inel")):
        E, L, _ = _kind_tables(ctx, kind, nsh)
        _write_mfp(os.path.join(folder, names[key]), E, L)
    if ctx.opt.include_photons:
        E, L, _ = _kind_tables(ctx, abi.TAB_PHOT, nsh)
        _write_mfp(os.path.join(folder, names["phot"]), E, L)
    for p, (kind, key) in enumerate(((abi.TAB_E_EL, "e_el"), (abi.TAB_H_EL, "h_el"))):
        E, L, _ = ctx.table(kind, 0)
        with open(os.path.join(folder, names[key]), "w") as f:
            if ctx.opt.kind_of_EMFP == 2:
                # DSF: total, emission and absorption paths (AI:1392-1398)
                t = ctx.desc.dsf[p]
                for i in range(len(E)):
                    f.write(fortran_f(E[i]) + fortran_es(L[i]) + fortran_es(t.dL_emit[i, -1])
                            + fortran_es(t.dL_absorb[i, -1]) + "\n")
            else:
                for i in range(len(E)):
                    f.write(fortran_f(E[i]) + fortran_e(L[i]) + "\n")
    for kind, s, i, path in _row_files(ctx, names, folder):
        hw, lam = ctx.row(kind, s, i)
        with open(path, "w") as f:
            for a, b in zip(hw, lam):
                f.write(fortran_es(a) + fortran_es(b) + "\n")
    return names


# ---------------------------------------------------------------- reading

class CacheMismatch(ValueError):
    """A cache file does not fit the context's grids (the reference computes
    the tables again in this case)."""


def _check_grid(path, got, want, rel):
    """The energy column against the built grid: equal to the field's
    rounding (16 decimals of F25.16, or rel for a converted column)."""
    if len(got) != len(want):
        raise CacheMismatch(f"{path}: {len(got)} energy nodes, the grid has {len(want)}")
    tol = np.maximum(rel * np.abs(want), 0.5e-16 + 2.3e-16 * np.abs(want))
    bad = np.abs(np.asarray(got) - want) > tol
    if bad.any():
        i = int(np.argmax(bad))
        raise CacheMismatch(f"{path}: energy node {i} is {got[i]!r}, the grid has {want[i]!r}")


def _read_mfp(path, nsh, E_grid, scale=1.0, rel=0.0):
    r = _rows(path)
    if any(len(x) != nsh + 2 for x in r):
        raise CacheMismatch(f"{path}: rows must hold the energy, {nsh} shells and the total")
    a = np.array(r)
    _check_grid(path, a[:, 0] * scale, E_grid, rel)
    return a[:, 1:1 + nsh].T.copy()


def seed_from_cache(ctx, folder, ion_only=False):
    """Overwrite the built tables of ctx with those in the cache folder and
    re-derive the fused totals.  The energy column must equal the built grid
    (the ion's, written in MeV, to the rounding of the conversion); rows must
    have the built lengths."""
    m = ctx.mat
    nsh = sum(m.n_shells[a] for a in range(m.n_atoms))
    names = file_names(m, ctx.ion, ctx.opt)
    E, _, _ = ctx.table(abi.TAB_SHI, 0)
    L = _read_mfp(os.path.join(folder, names["shi"]), nsh, E, 1.0e6, 4e-16)
    D = _read_mfp(os.path.join(folder, names["shi_dedx"]), nsh, E, 1.0e6, 4e-16)
    for s in range(nsh):
        ctx.set_table(abi.TAB_SHI, s, L=L[s], dEdx=D[s])
    if not ion_only:
        kinds = [(abi.TAB_E_INEL, "e_inel"), (abi.TAB_H_INEL, "h_inel")]
        if ctx.opt.include_photons:
            kinds.append((abi.TAB_PHOT, "phot"))
        for kind, key in kinds:
            E, _, _ = ctx.table(kind, 0)
            L = _read_mfp(os.path.join(folder, names[key]), nsh, E)
            for s in range(nsh):
                ctx.set_table(kind, s, L=L[s])
        for kind, key in ((abi.TAB_E_EL, "e_el"), (abi.TAB_H_EL, "h_el")):
            E, _, _ = ctx.table(kind, 0)
            path = os.path.join(folder, names[key])
            a = np.array(_rows(path))
            # a DSF file adds the emission and absorption paths, which the
            # library takes from the DSF tables' last entries, as they are made
            ncol = 4 if ctx.opt.kind_of_EMFP == 2 else 2
            if a.ndim != 2 or a.shape[1] != ncol:
                raise CacheMismatch(f"{path}: rows must hold the energy and the mean free path"
                                    + (" (total, emission, absorption)" if ncol == 4 else ""))
            _check_grid(path, a[:, 0], E, 0.0)
            ctx.set_table(kind, 0, L=a[:, 1])
        for kind, s, i, path in _row_files(ctx, names, folder):
            a = np.array(_rows(path))
            hw, _ = ctx.row(kind, s, i)
            if len(a) == 0 and len(hw) == 0:
                continue
            if a.ndim != 2 or a.shape[1] != 2:
                raise CacheMismatch(f"{path}: rows must hold the transfer energy and the cumulative path")
            if len(hw) != len(a):
                raise CacheMismatch(f"{path}: {len(a)} entries, the built row has {len(hw)}")
            ctx.set_row(kind, s, i, a[:, 0], a[:, 1])
    ctx.finish_tables()
    return names
