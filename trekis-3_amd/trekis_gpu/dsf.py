"""Dynamic-structure-factor (DSF) elastic scattering tables
(kind_of_EMFP = 2).

The reference reads one file per particle,
INPUT_DSF/<material>/<material>_{Electron,Hole}_DSF_Differential_EMFPs_<T>K.dat
(Reading_files_and_parameters.f90:316-322), and turns it into DSF_DEMFP:
for every particle energy, cumulative mean free paths over the transferred
energy, split into emission (dE >= 0, energy to the lattice) and absorption
(reading_DSF_cross_sections, RF:2516-2678).  read_dsf restates that
processing; the arrays it returns are what trk_dsf_set takes.

File layout: line 1 the number NTEPo of transferred energies per particle
energy; then lines "E dE DSF" [eV, eV, 1/(A eV)], NTEPo consecutive lines
per particle energy, dE decreasing within a block (the reader reverses it).

One deviation: the reference reads the data lines 2..N-1 of an N-line file
(RF:2562: its loop stops one line short), which leaves the last transferred
energy of the last particle energy unset; here every data line is read.
"""
import math
from dataclasses import dataclass

import numpy as np

from .inputs import _text


@dataclass
class DsfTable:
    """DSF_DEMFP of one particle (Objects.f90:173-178): particle energies E
    [eV] (nE,), and per energy the transferred energies dE [eV] with the
    cumulative mean free path dL and its emission and absorption parts
    dL_emit, dL_absorb [A], each (nE, nW)."""
    E: np.ndarray
    dE: np.ndarray
    dL: np.ndarray
    dL_emit: np.ndarray
    dL_absorb: np.ndarray

    def arrays(self):
        return tuple(np.ascontiguousarray(a, dtype=np.float64) for a in
                     (self.E, self.dE, self.dL, self.dL_emit, self.dL_absorb))


def _find_mono_2d(x, v):
    """Find_in_monotonous_2D_array on an increasing array (RF:3496-3559),
    1-based result: the bisection stops as soon as x[i_cur] <= v <= x[i_cur+1]."""
    n = len(x)
    i1, i2 = 1, n
    icur = (i1 + i2) // 2
    tv = x[icur - 1]
    if v < x[0]:
        icur = 0
    elif v >= x[n - 1]:
        icur = n - 1
    else:
        while not (v >= x[icur - 1] and v <= x[icur]):
            if tv <= v:
                i1 = icur
            else:
                i2 = icur
            icur = (i1 + i2) // 2
            tv = x[icur - 1]
    return icur + 1


def _approx(x, y, v, el1, el2):
    """Linear_approx_2d with both starting values given (RF:3272-3307)."""
    k = _find_mono_2d(x, v)
    if k == 1:
        return el2 + (y[0] - el2) / (x[0] - el1) * (v - el1)
    if y[k - 2] > 1.0e20:
        return y[k - 2]
    return y[k - 2] + (y[k - 1] - y[k - 2]) / (x[k - 1] - x[k - 2]) * (v - x[k - 2])


def process_dsf(E, dE_raw, dsf_raw):
    """The processing of reading_DSF_cross_sections (RF:2600-2670) on raw
    blocks: E (nE,), dE_raw and dsf_raw (nE, nW) in file order.  Every
    block is reversed, resampled on nW equal steps from -0.2 eV up to
    min(largest dE, 0.2 eV) by linear interpolation (zero at one step below
    -0.2 eV), and integrated by rectangles into 1/sum mean free paths (1e30
    where the sum is below 1e-10); emission sums only dE >= 0, absorption
    keeps the running total of the dE < 0 part."""
    E = np.asarray(E, dtype=np.float64)
    dE_raw = np.asarray(dE_raw, dtype=np.float64)
    dsf_raw = np.where(np.asarray(dsf_raw, dtype=np.float64) < 1.0e-10, 0.0, dsf_raw)
    nE, nW = dE_raw.shape
    if not np.all(np.diff(dE_raw, axis=1) < 0.0) or not np.all(np.diff(E) > 0.0):
        raise ValueError("DSF blocks must list decreasing transferred energies, at increasing particle energies")
    out_dE = np.zeros((nE, nW))
    out_L = np.zeros((nE, nW))
    out_Le = np.zeros((nE, nW))
    out_La = np.zeros((nE, nW))
    for i in range(nE):
        x = [float(v) for v in dE_raw[i, ::-1]]
        y = [float(v) for v in dsf_raw[i, ::-1]]
        Emin = -0.2
        e = Emin
        Emax = x[nW - 1]
        if Emax > 0.2:
            Emax = 0.2
        step = (Emax - Emin) / float(nW)
        s_all = 0.0
        s_emit = 0.0
        for j in range(nW):
            e = e + step
            loc = _approx(x, y, e, Emin - step, 0.0)
            s_all = s_all + loc * step
            out_dE[i, j] = e
            out_L[i, j] = 1.0 / s_all if abs(s_all) > 1.0e-10 else 1.0e30
            if e >= 0.0:
                s_emit = s_emit + loc * step
                out_Le[i, j] = 1.0 / s_emit if abs(s_emit) > 1.0e-10 else 1.0e30
                out_La[i, j] = 1.0e30 if j == 0 else out_La[i, j - 1]
            else:
                out_La[i, j] = 1.0 / s_all if abs(s_all) > 1.0e-10 else 1.0e30
                out_Le[i, j] = 1.0e30 if j == 0 else out_Le[i, j - 1]
    return DsfTable(E.copy(), out_dE, out_L, out_Le, out_La)


def parse_dsf(text):
    """(E, dE_raw, dsf_raw) of a DSF file's text, blocks in file order."""
    lines = [ln for ln in text.splitlines() if ln.strip()]
    if len(lines) < 3:
        raise ValueError("a DSF file holds the block length and at least two data lines")
    nW = int(float(lines[0].split()[0]))
    rows = np.array([[float(t.replace("d", "e").replace("D", "e")) for t in ln.split()[:3]] for ln in lines[1:]])
    nE = len(rows) // nW
    if nW < 2 or nE < 2:
        raise ValueError(f"a DSF file needs at least two blocks of at least two lines (NTEPo={nW}, {len(rows)} lines)")
    rows = rows[: nE * nW].reshape(nE, nW, 3)
    return rows[:, 0, 0].copy(), rows[:, :, 1].copy(), rows[:, :, 2].copy()


def read_dsf(path_or_text):
    """DsfTable of a DSF file (path or its text), processed as the reference
    reads it (reading_DSF_cross_sections, RF:2516-2678)."""
    return process_dsf(*parse_dsf(_text(path_or_text)))


def dsf_file_text(E, dE_raw, dsf_raw):
    """The text of a DSF file holding the given blocks (file order)."""
    nE, nW = dE_raw.shape
    out = [f"{nW}"]
    for i in range(nE):
        for j in range(nW):
            out.append(f"{E[i]:.17e} {dE_raw[i, j]:.17e} {dsf_raw[i, j]:.17e}")
    return "\n".join(out) + "\n"


def synthetic_dsf_blocks(Emin, Emax, temp, nE=40, nW=120, w0=0.04, sigma=0.015, lam0=20.0, lam_slope=5.0):
    """Raw DSF blocks of a synthetic crystal (no DSF data ships with the
    reference): a phonon peak of energy w0 [eV] and width sigma on both
    sides of zero transfer, the absorption side weighted by the Boltzmann
    factor exp(-w0/kT) (detailed balance at temp [K]), scaled so that the
    elastic mean free path is lam0 + lam_slope ln(1 + E) [A].  Particle
    energies log-spaced over [Emin, Emax]; transfers from +0.25 to -0.25 eV
    in file order."""
    E = np.geomspace(Emin, Emax, nE)
    w = np.linspace(0.25, -0.25, nW)
    kT = max(temp, 1.0) * 8.617333262e-5
    shape = np.exp(-((w - w0) / sigma) ** 2) + math.exp(-w0 / kT) * np.exp(-((w + w0) / sigma) ** 2)
    norm = np.trapezoid(shape[::-1], w[::-1])
    lam = lam0 + lam_slope * np.log1p(E)
    dsf = (1.0 / (lam * norm))[:, None] * shape[None, :]
    return E, np.tile(w, (nE, 1)), dsf


def synthetic_dsf(temp, e_Emax=1.0e5, h_Emax=30.0):
    """(electron, hole) DsfTables of the synthetic crystal, processed as the
    reference processes a file."""
    el = process_dsf(*synthetic_dsf_blocks(0.01, e_Emax, temp))
    ho = process_dsf(*synthetic_dsf_blocks(0.01, h_Emax, temp, lam0=10.0, lam_slope=2.0))
    return el, ho
