"""The FP32 table quadrature of the precision form (trk_options.quadrature = 1:
2000 geometric transfer-energy intervals with midpoints, 1000 Simpson
intervals in ln q, FP32 Kahan sums, FP64 cumulative normalisation).

Pinned two ways:
* against the exact double integral of the same Drude-Lorentz loss function
  (free-electron dispersion, m* = 1), computed here in FP64 with numpy and
  scipy: the inverse mean free path 1/lambda = int dw (1/(pi a0 E)) int
  Im[-1/eps(w, q)] dln q and the stopping power int w (...) dw agree to
  1e-3 relative -- the FP32 sums lose nothing a plausible mistake (a
  dropped Simpson weight, a wrong q limit or mass) would not show far
  above that;
* against the FP64 tables of the reference's quadrature (which the oracle
  reproduces bit for bit): the relative difference of every mean free path
  and stopping power over the energy grid is stated and bounded by the
  reference's own quadrature bias (its left-node Simpson weights: < 2 % for
  mean free paths and < 6 % for stopping powers at this grid)."""
import math

import numpy as np
import pytest
from scipy import integrate

import trekis_gpu as tg
from trekis_gpu import abi

from gpu_helpers import say

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")]

A0, PI = 0.5291772085936, math.pi


def _ctx(name, quadrature):
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(name)
    mat.prepare(lib)
    opt.quadrature = quadrature
    ctx = tg.Context(mat, ion, opt)
    ctx.build_tables()
    return ctx, mat, opt


def _shell_osc(mat, s):
    flat = [sh for a in mat.atoms for sh in a.shells]
    return flat[s].osc, max(flat[s].Ip, 1e-3)


def _exact_inel(osc, Ip, Ele, Mass=1.0):
    """(1/lambda, stopping power) of one shell for an electron of energy Ele,
    m* = 1, free-electron dispersion: Im = sum A G w / ((w^2 - (E0 + q~^2)^2)^2
    + G^2 w^2) with q~ = q / sqrt(2 m_e 1 eV) * hbar (so hbar^2 q^2 / 2 m_e = q~^2 eV)."""
    E0 = np.array([o[0] for o in osc])[:, None]
    A = np.array([o[1] for o in osc])[:, None]
    G = np.array([o[2] for o in osc])[:, None]

    def dsdw(w):
        lo = math.sqrt(Mass) * (math.sqrt(Ele) - math.sqrt(Ele - w))
        hi = math.sqrt(Mass) * (math.sqrt(Ele) + math.sqrt(Ele - w))
        u = np.linspace(math.log(lo), math.log(hi), 16001)
        q2 = np.exp(2.0 * u)[None, :]
        d = w * w - (E0 + q2) ** 2
        f = (A * G * w / (d * d + G * G * w * w)).sum(axis=0)
        h = u[1] - u[0]
        wts = np.ones_like(u)
        wts[1:-1:2], wts[2:-1:2] = 4.0, 2.0
        return (h / 3.0) * float((wts * f).sum()) / (PI * A0 * Ele)

    Emax = (Ele + Ip) / 2.0
    pts = [o[0] for o in osc if Ip < o[0] < Emax] or None
    inv = integrate.quad(dsdw, Ip, Emax, limit=400, epsrel=1e-9, points=pts)[0]
    sp = integrate.quad(lambda w: w * dsdw(w), Ip, Emax, limit=400, epsrel=1e-9, points=pts)[0]
    return inv, sp


def test_fp32_quadrature_against_exact_integral():
    ctx, mat, opt = _ctx("test", 1)
    E, _, _ = ctx.table(abi.TAB_E_INEL, 0)
    worst = 0.0
    for s in range(sum(mat.n_shells)):
        osc, Ip = _shell_osc(mat, s)
        _, L, D = ctx.table(abi.TAB_E_INEL, s)
        for Ele in (50.0, 200.0, 1000.0, 1900.0):
            i = int(np.argmin(np.abs(E - Ele)))
            if E[i] <= 2.0 * Ip or L[i] >= 1e15:
                continue
            inv, sp = _exact_inel(osc, Ip, E[i])
            rel_L = abs(L[i] * inv - 1.0)        # lambda = 1 / (M int), M = 1 for electrons
            rel_D = abs(D[i] / sp - 1.0)
            say(f"[q32] shell {s} E {E[i]:.1f} eV: lambda rel err {rel_L:.2e}, dE/dx rel err {rel_D:.2e}")
            worst = max(worst, rel_L)
            assert rel_L < 1e-3, (s, E[i], rel_L)
            assert rel_D < 1e-3, (s, E[i], rel_D)
    assert worst > 0.0


def test_fp32_tables_error_against_fp64_reference_tables():
    ref, mat, opt = _ctx("config2", 0)
    q32, _, _ = _ctx("config2", 1)
    nsh = sum(mat.n_shells)
    for kind in (abi.TAB_E_INEL, abi.TAB_H_INEL, abi.TAB_SHI):
        for s in range(nsh):
            E0, L0, D0 = ref.table(kind, s)
            E1, L1, D1 = q32.table(kind, s)
            assert np.array_equal(E0, E1)
            ok = (L0 < 1e14) & (L1 < 1e14) & (L0 > 0) & (D0 > 0)
            if not ok.any():
                continue
            rl = np.abs(L1[ok] / L0[ok] - 1.0)
            rd = np.abs(D1[ok] / D0[ok] - 1.0)
            say(f"[q32 vs fp64] {abi.TABLE_NAMES[kind]} shell {s}: lambda max {rl.max():.2e} median "
                f"{np.median(rl):.2e}; dE/dx max {rd.max():.2e} median {np.median(rd):.2e} over {int(ok.sum())} nodes")
            assert np.median(rl) < 2e-2 and np.median(rd) < 6e-2, (kind, s)
    # cumulative rows of the valence band: 2000 entries, never increasing,
    # ending at the table's mean free path
    vb = mat.n_shells[0] - 1
    E, L, _ = q32.table(abi.TAB_E_INEL, vb)
    i = int(np.searchsorted(E, 1000.0))
    hw, lam = q32.row(abi.TAB_E_INEL, vb, i)
    assert len(hw) == 2000 and np.all(np.diff(lam) <= 0.0)
    assert abs(lam[-1] / L[i] - 1.0) < 1e-12
