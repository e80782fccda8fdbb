"""The mixed-precision path (trk_options.precision = 1) against the FP64 path.

The mixed path samples in FP32 (logarithms, exponentials, trigonometric
functions of the free-flight, energy-transfer and angle sampling) while the
particle state, the energy ledger and the tallies stay FP64.  Both paths draw
the same per-particle Philox streams, so on the same iterations a history of
the mixed path follows the FP64 history until a rounding difference flips a
decision; the radial tallies then differ by far less than the Monte Carlo
noise of either.

Tolerances, per radial tally (time x radius; hole tallies summed over shells):
  - at the last time node and three intermediate ones (a quarter, half and
    three quarters of the window), the largest relative difference over the
    radial bins that hold at least 1% of the node's peak and where the FP64
    path counts at least MIN_COUNTS particles of the tally's kind (electrons
    for Out_ne, Out_Ee and Out_Elat, holes for Out_nh, Out_Eh, Out_Ehkin):
    1e-2, the accuracy the mixed form is specified to.  A history that takes
    another branch moves a bin by about one count, but a fast electron's
    whole cascade moves the far tail of the radial profile at once, hence
    the 1% of the peak.  Out_Elat is compared as the lattice energy
    deposited up to each time node (its per-node increments hold a few
    events per bin).  Energy tallies of the first femtoseconds near the
    track are carried by a few fast electrons, which one flipped history
    moves by more; they enter through the next bound only;
  - relative L1 difference over all bins and nodes: 5e-3.
The workloads: the Al-like metal, the Al2O3-like target with the Au ion,
the SiO2-like target with phonon scattering and photons, and the material
read from the example full-CDF file (test_cdf).
Photon tallies of these workloads hold only a few photon counts at the time
nodes and are not checked per bin.  The total deposited energy (Out_tot_E)
agrees to 1e-5 relative: the ion's energy transfers are sampled in FP32 and
its position is kept in FP32 too, so a flipped ionisation changes it, by far
less than its noise (~1e-3 for these iteration counts)."""
import numpy as np
import pytest

import trekis_gpu as tg

from gpu_helpers import setup, say

pytestmark = pytest.mark.gpu

RADIAL = {"Out_ne": "Out_ne", "Out_Ee": "Out_ne", "Out_Elat": "Out_ne",
          "Out_nh": "Out_nh", "Out_Eh": "Out_nh", "Out_Ehkin": "Out_nh"}
MIN_COUNTS = 1000
CASES = {"test_metal": 256, "config2": 512, "config5": 16, "test_cdf": 256, "opt_delta_cdf": 64, "opt_dsf": 128,
         "opt_combo": 128}


def _mixed_context(name):
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(name)
    mat.prepare(lib)
    opt.precision = 1
    ctx = tg.Context(mat, ion, opt)
    ctx.build_tables()
    return ctx


def _radial(t, k):
    a = t[k]
    if k == "Out_Elat":
        return np.cumsum(a, axis=0)      # lattice energy up to each time node, as the reference writes it out
    return a.sum(axis=(2, 3)) if a.ndim == 4 else a


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_radial_tallies_within_1pct(name):
    ctx, orc, mat, ion, opt, run, R, V = setup(name)
    nmc = CASES[name]
    say(f"[mixed] {name}: FP64 path, {nmc} iterations")
    ref = ctx.run_mc(nmc, run["seed"], R, V)
    say(f"[mixed] {name}: mixed path, {nmc} iterations")
    mctx = _mixed_context(name)
    mix = mctx.run_mc(nmc, run["seed"], R, V)
    ne_ref, ne_mix = int(ref.n_events.sum()), int(mix.n_events.sum())
    say(f"[mixed] {name}: events {ne_ref} vs {ne_mix}")
    assert abs(ne_mix - ne_ref) <= 1e-3 * ne_ref
    for k, kc in RADIAL.items():
        a, b = _radial(ref, k), _radial(mix, k)
        if not np.abs(a).any():
            assert not np.abs(b).any(), k
            continue
        l1 = float(np.abs(b - a).sum() / np.abs(a).sum())
        Nt = a.shape[0]
        nodes = sorted({Nt // 4, Nt // 2, (3 * Nt) // 4, Nt - 1})
        counts = (_radial(ref, kc) / V[None, :])[nodes]   # particles per bin, summed over iterations
        an, bn = a[nodes], b[nodes]
        peak = np.abs(an).max(axis=1, keepdims=True)
        sel = (counts >= MIN_COUNTS) & (np.abs(an) >= 1e-2 * peak) & (an != 0.0)
        assert sel.sum() > 0, k
        rel = np.where(sel, np.abs(bn - an) / np.where(an != 0.0, np.abs(an), 1.0), 0.0)
        i = np.unravel_index(np.argmax(rel), rel.shape)
        say(f"[mixed] {name} {k}: max rel {rel.max():.3e} over {int(sel.sum())} bins with >= {MIN_COUNTS} "
            f"counts at nodes {nodes} (node {nodes[i[0]]}, bin {i[1]}: {an[i]:.6e} vs {bn[i]:.6e}, "
            f"{counts[i]:.0f} counts), L1 {l1:.3e}, totals {float(a.sum()):.9e} vs {float(b.sum()):.9e}")
        assert rel.max() <= 1e-2, k
        assert l1 <= 5e-3, k
    te_a, te_b = float(ref["Out_tot_E"].sum()), float(mix["Out_tot_E"].sum())
    say(f"[mixed] {name} Out_tot_E: {te_a:.12e} vs {te_b:.12e}")
    assert abs(te_b - te_a) <= 1e-5 * abs(te_a)


def test_mixed_repeatable():
    """Fixed-point tally sums make the mixed path independent of the order in
    which lanes finish: two runs of the same iterations are identical."""
    ctx, orc, mat, ion, opt, run, R, V = setup("test_metal")
    mctx = _mixed_context("test_metal")
    a = mctx.run_mc(16, run["seed"], R, V)
    b = mctx.run_mc(16, run["seed"], R, V)
    assert np.array_equal(a.n_events, b.n_events)
    for k in list(RADIAL) + ["Out_tot_E", "Out_theta", "Out_Ee_vs_E"]:
        assert np.array_equal(a[k], b[k]), k


def test_mixed_emission_within_1pct():
    """Emission through the surface (opt_emission) on the mixed path: the
    numbers and kinetic energies of the emitted electrons at the last three
    quarter nodes within 1 % of the FP64 path's on the same 256 iterations,
    where each holds thousands of emissions."""
    ctx, orc, mat, ion, opt, run, R, V = setup("opt_emission")
    nmc = 256
    ref = ctx.run_mc(nmc, run["seed"], R, V)
    mix = _mixed_context("opt_emission").run_mc(nmc, run["seed"], R, V)
    Nt = ref["Out_Ne_Em"].shape[0]
    for k in ("Out_Ne_Em", "Out_E_Em"):
        for i in (Nt // 2, (3 * Nt) // 4, Nt - 1):
            a, b = float(ref[k][i]), float(mix[k][i])
            say(f"[mixed] opt_emission {k}[{i}]: {a:.9e} vs {b:.9e}")
            assert a > 0 and abs(b - a) <= 1e-2 * a, (k, i)
    assert ref["Out_Ne_Em"][-1] >= 1000
