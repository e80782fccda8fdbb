"""Batches: many ion impacts -- different ions and energies on two materials
-- transported in one persistent launch (trk_batch_*), replayed from a
hipGraph, with the launch shape chosen by the autotuner.

The FP64 batch must reproduce, for every impact, the oracle's run of that
impact alone (its iteration first_iter + i, on that impact's tables): every
event bit for bit and the tallies to the oracle tolerances.  The mixed batch
is compared with the FP64 batch on the same iterations by the radial-tally
rule of the mixed-precision tests."""
import numpy as np
import pytest

import trekis_gpu as tg
from trekis_gpu import abi
from trekis_gpu.batch import Batch, IonContext

from gpu_helpers import say, sort_trace, first_mismatch, TALLY_TOL

pytestmark = pytest.mark.gpu

# (material workload, Z, mass, MeV/u): two materials (the Al2O3-like oxide
# with Mott elastic scattering and photons, the Al-like metal with the
# phonon CDF), four ions each; electron grids end at 2 keV, so the energies
# stay below ~0.8 MeV/u
IMPACTS = [("test", 8, 16.00, 0.5), ("test", 6, 12.011, 0.3), ("test", 7, 14.007, 0.8), ("test", 2, 4.0026, 0.6),
           ("test_metal", 8, 16.00, 0.5), ("test_metal", 3, 6.94, 0.5), ("test_metal", 6, 12.011, 0.7),
           ("test_metal", 10, 20.18, 0.4)]
BATCH_TALLIES = [k for k in TALLY_TOL if k not in ("Out_Ee_vs_E", "Out_Eh_vs_E", "Out_theta", "Out_theta_h",
                                                    "Out_diff_coeff")]
_CACHE = {}


def _impacts(precision):
    if precision in _CACHE:
        return _CACHE[precision]
    lib = tg.load_library()
    mats = {}
    for name in dict.fromkeys(m for m, *_ in IMPACTS):
        mat, ion, opt, run = tg.config_material(name)
        mat.prepare(lib)
        opt.precision = precision
        ctx = tg.Context(mat, ion, opt)
        ctx.build_tables()
        mats[name] = (ctx, mat, opt, run)
    imps = []
    for name, Z, M, Eu in IMPACTS:
        ion = abi.Ion()
        ion.Zat, ion.Mass, ion.E, ion.Kind_Zeff, ion.fixed_Zeff, ion.Kind_ion = Z, M, Eu * 1e6 * M, 0, float(Z), 0
        c = IonContext(mats[name][0], ion)
        c.build_tables()
        imps.append(c)
    run = mats["test"][3]
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], 100.0)
    _CACHE[precision] = (mats, imps, R, V)
    return _CACHE[precision]


def test_ion_context_tables_equal_a_full_context():
    """An ion context shares the material tables and builds its own ion
    tables, equal bit for bit to those of a full context of that ion."""
    mats, imps, R, V = _impacts(0)
    lib = tg.load_library()
    mat, ion0, opt, run = tg.config_material("test")
    mat.prepare(lib)
    c = imps[2]
    full = tg.Context(mat, c.ion, opt)
    full.build_tables()
    nsh = sum(mat.n_shells)
    for kind in (abi.TAB_E_INEL, abi.TAB_H_INEL, abi.TAB_E_EL, abi.TAB_SHI, abi.TAB_SHI_DIFF):
        for s in ([0] if kind == abi.TAB_E_EL else range(nsh)):
            a, b = c.table(kind, s), full.table(kind, s)
            for x, y in zip(a, b):
                assert np.array_equal(x, y, equal_nan=True), (kind, s)
    assert np.array_equal(*[x.row(abi.TAB_E_INEL, 0, 150)[1] for x in (c, full)])


def test_batch_fp64_each_impact_equals_its_oracle_run():
    from oracle import Oracle
    mats, imps, R, V = _impacts(0)
    b = Batch(imps, R, V, capacity=1 << 20)
    cfg = b.config()
    say(f"[batch] {len(imps)} impacts in {cfg['groups']} groups, {cfg['workgroups']} workgroups of {cfg['block']}")
    assert cfg["groups"] == 2
    first, seed = 1000, 0x5EED0040
    tal = b.run(seed, first, trace_cap=2_000_000)
    ms, ev, rounds = b.timing()
    say(f"[batch] one launch: {ms:.2f} ms, {rounds} rounds, events {ev.tolist()}")
    assert rounds > 1 and ev.sum() > 0 and b.trace_total == len(b.trace)
    tr = b.trace
    total_events = np.zeros(8, dtype=np.int64)
    for i, c in enumerate(imps):
        name = IMPACTS[i][0]
        ctx, mat, opt, run = mats[name]
        orc = Oracle(mat, c.ion, opt)
        orc.inject_from(c)
        o = orc.run_mc(1, seed, R, V, tg.TallySet(orc.Nt, len(R), mat), first_iter=first + i, trace_cap=2_000_000)
        mine = tr[tr[:, 0] == first + i]
        say(f"[batch] impact {i} ({name}, Z={c.ion.Zat}): {len(mine)} events, oracle {len(o.trace)}")
        assert len(mine) == len(o.trace), i
        mm = first_mismatch(sort_trace(mine), sort_trace(o.trace))
        assert mm is None, (i, mm)
        kinds = np.bincount(mine[:, 3].astype(int), minlength=8)
        assert (kinds == o.n_events).all(), (i, kinds, o.n_events)
        total_events += kinds
        for k in BATCH_TALLIES:
            a, r = tal[i][k], o[k]
            rt = TALLY_TOL[k]
            scale = max(float(np.abs(r).max()), 1e-300)
            assert np.allclose(a, r, rtol=rt, atol=rt * scale), (i, k)
    assert (total_events == ev).all()
    # a second run replays the captured graph: same results; a run without
    # host tallies leaves them on the device for trk_batch_fetch
    tal2 = b.run(seed, first)
    b.run(seed, first, with_tallies=False)
    tal3 = b.fetch()
    for i in range(len(imps)):
        for k in BATCH_TALLIES:
            assert np.array_equal(tal[i][k], tal2[i][k]), (i, k)
            assert np.array_equal(tal[i][k], tal3[i][k]), (i, k)


def test_batch_mixed_within_1pct_of_fp64():
    """The mixed batch against the FP64 batch on the same 8 x 24 iterations:
    radial tallies summed over each material's impacts, by the rule of the
    mixed-precision tests (bins with >= 1000 particles and >= 1% of the
    node's peak, at four time nodes): within 1%."""
    ref_m, ref_i, R, V = _impacts(0)
    mix_m, mix_i, _, _ = _impacts(1)
    bf, bm = Batch(ref_i, R, V, capacity=1 << 20), Batch(mix_i, R, V, capacity=1 << 20)
    groups = {}
    for i, (name, *_r) in enumerate(IMPACTS):
        groups.setdefault(name, []).append(i)
    acc = {}
    for rep in range(24):
        first = 5000 + rep * len(IMPACTS)
        tf, tm = bf.run(0x5EED0041, first), bm.run(0x5EED0041, first)
        for name, idx in groups.items():
            for k in ("Out_ne", "Out_Ee", "Out_nh", "Out_Eh", "Out_Ehkin", "Out_Elat"):
                for which, t in (("f", tf), ("m", tm)):
                    s = sum(t[i][k] for i in idx)
                    acc[(name, k, which)] = acc.get((name, k, which), 0) + s
    worst = 0.0
    for name in groups:
        for k, kc in (("Out_ne", "Out_ne"), ("Out_Ee", "Out_ne"), ("Out_Elat", "Out_ne"), ("Out_nh", "Out_nh"),
                      ("Out_Eh", "Out_nh"), ("Out_Ehkin", "Out_nh")):
            def rad(x):
                return np.cumsum(x, axis=0) if k == "Out_Elat" else (x.sum(axis=(2, 3)) if x.ndim == 4 else x)
            a, m = rad(acc[(name, k, "f")]), rad(acc[(name, k, "m")])
            ct = acc[(name, kc, "f")]
            ct = (ct.sum(axis=(2, 3)) if ct.ndim == 4 else ct) / V[None, :]
            Nt = a.shape[0]
            nodes = sorted({Nt // 4, Nt // 2, (3 * Nt) // 4, Nt - 1})
            an, mn = a[nodes], m[nodes]
            peak = np.abs(an).max(axis=1, keepdims=True)
            sel = (ct[nodes] >= 1000) & (np.abs(an) >= 1e-2 * peak) & (an != 0.0)
            if not sel.any():
                continue
            rel = float((np.abs(mn - an)[sel] / np.abs(an)[sel]).max())
            say(f"[batch mixed] {name} {k}: max rel {rel:.3e} over {int(sel.sum())} bins")
            worst = max(worst, rel)
            assert rel <= 1e-2, (name, k, rel)
    assert worst > 0.0


def test_batch_autotuner_records_its_choice():
    mats, imps, R, V = _impacts(1)
    b = Batch(imps, R, V, capacity=1 << 20)
    cfg = b.tune(0x5EED0042, 9000)
    log = b.tune_log()
    say(f"[batch tune] chose {cfg}; timings {log}")
    assert len(log) == 9
    best = min(log, key=lambda e: e["ms"])
    assert (cfg["block"], cfg["round_events"]) == (best["block"], best["round_events"])
    t = b.run(0x5EED0042, 9000)
    assert sum(float(x["Out_tot_Ne"].sum()) for x in t) > 0


@pytest.mark.parametrize("name", ["opt_delta_cdf", "opt_dsf"])
def test_batch_refuses_the_general_transfers(name):
    """Impacts whose material uses the Delta-CDF model (kind_of_DR = 4) or
    DSF elastic scattering run through trk_run_mc; a batch refuses them with
    the reason."""
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(name)
    mat.prepare(lib)
    ctx = tg.Context(mat, ion, opt)
    ctx.build_tables()
    c = IonContext(ctx, ion)
    c.build_tables()
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], 100.0)
    with pytest.raises(tg.TrekisError, match="batches do not"):
        Batch([c], R, V, capacity=1 << 16)


@pytest.mark.parametrize("name", ["opt_dsf", "opt_beb", "opt_delta_cdf", "opt_emission", "opt_combo"])
def test_ion_context_on_general_material_runs_like_a_full_context(name):
    """An ion context on a material with DSF tables, BEB shells or the
    Delta-CDF model shares those tables (and the material's model data):
    trk_run_mc on it gives, bit for bit, the tallies of a full context of
    the same ion."""
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(name)
    mat.prepare(lib)
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    full = tg.Context(mat, ion, opt)
    full.build_tables()
    other = abi.Ion.from_buffer_copy(ion)
    other.E = ion.E * 0.5
    base = tg.Context(mat, other, opt)
    base.build_tables()
    c = IonContext(base, ion)
    c.build_tables()
    say(f"[ion ctx] {name}: full context and ion context, {run['nmc']} iterations")
    a = full.run_mc(run["nmc"], run["seed"], R, V)
    b = c.run_mc(run["nmc"], run["seed"], R, V)
    assert (a.n_events == b.n_events).all() and a.n_events.sum() > 0
    for k in a.arrays:
        assert np.array_equal(a[k], b[k]), k
