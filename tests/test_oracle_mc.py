"""Invariants of the oracle's Monte Carlo (do_Monte_Carlo restated in the
reference's global time order, Monte_Carlo.f90:395-640).

What pins it without the reference binary: energy conservation (the total
energy of electrons, holes, photons and atoms at every time node equals the
energy the ion has lost by then, MC:2155-2249 and the ledger of
Save_data_for_distributions), particle-number conservation (every
ionisation, Auger decay and photoabsorption makes one electron and one hole,
so Sum_j ne_j/V_j == Sum_j nh_j/V_j == Out_tot_Ne), the reference's event
order (non-decreasing event times within an iteration), per-particle event
numbering, and bitwise determinism of a seeded run.  With emission through
the surface (opt_emission), the rules of calculate_emission (MC:2477-2513)
are read back from the event trace: which collisions outside the surface
emit, the emitted fraction against the barrier's transmission, and the
emitted-electron tallies formed from the emissions the trace records."""
import ctypes
import math

import numpy as np
import pytest

import trekis_gpu as tg
from oracle import Oracle, dp

CONFIGS = ["test", "test_metal", "opt_emission", "opt_delta_cdf", "opt_beb", "opt_dsf", "opt_combo"]


@pytest.fixture(scope="module", params=CONFIGS)
def runs(request, lib):
    mat, ion, opt, run = tg.config_material(request.param)
    mat.prepare(lib)
    o = Oracle(mat, ion, opt)
    o.build_tables(8, shi_window=(ion.E * 0.8, ion.E * 1.01))
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    a = o.run_mc(run["nmc"], run["seed"], R, V, tg.TallySet(o.Nt, len(R), mat), trace_cap=3_000_000)
    b = o.run_mc(run["nmc"], run["seed"], R, V, tg.TallySet(o.Nt, len(R), mat), trace_cap=3_000_000)
    tim = np.zeros(o.Nt)
    n = ctypes.c_int32()
    o.lib.or_time_nodes(o.h, ctypes.byref(n), dp(tim), o.Nt)
    return dict(o=o, mat=mat, ion=ion, run=run, R=R, V=V, a=a, b=b, tim=tim, name=request.param)


def test_deterministic(runs):
    a, b = runs["a"], runs["b"]
    for k in a.arrays:
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    assert np.array_equal(a.trace, b.trace)
    assert (a.n_events == b.n_events).all()


def test_trace_is_complete_and_time_ordered(runs):
    a = runs["a"]
    tr = a.trace
    assert len(tr) == a.n_events.sum()
    for kind in range(8):
        assert (tr[:, 3] == kind).sum() == a.n_events[kind]
    for it in np.unique(tr[:, 0]):
        t = tr[tr[:, 0] == it, 4]
        assert np.all(np.diff(t) >= 0.0)
        assert t.max() < runs["tim"][-1]
    # each particle numbers its events 0, 1, 2, ... in time order
    ids = tr[:, 1].view(np.uint64)
    key = np.lexsort((tr[:, 2], ids, tr[:, 0]))
    ids_s, ev_s, it_s = ids[key], tr[key, 2], tr[key, 0]
    start = np.ones(len(key), bool)
    start[1:] = (ids_s[1:] != ids_s[:-1]) | (it_s[1:] != it_s[:-1])
    grp = np.cumsum(start) - 1
    first = np.flatnonzero(start)
    assert np.array_equal(ev_s, np.arange(len(key)) - first[grp])


def test_energy_conservation(runs):
    a, ion, tim = runs["a"], runs["ion"], runs["tim"]
    tr = a.trace
    lost = np.zeros(len(tim))
    for it in np.unique(tr[:, 0]):
        s = tr[(tr[:, 0] == it) & (tr[:, 3] == 0)]
        s = s[np.argsort(s[:, 2])]
        dep = -np.diff(np.concatenate([[ion.E], s[:, 5]]))
        for i, t in enumerate(tim):
            lost[i] += dep[s[:, 4] < t].sum()
    assert lost[-1] > 0
    np.testing.assert_allclose(a["Out_tot_E"], lost, rtol=1e-10)
    led = a["Out_E_e"] + a["Out_E_h"].sum(axis=(1, 2)) + a["Out_E_phot"] + a["Out_E_at"]
    np.testing.assert_allclose(led, a["Out_tot_E"], rtol=1e-12)


def test_number_conservation(runs):
    a, V = runs["a"], runs["V"]
    ne = (a["Out_ne"] / V[None, :]).sum(axis=1)
    nh = (a["Out_nh"] / V[None, :, None, None]).sum(axis=(1, 2, 3))
    np.testing.assert_allclose(ne, a["Out_tot_Ne"], rtol=1e-12)
    np.testing.assert_allclose(nh, ne, rtol=1e-12)
    ev = a.n_events
    # electrons made: one per SHI ionisation, electron or hole impact ionisation,
    # Auger decay and photoabsorption; none is destroyed (no emission)
    assert a["Out_tot_Ne"][-1] == ev[0] + ev[1] + ev[3] + ev[5] + ev[7]


def test_seed_changes_history(runs):
    o, run, R, V, mat = runs["o"], runs["run"], runs["R"], runs["V"], runs["mat"]
    c = o.run_mc(1, run["seed"] + 1, R, V, tg.TallySet(o.Nt, len(R), mat))
    d = o.run_mc(1, run["seed"], R, V, tg.TallySet(o.Nt, len(R), mat))
    assert not np.array_equal(c["Out_Ee"], d["Out_Ee"])


def test_iterations_are_independent(runs):
    """Iteration k of a run equals a run started at first_iter = k (the unit
    the GPU batches over)."""
    o, run, R, V, mat = runs["o"], runs["run"], runs["R"], runs["V"], runs["mat"]
    whole = o.run_mc(2, run["seed"], R, V, tg.TallySet(o.Nt, len(R), mat))
    p0 = o.run_mc(1, run["seed"], R, V, tg.TallySet(o.Nt, len(R), mat))
    p1 = o.run_mc(1, run["seed"], R, V, tg.TallySet(o.Nt, len(R), mat), first_iter=1)
    for k in ("Out_ne", "Out_Ee", "Out_tot_E", "Out_Elat", "Out_Eh"):
        np.testing.assert_allclose(p0[k] + p1[k], whole[k], rtol=1e-11, atol=1e-9)
    assert ((p0.n_events + p1.n_events) == whole.n_events).all()


def test_threaded_run_matches_sequential(runs):
    """The all-core CPU timing path (OpenMP over iterations, private tallies)
    gives the sequential run's events and tallies up to summation order."""
    o, run, R, V, mat = runs["o"], runs["run"], runs["R"], runs["V"], runs["mat"]
    a = o.run_mc(4, run["seed"], R, V, tg.TallySet(o.Nt, len(R), mat))
    b = o.run_mc_threads(4, run["seed"], R, V, tg.TallySet(o.Nt, len(R), mat), threads=4)
    assert (a.n_events == b.n_events).all()
    for k in ("Out_ne", "Out_Ee", "Out_nh", "Out_Elat", "Out_tot_E", "Out_theta"):
        np.testing.assert_allclose(b[k], a[k], rtol=1e-11, atol=1e-11 * abs(a[k]).max())


def test_emission_rules(runs):
    """Electron collisions at Z < 0 emit always at E >= 1.5 H, never at
    E <= W, and in between with the logistic transmission (the emitted
    count within 4 sigma of the summed probabilities); collisions inside
    never emit; an emitted electron has no later event; Out_Ne_Em and
    Out_E_Em at each node count and sum E - W over the emissions before it,
    and every iteration with an emission adds exactly one to the integral
    of Out_Ee_vs_E_Em over the energy grid Out_R/10."""
    if not runs["mat"].work_function > 0.0:
        pytest.skip("emission is on only in opt_emission and opt_combo")
    a, mat, R, tim = runs["a"], runs["mat"], runs["R"], runs["tim"]
    W, H = mat.work_function, mat.bar_height
    E1, g = ctypes.c_double(), ctypes.c_double()
    assert runs["o"].lib.or_barrier_parameters(W, mat.bar_length, H, ctypes.byref(E1), ctypes.byref(g)) == 0
    tr = a.trace
    el = tr[(tr[:, 3] == 1) | (tr[:, 3] == 2)]
    em = el[:, 10] == 1e30
    out = el[:, 9] < 0.0
    assert not em[~out].any()
    E = el[:, 5]
    assert em[out & (E >= 1.5 * H)].all() and (out & (E >= 1.5 * H)).sum() > 10
    assert not em[out & (E <= W)].any()
    mid = out & (E > W) & (E < 1.5 * H)
    P = 1.0 / (1.0 + np.exp(g.value * (E1.value - E[mid])))
    assert mid.sum() > 50
    assert abs(em[mid].sum() - P.sum()) < 4.0 * math.sqrt((P * (1 - P)).sum()) + 1.0
    # no event of an emitted electron after its emission
    ids = el[:, 1].view(np.uint64)
    for row in el[em]:
        same = (el[:, 0] == row[0]) & (ids == row[1:2].view(np.uint64)[0])
        assert el[same, 2].max() == row[2]
    RE = R / 10.0
    width = np.concatenate([[RE[0]], np.diff(RE)])
    for i, t in enumerate(tim):
        before = em & (el[:, 4] < t)
        assert a["Out_Ne_Em"][i] == before.sum()
        np.testing.assert_allclose(a["Out_E_Em"][i], (E[before] - W).sum(), rtol=1e-12, atol=0)
        n_iter = len(np.unique(el[before, 0]))
        np.testing.assert_allclose((a["Out_Ee_vs_E_Em"][i] * width).sum(), n_iter, rtol=1e-12)
    assert a["Out_Ne_Em"][-1] > 0
