"""Parity of the MI355X path with the oracle, through the C-ABI.

Tables: every energy grid, total inverse mean free path, stopping power and
cumulative transfer-energy row built by the GPU equals the oracle's bit for
bit (same nodes, same quadrature order, deterministic math on both sides).

Monte Carlo: on identical tables and seeds, every event of every particle
(kind, time, energies, position, next-event time, number of draws) equals
the oracle's bit for bit, whatever order the GPU processed the particles in;
the Out_* tallies agree to the tolerances in gpu_helpers.TALLY_TOL (the GPU
sums in 2^-40 eV fixed point, the oracle sequentially in FP64)."""
import numpy as np
import pytest

import trekis_gpu as tg
from trekis_gpu import abi, dsf

from gpu_helpers import setup, say, sort_trace, first_mismatch, compare_tallies

pytestmark = pytest.mark.gpu

SMALL = ["test", "test_metal", "test_sio2", "test_cdf", "test_polymer"]
FULL = ["config1", "config2"]       # the Al-like/Xe and Al2O3-like/Au workloads at full size
# the option branches the workloads above leave at their defaults: dispersion
# kinds 2 and 3, ion charge models 1-4, a 300 K target, the plasmon integration
# limit, logarithmic time grids, a Brandt-Kitagawa ion, emission through the
# surface, the Delta-CDF model, BEB shells, DSF elastic scattering and all of
# them together (configs.OPTION_WORKLOADS)
OPTIONS = ["opt_plasmon_pole", "opt_ritchie_warm", "opt_sg_warm", "opt_fixed_zeff", "opt_bk_ion", "opt_emission",
           "opt_delta_cdf", "opt_beb", "opt_dsf", "opt_combo"]
# table parity: test_sio2 has config5's tables (SiO2-like, U 20 MeV/u, phonon CDF, photons).
# opt_ritchie_warm's full oracle tables take minutes (a deterministic pow per
# q step); its rows are compared at sampled energies below instead.
TABLES = ["test", "test_metal", "test_cdf", "test_polymer"] + FULL + ["config5"] + \
    ["opt_plasmon_pole", "opt_sg_warm", "opt_fixed_zeff", "opt_bk_ion", "opt_delta_cdf", "opt_beb", "opt_dsf", "opt_combo"]


@pytest.mark.parametrize("name", TABLES)
def test_tables_bitwise(name):
    ctx, orc, mat, ion, opt, run, R, V = setup(name, build_oracle_tables=True)
    nsh = sum(mat.n_shells)
    kinds = [abi.TAB_E_INEL, abi.TAB_H_INEL, abi.TAB_E_EL, abi.TAB_H_EL, abi.TAB_SHI, abi.TAB_SHI_DIFF]
    if opt.include_photons:
        kinds.append(abi.TAB_PHOT)
    nrows = 0
    for kind in kinds:
        shells = [0] if kind in (abi.TAB_E_EL, abi.TAB_H_EL) else range(nsh)
        for s in shells:
            say(f"[tables] {name} kind {abi.TABLE_NAMES[kind]} shell {s}")
            Eg, Lg, Dg = ctx.table(kind, s)
            Eo, Lo, Do = orc.table(kind, s)
            assert np.array_equal(Eg, Eo), (kind, s)
            assert np.array_equal(Lg, Lo, equal_nan=True), (kind, s, np.flatnonzero(Lg != Lo)[:5])
            if kind != abi.TAB_SHI_DIFF:
                assert np.array_equal(Dg, Do, equal_nan=True), (kind, s)
            if kind in (abi.TAB_SHI, abi.TAB_SHI_DIFF, abi.TAB_PHOT):
                continue
            for i in range(len(Eg)):
                try:
                    hg, lg = ctx.row(kind, s, i)
                except tg.TrekisError:
                    continue
                ho, lo = orc.row(kind, s, i)
                assert np.array_equal(hg, ho) and np.array_equal(lg, lo), (kind, s, i)
                nrows += 1
    assert nrows > 100


def test_tables_sampled_rows_ritchie_dispersion():
    """Ritchie's extended dispersion (kind_of_DR 3, CS:382-385) at 300 K: the
    mean free path, stopping power and cumulative row of every shell at six
    energies spread over the grid, from the oracle's single-point quadrature,
    equal the GPU's table bit for bit."""
    import ctypes
    from oracle import dp
    ctx, orc, mat, ion, opt, run, R, V = setup("opt_ritchie_warm")
    nsh = sum(mat.n_shells)
    ncmp = 0
    for kind in (abi.TAB_E_INEL, abi.TAB_H_INEL):
        Eg, Lg, Dg = ctx.table(kind, 0)
        n = len(Eg)
        for i in sorted({40, n // 4, n // 2, (3 * n) // 4, n - 40, n - 1}):
            for s in range(nsh):
                try:
                    hg, lg = ctx.row(kind, s, i)
                except tg.TrekisError:
                    continue
                say(f"[tables] opt_ritchie_warm kind {abi.TABLE_NAMES[kind]} shell {s} node {i}")
                _, Ls, Ds = ctx.table(kind, s)
                L, D, cnt = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
                hw, lam = np.zeros(len(hg) + 64), np.zeros(len(hg) + 64)
                rc = orc.lib.or_inel_point(orc.h, kind, s, Eg[i], ctypes.byref(L), ctypes.byref(D), dp(hw), dp(lam),
                                           len(hw), ctypes.byref(cnt))
                assert rc == 0 and cnt.value == len(hg), (kind, s, i, rc, cnt.value, len(hg))
                assert L.value == Ls[i] and D.value == Ds[i], (kind, s, i, L.value, Ls[i], D.value, Ds[i])
                assert np.array_equal(hw[:len(hg)], hg) and np.array_equal(lam[:len(hg)], lg), (kind, s, i)
                ncmp += 1
    assert ncmp >= 12


@pytest.mark.parametrize("name", SMALL + FULL + OPTIONS)
def test_mc_replay_bitwise(name):
    ctx, orc, mat, ion, opt, run, R, V = setup(name)
    nmc = 1 if name in FULL else run["nmc"]
    say(f"[mc] {name}: GPU nmc={nmc}")
    g = ctx.run_mc(nmc, run["seed"], R, V, trace_cap=3_000_000)
    say(f"[mc] {name}: oracle")
    c = orc.run_mc(nmc, run["seed"], R, V, tg.TallySet(orc.Nt, len(R), mat), trace_cap=3_000_000)
    assert g.trace_total == len(c.trace) and len(c.trace) < 3_000_000
    assert (g.n_events == c.n_events).all(), (g.n_events, c.n_events)
    ta, tb = sort_trace(g.trace), sort_trace(c.trace)
    mm = first_mismatch(ta, tb)
    assert mm is None, mm
    bad = compare_tallies(g, c)
    assert not bad, bad


def test_mc_replay_synthetic_photon_tables():
    """The large single track's photon transport on synthetic attenuation
    tables (configs.synthetic_photoabsorption through the EPDL branch of
    Tot_Phot_IMFP, CS:855-877) instead of the CDF photon tables: the tables
    set on the device read back as set, photons are absorbed, and every
    event equals the oracle's replay on the same tables."""
    from oracle import Oracle
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material("test_sio2")
    mat.prepare(lib)
    ctx = tg.Context(mat, ion, opt)
    ctx.build_tables()
    E, L = tg.configs.apply_photon_tables(ctx, tg.configs.synthetic_photoabsorption(mat))
    for s in range(L.shape[0]):
        assert np.array_equal(ctx.table(abi.TAB_PHOT, s)[1], L[s]), s
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    orc = Oracle(mat, ion, opt)
    orc.inject_from(ctx)
    g = ctx.run_mc(run["nmc"], run["seed"], R, V, trace_cap=3_000_000)
    c = orc.run_mc(run["nmc"], run["seed"], R, V, tg.TallySet(orc.Nt, len(R), mat), trace_cap=3_000_000)
    say(f"[photons] synthetic tables: events {g.n_events.tolist()}")
    assert g.n_events[abi.EVENT_NAMES.index("photon")] > 0
    assert (g.n_events == c.n_events).all(), (g.n_events, c.n_events)
    mm = first_mismatch(sort_trace(g.trace), sort_trace(c.trace))
    assert mm is None, mm
    bad = compare_tallies(g, c)
    assert not bad, bad


@pytest.mark.parametrize("name", SMALL)
def test_mc_batching_and_replicas_exact(name):
    """Fixed-point tallies make the result independent of batch size, the
    number of tally replicas and the split of iterations over calls."""
    ctx, orc, mat, ion, opt, run, R, V = setup(name)
    base = ctx.run_mc(4, run["seed"], R, V)
    parts = [ctx.run_mc(2, run["seed"], R, V, first_iter=0), ctx.run_mc(2, run["seed"], R, V, first_iter=2)]
    opt2 = abi.Options.from_buffer_copy(opt)
    opt2.batch, opt2.tally_replicas = 1, 1
    ctx2 = tg.Context(mat, ion, opt2)
    ctx2.build_tables()
    one = ctx2.run_mc(4, run["seed"], R, V)
    again = ctx.run_mc(4, run["seed"], R, V)
    for k in base.arrays:
        assert np.array_equal(base[k], one[k]), k
        assert np.array_equal(base[k], again[k]), k
    for k in ("Out_tot_Ne", "Out_ne", "Out_nh"):
        assert np.allclose(parts[0][k] + parts[1][k], base[k], rtol=1e-13, atol=0), k
    assert ((parts[0].n_events + parts[1].n_events) == base.n_events).all()


def test_mc_replay_on_reference_radial_grid():
    """The reference's own 50 radii (1 A to 10 um, Radius_for_distributions):
    the GPU's tallies on that grid agree with the oracle's and every event
    is the same."""
    ctx, orc, mat, ion, opt, run, R0, V0 = setup("test")
    R, V = tg.reference_radial_grid(mat.layer)
    say("[mc] test on the reference radial grid")
    g = ctx.run_mc(run["nmc"], run["seed"], R, V, trace_cap=3_000_000)
    c = orc.run_mc(run["nmc"], run["seed"], R, V, tg.TallySet(orc.Nt, len(R), mat), trace_cap=3_000_000)
    assert (g.n_events == c.n_events).all(), (g.n_events, c.n_events)
    mm = first_mismatch(sort_trace(g.trace), sort_trace(c.trace))
    assert mm is None, mm
    bad = compare_tallies(g, c)
    assert not bad, bad


def test_input_tree_replay_bitwise():
    """A run set up from an input tree laid out like the reference's
    (trekis_gpu/examples: INPUT_PARAMETERS.txt, INPUT_CDF/, INPUT_DOS/) on
    the reference's radial grid: every event equals the oracle's and the
    tallies agree."""
    import os
    from oracle import Oracle
    from trekis_gpu import inputs
    lib = tg.load_library()
    path = os.path.join(os.path.dirname(os.path.abspath(tg.__file__)), "examples", "INPUT_PARAMETERS.txt")
    # radiative times per (element, shell), as the reference would take them
    # from EADL: the Al L shell and O K shell decay mostly radiatively here
    mat, ion, opt, nmc = inputs.load_inputs(path, radiat=[[5.0, 1.0e23], [2.0]])
    mat.prepare(lib)
    say(f"[mc] input tree: {mat.name}, Z={ion.Zat}, {nmc} iterations")
    ctx = tg.Context(mat, ion, opt)
    ctx.build_tables()
    orc = Oracle(mat, ion, opt)
    orc.inject_from(ctx)
    R, V = tg.reference_radial_grid(mat.layer)
    g = ctx.run_mc(nmc, 0x5EED0008, R, V, trace_cap=3_000_000)
    c = orc.run_mc(nmc, 0x5EED0008, R, V, tg.TallySet(orc.Nt, len(R), mat), trace_cap=3_000_000)
    assert (g.n_events == c.n_events).all(), (g.n_events, c.n_events)
    assert g.n_events[abi.EVENT_NAMES.index("radiative")] > 0
    mm = first_mismatch(sort_trace(g.trace), sort_trace(c.trace))
    assert mm is None, mm
    bad = compare_tallies(g, c)
    assert not bad, bad


def test_mc_replay_partly_staged():
    """With most tables left in HBM (a small LDS budget), the transport kernel
    that tests each table for staging replays the oracle bit for bit too."""
    import os
    ctx, orc, mat, ion, opt, run, R, V = setup("test")
    os.environ["TRK_LDS_LIMIT_BYTES"] = "16384"
    try:
        say("[mc] test: GPU with a 16 KB LDS budget")
        g = ctx.run_mc(run["nmc"], run["seed"], R, V, trace_cap=3_000_000)
    finally:
        del os.environ["TRK_LDS_LIMIT_BYTES"]
    c = orc.run_mc(run["nmc"], run["seed"], R, V, tg.TallySet(orc.Nt, len(R), mat), trace_cap=3_000_000)
    assert g.trace_total == len(c.trace)
    mm = first_mismatch(sort_trace(g.trace), sort_trace(c.trace))
    assert mm is None, mm
    assert not compare_tallies(g, c)


def test_error_paths():
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material("test")
    mat.prepare(lib)
    ctx = tg.Context(mat, ion, opt)
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    with pytest.raises(tg.TrekisError, match="call out of order"):
        ctx.run_mc(1, 1, R, V)
    ctx.build_tables()
    with pytest.raises(tg.TrekisError, match="invalid argument"):
        ctx.run_mc(1, 1, R, V, tallies=tg.TallySet(ctx.Nt, len(R) + 1, mat))
    # a generation larger than the particle lists: the kernels drop the
    # overflow, finish normally and the call reports it
    opt2 = abi.Options.from_buffer_copy(opt)
    opt2.gen_capacity = 64
    ctx2 = tg.Context(mat, ion, opt2)
    ctx2.build_tables()
    with pytest.raises(tg.TrekisError, match="capacity"):
        ctx2.run_mc(1, 1, R, V)
    bad = abi.Options.from_buffer_copy(opt)
    bad.precision = 3
    with pytest.raises(tg.TrekisError, match="invalid argument"):
        tg.Context(mat, ion, bad)


def test_dsf_error_paths():
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material("opt_dsf")
    mat.prepare(lib)
    tables = mat.dsf
    mat.dsf = None
    ctx = tg.Context(mat, ion, opt)
    # DSF elastic scattering without its tables
    with pytest.raises(tg.TrekisError, match="DSF"):
        ctx.build_tables()
    with pytest.raises(tg.TrekisError, match="invalid"):
        ctx.set_dsf(2, tables[0])
    bad = dsf.DsfTable(tables[0].E[::-1].copy(), tables[0].dE, tables[0].dL, tables[0].dL_emit, tables[0].dL_absorb)
    with pytest.raises(tg.TrekisError, match="increasing"):
        ctx.set_dsf(0, bad)
    neg = dsf.DsfTable(tables[0].E, tables[0].dE, -tables[0].dL, tables[0].dL_emit, tables[0].dL_absorb)
    with pytest.raises(tg.TrekisError, match="positive"):
        ctx.set_dsf(0, neg)
    ctx.set_dsf(0, tables[0])
    ctx.set_dsf(1, tables[1])
    ctx.build_tables()
    with pytest.raises(tg.TrekisError, match="before trk_build_tables"):
        ctx.set_dsf(0, tables[0])
    # a context without DSF scattering takes no DSF tables
    mat2, ion2, opt2, _ = tg.config_material("test")
    mat2.prepare(lib)
    with pytest.raises(tg.TrekisError, match="kind_of_EMFP = 2"):
        tg.Context(mat2, ion2, opt2).set_dsf(0, tables[0])
