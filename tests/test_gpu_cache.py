"""Device tables seeded from the reference's cache files (trk_table_set,
trk_row_set, trk_tables_finish behind trekis_gpu.cache.seed_from_cache).

A context builds its tables and writes them in the reference's layout; a
second context of the same material builds its tables, has every mean free
path and every cumulative row overwritten with other values (scaled by 1.5)
so that its runs differ, is seeded from the files, and then holds the first
one's tables bit for bit -- every mean free path, stopping power of the ion
and cumulative row -- and runs the same histories: every tally equal bit for
bit to the first context's run.  A run between a set and the finish is
refused.
"""
import numpy as np
import pytest

import trekis_gpu as tg
from trekis_gpu import abi, cache

from gpu_helpers import say

pytestmark = pytest.mark.gpu


def _ctx(name, quadrature):
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(name)
    mat.prepare(lib)
    opt.quadrature = quadrature
    ctx = tg.Context(mat, ion, opt)
    ctx.build_tables()
    return ctx, mat, run


def _all_tables(ctx, nsh):
    out = {}
    kinds = [abi.TAB_E_INEL, abi.TAB_H_INEL, abi.TAB_SHI] + ([abi.TAB_PHOT] if ctx.opt.include_photons else [])
    for kind in kinds:
        for s in range(nsh):
            out[(kind, s)] = ctx.table(kind, s)
    for kind in (abi.TAB_E_EL, abi.TAB_H_EL):
        out[(kind, 0)] = ctx.table(kind, 0)
    return out


@pytest.mark.parametrize("name", ["test_cdf", "test"])
def test_seeded_tables_and_run_equal_the_writer(name, tmp_path):
    a, mat, run = _ctx(name, 0)
    b, _, _ = _ctx(name, 0)
    nsh = sum(mat.n_shells)
    say(f"[cache] {name}: writing the reference's cache files")
    names = cache.write_cache(a, str(tmp_path))
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    nmc = 4
    ra = a.run_mc(nmc, run["seed"], R, V)
    # other tables in b: every mean free path and cumulative row scaled
    for (kind, s), (E, L, D) in _all_tables(b, nsh).items():
        b.set_table(kind, s, L=1.5 * L)
    for kind, s, i, _ in cache._row_files(b, names, str(tmp_path)):
        hw, lam = b.row(kind, s, i)
        b.set_row(kind, s, i, hw, 1.5 * lam)
    b.finish_tables()
    ta, tb = _all_tables(a, nsh), _all_tables(b, nsh)
    assert all(not np.array_equal(ta[k][1], tb[k][1]) for k in ta if np.isfinite(ta[k][1]).any())
    rx = b.run_mc(nmc, run["seed"], R, V)
    assert not all(np.array_equal(ra[f], rx[f]) for f in ra.arrays), "scaled tables should change the run"
    say(f"[cache] {name}: seeding a context whose tables were scaled")
    cache.seed_from_cache(b, str(tmp_path))
    tb = _all_tables(b, nsh)
    for k in ta:
        assert np.array_equal(ta[k][0], tb[k][0]), k
        assert np.array_equal(ta[k][1], tb[k][1]), k
        if k[0] == abi.TAB_SHI:
            assert np.array_equal(ta[k][2], tb[k][2]), k
    rows = 0
    for kind, s, i, _ in cache._row_files(a, names, str(tmp_path)):
        ha, la = a.row(kind, s, i)
        hb, lb = b.row(kind, s, i)
        assert np.array_equal(ha, hb) and np.array_equal(la, lb), (kind, s, i)
        rows += 1
    say(f"[cache] {name}: {len(ta)} tables and {rows} rows equal")
    rb = b.run_mc(nmc, run["seed"], R, V)
    for f in ra.arrays:
        assert np.array_equal(ra[f], rb[f]), f
    assert float(ra["Out_tot_Ne"].sum()) > 0


def test_run_between_set_and_finish_is_refused():
    a, mat, run = _ctx("test", 0)
    E, L, _ = a.table(abi.TAB_E_INEL, 0)
    a.set_table(abi.TAB_E_INEL, 0, L=L, E=E)
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    with pytest.raises(tg.TrekisError, match="trk_tables_finish"):
        a.run_mc(1, run["seed"], R, V)
    with pytest.raises(tg.TrekisError, match="differs from the built grid"):
        a.set_table(abi.TAB_E_INEL, 0, L=L, E=E * (1 + 1e-12))
    with pytest.raises(tg.TrekisError, match="energy nodes"):
        a.set_table(abi.TAB_E_INEL, 0, L=L[:-1])
    a.finish_tables()
    t = a.run_mc(1, run["seed"], R, V)
    assert float(t["Out_tot_Ne"].sum()) > 0
