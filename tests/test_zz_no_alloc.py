"""trk_run_mc_noalloc, the entry point that allocates no device memory (its
buffers are carved from a static device array), gives the results of
trk_run_mc on the same context: every Out_* tally and event count equal, on
two calls in a row, the second one smaller (fewer iterations, fewer radial
bins), so that what the first call left in the array must not leak into the
second."""
import numpy as np
import pytest

import trekis_gpu as tg
from trekis_gpu import abi

from gpu_helpers import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["test", "test_metal", "opt_emission", "opt_delta_cdf", "opt_beb", "opt_dsf", "opt_combo"])
def test_noalloc_equals_allocating_entry_point(name):
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(name)
    mat.prepare(lib)
    ctx = tg.Context(mat, ion, opt)
    ctx.build_tables()
    calls = [(run["nmc"], run["NR"], 0), (max(1, run["nmc"] // 2), run["NR"] // 2, 3)]
    for nmc, NR, first in calls:
        R, V = tg.radial_grid(NR, run["Rmin"], run["Rmax"], mat.layer)
        say(f"[noalloc] {name}: nmc={nmc} NR={NR} first_iter={first}")
        a = ctx.run_mc(nmc, run["seed"], R, V, first_iter=first)
        b = ctx.run_mc(nmc, run["seed"], R, V, first_iter=first, noalloc=True)
        assert (a.n_events == b.n_events).all(), (a.n_events, b.n_events)
        assert a.n_events.sum() > 0
        for k in a.arrays:
            assert np.array_equal(a[k], b[k]), (k, float(np.abs(a[k] - b[k]).max()))


def test_noalloc_refuses_what_does_not_fit():
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material("test")
    mat.prepare(lib)
    opt2 = abi.Options.from_buffer_copy(opt)
    opt2.gen_capacity = 1 << 22          # lists alone need ~2 GB, far above the scratch array
    ctx = tg.Context(mat, ion, opt2)
    ctx.build_tables()
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    with pytest.raises(tg.TrekisError, match="capacity exceeded"):
        ctx.run_mc(1, run["seed"], R, V, noalloc=True)
    with pytest.raises(tg.TrekisError, match="call out of order"):
        ctx.run_mc(1, run["seed"], R, V, noalloc=True, trace_cap=10)
