"""Shared set-up of the GPU parity tests: one Context and one oracle per
configuration, the oracle replaying on the tables the GPU built."""
import sys

import numpy as np

import trekis_gpu as tg
from trekis_gpu import abi
from oracle import Oracle

_CACHE = {}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(name, build_oracle_tables=False):
    """(ctx, oracle, mat, ion, opt, run, R, V) with GPU tables built; the oracle
    either builds its own tables (for table parity) or replays on the GPU's."""
    key = (name, build_oracle_tables)
    if key in _CACHE:
        return _CACHE[key]
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(name)
    mat.prepare(lib)
    say(f"[setup] {name}: building GPU tables")
    ctx = tg.Context(mat, ion, opt)
    ctx.build_tables()
    orc = Oracle(mat, ion, opt)
    if build_oracle_tables:
        say(f"[setup] {name}: building oracle tables")
        orc.build_tables(16)
    else:
        orc.inject_from(ctx)
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    out = (ctx, orc, mat, ion, opt, run, R, V)
    _CACHE[key] = out
    return out


def sort_trace(tr):
    ids = tr[:, 1].copy().view(np.uint64)
    key = np.lexsort((tr[:, 2], ids, tr[:, 0]))
    return tr[key]


def first_mismatch(a, b):
    """Index and field of the first differing record of two sorted traces."""
    n = min(len(a), len(b))
    neq = ~((a[:n] == b[:n]) | (np.isnan(a[:n]) & np.isnan(b[:n])))
    rows = np.flatnonzero(neq.any(axis=1))
    if len(rows) == 0:
        return None
    i = rows[0]
    return i, np.flatnonzero(neq[i]).tolist(), a[i].tolist(), b[i].tolist()


TALLY_TOL = {  # relative tolerance per tally (fixed-point 2^-40 eV sums on the GPU vs sequential FP64 sums)
    "Out_ne": 1e-12, "Out_nh": 1e-12, "Out_nphot": 1e-12, "Out_tot_Ne": 1e-12, "Out_tot_Nphot": 1e-12,
    "Out_Ee": 1e-10, "Out_Eh": 1e-10, "Out_Ehkin": 1e-10, "Out_Ephot": 1e-10, "Out_Elat": 1e-10,
    "Out_tot_E": 1e-10, "Out_E_e": 1e-10, "Out_E_phot": 1e-10, "Out_E_at": 1e-10, "Out_E_h": 1e-10,
    # normalised per iteration: count * (w / n) on the GPU, n additions of w / n in the oracle
    "Out_Ee_vs_E": 1e-11, "Out_Eh_vs_E": 1e-11, "Out_theta": 1e-11, "Out_theta_h": 1e-11, "Out_diff_coeff": 1e-11,
    # emitted electrons: count, kinetic energy sum, spectrum normalised per iteration
    "Out_Ne_Em": 1e-12, "Out_E_Em": 1e-10, "Out_Ee_vs_E_Em": 1e-11,
}


def compare_tallies(g, c):
    bad = []
    for k, rt in TALLY_TOL.items():
        a, b = g[k], c[k]
        scale = max(float(np.abs(b).max()), 1e-300)
        if not np.allclose(a, b, rtol=rt, atol=rt * scale):
            i = np.unravel_index(np.argmax(np.abs(a - b)), a.shape)
            bad.append((k, i, float(a[i]), float(b[i])))
    return bad
