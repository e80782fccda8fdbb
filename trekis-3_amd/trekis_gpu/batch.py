"""Batches of ion impacts in one persistent launch (trk_batch_* of the C-ABI).

The reference runs one ion on one target per program run.  A batch runs
many impacts -- different ions and energies on one or more materials, one
Monte Carlo iteration each -- in a single launch of the transport kernel:

    mat_ctx = Context(material, ion, options); mat_ctx.build_tables()
    impacts = [IonContext(mat_ctx, ion_i) for ion_i in ions]   # ion tables only
    b = Batch(impacts, Out_R, Out_V)
    b.tune(seed, first_iter)            # threads per workgroup, round depth
    tallies = b.run(seed, first_iter)   # one TallySet per impact

Impact i is iteration first_iter + i of its own run (its random streams are
that iteration's), so its tallies and events equal those of a run of that
impact alone (Context.run_mc(1, seed, first_iter=first_iter + i)).  The
per-iteration normalised tallies (spectra, angle histograms, diffusion
coefficient) are made by Context.run_mc only.
"""
import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint64, c_void_p

import numpy as np

from . import abi
from . import Context, TallySet, TrekisError, dptr, load_library


def _declare(lib):
    if getattr(lib, "_batch_declared", False):
        return lib
    P = POINTER
    sig = {
        "trk_ctx_create_ion": (c_int, [c_void_p, P(abi.Ion), P(c_void_p)]),
        "trk_batch_create": (c_int, [P(c_void_p), c_int32, abi.DP, abi.DP, c_int32, c_int32, P(c_void_p)]),
        "trk_batch_destroy": (None, [c_void_p]),
        "trk_batch_error": (c_int, [c_void_p, c_char_p, c_size_t]),
        "trk_batch_set_trace": (c_int, [c_void_p, abi.DP, c_int64, P(c_int64)]),
        "trk_batch_run": (c_int, [c_void_p, c_uint64, c_int64, P(P(abi.Tallies))]),
        "trk_batch_fetch": (c_int, [c_void_p, P(P(abi.Tallies))]),
        "trk_batch_tune": (c_int, [c_void_p, c_uint64, c_int64]),
        "trk_batch_config": (c_int, [c_void_p, P(c_int32), P(c_int32), P(c_int64), P(c_int32), P(c_int32)]),
        "trk_batch_tune_log": (c_int, [c_void_p, P(c_int32), P(c_int32), abi.DP, c_int32, P(c_int32)]),
        "trk_batch_timing": (c_int, [c_void_p, abi.DP, P(c_int64), P(c_int32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._batch_declared = True
    return lib


class IonContext(Context):
    """An ion on the material of another (built) context: the material's
    tables are shared, build_tables() computes the ion's own tables only.
    The material context must outlive it."""

    def __init__(self, material_ctx, ion):
        self.lib = _declare(load_library())
        self.base = material_ctx          # keeps the material context (and its tables) alive
        self.desc = material_ctx.desc
        self.mat = material_ctx.mat
        self.ion = ion
        self.opt = material_ctx.opt
        h = ctypes.c_void_p()
        rc = self.lib.trk_ctx_create_ion(material_ctx.h, ctypes.byref(self.ion), ctypes.byref(h))
        self.h = h
        if rc:
            msg = self.error()
            self.close()
            raise TrekisError(f"trk_ctx_create_ion: {abi.STATUS.get(rc, rc)}: {msg}")
        self.Nt = material_ctx.Nt


class Batch:
    """Impacts (Context or IonContext with built tables) run together."""

    def __init__(self, impacts, Out_R, Out_V, capacity=0):
        self.lib = _declare(load_library())
        self.impacts = list(impacts)
        self.R = np.ascontiguousarray(Out_R, dtype=np.float64)
        self.V = np.ascontiguousarray(Out_V, dtype=np.float64)
        arr = (c_void_p * len(self.impacts))(*[c.h.value for c in self.impacts])
        h = c_void_p()
        rc = self.lib.trk_batch_create(arr, len(self.impacts), dptr(self.R), dptr(self.V), len(self.R), capacity,
                                       ctypes.byref(h))
        self.h = h
        if rc:
            msg = self.error()
            self.close()
            raise TrekisError(f"trk_batch_create: {abi.STATUS.get(rc, rc)}: {msg}")
        self._trace = None

    def error(self):
        if not self.h:
            return ""
        buf = ctypes.create_string_buffer(512)
        self.lib.trk_batch_error(self.h, buf, 512)
        return buf.value.decode()

    def _check(self, rc, what):
        if rc:
            raise TrekisError(f"{what}: {abi.STATUS.get(rc, rc)}: {self.error()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.trk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def new_tallies(self):
        """One TallySet per impact, without the per-iteration normalised tallies."""
        out = []
        for c in self.impacts:
            t = TallySet(c.Nt, len(self.R), c.desc)
            for k in ("Out_Ee_vs_E", "Out_Eh_vs_E", "Out_theta", "Out_theta_h", "Out_diff_coeff",
                      "Out_Ne_Em", "Out_E_Em", "Out_Ee_vs_E_Em"):
                setattr(t.s, k, None)
            out.append(t)
        return out

    def run(self, seed, first_iter=0, tallies=None, trace_cap=0, with_tallies=True):
        tal = (tallies or self.new_tallies()) if with_tallies else None
        ntr = c_int64(0)
        trace = None
        if trace_cap:
            trace = np.zeros((trace_cap, 12))
            self._check(self.lib.trk_batch_set_trace(self.h, dptr(trace), trace_cap, ctypes.byref(ntr)), "set_trace")
        ptrs = None
        if tal is not None:
            ptrs = (POINTER(abi.Tallies) * len(tal))(*[ctypes.pointer(t.s) for t in tal])
        try:
            rc = self.lib.trk_batch_run(self.h, seed, first_iter, ptrs)
        finally:
            if trace_cap:
                self.lib.trk_batch_set_trace(self.h, None, 0, None)
        self._check(rc, "trk_batch_run")
        if trace_cap:
            self.trace = trace[: min(ntr.value, trace_cap)]
            self.trace_total = ntr.value
        return tal

    def fetch(self, tallies=None):
        """The tallies of the last run (one TallySet per impact), fetched from
        the device; added into `tallies` when given."""
        tal = tallies or self.new_tallies()
        ptrs = (POINTER(abi.Tallies) * len(tal))(*[ctypes.pointer(t.s) for t in tal])
        self._check(self.lib.trk_batch_fetch(self.h, ptrs), "trk_batch_fetch")
        return tal

    def tune(self, seed, first_iter=0):
        self._check(self.lib.trk_batch_tune(self.h, seed, first_iter), "trk_batch_tune")
        return self.config()

    def config(self):
        blk, dep, cap, ng = c_int32(), c_int32(), c_int32(), c_int32()
        wg = c_int64()
        self._check(self.lib.trk_batch_config(self.h, ctypes.byref(blk), ctypes.byref(dep), ctypes.byref(wg),
                                              ctypes.byref(cap), ctypes.byref(ng)), "trk_batch_config")
        return {"block": blk.value, "round_events": dep.value, "workgroups": wg.value, "capacity": cap.value,
                "groups": ng.value}

    def tune_log(self):
        n = c_int32()
        blk = (c_int32 * 64)()
        dep = (c_int32 * 64)()
        ms = (c_double * 64)()
        self._check(self.lib.trk_batch_tune_log(self.h, blk, dep, ms, 64, ctypes.byref(n)), "trk_batch_tune_log")
        return [{"block": blk[i], "round_events": dep[i], "ms": ms[i]} for i in range(min(n.value, 64))]

    def timing(self):
        """(milliseconds of the last run's launch, events by kind, transport rounds)."""
        ms = c_double()
        ev = (c_int64 * 8)()
        rounds = c_int32()
        self._check(self.lib.trk_batch_timing(self.h, ctypes.byref(ms), ev, ctypes.byref(rounds)), "trk_batch_timing")
        return ms.value, np.array(ev[:], dtype=np.int64), rounds.value
