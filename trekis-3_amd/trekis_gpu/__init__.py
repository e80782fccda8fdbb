"""Host-side mirror of the TREKIS-3 particle-history path on the MI355X.

The reference drives this path from Fortran: Read_input_file fills the
material, get_single_pole normalises the CDF oscillators, the table builders
run, then do_Monte_Carlo fills the Out_* tallies (Universal_MC_for_SHI_MAIN.f90
:128-276).  This module offers the same sequence over the C-ABI of
libtrekis_gpu.so: `Context(material, ion, options)` -> `build_tables()` ->
`run_mc(nmc, seed)`, returning the Out_* arrays with the reference's shapes.

The library is required: importing works anywhere, but creating a Context
without the built library or without a GPU raises.
"""
import ctypes
import os

import numpy as np

from . import abi
from .abi import Material, Ion, Options, Tallies

_LIB = None


def load_library(path=None):
    """Load libtrekis_gpu.so (built in-tree by `make -C trekis-3_amd`)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or abi.LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libtrekis_gpu.so not found at {p}: build it with __graft_entry__.build()")
    lib = abi.declare_library(ctypes.CDLL(p))
    if path is None:
        _LIB = lib
    return lib


def dptr(a):
    return a.ctypes.data_as(abi.DP)


class TrekisError(RuntimeError):
    pass


class Context:
    """One material + ion + options, with device tables and transport state."""

    def __init__(self, material, ion, options):
        self.lib = load_library()
        self.desc = material          # MaterialDesc (keeps the DOS arrays alive)
        self.mat = material.to_struct()
        self.ion = ion
        self.opt = options
        h = ctypes.c_void_p()
        rc = self.lib.trk_ctx_create(ctypes.byref(self.mat), ctypes.byref(self.ion), ctypes.byref(self.opt),
                                     ctypes.byref(h))
        self.h = h
        if rc:
            msg = self.error()
            self.close()
            raise TrekisError(f"trk_ctx_create: {abi.STATUS.get(rc, rc)}: {msg}")
        if options.kind_of_EMFP == 2 and getattr(material, "dsf", None) is not None:
            try:
                for particle, t in enumerate(material.dsf):
                    self.set_dsf(particle, t)
            except TrekisError:
                self.close()
                raise
        nt = ctypes.c_int32()
        self.lib.trk_time_nodes(self.h, ctypes.byref(nt))
        self.Nt = nt.value

    def set_dsf(self, particle, table):
        """DSF elastic tables (dsf.DsfTable) of electrons (particle 0) or
        holes (1); before build_tables."""
        E, dE, dL, Le, La = table.arrays()
        nE, nW = dE.shape
        self._check(self.lib.trk_dsf_set(self.h, particle, nE, nW, dptr(E), dptr(dE), dptr(dL), dptr(Le), dptr(La)),
                    "trk_dsf_set")

    def error(self):
        if not self.h:
            return ""
        buf = ctypes.create_string_buffer(512)
        self.lib.trk_last_error(self.h, buf, 512)
        return buf.value.decode()

    def _check(self, rc, what):
        if rc:
            raise TrekisError(f"{what}: {abi.STATUS.get(rc, rc)}: {self.error()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.trk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_tables(self):
        self._check(self.lib.trk_build_tables(self.h), "trk_build_tables")

    def table(self, kind, shell=0):
        n = ctypes.c_int32()
        self._check(self.lib.trk_table_size(self.h, kind, shell, ctypes.byref(n)), "trk_table_size")
        E = np.zeros(n.value)
        L = np.zeros(n.value)
        D = np.zeros(n.value)
        self._check(self.lib.trk_table_get(self.h, kind, shell, dptr(E), dptr(L), dptr(D), n.value), "trk_table_get")
        return E, L, D

    def row(self, kind, shell, i_E):
        n = ctypes.c_int32()
        self._check(self.lib.trk_row_size(self.h, kind, shell, i_E, ctypes.byref(n)), "trk_row_size")
        hw = np.zeros(n.value)
        lam = np.zeros(n.value)
        self._check(self.lib.trk_row_get(self.h, kind, shell, i_E, dptr(hw), dptr(lam), n.value), "trk_row_get")
        return hw, lam

    def set_table(self, kind, shell, L=None, dEdx=None, E=None):
        """Overwrite one shell's mean free paths / stopping powers of a built
        table (E, when given, must equal the built grid); finish_tables()
        must follow before the next run."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (E, L, dEdx)]
        n = next(len(a) for a in arrs if a is not None)
        self._check(self.lib.trk_table_set(self.h, kind, shell, *[dptr(a) if a is not None else None for a in arrs], n),
                    "trk_table_set")

    def set_row(self, kind, shell, i_E, hw, lam):
        hw = np.ascontiguousarray(hw, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        self._check(self.lib.trk_row_set(self.h, kind, shell, i_E, dptr(hw), dptr(lam), len(hw)), "trk_row_set")

    def finish_tables(self):
        self._check(self.lib.trk_tables_finish(self.h), "trk_tables_finish")

    def new_tallies(self, NR):
        return TallySet(self.Nt, NR, self.desc)

    def run_mc(self, nmc, seed, Out_R, Out_V, tallies=None, first_iter=0, trace_cap=0, noalloc=False):
        """do_Monte_Carlo: NMC iterations added into the Out_* tallies.  noalloc
        runs trk_run_mc_noalloc (no device allocation; no event trace)."""
        Out_R = np.ascontiguousarray(Out_R, dtype=np.float64)
        Out_V = np.ascontiguousarray(Out_V, dtype=np.float64)
        NR = len(Out_R)
        t = tallies or self.new_tallies(NR)
        trace = None
        ntr = ctypes.c_int64(0)
        if trace_cap:
            trace = np.zeros((trace_cap, 12))
            self.lib.trk_set_trace(self.h, dptr(trace), trace_cap, ctypes.byref(ntr))
        else:
            self.lib.trk_set_trace(self.h, None, 0, None)
        run = self.lib.trk_run_mc_noalloc if noalloc else self.lib.trk_run_mc
        try:
            rc = run(self.h, nmc, seed, first_iter, dptr(Out_R), dptr(Out_V), NR, ctypes.byref(t.s))
        finally:
            if trace_cap:
                # the library must not keep a pointer into this array
                self.lib.trk_set_trace(self.h, None, 0, None)
        self._check(rc, "trk_run_mc_noalloc" if noalloc else "trk_run_mc")
        if trace_cap:
            t.trace = trace[: min(ntr.value, trace_cap)]
            t.trace_total = ntr.value
        return t

    def set_profiling(self, on):
        self._check(self.lib.trk_set_profiling(self.h, 1 if on else 0), "trk_set_profiling")

    def kernel_time(self, kernel):
        """(milliseconds summed over launches, launches) of TRK_K_SHI / TRK_K_TRANSPORT."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self.lib.trk_kernel_time(self.h, kernel, ctypes.byref(ms), ctypes.byref(n)), "trk_kernel_time")
        return ms.value, n.value

    def timing(self):
        a = ctypes.c_double()
        b = ctypes.c_double()
        n = ctypes.c_int64()
        self.lib.trk_last_timing(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n))
        return a.value, b.value, n.value


class TallySet:
    """Caller-owned Out_* arrays (Fortran order) plus the C struct pointing at them."""

    FIELDS_NT_NR = ["Out_ne", "Out_Ee", "Out_nphot", "Out_Ephot", "Out_Ee_vs_E", "Out_Elat", "Out_Ee_vs_E_Em"]
    FIELDS_NT = ["Out_tot_Ne", "Out_tot_Nphot", "Out_tot_E", "Out_E_e", "Out_E_phot", "Out_E_at", "Out_diff_coeff",
                 "Out_Ne_Em", "Out_E_Em"]

    def __init__(self, Nt, NR, desc):
        na = desc.n_atoms
        ms = max(desc.n_shells)
        nd = desc.n_dos
        self.arrays = {}
        for f in self.FIELDS_NT_NR:
            self.arrays[f] = np.zeros((Nt, NR), order="F")
        for f in self.FIELDS_NT:
            self.arrays[f] = np.zeros(Nt)
        self.arrays["Out_Eh_vs_E"] = np.zeros((Nt, nd), order="F")
        for f in ["Out_nh", "Out_Eh", "Out_Ehkin"]:
            self.arrays[f] = np.zeros((Nt, NR, na, ms), order="F")
        self.arrays["Out_E_h"] = np.zeros((Nt, na, ms), order="F")
        self.arrays["Out_theta"] = np.zeros((Nt + 1, 180), order="F")
        self.arrays["Out_theta_h"] = np.zeros((Nt + 1, 180), order="F")
        s = Tallies()
        s.Nt, s.NR, s.n_dos, s.max_shells = Nt, NR, nd, ms
        for k, a in self.arrays.items():
            setattr(s, k, dptr(a))
        self.s = s
        self.trace = None

    def __getitem__(self, k):
        return self.arrays[k]

    @property
    def n_events(self):
        return np.array(self.s.n_events[:], dtype=np.int64)


from .configs import MaterialDesc, Shell, Atom, radial_grid, reference_radial_grid, config_material  # noqa: E402,F401
