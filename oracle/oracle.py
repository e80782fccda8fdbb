"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by the tests, by `__graft_entry__.smoke()` to check the GPU result, and
by the CPU-baseline leg of `bench.py`.  Never imported by the library.
"""
import ctypes
import os
import subprocess

import numpy as np

from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_uint64, c_void_p, c_char_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libtrk_oracle.so")


def build():
    subprocess.run(["make", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def load():
    import sys
    sys.path.insert(0, os.path.normpath(os.path.join(HERE, "..", "trekis-3_amd")))
    from trekis_gpu import abi
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    DP = abi.DP
    P = POINTER
    sig = {
        "or_create": (c_void_p, [P(abi.Material), P(abi.Ion), P(abi.Options)]),
        "or_destroy": (None, [c_void_p]),
        "or_error": (c_char_p, [c_void_p]),
        "or_build_tables": (c_int, [c_void_p, c_int]),
        "or_set_shi_window": (None, [c_void_p, c_double, c_double]),
        "or_finish_tables": (c_int, [c_void_p]),
        "or_table_size": (c_int, [c_void_p, c_int, c_int, P(c_int32)]),
        "or_table_get": (c_int, [c_void_p, c_int, c_int, DP, DP, DP, c_int32]),
        "or_row_size": (c_int, [c_void_p, c_int, c_int, c_int32, P(c_int32)]),
        "or_row_get": (c_int, [c_void_p, c_int, c_int, c_int32, DP, DP, c_int32]),
        "or_table_set": (c_int, [c_void_p, c_int, c_int, DP, DP, DP, c_int32]),
        "or_row_set": (c_int, [c_void_p, c_int, c_int, c_int32, DP, DP, c_int32]),
        "or_loss_func": (c_double, [c_void_p, c_int, c_double, c_double, c_double, c_double, c_double]),
        "or_diff_cs": (c_double, [c_void_p, c_int, c_double, c_double, c_double]),
        "or_diff_cs_phonon": (c_double, [c_void_p, c_double, c_double, c_double]),
        "or_shi_diff_cs": (c_double, [c_void_p, c_int, c_double, c_double, c_double, c_double]),
        "or_inel_point": (c_int, [c_void_p, c_int, c_int, c_double, DP, DP, DP, DP, c_int32, P(c_int32)]),
        "or_grid": (c_int, [P(abi.Material), c_double, c_double, DP, c_int32, P(c_int32)]),
        "or_define_dE": (c_double, [c_int, c_double, c_double, c_double, c_double]),
        "or_find_mono_1d": (c_int, [DP, c_int32, c_double]),
        "or_find_mono_2d": (c_int, [DP, c_int32, c_double]),
        "or_find_decr": (c_int, [DP, c_int32, c_double]),
        "or_find_linear": (c_int, [DP, c_int32, c_double]),
        "or_interpolate": (c_double, [c_int, c_double, c_double, c_double, c_double, c_double]),
        "or_sumrules": (c_int, [DP, DP, DP, c_int32, c_double, c_double, DP, DP]),
        "or_int_ritchi_x": (c_double, [c_double, c_double, c_double, c_double]),
        "or_int_ritchi_p_x": (c_double, [c_double, c_double, c_double, c_double]),
        "or_w_plasma": (c_double, [c_double, c_double]),
        "or_prepare_dos": (c_int, [DP, DP, c_int32, c_double, c_double, c_double, DP, DP, DP, DP, DP, DP, DP, c_int32]),
        "or_time_nodes": (c_int, [c_void_p, P(c_int32), DP, c_int32]),
        "or_constant": (c_double, [c_int]),
        "or_barrier_parameters": (c_int, [c_double, c_double, c_double, DP, DP]),
        "or_dsf_set": (c_int, [c_void_p, c_int, c_int32, c_int32, DP, DP, DP, DP, DP]),
        "or_dsf_transfer": (c_double, [c_void_p, c_int, c_double, c_double, c_double]),
        "or_delta_alpha": (c_double, [c_double, c_double, c_double, c_double]),
        "or_delta_cs": (c_double, [c_void_p, c_int, c_double, c_double, c_int, c_double]),
        "or_delta_eloss": (c_double, [c_void_p, c_int, c_double, c_double]),
        "or_delta_transfer": (c_double, [c_void_p, c_int, c_double, c_double, c_double]),
        "or_beb_sigma": (c_double, [c_double, c_double, c_double, c_double]),
        "or_beb_dsigma_int": (c_double, [c_double, c_double, c_double, c_double, c_double]),
        "or_beb_dsigma_w_int": (c_double, [c_double, c_double, c_double, c_double, c_double]),
        "or_beb_transfer": (c_double, [c_void_p, c_int, c_double, c_double, c_double, c_double]),
        "or_sample_transfer": (c_double, [c_void_p, c_int, c_int, c_double, c_double]),
        "or_mott_theta": (c_double, [c_void_p, c_int, c_double, c_double, DP]),
        "or_mott_W": (c_double, [c_void_p, c_int, c_double, c_double, c_double]),
        "or_cos_theta_W": (c_double, [c_void_p, c_double, c_double, c_double]),
        "or_auger_pick": (c_int, [c_void_p, c_double, c_double, c_int]),
        "or_new_angles": (None, [c_double, c_double, c_double, c_double, DP, DP]),
        "or_gap_walk": (c_double, [c_void_p, c_double, c_int]),
        "or_bk_charge": (c_double, [c_double, c_double, c_double]),
        "or_shi_zeff": (c_double, [c_void_p, c_double]),
        "or_shi_diff_cs_bk": (c_double, [c_void_p, c_int, c_double, c_double, c_double, c_double]),
        "or_run_mc": (c_int, [c_void_p, c_int64, c_uint64, c_int64, DP, DP, c_int32, P(abi.Tallies), DP, c_int64,
                              P(c_int64)]),
        "or_run_mc_threads": (c_int, [c_void_p, c_int64, c_uint64, c_int64, DP, DP, c_int32, P(abi.Tallies), c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def dp(a):
    from trekis_gpu import abi
    return a.ctypes.data_as(abi.DP)


class Oracle:
    """Sequential restatement of the path for one material/ion/options."""

    def __init__(self, material_desc, ion, opt):
        self.lib = load()
        self.desc = material_desc
        self.mat = material_desc.to_struct()
        self.ion = ion
        self.opt = opt
        self.h = self.lib.or_create(ctypes.byref(self.mat), ctypes.byref(ion), ctypes.byref(opt))
        if opt.kind_of_EMFP == 2 and getattr(material_desc, "dsf", None) is not None:
            for particle, t in enumerate(material_desc.dsf):
                E, dE, dL, Le, La = t.arrays()
                self.lib.or_dsf_set(self.h, particle, dE.shape[0], dE.shape[1], dp(E), dp(dE), dp(dL), dp(Le), dp(La))
        nt = c_int32()
        self.lib.or_time_nodes(self.h, ctypes.byref(nt), None, 0)
        self.Nt = nt.value

    def close(self):
        if self.h:
            self.lib.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_tables(self, threads=0, shi_window=None):
        if shi_window is not None:
            self.lib.or_set_shi_window(self.h, shi_window[0], shi_window[1])
        rc = self.lib.or_build_tables(self.h, threads)
        if rc:
            raise RuntimeError(self.lib.or_error(self.h))

    def table(self, kind, shell=0):
        n = c_int32()
        self.lib.or_table_size(self.h, kind, shell, ctypes.byref(n))
        E, L, D = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
        rc = self.lib.or_table_get(self.h, kind, shell, dp(E), dp(L), dp(D), n.value)
        if rc:
            raise RuntimeError("or_table_get")
        return E, L, D

    def row(self, kind, shell, i_E):
        n = c_int32()
        rc = self.lib.or_row_size(self.h, kind, shell, i_E, ctypes.byref(n))
        if rc:
            raise RuntimeError("or_row_size")
        hw, lam = np.zeros(n.value), np.zeros(n.value)
        self.lib.or_row_get(self.h, kind, shell, i_E, dp(hw), dp(lam), n.value)
        return hw, lam

    def inject_from(self, ctx, kinds=(0, 1, 2, 3, 4, 5, 6)):
        """Copy every table of a trekis_gpu.Context into the oracle (replay on
        identical tables)."""
        nsh = sum(self.desc.n_shells)
        for kind in kinds:
            if kind == 6:
                for s in range(nsh):
                    E, L, _ = ctx.table(kind, s)
                    self.lib.or_table_set(self.h, kind, s, dp(E), dp(L), dp(np.zeros_like(E)), len(E))
                continue
            shells = [0] if kind in (2, 3) else range(nsh)
            if kind == 4 and not self.opt.include_photons:
                continue
            for s in shells:
                E, L, D = ctx.table(kind, s)
                self.lib.or_table_set(self.h, kind, s, dp(E), dp(L), dp(D), len(E))
                for i in range(len(E)):
                    try:
                        hw, lam = ctx.row(kind, s, i)
                    except Exception:
                        break
                    self.lib.or_row_set(self.h, kind, s, i, dp(hw), dp(lam), len(hw))
        self.lib.or_finish_tables(self.h)

    def run_mc_threads(self, nmc, seed, Out_R, Out_V, tallies, threads, first_iter=0):
        rc = self.lib.or_run_mc_threads(self.h, nmc, seed, first_iter, dp(Out_R), dp(Out_V), len(Out_R),
                                        ctypes.byref(tallies.s), threads)
        if rc:
            raise RuntimeError(self.lib.or_error(self.h).decode())
        return tallies

    def run_mc(self, nmc, seed, Out_R, Out_V, tallies, first_iter=0, trace_cap=0):
        trace = np.zeros((max(trace_cap, 1), 12))
        n = c_int64()
        rc = self.lib.or_run_mc(self.h, nmc, seed, first_iter, dp(Out_R), dp(Out_V), len(Out_R), ctypes.byref(tallies.s),
                                dp(trace) if trace_cap else None, trace_cap, ctypes.byref(n))
        if rc:
            raise RuntimeError(self.lib.or_error(self.h).decode())
        tallies.trace = trace[: min(n.value, trace_cap)] if trace_cap else None
        return tallies
