"""ctypes mirror of include/trekis_gpu.h (the C-ABI of the MI355X library).

The struct layouts below must match the C header field for field; the tests
check sizeof() of each against the library.
"""
import ctypes
import os

from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_uint32, c_uint64, c_char_p, c_size_t, c_void_p

MAX_ATOMS = 8
MAX_SHELLS = 32
MAX_OSC = 128
MAX_PHONON = 16

TAB_E_INEL, TAB_H_INEL, TAB_E_EL, TAB_H_EL, TAB_PHOT, TAB_SHI, TAB_SHI_DIFF = range(7)
K_SHI, K_TRANSPORT = 0, 1
TABLE_NAMES = ["e_inel", "h_inel", "e_el", "h_el", "phot", "shi", "shi_diff"]
EVENT_NAMES = ["shi", "e_inel", "e_el", "h_inel", "h_el", "auger", "radiative", "photon"]

TRK_OK, TRK_ERR_ARG, TRK_ERR_DEVICE, TRK_ERR_STATE, TRK_ERR_NUMERIC, TRK_ERR_CAPACITY = range(6)
STATUS = {0: "ok", 1: "invalid argument", 2: "device error", 3: "call out of order",
          4: "NaN in a table search", 5: "capacity exceeded"}

DP = POINTER(c_double)


class Material(ctypes.Structure):
    _fields_ = [
        ("n_atoms", c_int32),
        ("Zat", c_int32 * MAX_ATOMS),
        ("Mass", c_double * MAX_ATOMS),
        ("Pers", c_double * MAX_ATOMS),
        ("n_shells", c_int32 * MAX_ATOMS),
        ("Ip", c_double * MAX_SHELLS),
        ("Nel", c_double * MAX_SHELLS),
        ("Auger", c_double * MAX_SHELLS),
        ("Radiat", c_double * MAX_SHELLS),
        ("Shl_num", c_int32 * MAX_SHELLS),
        ("n_osc", c_int32 * MAX_SHELLS),
        ("osc_E0", c_double * MAX_OSC),
        ("osc_A", c_double * MAX_OSC),
        ("osc_Gamma", c_double * MAX_OSC),
        ("n_phonon", c_int32),
        ("ph_E0", c_double * MAX_PHONON),
        ("ph_A", c_double * MAX_PHONON),
        ("ph_Gamma", c_double * MAX_PHONON),
        ("Dens", c_double),
        ("At_Dens", c_double),
        ("Vsound", c_double),
        ("Layer", c_double),
        ("N_VB_el", c_double),
        ("El_eff_mass", c_double),
        ("hole_mass", c_double),
        ("E_F", c_double),
        ("v_f", c_double),
        ("cut_off", c_double),
        ("temp", c_double),
        ("work_function", c_double),
        ("bar_length", c_double),
        ("bar_height", c_double),
        ("n_dos", c_int32),
        ("dos_E", DP),
        ("dos_DOS", DP),
        ("dos_int_DOS", DP),
        ("dos_k", DP),
        ("dos_Eff_m", DP),
        ("dos_k_inv", DP),
        ("dos_Eff_m_inv", DP),
        ("kocs", c_int32 * MAX_SHELLS),
        ("Ek", c_double * MAX_SHELLS),
    ]


class Ion(ctypes.Structure):
    _fields_ = [
        ("Zat", c_int32),
        ("Mass", c_double),
        ("E", c_double),
        ("Kind_Zeff", c_int32),
        ("fixed_Zeff", c_double),
        ("Kind_ion", c_int32),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("kind_of_EMFP", c_int32),
        ("CDF_elast_Zeff", c_int32),
        ("kind_of_DR", c_int32),
        ("plasmon_Emax", c_int32),
        ("include_photons", c_int32),
        ("dt_flag", c_int32),
        ("Tim", c_double),
        ("dt", c_double),
        ("e_grid_Emax", c_double),
        ("shi_grid_Emax", c_double),
        ("precision", c_int32),
        ("batch", c_int32),
        ("block", c_int32),
        ("tally_replicas", c_int32),
        ("gen_capacity", c_int32),
        ("quadrature", c_int32),
        ("reserved", c_int32 * 2),
    ]


class Tallies(ctypes.Structure):
    _fields_ = [
        ("Nt", c_int32),
        ("NR", c_int32),
        ("n_dos", c_int32),
        ("max_shells", c_int32),
        ("Out_ne", DP), ("Out_Ee", DP),
        ("Out_nphot", DP), ("Out_Ephot", DP),
        ("Out_Ee_vs_E", DP),
        ("Out_Eh_vs_E", DP),
        ("Out_Elat", DP),
        ("Out_nh", DP), ("Out_Eh", DP), ("Out_Ehkin", DP),
        ("Out_tot_Ne", DP), ("Out_tot_Nphot", DP), ("Out_tot_E", DP),
        ("Out_E_e", DP), ("Out_E_phot", DP), ("Out_E_at", DP),
        ("Out_E_h", DP),
        ("Out_theta", DP), ("Out_theta_h", DP),
        ("Out_diff_coeff", DP),
        ("Out_Ne_Em", DP), ("Out_E_Em", DP),
        ("Out_Ee_vs_E_Em", DP),
        ("n_events", c_int64 * 8),
        ("n_particles", c_int64),
    ]


def declare_library(lib):
    """Attach argument/return types of every exported symbol of trekis_gpu.h."""
    P = ctypes.POINTER
    sig = {
        "trk_ctx_create": (c_int, [P(Material), P(Ion), P(Options), P(c_void_p)]),
        "trk_ctx_destroy": (None, [c_void_p]),
        "trk_last_error": (c_int, [c_void_p, c_char_p, c_size_t]),
        "trk_build_tables": (c_int, [c_void_p]),
        "trk_table_size": (c_int, [c_void_p, c_int, c_int, P(c_int32)]),
        "trk_table_get": (c_int, [c_void_p, c_int, c_int, DP, DP, DP, c_int32]),
        "trk_row_size": (c_int, [c_void_p, c_int, c_int, c_int32, P(c_int32)]),
        "trk_row_get": (c_int, [c_void_p, c_int, c_int, c_int32, DP, DP, c_int32]),
        "trk_table_set": (c_int, [c_void_p, c_int, c_int, DP, DP, DP, c_int32]),
        "trk_dsf_set": (c_int, [c_void_p, c_int32, c_int32, c_int32, DP, DP, DP, DP, DP]),
        "trk_row_set": (c_int, [c_void_p, c_int, c_int, c_int32, DP, DP, c_int32]),
        "trk_tables_finish": (c_int, [c_void_p]),
        "trk_time_nodes": (c_int, [c_void_p, P(c_int32)]),
        "trk_run_mc": (c_int, [c_void_p, c_int64, c_uint64, c_int64, DP, DP, c_int32, P(Tallies)]),
        "trk_run_mc_noalloc": (c_int, [c_void_p, c_int64, c_uint64, c_int64, DP, DP, c_int32, P(Tallies)]),
        "trk_last_timing": (c_int, [c_void_p, DP, DP, P(c_int64)]),
        "trk_set_trace": (c_int, [c_void_p, DP, c_int64, P(c_int64)]),
        "trk_set_profiling": (c_int, [c_void_p, c_int]),
        "trk_kernel_time": (c_int, [c_void_p, c_int, DP, P(c_int64)]),
        "trk_det_log": (c_double, [c_double]),
        "trk_det_exp": (c_double, [c_double]),
        "trk_det_sin": (c_double, [c_double]),
        "trk_det_cos": (c_double, [c_double]),
        "trk_det_acos": (c_double, [c_double]),
        "trk_det_atan": (c_double, [c_double]),
        "trk_det_pow": (c_double, [c_double, c_double]),
        "trk_rng_uniform": (c_double, [c_uint64, c_uint64, c_uint32, c_uint32]),
        "trk_constant": (c_double, [c_int]),
        "trk_check_atomic_parameters": (c_int, [P(Material), c_int32, c_char_p, c_size_t]),
        "trk_sumrules": (c_int, [DP, DP, DP, c_int32, c_double, c_double, DP, DP]),
        "trk_w_plasma": (c_double, [c_double, c_double]),
        "trk_barrier_parameters": (c_int, [c_double, c_double, c_double, DP, DP]),
        "trk_delta_alpha": (c_double, [c_double, c_double, c_double, c_double]),
        "trk_dos_size": (c_int, [DP, c_int32, P(c_int32)]),
        "trk_prepare_dos": (c_int, [DP, DP, c_int32, c_double, c_double, c_double,
                                    DP, DP, DP, DP, DP, DP, DP, c_int32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libtrekis_gpu.so"))


def header_symbols():
    """Function names declared in include/trekis_gpu.h."""
    import re
    hdr = os.path.normpath(os.path.join(_HERE, "..", "..", "include", "trekis_gpu.h"))
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(trk_[a-z_0-9]+)\s*\(", text)))
