"""do_Monte_Carlo through the Fortran bridge (trekis-3_amd/fortran/
trekis_gpu_bridge.f90, compiled with the reference's Objects.f90): a driver
fills the reference's own types (Atom, Solid, Density_of_states, CDF, Ion,
Flag), allocates the Out_* arrays as Allocate_out_arrays does and calls the
bridge's do_Monte_Carlo, which flattens them, builds the tables on the GPU and
runs the transport.  Every Out_* array it returns equals, bit for bit, what
the same run gives through the C-ABI from Python."""
import os
import struct
import subprocess

import numpy as np
import pytest

import trekis_gpu as tg

from gpu_helpers import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "trekis-3_amd", "fortran", "_build", "bridge_driver")
ORDER = ["Out_ne", "Out_Ee", "Out_nphot", "Out_Ephot", "Out_Ee_vs_E", "Out_Eh_vs_E", "Out_Elat", "Out_nh", "Out_Eh",
         "Out_Ehkin", "Out_tot_Ne", "Out_tot_Nphot", "Out_tot_E", "Out_E_e", "Out_E_phot", "Out_E_at", "Out_E_h",
         "Out_theta", "Out_theta_h", "Out_diff_coeff", "Out_Ne_Em", "Out_E_Em", "Out_Ee_vs_E_Em"]


def _descriptor(m, ion, opt, nmc, seed, Nt, R, V, ms, dsf=None):
    """The driver's input stream (see bridge_driver.f90) from the C structs
    (and, with DSF elastic scattering, the electron and hole DSF tables)."""
    b = bytearray()
    i4 = lambda *v: b.extend(struct.pack(f"<{len(v)}i", *v))
    f8 = lambda *v: b.extend(struct.pack(f"<{len(v)}d", *v))
    i4(m.n_atoms)
    s = o = 0
    for a in range(m.n_atoms):
        i4(m.Zat[a]); f8(m.Mass[a], m.Pers[a]); i4(m.n_shells[a])
        for _ in range(m.n_shells[a]):
            f8(m.Ip[s], m.Nel[s], m.Auger[s], m.Radiat[s]); i4(m.Shl_num[s], m.n_osc[s], m.kocs[s] or 1)
            f8(m.Ek[s])
            for _ in range(m.n_osc[s]):
                f8(m.osc_E0[o], m.osc_A[o], m.osc_Gamma[o])
                o += 1
            s += 1
    i4(m.n_phonon)
    for k in range(m.n_phonon):
        f8(m.ph_E0[k], m.ph_A[k], m.ph_Gamma[k])
    f8(m.Dens, m.At_Dens, m.Vsound, m.Layer, m.N_VB_el, m.El_eff_mass, m.hole_mass, m.E_F, m.v_f, m.cut_off, m.temp,
       m.work_function, m.bar_length, m.bar_height)
    i4(m.n_dos)
    for f in ("dos_E", "dos_DOS", "dos_int_DOS", "dos_k", "dos_Eff_m", "dos_k_inv", "dos_Eff_m_inv"):
        f8(*[getattr(m, f)[k] for k in range(m.n_dos)])
    i4(ion.Zat); f8(ion.Mass, ion.E); i4(ion.Kind_Zeff); f8(ion.fixed_Zeff); i4(ion.Kind_ion)
    i4(opt.kind_of_EMFP, opt.CDF_elast_Zeff, opt.kind_of_DR, opt.plasmon_Emax, opt.include_photons, opt.dt_flag)
    f8(opt.Tim, opt.dt, opt.e_grid_Emax, opt.shi_grid_Emax)
    i4(nmc); b.extend(struct.pack("<q", seed)); i4(Nt, len(R), ms)
    f8(*R); f8(*V)
    if opt.kind_of_EMFP == 2:
        for t in dsf:
            E, dE, dL, Le, La = t.arrays()
            i4(*dE.shape)
            for i in range(len(E)):
                f8(E[i]); f8(*dE[i]); f8(*dL[i]); f8(*Le[i]); f8(*La[i])
    return bytes(b)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="the Fortran bridge driver is built where the reference's "
                                                       "Objects.f90 is (make -C trekis-3_amd/fortran)")
@pytest.mark.parametrize("name", ["test", "test_metal", "opt_emission", "opt_beb", "opt_dsf", "opt_combo"])
def test_fortran_do_monte_carlo_equals_c_abi(name, tmp_path):
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(name)
    mat.prepare(lib)
    ctx = tg.Context(mat, ion, opt)
    ctx.build_tables()
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    nmc, seed = 3, run["seed"]
    say(f"[fortran] {name}: C-ABI run")
    ref = ctx.run_mc(nmc, seed, R, V)
    m = ctx.mat
    ms = max(mat.n_shells)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(_descriptor(m, ion, opt, nmc, seed, ctx.Nt, R, V, ms, mat.dsf))
    say(f"[fortran] {name}: bridge driver")
    p = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=240)
    say(p.stdout.strip(), p.stderr.strip()[-2000:])
    assert p.returncode == 0
    raw = fout.read_bytes()
    err, num = struct.unpack("<2i", raw[:8])
    assert err == 0, (num, p.stdout)
    data = np.frombuffer(raw[8:], dtype="<f8")
    off = 0
    for k in ORDER:
        a = ref[k]
        n = a.size
        got = data[off:off + n].reshape(a.shape, order="F")
        off += n
        assert np.array_equal(got, a), (k, float(np.abs(got - a).max()))
    assert off == data.size
    assert float(ref["Out_tot_Ne"].sum()) > 0
