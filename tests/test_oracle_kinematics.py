"""Elastic scattering kinematics, the screened-Rutherford sampler and the
Auger / radiative shell choice of the oracle, pinned by closed forms.

* Mott_sample_mu (Cross_sections.f90:3573-3595): with the screening
  parameter nu, the screened-Rutherford distribution dsigma/dOmega ~
  1/(1 - mu + 2 nu)^2 has the CDF F(mu) = [1/(1 + 2nu - mu) - 1/(2 + 2nu)] /
  [1/(2nu) - 1/(2 + 2nu)]; the sampled mu must satisfy F(mu) = RN.
* transfered_E_from_theta (CS:3540-3557) and cos_theta_from_W (MC:1275-1299)
  are inverse relations of two-body elastic kinematics for a one-element
  target; W(pi) tends to the textbook 4 m M E / (M + m)^2 at low energy.
* Choose_for_Auger_shell (MC:1610-1647): random fractions map onto the
  eligible shells in proportion to their electron counts.
"""
import ctypes
import math

import numpy as np
import pytest

import trekis_gpu as tg
from oracle import Oracle

ME = 9.1093821545e-31
MP = 1836.1526724780 * ME


@pytest.fixture(scope="module")
def metal(lib):
    mat, ion, opt, run = tg.config_material("test_metal")
    mat.prepare(lib)
    return Oracle(mat, ion, opt), mat


@pytest.fixture(scope="module")
def oxide(lib):
    mat, ion, opt, run = tg.config_material("test")
    mat.prepare(lib)
    return Oracle(mat, ion, opt), mat


@pytest.mark.parametrize("Ee", [15.0, 200.0, 5000.0, 1.0e5])
def test_screened_rutherford_sampler(oxide, Ee):
    o, mat = oxide
    for atom in range(mat.n_atoms):
        nu = ctypes.c_double()
        for RN in np.linspace(0.001, 0.999, 37):
            th = o.lib.or_mott_theta(o.h, atom, Ee, float(RN), ctypes.byref(nu))
            mu = math.cos(th)
            n = nu.value
            F = (1.0 / (1.0 + 2 * n - mu) - 1.0 / (2.0 + 2 * n)) / (1.0 / (2 * n) - 1.0 / (2.0 + 2 * n))
            assert F == pytest.approx(RN, abs=1e-9)
        # screening grows with Z and falls with energy
    n_lo, n_hi = ctypes.c_double(), ctypes.c_double()
    o.lib.or_mott_theta(o.h, 1, Ee, 0.5, ctypes.byref(n_lo))      # O
    o.lib.or_mott_theta(o.h, 0, Ee, 0.5, ctypes.byref(n_hi))      # Al
    assert n_hi.value > n_lo.value


def test_screening_falls_with_energy(oxide):
    o, _ = oxide
    nus = []
    for E in (10.0, 100.0, 1000.0, 1e4):
        n = ctypes.c_double()
        o.lib.or_mott_theta(o.h, 0, E, 0.5, ctypes.byref(n))
        nus.append(n.value)
    assert all(a > b for a, b in zip(nus, nus[1:]))


@pytest.mark.parametrize("Ee", [20.0, 500.0, 20000.0])
def test_elastic_kinematics_inverse(metal, Ee):
    o, mat = metal
    for th in np.linspace(0.05, math.pi - 0.05, 15):
        W = o.lib.or_mott_W(o.h, 0, Ee, float(th), 1.0)
        assert 0.0 <= W < Ee
        mu = o.lib.or_cos_theta_W(o.h, Ee, W, ME)
        assert mu == pytest.approx(math.cos(th), abs=1e-6)


def test_backscatter_transfer_low_energy_limit(metal):
    o, mat = metal
    M = mat.atoms[0].mass * MP
    for Ee in (1.0, 10.0, 100.0):
        W = o.lib.or_mott_W(o.h, 0, Ee, math.pi, 1.0)
        assert W == pytest.approx(4.0 * ME * M * Ee / (M + ME) ** 2, rel=2e-4)


def test_auger_shell_choice_proportional_to_electrons(oxide):
    o, mat = oxide
    Ip = [s.Ip for a in mat.atoms for s in a.shells]
    Nel = [s.Nel for a in mat.atoms for s in a.shells]
    egap = mat.atoms[0].shells[-1].Ip
    for NRG, second in [(538.0, 0), (80.0, 0), (538.0 - 80.0, 1), (60.0, 1)]:
        elig = [s for s in range(len(Ip))
                if NRG > Ip[s] + 1e-3 and ((NRG - Ip[s]) >= egap or second)]
        K = 20000
        picks = [o.lib.or_auger_pick(o.h, NRG, (k + 0.5) / K, second) for k in range(K)]
        tot = sum(Nel[s] for s in elig)
        for s in range(len(Ip)):
            frac = picks.count(s) / K
            exp = Nel[s] / tot if s in elig else 0.0
            assert frac == pytest.approx(exp, abs=2.0 / K), (NRG, s)
