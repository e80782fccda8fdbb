"""DSF elastic scattering (kind_of_EMFP = 2): the reader's processing
(reading_DSF_cross_sections, Reading_files_and_parameters.f90:2516-2678)
and the oracle's transfer sampler (NRG_transfer_elastic_DSF,
Cross_sections.f90:3652-3795), pinned by what the mathematics fixes:

* the processing integrates the DSF by rectangles into inverse cumulative
  mean free paths: a constant DSF s gives 1/(s step j) at the j-th step, a
  linear one the closed-form rectangle sum; emission sums only dE >= 0,
  absorption stops at the last dE < 0, and a side without weight keeps
  1e30;
* the sampler is inverse-CDF sampling of the DSF: the emission fraction is
  the emission part's share of the DSF's integral, and the transfer drawn
  for a path draw u is the u-quantile of that side's peak (a Gaussian,
  quantile w0 + sigma sqrt(2) erfinv(2u - 1)), on the rows of the particle
  energies the grid holds; a swapped side, a reversed row or the wrong row
  fails it.

No DSF data ships with the reference: the tables here are synthetic."""
import math

import numpy as np
import pytest
from scipy.special import erfinv

import trekis_gpu as tg
from trekis_gpu import dsf
from oracle import Oracle


def blocks(fn, E=(1.0, 2.0), nW=120, wmax=0.25):
    """raw blocks in file order (transfers decreasing) of the DSF fn(i, w)"""
    w = np.linspace(wmax, -wmax, nW)
    return np.array(E), np.tile(w, (len(E), 1)), np.array([fn(i, w) for i in range(len(E))])


def test_constant_dsf_is_the_rectangle_sum():
    s = 0.7
    t = dsf.process_dsf(*blocks(lambda i, w: np.full_like(w, s)))
    nW = t.dE.shape[1]
    step = 0.4 / nW
    j = np.arange(1, nW + 1)
    assert np.allclose(t.dE[0], -0.2 + step * j, rtol=0, atol=1e-13)
    assert np.allclose(t.dL[0] * s * step * j, 1.0, rtol=1e-12)
    neg = t.dE[0] < 0.0
    k = int(neg.sum())
    # absorption: the running total over dE < 0, then constant
    assert np.allclose(t.dL_absorb[0, :k] * s * step * j[:k], 1.0, rtol=1e-12)
    assert np.all(t.dL_absorb[0, k:] == t.dL_absorb[0, k - 1])
    # emission: 1e30 until dE >= 0, then the sum over dE >= 0 only
    assert np.all(t.dL_emit[0, :k] == 1.0e30)
    assert np.allclose(t.dL_emit[0, k:] * s * step * (j[k:] - k), 1.0, rtol=1e-12)
    assert np.array_equal(t.dL[0], t.dL[1])


def test_linear_dsf_is_the_rectangle_sum():
    a, b = 2.0, 3.0
    t = dsf.process_dsf(*blocks(lambda i, w: a + b * w))
    nW = t.dE.shape[1]
    step = 0.4 / nW
    e = -0.2 + step * np.arange(1, nW + 1)
    expect = 1.0 / np.cumsum((a + b * e) * step)
    assert np.allclose(t.dL[0], expect, rtol=1e-11)


def test_sides_without_weight_stay_infinite():
    # only energy given to the lattice: no absorption path anywhere; the raw
    # rows run from +0.25 down to -0.25 and must be reversed to be searched
    t = dsf.process_dsf(*blocks(lambda i, w: np.where(w > 0.02, 1.0, 0.0)))
    assert np.all(t.dL_absorb == 1.0e30)
    assert np.all(np.isfinite(t.dL_emit[:, -1])) and t.dL_emit[0, -1] < 1.0e3
    # values below 1e-10 count as zero
    t = dsf.process_dsf(*blocks(lambda i, w: np.where(w > 0.02, 1.0, 5e-11)))
    assert np.all(t.dL_absorb == 1.0e30)


def test_file_round_trip():
    E, dE, s = dsf.synthetic_dsf_blocks(0.05, 50.0, 300.0, nE=5, nW=30)
    a = dsf.read_dsf(dsf.dsf_file_text(E, dE, s))
    b = dsf.process_dsf(E, dE, s)
    for x, y in zip(a.arrays(), b.arrays()):
        assert np.array_equal(x, y)
    with pytest.raises(ValueError):
        dsf.read_dsf("30\n1.0 0.1 1.0\n")
    with pytest.raises(ValueError, match="decreasing"):
        dsf.process_dsf(E, dE[:, ::-1], s)


W0 = (0.06, 0.10, 0.13)
SIG = 0.02
WEIGHT_ABS = (0.5, 0.25, 0.1)


def peaks(i, w):
    return np.exp(-0.5 * ((w - W0[i]) / SIG) ** 2) + WEIGHT_ABS[i] * np.exp(-0.5 * ((w + W0[i]) / SIG) ** 2)


@pytest.fixture(scope="module")
def sampler(lib):
    mat, ion, opt, run = tg.config_material("opt_dsf")
    t = dsf.process_dsf(*blocks(peaks, E=(1.0, 2.0, 3.0), nW=200, wmax=0.3))
    mat.dsf = (t, t)
    mat.prepare(lib)
    o = Oracle(mat, ion, opt)
    yield o
    o.close()


@pytest.mark.parametrize("i", [0, 1, 2])
@pytest.mark.parametrize("particle", [0, 1])
def test_sampler_is_the_inverse_cdf(sampler, i, particle):
    o = sampler
    Eel = (1.0, 2.0, 3.0)[i]
    frac = 1.0 / (1.0 + WEIGHT_ABS[i])
    f = lambda r1, r2: o.lib.or_dsf_transfer(o.h, particle, Eel, r1, r2)
    assert f(frac - 0.01, 0.5) > 0.0
    assert f(frac + 0.01, 0.5) < 0.0
    tol = 1.5 * 0.5 / 200 + 2e-3
    for u in (0.05, 0.2, 0.5, 0.8, 0.95):
        q = SIG * math.sqrt(2.0) * erfinv(2.0 * u - 1.0)
        assert abs(f(0.0, u) - (W0[i] + q)) < tol, (u, f(0.0, u))
        assert abs(f(0.999999, u) - (-W0[i] + q)) < tol, (u, f(0.999999, u))


def test_sampler_interpolates_between_rows(sampler):
    # half way between two particle energies the row is the mean of the two
    # cumulative paths: the median transfer lies between the two peaks
    o = sampler
    mid = o.lib.or_dsf_transfer(o.h, 0, 1.5, 0.0, 0.5)
    assert W0[0] + 0.005 < mid < W0[1] - 0.005
