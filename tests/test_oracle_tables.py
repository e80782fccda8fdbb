"""The oracle's table-building half, pinned without the reference binary.

The reference ships no tests or expected outputs, so each function is pinned
by what the mathematics fixes: sum rules of the Drude-Lorentz loss function,
quadrature of the closed-form antiderivatives, brute-force searches, exact
interpolation of the function class each interpolation mode is exact for,
and the integer node counts of the energy grids (get_grid_4CS,
Analytical_ion_dEdx.f90:2815-3101).  Host-side product functions that share
a restatement with the oracle are compared bitwise."""
import ctypes
import math

import numpy as np
import pytest
from scipy import integrate

import trekis_gpu as tg
from trekis_gpu import abi
from oracle import Oracle, dp

PI = math.pi
HBAR, ME, QE = 1.05457162853e-34, 9.1093821545e-31, 1.602176487e-19


def lf(A, E0, G, w):
    return A * G * w / ((w * w - E0 * E0) ** 2 + (G * w) ** 2)


@pytest.fixture(scope="module")
def test_oracle(lib):
    mat, ion, opt, run = tg.config_material("test")
    mat.prepare(lib)
    return Oracle(mat, ion, opt), mat


# ------------------------------------------------------------- searches
def _brute_up(A, v):           # 1-based index of the first element > v, clamped to [1, N]
    return min(max(1, 1 + int(np.sum(A <= v))), len(A))


@pytest.mark.parametrize("N", [2, 3, 7, 64, 301])
def test_find_monotonous(oracle_lib, N):
    r = np.random.default_rng(N)
    A = np.cumsum(r.uniform(0.1, 2.0, N))
    vs = np.concatenate([r.uniform(A[0] - 3, A[-1] + 3, 400), [A[0] - 1, A[-1], A[-1] + 1]])
    for v in vs:
        exp = _brute_up(A, v)
        assert oracle_lib.or_find_mono_1d(dp(A), N, float(v)) == exp
        assert oracle_lib.or_find_mono_2d(dp(A), N, float(v)) == exp
    for k in range(1, N - 1):      # node values: the 2D search may stop on either side
        got = oracle_lib.or_find_mono_2d(dp(A), N, float(A[k]))
        assert got in (k + 1, k + 2)
        assert oracle_lib.or_find_mono_1d(dp(A), N, float(A[k])) == k + 2
    assert oracle_lib.or_find_mono_1d(dp(A), N, float("nan")) == -1


@pytest.mark.parametrize("N", [2, 5, 64, 200])
def test_find_decreasing_and_linear(oracle_lib, N):
    r = np.random.default_rng(100 + N)
    D = np.cumsum(r.uniform(0.1, 2.0, N))[::-1].copy()
    I = D[::-1].copy()
    for v in r.uniform(D[-1] - 2, D[0] + 2, 300):
        exp = min(max(1, int(np.sum(D > v))), N)
        assert oracle_lib.or_find_decr(dp(D), N, float(v)) == exp
        lin = 1 + int(np.argmax(I >= v)) if (I >= v).any() else N
        assert oracle_lib.or_find_linear(dp(I), N, float(v)) == min(lin, N)


def test_find_decr_is_a_count_on_nonincreasing_rows(oracle_lib):
    """The transport kernel replaces the reference's bisection over a
    cumulative row whose entries never increase by a count: the bisection's
    result is max(1, number of entries > v) whatever probe path it takes,
    plateaus and node values included.  (A NaN value, which the oracle
    reports as an error, the device search maps to 1, as the count does.)"""
    r = np.random.default_rng(7)
    for trial in range(400):
        N = int(r.integers(1, 300))
        D = np.sort(r.integers(0, int(r.choice([3, 40, 1000])), N).astype(np.float64))[::-1].copy()
        vals = np.concatenate([r.uniform(D[-1] - 2, D[0] + 2, 20), D[r.integers(0, N, 5)]])
        for v in vals:
            exp = max(1, int(np.sum(D > v)))
            assert oracle_lib.or_find_decr(dp(D), N, float(v)) == exp, (trial, N, v)


def _count_lt_16way(A, v):
    """The probe schedule of the transport kernel's count of entries below v
    on a non-decreasing row: 15 probes per round narrow [lo, hi), then up to
    16 entries are counted one by one."""
    lo, hi = 0, len(A)
    guard = 0
    while guard < 8 and hi - lo > 16:
        n = hi - lo
        pos = [lo + ((j + 1) * n) // 16 for j in range(15)]
        lt = [A[p] < v for p in pos]
        nlo, nhi = lo, hi
        for j in range(15):
            if lt[j]:
                nlo = pos[j] + 1
        for j in range(14, -1, -1):
            if not lt[j]:
                nhi = pos[j]
        lo, hi = nlo, nhi
        guard += 1
    return lo + sum(1 for k in range(16) if lo + k < hi and A[lo + k] < v)


def test_first_at_least_is_a_count_on_nondecreasing_rows():
    """SHI_energy_transfer takes the first i (1-based, i < N) with
    1/L(i) >= Tot by a linear scan, or N (MC:1719-1780, trk_oracle_mc.c
    shi_transfer_rn).  On an inverse row that never decreases and holds no
    NaN that index is 1 + the number of entries below Tot among the first
    N-1, plateaus and node values included; the kernel's 16-way narrowing
    finds that count."""
    r = np.random.default_rng(13)
    for trial in range(300):
        N = int(r.integers(1, 3000)) if trial % 3 else int(r.integers(1, 40))
        inv = np.sort(r.integers(0, int(r.choice([3, 50, 10 ** 6])), N).astype(np.float64))
        vals = np.concatenate([r.uniform(inv[0] - 2, inv[-1] + 2, 12), inv[r.integers(0, N, 4)], [np.nan]])
        for Tot in vals:
            i = 1
            while i < N and inv[i - 1] < Tot:
                i += 1
            cnt = int(np.sum(inv[:N - 1] < Tot))
            assert i == cnt + 1, (trial, N, Tot)
            assert _count_lt_16way(inv[:N - 1], Tot) == cnt, (trial, N, Tot)


def test_bracket_guess_equals_search(oracle_lib):
    """A particle's last bracket h is taken without a search only when its
    value lies strictly inside [A(h-1), A(h)] of a strictly increasing grid;
    there both bisection variants return h."""
    r = np.random.default_rng(11)
    for N in (2, 3, 9, 200, 581):
        A = np.cumsum(r.uniform(0.01, 3.0, N))
        for h in range(2, N + 1):
            lo, hi = A[h - 2], A[h - 1]
            for v in (lo + (hi - lo) * f for f in (1e-9, 0.37, 1 - 1e-9)):
                if not (lo < v < hi):
                    continue
                assert oracle_lib.or_find_mono_2d(dp(A), N, float(v)) == h
                assert oracle_lib.or_find_mono_1d(dp(A), N, float(v)) == h


# ---------------------------------------------------------- interpolation
@pytest.mark.parametrize("flag,f", [
    (1, lambda E: 3.0 - 0.25 * E),
    (3, lambda E: 2.0 + 0.7 * np.log(E)),
    (4, lambda E: np.exp(0.3 - 0.02 * E)),
    (5, lambda E: 5.0 * E ** -1.7),
])
def test_interpolate_exact_for_its_class(oracle_lib, flag, f):
    r = np.random.default_rng(flag)
    for _ in range(200):
        E1, E2 = sorted(r.uniform(1.0, 50.0, 2))
        if E2 - E1 < 1e-3:
            continue
        E = r.uniform(E1, E2)
        got = oracle_lib.or_interpolate(flag, E1, E2, f(E1), f(E2), E)
        assert got == pytest.approx(f(E), rel=1e-12)
    assert oracle_lib.or_interpolate(1, 2.0, 2.0 + 1e-8, 1.0, 7.0, 2.0) == 7.0   # degenerate bracket


# ---------------------------------------------------------- loss function
@pytest.mark.parametrize("A,E0,G", [(1.0, 15.0, 0.6), (0.3, 24.0, 8.0), (2.0, 560.0, 120.0), (0.5, 14.0, 3.0)])
def test_loss_function_moments(test_oracle, A, E0, G):
    """f-sum rule int w Im[-1/eps] dw = pi A / 2 and the perfect-screening
    moment int Im[-1/eps]/w dw = pi A / (2 E0^2) of one Drude-Lorentz term;
    at q > 0 the same moments give the dispersed E0(q) = E0 + hbar^2 q^2/2m."""
    o, _ = test_oracle
    c = o.lib.or_loss_func
    for hq, E0q in [(0.0, E0), (3.0e9, None)]:
        mode = 1 if hq == 0.0 else 0
        f1 = integrate.quad(lambda w: w * c(o.h, mode, A, E0, G, w, hq), 0, np.inf, limit=400, epsabs=0, epsrel=1e-10)[0]
        f2 = integrate.quad(lambda w: c(o.h, mode, A, E0, G, w, hq) / w, 0, np.inf, limit=400, epsabs=0, epsrel=1e-10)[0]
        assert f1 == pytest.approx(PI * A / 2.0, rel=1e-7)
        E0eff = math.sqrt(PI * A / (2.0 * f2))
        if E0q is None:      # free-electron dispersion, CS:346-393 (q given in the reference's units)
            E0q = E0 + (HBAR * HBAR * hq * hq) / (2.0 * ME)
        assert E0eff == pytest.approx(E0q, rel=1e-6)


def test_loss_function_pointwise(test_oracle):
    o, _ = test_oracle
    r = np.random.default_rng(11)
    for _ in range(100):
        A, E0, G, w = r.uniform(0.01, 3), r.uniform(1, 600), r.uniform(0.1, 100), r.uniform(0.01, 2000)
        assert o.lib.or_loss_func(o.h, 1, A, E0, G, w, 0.0) == pytest.approx(lf(A, E0, G, w), rel=1e-14)


# ------------------------------------------------------------ sum rules
@pytest.mark.parametrize("A,E0,G", [(1.0, 15.0, 0.6), (0.3, 24.0, 60.0), (2.0, 560.0, 120.0)])
def test_int_ritchi_antiderivatives(oracle_lib, A, E0, G):
    for x1, x2 in [(1e-3, 10.0), (3.0, 40.0), (8.8, 1000.0)]:
        q1 = integrate.quad(lambda w: w * lf(A, E0, G, w), x1, x2, limit=400, epsrel=1e-12, points=[E0])[0]
        q2 = integrate.quad(lambda w: lf(A, E0, G, w) / w, x1, x2, limit=400, epsrel=1e-12, points=[E0])[0]
        d1 = oracle_lib.or_int_ritchi_x(A, E0, G, x2) - oracle_lib.or_int_ritchi_x(A, E0, G, x1)
        d2 = oracle_lib.or_int_ritchi_p_x(A, E0, G, x2) - oracle_lib.or_int_ritchi_p_x(A, E0, G, x1)
        assert d1 == pytest.approx(q1, rel=1e-8, abs=1e-14)
        assert d2 == pytest.approx(q2, rel=1e-8, abs=1e-14)
    full = oracle_lib.or_int_ritchi_x(A, E0, G, 1e10) - oracle_lib.or_int_ritchi_x(A, E0, G, 0.0)
    assert full == pytest.approx(PI * A / 2.0, rel=1e-8)
    # CS:819-823: at x = 0 both atan terms vanish and the reference returns sqrt(2) A
    assert oracle_lib.or_int_ritchi_p_x(A, E0, G, 0.0) == math.sqrt(2.0) * A


def test_sumrules_product_matches_oracle(lib, oracle_lib):
    r = np.random.default_rng(5)
    for n in (1, 2, 4, 7):
        A, E0, G = r.uniform(0.1, 2, n), r.uniform(5, 600, n), r.uniform(0.5, 100, n)
        xmin, Om = float(r.uniform(0, 20)), lib.trk_w_plasma(float(r.uniform(1e28, 1e29)), 0.0)
        a = [ctypes.c_double(), ctypes.c_double()]
        b = [ctypes.c_double(), ctypes.c_double()]
        lib.trk_sumrules(dp(A), dp(E0), dp(G), n, xmin, Om, ctypes.byref(a[0]), ctypes.byref(a[1]))
        oracle_lib.or_sumrules(dp(A), dp(E0), dp(G), n, xmin, Om, ctypes.byref(b[0]), ctypes.byref(b[1]))
        # product: std::complex (libm clog/csqrt); oracle: its own complex helpers
        assert a[0].value == pytest.approx(b[0].value, rel=1e-14)
        assert a[1].value == pytest.approx(b[1].value, rel=1e-14)
        assert lib.trk_w_plasma(1e29, 0.0) == oracle_lib.or_w_plasma(1e29, 0.0)


def test_normalised_shell_counts_electrons(lib):
    """After f-sum normalisation each core shell integrates to its Nel
    (sumrules, Cross_sections.f90:554-688)."""
    mat, *_ = tg.config_material("config2")
    mat.prepare(lib)
    for ia, atom in enumerate(mat.atoms):
        contrib = atom.pers / mat.sum_pers
        for js, sh in enumerate(atom.shells):
            if ia == 0 and js == len(atom.shells) - 1:
                continue
            A = np.array([o[1] for o in sh.osc]); E0 = np.array([o[0] for o in sh.osc]); G = np.array([o[2] for o in sh.osc])
            w = sum(integrate.quad(lambda x: x * lf(a, e, g, x), sh.Ip, np.inf, limit=400, epsrel=1e-11)[0]
                    for a, e, g in zip(A, E0, G))
            Om = lib.trk_w_plasma(1e6 * mat.At_Dens * contrib, 0.0)
            ne = 2.0 * QE * QE / (PI * Om * HBAR * HBAR) * w
            assert ne == pytest.approx(sh.Nel, rel=1e-6)


# ----------------------------------------------------------------- grids
GRID_COUNTS = {  # incident-energy grids of get_grid_4CS (AI:158, 210) and the SHI grid (AI:2270-2273)
    "config1": (476, 479, 581),
    "config2": (476, 485, 388),
}


@pytest.mark.parametrize("name", sorted(GRID_COUNTS))
def test_grid_node_counts(lib, oracle_lib, name):
    mat, ion, opt, _ = tg.config_material(name)
    mat.prepare(lib)
    m = mat.to_struct()
    emax = 175.6e6 * 2.0 / 1836.0
    M = ion.Mass * lib.trk_constant(3)
    egap = mat.atoms[0].shells[-1].Ip
    egap = max(egap, 0.1)
    shi_min = math.ceil((M + ME) ** 2 / (M * ME) * egap / 4.0)
    got = []
    for Emin, Emax in [(0.1, emax), (0.01, emax), (shi_min, 175.6e6 / 2.0 * ion.Mass)]:
        n = ctypes.c_int32()
        oracle_lib.or_grid(ctypes.byref(m), Emin, Emax, None, 0, ctypes.byref(n))
        g = np.zeros(n.value)
        oracle_lib.or_grid(ctypes.byref(m), Emin, Emax, dp(g), n.value, ctypes.byref(n))
        assert np.all(np.diff(g) > 0)
        ips = sorted({max(s.Ip, 0.1) for a in mat.atoms for s in a.shells})
        have = [bool(np.isclose(g, ip - 1e-3, rtol=0, atol=1e-9).any()) for ip in ips]
        if name == "config1" and Emin == 0.01:
            # AI:2910-2916: ten 0.01 steps from 0 end at 0.09999999999999999, the
            # next node is exactly 0.2 and 0.2 - 0.1 == 0.1 is not below the first
            # special point 0.1, so it and (in order) every later one is skipped
            assert not any(have)
        else:
            assert all(have[i] for i, ip in enumerate(ips) if Emin < ip < Emax)
        got.append(n.value)
    assert tuple(got) == GRID_COUNTS[name]


# ------------------------------------------------------------------ DOS
def _dos_arrays(fn, mat, rE, rD):
    M = ctypes.c_int32()
    tg.load_library().trk_dos_size(dp(rE), len(rE), ctypes.byref(M))
    out = [np.zeros(M.value) for _ in range(7)]
    rc = fn(dp(rE), dp(rD), len(rE), mat.N_VB_el, mat.At_Dens, mat.sum_pers, *[dp(a) for a in out], M.value)
    assert rc == 0
    return out


def test_dos_product_matches_oracle_and_free_electron_mass(lib, oracle_lib):
    mat, *_ = tg.config_material("config1")
    rE, rD = tg.configs.dos_free_electron(11.17)
    a = _dos_arrays(lib.trk_prepare_dos, mat, rE, rD)
    b = _dos_arrays(oracle_lib.or_prepare_dos, mat, rE, rD)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    E, DOS, intD, k, m, kI, mI = a
    assert intD[-1] == pytest.approx(mat.N_VB_el, rel=1e-12)
    # a parabolic band gives a constant effective mass; its value follows from
    # the electron density n = N_VB_el * At_Dens / sum_pers and the band width
    n = mat.N_VB_el * mat.At_Dens / mat.sum_pers * 1e6
    EF_n = HBAR ** 2 * (3 * PI * PI * n) ** (2.0 / 3.0) / (2 * ME) / QE
    # (Eff_m_inv pairs the count of the lowest i+1 bins with E(i) = i*dE, so it
    # approaches the band mass from above as (i+1)/i, RF:2431-2446)
    top = slice(len(E) // 2, len(E) - 1)
    assert np.allclose(mI[top], EF_n / 11.17, rtol=0.03)
    assert np.all(np.diff(mI[top]) < 0)
