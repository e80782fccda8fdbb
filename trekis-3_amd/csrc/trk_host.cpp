// trk_host.cpp -- host-side helpers of the library: energy grid, time grid,
// effective charges, sum rules and the derived DOS arrays.  These run once per
// context on the CPU; the device work is in trk_tables.hip and trk_mc.hip.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>
#include "trk_ctx.h"

namespace trk {

int set_error(trk_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

// Equilibrium_charge_SHI, Cross_sections.f90:2641-2680
double shi_zeff(const HostMat &h, double E)
{
    double vp = (E > 0.0) ? trk_sqrt(((2.0 * E) * kE) / (h.ion.Mass * kMp)) : 0.0;
    double Zt = h.Zt;
    double Zp = (double)h.ion.Zat;
    double v0 = trk_sqrt(((2.0 * kRy) * kE) / kMe);
    switch (h.ion.Kind_Zeff) {
    case 1:
        return Zp * (1.0 - trk_exp(-((vp / v0) / trk_pow(Zp, 0.66666666))));
    case 2: {
        double c1 = 0.6, c2 = 0.45;
        return Zp * trk_pow(1.0 + trk_pow(vp / (((trk_pow(Zp, c2) * v0) * 4.0) / 3.0), -1.0 / c1), -c1);
    }
    case 3: {
        double c1 = 1.0 - 0.26 * trk_exp(-(Zt / 11.0) - ((Zt - Zp) * (Zt - Zp)) / 9.0);
        double vpvo = (trk_pow(Zp, -0.543) * vp) / v0;
        double c2 = 1.0 + (0.03 * vpvo) * trk_log(Zt);
        double x = c1 * trk_pow((vpvo / c2) / 1.54, 1.0 + 1.83 / Zp);
        double x2 = x * x, x4 = x2 * x2;
        return (Zp * ((8.29 * x) + x4)) / ((((0.06 / x) + 4.0) + (7.4 * x)) + x4);
    }
    case 4:
        return h.ion.fixed_Zeff;
    default:
        return Zp * (1.0 - trk_exp(-(((vp * 125.0) / kC) / trk_pow(Zp, 0.66666666))));
    }
}

// Target charge for the phonon-CDF elastic cross section (Barkas-like,
// Equilibrium_charge_Target Cross_sections.f90:2601-2638, used at 2921-2923)
double zeff_target(const HostMat &h, double Ee)
{
    if (h.o.CDF_elast_Zeff != 0) return 1.0;
    double Zp = h.Zt;
    double vp = trk_sqrt(((2.0 * Ee) * kE) / kMe);
    return 1.0 + Zp * (1.0 - trk_exp(-(((vp * 125.0) / kC) / trk_pow(Zp, 0.66666666))));
}

// ---------------------------------------------------------------- energy grid
// get_grid_4CS / go_thru_grid / define_special_points / exclude_doubles /
// find_order_of_number_real, Analytical_IMFPs.f90:2815-3093.
static int digits_of(long long v)
{
    int d = 1;
    if (v < 0) { d++; v = -v; }
    while (v >= 10) { v /= 10; d++; }
    return d;
}

static int order_of(double num)
{
    if (num > 1e9) return 8 + digits_of((long long)std::ceil(num / 1.0e8));
    return digits_of((long long)std::ceil(num));
}

static double ten_to(int k)
{
    double r = 1.0;
    for (int i = 0; i < (k < 0 ? -k : k); ++i) r = r * 10.0;
    return (k < 0) ? 1.0 / r : r;
}

int grid_4cs(const trk_material &m, double Emin, double Emax, std::vector<double> &out)
{
    std::vector<double> sp;
    for (int a = 0, s = 0; a < m.n_atoms; ++a)
        for (int k = 0; k < m.n_shells[a]; ++k, ++s) sp.push_back(m.Ip[s]);
    std::stable_sort(sp.begin(), sp.end());
    // exclude_doubles (AI:2991-3030): pairs (i<k) with equal values mark index i
    {
        const int ns = (int)sp.size();
        std::vector<int> ind;
        for (int i = 0; i < ns - 1; ++i)
            for (int k = i + 1; k < ns; ++k)
                if (sp[i] == sp[k]) ind.push_back(i + 1);
        const int newn = ns - (int)ind.size();
        std::vector<double> keep;
        size_t ci = 0;
        for (int i = 1; i <= ns; ++i) {
            if (ci < ind.size() && i == ind[ci]) ci++;
            else if ((int)keep.size() < newn) keep.push_back(sp[i - 1]);
        }
        sp.swap(keep);
    }
    const double eps = 1.0e-3, dE_min = 0.01;
    out.clear();
    size_t SP = 0;
    double E = Emin - dE_min;
    while (E < Emax) {
        double dE;
        if (E < 0.1 - dE_min * 0.5) dE = dE_min;
        else if (E < 10.0 - dE_min * 0.5) dE = dE_min * 10.0;
        else if (E < 100.0 - dE_min * 0.5) dE = dE_min * 100.0;
        else dE = ten_to(order_of(E) - 2);
        E = E + dE;
        out.push_back(E);
        if (SP < sp.size() && (E >= sp[SP]) && ((E - dE) < sp[SP])) {
            out.push_back(sp[SP] - eps);
            out.push_back(sp[SP] + eps);
            SP++;
        }
    }
    std::stable_sort(out.begin(), out.end());
    return (int)out.size();
}

// set_time_grid (Monte_Carlo.f90:2118-2149) and the time_do bound (572-575)
void build_time_grid(HostMat &h)
{
    const double Tim = h.o.Tim, dt = h.o.dt;
    std::vector<double> tg;
    if (h.o.dt_flag <= 0) {
        int n = (int)std::ceil(Tim / dt) + 1;
        tg.assign(n, 0.0);
        tg[0] = dt;
        for (int i = 1; i <= n - 2; ++i) tg[i] = tg[i - 1] + dt;
        tg[n - 1] = Tim + dt;
    } else {
        int i = 0;
        double t = 0.01;
        while (t <= Tim) { i++; t = t * dt; }
        int n = i + 1;
        tg.assign(n, 0.0);
        tg[0] = 0.01;
        for (int k = 1; k <= n - 2; ++k) tg[k] = tg[k - 1] * dt;
        tg[n - 1] = Tim + dt;
    }
    h.tim_glob.clear();
    double cur = 0.0;
    size_t i = 0;
    while (cur <= Tim - 1.0e-6 && i < tg.size()) {
        cur = (tg[i] < Tim) ? tg[i] : Tim;
        h.tim_glob.push_back(cur);
        i++;
    }
    h.Nt = (int)h.tim_glob.size();
}

}  // namespace trk

// ---------------------------------------------------------------- C-ABI helpers
using cd = std::complex<double>;

// Int_Ritchi_x, Cross_sections.f90:763-792: closed-form antiderivative of
// x * Im[-1/eps] for one Drude-Lorentz term, via complex atan written as logs.
static double int_ritchi_x(double A, double E, double G, double x)
{
    const double sq2 = std::sqrt(2.0);
    const cd I(0.0, 1.0);
    double d = -(2.0 * E) * (2.0 * E) + G * G;
    cd Sc = (d >= 0.0) ? cd(std::sqrt(d), 0.0) : cd(0.0, std::sqrt(std::fabs(d)));
    cd Gc(G * G - 2.0 * E * E, 0.0);
    cd sp = sq2 * std::sqrt(Gc + G * Sc);
    cd sm = sq2 * std::sqrt(Gc - G * Sc);
    cd GS = G * Sc;
    cd Bc = (GS == cd(0.0, 0.0)) ? cd(0.0, 0.0) : Gc / GS;
    cd arg = (2.0 * x) / sm;
    cd arg2 = (2.0 * x) / sp;
    cd t1 = 0.5 * I * (std::log(1.0 - I * arg) - std::log(1.0 + I * arg));
    cd t2 = 0.5 * I * (std::log(1.0 - I * arg2) - std::log(1.0 + I * arg2));
    cd Ic = (A * G) * (t1 / sm * (1.0 - Bc) + t2 / sp * (1.0 + Bc));
    return Ic.real();
}

// Int_Ritchi_p_x, Cross_sections.f90:795-826 (f-sum antiderivative)
static double int_ritchi_p_x(double A, double E, double G, double x)
{
    const double sq2 = std::sqrt(2.0);
    const cd I(0.0, 1.0);
    double d = -(2.0 * E) * (2.0 * E) + G * G;
    cd Sc = (d >= 0.0) ? cd(std::sqrt(d), 0.0) : cd(0.0, std::sqrt(std::fabs(d)));
    cd Gc(G * G - 2.0 * E * E, 0.0);
    cd sp = std::sqrt(Gc + G * Sc);
    cd sm = std::sqrt(Gc - G * Sc);
    cd arg = (sq2 * x) / sm;
    cd arg2 = (sq2 * x) / sp;
    cd t1 = (0.5 * I * (std::log(1.0 - I * arg) - std::log(1.0 + I * arg))) / sm;
    cd t2 = (0.5 * I * (std::log(1.0 - I * arg2) - std::log(1.0 + I * arg2))) / sp;
    cd In = ((t1 - t2) == cd(0.0, 0.0)) ? cd(sq2 * A, 0.0) : (sq2 * A) / Sc * (t1 - t2);
    return In.real();
}

// Int_Ritchi_x as Reading_files_and_parameters.f90:2208-2239 has it, in
// explicit complex arithmetic (naive products and quotients, principal
// square root, log as log|z| + i arg z) so that it rounds the same on every
// host: the k-sum antiderivative of one Ritchie-Drude oscillator
namespace {
struct Cx { double re, im; };
Cx cx(double r, double i) { return Cx{r, i}; }
Cx cadd(Cx a, Cx b) { return cx(a.re + b.re, a.im + b.im); }
Cx csub(Cx a, Cx b) { return cx(a.re - b.re, a.im - b.im); }
Cx cmul(Cx a, Cx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
Cx cdiv(Cx a, Cx b)
{
    const double d = b.re * b.re + b.im * b.im;
    return cx((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
Cx cscale(Cx a, double s) { return cx(a.re * s, a.im * s); }
Cx csqrt_p(Cx z)
{
    const double r = std::sqrt(z.re * z.re + z.im * z.im);
    const double x = std::sqrt(0.5 * (r + std::fabs(z.re)));
    if (x == 0.0) return cx(0.0, 0.0);
    if (z.re >= 0.0) return cx(x, z.im / (2.0 * x));
    return cx(std::fabs(z.im) / (2.0 * x), (z.im >= 0.0) ? x : -x);
}
Cx clog_p(Cx z) { return cx(std::log(std::hypot(z.re, z.im)), std::atan2(z.im, z.re)); }

double int_ritchi_x_rf(double A, double E, double G, double x)
{
    const double sq2 = std::sqrt(2.0);
    const Cx I = cx(0.0, 1.0), one = cx(1.0, 0.0);
    const double g2 = G * G;
    const double d = -(2.0 * E) * (2.0 * E) + g2;
    const Cx Sc = (d >= 0.0) ? cx(std::sqrt(d), 0.0) : cx(0.0, std::sqrt(std::fabs(d)));
    const Cx Gc = cx(g2 - (2.0 * E) * E, 0.0);
    const Cx GammaSc = cscale(Sc, G);
    const Cx s_plus = cscale(csqrt_p(cadd(Gc, GammaSc)), sq2);
    const Cx s_minus = cscale(csqrt_p(csub(Gc, GammaSc)), sq2);
    const Cx Bc = cdiv(Gc, GammaSc);
    Cx Ic1 = cx(0.0, 0.0);
    if (!(s_minus.re == 0.0 && s_minus.im == 0.0)) {
        const Cx arg = cdiv(cx(2.0 * x, 0.0), s_minus);
        const Cx L = csub(clog_p(csub(one, cmul(I, arg))), clog_p(cadd(one, cmul(I, arg))));
        Ic1 = cmul(cdiv(cmul(cscale(I, 0.5), L), s_minus), csub(one, Bc));
    }
    const Cx arg2 = cdiv(cx(2.0 * x, 0.0), s_plus);
    const Cx L2 = csub(clog_p(csub(one, cmul(I, arg2))), clog_p(cadd(one, cmul(I, arg2))));
    const Cx Ic = cadd(Ic1, cmul(cdiv(cmul(cscale(I, 0.5), L2), s_plus), cadd(one, Bc)));
    return cscale(Ic, A * G).re;
}
}  // namespace

extern "C" {

double trk_det_log(double x) { return trk_log(x); }
double trk_det_exp(double x) { return trk_exp(x); }
double trk_det_sin(double x) { return trk_sin(x); }
double trk_det_cos(double x) { return trk_cos(x); }
double trk_det_acos(double x) { return trk_acos(x); }
double trk_det_atan(double x) { return trk_atan(x); }
double trk_det_pow(double x, double y) { return trk_pow(x, y); }

double trk_rng_uniform(uint64_t seed, uint64_t particle_id, uint32_t iteration, uint32_t draw)
{
    return trk_uniform(seed, particle_id, iteration, draw);
}

// check_atomic_parameters, Dealing_with_EADL.f90:343-372, for decay times a
// caller supplies: the valence band (Shl_num >= 63, or the last shell of the
// first atom, where this library keeps it) gets Radiat = Auger = 1e23 fs, and
// every shell gets Radiat = 1e23 fs when photons are off.  The reference reads
// an Auger time <= 0 or > 1e30 fs, and every other radiative time, from the
// EADL database, which this library does not have: such a shell is refused,
// as is any time that is not a positive finite number.
int trk_check_atomic_parameters(trk_material *m, int32_t include_photons, char *why, size_t n)
{
    auto fail = [&](const char *fmt, int s, double v) {
        if (why && n) snprintf(why, n, fmt, s, v);
        return (int)TRK_ERR_ARG;
    };
    if (!m || m->n_atoms < 1 || m->n_atoms > TRK_MAX_ATOMS) {
        if (why && n) snprintf(why, n, "n_atoms out of range");
        return TRK_ERR_ARG;
    }
    int nsh = 0;
    for (int a = 0; a < m->n_atoms; ++a) {
        if (m->n_shells[a] < 0 || nsh + m->n_shells[a] > TRK_MAX_SHELLS) {
            if (why && n) snprintf(why, n, "shell counts out of range");
            return TRK_ERR_ARG;
        }
        nsh += m->n_shells[a];
    }
    const int vb = m->n_shells[0] - 1;
    for (int s = 0; s < nsh; ++s) {
        const bool is_vb = (s == vb) || (m->Shl_num[s] >= 63);
        if (is_vb || !include_photons) {
            m->Radiat[s] = 1.0e23;                               // EADL:350-351
        } else if (!(m->Radiat[s] > 0.0) || !std::isfinite(m->Radiat[s])) {
            return fail("shell %d: radiative decay time %g fs must be positive and finite "
                        "(the reference reads it from the EADL database)", s, m->Radiat[s]);
        }
        if (is_vb) {
            m->Auger[s] = 1.0e23;                                // EADL:366-367
        } else if (!(m->Auger[s] > 0.0) || m->Auger[s] > 1.0e30 || !std::isfinite(m->Auger[s])) {
            return fail("shell %d: Auger decay time %g fs is outside (0, 1e30]; the reference then "
                        "reads it from the EADL database, which this library does not have", s, m->Auger[s]);
        }
    }
    if (why && n) why[0] = 0;
    return TRK_OK;
}

// Physical constants as the library folds them (for checking against other
// restatements of Universal_Constants.f90).
double trk_constant(int which)
{
    switch (which) {
    case 0: return trk::kPi;
    case 1: return trk::kE;
    case 2: return trk::kMe;
    case 3: return trk::kMp;
    case 4: return trk::kH;
    case 5: return trk::kAlpha;
    case 6: return trk::kC;
    case 7: return trk::kA0;
    default: return 0.0;
    }
}

double trk_w_plasma(double density_m3, double mass_kg)
{
    if (mass_kg > 0.0) return ((density_m3 * trk::kE) * trk::kE) / (trk::kE0 * mass_kg);
    return ((density_m3 * trk::kE) * trk::kE) / (trk::kE0 * trk::kMe);
}

// define_alpha, Reading_files_and_parameters.f90:2199-2206: the weight of an
// oscillator's delta peak, its k-sum above x_min
double trk_delta_alpha(double A, double E0, double Gamma, double x_min)
{
    const double low_lim = int_ritchi_x_rf(A, E0, Gamma, x_min);
    const double high_lim = int_ritchi_x_rf(A, E0, Gamma, 1.0e30);
    return high_lim - low_lim;
}

// barrier_parameters, Monte_Carlo.f90:2094-2116: the energy E1 at which the
// Eckart barrier of height H over a length L passes half of the electrons,
// and the slope gamma of the logistic 1/(1 + exp(gamma (E1 - E))) there
int trk_barrier_parameters(double W, double bar_length, double H, double *E1_out, double *gamma_out)
{
    using namespace trk;
    if (!E1_out || !gamma_out) return TRK_ERR_ARG;
    const double L = bar_length * 1.0e-10;
    const double twopi_h = (2.0 * kPi) * kH;
    const double B = ((2.0 * H) - W) + 2.0 * std::sqrt(H * H - H * W);
    const double ksi = 0.5 * std::sqrt((((((8.0 * kMe) * L) * L) * B) * kE) / (twopi_h * twopi_h) - 1.0);
    const double bb = std::cosh((2.0 * kPi) * ksi);
    const double delta = (((2.0 * kPi) * L) * std::sqrt((2.0 * kMe) * kE)) / twopi_h;
    const double E1 = H + (2.0 * std::sqrt(H * (H - W))) *
                              (std::acosh(bb) / (delta * (std::sqrt(H) + std::sqrt(H - W))) - 1.0);
    const double g1 = delta * (std::sqrt(E1) + std::sqrt(E1 - W));
    const double g2 = delta * (std::sqrt(E1) - std::sqrt(E1 - W));
    const double gamma = (g1 * std::sinh(g1) + (2.0 * g2) * std::sinh(g2)) /
                         (std::sqrt(E1 * (E1 - W)) * (bb + std::cosh(g1)));
    *E1_out = E1;
    *gamma_out = gamma;
    return (std::isfinite(E1) && std::isfinite(gamma) && W > 0.0 && bar_length > 0.0 && H > W) ? 0 : TRK_ERR_ARG;
}

// sumrules, Cross_sections.f90:728-749
int trk_sumrules(const double *A, const double *E0, const double *Gamma, int32_t n,
                 double x_min, double Omega, double *ksum, double *fsum)
{
    if (!A || !E0 || !Gamma || n <= 0 || !ksum || !fsum || Omega <= 0.0) return TRK_ERR_ARG;
    double ne = 0.0, f = 0.0;
    for (int j = 0; j < n; ++j) {
        ne = ne + int_ritchi_x(A[j], E0[j], Gamma[j], 1e10) - int_ritchi_x(A[j], E0[j], Gamma[j], x_min);
        f = f + int_ritchi_p_x(A[j], E0[j], Gamma[j], 1e10) - int_ritchi_p_x(A[j], E0[j], Gamma[j], x_min);
    }
    *ksum = ((2.0 * trk::kE) * trk::kE) / ((trk::kPi * Omega) * trk::kH * trk::kH) * ne;
    *fsum = (f * 2.0) / trk::kPi;
    return 0;
}

int trk_dos_size(const double *raw_E, int32_t n_raw, int32_t *M_out)
{
    if (!raw_E || n_raw < 2 || !M_out) return TRK_ERR_ARG;
    double span = (raw_E[n_raw - 1] - raw_E[n_raw - 1]) - (raw_E[0] - raw_E[n_raw - 1]);
    *M_out = (int32_t)std::floor(span / 0.1);
    return (*M_out > 1) ? 0 : TRK_ERR_ARG;
}

// reading_material_DOS, Reading_files_and_parameters.f90:2374-2450, with the
// lookup of Linear_approx_2d (RF:3272-3306).
int trk_prepare_dos(const double *raw_E, const double *raw_DOS, int32_t n_raw,
                    double N_VB_el, double At_Dens, double sum_Pers,
                    double *E, double *DOS, double *int_DOS, double *k, double *Eff_m,
                    double *k_inv, double *Eff_m_inv, int32_t M)
{
    using namespace trk;
    int32_t Mc = 0;
    if (trk_dos_size(raw_E, n_raw, &Mc) || Mc != M) return TRK_ERR_ARG;
    std::vector<double> T0(n_raw);
    for (int i = 0; i < n_raw; ++i) T0[i] = raw_E[i] - raw_E[n_raw - 1];
    const double dE = 0.1, El1 = T0[0] - dE, El2 = 0.0;
    double Ecur = T0[0];
    for (int i = 0; i < M; ++i) {
        Ecur = Ecur + dE;
        int N = find_mono_2d(T0.data(), n_raw, Ecur);
        double loc;
        if (N == 1) loc = El2 + ((raw_DOS[0] - El2) / (T0[0] - El1)) * (Ecur - El1);
        else if (raw_DOS[N - 2] > 1.0e20) loc = raw_DOS[N - 2];
        else loc = raw_DOS[N - 2] + ((raw_DOS[N - 1] - raw_DOS[N - 2]) / (T0[N - 1] - T0[N - 2])) * (Ecur - T0[N - 2]);
        E[i] = Ecur;
        DOS[i] = loc;
    }
    const double Elast = E[M - 1];
    for (int i = 0; i < M; ++i) E[i] = trk_fabs(E[i] - Elast);
    std::reverse(E, E + M);
    std::reverse(DOS, DOS + M);
    std::vector<double> Iinv(M);
    double S = 0.0, Si = 0.0;
    for (int i = 0; i < M; ++i) {
        S = S + DOS[i];
        int_DOS[i] = S;
        Si = Si + DOS[M - 1 - i];
        Iinv[i] = Si;
    }
    const double SUM = int_DOS[M - 1];
    for (int i = 0; i < M; ++i) {
        DOS[i] = (DOS[i] / SUM) * N_VB_el;
        int_DOS[i] = (int_DOS[i] / SUM) * N_VB_el;
    }
    const double kpre = (((3.0 * 2.0) * kPi) * kPi) / 2.0;
    S = 0.0;
    for (int i = 0; i < M; ++i) {
        S = S + DOS[i];
        k[i] = trk_pow((((kpre * S) * At_Dens) / sum_Pers) * 1.0e6, 1.0 / 3.0);
        Eff_m[i] = (E[i] < 1.0e-10) ? 1.0 : ((((kH * kH) * k[i]) * k[i]) / ((2.0 * E[i]) * kE)) / kMe;
    }
    const double SUMI = Iinv[M - 1];
    for (int i = 0; i < M; ++i) {
        double v = (Iinv[i] / SUMI) * N_VB_el;
        k_inv[i] = trk_pow((((kpre * v) * At_Dens) / sum_Pers) * 1.0e6, 1.0 / 3.0);
        Eff_m_inv[i] = (E[i] < 1.0e-10) ? 1.0 : ((((kH * kH) * k_inv[i]) * k_inv[i]) / ((2.0 * E[i]) * kE)) / kMe;
    }
    return 0;
}

}  // extern "C"
