/*
 * trk_oracle_delta.c -- TEST INFRASTRUCTURE (the oracle): the Delta-CDF model
 * of the inelastic cross sections (kind_of_DR = 4), restated from
 * Cross_sections.f90:1417-1428, 1449-1786, 2051-2123 and
 * Reading_files_and_parameters.f90:2199-2239 in the reference's order of
 * operations, for electrons and holes on the target's electrons (both with
 * the electron's mass, as the reference calls it).
 */
#include <math.h>
#include <stdlib.h>
#include "trk_oracle_internal.h"

#define SQRT trk_sqrt
#define LOG  trk_log
#define G_ME_EV (0.51099906 * 1.0e6)   /* g_me_eV, Universal_Constants.f90:31-32 */

/* ---------------------------------------------------------------- define_alpha */
typedef struct { double re, im; } cz;
static cz zmk(double r, double i) { cz z = { r, i }; return z; }
static cz zadd(cz a, cz b) { return zmk(a.re + b.re, a.im + b.im); }
static cz zsub(cz a, cz b) { return zmk(a.re - b.re, a.im - b.im); }
static cz zmul(cz a, cz b) { return zmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cz zdiv(cz a, cz b)
{
    double d = b.re * b.re + b.im * b.im;
    return zmk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
static cz zscale(cz a, double s) { return zmk(a.re * s, a.im * s); }
static cz zsqrt(cz z)
{
    double r = sqrt(z.re * z.re + z.im * z.im);
    double x = sqrt(0.5 * (r + fabs(z.re)));
    if (x == 0.0) return zmk(0.0, 0.0);
    if (z.re >= 0.0) return zmk(x, z.im / (2.0 * x));
    return zmk(fabs(z.im) / (2.0 * x), (z.im >= 0.0) ? x : -x);
}
static cz zlog(cz z) { return zmk(log(hypot(z.re, z.im)), atan2(z.im, z.re)); }

/* Int_Ritchi_x of Reading_files_and_parameters.f90:2208-2239 */
static double int_ritchi_x_rf(double A, double E, double Gamma, double x)
{
    double sq2 = sqrt(2.0);
    cz oneI = zmk(0.0, 1.0), one = zmk(1.0, 0.0);
    double g2 = Gamma * Gamma;
    cz Sc;
    if ((-(2.0 * E) * (2.0 * E) + g2) >= 0.0) Sc = zmk(sqrt(-(2.0 * E) * (2.0 * E) + g2), 0.0);
    else Sc = zmk(0.0, sqrt(fabs(-(2.0 * E) * (2.0 * E) + g2)));
    cz Gc = zmk(g2 - 2.0 * E * E, 0.0);
    cz GammaSc = zscale(Sc, Gamma);
    cz s_plus_c = zscale(zsqrt(zadd(Gc, GammaSc)), sq2);
    cz s_minus_c = zscale(zsqrt(zsub(Gc, GammaSc)), sq2);
    cz Bc = zdiv(Gc, GammaSc);
    cz Ic1 = zmk(0.0, 0.0);
    if (!(s_minus_c.re == 0.0 && s_minus_c.im == 0.0)) {
        cz arg = zdiv(zmk(2.0 * x, 0.0), s_minus_c);
        cz lg = zsub(zlog(zsub(one, zmul(oneI, arg))), zlog(zadd(one, zmul(oneI, arg))));
        Ic1 = zmul(zdiv(zmul(zscale(oneI, 0.5), lg), s_minus_c), zsub(one, Bc));
    }
    cz arg2 = zdiv(zmk(2.0 * x, 0.0), s_plus_c);
    cz lg2 = zsub(zlog(zsub(one, zmul(oneI, arg2))), zlog(zadd(one, zmul(oneI, arg2))));
    cz Ic = zadd(Ic1, zmul(zdiv(zmul(zscale(oneI, 0.5), lg2), s_plus_c), zadd(one, Bc)));
    Ic = zscale(Ic, A * Gamma);
    return Ic.re;
}

/* define_alpha, RF:2199-2206 */
double or_delta_alpha(double A, double E0, double Gamma, double x_min)
{
    double low_lim = int_ritchi_x_rf(A, E0, Gamma, x_min);
    double high_lim = int_ritchi_x_rf(A, E0, Gamma, 1.0e30);
    return high_lim - low_lim;
}

/* ---------------------------------------------------------------- the model */
/* rest_energy of the electron, CS:1557-1561 */
static double mc2(void) { return or_rest_energy(OR_ME); }

/* W_min, CS:1564-1589, equal masses */
static double W_min(double Ip, double E, double E0)
{
    double Wmin = Ip;
    double E0min = E0 * (1.0 - 0.25 * E0 / E);
    return (Wmin > E0min) ? Wmin : E0min;
}

/* W_max, CS:1592-1614, identical particles */
static double W_max(double E, double Ip) { return (E + Ip) * 0.50; }

/* find_Wmax_equal_Wmin, CS:1617-1647, identical particles */
static double Eeq_of(double Ip, double E0)
{
    return 5.0 / 4.0 * E0 - Ip / 2.0 + 0.25 * SQRT(17.0 * E0 * E0 - 12.0 * Ip * E0 + 4.0 * Ip * Ip);
}

/* P_prefactor, CS:1525-1534, with velosity_from_kinetic_energy, CS:1537-1549 */
static double P_prefactor(double E, double nat)
{
    double Erest = mc2();
    double fact = E / Erest + 1.0;
    double v = OR_CVEL * SQRT(1.0 - 1.0 / (fact * fact));
    double beta = v / OR_CVEL;
    return 1.0e24 / (OR_PI * OR_A0 * nat * G_ME_EV * (beta * beta));
}

/* integral_CS, CS:1706-1712 */
static double integral_CS(double alpha, double Mc2, double mtc2, double E0, double W)
{
    double Mc22 = 2.0 * Mc2;
    return alpha / (G_ME_EV * (Mc22 - E0)) * ((Mc22 - mtc2) * LOG(Mc22 + W - E0) + Mc22 * mtc2 / E0 * (LOG(W) - LOG(fabs(W - E0))));
}

/* integrated_delta_CDF_CS, CS:1690-1704 */
static double integrated_delta_CDF_CS(double alpha, double Mc2, double E0, double mtc2, double W, double Ip, double E)
{
    double Wmin = W_min(Ip, E, E0);
    if ((W < Wmin) || (E <= Ip)) return 0.0;
    return integral_CS(alpha, Mc2, mtc2, E0, W);
}

/* Find_linear_a_b, CS:1671-1688 (minimal_sufficient_E of equal masses is Ip) */
static void Find_linear_a_b(double alpha, double E0, double Ip, double nat, double Eeq, double *a, double *b)
{
    double Wmin_lim = W_min(Ip, Eeq, E0);
    double Wmax_lim = W_max(Eeq, Ip);
    double P = P_prefactor(Eeq, nat);
    double CS = -P * (integral_CS(alpha, mc2(), mc2(), E0, Wmax_lim) - integral_CS(alpha, mc2(), mc2(), E0, Wmin_lim));
    double IpMm = Ip;
    *a = CS / (Eeq - IpMm);
    *b = -CS * Ip / (Eeq - IpMm);
}

/* Integral_CDF_delta_CS, CS:1449-1522, for the oscillators of one shell;
 * has_emax: Emax_in is present */
double or_delta_cs(const or_ctx *c, int shell, double E, double Ip, int has_emax, double Emax_in)
{
    double Mc2 = mc2(), mtc2 = mc2();
    double CS = 0.0, P = 0.0;
    for (int i = c->osc_off[shell]; i < c->osc_off[shell + 1]; ++i) {
        double E0 = c->m.osc_E0[i];
        double Emin = W_min(Ip, E, E0);
        double Estart = Ip;
        double Eeq = Eeq_of(Ip, E0);
        if (E <= Estart) {
            CS = 0.0;
            P = 0.0;
        } else {
            double dEed = Eeq / 100.0;
            if (E <= Eeq + dEed) {
                double a, b;
                Find_linear_a_b(c->alpha[i], E0, Ip, c->m.At_Dens, Eeq + dEed, &a, &b);
                CS = a * E + b;
                P = 1.0;
            } else {
                double Emax = W_max(E, Ip);
                if (has_emax) {
                    if (Emax_in < Emin) Emax = Emin;
                    else if (Emax_in < Emax) Emax = Emax_in;
                }
                CS = CS - (integrated_delta_CDF_CS(c->alpha[i], Mc2, E0, mtc2, Emax, Ip, E) -
                           integrated_delta_CDF_CS(c->alpha[i], Mc2, E0, mtc2, Emin, Ip, E));
                P = P_prefactor(E, c->m.At_Dens);
            }
        }
    }
    return fabs(CS) * P;
}

/* int_energy_loss, CS:1715-1719 */
static double int_energy_loss(double alpha, double Mc2, double mtc2, double E0, double Wmax, double Wmin)
{
    return -alpha * (mtc2 * LOG(fabs((Wmax - E0) / (Wmin - E0))) +
                     (2.0 * Mc2 - mtc2) * LOG(fabs((-Wmax + E0 - 2.0 * Mc2) / (-Wmin + E0 - 2.0 * Mc2)))) / G_ME_EV;
}

/* energy_loss_delta, CS:1722-1786, charge 1 */
double or_delta_eloss(const or_ctx *c, int shell, double E, double Ip)
{
    double Mc2 = mc2(), mtc2 = mc2(), Zeff = 1.0, nat = c->m.At_Dens;
    double S_cur = 0.0, P = 0.0;
    for (int i = c->osc_off[shell]; i < c->osc_off[shell + 1]; ++i) {
        double E0 = c->m.osc_E0[i];
        double Emin = W_min(Ip, E, E0);
        double Estart = Ip;
        double Eeq = Eeq_of(Ip, E0);
        if (E <= Estart) {
            S_cur = 0.0;
            P = 0.0;
        } else {
            double dEed = Eeq / 100.0;
            double Emax = W_max(E, Ip);
            if (E <= Eeq + dEed) {
                double a, b;
                Find_linear_a_b(c->alpha[i], E0, Ip, nat, Eeq + dEed, &a, &b);
                S_cur = 0.5 * a * (Emax - Estart);
                P = 1.0;
            } else {
                S_cur = S_cur - int_energy_loss(c->alpha[i], Mc2, mtc2, E0, Emax, Emin);
                P = P_prefactor(E, nat);
            }
        }
    }
    return fabs(S_cur) * (Zeff * Zeff) * P * nat * 1.0e-24;
}

/* TotIMFP's Delta-CDF branch, CS:952-956, with MFP_from_sigma, CS:1417-1428 */
void or_delta_tot_imfp(const or_ctx *c, int shell, double Ele, double Emin, double *L, double *dEdx)
{
    double sigma = or_delta_cs(c, shell, Ele, Emin, 0, 0.0);
    if (sigma > 1.0e-24) *L = 1.0 / (sigma * (c->m.At_Dens * 1.0e-24));
    else *L = 1.0e30;
    *dEdx = or_delta_eloss(c, shell, Ele, Emin);
}

/* get_inelastic_energy_transfer, CS:2051-2123, for the draw RN.  The
 * halving count is capped (a finite interval is within 1e-3 eV long before;
 * only a non-finite energy reaches the cap) */
double or_delta_transfer(const or_ctx *c, int shell, double Ee, double Ip, double RN)
{
    double eps = 1.0e-3;
    double E_left = Ip;
    double E_right = (Ip + Ee) * 0.5;
    double CS_tot = or_delta_cs(c, shell, Ee, Ip, 1, E_right);
    double CS_sampled = RN * CS_tot;
    double E_cur = (E_left + E_right) * 0.5;
    double CS_cur = or_delta_cs(c, shell, Ee, Ip, 1, E_cur);
    int k = 0;
    while (fabs(CS_cur - CS_sampled) / CS_sampled > eps) {
        if (k++ >= 200) break;
        if (CS_cur > CS_sampled) E_right = E_cur;
        else E_left = E_cur;
        E_cur = (E_left + E_right) / 2.0;
        if (fabs(E_left - E_right) < eps) break;
        CS_cur = or_delta_cs(c, shell, Ee, Ip, 1, E_cur);
    }
    return E_cur;
}
