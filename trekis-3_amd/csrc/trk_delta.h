// trk_delta.h -- the Delta-CDF model of the inelastic cross sections
// (kind_of_DR = 4): each Ritchie-Drude oscillator of a shell is replaced by a
// delta peak of weight alpha (its k-sum above the shell's edge, define_alpha,
// Reading_files_and_parameters.f90:2199-2206) that disperses as a free
// electron, so the cross section and the energy loss of an electron (or a
// hole, which the reference treats with the electron's mass here) on the
// target's electrons have closed forms (Cross_sections.f90:1449-1786).
//
// Used by the table kernels (mean free paths and stopping powers of the
// electron and hole inelastic tables) and by the event code (the transferred
// energy found by bisection on the cross section integrated up to it,
// get_inelastic_energy_transfer, CS:2051-2123).  FP64 with the deterministic
// math layer in both transport paths.  Same order of operations as the
// reference.
#pragma once

#include "trk_dev_common.h"

namespace trk {

constexpr double kMeEV = 0.51099906 * 1.0e6;   // g_me_eV, Universal_Constants.f90:31-32

// rest_energy of the electron [eV], CS:1557-1561
TRK_HD double delta_mc2() { return ((kMe * kC) * kC) / kE; }

// W_min for identical particles with a delta peak at E0, CS:1564-1589
TRK_HD double delta_wmin(double Ip, double E, double E0)
{
    const double E0min = E0 * (1.0 - (0.25 * E0) / E);
    return (Ip > E0min) ? Ip : E0min;
}

// find_Wmax_equal_Wmin for identical particles, CS:1617-1647: the energy
// below which the model is replaced by a linear extrapolation
TRK_HD double delta_eeq(double Ip, double E0)
{
    return ((5.0 / 4.0) * E0 - Ip / 2.0) + 0.25 * trk_sqrt((((17.0 * E0) * E0) - ((12.0 * Ip) * E0)) + ((4.0 * Ip) * Ip));
}

// P_prefactor with the electron's mass, CS:1525-1534 (velocity from the
// relativistic kinetic energy, CS:1537-1549)
TRK_HD double delta_prefactor(double E, double nat)
{
    const double fact = E / delta_mc2() + 1.0;
    const double v = kC * trk_sqrt(1.0 - 1.0 / (fact * fact));
    const double beta = v / kC;
    return 1.0e24 / ((((kPi * kA0) * nat) * kMeEV) * (beta * beta));
}

// integral_CS, CS:1706-1712: the antiderivative over the transferred energy W
TRK_HD double delta_integral(double alpha, double E0, double W)
{
    const double Mc2 = delta_mc2(), mtc2 = delta_mc2();
    const double Mc22 = 2.0 * Mc2;
    return (alpha / (kMeEV * (Mc22 - E0))) *
           ((Mc22 - mtc2) * trk_log((Mc22 + W) - E0) + ((Mc22 * mtc2) / E0) * (trk_log(W) - trk_log(trk_fabs(W - E0))));
}

// integrated_delta_CDF_CS, CS:1690-1704
TRK_HD double delta_integrated(double alpha, double E0, double W, double Ip, double E)
{
    const double Wmin = delta_wmin(Ip, E, E0);
    if ((W < Wmin) || (E <= Ip)) return 0.0;
    return delta_integral(alpha, E0, W);
}

// Find_linear_a_b, CS:1671-1688: the line through (Ip, 0) and the model's
// cross section at Eeq
TRK_HD void delta_linear(double alpha, double E0, double Ip, double nat, double Eeq, double &a, double &b)
{
    const double Wmin_lim = delta_wmin(Ip, Eeq, E0);
    const double Wmax_lim = (Eeq + Ip) * 0.50;
    const double P = delta_prefactor(Eeq, nat);
    const double CS = -(P * (delta_integral(alpha, E0, Wmax_lim) - delta_integral(alpha, E0, Wmin_lim)));
    const double IpMm = Ip;
    a = CS / (Eeq - IpMm);
    b = -((CS * Ip) / (Eeq - IpMm));
}

// Integral_CDF_delta_CS for an electron on electrons, CS:1449-1522 [A^2];
// Emax_in > 0 limits the transferred energy.  The oscillators [o0, o1) of the
// shell have peaks E0 and weights alpha.  As in the reference, a peak
// below its switching energy replaces the running sum by its linear value
// (and the prefactor by 1), and the prefactor of the last peak applies.
TRK_HD double delta_cs(const double *E0, const double *alpha, int o0, int o1, double E, double Ip, double nat,
                       double Emax_in)
{
    double CS = 0.0, P = 0.0;
    for (int i = o0; i < o1; ++i) {
        const double Emin = delta_wmin(Ip, E, E0[i]);
        const double Estart = Ip;
        const double Eeq = delta_eeq(Ip, E0[i]);
        if (E <= Estart) {
            CS = 0.0;
            P = 0.0;
        } else {
            const double dEed = Eeq / 100.0;
            if (E <= Eeq + dEed) {
                double a, b;
                delta_linear(alpha[i], E0[i], Ip, nat, Eeq + dEed, a, b);
                CS = a * E + b;
                P = 1.0;
            } else {
                double Emax = (E + Ip) * 0.50;
                if (Emax_in > 0.0) {
                    if (Emax_in < Emin) Emax = Emin;
                    else if (Emax_in < Emax) Emax = Emax_in;
                }
                CS = CS - (delta_integrated(alpha[i], E0[i], Emax, Ip, E) - delta_integrated(alpha[i], E0[i], Emin, Ip, E));
                P = delta_prefactor(E, nat);
            }
        }
    }
    return trk_fabs(CS) * P;
}

// int_energy_loss, CS:1715-1719
TRK_HD double delta_int_loss(double alpha, double E0, double Wmax, double Wmin)
{
    const double Mc2 = delta_mc2(), mtc2 = delta_mc2();
    const double t1 = mtc2 * trk_log(trk_fabs((Wmax - E0) / (Wmin - E0)));
    const double t2 = ((2.0 * Mc2) - mtc2) *
                      trk_log(trk_fabs((((-Wmax) + E0) - 2.0 * Mc2) / (((-Wmin) + E0) - 2.0 * Mc2)));
    return -((alpha * (t1 + t2)) / kMeEV);
}

// energy_loss_delta for an electron (charge 1) on electrons, CS:1722-1786 [eV/A]
TRK_HD double delta_eloss(const double *E0, const double *alpha, int o0, int o1, double E, double Ip, double nat)
{
    double S = 0.0, P = 0.0;
    for (int i = o0; i < o1; ++i) {
        const double Emin = delta_wmin(Ip, E, E0[i]);
        const double Estart = Ip;
        const double Eeq = delta_eeq(Ip, E0[i]);
        if (E <= Estart) {
            S = 0.0;
            P = 0.0;
        } else {
            const double dEed = Eeq / 100.0;
            const double Emax = (E + Ip) * 0.50;
            if (E <= Eeq + dEed) {
                double a, b;
                delta_linear(alpha[i], E0[i], Ip, nat, Eeq + dEed, a, b);
                S = (0.5 * a) * (Emax - Estart);
                P = 1.0;
            } else {
                S = S - delta_int_loss(alpha[i], E0[i], Emax, Emin);
                P = delta_prefactor(E, nat);
            }
        }
    }
    return (((trk_fabs(S) * (1.0 * 1.0)) * P) * nat) * 1.0e-24;
}

// MFP_from_sigma, CS:1417-1428 [A]
TRK_HD double delta_mfp(double sigma, double nat)
{
    if (sigma > 1.0e-24) return 1.0 / (sigma * (nat * 1.0e-24));
    return 1.0e30;
}

// Halvings of the transfer search: a finite interval of at most 1e9 eV is
// below its 1e-3 eV tolerance after 40; the cap only ends a search on a
// non-finite energy (whose NaN the transport then reports).
constexpr int kDeltaMaxHalvings = 200;

// get_inelastic_energy_transfer, CS:2051-2123, for the draw RN: the
// transferred energy in [Ip, (Ip + Ee)/2] whose integrated cross section is
// RN times the total, to 1e-3 relative or 1e-3 eV.  Not inlined: the
// transport kernels call it only in the Delta-CDF model, and its loops
// inlined there would take registers from every other event.
__host__ __device__ __noinline__ inline double delta_transfer(const double *E0, const double *alpha, int o0, int o1, double Ee, double Ip, double nat,
                             double RN)
{
    const double eps = 1.0e-3;
    double E_left = Ip;
    double E_right = (Ip + Ee) * 0.5;
    const double CS_tot = delta_cs(E0, alpha, o0, o1, Ee, Ip, nat, E_right);
    const double CS_sampled = RN * CS_tot;
    double E_cur = (E_left + E_right) * 0.5;
    double CS_cur = delta_cs(E0, alpha, o0, o1, Ee, Ip, nat, E_cur);
    for (int k = 0; k < kDeltaMaxHalvings && trk_fabs(CS_cur - CS_sampled) / CS_sampled > eps; ++k) {
        if (CS_cur > CS_sampled) E_right = E_cur;
        else E_left = E_cur;
        E_cur = (E_left + E_right) / 2.0;
        if (trk_fabs(E_left - E_right) < eps) break;
        CS_cur = delta_cs(E0, alpha, o0, o1, Ee, Ip, nat, E_cur);
    }
    return E_cur;
}

}  // namespace trk
