// trk_beb.h -- the binary-encounter-Bethe (BEB) cross section of Kim and
// Rudd (Phys. Rev. A 50, 3954 (1994)) for electrons and holes on a shell
// with binding energy B, mean kinetic energy U and N electrons, as
// Cross_sections.f90:3888-4044 evaluates it (the energies normalised by B:
// t = T/B, u = U/B, w = W/B), and the transfer search of
// Electron_NRG_transfer_BEB (CS:2126-2161).  FP64 with the deterministic
// math layer: the table builder (host) and the event code (device) use the
// same functions.  Same order of operations as the reference.
#pragma once

#include "trk_dev_common.h"

namespace trk {

// S of Eq. (4): 4 pi a0^2 N (Ry/B)^2 [A^2]
TRK_HD double beb_S(double B, double N) { return ((((4.0 * kPi) * kA0) * kA0) * N) * (kRy / B) * (kRy / B); }

// Sigma_BEB, CS:3888-3904: the total cross section [A^2]
TRK_HD double beb_sigma(double T, double B, double U, double N)
{
    if (T <= B) return 0.0;
    const double S = beb_S(B, N);
    const double t0 = T / B;
    const double u0 = U / B;
    const double lt = trk_log(t0);
    return (S / ((t0 + u0) + 1.0)) * ((((lt * 0.5) * (1.0 - 1.0 / (t0 * t0))) + (1.0 - 1.0 / t0)) - lt / (t0 + 1.0));
}

// dSigma_dw_int, CS:3983-3986: the antiderivative of the differential cross
// section in w
TRK_HD double beb_dw_int(double S, double t0, double u0, double w0)
{
    return (S / ((t0 + u0) + 1.0)) *
           ((-(trk_log(w0 + 1.0) - trk_log(trk_fabs(t0 - w0))) / (t0 + 1.0) + (1.0 / (t0 - w0) - 1.0 / (w0 + 1.0))) +
            (trk_log(t0) * 0.5) * (1.0 / ((t0 - w0) * (t0 - w0)) - 1.0 / ((w0 + 1.0) * (w0 + 1.0))));
}

// dSigma_int_BEB, CS:3959-3981: the cross section of transfers below W [A^2]
TRK_HD double beb_dsigma_int(double T, double W, double B, double U, double N)
{
    const double S = beb_S(B, N);
    const double t0 = T / B, u0 = U / B, w0 = W / B;
    const double d0 = beb_dw_int(S, t0, u0, 0.0);
    return beb_dw_int(S, t0, u0, w0) - d0;
}

// dSigma_dw_w_int, CS:4032-4044: the antiderivative of w times the
// differential cross section
TRK_HD double beb_dw_w_int(double S, double t0, double u0, double w0)
{
    const double tw = 1.0 / ((t0 + u0) + 1.0);
    const double logwt = trk_log((w0 + 1.0) * trk_fabs(w0 - t0));
    const double overwt = 1.0 / (w0 - t0);
    const double overw1 = 1.0 / (w0 + 1.0);
    const double overt1 = 1.0 / (t0 + 1.0);
    const double A = ((tw * overt1) * overt1) * ((((t0 * t0) * trk_log(trk_fabs(w0 - t0))) + t0 * logwt) + trk_log(w0 + 1.0));
    const double Bt = tw * ((-(t0 * overwt) + overw1) + logwt);
    const double C = (tw * trk_log(t0)) * (((0.5 * ((overw1 * overw1) + ((t0 * overwt) * overwt))) + overwt) - overw1);
    return S * ((A + Bt) + C);
}

// dSigma_w_int_BEB, CS:4008-4030: the energy-weighted cross section of
// transfers below W [A^2 eV]
TRK_HD double beb_dsigma_w_int(double T, double W, double B, double U, double N)
{
    const double S = beb_S(B, N);
    const double t0 = T / B, u0 = U / B, w0 = W / B;
    const double d0 = beb_dw_w_int(S, t0, u0, 0.0);
    return B * (beb_dw_w_int(S, t0, u0, w0) - d0);
}

// Electron_NRG_transfer_BEB, CS:2126-2161: the transferred energy whose
// cumulative mean free path 1/(Mass n_a sigma(<W)) is L_need to 1e-3, by
// bisection from 0 up to (T - B)/2 (1000 halvings at most; as in the
// reference the first step moves the lower end to 0), plus B.  n_a: the
// density of the shell's atoms [1/A^3].
__host__ __device__ __noinline__ inline double beb_transfer(double T, double B, double U, double N, double Mass,
                                                           double n_a, double Emin, double L_need)
{
    double Emin1 = Emin;
    double Emax1 = (T - B) / 2.0;
    int coun = 0;
    double E = 0.0;
    double Sigma_cur = beb_dsigma_int(T, E, B, U, N);
    double L_cur = 1.0 / ((Mass * n_a) * Sigma_cur);
    while (trk_fabs(L_cur - L_need) / L_need > 0.001) {
        coun = coun + 1;
        Sigma_cur = beb_dsigma_int(T, E, B, U, N);
        L_cur = 1.0 / ((Mass * n_a) * Sigma_cur);
        if (L_cur > L_need) Emin1 = E;
        else Emax1 = E;
        E = (Emax1 + Emin1) / 2.0;
        if (coun >= 1000) break;
    }
    return E + B;
}

}  // namespace trk
