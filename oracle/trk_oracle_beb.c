/*
 * trk_oracle_beb.c -- TEST INFRASTRUCTURE (the oracle): the
 * binary-encounter-Bethe cross section of electrons and holes on a shell
 * (KOCS = 2), restated from Cross_sections.f90:3888-4044 (Sigma_BEB,
 * dSigma_int_BEB, dSigma_dw_int, dSigma_w_int_BEB, dSigma_dw_w_int), TotIMFP's
 * BEB branch (CS:1035-1041) and Electron_NRG_transfer_BEB (CS:2126-2161), in
 * the reference's order of operations.
 */
#include <math.h>
#include "trk_oracle_internal.h"

#define LOG trk_log

/* Sigma_BEB, CS:3888-3904 [A^2] */
double or_beb_sigma(double T, double B, double U, double N)
{
    if (T <= B) return 0.0;
    double S = 4.0 * OR_PI * OR_A0 * OR_A0 * N * (OR_RY / B) * (OR_RY / B);
    double t0 = T / B;
    double u0 = U / B;
    return S / (t0 + u0 + 1.0) * (LOG(t0) * 0.5 * (1.0 - 1.0 / (t0 * t0)) + (1.0 - 1.0 / t0) - LOG(t0) / (t0 + 1.0));
}

/* dSigma_dw_int, CS:3983-3986 */
static double dSigma_dw_int(double S, double t0, double u0, double w0)
{
    return S / (t0 + u0 + 1.0) * (-(LOG(w0 + 1.0) - LOG(fabs(t0 - w0))) / (t0 + 1.0) + (1.0 / (t0 - w0) - 1.0 / (w0 + 1.0)) +
                                  LOG(t0) * 0.5 * (1.0 / ((t0 - w0) * (t0 - w0)) - 1.0 / ((w0 + 1.0) * (w0 + 1.0))));
}

/* dSigma_int_BEB, CS:3959-3981 [A^2] */
double or_beb_dsigma_int(double T, double w, double B, double U, double N)
{
    double S = 4.0 * OR_PI * OR_A0 * OR_A0 * N * (OR_RY / B) * (OR_RY / B);
    double t0 = T / B, u0 = U / B, w0 = w / B;
    double dSigma0 = dSigma_dw_int(S, t0, u0, 0.0);
    return dSigma_dw_int(S, t0, u0, w0) - dSigma0;
}

/* dSigma_dw_w_int, CS:4032-4044 */
static double dSigma_dw_w_int(double S, double t0, double u0, double w0)
{
    double tw = 1.0 / (t0 + u0 + 1.0);
    double logwt = LOG((w0 + 1.0) * fabs(w0 - t0));
    double overwt = 1.0 / (w0 - t0);
    double overw1 = 1.0 / (w0 + 1.0);
    double overt1 = 1.0 / (t0 + 1.0);
    double A = tw * overt1 * overt1 * (t0 * t0 * LOG(fabs(w0 - t0)) + t0 * logwt + LOG(w0 + 1.0));
    double B = tw * (-t0 * overwt + overw1 + logwt);
    double C = tw * LOG(t0) * (0.5 * (overw1 * overw1 + t0 * overwt * overwt) + overwt - overw1);
    return S * (A + B + C);
}

/* dSigma_w_int_BEB, CS:4008-4030 [A^2 eV] */
double or_beb_dsigma_w_int(double T, double w, double B, double U, double N)
{
    double S = 4.0 * OR_PI * OR_A0 * OR_A0 * N * (OR_RY / B) * (OR_RY / B);
    double t0 = T / B, u0 = U / B, w0 = w / B;
    double dSigma0 = dSigma_dw_w_int(S, t0, u0, 0.0);
    return B * (dSigma_dw_w_int(S, t0, u0, w0) - dSigma0);
}

/* the density of the shell's atoms [1/A^3]: temp1 of CS:1037, 2144 */
static double shell_density(const or_ctx *c, int shell)
{
    return c->m.At_Dens * 1.0e-24 * (c->m.Pers[c->atom_of[shell]]) / c->sumPers;
}

/* TotIMFP's BEB branch, CS:1035-1041 */
void or_beb_tot_imfp(const or_ctx *c, int shell, double Ele, double Mass, double *L, double *dEdx)
{
    double B = c->m.Ip[shell], U = c->m.Ek[shell], N = c->m.Nel[shell];
    double Sigma = or_beb_sigma(Ele, B, U, N);
    double temp1 = shell_density(c, shell);
    *L = 1.0 / (Mass * temp1 * Sigma);
    double ddEdx = or_beb_dsigma_w_int(Ele, (Ele - 1.0) / 2.0, B, U, N);
    *dEdx = Mass * temp1 * ddEdx;
}

/* Electron_NRG_transfer_BEB, CS:2126-2161 */
double or_beb_transfer(const or_ctx *c, int shell, double Ele, double L_need, double Mass, double Emin)
{
    double B = c->m.Ip[shell], U = c->m.Ek[shell], N = c->m.Nel[shell];
    double Emin1 = Emin;
    double Emax1 = (Ele - B) / 2.0;
    int coun = 0;
    double E = 0.0;
    double Sigma_cur = or_beb_dsigma_int(Ele, E, B, U, N);
    double temp1 = shell_density(c, shell);
    double L_cur = 1.0 / (Mass * temp1 * Sigma_cur);
    while (fabs(L_cur - L_need) / L_need > 0.001) {
        coun = coun + 1;
        Sigma_cur = or_beb_dsigma_int(Ele, E, B, U, N);
        temp1 = shell_density(c, shell);
        L_cur = 1.0 / (Mass * temp1 * Sigma_cur);
        if (L_cur > L_need) Emin1 = E;
        else Emax1 = E;
        E = (Emax1 + Emin1) / 2.0;
        if (coun >= 1000) break;
    }
    return E + B;
}
