/*
 * trk_oracle_dsf.c -- TEST INFRASTRUCTURE (the oracle): elastic scattering
 * of electrons and holes from dynamic structure factors (kind_of_EMFP = 2).
 * Holds the DSF_DEMFP / DSF_DEMFP_H tables the caller hands over, builds the
 * elastic mean free path tables from their last entries
 * (Analytical_IMFPs.f90:905-911, 1181-1184) and samples transfers as
 * NRG_transfer_elastic_DSF (Cross_sections.f90:3652-3795) with
 * Linear_approx_2x1d_DSF (Reading_files_and_parameters.f90:3718-3745): the
 * interpolated row is formed as a whole array, the search runs on its
 * negated copy, as the reference does.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "trk_oracle_internal.h"

static double *dup(const double *a, size_t n)
{
    double *d = (double *)malloc(sizeof(double) * (n ? n : 1));
    if (n) memcpy(d, a, sizeof(double) * n);
    return d;
}

void or_dsf_free(or_ctx *c)
{
    for (int p = 0; p < 2; ++p) {
        free(c->dsfE[p]); free(c->dsfdE[p]); free(c->dsfL[p]); free(c->dsfLe[p]); free(c->dsfLa[p]);
        c->dsfE[p] = c->dsfdE[p] = c->dsfL[p] = c->dsfLe[p] = c->dsfLa[p] = NULL;
        c->dsf_nE[p] = c->dsf_nW[p] = 0;
    }
}

/* the tables of one particle (0 electron, 1 hole), E-major nE x nW rows */
int or_dsf_set(or_ctx *c, int particle, int32_t nE, int32_t nW, const double *E, const double *dE,
               const double *dL, const double *dL_emit, const double *dL_absorb)
{
    if (particle < 0 || particle > 1 || nE < 2 || nW < 2) return 1;
    const size_t cells = (size_t)nE * (size_t)nW;
    free(c->dsfE[particle]); free(c->dsfdE[particle]); free(c->dsfL[particle]);
    free(c->dsfLe[particle]); free(c->dsfLa[particle]);
    c->dsf_nE[particle] = nE;
    c->dsf_nW[particle] = nW;
    c->dsfE[particle] = dup(E, (size_t)nE);
    c->dsfdE[particle] = dup(dE, cells);
    c->dsfL[particle] = dup(dL, cells);
    c->dsfLe[particle] = dup(dL_emit, cells);
    c->dsfLa[particle] = dup(dL_absorb, cells);
    return 0;
}

/* Elastic_MFP%Total on the DSF energies: L = the last entry of each row,
 * dEdx left at zero (the reference never fills it for DSF) */
void or_dsf_elastic_table(const or_ctx *c, int particle, double **E, double **L, double **dEdx, int32_t *n)
{
    const int nE = c->dsf_nE[particle], nW = c->dsf_nW[particle];
    *n = nE;
    *E = dup(c->dsfE[particle], (size_t)nE);
    *L = (double *)malloc(sizeof(double) * (size_t)nE);
    *dEdx = (double *)calloc((size_t)nE, sizeof(double));
    for (int i = 0; i < nE; ++i) (*L)[i] = c->dsfL[particle][(size_t)i * nW + (nW - 1)];
}

/* linear_interpolation, CS:3796-3800 */
static double lin(double y1, double y2, double x1, double x2, double x)
{
    return y1 + (y2 - y1) / (x2 - x1) * (x - x1);
}

/* Linear_approx_2x1d_DSF, RF:3718-3745 */
static double approx_2x1d_dsf(const double *dL, const double *dEr, int n, double L_need)
{
    double *neg = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) neg[i] = -dL[i];
    int Number = or_find_mono_1d(neg, n, -L_need);
    free(neg);
    double V;
    if (Number == 1) {
        V = dEr[0] + (dEr[1] - dEr[0]) / (dL[1] - dL[0]) * (L_need - dL[0]);
    } else if (fabs(dL[Number - 1] - dL[Number - 2]) < 1.0e-9) {
        V = dEr[Number - 2];
    } else if (dL[Number - 2] > 1.0e20) {
        V = dEr[Number - 2];
    } else {
        V = dEr[Number - 2] + (dEr[Number - 1] - dEr[Number - 2]) / (dL[Number - 1] - dL[Number - 2]) * (L_need - dL[Number - 2]);
    }
    return V;
}

/* NRG_transfer_elastic_DSF, CS:3652-3795, with its two draws RN1 (emission
 * or absorption) and RN2 (the path to reach) */
double or_dsf_transfer(const or_ctx *c, int particle, double Eel, double RN1, double RN2)
{
    const int nE = c->dsf_nE[particle], nW = c->dsf_nW[particle];
    const double *Eg = c->dsfE[particle];
    /* 0) the row below the energy */
    int NumE = or_find_mono_1d(Eg, nE, Eel);
    if (NumE > 1) NumE = NumE - 1;
    /* 1) emission or absorption from the interpolated mean free paths */
    int iM = or_find_mono_1d(Eg, nE, Eel);
    if (iM > 1) iM = iM - 1;
    double Lt1 = c->dsfL[particle][(size_t)(iM - 1) * nW + (nW - 1)], Lt2 = c->dsfL[particle][(size_t)iM * nW + (nW - 1)];
    double Le1 = c->dsfLe[particle][(size_t)(iM - 1) * nW + (nW - 1)], Le2 = c->dsfLe[particle][(size_t)iM * nW + (nW - 1)];
    double La1 = c->dsfLa[particle][(size_t)(iM - 1) * nW + (nW - 1)], La2 = c->dsfLa[particle][(size_t)iM * nW + (nW - 1)];
    double EMFP_tot = lin(Lt1, Lt2, Eg[iM - 1], Eg[iM], Eel);
    double EMFP_emit = lin(Le1, Le2, Eg[iM - 1], Eg[iM], Eel);
    double EMFP_absorb = lin(La1, La2, Eg[iM - 1], Eg[iM], Eel);
    int emission = RN1 < EMFP_tot / EMFP_emit;
    /* 2) the row interpolated in the particle's energy, clamped at 0 */
    const double *Lsrc = emission ? c->dsfLe[particle] : c->dsfLa[particle];
    const double *dEsrc = c->dsfdE[particle];
    double *rowL = (double *)malloc(sizeof(double) * (size_t)nW);
    double *rowE = (double *)malloc(sizeof(double) * (size_t)nW);
    const double *La = Lsrc + (size_t)(NumE - 1) * nW, *Ea = dEsrc + (size_t)(NumE - 1) * nW;
    if (NumE == nE) {
        for (int j = 0; j < nW; ++j) { rowL[j] = La[j]; rowE[j] = Ea[j]; }
    } else {
        const double *Lb = La + nW, *Eb = Ea + nW;
        double Value1 = (Eel - Eg[NumE - 1]) / (Eg[NumE] - Eg[NumE - 1]);
        for (int j = 0; j < nW; ++j) {
            rowL[j] = La[j] + (Lb[j] - La[j]) * Value1;
            rowE[j] = Ea[j] + (Eb[j] - Ea[j]) * Value1;
        }
    }
    for (int j = 0; j < nW; ++j)
        if (rowL[j] < 0.0) rowL[j] = 0.0;
    double L_need = (emission ? EMFP_emit : EMFP_absorb) / RN2;
    double dE = approx_2x1d_dsf(rowL, rowE, nW, L_need);
    free(rowL);
    free(rowE);
    if (fabs(dE) > 1.0 && emission && dE > Eel) dE = Eel;
    return dE;
}
