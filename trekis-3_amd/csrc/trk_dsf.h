// trk_dsf.h -- elastic scattering of electrons and holes from dynamic
// structure factors (kind_of_EMFP = 2): the tables the caller hands over
// (DSF_DEMFP / DSF_DEMFP_H as the reference's reader leaves them) and the
// transfer sampling of NRG_transfer_elastic_DSF (Cross_sections.f90:
// 3652-3795).  A transfer dE > 0 heats the lattice (emission), dE < 0 takes
// energy from it (absorption).  FP64 with the reference's order of
// operations; the interpolated row of the sampling is formed entry by entry
// inside its search instead of in a temporary array.
#pragma once

#include "trk_dev_common.h"

namespace trk {

// One particle's DSF tables in device memory.  E and Lt are the elastic
// table's grid and total mean free paths (Elastic_MFP%Total), Le and La its
// emission and absorption parts (Analytical_IMFPs.f90:905-911); dE, dLe,
// dLa the nE x nW rows (E-major).  Plain data (kernel arguments and the
// batch's constant memory hold it): zero-initialise with DsfDev{}.
struct DsfDev {
    int nE, nW;
    const double *E, *Lt, *Le, *La;
    const double *dE, *dLe, *dLa;
};

// linear_interpolation, CS:3796-3800
TRK_HD double dsf_lin(double y1, double y2, double x1, double x2, double x)
{
    return y1 + ((y2 - y1) / (x2 - x1)) * (x - x1);
}

// entry j (0-based) of the row interpolated at Value1 between rows a and b
// (a == b: row a itself), the mean free path clamped at 0 (CS:3709-3740)
TRK_HD double dsf_row_L(const double *L, int a, int b, int nW, int j, double V)
{
    double v = (a == b) ? L[(size_t)a * nW + j] : L[(size_t)a * nW + j] + (L[(size_t)b * nW + j] - L[(size_t)a * nW + j]) * V;
    return (v < 0.0) ? 0.0 : v;
}
TRK_HD double dsf_row_dE(const double *dE, int a, int b, int nW, int j, double V)
{
    return (a == b) ? dE[(size_t)a * nW + j] : dE[(size_t)a * nW + j] + (dE[(size_t)b * nW + j] - dE[(size_t)a * nW + j]) * V;
}

// NRG_transfer_elastic_DSF for the two draws RN1 (emission or absorption)
// and RN2 (the mean free path to reach), with Linear_approx_2x1d_DSF
// (Reading_files_and_parameters.f90:3718-3745) on the interpolated row: the
// search of Find_in_monotonous_1D_array on the negated, decreasing
// cumulative path.  Not inlined: only the general transport kernel calls it.
__host__ __device__ __noinline__ inline double dsf_transfer(const DsfDev &t, double Eel, double RN1, double RN2)
{
    const int nE = t.nE, nW = t.nW;
    int NumE = find_mono_1d(t.E, nE, Eel);
    if (NumE > 1) NumE = NumE - 1;
    int iM = find_mono_1d(t.E, nE, Eel);
    if (iM > 1) iM = iM - 1;
    const double x1 = t.E[iM - 1], x2 = t.E[iM];
    const double EMFP_tot = dsf_lin(t.Lt[iM - 1], t.Lt[iM], x1, x2, Eel);
    const double EMFP_emit = dsf_lin(t.Le[iM - 1], t.Le[iM], x1, x2, Eel);
    const double EMFP_absorb = dsf_lin(t.La[iM - 1], t.La[iM], x1, x2, Eel);
    const bool emission = RN1 < EMFP_tot / EMFP_emit;
    // rows a, b (0-based) and the weight of b
    const int a = NumE - 1;
    const int b = (NumE == nE) ? a : NumE;
    const double V = (NumE == nE) ? 0.0 : (Eel - t.E[NumE - 1]) / (t.E[NumE] - t.E[NumE - 1]);
    const double *L = emission ? t.dLe : t.dLa;
    const double L_need = (emission ? EMFP_emit : EMFP_absorb) / RN2;
    // Find_in_monotonous_1D_array on -L(row), value -L_need
    const double v = -L_need;
    int Number;
    {
        int i1 = 1, i2 = nW;
        int icur = (i1 + i2) / 2;
        double tv = -dsf_row_L(L, a, b, nW, icur - 1, V);
        if (v != v) {
            icur = 0;
        } else if (v < -dsf_row_L(L, a, b, nW, 0, V)) {
            icur = 0;
        } else if (v >= -dsf_row_L(L, a, b, nW, nW - 1, V)) {
            icur = nW - 1;
        } else {
            while (i1 != i2 - 1) {
                if (tv <= v) i1 = icur;
                else i2 = icur;
                icur = (i1 + i2) / 2;
                tv = -dsf_row_L(L, a, b, nW, icur - 1, V);
            }
        }
        Number = icur + 1;
    }
    double dE;
    if (Number == 1) {
        const double L1 = dsf_row_L(L, a, b, nW, 0, V), L2 = dsf_row_L(L, a, b, nW, 1, V);
        const double e1 = dsf_row_dE(t.dE, a, b, nW, 0, V), e2 = dsf_row_dE(t.dE, a, b, nW, 1, V);
        dE = e1 + ((e2 - e1) / (L2 - L1)) * (L_need - L1);
    } else {
        const double Lm = dsf_row_L(L, a, b, nW, Number - 2, V), Ln = dsf_row_L(L, a, b, nW, Number - 1, V);
        const double em = dsf_row_dE(t.dE, a, b, nW, Number - 2, V);
        if (trk_fabs(Ln - Lm) < 1.0e-9) {
            dE = em;
        } else if (Lm > 1.0e20) {
            dE = em;
        } else {
            const double en = dsf_row_dE(t.dE, a, b, nW, Number - 1, V);
            dE = em + ((en - em) / (Ln - Lm)) * (L_need - Lm);
        }
    }
    if (trk_fabs(dE) > 1.0 && emission && dE > Eel) dE = Eel;
    return dE;
}

}  // namespace trk
