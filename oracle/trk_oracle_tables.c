/*
 * trk_oracle_tables.c -- CPU restatement of the TREKIS-3 table builders.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP table kernels).
 *
 * Each function follows the reference routine cited above it, keeping the
 * reference's node placement, loop structure and order of floating-point
 * operations (Fortran evaluates a*b/c*d left to right; the C below spells
 * every such chain out with parentheses).  Reference citations use
 * CS = Source_files/Cross_sections.f90, AI = Source_files/Analytical_IMFPs.f90,
 * RF = Source_files/Reading_files_and_parameters.f90,
 * MC = Source_files/Monte_Carlo.f90.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "trk_oracle_internal.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#define SQRT trk_sqrt
#define LOG  trk_log
#define EXP  trk_exp
#define POW  trk_pow

static double dmax(double a, double b) { return (a > b) ? a : b; }
static double dmin(double a, double b) { return (a < b) ? a : b; }

double or_constant(int which)
{
    switch (which) {
    case 0: return OR_PI;
    case 1: return OR_E;
    case 2: return OR_ME;
    case 3: return OR_MP;
    case 4: return OR_H;
    case 5: return OR_ALPHA;
    case 6: return OR_CVEL;
    case 7: return OR_A0;
    default: return 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Array searches (RF:3348-3559) -- exact replicas of the bisection paths;  */
/* indices are 1-based as in the reference.                                 */

/* Find_in_monotonous_1D_array, RF:3433-3494 */
int or_find_mono_1d(const double *A, int32_t N, double v)
{
    int i1 = 1, i2 = N;
    int icur = (i1 + i2) / 2;
    double tv = A[icur - 1];
    if (v != v) return -1;
    if (v < A[0]) {
        icur = 0;
    } else if (v >= A[N - 1]) {
        icur = N - 1;
    } else {
        for (;;) {
            if (i1 == i2 - 1) break;
            if (tv <= v) {
                i1 = icur;
                icur = (i1 + i2) / 2;
                tv = A[icur - 1];
            } else {
                i2 = icur;
                icur = (i1 + i2) / 2;
                tv = A[icur - 1];
            }
        }
    }
    return icur + 1;
}

/* Find_in_monotonous_2D_array, RF:3496-3559 (different exit test) */
int or_find_mono_2d(const double *A, int32_t N, double v)
{
    int i1 = 1, i2 = N;
    int icur = (i1 + i2) / 2;
    double tv = A[icur - 1];
    if (v != v) return -1;
    if (v < A[0]) {
        icur = 0;
    } else if (v >= A[N - 1]) {
        icur = N - 1;
    } else {
        for (;;) {
            if (v >= A[icur - 1] && v <= A[icur]) break;
            if (tv <= v) {
                i1 = icur;
                icur = (i1 + i2) / 2;
                tv = A[icur - 1];
            } else {
                i2 = icur;
                icur = (i1 + i2) / 2;
                tv = A[icur - 1];
            }
        }
    }
    return icur + 1;
}

/* Find_in_monoton_array_decreasing, RF:3376-3429 */
int or_find_decr(const double *A, int32_t N, double v)
{
    int i1 = 1, i2 = N;
    int icur = (i1 + i2) / 2;
    double tv = A[icur - 1];
    if (v != v) return -1;
    if (v < A[N - 1]) {
        icur = N;
    } else if (v > A[0]) {
        icur = 1;
    } else {
        while (abs(i1 - i2) > 1) {
            if (tv > v) {
                i1 = icur;
                icur = (i1 + i2) / 2;
                tv = A[icur - 1];
            } else {
                i2 = icur;
                icur = (i1 + i2) / 2;
                tv = A[icur - 1];
            }
        }
    }
    return icur;
}

/* Find_in_1D_array, RF:3348-3358: linear scan.  The reference reads past
 * the end when the value exceeds every element; here the scan stops at N. */
int or_find_linear(const double *A, int32_t N, double v)
{
    int i = 1;
    while (i < N && A[i - 1] < v) i++;
    return i;
}

/* Interpolate, CS:4051-4086 */
double or_interpolate(int flag, double E1, double E2, double S1, double S2, double E)
{
    if (trk_fabs(E2 - E1) < 1.0e-6) return (S1 > S2) ? S1 : S2;
    if (E == E1) return S1;
    switch (flag) {
    case 3: {
        double E2l = LOG(E2), E1l = LOG(E1), El = LOG(E);
        return S1 + ((S2 - S1) / (E2l - E1l)) * (El - E1l);
    }
    case 4: {
        double S1l = LOG(S1), S2l = LOG(S2);
        double o = S1l + ((S2l - S1l) / (E2 - E1)) * (E - E1);
        return EXP(o);
    }
    case 5: {
        double E2l = LOG(E2), E1l = LOG(E1), El = LOG(E);
        double S1l = LOG(S1), S2l = LOG(S2);
        double o = S1l + ((S2l - S1l) / (E2l - E1l)) * (El - E1l);
        return EXP(o);
    }
    default:
        return S1 + ((S2 - S1) / (E2 - E1)) * (E - E1);
    }
}

/* ------------------------------------------------------------------------ */
/* Loss function (CS:446-486) with dispersion (CS:346-393).                 */

static double lf_core(double A, double G1, double E0, double dE)
{
    double dE2 = dE * dE;
    double E02 = E0 * E0;
    return ((A * G1) * dE) / (((dE2 - E02) * (dE2 - E02)) + ((G1 * G1) * dE2));
}

/* mode 0: electron/hole in the target (mass from DOS or constant),
 * mode 1: photon (q = 0), mode 2: phonon (scattering on atoms, Mtarget). */
static double or_loss(const or_ctx *c, int mode, double A, double E, double G, double dE, double dq)
{
    double E0, G1;
    if (mode == 1) {
        G1 = G;
        E0 = E;
    } else {
        double hq2 = ((OR_H * OR_H) * dq) * dq;
        if (mode == 2) {
            E0 = E + hq2 / (2.0 * c->Mtarget);
            G1 = G;
        } else {
            double Mass;
            if (c->m.El_eff_mass == 0.0) {
                const double *k, *Em;
                int nk = c->m.n_dos;
                if (c->Egap < 0.2) { k = c->doskI; Em = c->dosMI; }
                else { k = c->dosk; Em = c->dosM; }
                double qlim = trk_fabs(dq) * SQRT(OR_E);
                if (qlim <= k[nk - 1]) {
                    int j = or_find_mono_1d(k, nk, qlim);
                    Mass = Em[j - 1];
                } else {
                    Mass = 1.0;
                }
            } else if (c->m.El_eff_mass > 0.0) {
                Mass = c->m.El_eff_mass;
            } else {
                Mass = 1.0;
            }
            double sqq = hq2 / ((2.0 * Mass) * OR_ME);
            switch (c->o.kind_of_DR) {
            case 2:
                E0 = SQRT(((E * E) + (((c->m.v_f * c->m.v_f) * hq2) * 0.3333333333333)) + (sqq * sqq));
                G1 = G;
                break;
            case 3:
                E0 = POW(POW(E, 0.666666666666) + POW(sqq, 0.666666666666), 1.5);
                G1 = SQRT((G * G) + (sqq * sqq));
                break;
            default:
                E0 = E + sqq;
                G1 = G;
            }
        }
    }
    return lf_core(A, G1, E0, dE);
}

double or_loss_func(const or_ctx *c, int mode, double A, double E0, double G, double hw, double hq)
{
    return or_loss(c, mode, A, E0, G, hw, hq);
}

/* Imewq, CS:397-442: sum over the oscillators of one shell. */
double or_imewq(const or_ctx *c, double hw, double hq, int shell, int photon)
{
    double ImE = 0.0;
    for (int i = c->osc_off[shell]; i < c->osc_off[shell + 1]; ++i)
        ImE = ImE + or_loss(c, photon ? 1 : 0, c->m.osc_A[i], c->m.osc_E0[i], c->m.osc_Gamma[i], hw, hq);
    return ImE;
}

/* Imewq_phonon, CS:3409-3433 */
static double or_imewq_phonon(const or_ctx *c, double hw, double hq)
{
    double ImE = 0.0;
    for (int i = 0; i < c->m.n_phonon; ++i)
        ImE = ImE + or_loss(c, 2, c->m.ph_A[i], c->m.ph_E0[i], c->m.ph_Gamma[i], hw, hq);
    return ImE;
}

/* ------------------------------------------------------------------------ */
/* q-integrals                                                              */

/* Diff_cross_section, CS:2217-2282 (Simpson over geometric q nodes with the
 * left-node 1/q weight and zero first left value). */
double or_diff_cs(const or_ctx *c, int shell, double Ee, double dE, double Mass)
{
    double qmin, qmax;
    double pre = SQRT((2.0 * Mass) * OR_ME) / OR_H;
    if (dE > Ee) {
        qmin = pre * SQRT(Ee);
        qmax = pre * SQRT(Ee);
    } else {
        qmin = pre * (SQRT(Ee) - SQRT(Ee - dE));
        qmax = pre * (SQRT(Ee) + SQRT(Ee - dE));
    }
    double dLs = 0.0, hq = qmin, dLs0 = 0.0;
    while (hq < qmax) {
        double dq = hq / 100.0;
        double a = hq + dq / 2.0;
        double temp1 = or_imewq(c, dE, a, shell, 0);
        double b = hq + dq;
        double dL = or_imewq(c, dE, b, shell, 0);
        dLs = dLs + ((dq / 6.0) * ((dLs0 + 4.0 * temp1) + dL)) / hq;
        dLs0 = dL;
        hq = hq + dq;
    }
    double T_fact;
    if (c->m.temp > 0.0) T_fact = 1.0 / (1.0 - EXP(-((dE / c->m.temp) * OR_KB)));
    else T_fact = 1.0;
    return ((1.0 / ((OR_PI * OR_A0) * Ee)) * dLs) * T_fact;
}

/* Diff_cross_section_phonon, CS:3142-3300 (screening = 1: CDF_elast_Zeff 0/1) */
double or_diff_cs_phonon(const or_ctx *c, double Ee, double dE, double Mass)
{
    const double eps = 1.0e-12;
    double qmin, qmax;
    double pre = SQRT((2.0 * Mass) * OR_ME) / OR_H;
    if (trk_fabs(dE) < eps) return 1.31e30;
    else if (dE > (Ee - eps)) qmin = pre * SQRT(Ee);
    else qmin = pre * (SQRT(Ee) - SQRT(trk_fabs(Ee - dE)));
    qmax = ((1.0 * SQRT((2.0 * Mass) * OR_ME)) / OR_H) * (SQRT(Ee) + SQRT(trk_fabs(Ee - dE)));
    double dLs = 0.0, hq = qmin, dLs0 = 0.0;
    while (trk_fabs(hq) < trk_fabs(qmax)) {
        double dq = hq / 100.0;
        double a = hq + dq / 2.0;
        double temp1 = or_imewq_phonon(c, dE, a);
        double b = hq + dq;
        double dL = or_imewq_phonon(c, dE, b);
        double Pot = 1.0 / hq;
        dLs = dLs + ((dq / 6.0) * ((dLs0 + 4.0 * temp1) + dL)) * Pot;
        dLs0 = dL;
        hq = hq + dq;
    }
    if (c->m.temp > 1.0e-6)
        return ((1.0 / ((OR_PI * OR_A0) * Ee)) * dLs) / (1.0 - EXP(-((dE / c->m.temp) * OR_KB)));
    return (1.0 / ((OR_PI * OR_A0) * Ee)) * dLs;
}

/* SHI_Diff_cross_section, CS:2683-2744 */
double or_shi_diff_cs(const or_ctx *c, int shell, double Ee, double MSHI, double Emax, double hw)
{
    double qmin, qmax;
    if (Ee > 0.0) qmin = (hw / OR_H) / SQRT((2.0 * Ee) / MSHI);
    else qmin = 0.0;
    if (Emax > 0.0) qmax = (SQRT(2.0 * OR_ME) / OR_H) * SQRT(Emax);
    else return 0.0;
    double dLs = 0.0, hq = qmin, dLs0 = 0.0;
    while (hq < qmax) {
        double dq = hq / 100.0;
        double a = hq + dq / 2.0;
        double temp1 = or_imewq(c, hw, a, shell, 0);
        double b = hq + dq;
        double dL = or_imewq(c, hw, b, shell, 0);
        dLs = dLs + ((dq / 6.0) * ((dLs0 + 4.0 * temp1) + dL)) / hq;
        dLs0 = dL;
        hq = hq + dq;
    }
    double T_fact;
    if (c->m.temp > 0.0) T_fact = 1.0 / (1.0 - EXP(-((hw / c->m.temp) * OR_KB)));
    else T_fact = 1.0;
    return dLs * T_fact;
}

/* ------------------------------------------------------------------------ */
/* omega-grid helpers (CS:1251-1412)                                         */

/* define_dE, CS:1374-1412 (new grid, both bounds present) */
double or_define_dE(int n, double E, double E0_min, double E0_max, double dE_min)
{
    double dE;
    if ((E > E0_min) && (E < E0_max)) dE = (E0_max - E0_min) / (double)n;
    else dE = E / (double)n;
    return dmax(dE, dE_min);
}

/* get_E_low, CS:1251-1274 */
static void get_E_low(double lo_b, double hi_b, double Emin, double Emax, double *lo, double *hi)
{
    if ((Emin > hi_b) || (Emax < lo_b)) {
        *lo = Emin;
        *hi = Emax;
    } else {
        *lo = dmax(lo_b, Emin);
        *hi = dmin(hi_b, Emax);
        if (trk_fabs(*lo - *hi) < 1.0e-3) {
            *lo = Emin;
            *hi = Emax;
        } else {
            double dE = hi_b - lo_b;
            *lo = dmin(*lo, *hi - dE);
            *lo = dmax(*lo, Emin);
        }
    }
}

/* get_diff_CS_grid_size, CS:1278-1343: node count from E_start = E0_min */
static int grid_count(int n, double Emax, double E0_min, double E0_max, double dE_min)
{
    double E = E0_min;
    int i = 0;
    while (E <= Emax) {
        i++;
        E = E + or_define_dE(n, E, E0_min, E0_max, dE_min);
    }
    return (i > 10) ? i : 10;
}

static void osc_window(const or_ctx *c, int shell, double *lo_b, double *hi_b)
{
    double lo = 0.0, hi = 0.0;
    for (int i = c->osc_off[shell]; i < c->osc_off[shell + 1]; ++i) {
        double l = c->m.osc_E0[i] - 5.0 * c->m.osc_Gamma[i];
        double h = c->m.osc_E0[i] + 5.0 * c->m.osc_Gamma[i];
        if (i == c->osc_off[shell] || l < lo) lo = l;
        if (i == c->osc_off[shell] || h > hi) hi = h;
    }
    *lo_b = lo;
    *hi_b = hi;
}

static double hole_mass_at(const or_ctx *c, double Ele)
{
    if (c->m.hole_mass >= 0.0) return c->m.hole_mass;
    int j = or_find_mono_1d(c->dosE, c->m.n_dos, Ele);
    return c->dosM[j - 1];
}

/* TotIMFP, CS:881-1050 (Ritchie CDF branch, CS_method = 1 tables).
 * kind: 0 electron, 1 hole, 2 photon-kind (unused here).  When `row` is not
 * NULL the cumulative row is produced (hw after each step, converted to
 * lambda_cum = 1/(Mass*int) with 1e15 sentinels, CS:1018-1038).  Row length
 * is max(fill count, count from E_low, 10): the reference allocates by
 * counting from E_low (CS:1120) but fills from Emin (CS:967-1002), which
 * overruns the allocation when E_low > Emin; the fill count is kept here. */
static void tot_imfp(const or_ctx *c, double Ele, int shell, int kind, double *Sigma, double *dEdx, or_row *row)
{
    double Emin = c->m.Ip[shell];
    double Egap = c->Egap;
    double Emax, Mass;
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    if (kind == 1) {
        Mass = hole_mass_at(c, Ele);
        Emax = ((4.0 * Ele) * Mass) / ((Mass + 1.0) * (Mass + 1.0));
    } else {
        Emax = (Ele + Emin) / 2.0;
        Mass = 1.0;
    }
    if (c->o.plasmon_Emax) {
        if (Emin == Egap) {
            double Ep = SQRT((((((c->m.N_VB_el * c->m.At_Dens) * 1.0e6) * OR_H) * OR_H) / (OR_ME * OR_E0)) + (Egap * Egap));
            if (Ep >= Emax) Emax = Ep;
            if (Ele < Emax) Emax = Ele;
        }
    }
    const int n = 100;
    double lo_b, hi_b;
    osc_window(c, shell, &lo_b, &hi_b);
    double E_low = dmax(lo_b, Emin);
    double E_high = dmin(hi_b, Emax);

    /* node count for the row allocation */
    int cap = 0;
    double *hw = NULL, *cum = NULL;
    if (row) {
        int cnt = 0;
        double E = Emin;
        while (E <= Emax) {
            cnt++;
            E = E + or_define_dE(n, E, E_low, E_high, 1.0e-3);
        }
        int nref = grid_count(n, Emax, E_low, E_high, 1.0e-3);
        cap = cnt;
        if (nref > cap) cap = nref;
        if (cap < 10) cap = 10;
        hw = (double *)calloc((size_t)cap, sizeof(double));
        cum = (double *)calloc((size_t)cap, sizeof(double));
    }

    int i = 0;
    double E = Emin;
    double Ltot1 = 0.0;
    double Ltot0 = or_diff_cs(c, shell, Ele, E, Mass);
    double ddEdx = 0.0;
    while (E <= Emax) {
        i++;
        double dE = or_define_dE(n, E, E_low, E_high, 1.0e-3);
        double a = E + dE / 2.0;
        double temp1 = or_diff_cs(c, shell, Ele, a, Mass);
        double b = E + dE;
        double dL = or_diff_cs(c, shell, Ele, b, Mass);
        double temp2 = (dE / 6.0) * ((Ltot0 + 4.0 * temp1) + dL);
        Ltot1 = Ltot1 + temp2;
        ddEdx = ddEdx + E * temp2;
        Ltot0 = dL;
        E = E + dE;
        if (row) {
            hw[i - 1] = E;
            cum[i - 1] = Ltot1;
        }
    }
    if (Ltot1 < 1.0e-15) {
        *Sigma = 1.0e20;
        *dEdx = 0.0;
    } else {
        *Sigma = 1.0 / (Mass * Ltot1);
        *dEdx = Mass * ddEdx;
    }
    /* the Delta-CDF model's closed forms in place of the quadrature's totals
     * (CS:952-956); the rows stay the quadrature's */
    if (c->o.kind_of_DR == 4 && kind <= 1) or_delta_tot_imfp(c, shell, Ele, Emin, Sigma, dEdx);
    /* a BEB shell's closed forms (TotIMFP's KOCS branch, CS:950, 1035-1041) */
    if (c->m.kocs[shell] == 2 && kind <= 1) or_beb_tot_imfp(c, shell, Ele, Mass, Sigma, dEdx);
    if (row) {
        for (int k = 0; k < cap; ++k) {
            if (cum[k] < 1.0e-15) cum[k] = 1.0e15;
            else cum[k] = 1.0 / (Mass * cum[k]);
        }
        row->n = cap;
        row->hw = hw;
        row->lam = cum;
    }
}

/* Equilibrium_charge_Target Barkas branch, CS:2601-2638 */
static double zeff_target(const or_ctx *c, double Ee)
{
    if (c->o.CDF_elast_Zeff != 0) return 1.0;
    double Zp = c->Zt;
    double vp = SQRT(((2.0 * Ee) * OR_E) / OR_ME);
    return 1.0 + Zp * (1.0 - EXP(-(((vp * 125.0) / OR_CVEL) / POW(Zp, 0.66666666))));
}

/* Tot_EMFP, CS:2966-3139, with the phonon-CDF diff. cross section. */
static void tot_emfp(const or_ctx *c, double Ele, int kind, double *Sigma, double *dEdx, or_row *row)
{
    double qd = POW(((6.0 * OR_PI) * OR_PI) * (c->m.At_Dens * 1.0e6), 0.33333333);
    double Edebay = ((OR_H * c->m.Vsound) * qd) / OR_E;
    Edebay = Edebay * 3.0;
    double Emin = 1.0 * 0.1e-8;
    double Mtarget = c->Mtarget;
    double Mass = (kind == 1) ? hole_mass_at(c, Ele) : 1.0;
    double Zeff = zeff_target(c, Ele);
    double Emax = ((((4.0 * Ele) * Mass) * OR_ME) * Mtarget) /
                  ((Mtarget + Mass * OR_ME) * (Mtarget + Mass * OR_ME));
    if (Edebay >= Emax) Emax = Edebay;
    if (Ele < Emax) Emax = Ele;
    Emax = 1.0 * Emax;
    const int n = 20;
    double lo_b = 0.0, hi_b = 0.0;
    for (int i = 0; i < c->m.n_phonon; ++i) {
        double l = c->m.ph_E0[i] - 5.0 * c->m.ph_Gamma[i];
        double h = c->m.ph_E0[i] + 5.0 * c->m.ph_Gamma[i];
        if (i == 0 || l < lo_b) lo_b = l;
        if (i == 0 || h > hi_b) hi_b = h;
    }
    double E_low, E_high;
    get_E_low(lo_b, hi_b, Emin, Emax, &E_low, &E_high);

    int cap = 0;
    double *hw = NULL, *cum = NULL;
    if (row) {
        int cnt = 0;
        double E = E_low;
        while (trk_fabs(E) <= trk_fabs(Emax)) {
            cnt++;
            E = E + 1.0 * or_define_dE(n, E, E_low, E_high, 1.0e-5);
        }
        int nref = grid_count(n, Emax, E_low, E_high, 1.0e-5);
        cap = (cnt > nref) ? cnt : nref;
        if (cap < 10) cap = 10;
        hw = (double *)calloc((size_t)cap, sizeof(double));
        cum = (double *)calloc((size_t)cap, sizeof(double));
    }
    int i = 0;
    double E = E_low;
    double Ltot1 = 0.0;
    double Ltot0 = or_diff_cs_phonon(c, Ele, E, Mass);
    double ddEdx = 0.0;
    while (trk_fabs(E) <= trk_fabs(Emax)) {
        i++;
        double dE = or_define_dE(n, E, E_low, E_high, 1.0e-5);
        dE = 1.0 * dE;
        double a = E + dE / 2.0;
        double temp1 = or_diff_cs_phonon(c, Ele, a, Mass);
        double b = E + dE;
        double dL = or_diff_cs_phonon(c, Ele, b, Mass);
        double temp2 = (trk_fabs(dE) / 6.0) * ((Ltot0 + 4.0 * temp1) + dL);
        Ltot1 = Ltot1 + temp2;
        ddEdx = ddEdx + E * temp2;
        Ltot0 = dL;
        E = E + dE;
        if (row) {
            hw[i - 1] = E;
            cum[i - 1] = Ltot1;
        }
    }
    if ((Mass < 1.0e-20) || (Ltot1 < 1.0e-15)) *Sigma = 1.0e29;
    else *Sigma = 1.0 / (((Zeff * Zeff) * Mass) * Ltot1);
    if (*Sigma > 1.0e30) *Sigma = 1.0e30;
    *dEdx = ((Zeff * Zeff) * Mass) * ddEdx;
    if (row) {
        for (int k = 0; k < cap; ++k) {
            if (cum[k] < 1.0e-15) cum[k] = 1.0e20;
            else cum[k] = 1.0 / (((Zeff * Zeff) * Mass) * cum[k]);
        }
        row->n = cap;
        row->hw = hw;
        row->lam = cum;
    }
}

/* Atomic_elastic_sigma, CS:3495-3514 (Mott, [cm^2]) */
static double mott_sigma(const or_ctx *c, int atom, double Ee)
{
    double Zat = (double)c->m.Zat[atom];
    double mec2e = ((OR_ME * OR_CVEL) * OR_CVEL) / OR_E;
    double Zat137 = Zat / 137.0;
    double RyEe = OR_RY / Ee;
    double beta2 = (2.0 * Ee) / mec2e;
    double pc = ((1.7e-5 * POW(Zat, 2.0 / 3.0)) * (1.0 - beta2)) / beta2;
    double nc = pc * (1.13 + (((3.76 * (Zat137 * Zat137)) / beta2) * SQRT(Ee / (Ee + mec2e))));
    return (((((((OR_PI * OR_A0) * OR_A0) * Zat) * (Zat + 1.0)) / (nc * (nc + 1.0))) * RyEe) * RyEe) * 1.0e-16;
}

/* Elastic_cross_section, CS:2892-2963 (plus the disabled case AI:1125) */
static void elastic_cs(const or_ctx *c, double Ee, int kind, double *EMFP, double *dEdx, or_row *row)
{
    *EMFP = 1.34e16;
    *dEdx = 0.0;
    if (c->o.kind_of_EMFP == 1) {
        tot_emfp(c, Ee, kind, EMFP, dEdx, row);
    } else if (c->o.kind_of_EMFP == 0) {
        double Mass = (kind == 1) ? hole_mass_at(c, Ee) : 1.0;
        double S = 0.0;
        for (int i = 0; i < c->m.n_atoms; ++i) S = S + mott_sigma(c, i, Ee) * c->m.Pers[i];
        S = (S / c->sumPers) * Mass;
        *EMFP = 1.0e8 / (S * c->m.At_Dens);
    } else {
        *EMFP = 1.0e30;
    }
}

/* Tot_Phot_IMFP, CS:833-878 (CDF branch) */
static void tot_phot(const or_ctx *c, double Ele, int shell, double *Sigma, double *dEdx)
{
    if (Ele < c->m.Ip[shell]) {
        *Sigma = 1.0e30;
        *dEdx = Ele / *Sigma;
    } else {
        double ImE = or_imewq(c, Ele, 0.0, shell, 1);
        *Sigma = ((OR_CVEL * OR_H) / ((ImE * Ele) * OR_E)) * 1.0e10;
        *dEdx = Ele / *Sigma;
    }
}

/* Equilibrium_charge_SHI, CS:2641-2680 */
double or_shi_zeff(const or_ctx *c, double E)
{
    double vp = (E > 0.0) ? SQRT(((2.0 * E) * OR_E) / (c->ion.Mass * OR_MP)) : 0.0;
    double Zt = c->Zt;
    double Zp = (double)c->ion.Zat;
    double v0 = SQRT(((2.0 * OR_RY) * OR_E) / OR_ME);
    switch (c->ion.Kind_Zeff) {
    case 1:
        return Zp * (1.0 - EXP(-((vp / v0) / POW(Zp, 0.66666666))));
    case 2: {
        double c1 = 0.6, c2 = 0.45;
        return Zp * POW(1.0 + POW(vp / (((POW(Zp, c2) * v0) * 4.0) / 3.0), -1.0 / c1), -c1);
    }
    case 3: {
        double c1 = 1.0 - 0.26 * EXP(-(Zt / 11.0) - ((Zt - Zp) * (Zt - Zp)) / 9.0);
        double vpvo = (POW(Zp, -0.543) * vp) / v0;
        double c2 = 1.0 + (0.03 * vpvo) * LOG(Zt);
        double x = c1 * POW((vpvo / c2) / 1.54, 1.0 + 1.83 / Zp);
        double x2 = x * x, x4 = x2 * x2;
        return (Zp * ((8.29 * x) + x4)) / ((((0.06 / x) + 4.0) + (7.4 * x)) + x4);
    }
    case 4:
        return c->ion.fixed_Zeff;
    default:
        return Zp * (1.0 - EXP(-(((vp * 125.0) / OR_CVEL) / POW(Zp, 0.66666666))));
    }
}

/* Brand_Kitagawa, CS:2878-2889 */
static double bk_charge(double hq, double Zp, double Zeff)
{
    double a = 0.2400519147;
    double Z = (Zp - Zeff) / Zp;
    double Z23 = (Z == 0.0) ? 0.0 : POW(Z, 2.0 / 3.0);
    double kl = ((((hq * ((OR_A0 * 1.0e-10) * SQRT(OR_E))) * 2.0) * a) * Z23) / (POW(Zp, 2.0 / 3.0) * (1.0 - Z / 7.0));
    double kl2 = kl * kl;
    return (Zp * ((1.0 - Z) + kl2)) / (1.0 + kl2);
}

/* SHI_Diff_cross_section_BK, CS:2826-2875 */
static double shi_diff_cs_bk(const or_ctx *c, int shell, double Ee, double MSHI, double Emax, double hw, double Zeff)
{
    double Zp = (double)c->ion.Zat;
    double qmin = (hw / OR_H) / SQRT((2.0 * Ee) / MSHI);
    double qmax = (SQRT(2.0 * OR_ME) / OR_H) * SQRT(Emax);
    double dLs = 0.0, hq = qmin, dLs0 = 0.0;
    while (hq < qmax) {
        double dq = hq / 100.0;
        double a = hq + dq / 2.0;
        double rho = bk_charge(a, Zp, Zeff);
        double temp1 = (or_imewq(c, hw, a, shell, 0) * rho) * rho;
        double b = hq + dq;
        rho = bk_charge(b, Zp, Zeff);
        double dL = (or_imewq(c, hw, b, shell, 0) * rho) * rho;
        dLs = dLs + ((dq / 6.0) * ((dLs0 + 4.0 * temp1) + dL)) / hq;
        dLs0 = dL;
        hq = hq + dq;
    }
    if (c->m.temp > 0.0) return dLs / (1.0 - EXP(-((hw / c->m.temp) * OR_KB)));
    return dLs;
}

double or_shi_diff_cs_bk(const or_ctx *c, int shell, double Ee, double MSHI, double Emax, double hw)
{
    return shi_diff_cs_bk(c, shell, Ee, MSHI, Emax, hw, or_shi_zeff(c, Ee));
}

double or_bk_charge(double hq, double Zp, double Zeff) { return bk_charge(hq, Zp, Zeff); }

/* SHI_TotIMFP_BK, CS:2748-2823: the ion's mean free path with the
 * Brandt-Kitagawa charge inside the q integral (grid as SHI_TotIMFP; the
 * stopping power sums with the left node of each step). */
static void shi_tot_imfp_bk(const or_ctx *c, double Ele, int shell, double *Sigma, double *dEdx)
{
    double MSHI = OR_MP * c->ion.Mass;
    double Zeff = or_shi_zeff(c, Ele);
    double Egap = c->Egap;
    double Emin = c->m.Ip[shell];
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    double Emax = (((4.0 * Ele) * OR_ME) * MSHI) / ((MSHI + OR_ME) * (MSHI + OR_ME));
    if (c->o.plasmon_Emax) {
        if (Emin == Egap) {
            double Ep = SQRT((((((c->m.N_VB_el * c->m.At_Dens) * 1.0e6) * OR_H) * OR_H) / (OR_ME * OR_E0)) + (Egap * Egap));
            if (Ep >= Emax) Emax = Ep;
        }
    }
    const int n = 1000;
    double lo_b, hi_b;
    osc_window(c, shell, &lo_b, &hi_b);
    double E_low = dmax(lo_b, Emin);
    double E_high = dmin(hi_b, Emax);
    double E = Emin, Ltot1 = 0.0, ddEdx = 0.0;
    double Ltot0 = shi_diff_cs_bk(c, shell, Ele, MSHI, Emax, E, Zeff);
    while (E <= Emax) {
        double dE = or_define_dE(n, E, E_low, E_high, 1.0e-3);
        double a = E + dE / 2.0;
        double temp1 = shi_diff_cs_bk(c, shell, Ele, MSHI, Emax, a, Zeff);
        double b = E + dE;
        double dL = shi_diff_cs_bk(c, shell, Ele, MSHI, Emax, b, Zeff);
        double temp2 = (dE / 6.0) * ((Ltot0 + 4.0 * temp1) + dL);
        Ltot1 = Ltot1 + temp2;
        ddEdx = ddEdx + E * temp2;
        Ltot0 = dL;
        E = E + dE;
    }
    double pre = (1.0 / ((OR_PI * OR_A0) * Ele)) * MSHI;
    *Sigma = (pre / OR_ME) * Ltot1;
    if (*Sigma > 1.0e30) *Sigma = 1.0e30;
    *dEdx = (pre / OR_ME) * ddEdx;
}

/* SHI_TotIMFP, CS:2452-2597 (CDF branch).  If dn/dE/dL are given the
 * cumulative row dSedE (CS:2577-2582) is produced; its E entry is the left
 * node of each step, as in the reference. */
static void shi_tot_imfp(const or_ctx *c, double Ele, int shell, double *Sigma, double *dEdx,
                         int32_t *dn, double **dEo, double **dLo)
{
    double MSHI = OR_MP * c->ion.Mass;
    double Zeff = or_shi_zeff(c, Ele);
    double Egap = c->Egap;
    double Emin = c->m.Ip[shell];
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    double Emax = (((4.0 * Ele) * OR_ME) * MSHI) / ((MSHI + OR_ME) * (MSHI + OR_ME));
    if (c->o.plasmon_Emax) {
        if (Emin == Egap) {
            double Ep = SQRT((((((c->m.N_VB_el * c->m.At_Dens) * 1.0e6) * OR_H) * OR_H) / (OR_ME * OR_E0)) + (Egap * Egap));
            if (Ep >= Emax) Emax = Ep;
        }
    }
    const int n = 1000;
    double lo_b, hi_b;
    osc_window(c, shell, &lo_b, &hi_b);
    double E_low = dmax(lo_b, Emin);
    double E_high = dmin(hi_b, Emax);

    int cap = 0;
    double *rE = NULL, *rL = NULL;
    if (dn) {
        double E0 = c->m.Ip[shell];
        double E_low0 = dmax(lo_b, E0);
        double E_high0 = dmin(hi_b, Emax);
        int k0 = 0;
        while (E0 <= Emax) {
            k0++;
            E0 = E0 + or_define_dE(n, E0, E_low0, E_high0, 1.0e-3);
        }
        cap = k0;
        rE = (double *)calloc((size_t)(cap > 0 ? cap : 1), sizeof(double));
        rL = (double *)calloc((size_t)(cap > 0 ? cap : 1), sizeof(double));
    }
    double Ltot1 = 0.0;
    double E = Emin;
    double Ltot0 = or_shi_diff_cs(c, shell, Ele, MSHI, Emax, E);
    double ddEdx = 0.0;
    int i = 0;
    while (E <= Emax) {
        i++;
        if (dn && i > cap) break;
        double dE = or_define_dE(n, E, E_low, E_high, 1.0e-3);
        double a = E + dE / 2.0;
        double temp1 = or_shi_diff_cs(c, shell, Ele, MSHI, Emax, a);
        double b = E + dE;
        double dL = or_shi_diff_cs(c, shell, Ele, MSHI, Emax, b);
        double temp2 = (dE / 6.0) * ((Ltot0 + 4.0 * temp1) + dL);
        Ltot1 = Ltot1 + temp2;
        ddEdx = ddEdx + (dE / 6.0) * (((E * Ltot0) + ((a * 4.0) * temp1)) + (b * dL));
        Ltot0 = dL;
        if (dn) {
            rE[i - 1] = E;
            rL[i - 1] = (((OR_PI * OR_A0) * Ele) * OR_ME) / (((MSHI * Zeff) * Zeff) * Ltot1);
        }
        E = E + dE;
    }
    double pre = (1.0 / ((OR_PI * OR_A0) * Ele)) * MSHI;
    *Sigma = (((pre / OR_ME) * Zeff) * Zeff) * Ltot1;
    if (*Sigma > 1.0e30) *Sigma = 1.0e30;
    *dEdx = (((pre / OR_ME) * Zeff) * Zeff) * ddEdx;
    if (dn) {
        *dn = cap;
        *dEo = rE;
        *dLo = rL;
    }
}

/* ------------------------------------------------------------------------ */
/* Energy grid: get_grid_4CS / go_thru_grid / define_special_points,        */
/* AI:2815-2987, find_order_of_number_real AI:3077-3093.                     */

static int ndigits_i64(int64_t v)
{
    int d = 1;
    if (v < 0) { d++; v = -v; }
    while (v >= 10) { v /= 10; d++; }
    return d;
}

static int find_order_real(double num)
{
    if (num > 1e9) {
        double r = num / 1.0e8;
        return 8 + ndigits_i64((int64_t)ceil(r));
    }
    return ndigits_i64((int64_t)ceil(num));
}

static double pow10i(int k)
{
    double r = 1.0;
    if (k >= 0) { for (int i = 0; i < k; ++i) r = r * 10.0; return r; }
    for (int i = 0; i < -k; ++i) r = r * 10.0;
    return 1.0 / r;
}

static void bubble_sort(double *a, int n)
{
    for (int j = n - 1; j >= 1; --j) {
        int swapped = 0;
        for (int i = 0; i < j; ++i) {
            if (a[i] > a[i + 1]) {
                double t = a[i]; a[i] = a[i + 1]; a[i + 1] = t;
                swapped = 1;
            }
        }
        if (!swapped) break;
    }
}

static int special_points(const trk_material *m, double *sp)
{
    int ns = 0;
    for (int a = 0, s = 0; a < m->n_atoms; ++a)
        for (int k = 0; k < m->n_shells[a]; ++k, ++s) sp[ns++] = m->Ip[s];
    bubble_sort(sp, ns);
    /* exclude_doubles, AI:2991-3030 */
    int ind[TRK_MAX_SHELLS * TRK_MAX_SHELLS];
    int ci = 0;
    for (int i = 0; i < ns - 1; ++i)
        for (int k = i + 1; k < ns; ++k)
            if (sp[i] == sp[k]) ind[ci++] = (i < k ? i : k) + 1;
    int ndub = ci;
    double copy[TRK_MAX_SHELLS];
    memcpy(copy, sp, sizeof(double) * (size_t)ns);
    int coun = 0, cidx = 0;
    int newn = ns - ndub;
    for (int i = 1; i <= ns; ++i) {
        if (cidx < ci && i == ind[cidx]) {
            cidx++;
        } else if (coun < newn) {
            sp[coun++] = copy[i - 1];
        }
    }
    return coun;
}

static int go_thru_grid(double Emin, double Emax, const double *sp, int nsp, double *arr, int cap)
{
    const double E_sp_eps = 1.0e-3, scale = 1.0;
    const double dE_min = 0.01 * scale;
    int SP = 1, N = 0;
    double E_cur = Emin - dE_min;
    while (E_cur < Emax) {
        N++;
        double dE;
        if (E_cur < 0.1 - dE_min * 0.5) dE = dE_min;
        else if (E_cur < 10.0 - dE_min * 0.5) dE = dE_min * 10.0;
        else if (E_cur < 100.0 - dE_min * 0.5) dE = dE_min * 100.0;
        else dE = pow10i(find_order_real(E_cur) - 2);
        E_cur = E_cur + dE;
        if (arr) {
            if (N > cap) return -1;
            arr[N - 1] = E_cur;
        }
        int here = 0;
        if (SP <= nsp) {
            if ((E_cur >= sp[SP - 1]) && ((E_cur - dE) < sp[SP - 1])) {
                SP++;
                here = 1;
            }
        }
        if (here) {
            if (arr) {
                if (N + 2 > cap) return -1;
                N++;
                arr[N - 1] = sp[SP - 2] - E_sp_eps;
                N++;
                arr[N - 1] = sp[SP - 2] + E_sp_eps;
            } else {
                N += 2;
            }
        }
    }
    if (arr) bubble_sort(arr, N);
    return N;
}

int or_grid(const trk_material *m, double Emin, double Emax, double *out, int32_t cap, int32_t *n)
{
    double sp[TRK_MAX_SHELLS];
    int nsp = special_points(m, sp);
    int N = go_thru_grid(Emin, Emax, sp, nsp, NULL, 0);
    *n = N;
    if (!out) return 0;
    if (N > cap) return TRK_ERR_CAPACITY;
    go_thru_grid(Emin, Emax, sp, nsp, out, cap);
    return 0;
}

static double *grid_alloc(const or_ctx *c, double Emin, double Emax, int32_t *n)
{
    int32_t N;
    or_grid(&c->m, Emin, Emax, NULL, 0, &N);
    double *g = (double *)malloc(sizeof(double) * (size_t)N);
    or_grid(&c->m, Emin, Emax, g, N, n);
    return g;
}

/* ------------------------------------------------------------------------ */
/* Sum rules, CS:712-826                                                     */

double or_w_plasma(double dens, double mass)
{
    if (mass > 0.0) return ((dens * OR_E) * OR_E) / (OR_E0 * mass);
    return ((dens * OR_E) * OR_E) / (OR_E0 * OR_ME);
}

/* complex helpers for the closed-form antiderivatives */
typedef struct { double re, im; } cplx;
static cplx cmk(double r, double i) { cplx z = { r, i }; return z; }
static cplx cadd(cplx a, cplx b) { return cmk(a.re + b.re, a.im + b.im); }
static cplx csub(cplx a, cplx b) { return cmk(a.re - b.re, a.im - b.im); }
static cplx cmul(cplx a, cplx b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cplx cdivz(cplx a, cplx b)
{
    double d = b.re * b.re + b.im * b.im;
    return cmk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
static cplx csqrtz(cplx z)
{
    double r = sqrt(z.re * z.re + z.im * z.im);
    double x = sqrt(0.5 * (r + fabs(z.re)));
    if (x == 0.0) return cmk(0.0, 0.0);
    if (z.re >= 0.0) return cmk(x, z.im / (2.0 * x));
    return cmk(fabs(z.im) / (2.0 * x), (z.im >= 0.0) ? x : -x);
}
static cplx clogz(cplx z) { return cmk(log(hypot(z.re, z.im)), atan2(z.im, z.re)); }
static cplx cscale(cplx a, double s) { return cmk(a.re * s, a.im * s); }

/* Int_Ritchi_x, CS:763-792 */
double or_int_ritchi_x(double A, double E, double G, double x)
{
    double sq2 = sqrt(2.0);
    cplx I = cmk(0.0, 1.0);
    cplx Sc;
    double d = -(2.0 * E) * (2.0 * E) + G * G;
    if (d >= 0.0) Sc = cmk(sqrt(d), 0.0);
    else Sc = cmk(0.0, sqrt(fabs(d)));
    cplx Gc = cmk(G * G - 2.0 * E * E, 0.0);
    cplx s_plus = cscale(csqrtz(cadd(Gc, cscale(Sc, G))), sq2);
    cplx s_minus = cscale(csqrtz(csub(Gc, cscale(Sc, G))), sq2);
    cplx GS = cscale(Sc, G);
    cplx Bc = (GS.re == 0.0 && GS.im == 0.0) ? cmk(0.0, 0.0) : cdivz(Gc, GS);
    cplx arg = cdivz(cmk(2.0 * x, 0.0), s_minus);
    cplx arg2 = cdivz(cmk(2.0 * x, 0.0), s_plus);
    cplx one = cmk(1.0, 0.0);
    cplx t1 = cmul(cscale(I, 0.5), csub(clogz(csub(one, cmul(I, arg))), clogz(cadd(one, cmul(I, arg)))));
    cplx t2 = cmul(cscale(I, 0.5), csub(clogz(csub(one, cmul(I, arg2))), clogz(cadd(one, cmul(I, arg2)))));
    cplx a1 = cmul(cdivz(t1, s_minus), csub(one, Bc));
    cplx a2 = cmul(cdivz(t2, s_plus), cadd(one, Bc));
    cplx Ic = cscale(cadd(a1, a2), A * G);
    return Ic.re;
}

/* Int_Ritchi_p_x, CS:795-826 */
double or_int_ritchi_p_x(double A, double E, double G, double x)
{
    double sq2 = sqrt(2.0);
    cplx I = cmk(0.0, 1.0);
    cplx Sc;
    double d = -(2.0 * E) * (2.0 * E) + G * G;
    if (d >= 0.0) Sc = cmk(sqrt(d), 0.0);
    else Sc = cmk(0.0, sqrt(fabs(d)));
    cplx Gc = cmk(G * G - 2.0 * E * E, 0.0);
    cplx s_plus = csqrtz(cadd(Gc, cscale(Sc, G)));
    cplx s_minus = csqrtz(csub(Gc, cscale(Sc, G)));
    cplx arg = cdivz(cmk(sq2 * x, 0.0), s_minus);
    cplx arg2 = cdivz(cmk(sq2 * x, 0.0), s_plus);
    cplx one = cmk(1.0, 0.0);
    cplx t1 = cdivz(cmul(cscale(I, 0.5), csub(clogz(csub(one, cmul(I, arg))), clogz(cadd(one, cmul(I, arg))))), s_minus);
    cplx t2 = cdivz(cmul(cscale(I, 0.5), csub(clogz(csub(one, cmul(I, arg2))), clogz(cadd(one, cmul(I, arg2))))), s_plus);
    cplx diff = csub(t1, t2);
    cplx In;
    if (diff.re == 0.0 && diff.im == 0.0) In = cmk(sq2 * A, 0.0);
    else In = cmul(cdivz(cmk(sq2 * A, 0.0), Sc), diff);
    return In.re;
}

/* sumrules, CS:728-749 */
int or_sumrules(const double *A, const double *E0, const double *G, int32_t n,
                double x_min, double Omega, double *ksum, double *fsum)
{
    double ne = 0.0, f = 0.0;
    for (int j = 0; j < n; ++j) {
        ne = ne + or_int_ritchi_x(A[j], E0[j], G[j], 1e10) - or_int_ritchi_x(A[j], E0[j], G[j], x_min);
        f = f + or_int_ritchi_p_x(A[j], E0[j], G[j], 1e10) - or_int_ritchi_p_x(A[j], E0[j], G[j], x_min);
    }
    *ksum = ((2.0 * OR_E) * OR_E) / ((OR_PI * Omega) * OR_H * OR_H) * ne;
    *fsum = (f * 2.0) / OR_PI;
    return 0;
}

/* reading_material_DOS, RF:2374-2450 (derived arrays only), with the
 * linear lookup Linear_approx_2d, RF:3272-3306. */
int or_prepare_dos(const double *raw_E, const double *raw_DOS, int32_t n_raw,
                   double N_VB_el, double At_Dens, double sum_Pers,
                   double *E, double *DOS, double *int_DOS, double *k, double *Eff_m,
                   double *k_inv, double *Eff_m_inv, int32_t M)
{
    double *T0 = (double *)malloc(sizeof(double) * (size_t)n_raw);
    for (int i = 0; i < n_raw; ++i) T0[i] = raw_E[i] - raw_E[n_raw - 1];
    const double dE = 0.1;
    int Mc = (int)floor((T0[n_raw - 1] - T0[0]) / dE);
    if (Mc != M) { free(T0); return TRK_ERR_ARG; }
    double El1 = T0[0] - dE, El2 = 0.0;
    double Ecur = T0[0];
    for (int i = 0; i < M; ++i) {
        Ecur = Ecur + dE;
        double loc;
        int N = or_find_mono_2d(T0, n_raw, Ecur);
        if (N == 1) {
            loc = El2 + ((raw_DOS[0] - El2) / (T0[0] - El1)) * (Ecur - El1);
        } else if (raw_DOS[N - 2] > 1.0e20) {
            loc = raw_DOS[N - 2];
        } else {
            loc = raw_DOS[N - 2] + ((raw_DOS[N - 1] - raw_DOS[N - 2]) / (T0[N - 1] - T0[N - 2])) * (Ecur - T0[N - 2]);
        }
        E[i] = Ecur;
        DOS[i] = loc;
    }
    double Elast = E[M - 1];
    for (int i = 0; i < M; ++i) E[i] = trk_fabs(E[i] - Elast);
    for (int i = 0; i < M / 2; ++i) {
        double t = E[i]; E[i] = E[M - 1 - i]; E[M - 1 - i] = t;
        t = DOS[i]; DOS[i] = DOS[M - 1 - i]; DOS[M - 1 - i] = t;
    }
    double *Dinv = (double *)malloc(sizeof(double) * (size_t)M);
    double *Iinv = (double *)malloc(sizeof(double) * (size_t)M);
    for (int i = 0; i < M; ++i) Dinv[i] = DOS[M - 1 - i];
    double S = 0.0, Si = 0.0;
    for (int i = 0; i < M; ++i) {
        S = S + DOS[i];
        int_DOS[i] = S;
        Si = Si + Dinv[i];
        Iinv[i] = Si;
    }
    double SUM = int_DOS[M - 1];
    for (int i = 0; i < M; ++i) {
        DOS[i] = (DOS[i] / SUM) * N_VB_el;
        int_DOS[i] = (int_DOS[i] / SUM) * N_VB_el;
    }
    const double kpre = (((3.0 * 2.0) * OR_PI) * OR_PI) / 2.0;
    S = 0.0;
    for (int i = 0; i < M; ++i) {
        S = S + DOS[i];
        k[i] = POW((((kpre * S) * At_Dens) / sum_Pers) * 1.0e6, 1.0 / 3.0);
        if (E[i] < 1.0e-10) Eff_m[i] = 1.0;
        else Eff_m[i] = ((((OR_H * OR_H) * k[i]) * k[i]) / ((2.0 * E[i]) * OR_E)) / OR_ME;
    }
    double SUMI = Iinv[M - 1];
    for (int i = 0; i < M; ++i) {
        Iinv[i] = (Iinv[i] / SUMI) * N_VB_el;
        k_inv[i] = POW((((kpre * Iinv[i]) * At_Dens) / sum_Pers) * 1.0e6, 1.0 / 3.0);
        if (E[i] < 1.0e-10) Eff_m_inv[i] = 1.0;
        else Eff_m_inv[i] = ((((OR_H * OR_H) * k_inv[i]) * k_inv[i]) / ((2.0 * E[i]) * OR_E)) / OR_ME;
    }
    free(T0);
    free(Dinv);
    free(Iinv);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Context                                                                   */

static double *dupd(const double *p, int n)
{
    if (!p || n <= 0) return NULL;
    double *q = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(q, p, sizeof(double) * (size_t)n);
    return q;
}

or_ctx *or_create(const trk_material *mat, const trk_ion *ion, const trk_options *opt)
{
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    c->m = *mat;
    c->ion = *ion;
    c->o = *opt;
    int nd = mat->n_dos;
    c->dosE = dupd(mat->dos_E, nd);
    c->dos = dupd(mat->dos_DOS, nd);
    c->dosI = dupd(mat->dos_int_DOS, nd);
    c->dosk = dupd(mat->dos_k, nd);
    c->dosM = dupd(mat->dos_Eff_m, nd);
    c->doskI = dupd(mat->dos_k_inv, nd);
    c->dosMI = dupd(mat->dos_Eff_m_inv, nd);
    int s = 0, o = 0;
    for (int a = 0; a < mat->n_atoms; ++a) {
        c->shl_off[a] = s;
        for (int k = 0; k < mat->n_shells[a]; ++k, ++s) {
            c->atom_of[s] = a;
            c->osc_off[s] = o;
            o += mat->n_osc[s];
        }
    }
    c->shl_off[mat->n_atoms] = s;
    c->osc_off[s] = o;
    /* Delta-CDF weights: each oscillator's k-sum above its shell's edge
     * (define_alpha, RF:1385-1388, 1572-1574) */
    if (opt->kind_of_DR == 4)
        for (int sh = 0; sh < s; ++sh)
            for (int k = c->osc_off[sh]; k < c->osc_off[sh + 1]; ++k)
                c->alpha[k] = or_delta_alpha(mat->osc_A[k], mat->osc_E0[k], mat->osc_Gamma[k], mat->Ip[sh]);
    c->nsh = s;
    c->vb = mat->n_shells[0] - 1;
    c->Egap = mat->Ip[c->vb];
    /* check_atomic_parameters, EADL:350-351, 366-367: no radiative decay of
     * the valence band or without photons, no Auger decay of the valence band */
    for (int k = 0; k < s; ++k) {
        int vbk = (k == c->vb) || (mat->Shl_num[k] >= 63);
        if (vbk || !opt->include_photons) c->m.Radiat[k] = 1.0e23;
        if (vbk) c->m.Auger[k] = 1.0e23;
    }
    double sm = 0.0, sp = 0.0, sz = 0.0;
    for (int a = 0; a < mat->n_atoms; ++a) {
        sm = sm + mat->Mass[a] * mat->Pers[a];
        sp = sp + mat->Pers[a];
        sz = sz + (double)mat->Zat[a] * mat->Pers[a];
    }
    c->Mtarget = (OR_MP * sm) / sp;
    c->sumPers = sp;
    c->Zt = sz / sp;
    c->e_emax = (opt->e_grid_Emax > 0.0) ? opt->e_grid_Emax : ((175.6e6 * 2.0) / 1836.0);
    c->shi_Zeff0 = or_shi_zeff(c, ion->E);
    if (mat->work_function > 0.0 &&
        or_barrier_parameters(mat->work_function, mat->bar_length, mat->bar_height, &c->em_E1, &c->em_gamma) == 0) {
        c->em_W = mat->work_function;
        c->em_H = mat->bar_height;
    }
    or_build_time_grid(c);
    return c;
}

static void free_table(or_table *t)
{
    free(t->E);
    for (int s = 0; s < TRK_MAX_SHELLS; ++s) {
        free(t->L[s]);
        free(t->dEdx[s]);
        if (t->rows[s]) {
            for (int i = 0; i < t->n; ++i) { free(t->rows[s][i].hw); free(t->rows[s][i].lam); }
            free(t->rows[s]);
        }
        free(t->dE[s]);
        free(t->dL[s]);
    }
    memset(t, 0, sizeof(*t));
}

void or_destroy(or_ctx *c)
{
    if (!c) return;
    for (int k = 0; k < OR_NTAB; ++k) free_table(&c->tab[k]);
    free(c->dosE); free(c->dos); free(c->dosI); free(c->dosk); free(c->dosM); free(c->doskI); free(c->dosMI);
    free(c->El_IMFP); free(c->El_EMFP); free(c->Hole_IMFP); free(c->Hole_EMFP); free(c->Phot_IMFP); free(c->SHI_path);
    free(c->tim_glob);
    or_dsf_free(c);
    free(c);
}

const char *or_error(const or_ctx *c) { return c->err; }

/* Test-only shortcut: evaluate the ion table only at grid nodes bracketing
 * [lo, hi] (the energies an ion can visit in a short track segment).  The
 * Monte Carlo reads no other node, so its histories are unchanged. */
void or_set_shi_window(or_ctx *c, double lo, double hi)
{
    c->shi_lo = lo;
    c->shi_hi = hi;
}

/* inelastic point (one energy, one shell) for unit tests */
int or_inel_point(const or_ctx *c, int kind, int shell, double Ele, double *L, double *dEdx,
                  double *hw, double *lam, int32_t cap, int32_t *n)
{
    or_row r = { 0, NULL, NULL };
    double S, D;
    int want = (hw != NULL);
    if (kind == TRK_TAB_E_INEL) tot_imfp(c, Ele, shell, 0, &S, &D, want ? &r : NULL);
    else if (kind == TRK_TAB_H_INEL) tot_imfp(c, Ele, shell, 1, &S, &D, want ? &r : NULL);
    else if (kind == TRK_TAB_E_EL) elastic_cs(c, Ele, 0, &S, &D, want ? &r : NULL);
    else if (kind == TRK_TAB_H_EL) elastic_cs(c, Ele, 1, &S, &D, want ? &r : NULL);
    else if (kind == TRK_TAB_PHOT) tot_phot(c, Ele, shell, &S, &D);
    else if (kind == TRK_TAB_SHI) {
        if (c->ion.Kind_ion == 1) shi_tot_imfp_bk(c, Ele, shell, &S, &D);
        else shi_tot_imfp(c, Ele, shell, &S, &D, NULL, NULL, NULL);
        S = (S > 1.0e-10) ? 1.0 / S : 1.0e28;
    } else return TRK_ERR_ARG;
    *L = S;
    *dEdx = D;
    if (want) {
        *n = r.n;
        for (int i = 0; i < r.n && i < cap; ++i) { hw[i] = r.hw[i]; lam[i] = r.lam[i]; }
        free(r.hw);
        free(r.lam);
    }
    return 0;
}

/* Fuse the per-shell tables into totals: How_many_electrons, MC:1902-2057 */
static void fuse(or_ctx *c)
{
    or_table *te = &c->tab[TRK_TAB_E_INEL], *th = &c->tab[TRK_TAB_H_INEL];
    or_table *tee = &c->tab[TRK_TAB_E_EL], *the = &c->tab[TRK_TAB_H_EL];
    or_table *tp = &c->tab[TRK_TAB_PHOT], *ts = &c->tab[TRK_TAB_SHI];
    free(c->El_IMFP); free(c->El_EMFP); free(c->Hole_IMFP); free(c->Hole_EMFP); free(c->Phot_IMFP); free(c->SHI_path);
    c->El_IMFP = (double *)calloc((size_t)te->n, sizeof(double));
    for (int i = 0; i < te->n; ++i) {
        double s = 0.0;
        for (int k = 0; k < c->nsh; ++k) if (te->L[k][i] > 1.0e-10) s = s + 1.0 / te->L[k][i];
        c->El_IMFP[i] = (s < 1.0e-10) ? 1.0e30 : 1.0 / s;
    }
    c->Hole_IMFP = (double *)calloc((size_t)th->n, sizeof(double));
    for (int i = 0; i < th->n; ++i) {
        double s = 0.0;
        for (int k = 0; k < c->nsh; ++k) s = s + 1.0 / th->L[k][i];
        c->Hole_IMFP[i] = (s < 1.0e-10) ? 1.0e30 : 1.0 / s;
    }
    c->El_EMFP = dupd(tee->L[0], tee->n);
    c->Hole_EMFP = dupd(the->L[0], the->n);
    if (tp->n > 0) {
        c->Phot_IMFP = (double *)calloc((size_t)tp->n, sizeof(double));
        for (int i = 0; i < tp->n; ++i) {
            double s = 0.0;
            for (int k = 0; k < c->nsh; ++k) s = s + 1.0 / tp->L[k][i];
            c->Phot_IMFP[i] = (s < 1.0e-10) ? 1.0e30 : 1.0 / s;
        }
    }
    c->SHI_path = (double *)calloc((size_t)ts->n, sizeof(double));
    for (int i = 0; i < ts->n; ++i) {
        double s = 0.0;
        for (int k = 0; k < c->nsh; ++k) s = s + 1.0 / ts->L[k][i];
        c->SHI_path[i] = (s < 1.0e-10) ? 1.0e30 : 1.0 / s;
    }
}

int or_build_tables(or_ctx *c, int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (int k = 0; k < OR_NTAB; ++k) free_table(&c->tab[k]);
    const int nsh = c->nsh;
    /* electron inelastic: AI:158, AI:1976-2143 */
    or_table *te = &c->tab[TRK_TAB_E_INEL];
    te->E = grid_alloc(c, 0.1, c->e_emax, &te->n);
    te->nsh = nsh;
    /* hole inelastic: AI:147-158 with Emax = max(DOS E) + 1 */
    double dosmax = c->dosE[0];
    for (int i = 1; i < c->m.n_dos; ++i) if (c->dosE[i] > dosmax) dosmax = c->dosE[i];
    or_table *th = &c->tab[TRK_TAB_H_INEL];
    th->E = grid_alloc(c, 0.1, dosmax + 1.0, &th->n);
    th->nsh = nsh;
    /* elastic: AI:210 */
    or_table *tee = &c->tab[TRK_TAB_E_EL];
    or_table *the = &c->tab[TRK_TAB_H_EL];
    tee->nsh = 1;
    the->nsh = 1;
    if (c->o.kind_of_EMFP == 2) {
        /* DSF: Elastic_MFP%Total on the DSF energies, AI:905-911, 1181-1184 */
        if (c->dsf_nE[0] < 2 || c->dsf_nE[1] < 2) {
            snprintf(c->err, sizeof c->err, "DSF elastic scattering needs the electron and hole DSF tables");
            return 1;
        }
        or_dsf_elastic_table(c, 0, &tee->E, &tee->L[0], &tee->dEdx[0], &tee->n);
        or_dsf_elastic_table(c, 1, &the->E, &the->L[0], &the->dEdx[0], &the->n);
    } else {
        tee->E = grid_alloc(c, 0.01, c->e_emax, &tee->n);
        the->E = grid_alloc(c, 0.01, dosmax + 1.0, &the->n);
    }
    /* photons */
    or_table *tp = &c->tab[TRK_TAB_PHOT];
    if (c->o.include_photons) {
        tp->E = grid_alloc(c, 0.1, c->e_emax, &tp->n);
        tp->nsh = nsh;
    }
    /* SHI: AI:2270-2273 */
    or_table *ts = &c->tab[TRK_TAB_SHI];
    {
        double M = c->ion.Mass * OR_MP;
        double Emin = ceil(((((M + OR_ME) * (M + OR_ME)) / (M * OR_ME)) * c->Egap) / 4.0);
        double Emax = (c->o.shi_grid_Emax > 0.0) ? c->o.shi_grid_Emax : (175.6e6 / 2.0) * c->ion.Mass;
        ts->E = grid_alloc(c, Emin, Emax, &ts->n);
        ts->nsh = nsh;
    }
    for (int s = 0; s < nsh; ++s) {
        te->L[s] = (double *)calloc((size_t)te->n, sizeof(double));
        te->dEdx[s] = (double *)calloc((size_t)te->n, sizeof(double));
        te->rows[s] = (or_row *)calloc((size_t)te->n, sizeof(or_row));
        th->L[s] = (double *)calloc((size_t)th->n, sizeof(double));
        th->dEdx[s] = (double *)calloc((size_t)th->n, sizeof(double));
        ts->L[s] = (double *)calloc((size_t)ts->n, sizeof(double));
        ts->dEdx[s] = (double *)calloc((size_t)ts->n, sizeof(double));
        if (tp->n) {
            tp->L[s] = (double *)calloc((size_t)tp->n, sizeof(double));
            tp->dEdx[s] = (double *)calloc((size_t)tp->n, sizeof(double));
        }
    }
    th->rows[c->vb] = (or_row *)calloc((size_t)th->n, sizeof(or_row));
    if (c->o.kind_of_EMFP != 2) {
        tee->L[0] = (double *)calloc((size_t)tee->n, sizeof(double));
        tee->dEdx[0] = (double *)calloc((size_t)tee->n, sizeof(double));
        the->L[0] = (double *)calloc((size_t)the->n, sizeof(double));
        the->dEdx[0] = (double *)calloc((size_t)the->n, sizeof(double));
    }
    if (c->o.kind_of_EMFP == 1) {
        tee->rows[0] = (or_row *)calloc((size_t)tee->n, sizeof(or_row));
        the->rows[0] = (or_row *)calloc((size_t)the->n, sizeof(or_row));
    }

    int64_t work = (int64_t)te->n * nsh;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t w = 0; w < work; ++w) {
        int i = (int)(w / nsh), s = (int)(w % nsh);
        tot_imfp(c, te->E[i], s, 0, &te->L[s][i], &te->dEdx[s][i], &te->rows[s][i]);
    }
    work = (int64_t)th->n * nsh;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t w = 0; w < work; ++w) {
        int i = (int)(w / nsh), s = (int)(w % nsh);
        tot_imfp(c, th->E[i], s, 1, &th->L[s][i], &th->dEdx[s][i], (s == c->vb) ? &th->rows[s][i] : NULL);
    }
    const int n_el = (c->o.kind_of_EMFP == 2) ? 0 : tee->n, n_hel = (c->o.kind_of_EMFP == 2) ? 0 : the->n;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n_el; ++i)
        elastic_cs(c, tee->E[i], 0, &tee->L[0][i], &tee->dEdx[0][i], tee->rows[0] ? &tee->rows[0][i] : NULL);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n_hel; ++i)
        elastic_cs(c, the->E[i], 1, &the->L[0][i], &the->dEdx[0][i], the->rows[0] ? &the->rows[0][i] : NULL);
    if (tp->n) {
        for (int s = 0; s < nsh; ++s)
            for (int i = 0; i < tp->n; ++i) tot_phot(c, tp->E[i], s, &tp->L[s][i], &tp->dEdx[s][i]);
    }
    work = (int64_t)ts->n * nsh;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t w = 0; w < work; ++w) {
        int i = (int)(w / nsh), s = (int)(w % nsh);
        double S, D;
        if (c->shi_hi > 0.0) {
            /* window mode: nodes the ion can visit (two brackets around
             * [shi_lo, shi_hi]); the rest are filled with NaN */
            int inside = (ts->E[i] >= c->shi_lo && ts->E[i] <= c->shi_hi);
            if (!inside && i + 2 < ts->n && ts->E[i + 2] >= c->shi_lo && ts->E[i] <= c->shi_hi) inside = 1;
            if (!inside && i >= 2 && ts->E[i - 2] <= c->shi_hi && ts->E[i] >= c->shi_lo) inside = 1;
            if (!inside) { ts->L[s][i] = TRK_NAN; ts->dEdx[s][i] = TRK_NAN; continue; }
        }
        if (c->ion.Kind_ion == 1) shi_tot_imfp_bk(c, ts->E[i], s, &S, &D);
        else shi_tot_imfp(c, ts->E[i], s, &S, &D, NULL, NULL, NULL);
        ts->L[s][i] = (S > 1.0e-10) ? 1.0 / S : 1.0e28;
        ts->dEdx[s][i] = D;
    }
    /* diff_SHI_MFP at the initial ion energy, MAIN:231-238 */
    or_table *td = &c->tab[TRK_TAB_SHI_DIFF];
    td->nsh = nsh;
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < nsh; ++s) {
        double S, D;
        shi_tot_imfp(c, c->ion.E, s, &S, &D, &td->dn[s], &td->dE[s], &td->dL[s]);
    }
    fuse(c);
    c->tables_ready = 1;
    return 0;
}

int or_finish_tables(or_ctx *c)
{
    fuse(c);
    c->tables_ready = 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* table access                                                              */

int or_table_size(const or_ctx *c, int kind, int shell, int32_t *n)
{
    if (kind < 0 || kind >= OR_NTAB) return TRK_ERR_ARG;
    const or_table *t = &c->tab[kind];
    if (kind == TRK_TAB_SHI_DIFF) { *n = t->dn[shell]; return 0; }
    *n = t->n;
    return 0;
}

int or_table_get(const or_ctx *c, int kind, int shell, double *E, double *L, double *dEdx, int32_t n)
{
    const or_table *t = &c->tab[kind];
    if (kind == TRK_TAB_SHI_DIFF) {
        if (n != t->dn[shell]) return TRK_ERR_ARG;
        if (E) memcpy(E, t->dE[shell], sizeof(double) * (size_t)n);
        if (L) memcpy(L, t->dL[shell], sizeof(double) * (size_t)n);
        return 0;
    }
    if (n != t->n) return TRK_ERR_ARG;
    if (E) memcpy(E, t->E, sizeof(double) * (size_t)n);
    if (L && t->L[shell]) memcpy(L, t->L[shell], sizeof(double) * (size_t)n);
    if (dEdx && t->dEdx[shell]) memcpy(dEdx, t->dEdx[shell], sizeof(double) * (size_t)n);
    return 0;
}

int or_row_size(const or_ctx *c, int kind, int shell, int32_t i_E, int32_t *n)
{
    const or_table *t = &c->tab[kind];
    if (!t->rows[shell] || i_E < 0 || i_E >= t->n) return TRK_ERR_ARG;
    *n = t->rows[shell][i_E].n;
    return 0;
}

int or_row_get(const or_ctx *c, int kind, int shell, int32_t i_E, double *hw, double *lam, int32_t n)
{
    const or_table *t = &c->tab[kind];
    if (!t->rows[shell] || i_E < 0 || i_E >= t->n) return TRK_ERR_ARG;
    const or_row *r = &t->rows[shell][i_E];
    if (n != r->n) return TRK_ERR_ARG;
    memcpy(hw, r->hw, sizeof(double) * (size_t)n);
    memcpy(lam, r->lam, sizeof(double) * (size_t)n);
    return 0;
}

int or_table_set(or_ctx *c, int kind, int shell, const double *E, const double *L, const double *dEdx, int32_t n)
{
    or_table *t = &c->tab[kind];
    if (kind == TRK_TAB_SHI_DIFF) {
        free(t->dE[shell]); free(t->dL[shell]);
        t->dn[shell] = n;
        t->dE[shell] = dupd(E, n);
        t->dL[shell] = dupd(L, n);
        if (shell + 1 > t->nsh) t->nsh = shell + 1;
        return 0;
    }
    if (t->E && t->n != n) return TRK_ERR_ARG;
    if (!t->E) { t->E = dupd(E, n); t->n = n; }
    free(t->L[shell]); free(t->dEdx[shell]);
    t->L[shell] = dupd(L, n);
    t->dEdx[shell] = dupd(dEdx, n);
    if (shell + 1 > t->nsh) t->nsh = shell + 1;
    return 0;
}

int or_row_set(or_ctx *c, int kind, int shell, int32_t i_E, const double *hw, const double *lam, int32_t n)
{
    or_table *t = &c->tab[kind];
    if (i_E < 0 || i_E >= t->n) return TRK_ERR_ARG;
    if (!t->rows[shell]) t->rows[shell] = (or_row *)calloc((size_t)t->n, sizeof(or_row));
    or_row *r = &t->rows[shell][i_E];
    free(r->hw); free(r->lam);
    r->n = n;
    r->hw = dupd(hw, n);
    r->lam = dupd(lam, n);
    return 0;
}
