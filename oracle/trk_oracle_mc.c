/*
 * trk_oracle_mc.c -- CPU restatement of the TREKIS-3 Monte Carlo iteration.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP transport kernels and
 * the CPU baseline of the benchmark).
 *
 * Follows Monte_Carlo_modelling (MC:427-679): one iteration is a
 * discrete-event simulation in global time order; after every time-grid node
 * the snapshot statistics of Calculated_statistics (MC:881-1110) are taken.
 * The reference finds the next event with minloc over all particles
 * (Find_min_time_particle, MC:531-563); the same order -- smallest tn, ties
 * broken by particle kind (ion, electrons, holes, photons) and then by array
 * index -- is produced here with a binary heap.
 *
 * Random numbers: every draw comes from the Philox stream of the particle
 * that owns the event (include/trk_philox.h), in the reference's draw order
 * (each random_number call site of the reference is cited inline).  Photon slots are
 * never reused (the reference overwrites live photons, MC:2844-2854 vs
 * MC:2963-2964; here each photon keeps its own slot).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "trk_oracle_internal.h"

#define SQRT trk_sqrt
#define LOG  trk_log
#define EXP  trk_exp
#define SIN  trk_sin
#define COS  trk_cos
#define ACOS trk_acos

enum { K_SHI = 0, K_EL = 1, K_HO = 2, K_PH = 3 };

typedef struct {
    double E, t0, tn, X, Y, Z, L, theta, phi;
    uint64_t id;
    uint32_t nd, ne;
} p_el;                       /* electrons and photons */

typedef struct {
    double E, t0, tn, X, Y, Z, L, theta, phi, Mass, Ehkin;
    int sh;                   /* flat shell index (KOA, Shl) */
    uint64_t id;
    uint32_t nd, ne;
} p_ho;

typedef struct {
    double E, t0, tn, X, Y, Z, L, Zeff;
    uint64_t id;
    uint32_t nd, ne;
} p_shi;

typedef struct { double t; int kind; int idx; } hnode;

typedef struct {
    or_ctx *c;
    uint64_t seed;
    uint32_t iter;
    p_shi shi;
    p_el *el; int nel, capel;
    p_ho *ho; int capho;              /* holes share the electron count (pairs) */
    p_el *ph; int nph, capph, nph_alive;
    hnode *heap; int nh, caph;
    double At_NRG;
    int inode;                        /* current time-node index (1-based) */
    const double *R, *V; int NR;
    trk_tallies *out;
    double *trace; int64_t tcap, tn;
    int64_t n_events[8];
    double *em; int nem, capem;       /* kinetic energies of the emitted electrons (Em_electrons) */
} or_mc;

/* ---------------------------------------------------------------- heap */
static int hless(const hnode *a, const hnode *b)
{
    if (a->t != b->t) return a->t < b->t;
    if (a->kind != b->kind) return a->kind < b->kind;
    return a->idx < b->idx;
}
static void hpush(or_mc *S, double t, int kind, int idx)
{
    if (S->nh == S->caph) {
        S->caph = S->caph ? 2 * S->caph : 1024;
        S->heap = (hnode *)realloc(S->heap, sizeof(hnode) * (size_t)S->caph);
    }
    int i = S->nh++;
    hnode n = { t, kind, idx };
    while (i > 0) {
        int p = (i - 1) / 2;
        if (!hless(&n, &S->heap[p])) break;
        S->heap[i] = S->heap[p];
        i = p;
    }
    S->heap[i] = n;
}
static hnode hpop(or_mc *S)
{
    hnode top = S->heap[0];
    hnode last = S->heap[--S->nh];
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = l + 1, m = i;
        const hnode *mv = &last;
        if (l < S->nh && hless(&S->heap[l], mv)) { m = l; mv = &S->heap[l]; }
        if (r < S->nh && hless(&S->heap[r], mv)) { m = r; mv = &S->heap[r]; }
        if (m == i) break;
        S->heap[i] = S->heap[m];
        i = m;
    }
    if (S->nh > 0) S->heap[i] = last;
    return top;
}
static double cur_tn(or_mc *S, int kind, int idx)
{
    switch (kind) {
    case K_SHI: return S->shi.tn;
    case K_EL: return S->el[idx].tn;
    case K_HO: return S->ho[idx].tn;
    default: return S->ph[idx].tn;
    }
}
/* next valid event: entries whose stored time no longer matches are stale */
static int hpeek(or_mc *S, hnode *out)
{
    while (S->nh > 0) {
        hnode h = S->heap[0];
        if (cur_tn(S, h.kind, h.idx) == h.t) { *out = h; return 1; }
        hpop(S);
    }
    return 0;
}

/* ---------------------------------------------------------------- rng */
static double rn(or_mc *S, uint64_t id, uint32_t *nd)
{
    return trk_uniform(S->seed, id, S->iter, (*nd)++);
}

/* ---------------------------------------------------------------- tables */
/* Next_free_path_2d, MC:1874-1899 */
double or_nfp_2d(or_ctx *c, double E, const double *Earr, const double *Larr, int32_t n)
{
    if (n < 1) return 1.0e30;   /* no table (no photons): an infinite path */
    int N = or_find_mono_2d(Earr, n, E);
    if (N < 0) { c->err_numeric = 1; return 1.0e30; }
    if (N == 1) {
        double MFP = or_interpolate(1, Earr[0], Earr[1], Larr[0], Larr[1], E);
        if (MFP < Larr[0]) MFP = Larr[0];
        return MFP;
    }
    int Nl = N - 1;
    if (Larr[Nl - 1] >= 1.0e16) return Larr[Nl - 1];
    return or_interpolate(5, Earr[Nl - 1], Earr[N - 1], Larr[Nl - 1], Larr[N - 1], E);
}

/* Next_free_path_1d, MC:1835-1871 (the N=1 floor compares with the energy
 * array, MC:1855, as the reference does) */
double or_nfp_1d(or_ctx *c, double E, const double *Earr, const double *Larr, int32_t n)
{
    int N = or_find_mono_1d(Earr, n, E);
    if (N < 0) { c->err_numeric = 1; return 1.0e30; }
    if (N == 1) {
        double MFP = or_interpolate(1, Earr[0], Earr[1], Larr[0], Larr[1], E);
        if (MFP < Earr[0]) MFP = Earr[0];
        return MFP;
    }
    int Nl = N - 1;
    if (Larr[Nl - 1] >= 1.0e16) return Larr[Nl - 1];
    return or_interpolate(5, Earr[Nl - 1], Earr[N - 1], Larr[Nl - 1], Larr[N - 1], E);
}

/* Which_shell, MC:1786-1832 */
static int which_shell(or_mc *S, const or_table *t, double E, uint64_t id, uint32_t *nd)
{
    or_ctx *c = S->c;
    double tmp[TRK_MAX_SHELLS];
    double tot = 0.0;
    for (int s = 0; s < c->nsh; ++s) {
        int N = or_find_mono_1d(t->E, t->n, E);
        double MFP;
        if (N < 0) { c->err_numeric = 1; N = 1; }
        if (N == 1) {
            MFP = 1.0e20;
        } else {
            const double *L = t->L[s];
            if ((L[N - 2] == L[N - 1]) || (L[N - 2] > 1.0e20)) MFP = L[N - 2];
            else MFP = or_interpolate(5, t->E[N - 2], t->E[N - 1], L[N - 2], L[N - 1], E);
        }
        tmp[s] = 1.0 / MFP;
        tot = tot + tmp[s];
    }
    double RN = rn(S, id, nd);
    tot = RN * tot;
    double sum = 0.0;
    int s;
    for (s = 0; s < c->nsh; ++s) {
        sum = sum + tmp[s];
        if (sum >= tot) break;
    }
    if (s == c->nsh) s = c->nsh - 1;
    return s;
}

/* interpolate_transferred_energy, CS:1968-2045 */
double or_interp_transfer(or_ctx *c, const or_table *t, int shell, double Ele, double L_need)
{
    int iE = or_find_mono_1d(t->E, t->n, Ele);
    if (iE < 0) { c->err_numeric = 1; return TRK_NAN; }
    if (iE > 1) {
        if (trk_fabs(t->E[iE - 2] - Ele) < 1.0e-6) iE = iE - 1;
    }
    const or_row *r = &t->rows[shell][iE - 1];
    int ih = or_find_decr(r->lam, r->n, L_need);
    if (ih < 0) { c->err_numeric = 1; return TRK_NAN; }
    double hw1;
    if ((ih == 1) || (ih == r->n)) hw1 = r->hw[ih - 1];
    else hw1 = or_interpolate(5, r->lam[ih - 1], r->lam[ih], r->hw[ih - 1], r->hw[ih], L_need);
    double out = hw1;
    if (iE > 1) {
        iE = iE - 1;
        const or_row *r2 = &t->rows[shell][iE - 1];
        ih = or_find_decr(r2->lam, r2->n, L_need);
        double hw2;
        if ((ih == 1) || (ih == r2->n)) hw2 = r2->hw[ih - 1];
        else hw2 = or_interpolate(5, r2->lam[ih - 1], r2->lam[ih], r2->hw[ih - 1], r2->hw[ih], L_need);
        if ((hw1 < 1.0e-10) || (hw2 < 1.0e-10)) out = or_interpolate(1, t->E[iE - 1], t->E[iE], hw1, hw2, Ele);
        else out = or_interpolate(5, t->E[iE - 1], t->E[iE], hw1, hw2, Ele);
    }
    return out;
}

static double hole_mass_tab(const or_ctx *c, double Ele)
{
    if (c->m.hole_mass >= 0.0) return c->m.hole_mass;
    int j = or_find_mono_1d(c->dosE, c->m.n_dos, Ele);
    return c->dosM[j - 1];
}

/* Electron_energy_transfer_inelastic, CS:1793-1871 (tabulated branch), for a
 * given random number */
static double transfer_inel_rn(or_ctx *c, double Ele, int shell, double L_tot, int hole, double RN)
{
    double L_need = L_tot / RN;
    double Emin = c->m.Ip[shell];
    double Egap = c->Egap;
    double Emax, Mass;
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    if (!hole) {
        Emax = (Ele + Emin) / 2.0;
        Mass = 1.0;
    } else {
        Mass = hole_mass_tab(c, Ele);
        Emax = ((4.0 * Ele) * Mass) / ((Mass + 1.0) * (Mass + 1.0));
    }
    if (c->o.plasmon_Emax) {
        if (Emin == Egap) {
            double Ep = SQRT((((((c->m.N_VB_el * c->m.At_Dens) * 1.0e6) * OR_H) * OR_H) / (OR_ME * OR_E0)) + (Egap * Egap));
            if (Ep >= Emax) Emax = Ep;
            if (Ele < Emax) Emax = Ele;
        }
    }
    double E;
    if (!hole) E = or_interp_transfer(c, &c->tab[TRK_TAB_E_INEL], shell, Ele, L_need);
    else E = or_interp_transfer(c, &c->tab[TRK_TAB_H_INEL], c->vb, Ele, L_need);
    if (E < Emin) E = Emin;
    if (E > Emax) E = Emax;
    if (E != E) E = Emin;
    return E;
}

/* Electron_energy_transfer_inelastic with Electron_NRG_transfer_CDF's
 * Delta-CDF branch (CS:1809, 1894-1896): the first draw only sets L_need;
 * get_inelastic_energy_transfer draws again */
static double transfer_inel_delta(or_mc *S, double Ele, int shell, int hole, uint64_t id, uint32_t *nd)
{
    or_ctx *c = S->c;
    double Emin = c->m.Ip[shell];
    double Egap = c->Egap;
    double Emax, Mass;
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    if (!hole) {
        Emax = (Ele + Emin) / 2.0;
    } else {
        Mass = hole_mass_tab(c, Ele);
        Emax = ((4.0 * Ele) * Mass) / ((Mass + 1.0) * (Mass + 1.0));
    }
    if (c->o.plasmon_Emax && Emin == Egap) {
        double Ep = SQRT((((((c->m.N_VB_el * c->m.At_Dens) * 1.0e6) * OR_H) * OR_H) / (OR_ME * OR_E0)) + (Egap * Egap));
        if (Ep >= Emax) Emax = Ep;
        if (Ele < Emax) Emax = Ele;
    }
    double E = or_delta_transfer(c, shell, Ele, Emin, rn(S, id, nd));
    if (E < Emin) E = Emin;
    if (E > Emax) E = Emax;
    if (E != E) E = Emin;
    return E;
}

/* Electron_energy_transfer_inelastic on a BEB shell (KOCS = 2, CS:1841-1848):
 * Electron_NRG_transfer_BEB's search for L_need */
static double transfer_inel_beb(or_ctx *c, double Ele, int shell, double L_tot, int hole, double RN)
{
    double L_need = L_tot / RN;
    double Emin = c->m.Ip[shell];
    double Egap = c->Egap;
    double Emax, Mass;
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    if (!hole) {
        Emax = (Ele + Emin) / 2.0;
        Mass = 1.0;
    } else {
        Mass = hole_mass_tab(c, Ele);
        Emax = ((4.0 * Ele) * Mass) / ((Mass + 1.0) * (Mass + 1.0));
    }
    if (c->o.plasmon_Emax && Emin == Egap) {
        double Ep = SQRT((((((c->m.N_VB_el * c->m.At_Dens) * 1.0e6) * OR_H) * OR_H) / (OR_ME * OR_E0)) + (Egap * Egap));
        if (Ep >= Emax) Emax = Ep;
        if (Ele < Emax) Emax = Ele;
    }
    double E = or_beb_transfer(c, shell, Ele, L_need, Mass, Emin);
    if (E < Emin) E = Emin;
    if (E > Emax) E = Emax;
    if (E != E) E = Emin;
    return E;
}

static double e_transfer_inel(or_mc *S, double Ele, int shell, double L_tot, int hole, uint64_t id, uint32_t *nd)
{
    double RN = rn(S, id, nd);
    if (S->c->m.kocs[shell] == 2) return transfer_inel_beb(S->c, Ele, shell, L_tot, hole, RN);
    if (S->c->o.kind_of_DR == 4) return transfer_inel_delta(S, Ele, shell, hole, id, nd);
    return transfer_inel_rn(S->c, Ele, shell, L_tot, hole, RN);
}

/* Electron_energy_transfer_elastic, CS:2285-2428 (tabulated branch) */
static double e_transfer_el(or_mc *S, double Ele, double L_tot, int hole, uint64_t id, uint32_t *nd)
{
    or_ctx *c = S->c;
    double RN = rn(S, id, nd);
    double L_need = L_tot / RN;
    double hw = or_interp_transfer(c, &c->tab[hole ? TRK_TAB_H_EL : TRK_TAB_E_EL], 0, Ele, L_need);
    if (hw >= Ele) hw = Ele;
    return hw;
}

double or_rest_energy(double M0) { return ((M0 * OR_CVEL) * OR_CVEL) / OR_E; }

/* Mott_sample_mu + screening_parameter + transfered_E_from_theta,
 * CS:3517-3613 (NRG_transfer_elastic_atomic) */
static double mott_transfer(or_mc *S, int atom, double Ee, double M_eff, uint64_t id, uint32_t *nd)
{
    or_ctx *c = S->c;
    double Z = (double)c->m.Zat[atom];
    double Mat = c->m.Mass[atom] * OR_MP;
    double RN = rn(S, id, nd);
    /* Mott_sample_mu with the electron mass (the mass argument is not passed) */
    double me = OR_ME;
    double Erest = or_rest_energy(me);
    double fact = Ee / Erest + 1.0;
    double v = OR_CVEL * SQRT(1.0 - 1.0 / (fact * fact));
    double theta;
    if (v < 1.0e-6) {
        theta = 0.0;
    } else {
        double beta = v / OR_CVEL;
        double beta2 = beta * beta;
        double tau = Ee / Erest;
        double nu = (((1.7e-5 * trk_pow(Z, 2.0 / 3.0)) * (1.0 - beta2)) / beta2) *
                    (1.13 + (((((3.76 * OR_ALPHA) * OR_ALPHA) / beta2) * Z) * Z) * SQRT(tau / (1.0 + tau)));
        double mu = (RN * ((2.0 * nu) + 1.0) - nu) / (RN + nu);
        theta = ACOS(mu);
    }
    double mc2 = or_rest_energy(OR_ME * M_eff);
    double Mct2 = or_rest_energy(Mat);
    double ct = COS(theta);
    double ct2 = ct * ct;
    double st2 = 1.0 - ct2;
    double Emc = Ee + mc2;
    double E2mc = Ee + 2.0 * mc2;
    double EmcMc = Emc + Mct2;
    double W1 = ((Emc * st2) + Mct2) - (ct * SQRT((Mct2 * Mct2) - ((mc2 * mc2) * st2)));
    double W2 = (Ee * E2mc) / ((EmcMc * EmcMc) - ((Ee * E2mc) * ct2));
    return W1 * W2;
}

/* cos_theta_from_W, MC:1275-1299 */
static double cos_theta_W(or_mc *S, double E, double W, double M_in, double mt, uint64_t id, uint32_t *nd)
{
    double Ein = or_rest_energy(M_in);
    double Et = or_rest_energy(mt);
    double E2mc = E + 2.0 * Ein;
    double EmW = E - W;
    double W1 = (E * E2mc) - (W * ((E + Ein) + Et));
    double W2 = ((E * E2mc) * EmW) * (E2mc - W);
    double mu;
    if (W2 > 0.0) mu = W1 / SQRT(W2);
    else mu = 0.0;
    if (trk_fabs(mu) > 1.0) {
        double RN = rn(S, id, nd);
        double th = OR_PI * RN;
        mu = COS(th);
    }
    return mu;
}

/* Test hooks for the elastic kinematics and the decay-shell choice. */
double or_mott_theta(const or_ctx *c, int atom, double Ee, double RN, double *nu_out)
{
    double Z = (double)c->m.Zat[atom];
    double Erest = or_rest_energy(OR_ME);
    double fact = Ee / Erest + 1.0;
    double v = OR_CVEL * SQRT(1.0 - 1.0 / (fact * fact));
    double beta = v / OR_CVEL, beta2 = beta * beta, tau = Ee / Erest;
    double nu = (((1.7e-5 * trk_pow(Z, 2.0 / 3.0)) * (1.0 - beta2)) / beta2) *
                (1.13 + (((((3.76 * OR_ALPHA) * OR_ALPHA) / beta2) * Z) * Z) * SQRT(tau / (1.0 + tau)));
    if (nu_out) *nu_out = nu;
    return ACOS((RN * ((2.0 * nu) + 1.0) - nu) / (RN + nu));
}

double or_mott_W(const or_ctx *c, int atom, double Ee, double theta, double M_eff)
{
    double Mct2 = or_rest_energy(c->m.Mass[atom] * OR_MP);
    double mc2 = or_rest_energy(OR_ME * M_eff);
    double ct = COS(theta), ct2 = ct * ct, st2 = 1.0 - ct2;
    double Emc = Ee + mc2, E2mc = Ee + 2.0 * mc2, EmcMc = Emc + Mct2;
    double W1 = ((Emc * st2) + Mct2) - (ct * SQRT((Mct2 * Mct2) - ((mc2 * mc2) * st2)));
    double W2 = (Ee * E2mc) / ((EmcMc * EmcMc) - ((Ee * E2mc) * ct2));
    return W1 * W2;
}

double or_cos_theta_W(const or_ctx *c, double E, double W, double M_in)
{
    double Ein = or_rest_energy(M_in), Et = or_rest_energy(c->Mtarget);
    double E2mc = E + 2.0 * Ein, EmW = E - W;
    double W1 = (E * E2mc) - (W * ((E + Ein) + Et));
    double W2 = ((E * E2mc) * EmW) * (E2mc - W);
    return (W2 > 0.0) ? W1 / SQRT(W2) : 0.0;
}

/* New_Angles_both, MC:1328-1360 */
static void new_angles(double phi0, double theta0, double theta, double psi, double *phi1, double *theta1)
{
    double p = phi0 + ((theta * COS(theta0)) * SIN(psi));
    double t = theta0 + (theta * COS(psi));
    while (t < 0.0) { t = trk_fabs(t); p = p + OR_PI; }
    while (t > OR_PI) { t = (2.0 * OR_PI) - t; p = p - OR_PI; }
    if (p > 2.0 * OR_PI) p = p - (((double)floor(p / (2.0 * OR_PI))) * 2.0) * OR_PI;
    if (p < 0.0) p = p + (((double)ceil(trk_fabs(p) / (2.0 * OR_PI))) * 2.0) * OR_PI;
    *phi1 = p;
    *theta1 = t;
}

void or_new_angles(double phi0, double theta0, double theta, double psi, double *phi1, double *theta1)
{
    new_angles(phi0, theta0, theta, psi, phi1, theta1);
}

/* Electron_recieves_E, MC:1653-1716 */
static double receives_E(or_mc *S, double dE, int shell, uint64_t id, uint32_t *nd)
{
    or_ctx *c = S->c;
    double E = dE - c->m.Ip[shell];
    if (shell == c->vb) {
        if (c->m.n_dos <= 0 || dE <= c->m.Ip[shell]) return E;
        int N = c->m.n_dos;
        int M;
        if (E < c->dosE[N - 1]) M = or_find_mono_1d(c->dosE, N, E);
        else M = N + 1;
        if (M > 1) {
            double Sum = c->dosI[M - 2];
            double RN = rn(S, id, nd);
            double Tot = RN * Sum;
            int Nt = or_find_mono_1d(c->dosI, N, Tot);
            double Ed;
            if (Nt > 1)
                Ed = c->dosE[Nt - 2] + ((c->dosE[Nt - 1] - c->dosE[Nt - 2]) * (Tot - c->dosI[Nt - 2])) /
                                       (c->dosI[Nt - 1] - c->dosI[Nt - 2]);
            else
                Ed = c->dosE[Nt - 1];
            return E - Ed;
        }
        return E;
    }
    return E;
}

/* From_where_in_VB, MC:1522-1573 */
static double from_where_in_vb(or_mc *S, int has_E, double Elim, uint64_t id, uint32_t *nd)
{
    or_ctx *c = S->c;
    if (c->m.n_dos <= 0) return 0.0;
    int N = c->m.n_dos;
    if (!has_E) {
        double Sum = c->dosI[N - 1];
        double RN = rn(S, id, nd);
        double Tot = RN * Sum;
        int Nt = or_find_mono_1d(c->dosI, N, Tot);
        if (Nt > 1)
            return c->dosE[Nt - 2] + ((c->dosE[Nt - 1] - c->dosE[Nt - 2]) * (Tot - c->dosI[Nt - 2])) /
                                     (c->dosI[Nt - 1] - c->dosI[Nt - 2]);
        return c->dosE[Nt - 1];
    }
    int M;
    if (Elim < c->dosE[N - 1]) M = or_find_mono_1d(c->dosE, N, Elim);
    else M = N + 1;
    if (M > 1) {
        double Sum = c->dosI[M - 2];
        double RN = rn(S, id, nd);
        double Tot = RN * Sum;
        int Nt = or_find_mono_1d(c->dosI, N, Tot);
        if (Nt > 1)
            return c->dosE[Nt - 2] + ((c->dosE[Nt - 1] - c->dosE[Nt - 2]) * (Tot - c->dosI[Nt - 2])) /
                                     (c->dosI[Nt - 1] - c->dosI[Nt - 2]);
        return c->dosE[Nt - 1];
    }
    return 0.0;
}

/* count_for_Auger_shells, MC:1576-1608 */
static double auger_count(const or_ctx *c, double NRG, int second)
{
    double coun = 0.0;
    for (int s = 0; s < c->nsh; ++s) {
        double Ed = second ? 1.0e10 : (NRG - c->m.Ip[s]);
        if ((NRG > c->m.Ip[s] + 1.0e-3) && (Ed >= c->Egap)) coun = coun + c->m.Nel[s];
    }
    return coun;
}

/* Choose_for_Auger_shell, MC:1610-1647; returns -1 if nothing qualifies */
static int auger_choose(const or_ctx *c, double NRG, double Shel, int second, int dflt)
{
    double cs = 0.0;
    int found = dflt;
    for (int s = 0; s < c->nsh; ++s) {
        double Ed = second ? 1.0e10 : (NRG - c->m.Ip[s]);
        if ((NRG > c->m.Ip[s] + 1.0e-3) && (Ed >= c->Egap)) {
            cs = cs + c->m.Nel[s];
            if (cs >= Shel) found = s;
        }
        if (cs >= Shel) break;
    }
    return found;
}

/* Test hook: the shell the cumulative choice picks for a random fraction. */
int or_auger_pick(const or_ctx *c, double NRG, double frac, int second)
{
    double coun = auger_count(c, NRG, second);
    if (!(coun > 0.0)) return -1;
    return auger_choose(c, NRG, frac * coun, second, -1);
}


/* ---------------------------------------------------------------- kinematics */
static double vel_el(double E) { return SQRT(((2.0 * E) * OR_E) / OR_ME); }

static double vel_hole(const p_ho *h)
{
    if (h->Mass < 1.0e6) {
        if ((h->Ehkin < -1.0e-6) || (h->Mass < 1.0e-10)) return 0.0;
        if (trk_fabs(h->Ehkin) < 1.0e-6) return 0.0;
        return SQRT(((2.0 * h->Ehkin) * OR_E) / (h->Mass * OR_ME));
    }
    return 0.0;
}

static double t_next(double t0, double MFP, double V)
{
    if (V > 1.0e-10) return t0 + (MFP / V) * 1.0e5;
    return 1.0e25;
}

static void cut_el(const or_ctx *c, p_el *e) { if (e->E < c->m.cut_off) e->tn = 1.0e20; }
static void cut_ho(const or_ctx *c, p_ho *h) { if ((h->Mass < 1.0e15) && (h->Ehkin < c->m.cut_off)) h->tn = 1.0e20; }

/* Hole_parameters + Assign_holes_mass + Get_time_of_next_event (holes),
 * MC:724-848.  `S_owner` draws come from the stream of the event owner. */
static void hole_params(or_mc *S, p_ho *h, double Eh, uint64_t id, uint32_t *nd)
{
    or_ctx *c = S->c;
    if (h->sh == c->vb) {
        double Etemp = h->Ehkin;
        h->Ehkin = Eh - c->Egap;
        h->E = c->Egap;
        if (c->m.hole_mass > 0.0) h->Mass = c->m.hole_mass;
        else {
            int Mn = or_find_mono_1d(c->dosE, c->m.n_dos, h->Ehkin);
            h->Mass = c->dosM[Mn - 1];
        }
        if (h->Mass < 1.0e3) {
            double HIMFP = or_nfp_2d(c, h->Ehkin, c->tab[TRK_TAB_H_INEL].E, c->Hole_IMFP, c->tab[TRK_TAB_H_INEL].n);
            double HEMFP;
            if (Etemp == (Eh - c->Egap)) HEMFP = 1.0e30;
            else HEMFP = or_nfp_2d(c, h->Ehkin, c->tab[TRK_TAB_H_EL].E, c->Hole_EMFP, c->tab[TRK_TAB_H_EL].n);
            double RN = rn(S, id, nd);
            double MFP = -LOG(RN) / (1.0 / HIMFP + 1.0 / HEMFP);
            h->tn = t_next(h->t0, MFP, vel_hole(h));
            h->L = MFP;
        } else {
            h->L = 1.0e30;
            h->tn = 1.0e30;
        }
    } else {
        h->Mass = 1.0e29;
        double RN = rn(S, id, nd);
        double nu = 1.0 / c->m.Auger[h->sh] + 1.0 / c->m.Radiat[h->sh];
        h->tn = h->t0 - LOG(RN) / nu;
        h->L = 1.0e30;
        h->E = Eh;
        h->Ehkin = 0.0;
    }
}

/* ---------------------------------------------------------------- storage */
static int new_pair(or_mc *S, uint64_t parent, uint32_t ev, uint32_t slot_e, uint32_t slot_h)
{
    if (S->nel == S->capel) {
        S->capel = S->capel ? 2 * S->capel : 4096;
        S->el = (p_el *)realloc(S->el, sizeof(p_el) * (size_t)S->capel);
        S->ho = (p_ho *)realloc(S->ho, sizeof(p_ho) * (size_t)S->capel);
    }
    int k = S->nel++;
    p_el *e = &S->el[k];
    p_ho *h = &S->ho[k];
    memset(e, 0, sizeof(*e));
    memset(h, 0, sizeof(*h));
    /* initial slot values, MC:1948-1951 / OBJ:471-474 */
    e->tn = 1.0e20;
    h->tn = 1.0e21; h->L = 1.0e30; h->Mass = 1.0e30; h->sh = -1;
    e->id = trk_child_id(parent, ev, slot_e);
    h->id = trk_child_id(parent, ev, slot_h);
    return k;
}

static int new_photon(or_mc *S, uint64_t parent, uint32_t ev, uint32_t slot)
{
    if (S->nph == S->capph) {
        S->capph = S->capph ? 2 * S->capph : 256;
        S->ph = (p_el *)realloc(S->ph, sizeof(p_el) * (size_t)S->capph);
    }
    int k = S->nph++;
    p_el *p = &S->ph[k];
    memset(p, 0, sizeof(*p));
    p->tn = 1.0e20;
    p->id = trk_child_id(parent, ev, slot);
    S->nph_alive++;
    return k;
}

static void trace_ev(or_mc *S, int kind, uint64_t id, uint32_t ev, double t, double E, double E2,
                     double X, double Y, double Z, double tn, uint32_t nd)
{
    if (kind >= 0 && kind < 8) S->n_events[kind]++;
    if (!S->trace) return;
    if (S->tn >= S->tcap) return;
    double *r = S->trace + 12 * S->tn++;
    double idd;
    memcpy(&idd, &id, sizeof(double));
    r[0] = (double)S->iter; r[1] = idd; r[2] = (double)ev; r[3] = (double)kind;
    r[4] = t; r[5] = E; r[6] = E2; r[7] = X; r[8] = Y; r[9] = Z; r[10] = tn; r[11] = (double)nd;
}

/* Out_Elat event tally, MC:2435-2436 / 2718-2719 */
static void tally_elat(or_mc *S, double X, double Y, double dE)
{
    S->At_NRG = S->At_NRG + dE;
    double R = SQRT(X * X + Y * Y);
    int j = or_find_mono_1d(S->R, S->NR, R);
    if (j < 0) { S->c->err_numeric = 1; return; }
    if (S->out->Out_Elat) {
        int Nt = S->out->Nt;
        S->out->Out_Elat[(S->inode - 1) + (int64_t)Nt * (j - 1)] += dE * S->V[j - 1];
    }
}

/* ---------------------------------------------------------------- SHI */
/* SHI_energy_transfer, MC:1719-1780 (CDF branch) */
static double shi_transfer_rn(const or_ctx *c, int shell, double RN)
{
    const or_table *t = &c->tab[TRK_TAB_SHI_DIFF];
    int N = t->dn[shell];
    const double *E = t->dE[shell], *L = t->dL[shell];
    double Ecur = c->m.Ip[shell];
    int M = or_find_mono_1d(E, N, c->m.Ip[shell]);
    double dL;
    if (M > 1) {
        if (L[M - 2] > 1.0e-10) dL = or_interpolate(5, E[M - 2], E[M - 1], L[M - 2], L[M - 1], Ecur);
        else dL = or_interpolate(1, E[M - 2], E[M - 1], L[M - 2], L[M - 1], Ecur);
    } else {
        dL = L[0];
    }
    double Tot;
    if ((dL > 0.0) && (L[N - 1] > 0.0)) Tot = 1.0 / dL + RN * (1.0 / L[N - 1] - 1.0 / dL);
    else Tot = 1.5e21;
    int Nt;
    if (Tot < 1.0e20) {
        /* Find_in_1D_array on 1/L, RF:3348-3358 */
        int i = 1;
        while (i < N && 1.0 / L[i - 1] < Tot) i++;
        Nt = i;
    } else {
        Nt = M;
    }
    if (Nt > M)
        return or_interpolate(5, 1.0 / L[Nt - 2], 1.0 / L[Nt - 1], E[Nt - 2], E[Nt - 1], Tot);
    return c->m.Ip[shell];
}

static double shi_transfer(or_mc *S, int shell, uint64_t id, uint32_t *nd)
{
    double RN = rn(S, id, nd);
    return shi_transfer_rn(S->c, shell, RN);
}

/* Test hooks: the inverse-CDF samplers for a given random number.  kind 0
 * electron inelastic (shell's own mean free path as the reference passes it,
 * MC:2329-2331), 1 VB-hole inelastic, 2 ion (SHI_energy_transfer). */
double or_sample_transfer(or_ctx *c, int kind, int shell, double Ele, double RN)
{
    if (kind == 2) return shi_transfer_rn(c, shell, RN);
    const or_table *t = &c->tab[kind == 0 ? TRK_TAB_E_INEL : TRK_TAB_H_INEL];
    int s = (kind == 0) ? shell : c->vb;
    double L = or_nfp_1d(c, Ele, t->E, t->L[s], t->n);
    return transfer_inel_rn(c, Ele, s, L, kind == 1, RN);
}

/* SHI_Monte_Carlo, MC:2153-2249 */
static void shi_event(or_mc *S)
{
    or_ctx *c = S->c;
    p_shi *I = &S->shi;
    uint32_t ev = I->ne++;
    double t_ev = I->tn;
    int sh = which_shell(S, &c->tab[TRK_TAB_SHI], I->E, I->id, &I->nd);
    double dE = shi_transfer(S, sh, I->id, &I->nd);
    double SHI_IMFP = or_nfp_2d(c, I->E, c->tab[TRK_TAB_SHI].E, c->SHI_path, c->tab[TRK_TAB_SHI].n);
    double RN = rn(S, I->id, &I->nd);
    SHI_IMFP = -SHI_IMFP * LOG(RN);
    double Z = I->Z + I->L;
    I->E = I->E - dE;
    I->t0 = I->tn;
    I->Z = Z;
    I->L = SHI_IMFP;
    {
        double V = SQRT(((2.0 * I->E) * OR_E) / (c->ion.Mass * OR_MP));
        I->tn = t_next(I->t0, SHI_IMFP, V);
    }
    I->Zeff = or_shi_zeff(c, I->E);

    int k = new_pair(S, I->id, ev, 0, 1);
    p_el *e = &S->el[k];
    p_ho *h = &S->ho[k];
    double dE_cur = receives_E(S, dE, sh, I->id, &I->nd);
    /* Update_electron_angles_SHI, MC:1170-1187 */
    double MSHI = c->ion.Mass * OR_MP;
    double theta;
    if (I->E <= 0.0) theta = OR_PI / 2.0;
    else theta = ACOS(SQRT(((((MSHI + OR_ME) * (MSHI + OR_ME)) / ((4.0 * MSHI) * OR_ME)) * dE) / I->E));
    double phi = (2.0 * OR_PI) * rn(S, I->id, &I->nd);
    double IMFP = or_nfp_2d(c, dE_cur, c->tab[TRK_TAB_E_INEL].E, c->El_IMFP, c->tab[TRK_TAB_E_INEL].n);
    double EMFP = or_nfp_2d(c, dE_cur, c->tab[TRK_TAB_E_EL].E, c->El_EMFP, c->tab[TRK_TAB_E_EL].n);
    RN = rn(S, I->id, &I->nd);
    double MFP = -LOG(RN) / (1.0 / IMFP + 1.0 / EMFP);
    /* Impact_parameter, MC:1113-1126 */
    double A = 1.0 + MSHI / OR_ME;
    double L = (((OR_A0 * I->Zeff) * OR_RY) / I->E) * SQRT(((((4.0 * I->E) / dE) * MSHI) / OR_ME) - A * A);
    double X = I->X + L * SIN(phi);
    double Y = I->Y + L * COS(phi);
    e->E = dE_cur; e->t0 = I->t0; e->X = X; e->Y = Y; e->Z = Z; e->L = MFP; e->theta = theta; e->phi = phi;
    e->tn = t_next(e->t0, MFP, vel_el(e->E));
    cut_el(c, e);
    /* hole, Update_holes_angles_SHI MC:1132-1140 */
    double hth = OR_PI * rn(S, I->id, &I->nd);
    double hph = (2.0 * OR_PI) * rn(S, I->id, &I->nd);
    h->t0 = I->t0; h->X = X; h->Y = Y; h->Z = Z; h->sh = sh; h->theta = hth; h->phi = hph;
    hole_params(S, h, dE - dE_cur, I->id, &I->nd);
    cut_ho(c, h);
    if (Z >= c->m.Layer) I->tn = 1.0e16;
    trace_ev(S, 0, I->id, ev, t_ev, I->E, I->Zeff, I->X, I->Y, I->Z, I->tn, I->nd);
    hpush(S, I->tn, K_SHI, 0);
    hpush(S, e->tn, K_EL, k);
    hpush(S, h->tn, K_HO, k);
}

/* ---------------------------------------------------------------- electron */
/* barrier_parameters, MC:2094-2116 (g_h is the reduced Planck constant) */
int or_barrier_parameters(double W, double bar_length, double H, double *E1, double *gamma)
{
    double Em_L = bar_length * 1.0e-10;
    double Em_B = 2.0 * H - W + 2.0 * sqrt(H * H - H * W);
    double Em_ksi = 0.5 * sqrt(8.0 * OR_ME * Em_L * Em_L * Em_B * OR_E / ((2.0 * OR_PI * OR_H) * (2.0 * OR_PI * OR_H)) - 1.0);
    double Em_bb = cosh(2.0 * OR_PI * Em_ksi);
    double Em_delta = 2.0 * OR_PI * Em_L * sqrt(2.0 * OR_ME * OR_E) / (2.0 * OR_PI * OR_H);
    double e1 = H + 2.0 * sqrt(H * (H - W)) * (acosh(Em_bb) / (Em_delta * (sqrt(H) + sqrt(H - W))) - 1.0);
    double g1 = Em_delta * (sqrt(e1) + sqrt(e1 - W));
    double g2 = Em_delta * (sqrt(e1) - sqrt(e1 - W));
    *E1 = e1;
    *gamma = (g1 * sinh(g1) + 2.0 * g2 * sinh(g2)) / (sqrt(e1 * (e1 - W)) * (Em_bb + cosh(g1)));
    return (isfinite(*E1) && isfinite(*gamma) && W > 0.0 && bar_length > 0.0 && H > W) ? 0 : 1;
}

/* calculate_emission, MC:2477-2513, for an electron whose collision point
 * has Z < 0: emitted (no further event, its kinetic energy outside E - W
 * recorded) or, if it moves outward, turned back */
static void emission(or_mc *S, p_el *e, uint32_t *nd)
{
    const or_ctx *c = S->c;
    int out = 0;
    double Ekin = e->E - c->em_W;
    if (e->E >= 1.5 * c->em_H) {
        out = 1;
    } else {
        double RN = rn(S, e->id, nd);
        double Em_Penetr = 1.0 / (1.0 + EXP(c->em_gamma * (c->em_E1 - e->E)));
        if ((Ekin > 0.0) && (RN < Em_Penetr)) out = 1;
        else if (COS(e->theta) < 0.0) e->theta = OR_PI - e->theta;
    }
    if (out) {
        e->tn = 1.0e30;
        e->L = 1.0e30;
        if (S->nem == S->capem) {
            S->capem = S->capem ? 2 * S->capem : 256;
            S->em = (double *)realloc(S->em, sizeof(double) * (size_t)S->capem);
        }
        S->em[S->nem++] = Ekin;
    }
}

/* Electron_Monte_Carlo, MC:2253-2474 */
static void electron_event(or_mc *S, int idx)
{
    or_ctx *c = S->c;
    p_el *e = &S->el[idx];
    uint64_t id = e->id;
    uint32_t nd = e->nd;
    uint32_t ev = e->ne++;
    double t_ev = e->tn;
    double Eel = e->E;
    double IMFP = or_nfp_2d(c, Eel, c->tab[TRK_TAB_E_INEL].E, c->El_IMFP, c->tab[TRK_TAB_E_INEL].n);
    double EMFP = or_nfp_2d(c, Eel, c->tab[TRK_TAB_E_EL].E, c->El_EMFP, c->tab[TRK_TAB_E_EL].n);
    double RN = rn(S, id, &nd);
    double L = e->L, th0 = e->theta, ph0 = e->phi;
    double X = e->X + (L * SIN(th0)) * SIN(ph0);
    double Y = e->Y + (L * SIN(th0)) * COS(ph0);
    double Z = e->Z + L * COS(th0);
    double dE, theta, phi;
    int kind;
    if (RN * (1.0 / IMFP + 1.0 / EMFP) < 1.0 / IMFP) {
        kind = 1;
        int sh = which_shell(S, &c->tab[TRK_TAB_E_INEL], Eel, id, &nd);
        int k = new_pair(S, id, ev, 0, 1);
        e = &S->el[idx];
        double IMFPs = or_nfp_1d(c, Eel, c->tab[TRK_TAB_E_INEL].E, c->tab[TRK_TAB_E_INEL].L[sh], c->tab[TRK_TAB_E_INEL].n);
        dE = e_transfer_inel(S, Eel, sh, IMFPs, 0, id, &nd);
        /* Update_electron_angles_El, MC:1189-1205 */
        theta = ACOS((Eel - dE) / SQRT(Eel * (Eel - dE)));
        if (theta != theta) theta = rn(S, id, &nd) * OR_PI;
        phi = (2.0 * OR_PI) * rn(S, id, &nd);
        double dE_cur = receives_E(S, dE, sh, id, &nd);
        double IM2 = or_nfp_2d(c, dE_cur, c->tab[TRK_TAB_E_INEL].E, c->El_IMFP, c->tab[TRK_TAB_E_INEL].n);
        double EM2 = or_nfp_2d(c, dE_cur, c->tab[TRK_TAB_E_EL].E, c->El_EMFP, c->tab[TRK_TAB_E_EL].n);
        double RN2 = rn(S, id, &nd);
        double MFP = -LOG(RN2) / (1.0 / IM2 + 1.0 / EM2);
        double theta2 = OR_PI / 2.0 - theta;
        double phi2 = phi + OR_PI;
        double phi1, theta1;
        new_angles(ph0, th0, theta2, phi2, &phi1, &theta1);
        p_el *ne = &S->el[k];
        p_ho *nh = &S->ho[k];
        ne->E = dE_cur; ne->t0 = e->tn; ne->X = X; ne->Y = Y; ne->Z = Z; ne->L = MFP; ne->theta = theta1; ne->phi = phi1;
        ne->tn = t_next(ne->t0, MFP, vel_el(ne->E));
        cut_el(c, ne);
        double hth = OR_PI * rn(S, id, &nd);
        double hph = (2.0 * OR_PI) * rn(S, id, &nd);
        nh->t0 = e->tn; nh->X = X; nh->Y = Y; nh->Z = Z; nh->sh = sh; nh->theta = hth; nh->phi = hph;
        hole_params(S, nh, dE - dE_cur, id, &nd);
        cut_ho(c, nh);
        hpush(S, ne->tn, K_EL, k);
        hpush(S, nh->tn, K_HO, k);
    } else {
        kind = 2;
        double EMFPt = or_nfp_1d(c, Eel, c->tab[TRK_TAB_E_EL].E, c->tab[TRK_TAB_E_EL].L[0], c->tab[TRK_TAB_E_EL].n);
        if (c->o.kind_of_EMFP == 2) {
            /* NRG_transfer_elastic_DSF, MC:2387-2390: two draws inside */
            double r1 = rn(S, id, &nd);
            double r2 = rn(S, id, &nd);
            dE = or_dsf_transfer(c, 0, Eel, r1, r2);
        } else if (c->o.kind_of_EMFP == 1) {
            dE = e_transfer_el(S, Eel, EMFPt, 0, id, &nd);
        } else {
            dE = 0.0;
            for (int a = 0; a < c->m.n_atoms; ++a) dE = dE + mott_transfer(S, a, Eel, 1.0, id, &nd) * c->m.Pers[a];
            dE = dE / c->sumPers;
        }
        /* Update_particle_angles_lat, MC:1252-1272 */
        double mu = cos_theta_W(S, Eel, dE, 1.0 * OR_ME, c->Mtarget, id, &nd);
        theta = ACOS(mu);
        phi = (2.0 * OR_PI) * rn(S, id, &nd);
        tally_elat(S, X, Y, dE);
    }
    double IM3 = or_nfp_2d(c, Eel - dE, c->tab[TRK_TAB_E_INEL].E, c->El_IMFP, c->tab[TRK_TAB_E_INEL].n);
    double EM3 = or_nfp_2d(c, Eel - dE, c->tab[TRK_TAB_E_EL].E, c->El_EMFP, c->tab[TRK_TAB_E_EL].n);
    double RN3 = rn(S, id, &nd);
    double MFP = -LOG(RN3) / (1.0 / IM3 + 1.0 / EM3);
    double phi1, theta1;
    new_angles(ph0, th0, theta, phi, &phi1, &theta1);
    e->E = Eel - dE; e->t0 = e->tn; e->X = X; e->Y = Y; e->Z = Z; e->L = MFP; e->theta = theta1; e->phi = phi1;
    e->tn = t_next(e->t0, MFP, vel_el(e->E));
    cut_el(c, e);
    if (c->em_W > 0.0 && Z < 0.0) emission(S, e, &nd);
    e->nd = nd;
    trace_ev(S, kind, id, ev, t_ev, e->E, 0.0, X, Y, Z, e->tn, nd);
    hpush(S, e->tn, K_EL, idx);
}

/* ---------------------------------------------------------------- holes */
/* The gap walk of check_hole_parameters, MC:696-719: a hole energy whose
 * DOS bracket is empty (< 1e-4) moves to the nearest nonempty level below
 * (down = 1, the excess goes to the electron) or above (down = 0, to the
 * lattice); the walk stops at the ends of the DOS array.  Returns 1 and the
 * level in *Eloc when the bracket was empty. */
static int gap_walk(const or_ctx *c, double Ehole, int down, double *Eloc)
{
    int m = or_find_mono_1d(c->dosE, c->m.n_dos, Ehole);
    if (!(c->dos[m - 1] < 1.0e-4)) return 0;
    if (down) {
        while (m > 1 && c->dos[m - 1] < 1.0e-4) m = m - 1;
    } else {
        while (m < c->m.n_dos && c->dos[m - 1] < 1.0e-4) m = m + 1;
    }
    *Eloc = c->dosE[m - 1];
    return 1;
}

double or_gap_walk(const or_ctx *c, double Ehole, int down)
{
    double Eloc = Ehole;
    gap_walk(c, Ehole, down, &Eloc);
    return Eloc;
}

/* check_hole_parameters, MC:682-721 (walks clamped to the DOS array) */
static double check_hole(or_mc *S, double Eel, double *dE, int with_el, int kel)
{
    double Ehole = Eel - *dE;
    double Eloc;
    if (gap_walk(S->c, Ehole, with_el, &Eloc)) {
        if (with_el) S->el[kel].E = S->el[kel].E + (Ehole - Eloc);
        else *dE = *dE + (Ehole - Eloc);
        Ehole = Eloc;
    }
    return Ehole;
}

/* Hole_Monte_Carlo, MC:2517-2869 */
static void hole_event(or_mc *S, int idx, double t_cur)
{
    or_ctx *c = S->c;
    p_ho *h = &S->ho[idx];
    uint64_t id = h->id;
    uint32_t nd = h->nd;
    uint32_t ev = h->ne++;
    double t_ev = h->tn;
    double Egap = c->Egap;
    int kind;
    if (h->sh == c->vb) {
        double Eel = h->Ehkin;
        double HIMFP = or_nfp_2d(c, Eel, c->tab[TRK_TAB_H_INEL].E, c->Hole_IMFP, c->tab[TRK_TAB_H_INEL].n);
        double HEMFP = or_nfp_2d(c, Eel, c->tab[TRK_TAB_H_EL].E, c->Hole_EMFP, c->tab[TRK_TAB_H_EL].n);
        double RN = rn(S, id, &nd);
        double L = h->L, th0 = h->theta, ph0 = h->phi;
        double X = h->X + (L * SIN(th0)) * SIN(ph0);
        double Y = h->Y + (L * SIN(th0)) * COS(ph0);
        double Z = h->Z + L * COS(th0);
        double dE, Ehole, hth1, hph1;
        if ((RN * (1.0 / HIMFP + 1.0 / HEMFP) < 1.0 / HIMFP) && (HIMFP < 1.0e15)) {
            kind = 3;
            int sh = which_shell(S, &c->tab[TRK_TAB_H_INEL], Eel, id, &nd);
            int k = new_pair(S, id, ev, 0, 1);
            h = &S->ho[idx];
            double HIs = or_nfp_1d(c, Eel, c->tab[TRK_TAB_H_INEL].E, c->tab[TRK_TAB_H_INEL].L[sh], c->tab[TRK_TAB_H_INEL].n);
            dE = e_transfer_inel(S, Eel, sh, HIs, 1, id, &nd);
            /* Update_holes_angles_el, MC:1142-1168 */
            double Mh = h->Mass * OR_ME;
            double E11 = Eel - dE;
            double hth = ACOS(SQRT(((((Mh + OR_ME) * (Mh + OR_ME)) / ((4.0 * Mh) * OR_ME)) * dE) / Eel));
            double hph = (2.0 * OR_PI) * rn(S, id, &nd);
            if (hth != hth) hth = OR_PI * rn(S, id, &nd);
            hth1 = ACOS(((Eel * (Mh - OR_ME)) + (E11 * (Mh + OR_ME))) / ((2.0 * Mh) * SQRT(Eel * E11)));
            hph1 = hph + OR_PI;
            if (hth1 != hth1) hth1 = OR_PI * rn(S, id, &nd);
            double dE_cur = receives_E(S, dE, sh, id, &nd);
            double IM = or_nfp_2d(c, dE_cur, c->tab[TRK_TAB_E_INEL].E, c->El_IMFP, c->tab[TRK_TAB_E_INEL].n);
            double EM = or_nfp_2d(c, dE_cur, c->tab[TRK_TAB_E_EL].E, c->El_EMFP, c->tab[TRK_TAB_E_EL].n);
            double RN2 = rn(S, id, &nd);
            double MFP = -LOG(RN2) / (1.0 / IM + 1.0 / EM);
            double theta2 = hth;
            (void)rn(S, id, &nd);        /* drawn and discarded, MC:2611 */
            double phi2 = hph;
            double phi1, theta1;
            new_angles(ph0, th0, theta2, phi2, &phi1, &theta1);
            p_el *ne = &S->el[k];
            p_ho *nh = &S->ho[k];
            ne->E = dE_cur; ne->t0 = h->tn; ne->X = X; ne->Y = Y; ne->Z = Z; ne->L = MFP; ne->theta = theta1; ne->phi = phi1;
            ne->tn = t_next(ne->t0, MFP, vel_el(ne->E));
            cut_el(c, ne);
            double nth = OR_PI * rn(S, id, &nd);
            double nph = (2.0 * OR_PI) * rn(S, id, &nd);
            nh->t0 = h->tn; nh->X = X; nh->Y = Y; nh->Z = Z; nh->sh = sh; nh->theta = nth; nh->phi = nph;
            hole_params(S, nh, dE - dE_cur, id, &nd);
            cut_ho(c, nh);
            Ehole = check_hole(S, Eel, &dE, 1, k);
            hpush(S, ne->tn, K_EL, k);
            hpush(S, nh->tn, K_HO, k);
        } else {
            kind = 4;
            double HEt = or_nfp_1d(c, Eel, c->tab[TRK_TAB_H_EL].E, c->tab[TRK_TAB_H_EL].L[0], c->tab[TRK_TAB_H_EL].n);
            if (c->o.kind_of_EMFP == 2) {
                /* NRG_transfer_elastic_DSF on the holes' tables, MC:2668-2677 */
                double r1 = rn(S, id, &nd);
                double r2 = rn(S, id, &nd);
                dE = or_dsf_transfer(c, 1, Eel, r1, r2);
            } else if (c->o.kind_of_EMFP == 1) {
                dE = e_transfer_el(S, Eel, HEt, 1, id, &nd);
            } else {
                dE = 0.0;
                for (int a = 0; a < c->m.n_atoms; ++a) dE = dE + mott_transfer(S, a, Eel, h->Mass, id, &nd) * c->m.Pers[a];
                dE = dE / c->sumPers;
            }
            double mu = cos_theta_W(S, Eel, dE, h->Mass * OR_ME, c->Mtarget, id, &nd);
            hth1 = ACOS(mu);
            hph1 = (2.0 * OR_PI) * rn(S, id, &nd);
            Ehole = check_hole(S, Eel, &dE, 0, 0);
            tally_elat(S, X, Y, dE);
        }
        double hph2, hth2;
        new_angles(ph0, th0, hth1, hph1, &hph2, &hth2);
        h->t0 = h->tn; h->X = X; h->Y = Y; h->Z = Z; h->theta = hth2; h->phi = hph2;
        hole_params(S, h, Ehole + Egap, id, &nd);
        cut_ho(c, h);
        h->nd = nd;
        trace_ev(S, kind, id, ev, t_ev, h->Ehkin, h->E, X, Y, Z, h->tn, nd);
        hpush(S, h->tn, K_HO, idx);
        return;
    }
    /* deep shell: Auger or radiative decay, MC:2739-2868 */
    double RN = rn(S, id, &nd);
    double tA = c->m.Auger[h->sh], tR = c->m.Radiat[h->sh];
    if (RN * (1.0 / tA + 1.0 / tR) < 1.0 / tA) {
        kind = 5;
        /* Auger_decay, MC:1366-1443 */
        int SHL = h->sh;
        int Sh1 = SHL, Sh2 = -1;
        double Ee = -1.0e-10, E_new1, E_new2 = 0.0;
        double coun = auger_count(c, c->m.Ip[SHL], 0);
        double Shel = rn(S, id, &nd) * coun;
        Sh1 = auger_choose(c, c->m.Ip[SHL], Shel, 0, Sh1);
        double dEc = 0.0;
        if (c->m.n_dos > 0 && Sh1 == c->vb) dEc = from_where_in_vb(S, 0, 0.0, id, &nd);
        E_new1 = dEc + c->m.Ip[Sh1];
        double Ediff = c->m.Ip[SHL] - E_new1;
        coun = auger_count(c, Ediff, 1);
        if (coun > 0.0) {
            Shel = rn(S, id, &nd) * coun;
            Sh2 = auger_choose(c, Ediff, Shel, 1, -1);
            dEc = 0.0;
            if (Sh2 >= 0 && c->m.n_dos > 0 && Sh2 == c->vb)
                dEc = from_where_in_vb(S, 1, Ediff - c->m.Ip[Sh2], id, &nd);
            E_new2 = dEc + ((Sh2 >= 0) ? c->m.Ip[Sh2] : 0.0);
            Ee = Ediff - E_new2;
        }
        if (Sh2 >= 0) {
            double hth = OR_PI * rn(S, id, &nd);
            double hph = (2.0 * OR_PI) * rn(S, id, &nd);
            h->t0 = h->tn; h->sh = Sh1; h->theta = hth; h->phi = hph;
            hole_params(S, h, E_new1, id, &nd);
            cut_ho(c, h);
            int k = new_pair(S, id, ev, 1, 0);
            h = &S->ho[idx];
            p_ho *nh = &S->ho[k];
            p_el *ne = &S->el[k];
            double nth = OR_PI * rn(S, id, &nd);
            double nph = (2.0 * OR_PI) * rn(S, id, &nd);
            nh->t0 = h->t0; nh->X = h->X; nh->Y = h->Y; nh->Z = h->Z; nh->sh = Sh2; nh->theta = nth; nh->phi = nph;
            hole_params(S, nh, E_new2, id, &nd);
            cut_ho(c, nh);
            double IM = or_nfp_2d(c, Ee, c->tab[TRK_TAB_E_INEL].E, c->El_IMFP, c->tab[TRK_TAB_E_INEL].n);
            double EM = or_nfp_2d(c, Ee, c->tab[TRK_TAB_E_EL].E, c->El_EMFP, c->tab[TRK_TAB_E_EL].n);
            double RN2 = rn(S, id, &nd);
            double MFP = -LOG(RN2) / (1.0 / IM + 1.0 / EM);
            double phi1 = (2.0 * OR_PI) * rn(S, id, &nd);
            double theta1 = OR_PI * rn(S, id, &nd);
            ne->E = Ee; ne->t0 = h->t0; ne->X = h->X; ne->Y = h->Y; ne->Z = h->Z; ne->L = MFP; ne->theta = theta1; ne->phi = phi1;
            ne->tn = t_next(ne->t0, MFP, vel_el(ne->E));
            cut_el(c, ne);
            hpush(S, ne->tn, K_EL, k);
            hpush(S, nh->tn, K_HO, k);
        } else {
            h->t0 = t_cur;
            h->tn = 1.0e21;
            kind = 5;
        }
    } else {
        kind = 6;
        /* Radiative_decay, MC:2969-2997 */
        int SHL = h->sh;
        double coun = auger_count(c, c->m.Ip[SHL], 1);
        double Shel = rn(S, id, &nd) * coun;
        int Sh1 = auger_choose(c, c->m.Ip[SHL], Shel, 1, SHL);
        double dEc = 0.0;
        if (c->m.n_dos > 0 && Sh1 == c->vb) dEc = from_where_in_vb(S, 0, 0.0, id, &nd);
        double E_new1 = dEc + c->m.Ip[Sh1];
        double dE = c->m.Ip[SHL] - E_new1;
        double hth = OR_PI * rn(S, id, &nd);
        double hph = (2.0 * OR_PI) * rn(S, id, &nd);
        h->t0 = h->tn; h->sh = Sh1; h->theta = hth; h->phi = hph;
        hole_params(S, h, E_new1, id, &nd);
        cut_ho(c, h);
        int k = new_photon(S, id, ev, 0);
        h = &S->ho[idx];
        p_el *p = &S->ph[k];
        double IM = or_nfp_2d(c, dE, c->tab[TRK_TAB_PHOT].E, c->Phot_IMFP, c->tab[TRK_TAB_PHOT].n);
        double RN2 = rn(S, id, &nd);
        double MFP = -LOG(RN2) * IM;
        double phi1 = (2.0 * OR_PI) * rn(S, id, &nd);
        double theta1 = OR_PI * rn(S, id, &nd);
        p->E = dE; p->t0 = h->t0; p->X = h->X; p->Y = h->Y; p->Z = h->Z; p->L = MFP; p->theta = theta1; p->phi = phi1;
        p->tn = t_next(p->t0, MFP, OR_CVEL);
        hpush(S, p->tn, K_PH, k);
    }
    h = &S->ho[idx];
    h->nd = nd;
    trace_ev(S, kind, id, ev, t_ev, h->Ehkin, h->E, h->X, h->Y, h->Z, h->tn, nd);
    hpush(S, h->tn, K_HO, idx);
}

/* ---------------------------------------------------------------- photons */
/* Photon_Monte_Carlo, MC:2873-2965 */
static void photon_event(or_mc *S, int idx)
{
    or_ctx *c = S->c;
    p_el *p = &S->ph[idx];
    uint64_t id = p->id;
    uint32_t nd = p->nd;
    uint32_t ev = p->ne++;
    double t_ev = p->tn;
    double Eel = p->E;
    double L = p->L, th0 = p->theta, ph0 = p->phi;
    double X = p->X + (L * SIN(th0)) * SIN(ph0);
    double Y = p->Y + (L * SIN(th0)) * COS(ph0);
    double Z = p->Z + L * COS(th0);
    int sh = which_shell(S, &c->tab[TRK_TAB_PHOT], Eel, id, &nd);
    int k = new_pair(S, id, ev, 0, 1);
    p = &S->ph[idx];
    double dE_cur = receives_E(S, Eel, sh, id, &nd);
    double IM = or_nfp_2d(c, dE_cur, c->tab[TRK_TAB_E_INEL].E, c->El_IMFP, c->tab[TRK_TAB_E_INEL].n);
    double EM = or_nfp_2d(c, dE_cur, c->tab[TRK_TAB_E_EL].E, c->El_EMFP, c->tab[TRK_TAB_E_EL].n);
    double RN = rn(S, id, &nd);
    double MFP = -LOG(RN) / (1.0 / IM + 1.0 / EM);
    double phi1, theta1;
    new_angles(p->phi, p->theta, OR_PI / 2.0, 0.0, &phi1, &theta1);
    p_el *ne = &S->el[k];
    p_ho *nh = &S->ho[k];
    ne->E = dE_cur; ne->t0 = p->tn; ne->X = X; ne->Y = Y; ne->Z = Z; ne->L = MFP; ne->theta = theta1; ne->phi = phi1;
    ne->tn = t_next(ne->t0, MFP, vel_el(ne->E));
    cut_el(c, ne);
    double hth = OR_PI * rn(S, id, &nd);
    double hph = (2.0 * OR_PI) * rn(S, id, &nd);
    nh->t0 = p->tn; nh->X = X; nh->Y = Y; nh->Z = Z; nh->sh = sh; nh->theta = hth; nh->phi = hph;
    hole_params(S, nh, Eel - dE_cur, id, &nd);
    cut_ho(c, nh);
    hpush(S, ne->tn, K_EL, k);
    hpush(S, nh->tn, K_HO, k);
    S->nph_alive--;
    p->E = 0.0; p->t0 = 1.0e27; p->tn = 1.0e27; p->X = 0.0; p->Y = 0.0; p->Z = 0.0; p->L = 1.0e26; p->theta = 0.0; p->phi = 0.0;
    p->nd = nd;
    trace_ev(S, 7, id, ev, t_ev, Eel, 0.0, X, Y, Z, p->tn, nd);
}

/* ---------------------------------------------------------------- statistics */
#define IX2(i, j, N1) ((int64_t)(i) + (int64_t)(N1) * (int64_t)(j))

/* Calculated_statistics, MC:881-1110 (the field parts are zero) */
static void statistics(or_mc *S, int i, double tim)
{
    or_ctx *c = S->c;
    trk_tallies *o = S->out;
    const int Nt = o->Nt, NR = S->NR, i0 = i - 1;
    const double cutoff = (c->m.cut_off > 0.0) ? c->m.cut_off : 0.0;
    double Se = 0.0, Sh = 0.0, Shk = 0.0, Sph = 0.0;
    for (int k = 0; k < S->nel; ++k) Se = Se + S->el[k].E;
    for (int k = 0; k < S->nel; ++k) Sh = Sh + S->ho[k].E;
    for (int k = 0; k < S->nel; ++k) Shk = Shk + S->ho[k].Ehkin;
    if (o->Out_tot_Ne) o->Out_tot_Ne[i0] += (double)S->nel;
    if (c->o.include_photons) {
        for (int k = 0; k < S->nph; ++k) Sph = Sph + S->ph[k].E;
        if (o->Out_tot_Nphot) o->Out_tot_Nphot[i0] += (double)S->nph_alive;
        if (o->Out_E_phot) o->Out_E_phot[i0] += Sph;
        if (o->Out_tot_E) o->Out_tot_E[i0] += ((((Se + Sh) + Shk) + Sph) + S->At_NRG) + 0.0;
    } else {
        if (o->Out_tot_E) o->Out_tot_E[i0] += (((Se + Sh) + Shk) + S->At_NRG) + 0.0;
    }
    if (o->Out_E_e) o->Out_E_e[i0] += Se;
    if (o->Out_E_at) o->Out_E_at[i0] += S->At_NRG;
    /* emitted electrons, MC:958-959, 1101-1109, on the energy grid Out_R/10 */
    if (c->em_W > 0.0) {
        double Sem = 0.0;
        for (int k = 0; k < S->nem; ++k) Sem = Sem + S->em[k];
        if (o->Out_Ne_Em) o->Out_Ne_Em[i0] += (double)S->nem;
        if (o->Out_E_Em) o->Out_E_Em[i0] += Sem;
        if (o->Out_Ee_vs_E_Em && S->nem > 0) {
            double *Out_E = (double *)malloc(sizeof(double) * (size_t)NR);
            for (int j = 0; j < NR; ++j) Out_E[j] = S->R[j] / 10.0;
            for (int k = 0; k < S->nem; ++k) {
                int j = or_find_mono_1d(Out_E, NR, S->em[k]);
                if (j > 1) o->Out_Ee_vs_E_Em[IX2(i0, j - 1, Nt)] += (1.0 / (Out_E[j - 1] - Out_E[j - 2])) / (double)S->nem;
                else o->Out_Ee_vs_E_Em[IX2(i0, j - 1, Nt)] += (1.0 / Out_E[j - 1]) / (double)S->nem;
            }
            free(Out_E);
        }
    }
    if (o->Out_E_h) {
        int ms = o->max_shells;
        for (int a = 0; a < c->m.n_atoms; ++a)
            for (int j = 0; j < c->m.n_shells[a]; ++j) {
                int s = c->shl_off[a] + j;
                double s1 = 0.0, s2 = 0.0;
                for (int k = 0; k < S->nel; ++k) if (S->ho[k].sh == s) s1 = s1 + S->ho[k].E;
                for (int k = 0; k < S->nel; ++k) if (S->ho[k].sh == s) s2 = s2 + S->ho[k].Ehkin;
                o->Out_E_h[i0 + (int64_t)Nt * (a + (int64_t)c->m.n_atoms * j)] += s1 + s2;
                (void)ms;
            }
    }
    if (c->o.include_photons && S->nph > 0) {
        for (int k = 0; k < S->nph; ++k) {
            p_el *p = &S->ph[k];
            if (p->E > 0.0) {
                double L0 = ((OR_CVEL * (tim - p->t0)) * 1.0e-5);
                if (L0 < 0.0) L0 = 0.0;
                double X = p->X + (L0 * SIN(p->theta)) * SIN(p->phi);
                double Y = p->Y + (L0 * SIN(p->theta)) * COS(p->phi);
                double R = SQRT(X * X + Y * Y);
                int j = or_find_mono_1d(S->R, NR, R);
                if (o->Out_nphot) o->Out_nphot[IX2(i0, j - 1, Nt)] += S->V[j - 1];
                if (o->Out_Ephot) o->Out_Ephot[IX2(i0, j - 1, Nt)] += p->E * S->V[j - 1];
            }
        }
    }
    int NVBtot = 0;
    for (int k = 0; k < S->nel; ++k) if (S->ho[k].sh == c->vb) NVBtot++;
    int NVB = 0;
    for (int k = 0; k < S->nel; ++k) {
        p_el *e = &S->el[k];
        double L0, th0, ph0;
        if (e->E > cutoff) {
            double V = vel_el(e->E);
            L0 = ((V * (tim - e->t0)) * 1.0e-5);
            if (L0 < 0.0) L0 = 0.0;
            th0 = e->theta;
            ph0 = e->phi;
        } else {
            L0 = 0.0; th0 = 0.0; ph0 = 0.0;
        }
        double X = e->X + (L0 * SIN(th0)) * SIN(ph0);
        double Y = e->Y + (L0 * SIN(th0)) * COS(ph0);
        double R = SQRT(X * X + Y * Y);
        int j = or_find_mono_1d(S->R, NR, R);
        if (j < 0) { c->err_numeric = 1; j = 1; }
        if (o->Out_ne) o->Out_ne[IX2(i0, j - 1, Nt)] += S->V[j - 1];
        if (o->Out_Ee) o->Out_Ee[IX2(i0, j - 1, Nt)] += e->E * S->V[j - 1];
        j = or_find_mono_1d(S->R, NR, e->E);
        if (j < 0) j = 1;
        if (o->Out_Ee_vs_E) {
            if (j > 1) o->Out_Ee_vs_E[IX2(i0, j - 1, Nt)] += (1.0 / (S->R[j - 1] - S->R[j - 2])) / (double)S->nel;
            else o->Out_Ee_vs_E[IX2(i0, j - 1, Nt)] += (1.0 / S->R[j - 1]) / (double)S->nel;
        }
        p_ho *h = &S->ho[k];
        if (h->sh == c->vb && o->Out_Eh_vs_E && c->m.n_dos > 1) {
            int jj = or_find_mono_1d(c->dosE, c->m.n_dos, h->Ehkin);
            if (jj < 0) jj = 1;
            if (jj > 1) o->Out_Eh_vs_E[IX2(i0, jj - 1, Nt)] += (1.0 / (c->dosE[jj - 1] - c->dosE[jj - 2])) / (double)NVBtot;
            else o->Out_Eh_vs_E[IX2(i0, jj - 1, Nt)] += (1.0 / (c->dosE[jj] - c->dosE[jj - 1])) / (double)NVBtot;
        }
        double xx = (th0 / OR_PI) * 180.0;
        if (e->E > 0.0 && o->Out_theta) {
            double grid[180];
            for (int g = 0; g < 180; ++g) grid[g] = (double)(g + 1);
            int jt = or_find_mono_1d(grid, 180, xx);
            if (jt < 0) jt = 1;
            o->Out_theta[IX2(i0, jt - 1, Nt + 1)] += 1.0 / (double)S->nel;
        }
        double Xh, Yh;
        if ((h->Mass < 1.0e3) && (h->Ehkin > cutoff)) {
            double V = vel_hole(h);
            L0 = ((V * (tim - h->t0)) * 1.0e-5);
            if (L0 < 0.0) L0 = 0.0;
            Xh = h->X + (L0 * SIN(h->theta)) * SIN(h->phi);
            Yh = h->Y + (L0 * SIN(h->theta)) * COS(h->phi);
            double xh = (h->theta / OR_PI) * 180.0;
            if (o->Out_theta_h) {
                double grid[180];
                for (int g = 0; g < 180; ++g) grid[g] = (double)(g + 1);
                int jt = or_find_mono_1d(grid, 180, xh);
                if (jt < 0) jt = 1;
                o->Out_theta_h[IX2(i0, jt - 1, Nt + 1)] += 1.0 / (double)NVBtot;
            }
        } else {
            Xh = h->X;
            Yh = h->Y;
        }
        double Rh = SQRT(Xh * Xh + Yh * Yh);
        int jh = or_find_mono_1d(S->R, NR, Rh);
        if (jh < 0) { c->err_numeric = 1; jh = 1; }
        int a = c->atom_of[h->sh];
        int sl = h->sh - c->shl_off[a];
        int64_t base = IX2(i0, jh - 1, Nt) + (int64_t)Nt * NR * (a + (int64_t)c->m.n_atoms * sl);
        if (o->Out_nh) o->Out_nh[base] += S->V[jh - 1];
        if (o->Out_Eh) o->Out_Eh[base] += h->E * S->V[jh - 1];
        if (o->Out_Ehkin) o->Out_Ehkin[base] += h->Ehkin * S->V[jh - 1];
        if (h->Mass < 1.0e3) {
            if (h->L < 1.0e3) {
                NVB++;
                double V = vel_hole(h);
                if (o->Out_diff_coeff) o->Out_diff_coeff[i0] = o->Out_diff_coeff[i0] + (((1.0 / 3.0) * V) * h->L) * 1.0e-6;
            }
        }
    }
    /* MC:1094: the running sum over all earlier iterations is divided again */
    if (o->Out_diff_coeff && NVB > 0) o->Out_diff_coeff[i0] = o->Out_diff_coeff[i0] / (double)NVB;
}

/* ---------------------------------------------------------------- driver */
/* set_time_grid MC:2118-2149 and the time_do loop bound MC:572-575 */
int or_build_time_grid(or_ctx *c)
{
    double Tim = c->o.Tim, dt = c->o.dt;
    int n;
    double *tg;
    if (c->o.dt_flag <= 0) {
        n = (int)ceil(Tim / dt) + 1;
        tg = (double *)calloc((size_t)n, sizeof(double));
        tg[0] = dt;
        for (int i = 1; i <= n - 2; ++i) tg[i] = tg[i - 1] + dt;
        tg[n - 1] = Tim + dt;
    } else {
        int i = 0;
        double t = 0.01;
        while (t <= Tim) { i++; t = t * dt; }
        n = i + 1;
        tg = (double *)calloc((size_t)n, sizeof(double));
        tg[0] = 0.01;
        for (int k = 1; k <= n - 2; ++k) tg[k] = tg[k - 1] * dt;
        tg[n - 1] = Tim + dt;
    }
    free(c->tim_glob);
    c->tim_glob = (double *)calloc((size_t)n, sizeof(double));
    double tg_cur = 0.0;
    int i = 0;
    while (tg_cur <= Tim - 1.0e-6 && i < n) {
        tg_cur = (tg[i] < Tim) ? tg[i] : Tim;
        c->tim_glob[i] = tg_cur;
        i++;
    }
    c->Nt = i;
    free(tg);
    return 0;
}

int or_time_nodes(const or_ctx *c, int32_t *Nt, double *tim_glob, int32_t cap)
{
    *Nt = c->Nt;
    if (tim_glob) for (int i = 0; i < c->Nt && i < cap; ++i) tim_glob[i] = c->tim_glob[i];
    return 0;
}

/* Monte_Carlo_modelling, MC:427-679 */
static void one_iteration(or_mc *S)
{
    or_ctx *c = S->c;
    S->nel = 0;
    S->nph = 0;
    S->nph_alive = 0;
    S->nh = 0;
    S->nem = 0;
    S->At_NRG = 0.0;
    p_shi *I = &S->shi;
    memset(I, 0, sizeof(*I));
    I->E = c->ion.E;
    I->Zeff = c->shi_Zeff0;
    I->id = 0;
    double SHI_IMFP = or_nfp_2d(c, I->E, c->tab[TRK_TAB_SHI].E, c->SHI_path, c->tab[TRK_TAB_SHI].n);
    double RN = rn(S, I->id, &I->nd);
    SHI_IMFP = -SHI_IMFP * LOG(RN);
    double V = SQRT(((2.0 * I->E) * OR_E) / (c->ion.Mass * OR_MP));
    I->tn = t_next(0.0, SHI_IMFP, V);
    I->L = SHI_IMFP;
    hpush(S, I->tn, K_SHI, 0);
    double Tim = c->o.Tim;
    for (int i = 1; i <= c->Nt; ++i) {
        double tim_glob = c->tim_glob[i - 1];
        double lim = (tim_glob < Tim) ? tim_glob : Tim;
        S->inode = i;
        hnode h;
        while (hpeek(S, &h) && h.t < lim) {
            hpop(S);
            double t_cur = h.t;
            switch (h.kind) {
            case K_SHI: shi_event(S); break;
            case K_EL: electron_event(S, h.idx); break;
            case K_HO: hole_event(S, h.idx, t_cur); break;
            default: photon_event(S, h.idx); break;
            }
        }
        statistics(S, i, tim_glob);
    }
}

int or_run_mc(or_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter,
              const double *Out_R, const double *Out_V, int32_t NR, trk_tallies *out,
              double *trace, int64_t trace_cap, int64_t *trace_n)
{
    if (!c->tables_ready) { snprintf(c->err, sizeof c->err, "tables not built"); return TRK_ERR_STATE; }
    if (out->Nt != c->Nt) { snprintf(c->err, sizeof c->err, "Nt mismatch %d vs %d", out->Nt, c->Nt); return TRK_ERR_ARG; }
    or_mc S;
    memset(&S, 0, sizeof S);
    S.c = c;
    S.seed = seed;
    S.R = Out_R;
    S.V = Out_V;
    S.NR = NR;
    S.out = out;
    S.trace = trace;
    S.tcap = trace_cap;
    int64_t nparticles = 0;
    for (int64_t it = 0; it < nmc; ++it) {
        S.iter = (uint32_t)(first_iter + it);
        one_iteration(&S);
        nparticles += 2 * (int64_t)S.nel + S.nph;
    }
    for (int k = 0; k < 8; ++k) out->n_events[k] += S.n_events[k];
    out->n_particles += nparticles / 2 + 0;
    if (trace_n) *trace_n = S.tn;
    free(S.el); free(S.ho); free(S.ph); free(S.heap); free(S.em);
    if (c->err_numeric) { snprintf(c->err, sizeof c->err, "NaN reached a table search"); return TRK_ERR_NUMERIC; }
    return 0;
}

/* The same iterations spread over OpenMP threads, each with private tally
 * copies summed at the end (the reference's OpenMP reduction over
 * iterations, MC:395-415).  Sums are not in iteration order, so results are
 * not bitwise those of or_run_mc; Out_diff_coeff (an order-dependent
 * recurrence) is not produced.  Used to time the CPU path on all cores. */
int or_run_mc_threads(or_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter,
                      const double *Out_R, const double *Out_V, int32_t NR, trk_tallies *out, int threads)
{
    if (!c->tables_ready) { snprintf(c->err, sizeof c->err, "tables not built"); return TRK_ERR_STATE; }
    if (out->Nt != c->Nt) { snprintf(c->err, sizeof c->err, "Nt mismatch %d vs %d", out->Nt, c->Nt); return TRK_ERR_ARG; }
    const int Nt = out->Nt, na = c->m.n_atoms, ms = out->max_shells, nd = out->n_dos;
    const size_t n2 = (size_t)Nt * NR, nh = n2 * na * ms, n1 = (size_t)Nt, ns = (size_t)Nt * na * ms;
    const size_t nth = (size_t)(Nt + 1) * 180, ndos = (size_t)Nt * (nd > 0 ? nd : 1);
    int64_t ev[8] = {0};
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
    {
        trk_tallies L = *out;
        double **dst[] = {&L.Out_ne, &L.Out_Ee, &L.Out_nphot, &L.Out_Ephot, &L.Out_Ee_vs_E, &L.Out_Elat,
                          &L.Out_nh, &L.Out_Eh, &L.Out_Ehkin, &L.Out_tot_Ne, &L.Out_tot_Nphot, &L.Out_tot_E,
                          &L.Out_E_e, &L.Out_E_phot, &L.Out_E_at, &L.Out_E_h, &L.Out_theta, &L.Out_theta_h,
                          &L.Out_Eh_vs_E, &L.Out_Ne_Em, &L.Out_E_Em, &L.Out_Ee_vs_E_Em};
        double *const src[] = {out->Out_ne, out->Out_Ee, out->Out_nphot, out->Out_Ephot, out->Out_Ee_vs_E,
                               out->Out_Elat, out->Out_nh, out->Out_Eh, out->Out_Ehkin, out->Out_tot_Ne,
                               out->Out_tot_Nphot, out->Out_tot_E, out->Out_E_e, out->Out_E_phot, out->Out_E_at,
                               out->Out_E_h, out->Out_theta, out->Out_theta_h, out->Out_Eh_vs_E,
                               out->Out_Ne_Em, out->Out_E_Em, out->Out_Ee_vs_E_Em};
        const size_t len[] = {n2, n2, n2, n2, n2, n2, nh, nh, nh, n1, n1, n1, n1, n1, n1, ns, nth, nth, ndos, n1, n1, n2};
        const int na_ = (int)(sizeof len / sizeof len[0]);
        for (int k = 0; k < na_; ++k) *dst[k] = src[k] ? (double *)calloc(len[k], sizeof(double)) : NULL;
        L.Out_diff_coeff = NULL;
        for (int k = 0; k < 8; ++k) L.n_events[k] = 0;
        or_mc S;
        memset(&S, 0, sizeof S);
        S.c = c; S.seed = seed; S.R = Out_R; S.V = Out_V; S.NR = NR; S.out = &L;
#pragma omp for schedule(dynamic, 1)
        for (int64_t it = 0; it < nmc; ++it) {
            S.iter = (uint32_t)(first_iter + it);
            one_iteration(&S);
        }
#pragma omp critical
        {
            for (int k = 0; k < na_; ++k)
                if (src[k])
                    for (size_t i = 0; i < len[k]; ++i) src[k][i] += (*dst[k])[i];
            for (int k = 0; k < 8; ++k) ev[k] += S.n_events[k];
        }
        for (int k = 0; k < na_; ++k) free(*dst[k]);
        free(S.el); free(S.ho); free(S.ph); free(S.heap); free(S.em);
    }
    for (int k = 0; k < 8; ++k) out->n_events[k] += ev[k];
    if (c->err_numeric) { snprintf(c->err, sizeof c->err, "NaN reached a table search"); return TRK_ERR_NUMERIC; }
    return 0;
}
