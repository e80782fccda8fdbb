/*
 * Internal definitions of the oracle (TEST INFRASTRUCTURE ONLY; see
 * trk_oracle.h).  Physical constants restate Universal_Constants.f90:28-60
 * with the same defining expressions, so each one rounds to the same double.
 */
#ifndef TRK_ORACLE_INTERNAL_H
#define TRK_ORACLE_INTERNAL_H

#include <stdint.h>
#include <stdio.h>
#include "trk_oracle.h"
#include "trk_detmath.h"
#include "trk_philox.h"

/* Universal_Constants.f90:28-60 */
#define OR_PI    3.1415926535897932384626433832795
#define OR_E     1.602176487e-19
#define OR_ME    9.1093821545e-31
#define OR_CVEL  299792458.0
#define OR_MP    (1836.1526724780 * OR_ME)
#define OR_H     1.05457162853e-34
#define OR_KB    11604.0
#define OR_E0    8.854187817620e-12
#define OR_RY    13.6056981
#define OR_A0    0.5291772085936
#define OR_ALPHA (OR_E * OR_E / (OR_H * OR_CVEL * 4.0 * OR_PI * OR_E0))

#define OR_NTAB 7

typedef struct {
    int32_t n;
    double *hw;
    double *lam;
} or_row;

typedef struct {
    int32_t n;                       /* grid size (SHI_DIFF: unused) */
    int32_t nsh;                     /* number of "shells" in the table */
    double *E;                       /* grid */
    double *L[TRK_MAX_SHELLS];
    double *dEdx[TRK_MAX_SHELLS];
    or_row *rows[TRK_MAX_SHELLS];    /* n rows per shell or NULL */
    /* SHI_DIFF rows: own length per shell */
    int32_t dn[TRK_MAX_SHELLS];
    double *dE[TRK_MAX_SHELLS];
    double *dL[TRK_MAX_SHELLS];
} or_table;

struct or_ctx {
    trk_material m;
    trk_ion ion;
    trk_options o;
    /* copies of caller DOS arrays */
    double *dosE, *dos, *dosI, *dosk, *dosM, *doskI, *dosMI;
    int nsh;                         /* total shells */
    int shl_off[TRK_MAX_ATOMS + 1];
    int osc_off[TRK_MAX_SHELLS + 1];
    int atom_of[TRK_MAX_SHELLS];
    int vb;                          /* flat index of the VB shell */
    double Egap;                     /* Ip of the last shell of atom 0 */
    double Mtarget;                  /* mean atomic mass [kg] */
    double Zt;                       /* mean atomic number */
    double sumPers;
    double e_emax;                   /* electron grid top */
    double shi_Zeff0;                /* ion charge at the initial energy */
    double shi_lo, shi_hi;           /* optional evaluation window of the ion table */
    or_table tab[OR_NTAB];
    int tables_ready;
    /* fused totals (How_many_electrons) */
    double *El_IMFP, *El_EMFP, *Hole_IMFP, *Hole_EMFP, *Phot_IMFP, *SHI_path;
    /* surface barrier (emission when em_W > 0): work function, height and
     * the logistic transmission's E1 and gamma */
    double em_W, em_H, em_E1, em_gamma;
    /* Delta-CDF weight of every oscillator (kind_of_DR = 4) */
    double alpha[TRK_MAX_OSC];
    /* DSF tables of electrons [0] and holes [1] (kind_of_EMFP = 2) */
    int32_t dsf_nE[2], dsf_nW[2];
    double *dsfE[2], *dsfdE[2], *dsfL[2], *dsfLe[2], *dsfLa[2];
    /* time grid */
    int32_t Nt;
    double *tim_glob;
    int err_numeric;
    char err[256];
};

/* shared helpers between the table and MC halves */
double or_imewq(const or_ctx *c, double hw, double hq, int shell, int photon);
double or_shi_zeff(const or_ctx *c, double E);
double or_rest_energy(double M0);
/* DSF elastic scattering (trk_oracle_dsf.c) */
void or_dsf_free(or_ctx *c);
void or_dsf_elastic_table(const or_ctx *c, int particle, double **E, double **L, double **dEdx, int32_t *n);
/* the Delta-CDF model (trk_oracle_delta.c) */
double or_delta_alpha(double A, double E0, double Gamma, double x_min);
double or_delta_cs(const or_ctx *c, int shell, double E, double Ip, int has_emax, double Emax_in);
double or_delta_eloss(const or_ctx *c, int shell, double E, double Ip);
void   or_delta_tot_imfp(const or_ctx *c, int shell, double Ele, double Emin, double *L, double *dEdx);
double or_delta_transfer(const or_ctx *c, int shell, double Ee, double Ip, double RN);
/* the BEB cross section (trk_oracle_beb.c) */
double or_beb_sigma(double T, double B, double U, double N);
double or_beb_dsigma_int(double T, double w, double B, double U, double N);
double or_beb_dsigma_w_int(double T, double w, double B, double U, double N);
void   or_beb_tot_imfp(const or_ctx *c, int shell, double Ele, double Mass, double *L, double *dEdx);
double or_beb_transfer(const or_ctx *c, int shell, double Ele, double L_need, double Mass, double Emin);
double or_nfp_1d(or_ctx *c, double E, const double *Earr, const double *Larr, int32_t n);
double or_nfp_2d(or_ctx *c, double E, const double *Earr, const double *Larr, int32_t n);
double or_interp_transfer(or_ctx *c, const or_table *t, int shell, double Ele, double L_need);
int    or_build_time_grid(or_ctx *c);

#endif
