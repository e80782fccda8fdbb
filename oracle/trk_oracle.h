/*
 * trk_oracle.h -- CPU restatement of the TREKIS-3 particle-history path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * implementation: only the tests, the smoke check and the CPU-baseline leg of
 * the benchmark load it.  Nothing in the shipped library links it or calls it.
 *
 * It restates, in plain sequential C, what the reference Fortran computes
 * (file:line citations on every function in the .c files), in the
 * reference's own order of operations, and takes the same input descriptors
 * as the C-ABI (include/trekis_gpu.h).  Elementary functions come from
 * include/trk_detmath.h and random numbers from include/trk_philox.h, the
 * same platform layer the kernels use, so that a GPU history and an oracle
 * history drawn from the same stream agree bit for bit.
 */
#ifndef TRK_ORACLE_H
#define TRK_ORACLE_H

#include <stdint.h>
#include "trekis_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx *or_create(const trk_material *mat, const trk_ion *ion, const trk_options *opt);
void    or_destroy(or_ctx *c);
const char *or_error(const or_ctx *c);

/* Build all tables (threads: OpenMP threads over energy nodes, 0 = default). */
int or_build_tables(or_ctx *c, int threads);
void or_set_shi_window(or_ctx *c, double lo, double hi);
/* Build only the grid and the fused arrays from tables already injected. */
int or_finish_tables(or_ctx *c);

/* Table access, same conventions as the C-ABI trk_table_* / trk_row_*. */
int or_table_size(const or_ctx *c, int kind, int shell, int32_t *n);
int or_table_get(const or_ctx *c, int kind, int shell, double *E, double *L, double *dEdx, int32_t n);
int or_row_size(const or_ctx *c, int kind, int shell, int32_t i_E, int32_t *n);
int or_row_get(const or_ctx *c, int kind, int shell, int32_t i_E, double *hw, double *lam, int32_t n);
/* Inject a table (grid + L + dEdx) or a row; used to replay the Monte Carlo
 * on tables built elsewhere.  Rows must be set after the table of that kind. */
int or_table_set(or_ctx *c, int kind, int shell, const double *E, const double *L, const double *dEdx, int32_t n);
int or_row_set(or_ctx *c, int kind, int shell, int32_t i_E, const double *hw, const double *lam, int32_t n);

/* Single-point evaluations for unit tests. */
double or_loss_func(const or_ctx *c, int mode, double A, double E0, double G, double hw, double hq);
double or_diff_cs(const or_ctx *c, int shell, double Ele, double hw, double Mass);
double or_diff_cs_phonon(const or_ctx *c, double Ele, double hw, double Mass);
double or_shi_diff_cs(const or_ctx *c, int shell, double Ele, double MSHI, double Emax, double hw);
/* Brandt-Kitagawa: the ion's effective charge at momentum hq, and the
 * ion's differential inverse mean free path with it (tests) */
double or_bk_charge(double hq, double Zp, double Zeff);
double or_shi_zeff(const or_ctx *c, double E);   /* Equilibrium_charge_SHI at energy E */
double or_shi_diff_cs_bk(const or_ctx *c, int shell, double Ele, double MSHI, double Emax, double hw);
int    or_inel_point(const or_ctx *c, int kind, int shell, double Ele, double *L, double *dEdx,
                     double *hw, double *lam, int32_t cap, int32_t *n);
int    or_grid(const trk_material *mat, double Emin, double Emax, double *out, int32_t cap, int32_t *n);
double or_define_dE(int n, double E, double E0_min, double E0_max, double dE_min);
int    or_find_mono_1d(const double *A, int32_t N, double v);
int    or_find_mono_2d(const double *A, int32_t N, double v);
int    or_find_decr(const double *A, int32_t N, double v);
int    or_find_linear(const double *A, int32_t N, double v);
double or_interpolate(int flag, double E1, double E2, double S1, double S2, double E);
int    or_sumrules(const double *A, const double *E0, const double *G, int32_t n,
                   double x_min, double Omega, double *ksum, double *fsum);
double or_int_ritchi_x(double A, double E, double G, double x);
double or_int_ritchi_p_x(double A, double E, double G, double x);
double or_w_plasma(double dens, double mass);
int    or_prepare_dos(const double *raw_E, const double *raw_DOS, int32_t n_raw,
                      double N_VB_el, double At_Dens, double sum_Pers,
                      double *E, double *DOS, double *int_DOS, double *k, double *Eff_m,
                      double *k_inv, double *Eff_m_inv, int32_t M);
int    or_time_nodes(const or_ctx *c, int32_t *Nt, double *tim_glob, int32_t cap);
/* inverse-CDF energy-transfer samplers for a given random number (tests) */
double or_sample_transfer(or_ctx *c, int kind, int shell, double Ele, double RN);
/* elastic kinematics and decay-shell choice (tests) */
double or_mott_theta(const or_ctx *c, int atom, double Ee, double RN, double *nu_out);
double or_mott_W(const or_ctx *c, int atom, double Ee, double theta, double M_eff);
double or_cos_theta_W(const or_ctx *c, double E, double W, double M_in);
int    or_auger_pick(const or_ctx *c, double NRG, double frac, int second);
/* angle composition (New_Angles_both) and the hole's gap walk (tests) */
void   or_new_angles(double phi0, double theta0, double theta, double psi, double *phi1, double *theta1);
double or_gap_walk(const or_ctx *c, double Ehole, int down);
double or_constant(int which);
/* barrier_parameters, MC:2094-2116; returns 0 when E1 and gamma are finite */
int    or_barrier_parameters(double W, double bar_length, double H, double *E1, double *gamma);
/* DSF elastic tables of electrons (particle 0) or holes (1), set before
 * or_build_tables, and the transfer sampler for two given draws */
int    or_dsf_set(or_ctx *c, int particle, int32_t nE, int32_t nW, const double *E, const double *dE,
                  const double *dL, const double *dL_emit, const double *dL_absorb);
double or_dsf_transfer(const or_ctx *c, int particle, double Eel, double RN1, double RN2);

/* Monte Carlo (do_Monte_Carlo in the reference's global time order).  If
 * trace != NULL each processed event appends 12 doubles, the record layout of
 * trk_set_trace: {iteration, particle id (bits), event no, kind, time of the
 * event, E, E2, X, Y, Z, time of the next event, draws used so far}. */
int or_run_mc(or_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter,
              const double *Out_R, const double *Out_V, int32_t NR, trk_tallies *out,
              double *trace, int64_t trace_cap, int64_t *trace_n);

/* The same on OpenMP threads over iterations (timing of the CPU path on all
 * cores; sums not in iteration order, Out_diff_coeff not produced). */
int or_run_mc_threads(or_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter,
                      const double *Out_R, const double *Out_V, int32_t NR, trk_tallies *out, int threads);

#ifdef __cplusplus
}
#endif
#endif
