/*
 * trekis_gpu.h -- C-ABI of the MI355X particle-history library.
 *
 * This is the drop-in boundary for the TREKIS-3 hot path.  The Fortran driver
 * keeps reading inputs and sorting outputs; it flattens its derived types into
 * the plain structs below and calls these entry points in place of
 *   - the table builders  Analytical_electron_dEdx / Analytical_ion_dEdx /
 *     SHI_TotIMFP        (Analytical_IMFPs.f90:100-117, 2242-2254;
 *                         Cross_sections.f90:2452; called at
 *                         Universal_MC_for_SHI_MAIN.f90:166, 184-198, 231-238)
 *   - do_Monte_Carlo     (Monte_Carlo.f90:39-92, called at
 *                         Universal_MC_for_SHI_MAIN.f90:271-276)
 *   - get_single_pole/sumrules f-sum normalisation (Cross_sections.f90:554-749)
 *   - the VB density-of-states derived arrays (Reading_files_and_parameters.f90:2395-2450)
 * Every signature uses plain pointers, sizes and fixed-size arrays; no C++ or
 * torch types cross it.  All entry points return 0 on success and a nonzero
 * trk_status otherwise; the message is available from trk_last_error().  The
 * library never aborts the process.
 *
 * Units follow the reference: energies [eV], lengths [Angstrom], times [fs],
 * masses of atoms and ions in proton masses (Objects.f90:44-53, 249-265),
 * effective masses in electron masses, densities [g/cm^3] and [1/cm^3].
 */
#ifndef TREKIS_GPU_H
#define TREKIS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRK_MAX_ATOMS   8
#define TRK_MAX_SHELLS  32     /* over all atoms */
#define TRK_MAX_OSC     128    /* Drude-Lorentz oscillators over all shells */
#define TRK_MAX_PHONON  16

enum trk_status {
    TRK_OK = 0,
    TRK_ERR_ARG = 1,        /* invalid argument or inconsistent descriptor */
    TRK_ERR_DEVICE = 2,     /* HIP runtime error */
    TRK_ERR_STATE = 3,      /* call out of order (e.g. MC before tables) */
    TRK_ERR_NUMERIC = 4,    /* NaN in a particle state or table (the reference stops there) */
    TRK_ERR_CAPACITY = 5    /* particle or table storage exhausted */
};

/* Target material, flattened from the reference types Atom, Solid, CDF and
 * Density_of_states (Objects.f90:92-130, 208-215, 249-265).  Shells are
 * stored atom-major: shell k of atom a has flat index sum_{b<a} n_shells[b] + k.
 * The valence band is the last shell of atom 0 (Find_VB_numbers,
 * Reading_files_and_parameters.f90:3252-3270 with Shl_num >= 63). */
typedef struct trk_material {
    int32_t n_atoms;
    int32_t Zat[TRK_MAX_ATOMS];
    double  Mass[TRK_MAX_ATOMS];        /* [proton masses] */
    double  Pers[TRK_MAX_ATOMS];        /* atoms of this kind per formula unit */
    int32_t n_shells[TRK_MAX_ATOMS];    /* >= 1 for atom 0 (its last shell is the VB); 0 allowed for the others */

    double  Ip[TRK_MAX_SHELLS];         /* [eV] ionisation potential (>= 0.1, RF:1554) */
    double  Nel[TRK_MAX_SHELLS];        /* electrons in the shell */
    double  Auger[TRK_MAX_SHELLS];      /* [fs] Auger decay time */
    double  Radiat[TRK_MAX_SHELLS];     /* [fs] radiative decay time */
    int32_t Shl_num[TRK_MAX_SHELLS];    /* EADL designator, >= 63 marks the VB */
    int32_t n_osc[TRK_MAX_SHELLS];      /* oscillators per shell, flattened shell-major below */
    double  osc_E0[TRK_MAX_OSC];
    double  osc_A[TRK_MAX_OSC];
    double  osc_Gamma[TRK_MAX_OSC];

    int32_t n_phonon;                   /* phonon (quasi-elastic) CDF oscillators */
    double  ph_E0[TRK_MAX_PHONON];
    double  ph_A[TRK_MAX_PHONON];
    double  ph_Gamma[TRK_MAX_PHONON];

    double  Dens;                       /* [g/cm^3] */
    double  At_Dens;                    /* [1/cm^3] */
    double  Vsound;                     /* [m/s] */
    double  Layer;                      /* [A] track segment analysed */
    double  N_VB_el;                    /* VB electrons per formula unit */
    double  El_eff_mass;                /* 0: from DOS; >0 constant [me] */
    double  hole_mass;                  /* >=0 constant [me]; <0 from DOS */
    double  E_F;                        /* [eV] */
    double  v_f;                        /* sqrt(2 E_F / me), reference units */
    double  cut_off;                    /* [eV] */
    double  temp;                       /* [K] */
    double  work_function;              /* [eV]; > 0 turns on emission through the surface Z = 0
                                           (calculate_emission, Monte_Carlo.f90:2477-2513) */
    double  bar_length;                 /* [A] surface barrier length (needed when work_function > 0) */
    double  bar_height;                 /* [eV] surface barrier height (> work_function) */

    int32_t n_dos;                      /* DOS arrays, caller-owned, length n_dos */
    const double *dos_E;                /* [eV] from the VB top, increasing */
    const double *dos_DOS;
    const double *dos_int_DOS;
    const double *dos_k;                /* [1/m] */
    const double *dos_Eff_m;            /* [me] */
    const double *dos_k_inv;
    const double *dos_Eff_m_inv;

    /* Inelastic cross section of electrons and holes per shell (KOCS,
     * Objects.f90:273): 0 or 1 the shell's CDF, 2 the binary-encounter-Bethe
     * formula of Kim and Rudd with the shell's mean kinetic energy Ek [eV]
     * (Sigma_BEB / dSigma_int_BEB, Cross_sections.f90:3888-4044; transfers by
     * Electron_NRG_transfer_BEB, CS:2126-2161).  The ion always uses the CDF
     * (the reference marks its BEB branch as not working, MC:1772), so every
     * shell keeps at least one oscillator. */
    int32_t kocs[TRK_MAX_SHELLS];
    double  Ek[TRK_MAX_SHELLS];
} trk_material;

/* Swift heavy ion (Objects.f90:44-53). */
typedef struct trk_ion {
    int32_t Zat;
    double  Mass;                       /* [proton masses] */
    double  E;                          /* [eV] total kinetic energy */
    int32_t Kind_Zeff;                  /* 0 Barkas, 1 Bohr, 2 Nikolaev-Dmitriev, 3 Schiwietz-Grande, 4 fixed */
    double  fixed_Zeff;
    int32_t Kind_ion;                   /* 0 point charge, 1 Brandt-Kitagawa charge in the ion's mean free
                                           paths (SHI_Total_IMFP, Cross_sections.f90:2431-2446) */
} trk_ion;

/* Numerical options (the Flag type, Objects.f90:135-180, and run controls). */
typedef struct trk_options {
    int32_t kind_of_EMFP;               /* -1 no elastic, 0 Mott, 1 phonon CDF, 2 DSF (tables from
                                           trk_dsf_set before trk_build_tables) */
    int32_t CDF_elast_Zeff;             /* 0 Barkas-like target charge, 1 unit charge */
    int32_t kind_of_DR;                 /* 1 free electron, 2 plasmon pole, 3 Ritchie, 4 Delta-CDF (the
                                           electron and hole inelastic cross sections of delta peaks,
                                           Cross_sections.f90:952-956, 1894-1896; loss functions elsewhere
                                           disperse as 1) */
    int32_t plasmon_Emax;               /* use plasmon energy as integration limit */
    int32_t include_photons;            /* radiative decay and photon transport */
    int32_t dt_flag;                    /* 0 linear, 1 logarithmic time grid */
    double  Tim;                        /* [fs] total time */
    double  dt;                         /* [fs] time step (log grid: ratio) */
    double  e_grid_Emax;                /* [eV] top of the electron/photon energy grid;
                                           <= 0 selects the reference value 175.6e6*2/1836 (AI:143) */
    double  shi_grid_Emax;              /* [eV] top of the ion energy grid;
                                           <= 0 selects the reference value 175.6e6/2*M (AI:2272) */
    int32_t precision;                  /* 0 FP64 everywhere (replays the oracle bit for bit);
                                           1 mixed: FP32 elementary functions in the event code,
                                           FP64 state, tables, arithmetic and tallies */
    int32_t batch;                      /* iterations per device batch, 0 = automatic */
    int32_t block;                      /* threads per workgroup for transport, 0 = automatic */
    int32_t tally_replicas;             /* copies of the tally arrays (contention), 0 = automatic */
    int32_t gen_capacity;               /* particle pairs per generation buffer, 0 = automatic */
    int32_t quadrature;                 /* CDF tables (inelastic electron/hole, ion):
                                           0 the reference's quadrature in FP64 (replays the oracle bit for bit);
                                           1 FP32 with compensated (Kahan) sums: 2000 geometric transfer-energy
                                             intervals (4001 points) x 1000 Simpson intervals in ln q (2001
                                             points), FP64 only for the cumulative normalisation; needs a
                                             constant electron effective mass (El_eff_mass > 0) */
    int32_t reserved[2];
} trk_options;

/* Tallies: caller-owned arrays in Fortran column-major order with the
 * dimensions of Allocate_out_arrays (Sorting_output_data.f90:1144-1250).
 * trk_run_mc ADDS unnormalised sums over iterations into them (the driver
 * divides by NMC afterwards, Universal_MC_for_SHI_MAIN.f90:281-314).
 * Any pointer may be NULL to skip that tally. */
typedef struct trk_tallies {
    int32_t Nt;                         /* time nodes (must equal trk_time_nodes()) */
    int32_t NR;                         /* radial bins */
    int32_t n_dos;                      /* VB-hole spectrum bins (size of Mat_DOS%E) */
    int32_t max_shells;                 /* shells of atom 0 (size(target_atoms(1)%Ip)) */
    double *Out_ne, *Out_Ee;            /* (Nt, NR) */
    double *Out_nphot, *Out_Ephot;      /* (Nt, NR) */
    double *Out_Ee_vs_E;                /* (Nt, NR) spectrum binned on Out_R */
    double *Out_Eh_vs_E;                /* (Nt, n_dos) */
    double *Out_Elat;                   /* (Nt, NR) */
    double *Out_nh, *Out_Eh, *Out_Ehkin;/* (Nt, NR, n_atoms, max_shells) */
    double *Out_tot_Ne, *Out_tot_Nphot, *Out_tot_E;   /* (Nt) */
    double *Out_E_e, *Out_E_phot, *Out_E_at;          /* (Nt) */
    double *Out_E_h;                    /* (Nt, n_atoms, max_shells) */
    double *Out_theta, *Out_theta_h;    /* (Nt+1, 180) */
    double *Out_diff_coeff;             /* (Nt) */
    double *Out_Ne_Em, *Out_E_Em;       /* (Nt) electrons emitted through the surface, their kinetic energy */
    double *Out_Ee_vs_E_Em;             /* (Nt, NR) their spectrum, binned on Out_R/10 */
    int64_t n_events[8];                /* out: SHI, e-inel, e-el, h-inel, h-el, Auger, radiative, photon */
    int64_t n_particles;                /* out: electrons + holes + photons created */
} trk_tallies;

typedef struct trk_ctx trk_ctx;

/* Context lifecycle.  Copies all inputs; binds the current HIP device. */
int  trk_ctx_create(const trk_material *mat, const trk_ion *ion, const trk_options *opt, trk_ctx **out);
void trk_ctx_destroy(trk_ctx *ctx);
int  trk_last_error(const trk_ctx *ctx, char *buf, size_t n);

/* Build every table of the path on the device (replaces Cross_sections.f90
 * TotIMFP 881-1050, Tot_EMFP 2966-3139, Elastic_cross_section 2892-2963,
 * Tot_Phot_IMFP 833-878, SHI_TotIMFP 2452-2597 and the drivers in
 * Analytical_IMFPs.f90:1830-2239, 2510-2751), then fuse the per-shell
 * tables into totals (How_many_electrons, Monte_Carlo.f90:1902-2057). */
int  trk_build_tables(trk_ctx *ctx);

/* Table kinds for inspection / host injection. */
enum trk_table_kind {
    TRK_TAB_E_INEL = 0,   /* Total_el_MFPs  + aidCS%EIdCS rows */
    TRK_TAB_H_INEL = 1,   /* Total_Hole_MFPs + aidCS%HIdCS rows (VB only) */
    TRK_TAB_E_EL   = 2,   /* Elastic_MFP%Total + aidCS%EEdCS rows */
    TRK_TAB_H_EL   = 3,   /* Elastic_Hole_MFP%Total + aidCS%HEdCS rows */
    TRK_TAB_PHOT   = 4,   /* Total_Photon_MFPs */
    TRK_TAB_SHI    = 5,   /* SHI_MFP */
    TRK_TAB_SHI_DIFF = 6  /* diff_SHI_MFP at the initial ion energy */
};

/* Grid size of a table kind (for TRK_TAB_SHI_DIFF: row length of `shell`). */
int  trk_table_size(const trk_ctx *ctx, int kind, int shell, int32_t *n_out);
/* Copy E grid, L (mean free path) and dEdx of one shell to the host. */
int  trk_table_get(const trk_ctx *ctx, int kind, int shell, double *E, double *L, double *dEdx, int32_t n);
/* Length of cumulative row i_E of a row-carrying table kind and shell. */
int  trk_row_size(const trk_ctx *ctx, int kind, int shell, int32_t i_E, int32_t *n_out);
/* Copy one cumulative row (hw [eV], lambda_cum [A]) to the host. */
int  trk_row_get(const trk_ctx *ctx, int kind, int shell, int32_t i_E, double *hw, double *lam, int32_t n);

/* Tables read from the reference's cache files instead of computed
 * (Analytical_IMFPs.f90:259-383 reads OUTPUT_*_IMFPs*.dat through
 * read_write_IMFP 585-630, and 663-711 / 745-790 / 1450-1510 / 1540-1600 the
 * row files of the diff_CS folder through read(FN_diff,*)).  After trk_build_tables, each
 * call overwrites one shell's mean free path (L) and stopping power (dEdx)
 * of a table kind -- E, when not NULL, must equal the built grid bit for bit
 * -- or one cumulative row (hw, lambda_cum) of the same length as built;
 * a NULL array is left as it is.  An ion context accepts only the ion's
 * kinds (TRK_TAB_SHI, TRK_TAB_SHI_DIFF); set the others on its material.
 * trk_tables_finish then re-derives the fused totals, logarithms, row flags
 * and the ion's inverse transfer rows; trk_run_mc refuses to run between a
 * set and the finish. */
int  trk_table_set(trk_ctx *ctx, int kind, int shell, const double *E, const double *L, const double *dEdx,
                   int32_t n);
int  trk_row_set(trk_ctx *ctx, int kind, int shell, int32_t i_E, const double *hw, const double *lam, int32_t n);
int  trk_tables_finish(trk_ctx *ctx);

/* Elastic scattering from dynamic structure factors (kind_of_EMFP = 2):
 * the differential mean free paths DSF_DEMFP (electrons, particle 0) or
 * DSF_DEMFP_H (holes, particle 1) as the reference's reader leaves them
 * (reading_DSF_cross_sections, Reading_files_and_parameters.f90:2516-2680):
 * nE particle energies E (increasing) with nW transferred energies each,
 * E-major arrays of the transferred energy dE, the cumulative mean free
 * path dL and its emission (dE >= 0) and absorption parts dL_emit,
 * dL_absorb [A].  The elastic mean free paths are their last entries
 * (Analytical_IMFPs.f90:905-911); transfers are sampled as
 * NRG_transfer_elastic_DSF does (Cross_sections.f90:3652-3795).  Call for
 * both particles after trk_ctx_create and before trk_build_tables. */
int  trk_dsf_set(trk_ctx *ctx, int32_t particle, int32_t nE, int32_t nW, const double *E, const double *dE,
                 const double *dL, const double *dL_emit, const double *dL_absorb);

/* Number of snapshot time nodes of the run (time_do loop, Monte_Carlo.f90:572-665). */
int  trk_time_nodes(const trk_ctx *ctx, int32_t *Nt_out);

/* Run NMC iterations (do_Monte_Carlo).  Out_R/Out_V are the radial grid of
 * Radius_for_distributions (Sorting_output_data.f90:1363-1435) chosen by the
 * caller.  Iterations are numbered first_iter .. first_iter+nmc-1 for the
 * random streams. */
int  trk_run_mc(trk_ctx *ctx, int64_t nmc, uint64_t seed, int64_t first_iter,
                const double *Out_R, const double *Out_V, int32_t NR, trk_tallies *out);

/* trk_run_mc without device allocation.  No counterpart in the reference:
 * this entry point is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as trk_run_mc, on a context whose
 * tables are built (trk_ctx_create and trk_build_tables allocate; this call
 * does not: no hipMalloc, hipFree or event creation).  Every device buffer of
 * the run -- particle lists, one tally replica, the per-iteration histograms
 * and the caller's Out_* arrays staged for a device-side finish -- is carved
 * from a static device array of TRK_NOALLOC_SCRATCH_BYTES that each GPU the
 * library runs on holds until the process ends.  The list capacity is the
 * largest power of two up to 2^22 whose layout fits (options.gen_capacity
 * fixes it); a run that does not fit returns TRK_ERR_CAPACITY before any
 * launch, and event traces (trk_set_trace) return TRK_ERR_STATE.  Every call
 * on one GPU uses that GPU's array and clears what it needs; calls that
 * overlap on one GPU (from several threads or contexts) race and must not. */
#define TRK_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int  trk_run_mc_noalloc(trk_ctx *ctx, int64_t nmc, uint64_t seed, int64_t first_iter,
                        const double *Out_R, const double *Out_V, int32_t NR, trk_tallies *out);

/* --- batches of impacts ---------------------------------------------------
 * The reference runs one ion on one target per program run, its MC
 * iterations spread over OpenMP threads (do_Monte_Carlo's iteration loop,
 * Monte_Carlo.f90:395-415).  A batch runs many impacts -- different ions,
 * energies and materials, one MC iteration each -- in one persistent launch.
 *
 * trk_ctx_create_ion: a context for another ion on the material of a built
 * context; it shares that context's material arrays and tables, and its
 * trk_build_tables computes only the ion's own tables (SHI_TotIMFP,
 * Cross_sections.f90:2452-2597, for its grid and initial energy).  The
 * material context must outlive it.  It is an ordinary context otherwise
 * (trk_run_mc works on it). */
int  trk_ctx_create_ion(const trk_ctx *material, const trk_ion *ion, trk_ctx **out);

typedef struct trk_batch trk_batch;

/* impacts: contexts with built tables, all of one precision and device;
 * those created on one material context (or that context itself) form a
 * group that shares its tables (at most 8 groups).  Impact i of a run is
 * iteration first_iter + i, so its events and tallies equal those of
 * trk_run_mc(impacts[i], 1, seed, first_iter + i, ...).  capacity:
 * particles of one kind a group's list holds per round (photons a quarter;
 * 0 = 2^22).  Out_R/Out_V: the radial grid of every impact. */
int  trk_batch_create(trk_ctx *const *impacts, int32_t n, const double *Out_R, const double *Out_V, int32_t NR,
                      int32_t capacity, trk_batch **out);
void trk_batch_destroy(trk_batch *b);
int  trk_batch_error(const trk_batch *b, char *buf, size_t n);
/* Run every impact once (the step sequence is a hipGraph captured on the
 * first run and replayed after).  outs[i] (may be NULL) receives impact i's
 * tallies as trk_run_mc would add them; the per-iteration normalised ones
 * (Out_Ee_vs_E, Out_Eh_vs_E, Out_theta, Out_theta_h, Out_diff_coeff) must
 * be NULL, and n_events is not filled (see trk_batch_timing). */
int  trk_batch_run(trk_batch *b, uint64_t seed, int64_t first_iter, trk_tallies *const *outs);
/* The tallies of the last trk_batch_run (or tuning run), added into outs as
 * trk_batch_run's outs would be: a caller that runs with outs = NULL keeps
 * the device-side work and fetches (over PCIe) only the runs it wants. */
int  trk_batch_fetch(trk_batch *b, trk_tallies *const *outs);
/* Event trace of the next runs, records as in trk_set_trace. */
int  trk_batch_set_trace(trk_batch *b, double *buf, int64_t cap_records, int64_t *n_out);
/* Autotuner: times the batch (seed, first_iter) with every candidate of
 * threads per workgroup (512, 640, 768) and events a particle may take per
 * round (256, 512, 1024), keeps the fastest and records each timing
 * (trk_batch_tune_log).  No counterpart in the reference. */
int  trk_batch_tune(trk_batch *b, uint64_t seed, int64_t first_iter);
int  trk_batch_config(const trk_batch *b, int32_t *block, int32_t *round_events, int64_t *workgroups,
                      int32_t *capacity, int32_t *groups);
int  trk_batch_tune_log(const trk_batch *b, int32_t *block, int32_t *round_events, double *ms, int32_t cap, int32_t *n);
/* Device time of the last run's graph (ms), its events by kind (as
 * trk_tallies.n_events) and the transport rounds it took. */
int  trk_batch_timing(const trk_batch *b, double *ms, int64_t *n_events, int32_t *rounds);

/* Timing of the last trk_run_mc: seconds from the first ion launch to the end
 * of the last generation of each batch (HIP events on the library's stream),
 * wall seconds of the whole call, and kernel launches.  The reference times
 * only the whole program (MAIN:320-330); this and the two calls below are
 * instrumentation with no counterpart there. */
int  trk_last_timing(const trk_ctx *ctx, double *transport_s, double *total_s, int64_t *launches);

/* Per-kernel device time: when on, every launch of the ion kernel (TRK_K_SHI)
 * and of the transport kernel (TRK_K_TRANSPORT) is bracketed by HIP events on
 * the library's stream and its duration summed.  Turning it on clears the sums. */
enum trk_kernel_id { TRK_K_SHI = 0, TRK_K_TRANSPORT = 1 };
int  trk_set_profiling(trk_ctx *ctx, int on);
int  trk_kernel_time(const trk_ctx *ctx, int kernel, double *ms_total, int64_t *launches);

/* Record the full event trace of iterations [first_iter, first_iter+nmc) into
 * caller buffers during the next trk_run_mc (replay checking).  Each record is
 * 12 doubles: {iteration, particle id (bits), event no, kind, time of the
 * event, E, E2, X, Y, Z, time of the next event, draws used so far}; E/E2 are
 * the ion's energy/charge, the electron's energy, the hole's kinetic/potential
 * energy or the absorbed photon's energy.  Records are in completion order. */
int  trk_set_trace(trk_ctx *ctx, double *buf, int64_t cap_records, int64_t *n_out);

/* --- helpers the Fortran bridge uses ------------------------------------- */

/* Deterministic elementary functions and Philox draws shared with the
 * instrumented host code (bit-identical to what the kernels compute). */
double trk_det_log(double x);
double trk_det_exp(double x);
double trk_det_sin(double x);
double trk_det_cos(double x);
double trk_det_acos(double x);
double trk_det_atan(double x);
double trk_det_pow(double x, double y);
double trk_rng_uniform(uint64_t seed, uint64_t particle_id, uint32_t iteration, uint32_t draw);
/* Physical constants as folded by the library (0 pi, 1 e, 2 m_e, 3 m_p, 4 hbar,
 * 5 alpha, 6 c, 7 a0), restating Universal_Constants.f90:28-60. */
double trk_constant(int which);

/* Decay-time rules of check_atomic_parameters (Dealing_with_EADL.f90:343-372)
 * applied in place to a caller's descriptor: Radiat = 1e23 fs for the valence
 * band (the last shell of atom 0, or Shl_num >= 63) and for every shell when
 * include_photons is 0; Auger = 1e23 fs for the valence band.  A deep shell
 * whose Auger time is <= 0 or > 1e30 fs (the reference's "read it from EADL"
 * marks), or any other time that is not positive and finite, is refused
 * (TRK_ERR_ARG, reason in why).  trk_ctx_create applies the same rules to its
 * copy of the descriptor. */
int  trk_check_atomic_parameters(trk_material *m, int32_t include_photons, char *why, size_t n);

/* Logistic transmission of the surface barrier, P(E) = 1/(1 + exp(gamma (E1 - E))),
 * fitted to the Eckart barrier of the given work function [eV], length [A] and
 * height [eV] (barrier_parameters, Monte_Carlo.f90:2094-2116).  TRK_ERR_ARG
 * when the parameters give no finite E1 and gamma. */
int  trk_barrier_parameters(double work_function, double bar_length, double bar_height,
                            double *E1, double *gamma);

/* Weight of an oscillator's delta peak in the Delta-CDF model
 * (kind_of_DR = 4): its k-sum integral above x_min [eV]
 * (define_alpha, Reading_files_and_parameters.f90:2199-2239). */
double trk_delta_alpha(double A, double E0, double Gamma, double x_min);

/* f-sum and k-sum rules of a set of Drude-Lorentz oscillators
 * (sumrules, Cross_sections.f90:728-749). */
int  trk_sumrules(const double *A, const double *E0, const double *Gamma, int32_t n,
                  double x_min, double Omega, double *ksum, double *fsum);
/* Squared plasma frequency for a density [1/m^3] and mass [kg]
 * (w_plasma, Cross_sections.f90:712-725); mass <= 0 means the electron mass. */
double trk_w_plasma(double density_m3, double mass_kg);

/* Derived VB arrays from a raw DOS (E, DOS) pair list
 * (reading_material_DOS, Reading_files_and_parameters.f90:2374-2450).
 * Output arrays must hold trk_dos_size() entries. */
int  trk_dos_size(const double *raw_E, int32_t n_raw, int32_t *M_out);
int  trk_prepare_dos(const double *raw_E, const double *raw_DOS, int32_t n_raw,
                     double N_VB_el, double At_Dens, double sum_Pers,
                     double *E, double *DOS, double *int_DOS, double *k, double *Eff_m,
                     double *k_inv, double *Eff_m_inv, int32_t M);

#ifdef __cplusplus
}
#endif
#endif /* TREKIS_GPU_H */
