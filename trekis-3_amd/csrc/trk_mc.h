// trk_mc.h -- device layouts of the event transport shared by the host
// driver (trk_mc.hip) and the two compilations of the event code
// (trk_mc_fp64.hip, trk_mc_mx.hip).
#pragma once

#include <hip/hip_runtime.h>
#include "trk_ctx.h"
#include "trk_delta.h"
#include "trk_beb.h"

namespace trk {

// ---------------------------------------------------------------- layouts
struct PList {            // SoA particle list
    double *E, *t0, *tn, *X, *Y, *Z, *L, *th, *ph, *Ehkin, *Mass;
    int *sh;
    uint64_t *id;
    uint32_t *nd, *ne, *iter;
};

struct Tally {            // fixed-point difference arrays, replicated R times
    int Nt, NR, nsh, R;
    unsigned long long *cnt_e, *E_e;               // [R][Nt][NR], [R][Nt][NR][2]
    unsigned long long *cnt_h, *E_h, *Ek_h;        // [R][Nt][NR][nsh] (x2 for energies)
    unsigned long long *cnt_p, *E_p;               // [R][Nt][NR]
    unsigned long long *Elat;                      // [R][Nt][NR][2] (not differenced)
};

// Per-iteration snapshot histograms for the tallies the reference normalises
// inside each iteration (Out_Ee_vs_E, Out_Eh_vs_E, Out_theta, Out_theta_h by
// that iteration's particle counts, Out_diff_coeff by its VB-hole count),
// difference arrays over time like the Tally arrays.  b = iteration - iter0.
struct Spec {
    int on, B, ND, iter0;
    int *nel, *nvb, *dcn;          // [B][Nt]
    int *eve;                      // [B][Nt][NR]   electron spectrum, bins of Out_R
    int *ehve;                     // [B][Nt][ND]   VB-hole spectrum, bins of the DOS grid
    int *th, *thh;                 // [B][Nt][180]  polar angle of electrons / moving holes
    unsigned long long *dcs;       // [B][Nt][2]    sum of v*L/3 of mobile holes, fixed point
};

// tables staged in each transport workgroup's LDS (see stage_tables)
constexpr int kStaged = 26;

// The ion of one impact: its parameters and its own tables (SHI_MFP, the
// fused SHI_path, diff_SHI_MFP at its initial energy and the mixed path's
// per-shell transfer constants).  A context has one; a batch keeps one per
// impact in device memory, read by the lanes that run an ion's collision.
struct IonDev {
    double mass, E0, Zeff0, fixed_zeff;   // [proton masses], [eV], charge at E0, fixed charge
    int Z, kind_zeff, nS, pad_;
    const double *ES, *LSsh, *shi_path, *lnES, *ln_LSsh, *ln_shi_path;
    const int *sd_n; const int64_t *sd_off; const double *sd_E, *sd_L, *sd_inv; const int *sd_nonmono;
    const int *sd_M; const double *sd_c; const int *sd_bk;
};

// Round state of one material group of a batch (device memory): particles
// of the current round by kind, the next round's allocation counters by list
// parity, the shared cursor and the parity of the current round's lists.
struct GroupCtl {
    int n_in[4];
    int out_cnt[2][4];
    int cursor;
    int cur;
    int overflow;
    int pad_;
};

// Control of a batch launch (device memory): the grid barrier, the round
// number, the end flag, and the error bits (barrier time-out).
struct BatchCtl {
    unsigned bar_count, bar_gen;
    int done, err;
    int round, n_groups;
    unsigned long long deadline_ticks;   // wall-clock ticks a barrier may wait
};
constexpr int kBatchErrBarrier = 1;
constexpr int kMaxGroups = 8;

struct McDev {
    MatDev m;
    // tables
    int nE; const double *E, *Lsh, *el_imfp; const int64_t *roff; const double *rhw, *rlam;
    int nEel; const double *Eel, *Lel; const int64_t *roff_el; const double *rhw_el, *rlam_el;
    int nH; const double *EH, *LHsh, *hole_imfp; const int64_t *roff_h; const double *rhw_h, *rlam_h;
    int nHel; const double *EHel, *LHel; const int64_t *roff_hel; const double *rhw_hel, *rlam_hel;
    // 1 per cumulative row whose lam never increases (counted search)
    const unsigned char *rmono, *rmono_h, *rmono_el, *rmono_hel;
    int nP; const double *EP, *LPsh, *phot_imfp;
    int nS; const double *ES, *LSsh, *shi_path;
    const int *sd_n; const int64_t *sd_off; const double *sd_E, *sd_L, *sd_inv; const int *sd_nonmono;
    // logs of grids and fused tables (interpolation without per-node logs)
    const double *lnE, *ln_el_imfp, *lnEel, *ln_Lel, *lnEH, *ln_hole_imfp, *lnEHel, *ln_LHel;
    // logs of the per-shell paths and of the cumulative transfer rows
    const double *ln_Lsh, *ln_LHsh;
    const double *ln_rhw, *ln_rlam, *ln_rhw_h, *ln_rlam_h, *ln_rhw_el, *ln_rlam_el, *ln_rhw_hel, *ln_rlam_hel;
    // run
    uint64_t seed;
    double T_last, Tim, cutoff_eff;
    int Nt; const double *tim_glob;
    double inv_dt;                      // 1/dt of a linear time grid, 0 otherwise
    int hints;                          // 1: the electron and hole energy grids, the time grid and Out_R are strictly increasing
    // tables copied into each transport workgroup's LDS (count and offset in
    // doubles; count 0 = read from global memory), see kStage* below
    int lds_n[kStaged], lds_off[kStaged];
    int NR; const double *R;
    int include_photons;
    // lists, by kind: electrons, holes, photons, ions
    PList in[4], out[4];
    int n_in[4];                        // particles of the current round
    int cap[4];                         // capacity of each list
    int *out_cnt;                       // [4] allocation counters of the next round
    int *cursor;                        // next item of the current round
    int *overflow;
    unsigned long long *counters;       // [16] event counters
    double *trace; unsigned long long *trace_n; long long trace_cap;
    Tally T;
    Spec S;
    // used by the mixed path only, and last, so that the offsets of the
    // fields above stay those the FP64 kernel was tuned with: per-shell
    // constants of shi_transfer (k_shi_consts), logs of the SHI grid, of
    // its per-shell paths and of the fused path
    const int *sd_M; const double *sd_c; const int *sd_bk;
    const double *lnES, *ln_LSsh, *ln_shi_path;
    // batch runs only: the impacts' ions (indexed by iteration - iter0), the
    // events a particle may take per round, this group's round state
    const IonDev *ions;
    const int *plane_of;        // tally plane of each impact in its group
    uint32_t iter0;
    int round_events;
    GroupCtl *gctl;
    // emission through the surface Z = 0 (em_W, the work function, > 0): the
    // barrier height, the E1 and gamma of its logistic transmission, and the
    // per-iteration tallies of the emitted electrons while S.on (indexed like
    // S): counts [B][Nt], kinetic energies [B][Nt][2] (fixed point) and
    // spectrum [B][Nt][NR] on the energy grid RE = Out_R/10, all difference
    // arrays over time
    double em_W, em_H, em_E1, em_gamma;
    const double *RE;
    int *em_n, *em_h;
    unsigned long long *em_e;
    // the oscillators' Delta-CDF weights (kind_of_DR = 4), indexed like m.oE0;
    // per shell, the electron and hole cross section (2: BEB) and the mean
    // kinetic energy of the shell's electrons (both read by the general
    // transport kernel only)
    const double *delta_alpha;
    const int *beb_kocs;
    const double *beb_Ek;
    // DSF elastic scattering (kind_of_EMFP = 2) of electrons [0] and holes [1]
    DsfDev dsf[2];
};

enum { EV_SHI = 0, EV_E_INEL, EV_E_EL, EV_H_INEL, EV_H_EL, EV_AUGER, EV_RAD, EV_PHOT };

// d.overflow bits: list capacities, the per-particle event limit, NaN state.
constexpr int kFlagPairs = 1, kFlagPhotons = 2, kFlagEvents = 4, kFlagNaN = 8;
// Events one particle may process in one launch.  A history needs a few
// thousand at most; the limit only guarantees that every wave ends.
constexpr uint32_t kMaxEvents = 1u << 22;

enum { K_EL = 0, K_HO = 1, K_PH = 2, K_ION = 3 };

enum { C_E_EL = 0, C_E_INEL, C_H_EL, C_H_INEL, C_DEEP, C_PHOT, C_ION, C_COUNT };

// A wave runs a category of events only when at least this many of its lanes
// want it (or when it is the most wanted one); the other lanes wait with
// their state untouched.  Rare, long branches (ionisation, decays) then run
// for many lanes at once instead of stalling the wave at almost every step.
constexpr int kBatchLanes = 16;
constexpr int kMaxWait = 16;

// Events one particle may take in one round.  A particle with more left is
// written to the next round's list; this bounds the tail of every launch by
// the cost of kRoundEvents events instead of the longest history.
constexpr uint32_t kRoundEvents = 512;

extern __shared__ double lds_tab[];

// Transport workgroup of each path: one per CU (the staged tables fill most
// of the LDS), 12 waves, 3 per SIMD (168 VGPRs).  At 16 waves the mixed
// path would have to spill to scratch, which costs more than the fourth
// wave gains.
constexpr int kTransportBlockFp64 = 768;
constexpr int kTransportBlockMx = 768;

// The event code is compiled twice, into namespaces fp64 (deterministic
// FP64, replays the oracle bit for bit) and mx (mixed precision); these are
// the kernels the host driver launches from each.
#define TRK_MC_KERNELS                                                          \
    template <bool kAll, bool kGeneral> __global__ void k_transport(McDev din); \
    __global__ void k_ion_init(McDev d, int nb, uint32_t first_iter);        \
    __global__ void k_shi_consts(IonDev iv, const double *Ip, int nsh, int *sd_M, double *sd_c); \
    __global__ void k_shi_buckets(IonDev iv, int nsh, const int *sd_M, const double *sd_c, int *sd_bk); \
    template <bool kAll> __global__ void k_batch(BatchCtl *ctl, int n_groups); \
    __global__ void k_ion_init_batch(int g, int nb, const int *idx);        \
    hipError_t batch_upload(const McDev *mc, int n_groups, hipStream_t st);
namespace fp64 { TRK_MC_KERNELS }
namespace mx { TRK_MC_KERNELS }
#undef TRK_MC_KERNELS

// The ion view of a context's own ion (host side).
IonDev ion_view_host(const trk_ctx *c);
hipError_t launch_shi_consts(const trk_ctx *c, const IonDev &iv, bool mixed, hipStream_t st);
// Transport descriptor of a context's tables (trk_mc.hip).
void fill_mcdev(const trk_ctx *c, McDev &d);
// LDS layout of the staged tables (trk_mc.hip); true when all are staged.
bool plan_lds(McDev &d, int Nt, int NR, int n_dos, int nsh, size_t limit_doubles, size_t *bytes);
// BEB shells or the Delta-CDF model: the transport kernel's general instantiation
bool uses_general_kernel(const HostMat &h);
// One tally plane converted and added into the caller's Out_* arrays (trk_mc.hip).
int finish_tallies_host(trk_ctx *c, const Tally &T, int plane_index, const double *Out_V, int NR, trk_tallies *out,
                        hipStream_t st);

// 128-bit tally words: (low, high) 64-bit halves of an exact integer
TRK_DEV __int128 fx_read(const unsigned long long *p)
{
    return (__int128)(long long)p[1] * ((__int128)1 << 32) + (__int128)p[0];
}

TRK_DEV void i128_store(unsigned long long *p, __int128 v)
{
    p[0] = (unsigned long long)v;
    p[1] = (unsigned long long)(v >> 64);
}

TRK_DEV __int128 i128_load(const unsigned long long *p)
{
    return (__int128)(((unsigned __int128)p[1] << 64) | (unsigned __int128)p[0]);
}

}  // namespace trk
