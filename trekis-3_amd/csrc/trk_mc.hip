// trk_mc.hip -- event transport of electrons, valence holes, deep holes and
// photons on the MI355X (replaces do_Monte_Carlo / Monte_Carlo_modelling,
// Monte_Carlo.f90:39-679, and the per-particle routines it calls).
//
// Design (MI355X-first, not the reference's loop):
//  * The reference runs one iteration as a global-time discrete-event
//    simulation and scans every particle for the next event
//    (Find_min_time_particle, MC:531-563), which is O(N) per event.  Histories
//    do not interact, so here every particle is transported independently:
//    a thread owns one particle and runs all of its events.  Each particle has
//    its own counter-based random stream, so the order in which the GPU
//    visits particles does not change any history.
//  * Particles are processed in generations.  Generation 0 is the ion of every
//    iteration of a batch (one thread per iteration); the electron/hole pairs
//    it creates form generation 1; the secondaries of generation g form g+1.
//    Children are appended with one wave-aggregated atomic per allocation, and
//    pairs are stored SoA in separate electron and hole lists so that a wave
//    holds particles of one kind.
//  * Snapshot tallies (Calculated_statistics, MC:881-1110) are not taken by
//    scanning all particles at every time node.  A particle's state is
//    constant between two of its events, so its contribution to the nodes
//    that fall in that interval is added to difference arrays over time
//    (one add at the first node, one subtraction after the last); a moving
//    particle adds one range per radial bin it crosses.  The host prefix-sums
//    the time axis once per run.  Adds are integer atomics on split fixed-point
//    accumulators (2^-40 eV resolution), which makes every tally independent
//    of the order of the adds and hence deterministic.
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "trk_ctx.h"

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

struct McDev {
    MatDev m;
    // tables
    int nE; const double *E, *Lsh, *el_imfp; const int64_t *roff; const double *rhw, *rlam;
    int nEel; const double *Eel, *Lel; const int64_t *roff_el; const double *rhw_el, *rlam_el;
    int nH; const double *EH, *LHsh, *hole_imfp; const int64_t *roff_h; const double *rhw_h, *rlam_h;
    int nHel; const double *EHel, *LHel; const int64_t *roff_hel; const double *rhw_hel, *rlam_hel;
    int nP; const double *EP, *LPsh, *phot_imfp;
    int nS; const double *ES, *LSsh, *shi_path;
    const int *sd_n; const int64_t *sd_off; const double *sd_E, *sd_L, *sd_inv; const int *sd_nonmono;
    // logs of grids and fused tables (interpolation without per-node logs)
    const double *lnE, *ln_el_imfp, *lnEel, *ln_Lel, *lnEH, *ln_hole_imfp, *lnEHel, *ln_LHel;
    // run
    uint64_t seed;
    double T_last, Tim, cutoff_eff;
    int Nt; const double *tim_glob;
    double inv_dt;                      // 1/dt of a linear time grid, 0 otherwise
    // tables copied into each transport workgroup's LDS (count and offset in
    // doubles; count 0 = read from global memory), see kStage* below
    int lds_n[24], lds_off[24];
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
};

enum { EV_SHI = 0, EV_E_INEL, EV_E_EL, EV_H_INEL, EV_H_EL, EV_AUGER, EV_RAD, EV_PHOT };

struct PState {
    double E, t0, tn, X, Y, Z, L, th, ph, Ehkin, Mass;
    int sh;
    uint64_t id;
    uint32_t nd, ne, iter;
    // electrons: total inelastic and elastic mean free paths at energy cE and
    // the elastic-grid bracket of cE, kept from the previous event's free-path
    // draw (the next event looks them up again at the same energy)
    double cE, cIM, cEM;
    int cN;
    // the particle's last Philox block (index + 1, 0 = none): draws 2k and
    // 2k+1 come from the same block, so every other draw is a lookup
    uint32_t cb, cw0, cw1, cw2, cw3;
};

// One Philox block of a particle's stream.
__device__ __forceinline__ void philox_block(uint32_t blk, uint64_t id, uint32_t iter, uint64_t seed, uint32_t *w)
{
    trk_u32x4 c;
    c.v[0] = blk; c.v[1] = (uint32_t)id; c.v[2] = (uint32_t)(id >> 32); c.v[3] = iter;
    trk_u32x4 r = trk_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    w[0] = r.v[0]; w[1] = r.v[1]; w[2] = r.v[2]; w[3] = r.v[3];
}

// Draw n of the particle's stream without consuming it; equals
// trk_uniform(seed, id, iter, n) bit for bit.
TRK_DEV double rn_at(const McDev &d, PState &p, uint32_t n)
{
    const uint32_t blk = n >> 1;
    if (p.cb != blk + 1u) {
        uint32_t w[4];
        philox_block(blk, p.id, p.iter, d.seed, w);
        p.cw0 = w[0]; p.cw1 = w[1]; p.cw2 = w[2]; p.cw3 = w[3];
        p.cb = blk + 1u;
    }
    const uint32_t a = (n & 1u) ? p.cw2 : p.cw0;
    const uint32_t b = (n & 1u) ? p.cw3 : p.cw1;
    const uint64_t u = ((((uint64_t)a) << 32) | (uint64_t)b) >> 11;
    return ((double)u + 0.5) * 0x1.0p-53;
}

TRK_DEV double rn(const McDev &d, PState &p) { return rn_at(d, p, p.nd++); }

TRK_DEV void load(const PList &l, int i, PState &p, int hole)
{
    p.E = l.E[i]; p.t0 = l.t0[i]; p.tn = l.tn[i]; p.X = l.X[i]; p.Y = l.Y[i]; p.Z = l.Z[i];
    p.L = l.L[i]; p.th = l.th[i]; p.ph = l.ph[i];
    if (hole) { p.Ehkin = l.Ehkin[i]; p.Mass = l.Mass[i]; p.sh = l.sh[i]; }
    else { p.Ehkin = 0.0; p.Mass = 1.0; p.sh = -1; }
    p.id = l.id[i]; p.nd = l.nd[i]; p.ne = l.ne[i]; p.iter = l.iter[i];
    p.cE = TRK_NAN; p.cIM = 0.0; p.cEM = 0.0; p.cN = 0;
    p.cb = 0u;
}

TRK_DEV void store(const PList &l, int i, const PState &p, int hole)
{
    l.E[i] = p.E; l.t0[i] = p.t0; l.tn[i] = p.tn; l.X[i] = p.X; l.Y[i] = p.Y; l.Z[i] = p.Z;
    l.L[i] = p.L; l.th[i] = p.th; l.ph[i] = p.ph;
    if (hole) { l.Ehkin[i] = p.Ehkin; l.Mass[i] = p.Mass; l.sh[i] = p.sh; }
    l.id[i] = p.id; l.nd[i] = p.nd; l.ne[i] = p.ne; l.iter[i] = p.iter;
}

// fresh slot values of the reference arrays (MC:1948-1951, Objects.f90:471-474)
TRK_DEV void init_child(PState &c, uint64_t parent, uint32_t ev, uint32_t slot, uint32_t iter, int hole)
{
    c.E = 0.0; c.t0 = 0.0; c.X = 0.0; c.Y = 0.0; c.Z = 0.0; c.L = 0.0; c.th = 0.0; c.ph = 0.0;
    if (hole) { c.tn = 1.0e21; c.L = 1.0e30; c.Mass = 1.0e30; c.Ehkin = 0.0; c.sh = -1; }
    else { c.tn = 1.0e20; c.Mass = 1.0; c.Ehkin = 0.0; c.sh = -1; }
    c.id = trk_child_id(parent, ev, slot);
    c.nd = 0; c.ne = 0; c.iter = iter;
    c.cb = 0u;
}

// ---------------------------------------------------------------- tallies
TRK_DEV void fx_add(unsigned long long *p, double v)
{
    long long x = (long long)__builtin_rint(v * 0x1.0p+40);
    unsigned long long lo = (unsigned long long)x & 0xFFFFFFFFull;
    long long hi = x >> 32;
    if (lo) atomicAdd(p, lo);
    if (hi) atomicAdd(p + 1, (unsigned long long)hi);
}

TRK_DEV void cnt_add(unsigned long long *p, long long v) { atomicAdd(p, (unsigned long long)v); }

// number of time nodes with tim_glob <= t
TRK_DEV int nodes_le(const McDev &d, double t)
{
    if (d.inv_dt > 0.0) {
        // linear grid: start next to t/dt and step to the exact count
        if (!(t >= d.tim_glob[0])) return 0;
        if (t >= d.tim_glob[d.Nt - 1]) return d.Nt;
        int i = (int)(t * d.inv_dt);
        if (i > d.Nt) i = d.Nt;
        if (i < 0) i = 0;
        while (i > 0 && d.tim_glob[i - 1] > t) --i;
        while (i < d.Nt && d.tim_glob[i] <= t) ++i;
        return i;
    }
    int lo = 0, hi = d.Nt;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (d.tim_glob[mid] <= t) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Range add over 0-based node rows [r0, r1] of one particle kind in radial bin j.
TRK_DEV void range_e(const McDev &d, int rep, int r0, int r1, int j, double E)
{
    const Tally &T = d.T;
    size_t base = ((size_t)rep * T.Nt) * T.NR;
    size_t a = base + (size_t)r0 * T.NR + j;
    cnt_add(T.cnt_e + a, 1);
    fx_add(T.E_e + 2 * a, E);
    if (r1 + 1 < T.Nt) {
        size_t b = base + (size_t)(r1 + 1) * T.NR + j;
        cnt_add(T.cnt_e + b, -1);
        fx_add(T.E_e + 2 * b, -E);
    }
}

TRK_DEV void range_p(const McDev &d, int rep, int r0, int r1, int j, double E)
{
    const Tally &T = d.T;
    size_t base = ((size_t)rep * T.Nt) * T.NR;
    size_t a = base + (size_t)r0 * T.NR + j;
    cnt_add(T.cnt_p + a, 1);
    fx_add(T.E_p + 2 * a, E);
    if (r1 + 1 < T.Nt) {
        size_t b = base + (size_t)(r1 + 1) * T.NR + j;
        cnt_add(T.cnt_p + b, -1);
        fx_add(T.E_p + 2 * b, -E);
    }
}

TRK_DEV void range_h(const McDev &d, int rep, int r0, int r1, int j, int s, double E, double Ek)
{
    const Tally &T = d.T;
    size_t base = ((size_t)rep * T.Nt) * T.NR;
    size_t a = (base + (size_t)r0 * T.NR + j) * T.nsh + s;
    cnt_add(T.cnt_h + a, 1);
    fx_add(T.E_h + 2 * a, E);
    fx_add(T.Ek_h + 2 * a, Ek);
    if (r1 + 1 < T.Nt) {
        size_t b = (base + (size_t)(r1 + 1) * T.NR + j) * T.nsh + s;
        cnt_add(T.cnt_h + b, -1);
        fx_add(T.E_h + 2 * b, -E);
        fx_add(T.Ek_h + 2 * b, -Ek);
    }
}

// Snapshot contribution of one state, valid from its creation time t0 until
// its next event tn (MC:998-1097): nodes i with t0 < tim_glob(i) <= tn.
// kind 0 electron, 1 hole, 2 photon.
TRK_DEV double vel_hole(const PState &h);

TRK_DEV void range_i(int *a, int stride, int Nt, int r0, int r1)
{
    atomicAdd(a + (size_t)r0 * stride, 1);
    if (r1 + 1 < Nt) atomicAdd(a + (size_t)(r1 + 1) * stride, -1);
}

// bin of a polar angle on the 1-degree grid 1..180 (find_mono_1d on it)
TRK_DEV int theta_bin(double th)
{
    double xx = (th / kPi) * 180.0;
    if (!(xx >= 1.0)) return 0;
    if (xx >= 180.0) return 179;
    int b = (int)xx;              // number of grid values <= xx
    return b < 179 ? b : 179;
}

// Calculated_statistics, MC:1022-1110: the per-iteration normalised tallies
TRK_DEV void spectra(const McDev &d, const PState &p, int kind, int r0, int r1)
{
    const Spec &S = d.S;
    const int Nt = d.Nt;
    const size_t b = (size_t)(p.iter - (uint32_t)S.iter0);
    if (kind == 0) {
        range_i(S.nel + b * Nt, 1, Nt, r0, r1);
        int j = find_mono_1d(d.R, d.NR, p.E) - 1;
        range_i(S.eve + (b * Nt) * d.NR + j, d.NR, Nt, r0, r1);
        if (p.E > 0.0) {
            double th0 = (p.E > d.cutoff_eff) ? p.th : 0.0;
            range_i(S.th + (b * Nt) * 180 + theta_bin(th0), 180, Nt, r0, r1);
        }
    } else if (kind == 1) {
        if (p.sh == d.m.vb) {
            range_i(S.nvb + b * Nt, 1, Nt, r0, r1);
            if (S.ND > 1) {
                int jj = find_mono_1d(d.m.dosE, d.m.n_dos, p.Ehkin) - 1;
                range_i(S.ehve + (b * Nt) * S.ND + jj, S.ND, Nt, r0, r1);
            }
        }
        if (p.Mass < 1.0e3 && p.Ehkin > d.cutoff_eff) range_i(S.thh + (b * Nt) * 180 + theta_bin(p.th), 180, Nt, r0, r1);
        if (p.Mass < 1.0e3 && p.L < 1.0e3) {
            range_i(S.dcn + b * Nt, 1, Nt, r0, r1);
            double v = (((1.0 / 3.0) * vel_hole(p)) * p.L) * 1.0e-6;
            fx_add(S.dcs + 2 * (b * Nt + r0), v);
            if (r1 + 1 < Nt) fx_add(S.dcs + 2 * (b * Nt + r1 + 1), -v);
        }
    }
}

TRK_DEV void segment(const McDev &d, const PState &p, int kind, int rep)
{
    int r0 = nodes_le(d, p.t0);                         // first node row with tim > t0
    int r1 = (p.tn < d.T_last) ? nodes_le(d, p.tn) - 1 : d.Nt - 1;
    if (r0 > r1) return;
    if (d.S.on) spectra(d, p, kind, r0, r1);
    double V = 0.0, st = 0.0, sp = 0.0, cp = 0.0;
    bool moving;
    if (kind == 0) {
        moving = p.E > d.cutoff_eff;
        if (moving) V = trk_sqrt(((2.0 * p.E) * kE) / kMe);
    } else if (kind == 1) {
        moving = (p.Mass < 1.0e3) && (p.Ehkin > d.cutoff_eff);
        if (moving) {
            if (p.Mass < 1.0e6 && !((p.Ehkin < -1.0e-6) || (p.Mass < 1.0e-10)) && !(trk_fabs(p.Ehkin) < 1.0e-6))
                V = trk_sqrt(((2.0 * p.Ehkin) * kE) / (p.Mass * kMe));
            else
                V = 0.0;
        }
    } else {
        if (!(p.E > 0.0)) return;
        moving = true;
        V = kC;
    }
    if (moving) { st = trk_sin(p.th); trk_sincos(p.ph, &sp, &cp); }
    else { st = trk_sin(0.0); trk_sincos(0.0, &sp, &cp); }
    int jprev = -1, rstart = r0;
    for (int r = r0; r <= r1; ++r) {
        double L0 = 0.0;
        if (moving) {
            L0 = ((V * (d.tim_glob[r] - p.t0)) * 1.0e-5);
            if (L0 < 0.0) L0 = 0.0;
        }
        double X = p.X + (L0 * st) * sp;
        double Y = p.Y + (L0 * st) * cp;
        double R = trk_sqrt(X * X + Y * Y);
        int j = find_mono_1d(d.R, d.NR, R) - 1;
        if (j < 0) j = 0;
        if (jprev >= 0 && j != jprev) {
            if (kind == 0) range_e(d, rep, rstart, r - 1, jprev, p.E);
            else if (kind == 1) range_h(d, rep, rstart, r - 1, jprev, p.sh, p.E, p.Ehkin);
            else range_p(d, rep, rstart, r - 1, jprev, p.E);
            rstart = r;
        }
        jprev = j;
        if (!moving) { r = r1; }
    }
    if (kind == 0) range_e(d, rep, rstart, r1, jprev, p.E);
    else if (kind == 1) range_h(d, rep, rstart, r1, jprev, p.sh, p.E, p.Ehkin);
    else range_p(d, rep, rstart, r1, jprev, p.E);
}

// Out_Elat: per elastic event in the grid_do pass of its time node (MC:2436, 2719)
TRK_DEV void tally_elat(const McDev &d, int rep, double t_ev, double X, double Y, double dE)
{
    int r = nodes_le(d, t_ev);
    if (r >= d.Nt) return;
    double R = trk_sqrt(X * X + Y * Y);
    int j = find_mono_1d(d.R, d.NR, R) - 1;
    if (j < 0) j = 0;
    const Tally &T = d.T;
    size_t a = (((size_t)rep * T.Nt) + r) * T.NR + j;
    fx_add(T.Elat + 2 * a, dE);
}

TRK_DEV void trace_ev(const McDev &d, int kind, const PState &p, uint32_t ev, double t_ev,
                      double E, double E2, double X, double Y, double Z)
{
    if (!d.trace) return;
    unsigned long long k = atomicAdd(d.trace_n, 1ull);
    if ((long long)k >= d.trace_cap) return;
    double *r = d.trace + 12 * k;
    r[0] = (double)p.iter;
    r[1] = __longlong_as_double((long long)p.id);
    r[2] = (double)ev; r[3] = (double)kind; r[4] = t_ev; r[5] = E; r[6] = E2;
    r[7] = X; r[8] = Y; r[9] = Z; r[10] = p.tn; r[11] = (double)p.nd;
}

// ---------------------------------------------------------------- lookups
// Next_free_path_2d (MC:1874-1899) and _1d (MC:1835-1871)
TRK_HD double nfp_2d(const double *Ea, const double *La, int n, double E)
{
    int N = find_mono_2d(Ea, n, E);
    if (N < 0) return 1.0e30;
    if (N == 1) {
        double M = interpolate(1, Ea[0], Ea[1], La[0], La[1], E);
        return (M < La[0]) ? La[0] : M;
    }
    int Nl = N - 1;
    if (La[Nl - 1] >= 1.0e16) return La[Nl - 1];
    return interpolate(5, Ea[Nl - 1], Ea[N - 1], La[Nl - 1], La[N - 1], E);
}

TRK_HD double nfp_1d(const double *Ea, const double *La, int n, double E)
{
    int N = find_mono_1d(Ea, n, E);
    if (N < 0) return 1.0e30;
    if (N == 1) {
        double M = interpolate(1, Ea[0], Ea[1], La[0], La[1], E);
        return (M < Ea[0]) ? Ea[0] : M;
    }
    int Nl = N - 1;
    if (La[Nl - 1] >= 1.0e16) return La[Nl - 1];
    return interpolate(5, Ea[Nl - 1], Ea[N - 1], La[Nl - 1], La[N - 1], E);
}

// Interpolate flag 5 with the logs of the nodes, the values and the argument
// given: the same operations on the same numbers as interpolate(5, ...).
TRK_HD double interp5_ln(double E1, double E2, double S1, double S2, double E, double lE1, double lE2,
                         double lS1, double lS2, double lE)
{
    if (trk_fabs(E2 - E1) < 1.0e-6) return (S1 > S2) ? S1 : S2;
    if (E == E1) return S1;
    double o = lS1 + ((lS2 - lS1) / (lE2 - lE1)) * (lE - lE1);
    return trk_exp(o);
}

// nfp_2d / nfp_1d on a table with precomputed logs (lEa, lLa); lE = trk_log(E).
// *Nout receives the bracket index of the search.
TRK_HD double nfp_2d_ln(const double *Ea, const double *La, const double *lEa, const double *lLa, int n, double E,
                        double lE, int *Nout)
{
    int N = find_mono_2d(Ea, n, E);
    *Nout = N;
    if (N == 1) {
        double M = interpolate(1, Ea[0], Ea[1], La[0], La[1], E);
        return (M < La[0]) ? La[0] : M;
    }
    int Nl = N - 1;
    if (La[Nl - 1] >= 1.0e16) return La[Nl - 1];
    return interp5_ln(Ea[Nl - 1], Ea[N - 1], La[Nl - 1], La[N - 1], E, lEa[Nl - 1], lEa[N - 1], lLa[Nl - 1],
                      lLa[N - 1], lE);
}

TRK_HD double nfp_1d_ln(const double *Ea, const double *La, const double *lEa, const double *lLa, int n, double E,
                        double lE)
{
    int N = find_mono_1d(Ea, n, E);
    if (N == 1) {
        double M = interpolate(1, Ea[0], Ea[1], La[0], La[1], E);
        return (M < Ea[0]) ? Ea[0] : M;
    }
    int Nl = N - 1;
    if (La[Nl - 1] >= 1.0e16) return La[Nl - 1];
    return interp5_ln(Ea[Nl - 1], Ea[N - 1], La[Nl - 1], La[N - 1], E, lEa[Nl - 1], lEa[N - 1], lLa[Nl - 1],
                      lLa[N - 1], lE);
}

// Which_shell, MC:1786-1832 (per-shell lambda on a shared grid)
TRK_DEV double shell_mfp(const double *Eg, const double *Lsh, int n, int N, int s, double E)
{
    if (N == 1) return 1.0e20;
    const double *L = Lsh + (size_t)s * n;
    if ((L[N - 2] == L[N - 1]) || (L[N - 2] > 1.0e20)) return L[N - 2];
    return interpolate(5, Eg[N - 2], Eg[N - 1], L[N - 2], L[N - 1], E);
}

TRK_DEV int which_shell(const McDev &d, const double *Eg, const double *Lsh, int n, double E, PState &p)
{
    const int nsh = d.m.nsh;
    int N = find_mono_1d(Eg, n, E);
    if (N < 0) N = 1;
    double tot = 0.0;
    if (nsh <= 8) {
        double tmp[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            tmp[s] = 0.0;
            if (s < nsh) {
                tmp[s] = 1.0 / shell_mfp(Eg, Lsh, n, N, s, E);
                tot = tot + tmp[s];
            }
        }
        tot = rn(d, p) * tot;
        double sum = 0.0;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (s < nsh) {
                sum = sum + tmp[s];
                if (sum >= tot) return s;
            }
        }
        return nsh - 1;
    }
    for (int s = 0; s < nsh; ++s) tot = tot + 1.0 / shell_mfp(Eg, Lsh, n, N, s, E);
    tot = rn(d, p) * tot;
    double sum = 0.0;
    for (int s = 0; s < nsh; ++s) {
        sum = sum + 1.0 / shell_mfp(Eg, Lsh, n, N, s, E);
        if (sum >= tot) return s;
    }
    return nsh - 1;
}

// interpolate_transferred_energy, Cross_sections.f90:1968-2045
TRK_DEV double interp_transfer(const double *Eg, int n, const int64_t *roff, const double *hwa,
                               const double *lama, int shell, double Ele, double L_need)
{
    int iE = find_mono_1d(Eg, n, Ele);
    if (iE < 0) return TRK_NAN;
    if (iE > 1 && trk_fabs(Eg[iE - 2] - Ele) < 1.0e-6) iE = iE - 1;
    size_t slot = (size_t)shell * n + (iE - 1);
    int64_t o = roff[slot];
    int len = (int)(roff[slot + 1] - o);
    if (len < 1) return TRK_NAN;
    const double *hw = hwa + o, *lam = lama + o;
    int ih = find_decr(lam, len, L_need);
    if (ih < 0) return TRK_NAN;
    double hw1 = ((ih == 1) || (ih == len)) ? hw[ih - 1] : interpolate(5, lam[ih - 1], lam[ih], hw[ih - 1], hw[ih], L_need);
    double out = hw1;
    if (iE > 1) {
        iE = iE - 1;
        slot = (size_t)shell * n + (iE - 1);
        o = roff[slot];
        len = (int)(roff[slot + 1] - o);
        if (len < 1) return TRK_NAN;
        hw = hwa + o;
        lam = lama + o;
        ih = find_decr(lam, len, L_need);
        double hw2 = ((ih == 1) || (ih == len)) ? hw[ih - 1] : interpolate(5, lam[ih - 1], lam[ih], hw[ih - 1], hw[ih], L_need);
        if ((hw1 < 1.0e-10) || (hw2 < 1.0e-10)) out = interpolate(1, Eg[iE - 1], Eg[iE], hw1, hw2, Ele);
        else out = interpolate(5, Eg[iE - 1], Eg[iE], hw1, hw2, Ele);
    }
    return out;
}

TRK_DEV double hole_mass_tab(const McDev &d, double Ele)
{
    if (d.m.hole_mass >= 0.0) return d.m.hole_mass;
    return d.m.dosM[find_mono_1d(d.m.dosE, d.m.n_dos, Ele) - 1];
}

// Electron_energy_transfer_inelastic, Cross_sections.f90:1793-1871
TRK_DEV double transfer_inel(const McDev &d, double Ele, int shell, double L_tot, int hole, PState &p)
{
    double RN = rn(d, p);
    double L_need = L_tot / RN;
    double Emin = d.m.Ip[shell];
    double Egap = d.m.Egap;
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    double Emax, Mass;
    if (!hole) {
        Emax = (Ele + Emin) / 2.0;
    } else {
        Mass = hole_mass_tab(d, Ele);
        Emax = ((4.0 * Ele) * Mass) / ((Mass + 1.0) * (Mass + 1.0));
    }
    if (d.m.plasmon_Emax && Emin == Egap) {
        double Ep = trk_sqrt((((((d.m.N_VB_el * d.m.At_Dens) * 1.0e6) * kH) * kH) / (kMe * kE0)) + (Egap * Egap));
        if (Ep >= Emax) Emax = Ep;
        if (Ele < Emax) Emax = Ele;
    }
    double E;
    if (!hole) E = interp_transfer(d.E, d.nE, d.roff, d.rhw, d.rlam, shell, Ele, L_need);
    else E = interp_transfer(d.EH, d.nH, d.roff_h, d.rhw_h, d.rlam_h, d.m.vb, Ele, L_need);
    if (E < Emin) E = Emin;
    if (E > Emax) E = Emax;
    if (E != E) E = Emin;
    return E;
}

// Electron_energy_transfer_elastic, Cross_sections.f90:2285-2428 (tabulated)
TRK_DEV double transfer_el(const McDev &d, double Ele, double L_tot, int hole, PState &p)
{
    double RN = rn(d, p);
    double L_need = L_tot / RN;
    double hw = hole ? interp_transfer(d.EHel, d.nHel, d.roff_hel, d.rhw_hel, d.rlam_hel, 0, Ele, L_need)
                     : interp_transfer(d.Eel, d.nEel, d.roff_el, d.rhw_el, d.rlam_el, 0, Ele, L_need);
    if (hw >= Ele) hw = Ele;
    return hw;
}

TRK_DEV double rest_energy(double M0) { return ((M0 * kC) * kC) / kE; }

// NRG_transfer_elastic_atomic with Mott_sample_mu, screening_parameter and
// transfered_E_from_theta, Cross_sections.f90:3517-3613, for every atom kind
// of the target weighted by its share (MC:2395-2406).  The terms that depend
// only on the electron energy are formed once, the per-atom constants Z^(2/3)
// and the atom's rest energy come precomputed; the arithmetic of each term is
// unchanged.
TRK_DEV double mott_transfer_all(const McDev &d, double Ee, double M_eff, PState &p)
{
    const double Erest = rest_energy(kMe);
    const double fact = Ee / Erest + 1.0;
    const double v = kC * trk_sqrt(1.0 - 1.0 / (fact * fact));
    const double beta = v / kC;
    const double beta2 = beta * beta;
    const double tau = Ee / Erest;
    const double A1 = 1.0 - beta2;
    const double A2 = ((3.76 * kAlpha) * kAlpha) / beta2;
    const double S = trk_sqrt(tau / (1.0 + tau));
    const double mc2 = rest_energy(kMe * M_eff);
    const double Emc = Ee + mc2;
    const double E2mc = Ee + 2.0 * mc2;
    double dE = 0.0;
    for (int a = 0; a < d.m.n_atoms; ++a) {
        const double Z = d.m.mott_Z[a];
        const double Mct2 = d.m.mott_Mct2[a];
        double RN = rn(d, p);
        double theta;
        if (v < 1.0e-6) {
            theta = 0.0;
        } else {
            double nu = (((1.7e-5 * d.m.mott_Z23[a]) * A1) / beta2) * (1.13 + ((A2 * Z) * Z) * S);
            double mu = (RN * ((2.0 * nu) + 1.0) - nu) / (RN + nu);
            theta = trk_acos(mu);
        }
        double ct = trk_cos(theta);
        double ct2 = ct * ct;
        double st2 = 1.0 - ct2;
        double EmcMc = Emc + Mct2;
        double W1 = ((Emc * st2) + Mct2) - (ct * trk_sqrt((Mct2 * Mct2) - ((mc2 * mc2) * st2)));
        double W2 = (Ee * E2mc) / ((EmcMc * EmcMc) - ((Ee * E2mc) * ct2));
        dE = dE + (W1 * W2) * d.m.mott_Pers[a];
    }
    return dE / d.m.sumPers;
}

TRK_DEV double cos_theta_W(const McDev &d, double E, double W, double M_in, PState &p)
{
    double Ein = rest_energy(M_in);
    double Et = rest_energy(d.m.Mtarget);
    double E2mc = E + 2.0 * Ein;
    double EmW = E - W;
    double W1 = (E * E2mc) - (W * ((E + Ein) + Et));
    double W2 = ((E * E2mc) * EmW) * (E2mc - W);
    double mu = (W2 > 0.0) ? W1 / trk_sqrt(W2) : 0.0;
    if (trk_fabs(mu) > 1.0) mu = trk_cos(kPi * rn(d, p));
    return mu;
}

// New_Angles_both, MC:1328-1360
// ct0 = trk_cos(theta0), computed by the caller for the flight before.
TRK_DEV void new_angles(double phi0, double theta0, double ct0, double theta, double psi, double &phi1, double &theta1)
{
    double sps, cps;
    trk_sincos(psi, &sps, &cps);
    double ph = phi0 + ((theta * ct0) * sps);
    double t = theta0 + (theta * cps);
    while (t < 0.0) { t = trk_fabs(t); ph = ph + kPi; }
    while (t > kPi) { t = (2.0 * kPi) - t; ph = ph - kPi; }
    if (ph > 2.0 * kPi) ph = ph - (((double)floor(ph / (2.0 * kPi))) * 2.0) * kPi;
    if (ph < 0.0) ph = ph + (((double)ceil(trk_fabs(ph) / (2.0 * kPi))) * 2.0) * kPi;
    phi1 = ph;
    theta1 = t;
}

// integrated-DOS level sampling shared by Electron_recieves_E (MC:1653-1716)
// and From_where_in_VB (MC:1522-1573)
TRK_DEV double dos_level(const McDev &d, double Sum, PState &p)
{
    const double *E = d.m.dosE, *I = d.m.dosI;
    const int N = d.m.n_dos;
    double RN = rn(d, p);
    double Tot = RN * Sum;
    int Nt = find_mono_1d(I, N, Tot);
    if (Nt > 1) return E[Nt - 2] + ((E[Nt - 1] - E[Nt - 2]) * (Tot - I[Nt - 2])) / (I[Nt - 1] - I[Nt - 2]);
    return E[Nt - 1];
}

TRK_DEV double receives_E(const McDev &d, double dE, int shell, PState &p)
{
    double E = dE - d.m.Ip[shell];
    if (shell != d.m.vb) return E;
    if (d.m.n_dos <= 0 || dE <= d.m.Ip[shell]) return E;
    const int N = d.m.n_dos;
    int M = (E < d.m.dosE[N - 1]) ? find_mono_1d(d.m.dosE, N, E) : N + 1;
    if (M > 1) return E - dos_level(d, d.m.dosI[M - 2], p);
    return E;
}

TRK_DEV double from_where_in_vb(const McDev &d, int has_E, double Elim, PState &p)
{
    if (d.m.n_dos <= 0) return 0.0;
    const int N = d.m.n_dos;
    if (!has_E) return dos_level(d, d.m.dosI[N - 1], p);
    int M = (Elim < d.m.dosE[N - 1]) ? find_mono_1d(d.m.dosE, N, Elim) : N + 1;
    if (M > 1) return dos_level(d, d.m.dosI[M - 2], p);
    return 0.0;
}

// count_for_Auger_shells / Choose_for_Auger_shell, MC:1576-1647
TRK_DEV double auger_count(const McDev &d, double NRG, int second)
{
    double c = 0.0;
    for (int s = 0; s < d.m.nsh; ++s) {
        double Ed = second ? 1.0e10 : (NRG - d.m.Ip[s]);
        if ((NRG > d.m.Ip[s] + 1.0e-3) && (Ed >= d.m.Egap)) c = c + d.m.Nel[s];
    }
    return c;
}

TRK_DEV int auger_choose(const McDev &d, double NRG, double Shel, int second, int dflt)
{
    double cs = 0.0;
    int found = dflt;
    for (int s = 0; s < d.m.nsh; ++s) {
        double Ed = second ? 1.0e10 : (NRG - d.m.Ip[s]);
        if ((NRG > d.m.Ip[s] + 1.0e-3) && (Ed >= d.m.Egap)) {
            cs = cs + d.m.Nel[s];
            if (cs >= Shel) found = s;
        }
        if (cs >= Shel) break;
    }
    return found;
}

// ---------------------------------------------------------------- kinematics
TRK_DEV double vel_el(double E) { return trk_sqrt(((2.0 * E) * kE) / kMe); }

TRK_DEV double vel_hole(const PState &h)
{
    if (h.Mass < 1.0e6) {
        if ((h.Ehkin < -1.0e-6) || (h.Mass < 1.0e-10)) return 0.0;
        if (trk_fabs(h.Ehkin) < 1.0e-6) return 0.0;
        return trk_sqrt(((2.0 * h.Ehkin) * kE) / (h.Mass * kMe));
    }
    return 0.0;
}

TRK_DEV double t_next(double t0, double MFP, double V) { return (V > 1.0e-10) ? t0 + (MFP / V) * 1.0e5 : 1.0e25; }
TRK_DEV void cut_el(const McDev &d, PState &e) { if (e.E < d.m.cut_off) e.tn = 1.0e20; }
TRK_DEV void cut_ho(const McDev &d, PState &h) { if ((h.Mass < 1.0e15) && (h.Ehkin < d.m.cut_off)) h.tn = 1.0e20; }

// Hole_parameters + Assign_holes_mass + Get_time_of_next_event, MC:724-848;
// draws come from the stream of the particle that owns the event (o).
TRK_DEV void hole_params(const McDev &d, PState &h, double Eh, PState &o)
{
    if (h.sh == d.m.vb) {
        double Etemp = h.Ehkin;
        h.Ehkin = Eh - d.m.Egap;
        h.E = d.m.Egap;
        if (d.m.hole_mass > 0.0) h.Mass = d.m.hole_mass;
        else h.Mass = d.m.dosM[find_mono_1d(d.m.dosE, d.m.n_dos, h.Ehkin) - 1];
        if (h.Mass < 1.0e3) {
            double HI = nfp_2d(d.EH, d.hole_imfp, d.nH, h.Ehkin);
            double HE = (Etemp == (Eh - d.m.Egap)) ? 1.0e30 : nfp_2d(d.EHel, d.LHel, d.nHel, h.Ehkin);
            double RN = rn(d, o);
            double MFP = -trk_log(RN) / (1.0 / HI + 1.0 / HE);
            h.tn = t_next(h.t0, MFP, vel_hole(h));
            h.L = MFP;
        } else {
            h.L = 1.0e30;
            h.tn = 1.0e30;
        }
    } else {
        h.Mass = 1.0e29;
        double RN = rn(d, o);
        double nu = 1.0 / d.m.Auger[h.sh] + 1.0 / d.m.Radiat[h.sh];
        h.tn = h.t0 - trk_log(RN) / nu;
        h.L = 1.0e30;
        h.E = Eh;
        h.Ehkin = 0.0;
    }
}

TRK_DEV double el_mfp(const McDev &d, double E, double RN)
{
    double lE = trk_log(E);
    int N;
    double IM = nfp_2d_ln(d.E, d.el_imfp, d.lnE, d.ln_el_imfp, d.nE, E, lE, &N);
    double EM = nfp_2d_ln(d.Eel, d.Lel, d.lnEel, d.ln_Lel, d.nEel, E, lE, &N);
    return -trk_log(RN) / (1.0 / IM + 1.0 / EM);
}

// el_mfp for the electron's own next flight: also keeps both mean free paths
// and the elastic bracket for the electron's next event at this energy.
TRK_DEV double el_mfp_keep(const McDev &d, double E, double RN, PState &e)
{
    double lE = trk_log(E);
    int N;
    double IM = nfp_2d_ln(d.E, d.el_imfp, d.lnE, d.ln_el_imfp, d.nE, E, lE, &N);
    double EM = nfp_2d_ln(d.Eel, d.Lel, d.lnEel, d.ln_Lel, d.nEel, E, lE, &N);
    e.cE = E; e.cIM = IM; e.cEM = EM; e.cN = N;
    return -trk_log(RN) / (1.0 / IM + 1.0 / EM);
}

// ---------------------------------------------------------------- child allocation
// d.overflow bits: list capacities, the per-particle event limit, NaN state.
constexpr int kFlagPairs = 1, kFlagPhotons = 2, kFlagEvents = 4, kFlagNaN = 8;
// Events one particle may process in one launch.  A history needs a few
// thousand at most; the limit only guarantees that every wave ends.
constexpr uint32_t kMaxEvents = 1u << 22;

// Slot allocation aggregated over the active lanes of a wave: one atomic per
// wave.  Every active lane must pass the same counter.
TRK_DEV int wave_alloc(int *counter)
{
    unsigned long long mask = __ballot(1);
    int lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    int rank = __popcll(mask & ((1ull << lane) - 1ull));
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    return base + rank;
}

enum { K_EL = 0, K_HO = 1, K_PH = 2, K_ION = 3 };

// Append one particle to the next round's list of its kind.
TRK_DEV void emit(const McDev &d, int kind, const PState &q)
{
    int k = wave_alloc(d.out_cnt + kind);
    if (k >= d.cap[kind]) { atomicOr(d.overflow, kind == K_PH ? kFlagPhotons : kFlagPairs); return; }
    store(d.out[kind], k, q, kind == K_HO || kind == K_ION);
}

TRK_DEV void emit_pair(const McDev &d, const PState &e, const PState &h)
{
    emit(d, K_EL, e);
    emit(d, K_HO, h);
}

TRK_DEV void emit_photon(const McDev &d, const PState &q) { emit(d, K_PH, q); }

// ---------------------------------------------------------------- events
// Electron_Monte_Carlo, MC:2253-2474
TRK_DEV int electron_event(const McDev &d, PState &e, int rep)
{
    uint32_t ev = e.ne++;
    double t_ev = e.tn;
    double Eel = e.E;
    double IMFP, EMFP;
    int Nel;
    if (Eel == e.cE) {
        IMFP = e.cIM; EMFP = e.cEM; Nel = e.cN;
    } else {
        double lE = trk_log(Eel);
        IMFP = nfp_2d_ln(d.E, d.el_imfp, d.lnE, d.ln_el_imfp, d.nE, Eel, lE, &Nel);
        EMFP = nfp_2d_ln(d.Eel, d.Lel, d.lnEel, d.ln_Lel, d.nEel, Eel, lE, &Nel);
    }
    double RN = rn(d, e);
    double L = e.L, th0 = e.th, ph0 = e.ph;
    double sth0, cth0, sph0, cph0;
    trk_sincos(th0, &sth0, &cth0);
    trk_sincos(ph0, &sph0, &cph0);
    double X = e.X + (L * sth0) * sph0;
    double Y = e.Y + (L * sth0) * cph0;
    double Z = e.Z + L * cth0;
    double dE, theta, phi;
    int kind;
    if (RN * (1.0 / IMFP + 1.0 / EMFP) < 1.0 / IMFP) {
        kind = EV_E_INEL;
        int sh = which_shell(d, d.E, d.Lsh, d.nE, Eel, e);
        double IMs = nfp_1d(d.E, d.Lsh + (size_t)sh * d.nE, d.nE, Eel);
        dE = transfer_inel(d, Eel, sh, IMs, 0, e);
        theta = trk_acos((Eel - dE) / trk_sqrt(Eel * (Eel - dE)));
        if (theta != theta) theta = rn(d, e) * kPi;
        phi = (2.0 * kPi) * rn(d, e);
        double dE_cur = receives_E(d, dE, sh, e);
        double MFP = el_mfp(d, dE_cur, rn(d, e));
        double phi1, theta1;
        new_angles(ph0, th0, cth0, kPi / 2.0 - theta, phi + kPi, phi1, theta1);
        PState ne, nh;
        init_child(ne, e.id, ev, 0, e.iter, 0);
        init_child(nh, e.id, ev, 1, e.iter, 1);
        ne.E = dE_cur; ne.t0 = e.tn; ne.X = X; ne.Y = Y; ne.Z = Z; ne.L = MFP; ne.th = theta1; ne.ph = phi1;
        ne.tn = t_next(ne.t0, MFP, vel_el(ne.E));
        cut_el(d, ne);
        double hth = kPi * rn(d, e);
        double hph = (2.0 * kPi) * rn(d, e);
        nh.t0 = e.tn; nh.X = X; nh.Y = Y; nh.Z = Z; nh.sh = sh; nh.th = hth; nh.ph = hph;
        hole_params(d, nh, dE - dE_cur, e);
        cut_ho(d, nh);
        emit_pair(d, ne, nh);
    } else {
        kind = EV_E_EL;
        // nfp_1d brackets like nfp_2d unless the energy sits on a node or
        // outside the grid; only then is the lookup repeated
        double EMt;
        if (Nel > 1 && d.Eel[Nel - 2] < Eel && Eel < d.Eel[Nel - 1]) EMt = EMFP;
        else EMt = nfp_1d_ln(d.Eel, d.Lel, d.lnEel, d.ln_Lel, d.nEel, Eel, trk_log(Eel));
        if (d.m.kind_of_EMFP == 1) {
            dE = transfer_el(d, Eel, EMt, 0, e);
        } else {
            dE = mott_transfer_all(d, Eel, 1.0, e);
        }
        double mu = cos_theta_W(d, Eel, dE, 1.0 * kMe, e);
        theta = trk_acos(mu);
        phi = (2.0 * kPi) * rn(d, e);
        tally_elat(d, rep, t_ev, X, Y, dE);
    }
    double MFP = el_mfp_keep(d, Eel - dE, rn(d, e), e);
    double phi1, theta1;
    new_angles(ph0, th0, cth0, theta, phi, phi1, theta1);
    e.E = Eel - dE; e.t0 = e.tn; e.X = X; e.Y = Y; e.Z = Z; e.L = MFP; e.th = theta1; e.ph = phi1;
    e.tn = t_next(e.t0, MFP, vel_el(e.E));
    cut_el(d, e);
    trace_ev(d, kind, e, ev, t_ev, e.E, 0.0, X, Y, Z);
    return kind;
}

// check_hole_parameters, MC:682-721 (DOS walks clamped to the array)
TRK_DEV double check_hole(const McDev &d, double Eel, double &dE, PState *el)
{
    double Ehole = Eel - dE;
    int m = find_mono_1d(d.m.dosE, d.m.n_dos, Ehole);
    if (d.m.dos[m - 1] < 1.0e-4) {
        if (el) {
            while (m > 1 && d.m.dos[m - 1] < 1.0e-4) m = m - 1;
            double Eloc = d.m.dosE[m - 1];
            el->E = el->E + (Ehole - Eloc);
            Ehole = Eloc;
        } else {
            while (m < d.m.n_dos && d.m.dos[m - 1] < 1.0e-4) m = m + 1;
            double Eloc = d.m.dosE[m - 1];
            dE = dE + (Ehole - Eloc);
            Ehole = Eloc;
        }
    }
    return Ehole;
}

// Hole_Monte_Carlo, MC:2517-2869
TRK_DEV int hole_event(const McDev &d, PState &h, int rep)
{
    uint32_t ev = h.ne++;
    double t_ev = h.tn;
    const double Egap = d.m.Egap;
    int kind;
    if (h.sh == d.m.vb) {
        double Eel = h.Ehkin;
        double HI, HE;
        if (Eel == h.cE) {
            HI = h.cIM; HE = h.cEM;
        } else {
            double lE = trk_log(Eel);
            int N;
            HI = nfp_2d_ln(d.EH, d.hole_imfp, d.lnEH, d.ln_hole_imfp, d.nH, Eel, lE, &N);
            HE = nfp_2d_ln(d.EHel, d.LHel, d.lnEHel, d.ln_LHel, d.nHel, Eel, lE, &N);
        }
        double RN = rn(d, h);
        double L = h.L, th0 = h.th, ph0 = h.ph;
        double sth0, cth0, sph0, cph0;
        trk_sincos(th0, &sth0, &cth0);
        trk_sincos(ph0, &sph0, &cph0);
        double X = h.X + (L * sth0) * sph0;
        double Y = h.Y + (L * sth0) * cph0;
        double Z = h.Z + L * cth0;
        double dE, Ehole, hth1, hph1;
        if ((RN * (1.0 / HI + 1.0 / HE) < 1.0 / HI) && (HI < 1.0e15)) {
            kind = EV_H_INEL;
            int sh = which_shell(d, d.EH, d.LHsh, d.nH, Eel, h);
            double HIs = nfp_1d(d.EH, d.LHsh + (size_t)sh * d.nH, d.nH, Eel);
            dE = transfer_inel(d, Eel, sh, HIs, 1, h);
            double Mh = h.Mass * kMe;
            double E11 = Eel - dE;
            double hth = trk_acos(trk_sqrt(((((Mh + kMe) * (Mh + kMe)) / ((4.0 * Mh) * kMe)) * dE) / Eel));
            double hph = (2.0 * kPi) * rn(d, h);
            if (hth != hth) hth = kPi * rn(d, h);
            hth1 = trk_acos(((Eel * (Mh - kMe)) + (E11 * (Mh + kMe))) / ((2.0 * Mh) * trk_sqrt(Eel * E11)));
            hph1 = hph + kPi;
            if (hth1 != hth1) hth1 = kPi * rn(d, h);
            double dE_cur = receives_E(d, dE, sh, h);
            double MFP = el_mfp(d, dE_cur, rn(d, h));
            (void)rn(d, h);  // drawn and discarded, MC:2611
            double phi1, theta1;
            new_angles(ph0, th0, cth0, hth, hph, phi1, theta1);
            PState ne, nh;
            init_child(ne, h.id, ev, 0, h.iter, 0);
            init_child(nh, h.id, ev, 1, h.iter, 1);
            ne.E = dE_cur; ne.t0 = h.tn; ne.X = X; ne.Y = Y; ne.Z = Z; ne.L = MFP; ne.th = theta1; ne.ph = phi1;
            ne.tn = t_next(ne.t0, MFP, vel_el(ne.E));
            cut_el(d, ne);
            double nth = kPi * rn(d, h);
            double nph = (2.0 * kPi) * rn(d, h);
            nh.t0 = h.tn; nh.X = X; nh.Y = Y; nh.Z = Z; nh.sh = sh; nh.th = nth; nh.ph = nph;
            hole_params(d, nh, dE - dE_cur, h);
            cut_ho(d, nh);
            Ehole = check_hole(d, Eel, dE, &ne);
            emit_pair(d, ne, nh);
        } else {
            kind = EV_H_EL;
            double HEt = nfp_1d(d.EHel, d.LHel, d.nHel, Eel);
            if (d.m.kind_of_EMFP == 1) {
                dE = transfer_el(d, Eel, HEt, 1, h);
            } else {
                dE = mott_transfer_all(d, Eel, h.Mass, h);
            }
            double mu = cos_theta_W(d, Eel, dE, h.Mass * kMe, h);
            hth1 = trk_acos(mu);
            hph1 = (2.0 * kPi) * rn(d, h);
            Ehole = check_hole(d, Eel, dE, nullptr);
            tally_elat(d, rep, t_ev, X, Y, dE);
        }
        double hph2, hth2;
        new_angles(ph0, th0, cth0, hth1, hph1, hph2, hth2);
        h.t0 = h.tn; h.X = X; h.Y = Y; h.Z = Z; h.th = hth2; h.ph = hph2;
        hole_params(d, h, Ehole + Egap, h);
        cut_ho(d, h);
        trace_ev(d, kind, h, ev, t_ev, h.Ehkin, h.E, X, Y, Z);
        return kind;
    }
    // deep shell, MC:2739-2868
    double RN = rn(d, h);
    double tA = d.m.Auger[h.sh], tR = d.m.Radiat[h.sh];
    if (RN * (1.0 / tA + 1.0 / tR) < 1.0 / tA) {
        kind = EV_AUGER;
        int SHL = h.sh;
        int Sh2 = -1;
        double Ee = -1.0e-10, E_new2 = 0.0;
        double coun = auger_count(d, d.m.Ip[SHL], 0);
        double Shel = rn(d, h) * coun;
        int Sh1 = auger_choose(d, d.m.Ip[SHL], Shel, 0, SHL);
        double dEc = 0.0;
        if (d.m.n_dos > 0 && Sh1 == d.m.vb) dEc = from_where_in_vb(d, 0, 0.0, h);
        double E_new1 = dEc + d.m.Ip[Sh1];
        double Ediff = d.m.Ip[SHL] - E_new1;
        coun = auger_count(d, Ediff, 1);
        if (coun > 0.0) {
            Shel = rn(d, h) * coun;
            Sh2 = auger_choose(d, Ediff, Shel, 1, -1);
            dEc = 0.0;
            if (Sh2 >= 0 && d.m.n_dos > 0 && Sh2 == d.m.vb) dEc = from_where_in_vb(d, 1, Ediff - d.m.Ip[Sh2], h);
            E_new2 = dEc + ((Sh2 >= 0) ? d.m.Ip[Sh2] : 0.0);
            Ee = Ediff - E_new2;
        }
        if (Sh2 >= 0) {
            double hth = kPi * rn(d, h);
            double hph = (2.0 * kPi) * rn(d, h);
            h.t0 = h.tn; h.sh = Sh1; h.th = hth; h.ph = hph;
            hole_params(d, h, E_new1, h);
            cut_ho(d, h);
            PState ne, nh;
            init_child(nh, h.id, ev, 0, h.iter, 1);
            init_child(ne, h.id, ev, 1, h.iter, 0);
            double nth = kPi * rn(d, h);
            double nph = (2.0 * kPi) * rn(d, h);
            nh.t0 = h.t0; nh.X = h.X; nh.Y = h.Y; nh.Z = h.Z; nh.sh = Sh2; nh.th = nth; nh.ph = nph;
            hole_params(d, nh, E_new2, h);
            cut_ho(d, nh);
            double MFP = el_mfp(d, Ee, rn(d, h));
            double phi1 = (2.0 * kPi) * rn(d, h);
            double theta1 = kPi * rn(d, h);
            ne.E = Ee; ne.t0 = h.t0; ne.X = h.X; ne.Y = h.Y; ne.Z = h.Z; ne.L = MFP; ne.th = theta1; ne.ph = phi1;
            ne.tn = t_next(ne.t0, MFP, vel_el(ne.E));
            cut_el(d, ne);
            emit_pair(d, ne, nh);
        } else {
            h.t0 = t_ev;
            h.tn = 1.0e21;
        }
    } else {
        kind = EV_RAD;
        int SHL = h.sh;
        double coun = auger_count(d, d.m.Ip[SHL], 1);
        double Shel = rn(d, h) * coun;
        int Sh1 = auger_choose(d, d.m.Ip[SHL], Shel, 1, SHL);
        double dEc = 0.0;
        if (d.m.n_dos > 0 && Sh1 == d.m.vb) dEc = from_where_in_vb(d, 0, 0.0, h);
        double E_new1 = dEc + d.m.Ip[Sh1];
        double dE = d.m.Ip[SHL] - E_new1;
        double hth = kPi * rn(d, h);
        double hph = (2.0 * kPi) * rn(d, h);
        h.t0 = h.tn; h.sh = Sh1; h.th = hth; h.ph = hph;
        hole_params(d, h, E_new1, h);
        cut_ho(d, h);
        PState q;
        init_child(q, h.id, ev, 0, h.iter, 0);
        double IM = nfp_2d(d.EP, d.phot_imfp, d.nP, dE);
        double MFP = -trk_log(rn(d, h)) * IM;
        double phi1 = (2.0 * kPi) * rn(d, h);
        double theta1 = kPi * rn(d, h);
        q.E = dE; q.t0 = h.t0; q.X = h.X; q.Y = h.Y; q.Z = h.Z; q.L = MFP; q.th = theta1; q.ph = phi1;
        q.tn = t_next(q.t0, MFP, kC);
        emit_photon(d, q);
    }
    trace_ev(d, kind, h, ev, t_ev, h.Ehkin, h.E, h.X, h.Y, h.Z);
    return kind;
}

// Photon_Monte_Carlo, MC:2873-2965
TRK_DEV void photon_event(const McDev &d, PState &q)
{
    uint32_t ev = q.ne++;
    double t_ev = q.tn;
    double Eel = q.E;
    double L = q.L, th0 = q.th, ph0 = q.ph;
    double sth0, cth0, sph0, cph0;
    trk_sincos(th0, &sth0, &cth0);
    trk_sincos(ph0, &sph0, &cph0);
    double X = q.X + (L * sth0) * sph0;
    double Y = q.Y + (L * sth0) * cph0;
    double Z = q.Z + L * cth0;
    int sh = which_shell(d, d.EP, d.LPsh, d.nP, Eel, q);
    double dE_cur = receives_E(d, Eel, sh, q);
    double MFP = el_mfp(d, dE_cur, rn(d, q));
    double phi1, theta1;
    new_angles(q.ph, q.th, cth0, kPi / 2.0, 0.0, phi1, theta1);
    PState ne, nh;
    init_child(ne, q.id, ev, 0, q.iter, 0);
    init_child(nh, q.id, ev, 1, q.iter, 1);
    ne.E = dE_cur; ne.t0 = q.tn; ne.X = X; ne.Y = Y; ne.Z = Z; ne.L = MFP; ne.th = theta1; ne.ph = phi1;
    ne.tn = t_next(ne.t0, MFP, vel_el(ne.E));
    cut_el(d, ne);
    double hth = kPi * rn(d, q);
    double hph = (2.0 * kPi) * rn(d, q);
    nh.t0 = q.tn; nh.X = X; nh.Y = Y; nh.Z = Z; nh.sh = sh; nh.th = hth; nh.ph = hph;
    hole_params(d, nh, Eel - dE_cur, q);
    cut_ho(d, nh);
    emit_pair(d, ne, nh);
    q.E = 0.0; q.t0 = 1.0e27; q.tn = 1.0e27; q.X = 0.0; q.Y = 0.0; q.Z = 0.0; q.L = 1.0e26; q.th = 0.0; q.ph = 0.0;
    trace_ev(d, EV_PHOT, q, ev, t_ev, Eel, 0.0, X, Y, Z);
}

// ---------------------------------------------------------------- SHI
// SHI_energy_transfer, MC:1719-1780 (CDF branch)
TRK_DEV double shi_transfer(const McDev &d, int shell, PState &I)
{
    const int N = d.sd_n[shell];
    const double *E = d.sd_E + d.sd_off[shell], *L = d.sd_L + d.sd_off[shell];
    double RN = rn(d, I);
    if (N < 1) return TRK_NAN;
    double Ecur = d.m.Ip[shell];
    int M = find_mono_1d(E, N, d.m.Ip[shell]);
    double dL;
    if (M > 1) dL = (L[M - 2] > 1.0e-10) ? interpolate(5, E[M - 2], E[M - 1], L[M - 2], L[M - 1], Ecur)
                                          : interpolate(1, E[M - 2], E[M - 1], L[M - 2], L[M - 1], Ecur);
    else dL = L[0];
    double Tot = ((dL > 0.0) && (L[N - 1] > 0.0)) ? 1.0 / dL + RN * (1.0 / L[N - 1] - 1.0 / dL) : 1.5e21;
    int Nt;
    if (Tot < 1.0e20) {
        // SHI_energy_transfer scans linearly for the first i with
        // 1/L(i) >= Tot among i < N (Find_in_1D_array, RF:3348-3358); on a
        // row whose inverse never decreases a lower-bound bisection finds the
        // same index in log2(N) steps (the rows have ~10^3-10^4 nodes).
        const double *inv = d.sd_inv + d.sd_off[shell];
        if (!d.sd_nonmono[shell]) {
            int lo = 0, hi = N - 1;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (inv[mid] < Tot) lo = mid + 1;
                else hi = mid;
            }
            Nt = lo + 1;
        } else {
            int i = 1;
            while (i < N && inv[i - 1] < Tot) i++;
            Nt = i;
        }
    } else {
        Nt = M;
    }
    if (Nt > M) return interpolate(5, 1.0 / L[Nt - 2], 1.0 / L[Nt - 1], E[Nt - 2], E[Nt - 1], Tot);
    return d.m.Ip[shell];
}

TRK_DEV double shi_zeff_dev(const McDev &d, double E)
{
    double vp = (E > 0.0) ? trk_sqrt(((2.0 * E) * kE) / (d.m.ion_mass * kMp)) : 0.0;
    double Zp = (double)d.m.ion_Z;
    double Zt = d.m.Zt;
    double v0 = trk_sqrt(((2.0 * kRy) * kE) / kMe);
    switch (d.m.ion_kind_zeff) {
    case 1:
        return Zp * (1.0 - trk_exp(-((vp / v0) / trk_pow(Zp, 0.66666666))));
    case 2:
        return Zp * trk_pow(1.0 + trk_pow(vp / (((trk_pow(Zp, 0.45) * v0) * 4.0) / 3.0), -1.0 / 0.6), -0.6);
    case 3: {
        double c1 = 1.0 - 0.26 * trk_exp(-(Zt / 11.0) - ((Zt - Zp) * (Zt - Zp)) / 9.0);
        double vpvo = (trk_pow(Zp, -0.543) * vp) / v0;
        double c2 = 1.0 + (0.03 * vpvo) * trk_log(Zt);
        double x = c1 * trk_pow((vpvo / c2) / 1.54, 1.0 + 1.83 / Zp);
        double x2 = x * x, x4 = x2 * x2;
        return (Zp * ((8.29 * x) + x4)) / ((((0.06 / x) + 4.0) + (7.4 * x)) + x4);
    }
    case 4:
        return d.m.ion_fixed_zeff;
    default:
        return Zp * (1.0 - trk_exp(-(((vp * 125.0) / kC) / trk_pow(Zp, 0.66666666))));
    }
}

// The ion of every iteration starts at the surface with its first free
// flight drawn (MC:564); one thread per iteration writes it to the ion list.
__global__ void __launch_bounds__(64) k_ion_init(McDev d, int nb, uint32_t first_iter)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    PState I;
    I.E = d.m.ion_E; I.t0 = 0.0; I.X = 0.0; I.Y = 0.0; I.Z = 0.0; I.th = 0.0; I.ph = 0.0;
    I.Ehkin = 0.0; I.Mass = d.m.ion_Zeff0; I.sh = -1;   // Mass holds Zeff for the ion
    I.id = 0; I.nd = 0; I.ne = 0; I.iter = first_iter + (uint32_t)b;
    I.cb = 0u;
    const double MSHI = d.m.ion_mass * kMp;
    double SL = nfp_2d(d.ES, d.shi_path, d.nS, I.E);
    SL = -SL * trk_log(rn(d, I));
    I.tn = t_next(0.0, SL, trk_sqrt(((2.0 * I.E) * kE) / MSHI));
    I.L = SL;
    store(d.in[K_ION], b, I, 1);
}

// One ionising collision of the ion (SHI_Monte_Carlo, MC:2153-2249): shell,
// energy transfer, next free flight, then the delta electron and the hole.
TRK_DEV void shi_event(const McDev &d, PState &I)
{
    const double MSHI = d.m.ion_mass * kMp;
    uint32_t ev = I.ne++;
    double t_ev = I.tn;
    int sh = which_shell(d, d.ES, d.LSsh, d.nS, I.E, I);
    double dE = shi_transfer(d, sh, I);
    double SI = nfp_2d(d.ES, d.shi_path, d.nS, I.E);
    SI = -SI * trk_log(rn(d, I));
    double Z = I.Z + I.L;
    I.E = I.E - dE;
    I.t0 = I.tn;
    I.Z = Z;
    I.L = SI;
    I.tn = t_next(I.t0, SI, trk_sqrt(((2.0 * I.E) * kE) / MSHI));
    double Zeff = shi_zeff_dev(d, I.E);
    I.Mass = Zeff;
    PState ne, nh;
    init_child(ne, I.id, ev, 0, I.iter, 0);
    init_child(nh, I.id, ev, 1, I.iter, 1);
    double dE_cur = receives_E(d, dE, sh, I);
    double theta = (I.E <= 0.0) ? kPi / 2.0
                                : trk_acos(trk_sqrt(((((MSHI + kMe) * (MSHI + kMe)) / ((4.0 * MSHI) * kMe)) * dE) / I.E));
    double phi = (2.0 * kPi) * rn(d, I);
    double MFP = el_mfp(d, dE_cur, rn(d, I));
    double A = 1.0 + MSHI / kMe;
    double Lb = (((kA0 * Zeff) * kRy) / I.E) * trk_sqrt(((((4.0 * I.E) / dE) * MSHI) / kMe) - A * A);
    double sphi, cphi;
    trk_sincos(phi, &sphi, &cphi);
    double X = I.X + Lb * sphi;
    double Y = I.Y + Lb * cphi;
    ne.E = dE_cur; ne.t0 = I.t0; ne.X = X; ne.Y = Y; ne.Z = Z; ne.L = MFP; ne.th = theta; ne.ph = phi;
    ne.tn = t_next(ne.t0, MFP, vel_el(ne.E));
    cut_el(d, ne);
    double hth = kPi * rn(d, I);
    double hph = (2.0 * kPi) * rn(d, I);
    nh.t0 = I.t0; nh.X = X; nh.Y = Y; nh.Z = Z; nh.sh = sh; nh.th = hth; nh.ph = hph;
    hole_params(d, nh, dE - dE_cur, I);
    cut_ho(d, nh);
    if (Z >= d.m.Layer)
        I.tn = 1.0e16;
    emit_pair(d, ne, nh);
    trace_ev(d, EV_SHI, I, ev, t_ev, I.E, Zeff, I.X, I.Y, I.Z);
}

// Category of a live particle's next event, read from the draw that event
// will use first (the draw is only looked at here; the event consumes it):
// the electron el/inel test (MC:2301-2309), the VB-hole test (MC:2568-2586),
// deep-hole decay, photoabsorption and ion collisions.  Fills the free-path
// cache it needs.
enum { C_E_EL = 0, C_E_INEL, C_H_EL, C_H_INEL, C_DEEP, C_PHOT, C_ION, C_COUNT };

TRK_DEV int peek_category(const McDev &d, PState &p, int kind)
{
    if (kind == K_PH) return C_PHOT;
    if (kind == K_ION) return C_ION;
    if (kind == K_HO && p.sh != d.m.vb) return C_DEEP;
    const double RN = rn_at(d, p, p.nd);
    if (kind == K_EL) {
        if (!(p.E == p.cE)) {
            double lE = trk_log(p.E);
            int N;
            p.cIM = nfp_2d_ln(d.E, d.el_imfp, d.lnE, d.ln_el_imfp, d.nE, p.E, lE, &N);
            p.cEM = nfp_2d_ln(d.Eel, d.Lel, d.lnEel, d.ln_Lel, d.nEel, p.E, lE, &N);
            p.cN = N;
            p.cE = p.E;
        }
        return (RN * (1.0 / p.cIM + 1.0 / p.cEM) < 1.0 / p.cIM) ? C_E_INEL : C_E_EL;
    }
    if (!(p.Ehkin == p.cE)) {
        double lE = trk_log(p.Ehkin);
        int N;
        p.cIM = nfp_2d_ln(d.EH, d.hole_imfp, d.lnEH, d.ln_hole_imfp, d.nH, p.Ehkin, lE, &N);
        p.cEM = nfp_2d_ln(d.EHel, d.LHel, d.lnEHel, d.ln_LHel, d.nHel, p.Ehkin, lE, &N);
        p.cN = N;
        p.cE = p.Ehkin;
    }
    return ((RN * (1.0 / p.cIM + 1.0 / p.cEM) < 1.0 / p.cIM) && (p.cIM < 1.0e15)) ? C_H_INEL : C_H_EL;
}

// A wave runs a category of events only when at least this many of its lanes
// want it (or when it is the most wanted one); the other lanes wait with
// their state untouched.  Rare, long branches (ionisation, decays) then run
// for many lanes at once instead of stalling the wave at almost every step.
constexpr int kBatchLanes = 16;
constexpr int kMaxWait = 8;

// Events one particle may take in one round.  A particle with more left is
// written to the next round's list; this bounds the tail of every launch by
// the cost of kRoundEvents events instead of the longest history.
constexpr uint32_t kRoundEvents = 512;

// One round of transport.  Items [0, n_e) are electrons, then holes, photons
// and ions.  A lane takes the next item from a shared cursor whenever its
// particle is finished (below the cut-off, absorbed, past the last time
// node) or has used its event budget for the round, so waves stay full until
// the round's list is used up; every lane leaves once the cursor has passed
// the last item.  Children and unfinished particles go to the next round's
// lists.  A particle's snapshot tallies for the interval up to its next
// event are added once, when that interval is known (when the particle is
// loaded and after each event it takes in this round).
// The small tables every event searches (energy grids, fused free paths and
// their logs, per-shell free paths, radial and time grids, the VB DOS) are
// copied into LDS once per workgroup: a bisection step then costs an LDS
// access instead of an L2 round trip.  Order of the staged arrays:
//  0-3 E, El_IMFP, ln E, ln El_IMFP      4-7 elastic E, L, ln E, ln L
//  8-11 hole E, Hole_IMFP, logs          12-15 hole elastic E, L, logs
//  16 Out_R  17 time nodes  18-21 DOS E, DOS, int_DOS, Eff_m
//  22 per-shell electron L  23 per-shell hole L
extern __shared__ double lds_tab[];

TRK_DEV void stage_tables(const McDev &d, McDev &l)
{
    const double *src[24] = {d.E, d.el_imfp, d.lnE, d.ln_el_imfp, d.Eel, d.Lel, d.lnEel, d.ln_Lel,
                             d.EH, d.hole_imfp, d.lnEH, d.ln_hole_imfp, d.EHel, d.LHel, d.lnEHel, d.ln_LHel,
                             d.R, d.tim_glob, d.m.dosE, d.m.dos, d.m.dosI, d.m.dosM, d.Lsh, d.LHsh};
#pragma unroll
    for (int a = 0; a < 24; ++a) {
        const int n = d.lds_n[a];
        double *dst = lds_tab + d.lds_off[a];
        for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[a][i];
    }
    __syncthreads();
#define TRK_STAGED(a, field) if (d.lds_n[a]) l.field = lds_tab + d.lds_off[a]
    TRK_STAGED(0, E); TRK_STAGED(1, el_imfp); TRK_STAGED(2, lnE); TRK_STAGED(3, ln_el_imfp);
    TRK_STAGED(4, Eel); TRK_STAGED(5, Lel); TRK_STAGED(6, lnEel); TRK_STAGED(7, ln_Lel);
    TRK_STAGED(8, EH); TRK_STAGED(9, hole_imfp); TRK_STAGED(10, lnEH); TRK_STAGED(11, ln_hole_imfp);
    TRK_STAGED(12, EHel); TRK_STAGED(13, LHel); TRK_STAGED(14, lnEHel); TRK_STAGED(15, ln_LHel);
    TRK_STAGED(16, R); TRK_STAGED(17, tim_glob); TRK_STAGED(18, m.dosE); TRK_STAGED(19, m.dos);
    TRK_STAGED(20, m.dosI); TRK_STAGED(21, m.dosM); TRK_STAGED(22, Lsh); TRK_STAGED(23, LHsh);
#undef TRK_STAGED
}

constexpr int kTransportBlock = 768;   // 12 waves: one workgroup per CU at 3 waves per SIMD

__global__ void __launch_bounds__(kTransportBlock) k_transport(McDev din)
{
    McDev d = din;
    stage_tables(din, d);
    // ions first: they start every round at once and feed all other work
    const int n0 = d.n_in[K_ION], n1 = n0 + d.n_in[K_EL], n2 = n1 + d.n_in[K_HO], total = n2 + d.n_in[K_PH];
    unsigned long long c_ei = 0, c_ee = 0, c_hi = 0, c_he = 0, c_au = 0, c_ra = 0, c_ph = 0, c_shi = 0;
    PState p;
    int kind = 0, rep = 0, cat = 0;
    uint32_t budget = 0;
    int waited = 0;
    bool have = false;
    for (;;) {
        if (!have) {
            int t = wave_alloc(d.cursor);
            if (t >= total) break;
            if (t < n0) { kind = K_ION; load(d.in[K_ION], t, p, 1); }
            else if (t < n1) { kind = K_EL; load(d.in[K_EL], t - n0, p, 0); }
            else if (t < n2) { kind = K_HO; load(d.in[K_HO], t - n1, p, 1); }
            else { kind = K_PH; load(d.in[K_PH], t - n2, p, 0); }
            rep = (int)(p.iter % (uint32_t)d.T.R);
            budget = kRoundEvents;
            waited = 0;
            if (kind != K_ION) segment(d, p, kind, rep);
            if (p.tn < d.T_last) {
                if (p.E != p.E) atomicOr(d.overflow, kFlagNaN);
                else { have = true; cat = peek_category(d, p, kind); }
            }
        }
        // choose the categories this wave runs now: every category wanted by
        // kBatchLanes lanes, the most wanted one, the ions (few, and the root
        // of all other work) and any category a lane has waited kMaxWait
        // steps for
        int best = -1, bestn = 0;
        unsigned run = 1u << C_ION;
#pragma unroll
        for (int c = 0; c < C_COUNT; ++c) {
            int n = __popcll(__ballot(have && cat == c));
            if (n >= kBatchLanes) run |= 1u << c;
            if (n > bestn) { bestn = n; best = c; }
            if (__ballot(have && cat == c && waited >= kMaxWait)) run |= 1u << c;
        }
        if (best >= 0) run |= 1u << best;
        if (!have) continue;
        if (!((run >> cat) & 1u)) { ++waited; continue; }
        waited = 0;

        if (kind == K_EL) {
            int k = electron_event(d, p, rep);
            if (k == EV_E_INEL) c_ei++; else c_ee++;
        } else if (kind == K_HO) {
            int k = hole_event(d, p, rep);
            if (k == EV_H_INEL) c_hi++;
            else if (k == EV_H_EL) c_he++;
            else if (k == EV_AUGER) c_au++;
            else c_ra++;
        } else if (kind == K_PH) {
            photon_event(d, p);
            c_ph++;
            have = false;
            continue;
        } else {
            shi_event(d, p);
            c_shi++;
        }
        if (p.ne >= kMaxEvents) { atomicOr(d.overflow, kFlagEvents); have = false; continue; }
        if (p.E != p.E || p.tn != p.tn) { atomicOr(d.overflow, kFlagNaN); have = false; continue; }
        if (--budget == 0 && p.tn < d.T_last) {
            // continues next round; its next interval is tallied then.  The
            // append is made per kind so that every lane in one wave-wide
            // allocation uses the same list counter.
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (kind == k) emit(d, k, p);
            have = false;
            continue;
        }
        if (kind != K_ION) segment(d, p, kind, rep);
        if (!(p.tn < d.T_last)) { have = false; continue; }
        cat = peek_category(d, p, kind);
    }
    if (c_ei) atomicAdd(&d.counters[EV_E_INEL], c_ei);
    if (c_ee) atomicAdd(&d.counters[EV_E_EL], c_ee);
    if (c_hi) atomicAdd(&d.counters[EV_H_INEL], c_hi);
    if (c_he) atomicAdd(&d.counters[EV_H_EL], c_he);
    if (c_au) atomicAdd(&d.counters[EV_AUGER], c_au);
    if (c_ra) atomicAdd(&d.counters[EV_RAD], c_ra);
    if (c_ph) atomicAdd(&d.counters[EV_PHOT], c_ph);
    if (c_shi) atomicAdd(&d.counters[EV_SHI], c_shi);
}

// ---------------------------------------------------------------- per-iteration tallies
// Prefix sums over time of [B][Nt][ncol] difference arrays, one thread per column.
__global__ void k_prefix_i(int *a, int B, int Nt, int ncol)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int64_t)B * ncol) return;
    int64_t b = c / ncol, col = c % ncol;
    int *p = a + (size_t)b * Nt * ncol + col;
    int run = 0;
    for (int i = 0; i < Nt; ++i) { run += p[(size_t)i * ncol]; p[(size_t)i * ncol] = run; }
}

// Same for the fixed-point sums [B][Nt][2]; the running value is left in word 0.
__global__ void k_prefix_fx(unsigned long long *a, int B, int Nt)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    unsigned long long *p = a + (size_t)b * Nt * 2;
    long long run = 0;
    for (int i = 0; i < Nt; ++i) {
        run += (long long)p[2 * i + 1] * 4294967296ll + (long long)p[2 * i];
        p[2 * i] = (unsigned long long)run;
        p[2 * i + 1] = 0ull;
    }
}

// acc(i, j) += sum over iterations of count * (w_j / norm), in iteration order.
// wmode 0: w = 1 (angle histograms); 1: w from the radial grid used as an
// energy grid (MC:1024-1029); 2: w from the DOS grid (MC:1033-1040).
__global__ void k_norm_hist(double *acc, int accld, const int *cnt, const int *norm, int B, int Nt, int ncol, int wmode,
                            const double *grid)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Nt * ncol) return;
    int i = (int)(t % Nt), j = (int)(t / Nt);
    double w = 1.0;
    if (wmode == 1) w = (j > 0) ? 1.0 / (grid[j] - grid[j - 1]) : 1.0 / grid[0];
    else if (wmode == 2) w = (j > 0) ? 1.0 / (grid[j] - grid[j - 1]) : 1.0 / (grid[1] - grid[0]);
    double sum = acc[(size_t)i + (size_t)accld * j];
    for (int b = 0; b < B; ++b) {
        int c = cnt[((size_t)b * Nt + i) * ncol + j];
        if (c) sum = sum + (double)c * (w / (double)norm[(size_t)b * Nt + i]);
    }
    acc[(size_t)i + (size_t)accld * j] = sum;
}

// Out_diff_coeff recurrence of MC:1094-1097: every iteration divides the
// running value plus its own sum by its own count.
__global__ void k_diff_coeff(double *D, const unsigned long long *dcs, const int *dcn, int B, int Nt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt) return;
    double v = D[i];
    for (int b = 0; b < B; ++b) {
        int n = dcn[(size_t)b * Nt + i];
        if (n > 0) {
            long long x = (long long)dcs[((size_t)b * Nt + i) * 2];
            long long hi = x >> 40;
            double s = (double)hi + (double)(x - (hi << 40)) * 0x1.0p-40;
            v = (v + s) / (double)n;
        }
    }
    D[i] = v;
}

// ---------------------------------------------------------------- tally reduction
// Sums the R replicas of every tally cell and takes the prefix over time of
// the difference arrays, in exact integer arithmetic, leaving the result in
// replica 0: counts as int64, energies as int128 (low word, high word) in
// units of 2^-40 eV.  One thread per column: electrons and photons by radial
// bin, the lattice energy by radial bin (per node, no prefix), holes by
// (radial bin, shell).
TRK_DEV __int128 fx_read(const unsigned long long *p)
{
    return (__int128)(long long)p[1] * ((__int128)1 << 32) + (__int128)p[0];
}

TRK_DEV void i128_store(unsigned long long *p, __int128 v)
{
    p[0] = (unsigned long long)v;
    p[1] = (unsigned long long)(v >> 64);
}

__global__ void k_tally_reduce(Tally T)
{
    const int NR = T.NR, Nt = T.Nt, nsh = T.nsh, R = T.R;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t plane = (size_t)Nt * NR;
    if (t < 2 * NR) {                       // electrons, photons
        const int j = (int)(t % NR);
        unsigned long long *cnt = (t < NR) ? T.cnt_e : T.cnt_p;
        unsigned long long *en = (t < NR) ? T.E_e : T.E_p;
        long long rc = 0;
        __int128 re = 0;
        for (int i = 0; i < Nt; ++i) {
            const size_t a = (size_t)i * NR + j;
            long long sc = 0;
            __int128 se = 0;
            for (int r = 0; r < R; ++r) {
                sc += (long long)cnt[r * plane + a];
                se += fx_read(en + 2 * (r * plane + a));
            }
            rc += sc;
            re += se;
            cnt[a] = (unsigned long long)rc;
            i128_store(en + 2 * a, re);
        }
    } else if (t < 3 * NR) {                // lattice energy per node
        const int j = (int)(t - 2 * NR);
        for (int i = 0; i < Nt; ++i) {
            const size_t a = (size_t)i * NR + j;
            __int128 se = 0;
            for (int r = 0; r < R; ++r) se += fx_read(T.Elat + 2 * (r * plane + a));
            i128_store(T.Elat + 2 * a, se);
        }
    } else if (t < 3 * NR + (int64_t)NR * nsh) {   // holes
        const int64_t q = t - 3 * NR;
        const int j = (int)(q / nsh), sh = (int)(q % nsh);
        const size_t hplane = plane * nsh;
        long long rc = 0;
        __int128 rE = 0, rK = 0;
        for (int i = 0; i < Nt; ++i) {
            const size_t a = ((size_t)i * NR + j) * nsh + sh;
            long long sc = 0;
            __int128 sE = 0, sK = 0;
            for (int r = 0; r < R; ++r) {
                sc += (long long)T.cnt_h[r * hplane + a];
                sE += fx_read(T.E_h + 2 * (r * hplane + a));
                sK += fx_read(T.Ek_h + 2 * (r * hplane + a));
            }
            rc += sc; rE += sE; rK += sK;
            T.cnt_h[a] = (unsigned long long)rc;
            i128_store(T.E_h + 2 * a, rE);
            i128_store(T.Ek_h + 2 * a, rK);
        }
    }
}

// ---------------------------------------------------------------- host side
struct McState {
    PList L[2][4];            // [ping-pong][electrons, holes, photons, ions]
    int cap[4] = {0, 0, 0, 0};
    int *d_counts = nullptr;  // [9]: list counts of round A [4], of round B [4], transport cursor
    int64_t max_blocks = 0;   // resident transport workgroups on the device
    int ncu = 0, max_lds = 0;
    size_t lds_bytes = 0;     // staged tables per transport workgroup
    int *d_overflow = nullptr;
    unsigned long long *d_counters = nullptr;
    Tally T{};
    size_t tally_bytes = 0;
    double *d_tim = nullptr, *d_R = nullptr;
    int NR_alloc = 0, Nt_alloc = 0, R_alloc = 0;
    double *d_trace = nullptr;
    unsigned long long *d_trace_n = nullptr;
    int64_t trace_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // per-iteration tallies (allocated for spec_B iterations) and their accumulators
    Spec S{};
    int spec_B = 0, spec_Nt = 0, spec_NR = 0, spec_ND = 0;
    double *acc_eve = nullptr, *acc_ehve = nullptr, *acc_th = nullptr, *acc_thh = nullptr, *acc_dc = nullptr;
};

#define HIPC(x)                                                                          \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) return set_error(c, TRK_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

static int alloc_list(trk_ctx *c, PList &l, int cap, int hole)
{
    size_t n = (size_t)cap;
    double **dp[] = {&l.E, &l.t0, &l.tn, &l.X, &l.Y, &l.Z, &l.L, &l.th, &l.ph};
    for (double **p : dp) HIPC(hipMalloc((void **)p, n * sizeof(double)));
    if (hole) {
        HIPC(hipMalloc((void **)&l.Ehkin, n * sizeof(double)));
        HIPC(hipMalloc((void **)&l.Mass, n * sizeof(double)));
        HIPC(hipMalloc((void **)&l.sh, n * sizeof(int)));
    } else {
        l.Ehkin = l.Mass = nullptr;
        l.sh = nullptr;
    }
    HIPC(hipMalloc((void **)&l.id, n * sizeof(uint64_t)));
    HIPC(hipMalloc((void **)&l.nd, n * sizeof(uint32_t)));
    HIPC(hipMalloc((void **)&l.ne, n * sizeof(uint32_t)));
    HIPC(hipMalloc((void **)&l.iter, n * sizeof(uint32_t)));
    return 0;
}

static void free_list(PList &l)
{
    void *ps[] = {l.E, l.t0, l.tn, l.X, l.Y, l.Z, l.L, l.th, l.ph, l.Ehkin, l.Mass, l.sh, l.id, l.nd, l.ne, l.iter};
    for (void *p : ps) if (p) (void)hipFree(p);
    l = PList{};
}

static void free_spec(McState *s)
{
    void *ps[] = {s->S.nel, s->S.nvb, s->S.dcn, s->S.eve, s->S.ehve, s->S.th, s->S.thh, s->S.dcs,
                  s->acc_eve, s->acc_ehve, s->acc_th, s->acc_thh, s->acc_dc};
    for (void *p : ps) if (p) (void)hipFree(p);
    s->S = Spec{};
    s->acc_eve = s->acc_ehve = s->acc_th = s->acc_thh = s->acc_dc = nullptr;
    s->spec_B = 0;
}

void mc_free(McState *s)
{
    if (!s) return;
    free_spec(s);
    for (auto &a : s->L) for (auto &l : a) free_list(l);
    void *ps[] = {s->d_counts, s->d_overflow, s->d_counters, s->T.cnt_e, s->T.E_e, s->T.cnt_h, s->T.E_h,
                  s->T.Ek_h, s->T.cnt_p, s->T.E_p, s->T.Elat, s->d_tim, s->d_R, s->d_trace, s->d_trace_n};
    for (void *p : ps) if (p) (void)hipFree(p);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    delete s;
}

static int ensure_tally(trk_ctx *c, McState *s, int Nt, int NR, int nsh, int R)
{
    if (s->T.cnt_e && s->Nt_alloc == Nt && s->NR_alloc == NR && s->R_alloc == R && s->T.nsh == nsh) return 0;
    void *ps[] = {s->T.cnt_e, s->T.E_e, s->T.cnt_h, s->T.E_h, s->T.Ek_h, s->T.cnt_p, s->T.E_p, s->T.Elat};
    for (void *p : ps) if (p) (void)hipFree(p);
    size_t cell = (size_t)R * Nt * NR;
    Tally &T = s->T;
    T.Nt = Nt; T.NR = NR; T.nsh = nsh; T.R = R;
    HIPC(hipMalloc((void **)&T.cnt_e, cell * 8));
    HIPC(hipMalloc((void **)&T.E_e, cell * 16));
    HIPC(hipMalloc((void **)&T.cnt_h, cell * nsh * 8));
    HIPC(hipMalloc((void **)&T.E_h, cell * nsh * 16));
    HIPC(hipMalloc((void **)&T.Ek_h, cell * nsh * 16));
    HIPC(hipMalloc((void **)&T.cnt_p, cell * 8));
    HIPC(hipMalloc((void **)&T.E_p, cell * 16));
    HIPC(hipMalloc((void **)&T.Elat, cell * 16));
    s->tally_bytes = cell * (8 + 16 + nsh * 40 + 8 + 16 + 16);
    s->Nt_alloc = Nt; s->NR_alloc = NR; s->R_alloc = R;
    return 0;
}

static int ensure_spec(trk_ctx *c, McState *s, int B, int Nt, int NR, int ND)
{
    if (s->spec_B >= B && s->spec_Nt == Nt && s->spec_NR == NR && s->spec_ND == ND) return 0;
    free_spec(s);
    Spec &S = s->S;
    const size_t bt = (size_t)B * Nt;
    HIPC(hipMalloc((void **)&S.nel, bt * sizeof(int)));
    HIPC(hipMalloc((void **)&S.nvb, bt * sizeof(int)));
    HIPC(hipMalloc((void **)&S.dcn, bt * sizeof(int)));
    HIPC(hipMalloc((void **)&S.eve, bt * NR * sizeof(int)));
    HIPC(hipMalloc((void **)&S.ehve, bt * (ND > 1 ? ND : 1) * sizeof(int)));
    HIPC(hipMalloc((void **)&S.th, bt * 180 * sizeof(int)));
    HIPC(hipMalloc((void **)&S.thh, bt * 180 * sizeof(int)));
    HIPC(hipMalloc((void **)&S.dcs, bt * 2 * sizeof(unsigned long long)));
    HIPC(hipMalloc((void **)&s->acc_eve, (size_t)Nt * NR * sizeof(double)));
    HIPC(hipMalloc((void **)&s->acc_ehve, (size_t)Nt * (ND > 1 ? ND : 1) * sizeof(double)));
    HIPC(hipMalloc((void **)&s->acc_th, (size_t)(Nt + 1) * 180 * sizeof(double)));
    HIPC(hipMalloc((void **)&s->acc_thh, (size_t)(Nt + 1) * 180 * sizeof(double)));
    HIPC(hipMalloc((void **)&s->acc_dc, (size_t)Nt * sizeof(double)));
    s->spec_B = B; s->spec_Nt = Nt; s->spec_NR = NR; s->spec_ND = ND;
    return 0;
}

static void fill_mcdev(const trk_ctx *c, McDev &d)
{
    const TableSet &ts = c->ts;
    const TableDev *T = ts.tab;
    d.m = c->md;
    d.nE = T[TRK_TAB_E_INEL].n; d.E = T[TRK_TAB_E_INEL].E; d.Lsh = T[TRK_TAB_E_INEL].L; d.el_imfp = ts.el_imfp;
    d.roff = T[TRK_TAB_E_INEL].row_off; d.rhw = T[TRK_TAB_E_INEL].hw; d.rlam = T[TRK_TAB_E_INEL].lam;
    d.nEel = T[TRK_TAB_E_EL].n; d.Eel = T[TRK_TAB_E_EL].E; d.Lel = T[TRK_TAB_E_EL].L;
    d.roff_el = T[TRK_TAB_E_EL].row_off; d.rhw_el = T[TRK_TAB_E_EL].hw; d.rlam_el = T[TRK_TAB_E_EL].lam;
    d.nH = T[TRK_TAB_H_INEL].n; d.EH = T[TRK_TAB_H_INEL].E; d.LHsh = T[TRK_TAB_H_INEL].L; d.hole_imfp = ts.hole_imfp;
    d.roff_h = T[TRK_TAB_H_INEL].row_off; d.rhw_h = T[TRK_TAB_H_INEL].hw; d.rlam_h = T[TRK_TAB_H_INEL].lam;
    d.nHel = T[TRK_TAB_H_EL].n; d.EHel = T[TRK_TAB_H_EL].E; d.LHel = T[TRK_TAB_H_EL].L;
    d.roff_hel = T[TRK_TAB_H_EL].row_off; d.rhw_hel = T[TRK_TAB_H_EL].hw; d.rlam_hel = T[TRK_TAB_H_EL].lam;
    d.nP = T[TRK_TAB_PHOT].n; d.EP = T[TRK_TAB_PHOT].E; d.LPsh = T[TRK_TAB_PHOT].L; d.phot_imfp = ts.phot_imfp;
    d.nS = T[TRK_TAB_SHI].n; d.ES = T[TRK_TAB_SHI].E; d.LSsh = T[TRK_TAB_SHI].L; d.shi_path = ts.shi_path;
}

}  // namespace trk

namespace trk {

static __int128 fx_value(const unsigned long long *p) { return (__int128)(long long)p[1] * ((__int128)1 << 32) + (__int128)p[0]; }
static double fx_to_double(__int128 v)
{
    // exact integer in units of 2^-40 eV -> double (two roundings at most)
    long long hi = (long long)(v >> 40);
    __int128 lo = v - ((__int128)hi << 40);
    return (double)hi + (double)(long long)lo * 0x1.0p-40;
}

// Each generation is the set of particles born in the previous one; a chain of
// ionisations is at most a few hundred long.
constexpr int kMaxGenerations = 1 << 16;

int run_mc(trk_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter, const double *Out_R, const double *Out_V,
           int NR, trk_tallies *out)
{
    auto tstart = std::chrono::steady_clock::now();
    const HostMat &h = c->hm;
    if (!c->mc) c->mc = new McState();
    McState *s = c->mc;
    const int Nt = h.Nt, nsh = h.nsh;
    const int R = (c->hm.o.tally_replicas > 0) ? c->hm.o.tally_replicas : 8;
    // capacities: particles of each kind per round
    int cap = (c->hm.o.gen_capacity > 0) ? c->hm.o.gen_capacity : (1 << 22);
    if (s->cap[0] != cap) {
        for (int g = 0; g < 2; ++g)
            for (int k = 0; k < 3; ++k) free_list(s->L[g][k]);
        const int caps[3] = {cap, cap, cap / 4};
        for (int g = 0; g < 2; ++g)
            for (int k = 0; k < 3; ++k)
                if (alloc_list(c, s->L[g][k], caps[k], k == K_HO)) return TRK_ERR_DEVICE;
        for (int k = 0; k < 3; ++k) s->cap[k] = caps[k];
        if (!s->d_counts) HIPC(hipMalloc((void **)&s->d_counts, 16 * sizeof(int)));
        if (!s->d_overflow) HIPC(hipMalloc((void **)&s->d_overflow, sizeof(int)));
        if (!s->d_counters) HIPC(hipMalloc((void **)&s->d_counters, 16 * sizeof(unsigned long long)));
        if (!s->ev0) { HIPC(hipEventCreate(&s->ev0)); HIPC(hipEventCreate(&s->ev1)); }
        int per_cu = 0, dev = 0, ncu = 0;
        HIPC(hipGetDevice(&dev));
        HIPC(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        s->ncu = ncu;
        HIPC(hipDeviceGetAttribute(&s->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        (void)per_cu;
    }
    if (ensure_tally(c, s, Nt, NR, nsh, R)) return TRK_ERR_DEVICE;
    if (s->d_tim) (void)hipFree(s->d_tim);
    if (s->d_R) (void)hipFree(s->d_R);
    HIPC(hipMalloc((void **)&s->d_tim, sizeof(double) * Nt));
    HIPC(hipMalloc((void **)&s->d_R, sizeof(double) * NR));
    hipStream_t st = c->stream;
    HIPC(hipMemcpyAsync(s->d_tim, h.tim_glob.data(), sizeof(double) * Nt, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(s->d_R, Out_R, sizeof(double) * NR, hipMemcpyHostToDevice, st));
    Tally &T = s->T;
    size_t cell = (size_t)R * Nt * NR;
    HIPC(hipMemsetAsync(T.cnt_e, 0, cell * 8, st));
    HIPC(hipMemsetAsync(T.E_e, 0, cell * 16, st));
    HIPC(hipMemsetAsync(T.cnt_h, 0, cell * nsh * 8, st));
    HIPC(hipMemsetAsync(T.E_h, 0, cell * nsh * 16, st));
    HIPC(hipMemsetAsync(T.Ek_h, 0, cell * nsh * 16, st));
    HIPC(hipMemsetAsync(T.cnt_p, 0, cell * 8, st));
    HIPC(hipMemsetAsync(T.E_p, 0, cell * 16, st));
    HIPC(hipMemsetAsync(T.Elat, 0, cell * 16, st));
    HIPC(hipMemsetAsync(s->d_counters, 0, 16 * sizeof(unsigned long long), st));
    HIPC(hipMemsetAsync(s->d_overflow, 0, sizeof(int), st));

    // SHI_DIFF per-shell layout
    const TableDev &td = c->ts.tab[TRK_TAB_SHI_DIFF];
    int *d_sdn = nullptr;
    int64_t *d_sdoff = nullptr;
    HIPC(hipMalloc((void **)&d_sdn, sizeof(int) * nsh));
    HIPC(hipMalloc((void **)&d_sdoff, sizeof(int64_t) * (nsh + 1)));
    HIPC(hipMemcpyAsync(d_sdn, td.dn.data(), sizeof(int) * nsh, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d_sdoff, td.doff.data(), sizeof(int64_t) * (nsh + 1), hipMemcpyHostToDevice, st));

    McDev d{};
    fill_mcdev(c, d);
    d.sd_n = d_sdn; d.sd_off = d_sdoff; d.sd_E = td.dE; d.sd_L = td.dL; d.sd_inv = td.dInv; d.sd_nonmono = td.dNonMono;
    d.seed = seed;
    d.Tim = h.o.Tim;
    d.T_last = h.tim_glob[Nt - 1];
    d.cutoff_eff = (h.m.cut_off > 0.0) ? h.m.cut_off : 0.0;
    d.Nt = Nt; d.tim_glob = s->d_tim;
    d.inv_dt = (h.o.dt_flag <= 0 && h.o.dt > 0.0) ? 1.0 / h.o.dt : 0.0;
    {
        const TableSet &ts = c->ts;
        d.lnE = ts.ln_E[TRK_TAB_E_INEL]; d.ln_el_imfp = ts.ln_el_imfp;
        d.lnEel = ts.ln_E[TRK_TAB_E_EL]; d.ln_Lel = ts.ln_Lel;
        d.lnEH = ts.ln_E[TRK_TAB_H_INEL]; d.ln_hole_imfp = ts.ln_hole_imfp;
        d.lnEHel = ts.ln_E[TRK_TAB_H_EL]; d.ln_LHel = ts.ln_LHel;
    }
    d.NR = NR; d.R = s->d_R;
    d.include_photons = h.o.include_photons;
    d.overflow = s->d_overflow;
    d.counters = s->d_counters;
    d.T = T;
    if (c->trace_host && c->trace_cap > 0) {
        if (s->trace_cap < c->trace_cap) {
            if (s->d_trace) (void)hipFree(s->d_trace);
            HIPC(hipMalloc((void **)&s->d_trace, sizeof(double) * 12 * c->trace_cap));
            if (!s->d_trace_n) HIPC(hipMalloc((void **)&s->d_trace_n, sizeof(unsigned long long)));
            s->trace_cap = c->trace_cap;
        }
        HIPC(hipMemsetAsync(s->d_trace_n, 0, sizeof(unsigned long long), st));
        d.trace = s->d_trace; d.trace_n = s->d_trace_n; d.trace_cap = c->trace_cap;
    }

    // batch size: SHI events per iteration x pairs per SHI event must fit one generation
    double est_pairs = 0.0;
    {
        std::vector<double> sp(c->ts.tab[TRK_TAB_SHI].n);
        HIPC(hipMemcpy(sp.data(), c->ts.shi_path, sizeof(double) * sp.size(), hipMemcpyDeviceToHost));
        std::vector<double> sE(c->ts.tab[TRK_TAB_SHI].n);
        HIPC(hipMemcpy(sE.data(), c->ts.tab[TRK_TAB_SHI].E, sizeof(double) * sE.size(), hipMemcpyDeviceToHost));
        double lam = nfp_2d(sE.data(), sp.data(), (int)sp.size(), h.ion.E);
        est_pairs = (lam > 0.0) ? h.m.Layer / lam : 1e4;
    }
    int64_t B = (c->hm.o.batch > 0) ? c->hm.o.batch : (int64_t)((double)cap / (3.0 * est_pairs + 64.0));
    if (B < 1) B = 1;
    if (B > nmc) B = nmc;
    if (s->cap[K_ION] < B) {
        for (int g = 0; g < 2; ++g) {
            free_list(s->L[g][K_ION]);
            if (alloc_list(c, s->L[g][K_ION], (int)B, 1)) return TRK_ERR_DEVICE;
        }
        s->cap[K_ION] = (int)B;
    }
    for (int k = 0; k < 4; ++k) d.cap[k] = s->cap[k];

    // per-iteration normalised tallies, when the caller asks for any of them
    const bool want_spec = out->Out_Ee_vs_E || out->Out_Eh_vs_E || out->Out_theta || out->Out_theta_h || out->Out_diff_coeff;
    const int ND = h.m.n_dos;
    d.S = Spec{};
    if (want_spec) {
        if (ensure_spec(c, s, (int)B, Nt, NR, ND)) return TRK_ERR_DEVICE;
        HIPC(hipMemsetAsync(s->acc_eve, 0, (size_t)Nt * NR * sizeof(double), st));
        HIPC(hipMemsetAsync(s->acc_ehve, 0, (size_t)Nt * (ND > 1 ? ND : 1) * sizeof(double), st));
        HIPC(hipMemsetAsync(s->acc_th, 0, (size_t)(Nt + 1) * 180 * sizeof(double), st));
        HIPC(hipMemsetAsync(s->acc_thh, 0, (size_t)(Nt + 1) * 180 * sizeof(double), st));
        if (out->Out_diff_coeff)
            HIPC(hipMemcpyAsync(s->acc_dc, out->Out_diff_coeff, (size_t)Nt * sizeof(double), hipMemcpyHostToDevice, st));
        else
            HIPC(hipMemsetAsync(s->acc_dc, 0, (size_t)Nt * sizeof(double), st));
        d.S = s->S;
        d.S.on = 1;
        d.S.ND = ND;
    }

    // LDS plan: stage the arrays in order while they fit
    {
        const int nsz[24] = {d.nE, d.nE, d.nE, d.nE, d.nEel, d.nEel, d.nEel, d.nEel,
                             d.nH, d.nH, d.nH, d.nH, d.nHel, d.nHel, d.nHel, d.nHel,
                             NR, Nt, h.m.n_dos, h.m.n_dos, h.m.n_dos, h.m.n_dos, nsh * d.nE, nsh * d.nH};
        const bool present[24] = {true, true, d.lnE != nullptr, d.ln_el_imfp != nullptr, true, true,
                                  d.lnEel != nullptr, d.ln_Lel != nullptr, true, true, d.lnEH != nullptr,
                                  d.ln_hole_imfp != nullptr, true, true, d.lnEHel != nullptr, d.ln_LHel != nullptr,
                                  true, true, true, true, true, true, true, true};
        const size_t limit = (size_t)s->max_lds / sizeof(double);
        size_t off = 0;
        for (int a = 0; a < 24; ++a) {
            size_t n = present[a] ? (size_t)nsz[a] : 0;
            if (off + n > limit) n = 0;
            d.lds_n[a] = (int)n;
            d.lds_off[a] = (int)off;
            off += n;
        }
        s->lds_bytes = off * sizeof(double);
        if (s->lds_bytes > 65536)
            HIPC(hipFuncSetAttribute((const void *)k_transport, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)s->lds_bytes));
        int per_cu = 0;
        HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_transport, kTransportBlock, s->lds_bytes));
        if (per_cu < 1) per_cu = 1;
        s->max_blocks = (int64_t)per_cu * s->ncu;
    }

    double transport_s = 0.0;
    int64_t launches = 0;
    std::vector<int> hc(8);
    const bool verbose = getenv("TRK_VERBOSE") != nullptr;   // per-round list sizes on stderr
    for (int64_t b0 = 0; b0 < nmc; b0 += B) {
        int nb = (int)((nmc - b0 < B) ? nmc - b0 : B);
        int cur = 0, rounds = 0;
        if (want_spec) {
            const size_t bt = (size_t)nb * Nt;
            HIPC(hipMemsetAsync(s->S.nel, 0, bt * sizeof(int), st));
            HIPC(hipMemsetAsync(s->S.nvb, 0, bt * sizeof(int), st));
            HIPC(hipMemsetAsync(s->S.dcn, 0, bt * sizeof(int), st));
            HIPC(hipMemsetAsync(s->S.eve, 0, bt * NR * sizeof(int), st));
            HIPC(hipMemsetAsync(s->S.ehve, 0, bt * (ND > 1 ? ND : 1) * sizeof(int), st));
            HIPC(hipMemsetAsync(s->S.th, 0, bt * 180 * sizeof(int), st));
            HIPC(hipMemsetAsync(s->S.thh, 0, bt * 180 * sizeof(int), st));
            HIPC(hipMemsetAsync(s->S.dcs, 0, bt * 2 * sizeof(unsigned long long), st));
            d.S.B = nb;
            d.S.iter0 = (int)(first_iter + b0);
        }
        for (int k = 0; k < 4; ++k) d.in[k] = s->L[cur][k];
        HIPC(hipEventRecord(s->ev0, st));
        if (c->prof) HIPC(hipEventRecord(c->pe[0], st));
        k_ion_init<<<(nb + 63) / 64, 64, 0, st>>>(d, nb, (uint32_t)(first_iter + b0));
        HIPC(hipGetLastError());
        if (c->prof) HIPC(hipEventRecord(c->pe[1], st));
        launches++;
        int pending = 0;          // kernel whose events are still to be read
        int nin[4] = {0, 0, 0, nb};
        for (;;) {
            if (rounds > 0) {
                HIPC(hipMemcpyAsync(hc.data(), s->d_counts + 4 * cur, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
            }
            HIPC(hipStreamSynchronize(st));
            if (c->prof && pending >= 0) {
                float kms = 0.f;
                HIPC(hipEventElapsedTime(&kms, c->pe[0], c->pe[1]));
                c->k_ms[pending] += kms;
                c->k_launches[pending]++;
                pending = -1;
            }
            int ovf = 0;
            HIPC(hipMemcpy(&ovf, s->d_overflow, sizeof(int), hipMemcpyDeviceToHost));
            if (ovf & kFlagNaN) return set_error(c, TRK_ERR_NUMERIC, "a particle state became NaN (iterations %lld..%lld)",
                                                  (long long)(first_iter + b0), (long long)(first_iter + b0 + nb - 1));
            if (ovf & kFlagEvents) return set_error(c, TRK_ERR_CAPACITY, "a particle exceeded %u events", kMaxEvents);
            if (ovf)
                return set_error(c, TRK_ERR_CAPACITY,
                                 "particle list capacity exceeded (flags %d, round %d, batch %lld; capacities e %d h %d p %d ion %d)",
                                 ovf, rounds, (long long)B, s->cap[0], s->cap[1], s->cap[2], s->cap[3]);
            if (++rounds > kMaxGenerations) return set_error(c, TRK_ERR_CAPACITY, "more than %d transport rounds", kMaxGenerations);
            if (rounds > 1)
                for (int k = 0; k < 4; ++k) nin[k] = hc[k] < s->cap[k] ? hc[k] : s->cap[k];
            if (verbose)
                fprintf(stderr, "[trk] batch %lld round %d: electrons %d holes %d photons %d ions %d\n",
                        (long long)b0, rounds, nin[0], nin[1], nin[2], nin[3]);
            int64_t nthr = (int64_t)nin[0] + nin[1] + nin[2] + nin[3];
            if (nthr == 0) break;
            int nxt = 1 - cur;
            HIPC(hipMemsetAsync(s->d_counts + 4 * nxt, 0, 4 * sizeof(int), st));
            for (int k = 0; k < 4; ++k) { d.in[k] = s->L[cur][k]; d.out[k] = s->L[nxt][k]; d.n_in[k] = nin[k]; }
            d.out_cnt = s->d_counts + 4 * nxt;
            int64_t nblk = (nthr + kTransportBlock - 1) / kTransportBlock;
            if (nblk > s->max_blocks) nblk = s->max_blocks;
            d.cursor = s->d_counts + 8;
            HIPC(hipMemsetAsync(d.cursor, 0, sizeof(int), st));
            if (c->prof) HIPC(hipEventRecord(c->pe[0], st));
            k_transport<<<(unsigned)nblk, kTransportBlock, s->lds_bytes, st>>>(d);
            HIPC(hipGetLastError());
            if (c->prof) { HIPC(hipEventRecord(c->pe[1], st)); pending = 1; }
            launches++;
            cur = nxt;
        }
        if (want_spec) {
            const Spec &S = s->S;
            auto blk = [](int64_t n) { return (unsigned)((n + 255) / 256); };
            k_prefix_i<<<blk(nb), 256, 0, st>>>(S.nel, nb, Nt, 1);
            k_prefix_i<<<blk(nb), 256, 0, st>>>(S.nvb, nb, Nt, 1);
            k_prefix_i<<<blk(nb), 256, 0, st>>>(S.dcn, nb, Nt, 1);
            k_prefix_i<<<blk((int64_t)nb * NR), 256, 0, st>>>(S.eve, nb, Nt, NR);
            if (ND > 1) k_prefix_i<<<blk((int64_t)nb * ND), 256, 0, st>>>(S.ehve, nb, Nt, ND);
            k_prefix_i<<<blk((int64_t)nb * 180), 256, 0, st>>>(S.th, nb, Nt, 180);
            k_prefix_i<<<blk((int64_t)nb * 180), 256, 0, st>>>(S.thh, nb, Nt, 180);
            k_prefix_fx<<<blk(nb), 256, 0, st>>>(S.dcs, nb, Nt);
            k_norm_hist<<<blk((int64_t)Nt * NR), 256, 0, st>>>(s->acc_eve, Nt, S.eve, S.nel, nb, Nt, NR, 1, s->d_R);
            if (ND > 1)
                k_norm_hist<<<blk((int64_t)Nt * ND), 256, 0, st>>>(s->acc_ehve, Nt, S.ehve, S.nvb, nb, Nt, ND, 2, c->md.dosE);
            k_norm_hist<<<blk((int64_t)Nt * 180), 256, 0, st>>>(s->acc_th, Nt + 1, S.th, S.nel, nb, Nt, 180, 0, nullptr);
            k_norm_hist<<<blk((int64_t)Nt * 180), 256, 0, st>>>(s->acc_thh, Nt + 1, S.thh, S.nvb, nb, Nt, 180, 0, nullptr);
            k_diff_coeff<<<blk(Nt), 256, 0, st>>>(s->acc_dc, S.dcs, S.dcn, nb, Nt);
            HIPC(hipGetLastError());
        }
        HIPC(hipEventRecord(s->ev1, st));
        HIPC(hipEventSynchronize(s->ev1));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        transport_s += ms * 1e-3;
    }
    HIPC(hipStreamSynchronize(st));
    (void)hipFree(d_sdn);
    (void)hipFree(d_sdoff);

    // counters and trace
    unsigned long long ctr[16];
    HIPC(hipMemcpy(ctr, s->d_counters, sizeof(ctr), hipMemcpyDeviceToHost));
    int64_t tot_ev = 0;
    for (int k = 0; k < 8; ++k) { out->n_events[k] += (int64_t)ctr[k]; tot_ev += (int64_t)ctr[k]; }
    if (c->trace_host && c->trace_cap > 0) {
        unsigned long long tn = 0;
        HIPC(hipMemcpy(&tn, s->d_trace_n, sizeof(tn), hipMemcpyDeviceToHost));
        int64_t ncopy = (int64_t)tn < c->trace_cap ? (int64_t)tn : c->trace_cap;
        HIPC(hipMemcpy(c->trace_host, s->d_trace, sizeof(double) * 12 * ncopy, hipMemcpyDeviceToHost));
        if (c->trace_n_out) *c->trace_n_out = (int64_t)tn;
    }

    // finalize: replicas summed and time prefix taken on the device (exact
    // integers); the host converts, scales by Out_V and forms the totals
    {
        const int64_t ncol = 3 * (int64_t)NR + (int64_t)NR * nsh;
        k_tally_reduce<<<(unsigned)((ncol + 127) / 128), 128, 0, st>>>(T);
        HIPC(hipGetLastError());
    }
    const size_t plane = (size_t)Nt * NR;
    std::vector<unsigned long long> he(plane), hEe(2 * plane), hh(plane * nsh), hEh(2 * plane * nsh),
        hEk(2 * plane * nsh), hp(plane), hEp(2 * plane), hEl(2 * plane);
    HIPC(hipMemcpyAsync(he.data(), T.cnt_e, plane * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEe.data(), T.E_e, plane * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hh.data(), T.cnt_h, plane * nsh * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEh.data(), T.E_h, plane * nsh * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEk.data(), T.Ek_h, plane * nsh * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hp.data(), T.cnt_p, plane * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEp.data(), T.E_p, plane * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEl.data(), T.Elat, plane * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    auto i128 = [](const unsigned long long *p) { return (__int128)(((unsigned __int128)p[1] << 64) | p[0]); };
    const int na = h.m.n_atoms;
    __int128 at_cum = 0;
    std::vector<__int128> shE(nsh);
    for (int i = 0; i < Nt; ++i) {
        __int128 tot_e = 0, tot_Ee = 0, tot_p = 0, tot_Ep = 0, row_lat = 0;
        for (int s2 = 0; s2 < nsh; ++s2) shE[s2] = 0;
        for (int j = 0; j < NR; ++j) {
            const size_t a = (size_t)i * NR + j;
            const long long ne_ = (long long)he[a], np_ = (long long)hp[a];
            const __int128 Ee = i128(&hEe[2 * a]), Ep = i128(&hEp[2 * a]), El = i128(&hEl[2 * a]);
            tot_e += ne_; tot_Ee += Ee; tot_p += np_; tot_Ep += Ep; row_lat += El;
            const double V = Out_V[j];
            const size_t o = (size_t)i + (size_t)Nt * j;
            if (out->Out_ne) out->Out_ne[o] += (double)ne_ * V;
            if (out->Out_Ee) out->Out_Ee[o] += fx_to_double(Ee) * V;
            if (out->Out_nphot) out->Out_nphot[o] += (double)np_ * V;
            if (out->Out_Ephot) out->Out_Ephot[o] += fx_to_double(Ep) * V;
            if (out->Out_Elat) out->Out_Elat[o] += fx_to_double(El) * V;
            for (int s2 = 0; s2 < nsh; ++s2) {
                const size_t q = a * nsh + s2;
                const long long nh_ = (long long)hh[q];
                const __int128 Eh = i128(&hEh[2 * q]), Ek = i128(&hEk[2 * q]);
                shE[s2] += Eh + Ek;
                int at = c->hm.atom_of[s2], sl = s2 - c->hm.shl_off[at];
                size_t oh = (size_t)i + (size_t)Nt * (j + (size_t)NR * (at + (size_t)na * sl));
                if (out->Out_nh) out->Out_nh[oh] += (double)nh_ * V;
                if (out->Out_Eh) out->Out_Eh[oh] += fx_to_double(Eh) * V;
                if (out->Out_Ehkin) out->Out_Ehkin[oh] += fx_to_double(Ek) * V;
            }
        }
        at_cum += row_lat;
        __int128 hsum = 0;
        for (int s2 = 0; s2 < nsh; ++s2) {
            hsum += shE[s2];
            int at = c->hm.atom_of[s2], sl = s2 - c->hm.shl_off[at];
            if (out->Out_E_h) out->Out_E_h[(size_t)i + (size_t)Nt * (at + (size_t)na * sl)] += fx_to_double(shE[s2]);
        }
        if (out->Out_tot_Ne) out->Out_tot_Ne[i] += (double)(long long)tot_e;
        if (out->Out_E_e) out->Out_E_e[i] += fx_to_double(tot_Ee);
        if (out->Out_E_at) out->Out_E_at[i] += fx_to_double(at_cum);
        if (h.o.include_photons) {
            if (out->Out_tot_Nphot) out->Out_tot_Nphot[i] += (double)(long long)tot_p;
            if (out->Out_E_phot) out->Out_E_phot[i] += fx_to_double(tot_Ep);
        }
        if (out->Out_tot_E) out->Out_tot_E[i] += fx_to_double(tot_Ee + hsum + (h.o.include_photons ? tot_Ep : 0) + at_cum);
    }
    if (want_spec) {
        std::vector<double> buf;
        auto add_back = [&](double *dst, const double *src, size_t n) -> int {
            if (!dst) return 0;
            buf.resize(n);
            HIPC(hipMemcpy(buf.data(), src, n * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < n; ++k) dst[k] += buf[k];
            return 0;
        };
        if (add_back(out->Out_Ee_vs_E, s->acc_eve, (size_t)Nt * NR)) return TRK_ERR_DEVICE;
        if (ND > 1 && add_back(out->Out_Eh_vs_E, s->acc_ehve, (size_t)Nt * ND)) return TRK_ERR_DEVICE;
        if (add_back(out->Out_theta, s->acc_th, (size_t)(Nt + 1) * 180)) return TRK_ERR_DEVICE;
        if (add_back(out->Out_theta_h, s->acc_thh, (size_t)(Nt + 1) * 180)) return TRK_ERR_DEVICE;
        if (out->Out_diff_coeff)
            HIPC(hipMemcpy(out->Out_diff_coeff, s->acc_dc, (size_t)Nt * sizeof(double), hipMemcpyDeviceToHost));
    }
    c->last_transport_s = transport_s;
    c->last_total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tstart).count();
    c->last_launches = launches;
    (void)tot_ev;
    return 0;
}

}  // namespace trk
