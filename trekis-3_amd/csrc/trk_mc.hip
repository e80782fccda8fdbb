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
#include <mutex>
#include <vector>
#include "trk_mc.h"

namespace trk {


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

// The energy grid of the emitted electrons' spectrum, Out_E = Out_R/10 (MC:940-941).
__global__ void k_tenth(double *RE, const double *R, int NR)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < NR) RE[j] = R[j] / 10.0;
}

// Prefix over time of the emitted electrons' kinetic energies [B][Nt][2]:
// the split words of each node are combined and the running sum is left as
// one 128-bit integer (low, high) in units of 2^-40 eV.
__global__ void k_prefix_fx128(unsigned long long *a, int B, int Nt)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    unsigned long long *p = a + (size_t)b * Nt * 2;
    __int128 run = 0;
    for (int i = 0; i < Nt; ++i) {
        run += fx_read(p + 2 * i);
        i128_store(p + 2 * i, run);
    }
}

// Out_Ne_Em and Out_E_Em (MC:958-959): per node, the iterations' emitted
// counts (after their prefix) and kinetic energies (after k_prefix_fx128)
// summed exactly, then added to the accumulators.
__global__ void k_emit_sums(double *acc_n, double *acc_e, const int *n, const unsigned long long *e, int B, int Nt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt) return;
    long long cn = 0;
    __int128 ce = 0;
    for (int b = 0; b < B; ++b) {
        cn += n[(size_t)b * Nt + i];
        ce += i128_load(e + 2 * ((size_t)b * Nt + i));
    }
    const long long hi = (long long)(ce >> 40);
    const __int128 lo = ce - ((__int128)hi << 40);
    acc_n[i] += (double)cn;
    acc_e[i] += (double)hi + (double)(long long)lo * 0x1.0p-40;
}

// acc(i, j) += sum over iterations of count * (w_j / norm), in iteration order.
// wmode 0: w = 1 (angle histograms); 1: w from the radial grid used as an
// energy grid (MC:1024-1029); 2: w from the DOS grid (MC:1033-1040).
__global__ void k_norm_hist(double *acc, int accld, const int *cnt, const int *norm, int B, int Nt, int ncol, int wmode,
                            const double *grid)
{
    // neighbouring threads take neighbouring columns j of the same node i:
    // the per-iteration rows cnt[b][i][:] are read coalesced
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Nt * ncol) return;
    int j = (int)(t % ncol), i = (int)(t / ncol);
    double w = 1.0;
    if (wmode == 1) w = (j > 0) ? 1.0 / (grid[j] - grid[j - 1]) : 1.0 / grid[0];
    else if (wmode == 2) w = (j > 0) ? 1.0 / (grid[j] - grid[j - 1]) : 1.0 / (grid[1] - grid[0]);
    double sum = acc[(size_t)i + (size_t)accld * j];
    // iterations in order; their counts loaded eight at a time ahead of the sum
    for (int b0 = 0; b0 < B; b0 += 8) {
        int c[8], n[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            c[k] = 0;
            n[k] = 1;
            if (b0 + k < B) {
                c[k] = cnt[((size_t)(b0 + k) * Nt + i) * ncol + j];
                n[k] = norm[(size_t)(b0 + k) * Nt + i];
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (c[k]) sum = sum + (double)c[k] * (w / (double)n[k]);
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
    // the recurrence runs over the iterations in order; their counts and
    // sums are loaded sixteen at a time ahead of it
    for (int b0 = 0; b0 < B; b0 += 16) {
        int n[16];
        long long x[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            n[k] = 0;
            x[k] = 0;
            if (b0 + k < B) {
                n[k] = dcn[(size_t)(b0 + k) * Nt + i];
                x[k] = (long long)dcs[((size_t)(b0 + k) * Nt + i) * 2];
            }
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (n[k] > 0) {
                long long hi = x[k] >> 40;
                double s = (double)hi + (double)(x[k] - (hi << 40)) * 0x1.0p-40;
                v = (v + s) / (double)n[k];
            }
        }
    }
    D[i] = v;
}

// ---------------------------------------------------------------- tally reduction
// Sums the R replicas of every tally cell and takes the prefix over time of
// the difference arrays, in exact integer arithmetic, leaving the result in
// replica 0: counts as int64, energies as int128 (low word, high word) in
// units of 2^-40 eV.  Two kernels: the replica sum per cell, then the running
// sum per column (electrons and photons by radial bin, holes by radial bin
// and shell; the lattice energy per node has no prefix).
// (1) one thread per cell: the replicas summed into replica 0 (counts as
// int64, energies as int128)
__global__ void k_replica_sum(Tally T)
{
    const int NR = T.NR, Nt = T.Nt, nsh = T.nsh, R = T.R;
    const size_t plane = (size_t)Nt * NR, hplane = plane * nsh;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2 * (int64_t)plane) {                       // electrons, photons
        const size_t a = (size_t)(t < (int64_t)plane ? t : t - plane);
        unsigned long long *cnt = (t < (int64_t)plane) ? T.cnt_e : T.cnt_p;
        unsigned long long *en = (t < (int64_t)plane) ? T.E_e : T.E_p;
        long long sc = 0;
        __int128 se = 0;
        for (int r = 0; r < R; ++r) {
            sc += (long long)cnt[r * plane + a];
            se += fx_read(en + 2 * (r * plane + a));
        }
        cnt[a] = (unsigned long long)sc;
        i128_store(en + 2 * a, se);
    } else if (t < 3 * (int64_t)plane) {                // lattice energy
        const size_t a = (size_t)(t - 2 * plane);
        __int128 se = 0;
        for (int r = 0; r < R; ++r) se += fx_read(T.Elat + 2 * (r * plane + a));
        i128_store(T.Elat + 2 * a, se);
    } else if (t < 3 * (int64_t)plane + (int64_t)hplane) {   // holes
        const size_t a = (size_t)(t - 3 * plane);
        long long sc = 0;
        __int128 sE = 0, sK = 0;
        for (int r = 0; r < R; ++r) {
            sc += (long long)T.cnt_h[r * hplane + a];
            sE += fx_read(T.E_h + 2 * (r * hplane + a));
            sK += fx_read(T.Ek_h + 2 * (r * hplane + a));
        }
        T.cnt_h[a] = (unsigned long long)sc;
        i128_store(T.E_h + 2 * a, sE);
        i128_store(T.Ek_h + 2 * a, sK);
    }
}

// (2) one thread per column: the running sum over the time nodes of the
// difference arrays in replica 0 (the lattice energy is not differenced)
__global__ void k_time_prefix(Tally T)
{
    const int NR = T.NR, Nt = T.Nt, nsh = T.nsh;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2 * NR) {
        const int j = (int)(t % NR);
        unsigned long long *cnt = (t < NR) ? T.cnt_e : T.cnt_p;
        unsigned long long *en = (t < NR) ? T.E_e : T.E_p;
        long long rc = 0;
        __int128 re = 0;
#pragma unroll 8
        for (int i = 0; i < Nt; ++i) {
            const size_t a = (size_t)i * NR + j;
            rc += (long long)cnt[a];
            re += i128_load(en + 2 * a);
            cnt[a] = (unsigned long long)rc;
            i128_store(en + 2 * a, re);
        }
    } else if (t < 2 * NR + (int64_t)NR * nsh) {
        const int64_t q = t - 2 * NR;
        const int j = (int)(q / nsh), sh = (int)(q % nsh);
        long long rc = 0;
        __int128 rE = 0, rK = 0;
#pragma unroll 8
        for (int i = 0; i < Nt; ++i) {
            const size_t a = ((size_t)i * NR + j) * nsh + sh;
            rc += (long long)T.cnt_h[a];
            rE += i128_load(T.E_h + 2 * a);
            rK += i128_load(T.Ek_h + 2 * a);
            T.cnt_h[a] = (unsigned long long)rc;
            i128_store(T.E_h + 2 * a, rE);
            i128_store(T.Ek_h + 2 * a, rK);
        }
    }
}

// ---------------------------------------------------------------- host side
// Device buffers of the transport, allocated once per context and kept
// (grown when a run needs more).
struct McState {
    PList L[2][4];            // [ping-pong][electrons, holes, photons, ions]
    int cap[4] = {0, 0, 0, 0};
    int *d_counts = nullptr;  // [16]: list counts of round A [4], of round B [4], transport cursor
    int64_t max_blocks = 0;   // resident transport workgroups on the device
    int ncu = 0, max_lds = 0;
    size_t lds_bytes = 0;     // staged tables per transport workgroup
    bool all_staged = false;  // every staged array fits (the k_transport<true> instantiation runs)
    int *d_overflow = nullptr;
    unsigned long long *d_counters = nullptr;
    Tally T{};
    size_t tally_bytes = 0;
    int NR_alloc = 0, Nt_alloc = 0, R_alloc = 0;
    double *d_tim = nullptr, *d_R = nullptr;   // time nodes and radii of the run
    int tim_cap = 0, R_cap = 0;
    double *d_trace = nullptr;
    unsigned long long *d_trace_n = nullptr;
    int64_t trace_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // per-iteration tallies (allocated for spec_B iterations) and their accumulators
    Spec S{};
    int spec_B = 0, spec_Nt = 0, spec_NR = 0, spec_ND = 0;
    double *acc_eve = nullptr, *acc_ehve = nullptr, *acc_th = nullptr, *acc_thh = nullptr, *acc_dc = nullptr;
    // emitted electrons per iteration ([B][Nt] counts, [B][Nt][2] energies,
    // [B][Nt][NR] spectrum), their accumulators and the spectrum's grid
    int *em_n = nullptr, *em_h = nullptr;
    unsigned long long *em_e = nullptr;
    double *acc_emn = nullptr, *acc_eme = nullptr, *acc_emh = nullptr;
    double *d_RE = nullptr;
    int RE_cap = 0;
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
                  s->acc_eve, s->acc_ehve, s->acc_th, s->acc_thh, s->acc_dc,
                  s->em_n, s->em_h, s->em_e, s->acc_emn, s->acc_eme, s->acc_emh};
    for (void *p : ps) if (p) (void)hipFree(p);
    s->S = Spec{};
    s->acc_eve = s->acc_ehve = s->acc_th = s->acc_thh = s->acc_dc = nullptr;
    s->em_n = s->em_h = nullptr;
    s->em_e = nullptr;
    s->acc_emn = s->acc_eme = s->acc_emh = nullptr;
    s->spec_B = 0;
}

void mc_free(McState *s)
{
    if (!s) return;
    free_spec(s);
    for (auto &a : s->L) for (auto &l : a) free_list(l);
    void *ps[] = {s->d_counts, s->d_overflow, s->d_counters, s->T.cnt_e, s->T.E_e, s->T.cnt_h, s->T.E_h,
                  s->T.Ek_h, s->T.cnt_p, s->T.E_p, s->T.Elat, s->d_tim, s->d_R, s->d_RE, s->d_trace, s->d_trace_n};
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
    s->T = Tally{};
    size_t cell = (size_t)R * Nt * NR;
    Tally &T = s->T;
    HIPC(hipMalloc((void **)&T.cnt_e, cell * 8));
    HIPC(hipMalloc((void **)&T.E_e, cell * 16));
    HIPC(hipMalloc((void **)&T.cnt_h, cell * nsh * 8));
    HIPC(hipMalloc((void **)&T.E_h, cell * nsh * 16));
    HIPC(hipMalloc((void **)&T.Ek_h, cell * nsh * 16));
    HIPC(hipMalloc((void **)&T.cnt_p, cell * 8));
    HIPC(hipMalloc((void **)&T.E_p, cell * 16));
    HIPC(hipMalloc((void **)&T.Elat, cell * 16));
    T.Nt = Nt; T.NR = NR; T.nsh = nsh; T.R = R;
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
    HIPC(hipMalloc((void **)&s->em_n, bt * sizeof(int)));
    HIPC(hipMalloc((void **)&s->em_h, bt * NR * sizeof(int)));
    HIPC(hipMalloc((void **)&s->em_e, bt * 2 * sizeof(unsigned long long)));
    HIPC(hipMalloc((void **)&s->acc_emn, (size_t)Nt * sizeof(double)));
    HIPC(hipMalloc((void **)&s->acc_eme, (size_t)Nt * sizeof(double)));
    HIPC(hipMalloc((void **)&s->acc_emh, (size_t)Nt * NR * sizeof(double)));
    s->spec_B = B; s->spec_Nt = Nt; s->spec_NR = NR; s->spec_ND = ND;
    return 0;
}

// the per-iteration normalised tallies (and the emitted electrons', counted
// with them) are taken when the caller asks for any of them
static bool wants_per_iteration(const trk_ctx *c, const trk_tallies *out)
{
    const bool em = c->hm.m.work_function > 0.0 && (out->Out_Ne_Em || out->Out_E_Em || out->Out_Ee_vs_E_Em);
    return em || out->Out_Ee_vs_E || out->Out_Eh_vs_E || out->Out_theta || out->Out_theta_h || out->Out_diff_coeff;
}

// grow a device array of doubles to hold n values
static int ensure_doubles(trk_ctx *c, double **p, int *have, int n)
{
    if (*p && *have >= n) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    HIPC(hipMalloc((void **)p, sizeof(double) * (size_t)(n > 0 ? n : 1)));
    *have = n;
    return 0;
}

static int device_info(trk_ctx *c, McState *s)
{
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    HIPC(hipDeviceGetAttribute(&s->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPC(hipDeviceGetAttribute(&s->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return 0;
}

void fill_mcdev(const trk_ctx *c, McDev &d)
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
    d.rmono = T[TRK_TAB_E_INEL].row_mono; d.rmono_h = T[TRK_TAB_H_INEL].row_mono;
    d.rmono_el = T[TRK_TAB_E_EL].row_mono; d.rmono_hel = T[TRK_TAB_H_EL].row_mono;
    d.nP = T[TRK_TAB_PHOT].n; d.EP = T[TRK_TAB_PHOT].E; d.LPsh = T[TRK_TAB_PHOT].L; d.phot_imfp = ts.phot_imfp;
    d.nS = T[TRK_TAB_SHI].n; d.ES = T[TRK_TAB_SHI].E; d.LSsh = T[TRK_TAB_SHI].L; d.shi_path = ts.shi_path;
    const TableDev &td = T[TRK_TAB_SHI_DIFF];
    d.sd_n = ts.sd_n; d.sd_off = ts.sd_off; d.sd_E = td.dE; d.sd_L = td.dL; d.sd_inv = td.dInv;
    d.sd_nonmono = td.dNonMono; d.sd_M = ts.sd_M; d.sd_c = ts.sd_c; d.sd_bk = ts.sd_bk;
    d.lnE = ts.ln_E[TRK_TAB_E_INEL]; d.ln_el_imfp = ts.ln_el_imfp;
    d.lnEel = ts.ln_E[TRK_TAB_E_EL]; d.ln_Lel = ts.ln_Lel;
    d.lnEH = ts.ln_E[TRK_TAB_H_INEL]; d.ln_hole_imfp = ts.ln_hole_imfp;
    d.lnEHel = ts.ln_E[TRK_TAB_H_EL]; d.ln_LHel = ts.ln_LHel;
    d.ln_Lsh = ts.ln_L[TRK_TAB_E_INEL]; d.ln_LHsh = ts.ln_L[TRK_TAB_H_INEL];
    d.ln_rhw = ts.ln_hw[TRK_TAB_E_INEL]; d.ln_rlam = ts.ln_lam[TRK_TAB_E_INEL];
    d.ln_rhw_h = ts.ln_hw[TRK_TAB_H_INEL]; d.ln_rlam_h = ts.ln_lam[TRK_TAB_H_INEL];
    d.ln_rhw_el = ts.ln_hw[TRK_TAB_E_EL]; d.ln_rlam_el = ts.ln_lam[TRK_TAB_E_EL];
    d.ln_rhw_hel = ts.ln_hw[TRK_TAB_H_EL]; d.ln_rlam_hel = ts.ln_lam[TRK_TAB_H_EL];
    d.lnES = ts.ln_E[TRK_TAB_SHI]; d.ln_LSsh = ts.ln_L[TRK_TAB_SHI]; d.ln_shi_path = ts.ln_shi_path;
    const MatBuffers &mb = (c->base ? c->base : c)->mb;
    d.delta_alpha = mb.oAlpha;
    d.beb_kocs = mb.kocs;
    d.beb_Ek = mb.Ek;
    d.dsf[0] = ts.dsf[0];
    d.dsf[1] = ts.dsf[1];
    d.em_W = d.em_H = d.em_E1 = d.em_gamma = 0.0;
    const trk_material &m = c->hm.m;
    if (m.work_function > 0.0 &&
        trk_barrier_parameters(m.work_function, m.bar_length, m.bar_height, &d.em_E1, &d.em_gamma) == 0) {
        d.em_W = m.work_function;
        d.em_H = m.bar_height;
    }
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

// The tallies of plane `plane` of T (replicas summed and time prefix taken
// on the device, exact integers): converted, scaled by Out_V and added into
// the caller's Out_* arrays, with the totals per time node formed here.
int finish_tallies_host(trk_ctx *c, const Tally &T, int plane_index, const double *Out_V, int NR, trk_tallies *out,
                        hipStream_t st)
{
    const HostMat &h = c->hm;
    const int Nt = h.Nt, nsh = h.nsh;
    const size_t plane = (size_t)Nt * NR;
    const size_t p0 = (size_t)plane_index * plane;
    std::vector<unsigned long long> he(plane), hEe(2 * plane), hh(plane * nsh), hEh(2 * plane * nsh),
        hEk(2 * plane * nsh), hp(plane), hEp(2 * plane), hEl(2 * plane);
    HIPC(hipMemcpyAsync(he.data(), T.cnt_e + p0, plane * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEe.data(), T.E_e + 2 * p0, plane * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hh.data(), T.cnt_h + p0 * nsh, plane * nsh * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEh.data(), T.E_h + 2 * p0 * nsh, plane * nsh * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEk.data(), T.Ek_h + 2 * p0 * nsh, plane * nsh * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hp.data(), T.cnt_p + p0, plane * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEp.data(), T.E_p + 2 * p0, plane * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(hEl.data(), T.Elat + 2 * p0, plane * 16, hipMemcpyDeviceToHost, st));
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
    return 0;
}

// Each generation is the set of particles born in the previous one; a chain of
// ionisations is at most a few hundred long.
constexpr int kMaxGenerations = 1 << 16;

// Iterations per batch: the ion's expected ionisations over the layer times
// the pairs each makes must fit one generation's lists.
static int64_t batch_size(const trk_ctx *c, int cap, int64_t nmc)
{
    const double lam = c->ts.shi_path0;
    const double est_pairs = (lam > 0.0) ? c->hm.m.Layer / lam : 1e4;
    int64_t B = (c->hm.o.batch > 0) ? c->hm.o.batch : (int64_t)((double)cap / (3.0 * est_pairs + 64.0));
    if (B < 1) B = 1;
    if (B > nmc) B = nmc;
    return B;
}

// LDS layout of the tables a transport workgroup stages (in this order while
// they fit in `limit` doubles): sets d.lds_n / d.lds_off, the bytes in
// *bytes, and returns whether every array is staged.
bool plan_lds(McDev &d, int Nt, int NR, int n_dos, int nsh, size_t limit, size_t *bytes)
{
    const int nsz[kStaged] = {d.nE, d.nE, d.nE, d.nE, d.nEel, d.nEel, d.nEel, d.nEel,
                              d.nH, d.nH, d.nH, d.nH, d.nHel, d.nHel, d.nHel, d.nHel,
                              NR, Nt, n_dos, n_dos, n_dos, n_dos, nsh * d.nE, nsh * d.nH,
                              nsh * d.nE, nsh * d.nH};
    const bool present[kStaged] = {true, true, d.lnE != nullptr, d.ln_el_imfp != nullptr, true, true,
                                   d.lnEel != nullptr, d.ln_Lel != nullptr, true, true, d.lnEH != nullptr,
                                   d.ln_hole_imfp != nullptr, true, true, d.lnEHel != nullptr,
                                   d.ln_LHel != nullptr, true, true, true, true, true, true, true, true,
                                   d.ln_Lsh != nullptr, d.ln_LHsh != nullptr};
    size_t off = 0;
    bool all = true;
    for (int a = 0; a < kStaged; ++a) {
        size_t n = present[a] ? (size_t)nsz[a] : 0;
        if (off + n > limit) n = 0;
        if (!present[a] || n != (size_t)nsz[a] || n == 0) all = false;
        d.lds_n[a] = (int)n;
        d.lds_off[a] = (int)off;
        off += n;
    }
    *bytes = off * sizeof(double);
    return all;
}

// k_shi_consts then k_shi_buckets of the path's precision on a table set.
hipError_t launch_shi_consts(const trk_ctx *c, const IonDev &iv, bool mixed, hipStream_t st)
{
    const int nsh = c->hm.nsh;
    if (mixed) {
        mx::k_shi_consts<<<(nsh + 63) / 64, 64, 0, st>>>(iv, c->md.Ip, nsh, c->ts.sd_M, c->ts.sd_c);
        mx::k_shi_buckets<<<nsh, 128, 0, st>>>(iv, nsh, c->ts.sd_M, c->ts.sd_c, c->ts.sd_bk);
    } else {
        fp64::k_shi_consts<<<(nsh + 63) / 64, 64, 0, st>>>(iv, c->md.Ip, nsh, c->ts.sd_M, c->ts.sd_c);
        fp64::k_shi_buckets<<<nsh, 128, 0, st>>>(iv, nsh, c->ts.sd_M, c->ts.sd_c, c->ts.sd_bk);
    }
    return hipGetLastError();
}

// The ion of a context as the event code reads it.
IonDev ion_view_host(const trk_ctx *c)
{
    const TableSet &ts = c->ts;
    const TableDev &td = ts.tab[TRK_TAB_SHI_DIFF];
    IonDev v{};
    v.mass = c->md.ion_mass; v.E0 = c->md.ion_E; v.Zeff0 = c->md.ion_Zeff0; v.fixed_zeff = c->md.ion_fixed_zeff;
    v.Z = c->md.ion_Z; v.kind_zeff = c->md.ion_kind_zeff; v.nS = ts.tab[TRK_TAB_SHI].n;
    v.ES = ts.tab[TRK_TAB_SHI].E; v.LSsh = ts.tab[TRK_TAB_SHI].L; v.shi_path = ts.shi_path;
    v.lnES = ts.ln_E[TRK_TAB_SHI]; v.ln_LSsh = ts.ln_L[TRK_TAB_SHI]; v.ln_shi_path = ts.ln_shi_path;
    v.sd_n = ts.sd_n; v.sd_off = ts.sd_off; v.sd_E = td.dE; v.sd_L = td.dL; v.sd_inv = td.dInv;
    v.sd_nonmono = td.dNonMono; v.sd_M = ts.sd_M; v.sd_c = ts.sd_c; v.sd_bk = ts.sd_bk;
    return v;
}

// Buffers of the device finalize used by runs without allocation: the
// caller's Out_* arrays staged in device memory, the radial volumes and the
// per-node integer sums.
struct NoAllocBufs {
    double *stage = nullptr;
    size_t stage_n = 0;
    double *d_V = nullptr;
    unsigned long long *node_sums = nullptr;   // [Nt][4]: lattice energy of the node, the rest of tot_E
};

static int finalize_on_device(trk_ctx *c, McState *s, const NoAllocBufs &na, const double *Out_V, int NR,
                              trk_tallies *out, bool want_spec);

// the transport kernel instantiation with the general transfers runs when a
// shell uses the BEB cross section, the Delta-CDF model is on or elastic
// scattering comes from DSF tables
bool uses_general_kernel(const HostMat &h)
{
    if (h.o.kind_of_DR == 4 || h.o.kind_of_EMFP == 2) return true;
    for (int s = 0; s < h.nsh; ++s)
        if (h.m.kocs[s] == 2) return true;
    return false;
}

// The transport kernel of a run: precision, all tables staged in LDS, the
// general inelastic transfers (BEB shells, the Delta-CDF model)
static const void *transport_kernel(bool mixed, bool all, bool general)
{
    if (general)
        return mixed ? (all ? (const void *)mx::k_transport<true, true> : (const void *)mx::k_transport<false, true>)
                     : (all ? (const void *)fp64::k_transport<true, true> : (const void *)fp64::k_transport<false, true>);
    return mixed ? (all ? (const void *)mx::k_transport<true, false> : (const void *)mx::k_transport<false, false>)
                 : (all ? (const void *)fp64::k_transport<true, false> : (const void *)fp64::k_transport<false, false>);
}

// The run itself on buffers the caller of this function provides in s (and
// na, for a run without allocation: the tallies are then finished on the
// device instead of the host).
static int run_core(trk_ctx *c, McState *s, int64_t nmc, uint64_t seed, int64_t first_iter, const double *Out_R,
                    const double *Out_V, int NR, trk_tallies *out, int64_t B, bool want_spec,
                    const NoAllocBufs *nab = nullptr)
{
    auto tstart = std::chrono::steady_clock::now();
    const bool mixed = c->hm.o.precision == 1;   // FP32 elementary functions in the event code
    const bool general = uses_general_kernel(c->hm);   // BEB shells, the Delta-CDF model or DSF elastic scattering
    const HostMat &h = c->hm;
    const int Nt = h.Nt, nsh = h.nsh;
    hipStream_t st = c->stream;
    HIPC(hipMemcpyAsync(s->d_tim, h.tim_glob.data(), sizeof(double) * Nt, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(s->d_R, Out_R, sizeof(double) * NR, hipMemcpyHostToDevice, st));
    Tally &T = s->T;
    const int R = T.R;
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

    McDev d{};
    fill_mcdev(c, d);
    // per-shell constants of shi_transfer and the index of its rows by draw,
    // computed once per table set with the path's own arithmetic (the FP64
    // path still evaluates the constants in every event, to the same bits)
    if (!c->ts.sd_consts) {
        HIPC(launch_shi_consts(c, ion_view_host(c), mixed, st));
        c->ts.sd_consts = true;
    }
    d.seed = seed;
    d.Tim = h.o.Tim;
    d.T_last = h.tim_glob[Nt - 1];
    d.cutoff_eff = (h.m.cut_off > 0.0) ? h.m.cut_off : 0.0;
    d.Nt = Nt; d.tim_glob = s->d_tim;
    d.inv_dt = (h.o.dt_flag <= 0 && h.o.dt > 0.0) ? 1.0 / h.o.dt : 0.0;
    {
        // a particle's last bracket is tried before a search only on strictly
        // increasing grids, where it gives the search's own result: the
        // energy grids (checked when the tables were built), the time grid
        // (a flight with no node in (t0, tn] is recognised from the first
        // node after t0) and the radii Out_R
        d.hints = c->ts.grids_increasing ? 1 : 0;
        for (int i = 1; i < Nt; ++i)
            if (!(h.tim_glob[i - 1] < h.tim_glob[i])) d.hints = 0;
        for (int i = 1; i < NR; ++i)
            if (!(Out_R[i - 1] < Out_R[i])) d.hints = 0;
    }
    d.NR = NR; d.R = s->d_R;
    d.include_photons = h.o.include_photons;
    d.overflow = s->d_overflow;
    d.counters = s->d_counters;
    d.T = T;
    if (c->trace_host && c->trace_cap > 0) {
        HIPC(hipMemsetAsync(s->d_trace_n, 0, sizeof(unsigned long long), st));
        d.trace = s->d_trace; d.trace_n = s->d_trace_n; d.trace_cap = c->trace_cap;
    }
    for (int k = 0; k < 4; ++k) d.cap[k] = s->cap[k];

    // per-iteration normalised tallies, when the caller asks for any of them
    const int ND = h.m.n_dos;
    d.S = Spec{};
    if (want_spec) {
        HIPC(hipMemsetAsync(s->acc_emn, 0, (size_t)Nt * sizeof(double), st));
        HIPC(hipMemsetAsync(s->acc_eme, 0, (size_t)Nt * sizeof(double), st));
        HIPC(hipMemsetAsync(s->acc_emh, 0, (size_t)Nt * NR * sizeof(double), st));
        k_tenth<<<(NR + 255) / 256, 256, 0, st>>>(s->d_RE, s->d_R, NR);
        d.RE = s->d_RE; d.em_n = s->em_n; d.em_h = s->em_h; d.em_e = s->em_e;
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
        size_t limit = (size_t)s->max_lds / sizeof(double);
        // TRK_LDS_LIMIT_BYTES lowers the budget (tests of the partly staged kernel)
        if (const char *e = getenv("TRK_LDS_LIMIT_BYTES")) {
            const long long v = atoll(e);
            if (v >= 0 && (size_t)v / sizeof(double) < limit) limit = (size_t)v / sizeof(double);
        }
        const bool all = plan_lds(d, Nt, NR, h.m.n_dos, nsh, limit, &s->lds_bytes);
        s->all_staged = all;
        const void *kt = transport_kernel(mixed, all, general);
        if (s->lds_bytes > 65536)
            HIPC(hipFuncSetAttribute(kt, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->lds_bytes));
        int per_cu = 0;
        HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kt, mixed ? kTransportBlockMx : kTransportBlockFp64,
                                                          s->lds_bytes));
        if (per_cu < 1) per_cu = 1;
        s->max_blocks = (int64_t)per_cu * s->ncu;
    }

    double transport_s = 0.0;
    int64_t launches = 0;
    int hc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool verbose = getenv("TRK_VERBOSE") != nullptr;   // per-round list sizes on stderr
    for (int64_t b0 = 0; b0 < nmc; b0 += B) {
        int nb = (int)((nmc - b0 < B) ? nmc - b0 : B);
        int cur = 0, rounds = 0;
        auto tb0 = std::chrono::steady_clock::now();
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
            HIPC(hipMemsetAsync(s->em_n, 0, bt * sizeof(int), st));
            HIPC(hipMemsetAsync(s->em_h, 0, bt * NR * sizeof(int), st));
            HIPC(hipMemsetAsync(s->em_e, 0, bt * 2 * sizeof(unsigned long long), st));
            d.S.B = nb;
            d.S.iter0 = (int)(first_iter + b0);
        }
        for (int k = 0; k < 4; ++k) d.in[k] = s->L[cur][k];
        if (s->ev0) HIPC(hipEventRecord(s->ev0, st));
        if (c->prof) HIPC(hipEventRecord(c->pe[0], st));
        if (mixed) mx::k_ion_init<<<(nb + 63) / 64, 64, 0, st>>>(d, nb, (uint32_t)(first_iter + b0));
        else fp64::k_ion_init<<<(nb + 63) / 64, 64, 0, st>>>(d, nb, (uint32_t)(first_iter + b0));
        HIPC(hipGetLastError());
        if (c->prof) HIPC(hipEventRecord(c->pe[1], st));
        launches++;
        int pending = 0;          // kernel whose events are still to be read
        int nin[4] = {0, 0, 0, nb};
        for (;;) {
            if (rounds > 0) {
                HIPC(hipMemcpyAsync(hc, s->d_counts + 4 * cur, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
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
            const int tb = mixed ? kTransportBlockMx : kTransportBlockFp64;
            int64_t nblk = (nthr + tb - 1) / tb;
            int tbl = tb;             // threads per workgroup of this launch
            if (nblk < s->max_blocks) {
                // few particles (the ions' first round, the last rounds):
                // spread them over the resident workgroups, one or a few
                // waves each, rather than full workgroups on a few CUs (the
                // kernel stages its tables with blockDim.x threads)
                const int64_t waves = (nthr + 63) / 64;
                const int64_t w = (waves + s->max_blocks - 1) / s->max_blocks;
                tbl = (int)(64 * w);
                nblk = (nthr + tbl - 1) / tbl;
            }
            if (nblk > s->max_blocks) nblk = s->max_blocks;
            d.cursor = s->d_counts + 8;
            HIPC(hipMemsetAsync(d.cursor, 0, sizeof(int), st));
            if (c->prof) HIPC(hipEventRecord(c->pe[0], st));
            const unsigned g = (unsigned)nblk;
            const size_t lds = s->lds_bytes;
            if (general) {
                if (mixed && s->all_staged) mx::k_transport<true, true><<<g, tbl, lds, st>>>(d);
                else if (mixed) mx::k_transport<false, true><<<g, tbl, lds, st>>>(d);
                else if (s->all_staged) fp64::k_transport<true, true><<<g, tbl, lds, st>>>(d);
                else fp64::k_transport<false, true><<<g, tbl, lds, st>>>(d);
            } else {
                if (mixed && s->all_staged) mx::k_transport<true, false><<<g, tbl, lds, st>>>(d);
                else if (mixed) mx::k_transport<false, false><<<g, tbl, lds, st>>>(d);
                else if (s->all_staged) fp64::k_transport<true, false><<<g, tbl, lds, st>>>(d);
                else fp64::k_transport<false, false><<<g, tbl, lds, st>>>(d);
            }
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
            if (d.em_W > 0.0) {
                k_prefix_i<<<blk(nb), 256, 0, st>>>(s->em_n, nb, Nt, 1);
                k_prefix_i<<<blk((int64_t)nb * NR), 256, 0, st>>>(s->em_h, nb, Nt, NR);
                k_prefix_fx128<<<blk(nb), 256, 0, st>>>(s->em_e, nb, Nt);
                k_norm_hist<<<blk((int64_t)Nt * NR), 256, 0, st>>>(s->acc_emh, Nt, s->em_h, s->em_n, nb, Nt, NR, 1, s->d_RE);
                k_emit_sums<<<blk(Nt), 256, 0, st>>>(s->acc_emn, s->acc_eme, s->em_n, s->em_e, nb, Nt);
            }
            HIPC(hipGetLastError());
        }
        if (s->ev1) {
            HIPC(hipEventRecord(s->ev1, st));
            HIPC(hipEventSynchronize(s->ev1));
            float ms = 0.f;
            HIPC(hipEventElapsedTime(&ms, s->ev0, s->ev1));
            transport_s += ms * 1e-3;
        } else {
            HIPC(hipStreamSynchronize(st));
            transport_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - tb0).count();
        }
    }
    HIPC(hipStreamSynchronize(st));

    // counters and trace
    unsigned long long ctr[16];
    HIPC(hipMemcpy(ctr, s->d_counters, sizeof(ctr), hipMemcpyDeviceToHost));
    for (int k = 0; k < 8; ++k) out->n_events[k] += (int64_t)ctr[k];
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
        const int64_t ncell = (3 + (int64_t)nsh) * (int64_t)Nt * NR;
        k_replica_sum<<<(unsigned)((ncell + 255) / 256), 256, 0, st>>>(T);
        const int64_t ncol = 2 * (int64_t)NR + (int64_t)NR * nsh;
        k_time_prefix<<<(unsigned)((ncol + 63) / 64), 64, 0, st>>>(T);
        HIPC(hipGetLastError());
    }
    if (nab) {
        const int frc = finalize_on_device(c, s, *nab, Out_V, NR, out, want_spec);
        if (frc) return frc;
        c->last_transport_s = transport_s;
        c->last_total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tstart).count();
        c->last_launches = launches;
        return 0;
    }
    {
        const int frc = finish_tallies_host(c, T, 0, Out_V, NR, out, st);
        if (frc) return frc;
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
        if (d.em_W > 0.0) {
            if (add_back(out->Out_Ne_Em, s->acc_emn, (size_t)Nt)) return TRK_ERR_DEVICE;
            if (add_back(out->Out_E_Em, s->acc_eme, (size_t)Nt)) return TRK_ERR_DEVICE;
            if (add_back(out->Out_Ee_vs_E_Em, s->acc_emh, (size_t)Nt * NR)) return TRK_ERR_DEVICE;
        }
    }
    c->last_transport_s = transport_s;
    c->last_total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tstart).count();
    c->last_launches = launches;
    return 0;
}

// do_Monte_Carlo on buffers this context allocates on its first run and
// keeps for the next ones.
int run_mc(trk_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter, const double *Out_R, const double *Out_V,
           int NR, trk_tallies *out)
{
    const HostMat &h = c->hm;
    if (!c->mc) c->mc = new McState();
    McState *s = c->mc;
    const int Nt = h.Nt, nsh = h.nsh;
    const int R = (h.o.tally_replicas > 0) ? h.o.tally_replicas : 8;
    // capacities: particles of each kind per round
    const int cap = (h.o.gen_capacity > 0) ? h.o.gen_capacity : (1 << 22);
    if (s->cap[0] != cap) {
        for (int g = 0; g < 2; ++g)
            for (int k = 0; k < 3; ++k) free_list(s->L[g][k]);
        s->cap[0] = s->cap[1] = s->cap[2] = 0;
        const int caps[3] = {cap, cap, cap / 4};
        for (int g = 0; g < 2; ++g)
            for (int k = 0; k < 3; ++k)
                if (alloc_list(c, s->L[g][k], caps[k], k == K_HO)) return TRK_ERR_DEVICE;
        for (int k = 0; k < 3; ++k) s->cap[k] = caps[k];
    }
    if (!s->d_counts) HIPC(hipMalloc((void **)&s->d_counts, 16 * sizeof(int)));
    if (!s->d_overflow) HIPC(hipMalloc((void **)&s->d_overflow, sizeof(int)));
    if (!s->d_counters) HIPC(hipMalloc((void **)&s->d_counters, 16 * sizeof(unsigned long long)));
    if (!s->ev0) { HIPC(hipEventCreate(&s->ev0)); HIPC(hipEventCreate(&s->ev1)); }
    if (!s->ncu && device_info(c, s)) return TRK_ERR_DEVICE;
    if (ensure_tally(c, s, Nt, NR, nsh, R)) return TRK_ERR_DEVICE;
    if (ensure_doubles(c, &s->d_tim, &s->tim_cap, Nt) || ensure_doubles(c, &s->d_R, &s->R_cap, NR) ||
        ensure_doubles(c, &s->d_RE, &s->RE_cap, NR))
        return TRK_ERR_DEVICE;
    const int64_t B = batch_size(c, cap, nmc);
    if (s->cap[K_ION] < B) {
        for (int g = 0; g < 2; ++g) {
            free_list(s->L[g][K_ION]);
            s->cap[K_ION] = 0;
            if (alloc_list(c, s->L[g][K_ION], (int)B, 1)) return TRK_ERR_DEVICE;
        }
        s->cap[K_ION] = (int)B;
    }
    const bool want_spec = wants_per_iteration(c, out);
    if (want_spec && ensure_spec(c, s, (int)B, Nt, NR, h.m.n_dos)) return TRK_ERR_DEVICE;
    if (c->trace_host && c->trace_cap > 0) {
        if (s->trace_cap < c->trace_cap) {
            if (s->d_trace) (void)hipFree(s->d_trace);
            s->d_trace = nullptr;
            s->trace_cap = 0;
            HIPC(hipMalloc((void **)&s->d_trace, sizeof(double) * 12 * c->trace_cap));
            s->trace_cap = c->trace_cap;
        }
        if (!s->d_trace_n) HIPC(hipMalloc((void **)&s->d_trace_n, sizeof(unsigned long long)));
    }
    return run_core(c, s, nmc, seed, first_iter, Out_R, Out_V, NR, out, B, want_spec);
}


// ---------------------------------------------------------------- runs without allocation
// Every device buffer of trk_run_mc_noalloc (particle lists, tallies,
// per-iteration histograms, the staged Out_* arrays) is carved from this
// array.  Each GPU the library runs on holds its own copy until the process
// ends.  In host code its name is not a device address: the address on a
// device comes from hipGetSymbolAddress, made once per device with that
// device current, and only that address is used.
static __device__ __align__(256) unsigned char scratch[TRK_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices] = {};

static int scratch_base(trk_ctx *c, unsigned char **base)
{
    int dev = -1;
    HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(c, TRK_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(c, TRK_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *base = g_scratch_ptr[dev];
    return 0;
}

// Hands out 256-byte aligned pieces of a region; with base == nullptr it
// only counts the bytes a layout needs.
struct Carver {
    unsigned char *base;
    size_t off = 0;
    template <class T> T *take(size_t n)
    {
        off = (off + 255) & ~(size_t)255;
        T *p = base ? (T *)(base + off) : nullptr;
        off += (n ? n : 1) * sizeof(T);
        return p;
    }
};

static void carve_list(Carver &cv, PList &l, size_t n, int hole)
{
    l.E = cv.take<double>(n); l.t0 = cv.take<double>(n); l.tn = cv.take<double>(n);
    l.X = cv.take<double>(n); l.Y = cv.take<double>(n); l.Z = cv.take<double>(n);
    l.L = cv.take<double>(n); l.th = cv.take<double>(n); l.ph = cv.take<double>(n);
    if (hole) { l.Ehkin = cv.take<double>(n); l.Mass = cv.take<double>(n); l.sh = cv.take<int>(n); }
    else { l.Ehkin = l.Mass = nullptr; l.sh = nullptr; }
    l.id = cv.take<uint64_t>(n); l.nd = cv.take<uint32_t>(n); l.ne = cv.take<uint32_t>(n); l.iter = cv.take<uint32_t>(n);
}

// doubles of the caller's Out_* arrays the device finalize stages
static size_t stage_doubles(const trk_ctx *c, const trk_tallies *out, int NR, bool want_spec)
{
    const size_t Nt = (size_t)c->hm.Nt, na = (size_t)c->hm.m.n_atoms, ms = (size_t)(out->max_shells > 0 ? out->max_shells : 1);
    const size_t ND = (size_t)(c->hm.m.n_dos > 1 ? c->hm.m.n_dos : 1);
    const size_t plane = Nt * (size_t)NR;
    size_t n = 0;
    const double *planes[] = {out->Out_ne, out->Out_Ee, out->Out_nphot, out->Out_Ephot, out->Out_Elat};
    for (const double *p : planes) if (p) n += plane;
    const double *hol[] = {out->Out_nh, out->Out_Eh, out->Out_Ehkin};
    for (const double *p : hol) if (p) n += plane * na * ms;
    if (out->Out_E_h) n += Nt * na * ms;
    const double *nodes[] = {out->Out_tot_Ne, out->Out_tot_Nphot, out->Out_tot_E, out->Out_E_e, out->Out_E_phot, out->Out_E_at};
    for (const double *p : nodes) if (p) n += Nt;
    if (want_spec) {
        if (out->Out_Ee_vs_E) n += plane;
        if (out->Out_Eh_vs_E) n += Nt * ND;
        if (out->Out_theta) n += (Nt + 1) * 180;
        if (out->Out_theta_h) n += (Nt + 1) * 180;
        if (c->hm.m.work_function > 0.0) {
            if (out->Out_Ne_Em) n += Nt;
            if (out->Out_E_Em) n += Nt;
            if (out->Out_Ee_vs_E_Em) n += plane;
        }
    }
    return n;
}

// The layout of a run in a region: lists of `cap` electrons and holes
// (cap/4 photons) and B ions, both generations; one tally replica; the
// per-iteration histograms of B iterations; the device finalize's buffers.
static void carve_state(Carver &cv, McState &s, NoAllocBufs &na, const trk_ctx *c, int cap, int B, int NR,
                        bool want_spec, size_t stage_n)
{
    const int Nt = c->hm.Nt, nsh = c->hm.nsh;
    const int ND = c->hm.m.n_dos;
    const int caps[4] = {cap, cap, cap / 4, B};
    for (int g = 0; g < 2; ++g)
        for (int k = 0; k < 4; ++k) carve_list(cv, s.L[g][k], (size_t)caps[k], k == K_HO || k == K_ION);
    for (int k = 0; k < 4; ++k) s.cap[k] = caps[k];
    s.d_counts = cv.take<int>(16);
    s.d_overflow = cv.take<int>(1);
    s.d_counters = cv.take<unsigned long long>(16);
    const size_t cell = (size_t)Nt * NR;
    Tally &T = s.T;
    T.Nt = Nt; T.NR = NR; T.nsh = nsh; T.R = 1;
    T.cnt_e = cv.take<unsigned long long>(cell);
    T.E_e = cv.take<unsigned long long>(2 * cell);
    T.cnt_h = cv.take<unsigned long long>(cell * nsh);
    T.E_h = cv.take<unsigned long long>(2 * cell * nsh);
    T.Ek_h = cv.take<unsigned long long>(2 * cell * nsh);
    T.cnt_p = cv.take<unsigned long long>(cell);
    T.E_p = cv.take<unsigned long long>(2 * cell);
    T.Elat = cv.take<unsigned long long>(2 * cell);
    s.d_tim = cv.take<double>((size_t)Nt);
    s.d_R = cv.take<double>((size_t)NR);
    if (want_spec) {
        Spec &S = s.S;
        const size_t bt = (size_t)B * Nt;
        S.nel = cv.take<int>(bt); S.nvb = cv.take<int>(bt); S.dcn = cv.take<int>(bt);
        S.eve = cv.take<int>(bt * NR);
        S.ehve = cv.take<int>(bt * (ND > 1 ? ND : 1));
        S.th = cv.take<int>(bt * 180); S.thh = cv.take<int>(bt * 180);
        S.dcs = cv.take<unsigned long long>(bt * 2);
        s.acc_eve = cv.take<double>((size_t)Nt * NR);
        s.acc_ehve = cv.take<double>((size_t)Nt * (ND > 1 ? ND : 1));
        s.acc_th = cv.take<double>((size_t)(Nt + 1) * 180);
        s.acc_thh = cv.take<double>((size_t)(Nt + 1) * 180);
        s.acc_dc = cv.take<double>((size_t)Nt);
        s.em_n = cv.take<int>(bt);
        s.em_h = cv.take<int>(bt * NR);
        s.em_e = cv.take<unsigned long long>(bt * 2);
        s.acc_emn = cv.take<double>((size_t)Nt);
        s.acc_eme = cv.take<double>((size_t)Nt);
        s.acc_emh = cv.take<double>((size_t)Nt * NR);
        s.d_RE = cv.take<double>((size_t)NR);
        s.spec_B = B;
    }
    na.stage = cv.take<double>(stage_n);
    na.stage_n = stage_n;
    na.d_V = cv.take<double>((size_t)NR);
    na.node_sums = cv.take<unsigned long long>((size_t)Nt * 4);
}

// exact integer in units of 2^-40 eV -> double, as the host finalize converts
TRK_DEV double fx_to_double_dev(__int128 v)
{
    long long hi = (long long)(v >> 40);
    __int128 lo = v - ((__int128)hi << 40);
    return (double)hi + (double)(long long)lo * 0x1.0p-40;
}

// Destinations of the device finalize (staged Out_* arrays, nullptr = skip).
struct FinOut {
    double *ne, *Ee, *nph, *Eph, *Elat, *nh, *Eh, *Ehk, *E_h;
    double *tot_Ne, *tot_Nph, *tot_E, *E_e, *E_ph, *E_at;
};

// (1) one thread per (node, radial bin): the radial tallies, scaled by the
// bin's inverse volume and added to the staged arrays (Fortran order)
__global__ void k_fin_cells(Tally T, const double *V, FinOut o, int na, const int *atom_of, const int *shl_off)
{
    const int Nt = T.Nt, NR = T.NR, nsh = T.nsh;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Nt * NR) return;
    const int i = (int)(t % Nt), j = (int)(t / Nt);
    const size_t a = (size_t)i * NR + j;
    const double Vj = V[j];
    const size_t oo = (size_t)t;
    if (o.ne) o.ne[oo] += (double)(long long)T.cnt_e[a] * Vj;
    if (o.Ee) o.Ee[oo] += fx_to_double_dev(i128_load(T.E_e + 2 * a)) * Vj;
    if (o.nph) o.nph[oo] += (double)(long long)T.cnt_p[a] * Vj;
    if (o.Eph) o.Eph[oo] += fx_to_double_dev(i128_load(T.E_p + 2 * a)) * Vj;
    if (o.Elat) o.Elat[oo] += fx_to_double_dev(i128_load(T.Elat + 2 * a)) * Vj;
    for (int s2 = 0; s2 < nsh; ++s2) {
        const size_t q = a * nsh + s2;
        const int at = atom_of[s2], sl = s2 - shl_off[at];
        const size_t oh = (size_t)i + (size_t)Nt * (j + (size_t)NR * (at + (size_t)na * sl));
        if (o.nh) o.nh[oh] += (double)(long long)T.cnt_h[q] * Vj;
        if (o.Eh) o.Eh[oh] += fx_to_double_dev(i128_load(T.E_h + 2 * q)) * Vj;
        if (o.Ehk) o.Ehk[oh] += fx_to_double_dev(i128_load(T.Ek_h + 2 * q)) * Vj;
    }
}

// (2) one thread per node: totals over the radial bins and shells; the
// lattice energy of the node and the rest of the total energy are kept for
// the running sum of (3)
__global__ void k_fin_nodes(Tally T, FinOut o, int na, const int *atom_of, const int *shl_off, int photons,
                            unsigned long long *node_sums)
{
    const int Nt = T.Nt, NR = T.NR, nsh = T.nsh;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt) return;
    __int128 tot_e = 0, tot_Ee = 0, tot_p = 0, tot_Ep = 0, row_lat = 0, hsum = 0;
    for (int j = 0; j < NR; ++j) {
        const size_t a = (size_t)i * NR + j;
        tot_e += (long long)T.cnt_e[a];
        tot_Ee += i128_load(T.E_e + 2 * a);
        tot_p += (long long)T.cnt_p[a];
        tot_Ep += i128_load(T.E_p + 2 * a);
        row_lat += i128_load(T.Elat + 2 * a);
    }
    for (int s2 = 0; s2 < nsh; ++s2) {
        __int128 sh = 0;
        for (int j = 0; j < NR; ++j) {
            const size_t q = ((size_t)i * NR + j) * nsh + s2;
            sh += i128_load(T.E_h + 2 * q) + i128_load(T.Ek_h + 2 * q);
        }
        hsum += sh;
        const int at = atom_of[s2], sl = s2 - shl_off[at];
        if (o.E_h) o.E_h[(size_t)i + (size_t)Nt * (at + (size_t)na * sl)] += fx_to_double_dev(sh);
    }
    if (o.tot_Ne) o.tot_Ne[i] += (double)(long long)tot_e;
    if (o.E_e) o.E_e[i] += fx_to_double_dev(tot_Ee);
    if (photons) {
        if (o.tot_Nph) o.tot_Nph[i] += (double)(long long)tot_p;
        if (o.E_ph) o.E_ph[i] += fx_to_double_dev(tot_Ep);
    }
    i128_store(node_sums + 4 * (size_t)i, row_lat);
    i128_store(node_sums + 4 * (size_t)i + 2, tot_Ee + hsum + (photons ? tot_Ep : (__int128)0));
}

// (3) one thread: the lattice energy deposited up to each node and the total energy
__global__ void k_fin_prefix(FinOut o, int Nt, const unsigned long long *node_sums)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    __int128 at_cum = 0;
    for (int i = 0; i < Nt; ++i) {
        at_cum += i128_load(node_sums + 4 * (size_t)i);
        if (o.E_at) o.E_at[i] += fx_to_double_dev(at_cum);
        if (o.tot_E) o.tot_E[i] += fx_to_double_dev(i128_load(node_sums + 4 * (size_t)i + 2) + at_cum);
    }
}

__global__ void k_add(double *dst, const double *src, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = dst[k] + src[k];
}

static int finalize_on_device(trk_ctx *c, McState *s, const NoAllocBufs &na, const double *Out_V, int NR,
                              trk_tallies *out, bool want_spec)
{
    const HostMat &h = c->hm;
    const int Nt = h.Nt;
    const size_t nat = (size_t)h.m.n_atoms, ms = (size_t)(out->max_shells > 0 ? out->max_shells : 1);
    const size_t ND = (size_t)(h.m.n_dos > 1 ? h.m.n_dos : 1);
    const size_t plane = (size_t)Nt * NR;
    hipStream_t st = c->stream;
    FinOut fo{};
    struct Staged { double *host; size_t n; double **dev; };
    Staged arr[] = {
        {out->Out_ne, plane, &fo.ne}, {out->Out_Ee, plane, &fo.Ee}, {out->Out_nphot, plane, &fo.nph},
        {out->Out_Ephot, plane, &fo.Eph}, {out->Out_Elat, plane, &fo.Elat},
        {out->Out_nh, plane * nat * ms, &fo.nh}, {out->Out_Eh, plane * nat * ms, &fo.Eh},
        {out->Out_Ehkin, plane * nat * ms, &fo.Ehk}, {out->Out_E_h, (size_t)Nt * nat * ms, &fo.E_h},
        {out->Out_tot_Ne, (size_t)Nt, &fo.tot_Ne}, {out->Out_tot_Nphot, (size_t)Nt, &fo.tot_Nph},
        {out->Out_tot_E, (size_t)Nt, &fo.tot_E}, {out->Out_E_e, (size_t)Nt, &fo.E_e},
        {out->Out_E_phot, (size_t)Nt, &fo.E_ph}, {out->Out_E_at, (size_t)Nt, &fo.E_at}};
    double *spec_dev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const bool em = want_spec && h.m.work_function > 0.0;
    Staged sp[] = {{want_spec ? out->Out_Ee_vs_E : nullptr, plane, &spec_dev[0]},
                   {want_spec && h.m.n_dos > 1 ? out->Out_Eh_vs_E : nullptr, (size_t)Nt * ND, &spec_dev[1]},
                   {want_spec ? out->Out_theta : nullptr, (size_t)(Nt + 1) * 180, &spec_dev[2]},
                   {want_spec ? out->Out_theta_h : nullptr, (size_t)(Nt + 1) * 180, &spec_dev[3]},
                   {em ? out->Out_Ne_Em : nullptr, (size_t)Nt, &spec_dev[4]},
                   {em ? out->Out_E_Em : nullptr, (size_t)Nt, &spec_dev[5]},
                   {em ? out->Out_Ee_vs_E_Em : nullptr, plane, &spec_dev[6]}};
    size_t off = 0;
    for (Staged *group : {arr, sp}) {
        const size_t cnt = (group == arr) ? sizeof(arr) / sizeof(arr[0]) : sizeof(sp) / sizeof(sp[0]);
        for (size_t k = 0; k < cnt; ++k) {
            Staged &a = group[k];
            if (!a.host) continue;
            if (off + a.n > na.stage_n) return set_error(c, TRK_ERR_CAPACITY, "staging area too small");
            *a.dev = na.stage + off;
            off += a.n;
            HIPC(hipMemcpyAsync(*a.dev, a.host, a.n * sizeof(double), hipMemcpyHostToDevice, st));
        }
    }
    HIPC(hipMemcpyAsync(na.d_V, Out_V, sizeof(double) * NR, hipMemcpyHostToDevice, st));
    const MatDev &md = c->md;
    k_fin_cells<<<(unsigned)((plane + 255) / 256), 256, 0, st>>>(s->T, na.d_V, fo, (int)nat, md.atom_of, md.shl_off);
    k_fin_nodes<<<(unsigned)((Nt + 63) / 64), 64, 0, st>>>(s->T, fo, (int)nat, md.atom_of, md.shl_off,
                                                           h.o.include_photons, na.node_sums);
    k_fin_prefix<<<1, 64, 0, st>>>(fo, Nt, na.node_sums);
    const double *acc[7] = {s->acc_eve, s->acc_ehve, s->acc_th, s->acc_thh, s->acc_emn, s->acc_eme, s->acc_emh};
    for (int k = 0; k < 7; ++k)
        if (spec_dev[k]) k_add<<<(unsigned)((sp[k].n + 255) / 256), 256, 0, st>>>(spec_dev[k], acc[k], (int64_t)sp[k].n);
    HIPC(hipGetLastError());
    for (Staged *group : {arr, sp}) {
        const size_t cnt = (group == arr) ? sizeof(arr) / sizeof(arr[0]) : sizeof(sp) / sizeof(sp[0]);
        for (size_t k = 0; k < cnt; ++k)
            if (group[k].host)
                HIPC(hipMemcpyAsync(group[k].host, *group[k].dev, group[k].n * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    if (want_spec && out->Out_diff_coeff)
        HIPC(hipMemcpyAsync(out->Out_diff_coeff, s->acc_dc, (size_t)Nt * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return 0;
}

// do_Monte_Carlo on buffers carved from the scratch array: no device memory
// is allocated or freed.  The particle-list capacity is the largest power of
// two (2^22 down to 2^10) whose layout fits, unless the options fix it.
int run_mc_noalloc(trk_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter, const double *Out_R,
                   const double *Out_V, int NR, trk_tallies *out)
{
    if (c->trace_host && c->trace_cap > 0)
        return set_error(c, TRK_ERR_STATE, "event traces need trk_run_mc (the run without allocation records none)");
    int cur = -1;
    HIPC(hipGetDevice(&cur));
    if (cur != c->device) HIPC(hipSetDevice(c->device));
    unsigned char *base = nullptr;
    int rc = scratch_base(c, &base);
    if (rc) return rc;
    const bool want_spec = wants_per_iteration(c, out);
    const size_t stage_n = stage_doubles(c, out, NR, want_spec);
    const size_t limit = (size_t)TRK_NOALLOC_SCRATCH_BYTES;
    McState s{};
    NoAllocBufs na{};
    const int fixed = c->hm.o.gen_capacity;
    int cap = 0;
    int64_t B = 0;
    for (int cand = (fixed > 0 ? fixed : (1 << 22)); cand >= (fixed > 0 ? fixed : (1 << 10)); cand >>= 1) {
        const int64_t b = batch_size(c, cand, nmc);
        Carver cv{nullptr};
        carve_state(cv, s, na, c, cand, (int)b, NR, want_spec, stage_n);
        if (cv.off <= limit) { cap = cand; B = b; break; }
        if (fixed > 0) break;
    }
    if (cap == 0)
        return set_error(c, TRK_ERR_CAPACITY, "the run does not fit the %zu-byte scratch array of the entry point "
                         "without allocation (fewer time nodes, radial bins or iterations per batch would)", limit);
    s = McState{};
    na = NoAllocBufs{};
    Carver cv{base};
    carve_state(cv, s, na, c, cap, (int)B, NR, want_spec, stage_n);
    if (device_info(c, &s)) return TRK_ERR_DEVICE;
    rc = run_core(c, &s, nmc, seed, first_iter, Out_R, Out_V, NR, out, B, want_spec, &na);
    if (cur != c->device) (void)hipSetDevice(cur);
    return rc;
}

}  // namespace trk
