// trk_batch.hip -- many ion impacts in one persistent launch.
//
// The reference runs one ion on one target per program run and spreads its
// MC iterations over OpenMP threads (do_Monte_Carlo, Monte_Carlo.f90:395-415).
// A batch takes many impacts -- different ions, energies and materials, one
// MC iteration each -- and transports them all in one launch of k_batch:
//  * impacts on the same material context form a group and share its
//    tables (trk_ctx_create_ion); each impact has its own ion tables
//    (IonDev, read by the lanes that run its ion's collisions);
//  * every group has its own particle lists and tallies, one tally plane
//    per impact of the group, so each impact's Out_* arrays come out as if
//    it had run alone (impact i is iteration first_iter + i, with that
//    iteration's random streams);
//  * k_batch runs the transport rounds of all groups with a grid barrier
//    between rounds instead of a host synchronisation; in a round each
//    workgroup stages one group's tables in LDS and takes another group
//    when that one's items are used up;
//  * the per-run sequence (control reset, tally clearing, ions, transport,
//    time prefix of the tallies) is captured once in a hipGraph and
//    replayed by every run;
//  * threads per workgroup and the events a particle may take per round
//    (how deep a round goes into each history before it is queued for the
//    next) are chosen by trk_batch_tune, which times the candidates on the
//    batch itself and records every timing.
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <new>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include "trk_mc.h"

namespace trk {

struct BatchGroup {
    const trk_ctx *mat = nullptr;     // the material context whose tables the group uses
    std::vector<int> members;         // impact indices
    int *d_idx = nullptr;             // members on the device
    PList L[2][4];
    int cap[4] = {0, 0, 0, 0};
    Tally T{};
    GroupCtl *gctl = nullptr;
    double *d_tim = nullptr;
};

struct TuneEntry {
    int block, round_events;
    double ms;
};

}  // namespace trk

struct trk_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<trk_ctx *> impacts;
    std::vector<trk::BatchGroup> groups;
    bool mixed = false;
    int NR = 0;
    std::vector<double> R, V;
    double *d_R = nullptr;
    trk::IonDev *d_ions = nullptr;
    int *d_plane_of = nullptr;            // each impact's tally plane in its group
    std::vector<int> plane_of, group_of;
    trk::BatchCtl *d_ctl = nullptr;
    int *d_overflow = nullptr;
    unsigned long long *d_counters = nullptr;
    double *d_trace = nullptr;
    unsigned long long *d_trace_n = nullptr;
    int64_t trace_cap = 0, trace_alloc = 0;
    double *trace_host = nullptr;
    int64_t *trace_n_out = nullptr;
    trk::McDev mc[trk::kMaxGroups][2];
    int capacity = 0;
    int block = 768, round_events = 512;
    int64_t blocks = 0;
    size_t lds_bytes = 0;
    bool all_staged = false;
    int ncu = 0, max_lds = 0;
    unsigned long long deadline_ticks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    bool has_run = false;          // a run's finished tally planes are on the device
    int64_t last_events[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int rounds = 0;
    std::vector<trk::TuneEntry> tune_log;
    std::string err;
};

namespace trk {

static int berr(trk_batch *b, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

#define BHIP(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return berr(b, TRK_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

template <class T>
static int balloc(trk_batch *b, T **p, size_t n)
{
    BHIP(hipMalloc((void **)p, (n ? n : 1) * sizeof(T)));
    return 0;
}

static int batch_alloc_list(trk_batch *b, PList &l, int cap, int hole)
{
    const size_t n = (size_t)(cap > 0 ? cap : 1);
    double **dp[] = {&l.E, &l.t0, &l.tn, &l.X, &l.Y, &l.Z, &l.L, &l.th, &l.ph};
    for (double **p : dp) if (balloc(b, p, n)) return TRK_ERR_DEVICE;
    if (hole) {
        if (balloc(b, &l.Ehkin, n) || balloc(b, &l.Mass, n) || balloc(b, &l.sh, n)) return TRK_ERR_DEVICE;
    }
    if (balloc(b, &l.id, n) || balloc(b, &l.nd, n) || balloc(b, &l.ne, n) || balloc(b, &l.iter, n)) return TRK_ERR_DEVICE;
    return 0;
}

static void batch_free_list(PList &l)
{
    void *ps[] = {l.E, l.t0, l.tn, l.X, l.Y, l.Z, l.L, l.th, l.ph, l.Ehkin, l.Mass, l.sh, l.id, l.nd, l.ne, l.iter};
    for (void *p : ps) if (p) (void)hipFree(p);
    l = PList{};
}

// Control reset of a batch run (one thread): the barrier, the round counter,
// and each group's round state with its ions as the first round's items.
__global__ void k_batch_reset(BatchCtl *ctl, int n_groups, unsigned long long deadline, GroupCtl *g0, GroupCtl *g1,
                              GroupCtl *g2, GroupCtl *g3, GroupCtl *g4, GroupCtl *g5, GroupCtl *g6, GroupCtl *g7,
                              const int *n_ions, int *overflow, unsigned long long *counters, unsigned long long *trace_n)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    ctl->bar_count = 0u;
    ctl->bar_gen = 0u;
    ctl->done = 0;
    ctl->err = 0;
    ctl->round = 0;
    ctl->n_groups = n_groups;
    ctl->deadline_ticks = deadline;
    GroupCtl *gs[kMaxGroups] = {g0, g1, g2, g3, g4, g5, g6, g7};
    for (int g = 0; g < n_groups; ++g) {
        GroupCtl *gc = gs[g];
        for (int k = 0; k < 4; ++k) {
            gc->n_in[k] = 0;
            gc->out_cnt[0][k] = 0;
            gc->out_cnt[1][k] = 0;
        }
        gc->n_in[K_ION] = n_ions[g];
        gc->cursor = 0;
        gc->cur = 0;
        gc->overflow = 0;
    }
    *overflow = 0;
    for (int k = 0; k < 16; ++k) counters[k] = 0ull;
    if (trace_n) *trace_n = 0ull;
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// The lattice-energy cells of every plane from the split fixed-point form the
// kernels add in (low 32 bits, signed high part) to 128-bit integers.
__global__ void k_elat_to_i128(unsigned long long *Elat, int64_t ncell)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncell) i128_store(Elat + 2 * i, fx_read(Elat + 2 * i));
}

// Running sum over the time nodes of the difference arrays of every tally
// plane (one plane per impact), from the split fixed-point form of the
// energies to 128-bit integers, as the replica sum and time prefix of a
// single run leave them (the lattice energy is converted by k_elat_to_i128).
__global__ void k_time_prefix_planes(Tally T)
{
    const int NR = T.NR, Nt = T.Nt, nsh = T.nsh, R = T.R;
    const int64_t per = 2 * (int64_t)NR + (int64_t)NR * nsh;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * R) return;
    const int64_t r = t / per, q = t % per;
    const size_t plane = (size_t)Nt * NR, base = (size_t)r * plane;
    if (q < 2 * NR) {
        const int j = (int)(q % NR);
        unsigned long long *cnt = (q < NR) ? T.cnt_e : T.cnt_p;
        unsigned long long *en = (q < NR) ? T.E_e : T.E_p;
        long long rc = 0;
        __int128 re = 0;
        for (int i = 0; i < Nt; ++i) {
            const size_t a = base + (size_t)i * NR + j;
            rc += (long long)cnt[a];
            re += fx_read(en + 2 * a);
            cnt[a] = (unsigned long long)rc;
            i128_store(en + 2 * a, re);
        }
    } else {
        const int64_t u = q - 2 * NR;
        const int j = (int)(u / nsh), sh = (int)(u % nsh);
        long long rc = 0;
        __int128 rE = 0, rK = 0;
        for (int i = 0; i < Nt; ++i) {
            const size_t a = (base + (size_t)i * NR + j) * nsh + sh;
            rc += (long long)T.cnt_h[a];
            rE += fx_read(T.E_h + 2 * a);
            rK += fx_read(T.Ek_h + 2 * a);
            T.cnt_h[a] = (unsigned long long)rc;
            i128_store(T.E_h + 2 * a, rE);
            i128_store(T.Ek_h + 2 * a, rK);
        }
    }
}

static void batch_free(trk_batch *b)
{
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (BatchGroup &g : b->groups) {
        for (auto &a : g.L) for (auto &l : a) batch_free_list(l);
        void *ps[] = {g.d_idx, g.gctl, g.d_tim, g.T.cnt_e, g.T.E_e, g.T.cnt_h, g.T.E_h, g.T.Ek_h, g.T.cnt_p, g.T.E_p,
                      g.T.Elat};
        for (void *p : ps) if (p) (void)hipFree(p);
    }
    b->groups.clear();
    void *ps[] = {b->d_R, b->d_ions, b->d_plane_of, b->d_ctl, b->d_overflow, b->d_counters, b->d_trace, b->d_trace_n};
    for (void *p : ps) if (p) (void)hipFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

// Launch shape for the current block size: every workgroup must be resident
// at once (the grid barrier waits for all of them).
static int batch_shape(trk_batch *b)
{
    const void *kt = b->mixed ? (b->all_staged ? (const void *)mx::k_batch<true> : (const void *)mx::k_batch<false>)
                              : (b->all_staged ? (const void *)fp64::k_batch<true> : (const void *)fp64::k_batch<false>);
    if (b->lds_bytes > 65536)
        BHIP(hipFuncSetAttribute(kt, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds_bytes));
    int per_cu = 0;
    BHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kt, b->block, b->lds_bytes));
    if (per_cu < 1) return berr(b, TRK_ERR_CAPACITY, "a %d-thread batch workgroup with %zu bytes of LDS does not fit a CU",
                                b->block, b->lds_bytes);
    b->blocks = (int64_t)per_cu * b->ncu;
    return 0;
}

// Per-group transport descriptors by list parity (seed and iteration base
// are set by each run).
static void batch_descriptors(trk_batch *b)
{
    const int G = (int)b->groups.size();
    for (int g = 0; g < G; ++g) {
        BatchGroup &gr = b->groups[g];
        const trk_ctx *c = b->impacts[gr.members[0]];
        const HostMat &h = c->hm;
        for (int par = 0; par < 2; ++par) {
            McDev &d = b->mc[g][par];
            d = McDev{};
            fill_mcdev(c, d);
            d.Tim = h.o.Tim;
            d.T_last = h.tim_glob[h.Nt - 1];
            d.cutoff_eff = (h.m.cut_off > 0.0) ? h.m.cut_off : 0.0;
            d.Nt = h.Nt;
            d.tim_glob = gr.d_tim;
            d.inv_dt = (h.o.dt_flag <= 0 && h.o.dt > 0.0) ? 1.0 / h.o.dt : 0.0;
            d.hints = 1;
            for (int i : gr.members)
                if (!b->impacts[i]->ts.grids_increasing) d.hints = 0;
            for (int i = 1; i < h.Nt; ++i)
                if (!(h.tim_glob[i - 1] < h.tim_glob[i])) d.hints = 0;
            for (int i = 1; i < b->NR; ++i)
                if (!(b->R[i - 1] < b->R[i])) d.hints = 0;
            d.NR = b->NR;
            d.R = b->d_R;
            d.include_photons = h.o.include_photons;
            for (int k = 0; k < 4; ++k) {
                d.in[k] = gr.L[par][k];
                d.out[k] = gr.L[1 - par][k];
                d.n_in[k] = 0;
                d.cap[k] = gr.cap[k];
            }
            d.out_cnt = &gr.gctl->out_cnt[1 - par][0];
            d.cursor = &gr.gctl->cursor;
            d.overflow = b->d_overflow;
            d.counters = b->d_counters;
            d.T = gr.T;
            d.S = Spec{};
            if (b->trace_host && b->trace_cap > 0) {
                d.trace = b->d_trace;
                d.trace_n = b->d_trace_n;
                d.trace_cap = b->trace_cap;
            }
            d.ions = b->d_ions;
            d.plane_of = b->d_plane_of;
            d.round_events = b->round_events;
            d.gctl = gr.gctl;
            size_t bytes = 0;
            const size_t limit = ((size_t)b->max_lds - 256) / sizeof(double);
            const bool all = plan_lds(d, h.Nt, b->NR, h.m.n_dos, h.nsh, limit, &bytes);
            if (par == 0) {
                if (g == 0) { b->lds_bytes = 0; b->all_staged = true; }
                b->lds_bytes = bytes > b->lds_bytes ? bytes : b->lds_bytes;
                b->all_staged = b->all_staged && all;
            }
        }
    }
}

static int batch_capture(trk_batch *b)
{
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    const int G = (int)b->groups.size();
    hipStream_t st = b->stream;
    std::vector<int> nions(kMaxGroups, 0);
    for (int g = 0; g < G; ++g) nions[g] = (int)b->groups[g].members.size();
    int *d_nions = nullptr;
    // the counts live in the control block's spare words: no allocation per capture
    d_nions = (int *)&b->d_ctl[1];
    BHIP(hipMemcpy(d_nions, nions.data(), sizeof(int) * kMaxGroups, hipMemcpyHostToDevice));
    GroupCtl *gc[kMaxGroups] = {};
    for (int g = 0; g < G; ++g) gc[g] = b->groups[g].gctl;
    for (int g = G; g < kMaxGroups; ++g) gc[g] = gc[0];
    BHIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    k_batch_reset<<<1, 64, 0, st>>>(b->d_ctl, G, b->deadline_ticks, gc[0], gc[1], gc[2], gc[3], gc[4], gc[5], gc[6],
                                     gc[7], d_nions, b->d_overflow, b->d_counters,
                                     (b->trace_host && b->trace_cap > 0) ? b->d_trace_n : nullptr);
    for (int g = 0; g < G; ++g) {
        const Tally &T = b->groups[g].T;
        const int64_t cell = (int64_t)T.R * T.Nt * T.NR;
        struct { unsigned long long *p; int64_t n; } z[] = {
            {T.cnt_e, cell}, {T.E_e, 2 * cell}, {T.cnt_h, cell * T.nsh}, {T.E_h, 2 * cell * T.nsh},
            {T.Ek_h, 2 * cell * T.nsh}, {T.cnt_p, cell}, {T.E_p, 2 * cell}, {T.Elat, 2 * cell}};
        for (auto &a : z) k_fill_u64<<<(unsigned)((a.n + 255) / 256), 256, 0, st>>>(a.p, a.n);
    }
    for (int g = 0; g < G; ++g) {
        const int nb = (int)b->groups[g].members.size();
        if (b->mixed) mx::k_ion_init_batch<<<(nb + 63) / 64, 64, 0, st>>>(g, nb, b->groups[g].d_idx);
        else fp64::k_ion_init_batch<<<(nb + 63) / 64, 64, 0, st>>>(g, nb, b->groups[g].d_idx);
    }
    const unsigned nblk = (unsigned)b->blocks;
    if (b->mixed && b->all_staged) mx::k_batch<true><<<nblk, b->block, b->lds_bytes, st>>>(b->d_ctl, G);
    else if (b->mixed) mx::k_batch<false><<<nblk, b->block, b->lds_bytes, st>>>(b->d_ctl, G);
    else if (b->all_staged) fp64::k_batch<true><<<nblk, b->block, b->lds_bytes, st>>>(b->d_ctl, G);
    else fp64::k_batch<false><<<nblk, b->block, b->lds_bytes, st>>>(b->d_ctl, G);
    for (int g = 0; g < G; ++g) {
        const Tally &T = b->groups[g].T;
        const int64_t n = (2 * (int64_t)T.NR + (int64_t)T.NR * T.nsh) * T.R;
        k_time_prefix_planes<<<(unsigned)((n + 63) / 64), 64, 0, st>>>(T);
        const int64_t nc = (int64_t)T.R * T.Nt * T.NR;
        k_elat_to_i128<<<(unsigned)((nc + 255) / 256), 256, 0, st>>>(T.Elat, nc);
    }
    const hipError_t le = hipGetLastError();
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &graph);
    if (le != hipSuccess) { if (graph) (void)hipGraphDestroy(graph); return berr(b, TRK_ERR_DEVICE, "launch in capture: %s", hipGetErrorString(le)); }
    if (ce != hipSuccess) return berr(b, TRK_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    b->graph = graph;
    BHIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return 0;
}

static int batch_fetch(trk_batch *b, trk_tallies *const *outs);

// One replay of the batch graph; host-side tallies only when outs is given.
static int batch_launch(trk_batch *b, uint64_t seed, int64_t first_iter, trk_tallies *const *outs)
{
    const int G = (int)b->groups.size();
    hipStream_t st = b->stream;
    for (int g = 0; g < G; ++g)
        for (int par = 0; par < 2; ++par) {
            b->mc[g][par].seed = seed;
            b->mc[g][par].iter0 = (uint32_t)first_iter;
        }
    if (!b->exec && batch_capture(b)) return TRK_ERR_DEVICE;
    BHIP(b->mixed ? mx::batch_upload(&b->mc[0][0], G, st) : fp64::batch_upload(&b->mc[0][0], G, st));
    BHIP(hipEventRecord(b->ev0, st));
    BHIP(hipGraphLaunch(b->exec, st));
    BHIP(hipEventRecord(b->ev1, st));
    BHIP(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    BatchCtl ctl{};
    BHIP(hipMemcpy(&ctl, b->d_ctl, sizeof ctl, hipMemcpyDeviceToHost));
    b->rounds = ctl.round;
    if (ctl.err & kBatchErrBarrier)
        return berr(b, TRK_ERR_DEVICE, "a batch round barrier timed out (workgroups not all resident?)");
    if (ctl.round >= (1 << 16)) return berr(b, TRK_ERR_CAPACITY, "more than 65536 transport rounds");
    int ovf = 0;
    BHIP(hipMemcpy(&ovf, b->d_overflow, sizeof ovf, hipMemcpyDeviceToHost));
    if (ovf & kFlagNaN) return berr(b, TRK_ERR_NUMERIC, "a particle state became NaN");
    if (ovf & kFlagEvents) return berr(b, TRK_ERR_CAPACITY, "a particle exceeded %u events", kMaxEvents);
    for (int g = 0; g < G; ++g) {
        GroupCtl gc{};
        BHIP(hipMemcpy(&gc, b->groups[g].gctl, sizeof gc, hipMemcpyDeviceToHost));
        if (gc.overflow) ovf |= kFlagPairs;
    }
    if (ovf) return berr(b, TRK_ERR_CAPACITY, "particle list capacity %d exceeded (flags %d)", b->capacity, ovf);
    unsigned long long ctr[16];
    BHIP(hipMemcpy(ctr, b->d_counters, sizeof ctr, hipMemcpyDeviceToHost));
    for (int k = 0; k < 8; ++k) b->last_events[k] = (int64_t)ctr[k];
    if (b->trace_host && b->trace_cap > 0) {
        unsigned long long tn = 0;
        BHIP(hipMemcpy(&tn, b->d_trace_n, sizeof tn, hipMemcpyDeviceToHost));
        const int64_t ncopy = (int64_t)tn < b->trace_cap ? (int64_t)tn : b->trace_cap;
        BHIP(hipMemcpy(b->trace_host, b->d_trace, sizeof(double) * 12 * ncopy, hipMemcpyDeviceToHost));
        if (b->trace_n_out) *b->trace_n_out = (int64_t)tn;
    }
    b->has_run = true;
    return outs ? batch_fetch(b, outs) : 0;
}

// The host tallies of every impact of the last run, from its finished planes.
static int batch_fetch(trk_batch *b, trk_tallies *const *outs)
{
    const int n = (int)b->impacts.size();
    for (int i = 0; i < n; ++i) {
        if (!outs[i]) continue;
        trk_ctx *c = b->impacts[i];
        const Tally &T = b->groups[b->group_of[i]].T;
        const int rc = finish_tallies_host(c, T, b->plane_of[i], b->V.data(), b->NR, outs[i], b->stream);
        if (rc) return berr(b, rc, "tallies of impact %d: %s", i, c->err.c_str());
    }
    return 0;
}

static int check_outs(trk_batch *b, trk_tallies *const *outs)
{
    const int n = (int)b->impacts.size();
    for (int i = 0; i < n; ++i) {
        trk_tallies *o = outs[i];
        if (!o) continue;
        const trk_ctx *c = b->impacts[i];
        if (o->Nt != c->hm.Nt || o->NR != b->NR)
            return berr(b, TRK_ERR_ARG, "tallies of impact %d have Nt=%d NR=%d, the impact has %d and %d", i, o->Nt,
                        o->NR, c->hm.Nt, b->NR);
        if (o->Out_Ee_vs_E || o->Out_Eh_vs_E || o->Out_theta || o->Out_theta_h || o->Out_diff_coeff ||
            o->Out_Ne_Em || o->Out_E_Em || o->Out_Ee_vs_E_Em)
            return berr(b, TRK_ERR_ARG, "the per-iteration normalised tallies (spectra, angles, diffusion, "
                                        "emitted electrons) are made by trk_run_mc, not by batches: pass NULL for them");
        int ms = 0;
        for (int a = 0; a < c->hm.m.n_atoms; ++a) ms = c->hm.m.n_shells[a] > ms ? c->hm.m.n_shells[a] : ms;
        if ((o->Out_nh || o->Out_Eh || o->Out_Ehkin || o->Out_E_h) && o->max_shells < ms)
            return berr(b, TRK_ERR_ARG, "tallies of impact %d have max_shells=%d < %d", i, o->max_shells, ms);
    }
    return 0;
}

}  // namespace trk

using namespace trk;

extern "C" {

int trk_ctx_create_ion(const trk_ctx *mat, const trk_ion *ion, trk_ctx **out)
{
    if (!out) return TRK_ERR_ARG;
    *out = nullptr;
    if (!mat || !ion) return TRK_ERR_ARG;
    trk_ctx *c = new (std::nothrow) trk_ctx();
    if (!c) return TRK_ERR_ARG;
    *out = c;
    if (!mat->tables_ready) return set_error(c, TRK_ERR_STATE, "the material context's tables must be built first");
    if (mat->base) return set_error(c, TRK_ERR_ARG, "the material context is itself an ion context");
    if (ion->Zat < 1 || !(ion->Mass > 0.0) || !(ion->E > 0.0)) return set_error(c, TRK_ERR_ARG, "invalid ion");
    if (ion->Kind_ion != 0 && ion->Kind_ion != 1)
        return set_error(c, TRK_ERR_ARG, "Kind_ion must be 0 (point charge) or 1 (Brandt-Kitagawa)");
    c->base = mat;
    c->hm = mat->hm;
    c->hm.ion = *ion;
    c->hm.shi_Zeff0 = shi_zeff(c->hm, ion->E);
    c->md = mat->md;           // material arrays are the material context's
    c->md.ion_mass = ion->Mass; c->md.ion_E = ion->E; c->md.ion_Zeff0 = c->hm.shi_Zeff0;
    c->md.ion_Z = ion->Zat; c->md.ion_kind_zeff = ion->Kind_Zeff; c->md.ion_fixed_zeff = ion->fixed_Zeff;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev != mat->device)
        return set_error(c, TRK_ERR_STATE, "the material context belongs to device %d", mat->device);
    c->device = dev;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return set_error(c, TRK_ERR_DEVICE, "hipStreamCreateWithFlags failed");
    return 0;
}

int trk_batch_create(trk_ctx *const *impacts, int32_t n, const double *Out_R, const double *Out_V, int32_t NR,
                     int32_t capacity, trk_batch **out)
{
    if (!out) return TRK_ERR_ARG;
    *out = nullptr;
    trk_batch *b = new (std::nothrow) trk_batch();
    if (!b) return TRK_ERR_ARG;
    *out = b;
    if (!impacts || n < 1 || !Out_R || !Out_V || NR < 1) return berr(b, TRK_ERR_ARG, "invalid trk_batch_create arguments");
    for (int j = 1; j < NR; ++j)
        if (!(Out_R[j] > Out_R[j - 1])) return berr(b, TRK_ERR_ARG, "Out_R must be strictly increasing");
    b->impacts.assign(impacts, impacts + n);
    std::map<const trk_ctx *, int> gi;
    for (int i = 0; i < n; ++i) {
        trk_ctx *c = impacts[i];
        if (!c || !c->tables_ready) return berr(b, TRK_ERR_STATE, "impact %d has no built tables", i);
        if (c->tables_dirty || (c->base && c->base->tables_dirty))
            return berr(b, TRK_ERR_STATE, "impact %d: tables were set, trk_tables_finish must run first", i);
        if (c->hm.o.precision != impacts[0]->hm.o.precision) return berr(b, TRK_ERR_ARG, "impacts mix precisions");
        if (uses_general_kernel(c->hm))
            return berr(b, TRK_ERR_ARG, "impact %d uses the Delta-CDF model (kind_of_DR = 4), BEB shells or DSF "
                                        "elastic scattering, which trk_run_mc transports and batches do not", i);
        if (c->device != impacts[0]->device) return berr(b, TRK_ERR_ARG, "impacts on different devices");
        const trk_ctx *m = c->base ? c->base : c;
        auto it = gi.find(m);
        if (it == gi.end()) {
            if ((int)b->groups.size() >= kMaxGroups) return berr(b, TRK_ERR_CAPACITY, "more than %d materials in one batch", kMaxGroups);
            it = gi.emplace(m, (int)b->groups.size()).first;
            b->groups.emplace_back();
            b->groups.back().mat = m;
        }
        b->groups[it->second].members.push_back(i);
        b->group_of.push_back(it->second);
        b->plane_of.push_back((int)b->groups[it->second].members.size() - 1);
    }
    b->mixed = impacts[0]->hm.o.precision == 1;
    b->device = impacts[0]->device;
    int cur = -1;
    BHIP(hipGetDevice(&cur));
    if (cur != b->device) return berr(b, TRK_ERR_STATE, "the impacts belong to device %d", b->device);
    BHIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BHIP(hipEventCreate(&b->ev0));
    BHIP(hipEventCreate(&b->ev1));
    BHIP(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device));
    BHIP(hipDeviceGetAttribute(&b->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device));
    int khz = 0;
    BHIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
    b->deadline_ticks = (unsigned long long)(khz > 0 ? khz : 100000) * 1000ull * 60ull;   // 60 s per round
    b->NR = NR;
    b->R.assign(Out_R, Out_R + NR);
    b->V.assign(Out_V, Out_V + NR);
    b->capacity = capacity > 0 ? capacity : (1 << 22);
    if (balloc(b, &b->d_R, NR)) return TRK_ERR_DEVICE;
    BHIP(hipMemcpy(b->d_R, Out_R, sizeof(double) * NR, hipMemcpyHostToDevice));
    // the control block and, after it, kMaxGroups ion counts of the capture
    if (balloc(b, (char **)&b->d_ctl, 2 * sizeof(BatchCtl) + sizeof(int) * kMaxGroups)) return TRK_ERR_DEVICE;
    if (balloc(b, &b->d_overflow, 1) || balloc(b, &b->d_counters, 16)) return TRK_ERR_DEVICE;
    // the impacts' ions; the mixed path's per-shell transfer constants
    std::vector<IonDev> ions(n);
    for (int i = 0; i < n; ++i) {
        trk_ctx *c = impacts[i];
        ions[i] = ion_view_host(c);
        if (!c->ts.sd_consts) {
            BHIP(launch_shi_consts(c, ions[i], b->mixed, b->stream));
            c->ts.sd_consts = true;
        }
    }
    if (balloc(b, &b->d_ions, n) || balloc(b, &b->d_plane_of, n)) return TRK_ERR_DEVICE;
    BHIP(hipMemcpy(b->d_ions, ions.data(), sizeof(IonDev) * n, hipMemcpyHostToDevice));
    BHIP(hipMemcpy(b->d_plane_of, b->plane_of.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    // per group: lists, one tally plane per impact of the group, round state, time grid
    for (BatchGroup &g : b->groups) {
        const trk_ctx *c = b->impacts[g.members[0]];
        const int nm = (int)g.members.size();
        const int caps[4] = {b->capacity, b->capacity, b->capacity / 4 > 0 ? b->capacity / 4 : 1, nm};
        for (int par = 0; par < 2; ++par)
            for (int k = 0; k < 4; ++k)
                if (batch_alloc_list(b, g.L[par][k], caps[k], k == K_HO || k == K_ION)) return TRK_ERR_DEVICE;
        for (int k = 0; k < 4; ++k) g.cap[k] = caps[k];
        Tally &T = g.T;
        T.Nt = c->hm.Nt; T.NR = NR; T.nsh = c->hm.nsh; T.R = nm;
        const size_t cell = (size_t)nm * T.Nt * NR;
        if (balloc(b, &T.cnt_e, cell) || balloc(b, &T.E_e, 2 * cell) || balloc(b, &T.cnt_h, cell * T.nsh) ||
            balloc(b, &T.E_h, 2 * cell * T.nsh) || balloc(b, &T.Ek_h, 2 * cell * T.nsh) || balloc(b, &T.cnt_p, cell) ||
            balloc(b, &T.E_p, 2 * cell) || balloc(b, &T.Elat, 2 * cell))
            return TRK_ERR_DEVICE;
        if (balloc(b, &g.gctl, 1) || balloc(b, &g.d_tim, (size_t)T.Nt) || balloc(b, &g.d_idx, (size_t)nm)) return TRK_ERR_DEVICE;
        BHIP(hipMemcpy(g.d_tim, c->hm.tim_glob.data(), sizeof(double) * T.Nt, hipMemcpyHostToDevice));
        BHIP(hipMemcpy(g.d_idx, g.members.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
    }
    batch_descriptors(b);
    if (batch_shape(b)) return b->err.empty() ? TRK_ERR_DEVICE : TRK_ERR_CAPACITY;
    BHIP(hipStreamSynchronize(b->stream));
    return 0;
}

void trk_batch_destroy(trk_batch *b)
{
    if (!b) return;
    batch_free(b);
    delete b;
}

int trk_batch_error(const trk_batch *b, char *buf, size_t n)
{
    if (!b || !buf || n == 0) return TRK_ERR_ARG;
    std::strncpy(buf, b->err.c_str(), n - 1);
    buf[n - 1] = 0;
    return 0;
}

int trk_batch_set_trace(trk_batch *b, double *buf, int64_t cap_records, int64_t *n_out)
{
    if (!b || cap_records < 0) return TRK_ERR_ARG;
    if (buf && cap_records > b->trace_alloc) {
        if (b->d_trace) (void)hipFree(b->d_trace);
        b->d_trace = nullptr;
        b->trace_alloc = 0;
        if (balloc(b, &b->d_trace, (size_t)12 * cap_records)) return TRK_ERR_DEVICE;
        b->trace_alloc = cap_records;
    }
    if (buf && !b->d_trace_n && balloc(b, &b->d_trace_n, 1)) return TRK_ERR_DEVICE;
    b->trace_host = buf;
    b->trace_cap = buf ? cap_records : 0;
    b->trace_n_out = buf ? n_out : nullptr;
    // the descriptors and the captured sequence depend on tracing
    batch_descriptors(b);
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    return 0;
}

int trk_batch_run(trk_batch *b, uint64_t seed, int64_t first_iter, trk_tallies *const *outs)
{
    if (!b) return TRK_ERR_ARG;
    const int n = (int)b->impacts.size();
    if (first_iter < 0 || first_iter + n > 0xFFFFFFFFll) return berr(b, TRK_ERR_ARG, "invalid first_iter");
    for (int i = 0; i < n; ++i) {
        const trk_ctx *c = b->impacts[i];
        if (c->tables_dirty || (c->base && c->base->tables_dirty))
            return berr(b, TRK_ERR_STATE, "impact %d: tables were set after the batch was made; finish them and make "
                                          "the batch again", i);
    }
    if (outs) {
        const int rc = check_outs(b, outs);
        if (rc) return rc;
    }
    return batch_launch(b, seed, first_iter, outs);
}

int trk_batch_fetch(trk_batch *b, trk_tallies *const *outs)
{
    if (!b || !outs) return TRK_ERR_ARG;
    if (!b->has_run) return berr(b, TRK_ERR_STATE, "trk_batch_run must run before its tallies are fetched");
    const int rc = check_outs(b, outs);
    return rc ? rc : batch_fetch(b, outs);
}

int trk_batch_tune(trk_batch *b, uint64_t seed, int64_t first_iter)
{
    if (!b) return TRK_ERR_ARG;
    // candidates: threads per workgroup (waves per CU) x events per particle
    // per round; each timed on the batch itself after one run to warm up
    const int blocks[3] = {512, 640, 768};
    const int depths[3] = {256, 512, 1024};
    b->tune_log.clear();
    double best = 1e300;
    int bb = b->block, bd = b->round_events;
    for (int blk : blocks)
        for (int dep : depths) {
            b->block = blk;
            b->round_events = dep;
            batch_descriptors(b);
            if (batch_shape(b)) continue;
            if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
            if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
            int rc = batch_launch(b, seed, first_iter, nullptr);
            if (!rc) rc = batch_launch(b, seed, first_iter, nullptr);
            if (rc) return rc;
            b->tune_log.push_back(TuneEntry{blk, dep, b->last_ms});
            if (b->last_ms < best) { best = b->last_ms; bb = blk; bd = dep; }
        }
    b->block = bb;
    b->round_events = bd;
    batch_descriptors(b);
    if (batch_shape(b)) return TRK_ERR_CAPACITY;
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    return 0;
}

int trk_batch_config(const trk_batch *b, int32_t *block, int32_t *round_events, int64_t *workgroups, int32_t *capacity,
                     int32_t *groups)
{
    if (!b) return TRK_ERR_ARG;
    if (block) *block = b->block;
    if (round_events) *round_events = b->round_events;
    if (workgroups) *workgroups = b->blocks;
    if (capacity) *capacity = b->capacity;
    if (groups) *groups = (int32_t)b->groups.size();
    return 0;
}

int trk_batch_tune_log(const trk_batch *b, int32_t *block, int32_t *round_events, double *ms, int32_t cap, int32_t *n)
{
    if (!b || !n) return TRK_ERR_ARG;
    *n = (int32_t)b->tune_log.size();
    for (int i = 0; i < *n && i < cap; ++i) {
        if (block) block[i] = b->tune_log[i].block;
        if (round_events) round_events[i] = b->tune_log[i].round_events;
        if (ms) ms[i] = b->tune_log[i].ms;
    }
    return 0;
}

int trk_batch_timing(const trk_batch *b, double *ms, int64_t *n_events, int32_t *rounds)
{
    if (!b) return TRK_ERR_ARG;
    if (ms) *ms = b->last_ms;
    if (n_events) for (int k = 0; k < 8; ++k) n_events[k] = b->last_events[k];
    if (rounds) *rounds = b->rounds;
    return 0;
}

}  // extern "C"
