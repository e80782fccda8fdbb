// trk_ctx.h -- host-side context of the library (one per caller context).
#pragma once

#include <string>
#include <vector>
#include "trk_tables.h"

namespace trk {

// Host copy of the caller's descriptors plus derived scalars.
struct HostMat {
    trk_material m;
    trk_ion ion;
    trk_options o;
    std::vector<double> dosE, dos, dosI, dosk, dosM, doskI, dosMI;
    int nsh = 0;
    int shl_off[TRK_MAX_ATOMS + 1];
    int osc_off[TRK_MAX_SHELLS + 1];
    int atom_of[TRK_MAX_SHELLS];
    int vb = 0;
    double Egap = 0, Mtarget = 0, Zt = 0, sumPers = 0;
    double e_emax = 0;
    double shi_Zeff0 = 0;
    int Nt = 0;
    std::vector<double> tim_glob;
    // DSF elastic tables (kind_of_EMFP = 2) of electrons [0] and holes [1]:
    // nE energies x nW transfers, E-major
    int dsf_nW[2] = {0, 0};
    std::vector<double> dsfE[2], dsfdE[2], dsfL[2], dsfLe[2], dsfLa[2];
};

// Device copies of the material arrays behind MatDev.
struct MatBuffers {
    int *osc_off = nullptr, *atom_of = nullptr, *shl_off = nullptr, *Zat = nullptr;
    double *oA = nullptr, *oE0 = nullptr, *oG = nullptr;
    double *phA = nullptr, *phE0 = nullptr, *phG = nullptr;
    double *k = nullptr, *Effm = nullptr;
    double *dosE = nullptr, *dos = nullptr, *dosI = nullptr, *dosM = nullptr;
    double *Ip = nullptr, *Nel = nullptr, *Auger = nullptr, *Radiat = nullptr;
    double *Mass = nullptr, *Pers = nullptr;
    double *mott = nullptr;     // [4][n_atoms]: Z, Z^(2/3), rest energy, Pers
    double *oAlpha = nullptr;   // Delta-CDF weight of every oscillator (kind_of_DR = 4 only)
    int *kocs = nullptr;        // per shell: 1 CDF, 2 BEB (electrons and holes)
    double *Ek = nullptr;       // per shell: mean kinetic energy of its electrons [eV] (BEB)
};

struct McState;  // transport buffers, trk_mc.hip

}  // namespace trk

struct trk_ctx {
    trk::HostMat hm;
    trk::MatBuffers mb;
    trk::MatDev md;
    trk::TableSet ts;
    bool tables_ready = false;
    bool tables_dirty = false;     // tables set from host arrays, not yet finished
    hipStream_t stream = nullptr;
    int device = 0;
    std::string err;
    trk::McState *mc = nullptr;
    // an ion context (trk_ctx_create_ion): the material context whose
    // material arrays and tables this one uses (it must outlive this one)
    const trk_ctx *base = nullptr;
    // timing of the last run
    double last_transport_s = 0, last_total_s = 0;
    int64_t last_launches = 0;
    // per-kernel profiling (trk_set_profiling)
    bool prof = false;
    double k_ms[2] = {0.0, 0.0};
    int64_t k_launches[2] = {0, 0};
    hipEvent_t pe[2] = {nullptr, nullptr};
    // tracing
    double *trace_host = nullptr;
    int64_t trace_cap = 0;
    int64_t *trace_n_out = nullptr;
};

namespace trk {
int set_error(trk_ctx *c, int code, const char *fmt, ...);
double shi_zeff(const HostMat &h, double E);
double zeff_target(const HostMat &h, double Ee);
int grid_4cs(const trk_material &m, double Emin, double Emax, std::vector<double> &out);
void build_time_grid(HostMat &h);
}  // namespace trk
