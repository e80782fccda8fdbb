// trk_tables.h -- device table layout and the table-build interface.
#pragma once

#include <stdint.h>
#include <vector>
#include <string>
#include "trk_dev_common.h"
#include "trk_dsf.h"

namespace trk {

// Material view handed to kernels by value (pointers are device memory).
struct MatDev {
    int nsh, vb, n_atoms, n_dos;
    double Egap, Mtarget, Zt, sumPers;
    double El_eff_mass, v_f, temp, N_VB_el, At_Dens, cut_off;
    double hole_mass, Layer;
    int kind_of_DR, use_k, n_ph, kind_of_EMFP, plasmon_Emax;
    const int *osc_off;          // [nsh+1]
    const double *oA, *oE0, *oG; // oscillators
    const double *phA, *phE0, *phG;
    const double *k, *Effm;      // DOS k / Eff_m (or the inverted pair for metals)
    int nk;
    const double *dosE, *dos, *dosI, *dosM;
    // per shell
    const double *Ip, *Nel, *Auger, *Radiat;
    const int *atom_of, *shl_off;
    const int *Zat;
    const double *Mass, *Pers;
    // per-atom constants of the screened-Rutherford (Mott) sampling, computed
    // once on the host: Z, Z^(2/3), the atom's rest energy [eV], Pers
    const double *mott_Z, *mott_Z23, *mott_Mct2, *mott_Pers;
    double ion_mass, ion_E, ion_Zeff0;
    int ion_Z, ion_kind_zeff;
    double ion_fixed_zeff;
};

// One table kind on the device.  Per-shell arrays are shell-major [nsh][n].
// The ion's energy-transfer draw RN in [b/kShiBuckets, (b+1)/kShiBuckets)
// finds its entry of a shell's inverse-path row between the entries that
// the bucket's two ends find (k_shi_buckets).
constexpr int kShiBuckets = 64;

struct TableDev {
    int n = 0;
    int nsh = 0;
    double *E = nullptr;
    double *L = nullptr;
    double *dEdx = nullptr;
    // cumulative rows: row (shell s, energy i) at offset row_off[s*n+i]
    int64_t *row_off = nullptr;   // [nsh*n+1], device
    unsigned char *row_mono = nullptr;  // [nsh*n], 1 if the row's lam never increases (no NaN)
    double *hw = nullptr;
    double *lam = nullptr;
    std::vector<int64_t> row_off_h;   // host copy of offsets
    std::vector<int> has_rows;        // per shell
    // SHI_DIFF: per-shell rows of own length (E, L)
    std::vector<int> dn;
    std::vector<int64_t> doff;
    double *dE = nullptr;
    double *dL = nullptr;
    double *dInv = nullptr;           // 1/dL, for the ion's transfer-energy search
    int *dNonMono = nullptr;          // [nsh] nonzero if 1/dL of the shell's row ever decreases
};

struct TableSet {
    TableDev tab[7];
    double *el_imfp = nullptr;     // fused El_IMFP(2,:) lambda column
    double *hole_imfp = nullptr;
    double *phot_imfp = nullptr;
    double *shi_path = nullptr;
    // trk_log of every grid node and of every fused / electron-elastic / hole-
    // elastic mean free path: the log-log interpolation of a lookup then
    // needs only the log of the particle energy (same values, same result)
    double *ln_E[7] = {};          // per kind, [n]
    double *ln_el_imfp = nullptr, *ln_hole_imfp = nullptr, *ln_Lel = nullptr, *ln_LHel = nullptr;
    double *ln_shi_path = nullptr;  // log of the ion's fused path
    // per kind: logs of the per-shell mean free paths [nsh*n] (inelastic
    // kinds) and of the cumulative rows' transfer energies and lambdas
    // (same layout as hw / lam)
    double *ln_L[7] = {}, *ln_hw[7] = {}, *ln_lam[7] = {};
    // the material kinds (electron/hole inelastic and elastic, photon) and
    // their derived arrays belong to another table set (an ion context built
    // on a material context): free_tables leaves them alone
    bool shared_material = false;
    // run constants fixed by the tables, set at the end of build_tables
    bool grids_increasing = false;  // electron, hole and ion energy grids strictly increasing
    double shi_path0 = 0.0;         // the ion's fused mean free path at its initial energy [A]
    int *sd_n = nullptr;            // [nsh] SHI_DIFF row length per shell (device)
    int64_t *sd_off = nullptr;      // [nsh+1] SHI_DIFF row offsets (device)
    int *sd_M = nullptr;            // [2*nsh] transfer constants of the path's precision (k_shi_consts)
    double *sd_c = nullptr;         // [2*nsh]
    int *sd_bk = nullptr;           // [nsh][kShiBuckets + 1] transfer-row index by draw (k_shi_buckets)
    bool sd_consts = false;         // sd_M / sd_c / sd_bk filled
    int64_t launches = 0;
    double build_seconds = 0.0;
    int64_t dcs_evals = 0;         // number of q-integrals evaluated
    // DSF elastic scattering (kind_of_EMFP = 2) of electrons [0] and holes
    // [1]: the sampling tables (their E and Lt alias the elastic tables') and
    // the one allocation per particle behind dE, dLe, dLa, Le, La
    DsfDev dsf[2] = {};
    double *dsf_buf[2] = {nullptr, nullptr};
};

struct HostMat;  // defined in trk_ctx.h

// mat: when given, the material kinds are taken from that (built) table set
// and only the ion's tables are computed.  delta_alpha: the oscillators'
// Delta-CDF weights (device memory) when kind_of_DR = 4.
int build_tables(const HostMat &hm, const MatDev &md, TableSet &ts, hipStream_t st, std::string &err,
                 const TableSet *mat = nullptr, const double *delta_alpha = nullptr);
void free_tables(TableSet &ts);
int refresh_derived(const HostMat &hm, TableSet &ts, hipStream_t st, std::string &err);

}  // namespace trk
