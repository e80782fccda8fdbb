// trk_api.cpp -- the extern "C" entry points declared in include/trekis_gpu.h.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include "trk_ctx.h"

namespace trk {
int run_mc(trk_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter, const double *Out_R, const double *Out_V,
           int NR, trk_tallies *out);
int run_mc_noalloc(trk_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter, const double *Out_R,
                   const double *Out_V, int NR, trk_tallies *out);
void mc_free(McState *s);
}  // namespace trk

using namespace trk;

#define HIPC(x)                                                                          \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) return set_error(c, TRK_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

template <class T>
static int upload(trk_ctx *c, T **dst, const T *src, size_t n)
{
    size_t bytes = (n ? n : 1) * sizeof(T);
    HIPC(hipMalloc((void **)dst, bytes));
    if (n) HIPC(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

static int validate(const trk_material *m, const trk_ion *ion, const trk_options *o, std::string &why)
{
    if (m->n_atoms < 1 || m->n_atoms > TRK_MAX_ATOMS) { why = "n_atoms out of range"; return 1; }
    int nsh = 0, nosc = 0;
    for (int a = 0; a < m->n_atoms; ++a) {
        // other atoms may give all their electrons to the valence band (the
        // reference's hydrocarbon files list no shell for hydrogen)
        if (m->n_shells[a] < (a == 0 ? 1 : 0)) { why = "the first atom needs at least one shell (the valence band)"; return 1; }
        nsh += m->n_shells[a];
    }
    if (nsh > TRK_MAX_SHELLS) { why = "too many shells"; return 1; }
    for (int s = 0; s < nsh; ++s) {
        if (m->n_osc[s] < 1) { why = "every shell needs at least one CDF oscillator"; return 1; }
        nosc += m->n_osc[s];
        if (!(m->Ip[s] > 0.0)) { why = "ionisation potentials must be positive"; return 1; }
    }
    if (nosc > TRK_MAX_OSC) { why = "too many oscillators"; return 1; }
    for (int s = 0; s < nsh; ++s) {
        if (m->kocs[s] < 0 || m->kocs[s] > 2) { why = "kocs must be 0 or 1 (CDF) or 2 (BEB)"; return 1; }
        if (m->kocs[s] == 2 && !(m->Ek[s] >= 0.0 && m->Ek[s] < 1.0e30))
            { why = "a BEB shell needs the mean kinetic energy Ek of its electrons (>= 0, finite)"; return 1; }
    }
    if (m->n_phonon < 0 || m->n_phonon > TRK_MAX_PHONON) { why = "n_phonon out of range"; return 1; }
    if (o->kind_of_EMFP == 1 && m->n_phonon < 1) { why = "phonon-CDF elastic scattering needs phonon oscillators"; return 1; }
    if (o->kind_of_EMFP < -1 || o->kind_of_EMFP > 2)
        { why = "kind_of_EMFP must be -1 (none), 0 (Mott), 1 (phonon CDF) or 2 (DSF)"; return 1; }
    if (m->n_dos < 2 || !m->dos_E || !m->dos_DOS || !m->dos_int_DOS || !m->dos_Eff_m) { why = "a VB density of states is required"; return 1; }
    if (m->El_eff_mass == 0.0 && (!m->dos_k || !m->dos_k_inv || !m->dos_Eff_m_inv)) { why = "El_eff_mass = 0 needs k and Eff_m arrays"; return 1; }
    if (m->work_function > 0.0) {
        // the reader turns emission off when a barrier dimension is not
        // positive (RF:378-381); a caller passing one is refused instead
        double E1 = 0.0, g = 0.0;
        if (!(m->bar_length > 0.0) || !(m->bar_height > m->work_function) ||
            trk_barrier_parameters(m->work_function, m->bar_length, m->bar_height, &E1, &g) != 0) {
            why = "emission (work_function > 0) needs bar_length > 0 and bar_height > work_function, "
                  "with a finite barrier transmission (barrier_parameters, MC:2094-2116)";
            return 1;
        }
    }
    if (!(m->At_Dens > 0.0) || !(m->Layer > 0.0)) { why = "At_Dens and Layer must be positive"; return 1; }
    if (ion->Zat < 1 || !(ion->Mass > 0.0) || !(ion->E > 0.0)) { why = "invalid ion"; return 1; }
    if (ion->Kind_ion != 0 && ion->Kind_ion != 1) { why = "Kind_ion must be 0 (point charge) or 1 (Brandt-Kitagawa)"; return 1; }
    if (!(o->Tim > 0.0) || !(o->dt > 0.0)) { why = "Tim and dt must be positive"; return 1; }
    if (o->dt_flag >= 1 && o->dt < 2.0) { why = "a logarithmic time grid needs dt >= 2 (RF:299)"; return 1; }
    if (o->kind_of_DR < 1 || o->kind_of_DR > 4)
        { why = "kind_of_DR must be 1 (free electron), 2 (plasmon pole), 3 (Ritchie) or 4 (Delta-CDF)"; return 1; }
    if (o->precision != 0 && o->precision != 1) { why = "precision must be 0 (FP64) or 1 (mixed)"; return 1; }
    if (o->quadrature != 0 && o->quadrature != 1) { why = "quadrature must be 0 (reference, FP64) or 1 (FP32 compensated)"; return 1; }
    if (o->quadrature == 1 && !(m->El_eff_mass > 0.0))
        { why = "the FP32 table quadrature needs a constant electron effective mass (El_eff_mass > 0)"; return 1; }
    return 0;
}

extern "C" {

int trk_ctx_create(const trk_material *mat, const trk_ion *ion, const trk_options *opt, trk_ctx **out)
{
    if (!out) return TRK_ERR_ARG;
    *out = nullptr;
    if (!mat || !ion || !opt) return TRK_ERR_ARG;
    trk_ctx *c = new (std::nothrow) trk_ctx();
    if (!c) return TRK_ERR_ARG;
    *out = c;
    std::string why;
    if (validate(mat, ion, opt, why)) return set_error(c, TRK_ERR_ARG, "%s", why.c_str());
    HostMat &h = c->hm;
    h.m = *mat;
    {
        // the reference's decay-time rules on the library's copy: with them
        // no radiative decay is drawn when photons are off (there is then no
        // photon table to take the photon's free path from)
        char buf[256];
        if (trk_check_atomic_parameters(&h.m, opt->include_photons, buf, sizeof buf))
            return set_error(c, TRK_ERR_ARG, "%s", buf);
    }
    h.ion = *ion;
    h.o = *opt;
    const int nd = mat->n_dos;
    auto cp = [nd](const double *p, std::vector<double> &v) { v.assign(p ? p : (const double *)nullptr, p ? p + nd : nullptr); };
    cp(mat->dos_E, h.dosE); cp(mat->dos_DOS, h.dos); cp(mat->dos_int_DOS, h.dosI); cp(mat->dos_k, h.dosk);
    cp(mat->dos_Eff_m, h.dosM); cp(mat->dos_k_inv, h.doskI); cp(mat->dos_Eff_m_inv, h.dosMI);
    h.m.dos_E = h.m.dos_DOS = h.m.dos_int_DOS = h.m.dos_k = h.m.dos_Eff_m = h.m.dos_k_inv = h.m.dos_Eff_m_inv = nullptr;
    int s = 0, o = 0;
    for (int a = 0; a < mat->n_atoms; ++a) {
        h.shl_off[a] = s;
        for (int k = 0; k < mat->n_shells[a]; ++k, ++s) {
            h.atom_of[s] = a;
            h.osc_off[s] = o;
            o += mat->n_osc[s];
        }
    }
    h.shl_off[mat->n_atoms] = s;
    h.osc_off[s] = o;
    h.nsh = s;
    h.vb = mat->n_shells[0] - 1;
    h.Egap = mat->Ip[h.vb];
    double sm = 0, sp = 0, sz = 0;
    for (int a = 0; a < mat->n_atoms; ++a) {
        sm = sm + mat->Mass[a] * mat->Pers[a];
        sp = sp + mat->Pers[a];
        sz = sz + (double)mat->Zat[a] * mat->Pers[a];
    }
    h.Mtarget = (kMp * sm) / sp;
    h.sumPers = sp;
    h.Zt = sz / sp;
    h.e_emax = (opt->e_grid_Emax > 0.0) ? opt->e_grid_Emax : ((175.6e6 * 2.0) / 1836.0);
    h.shi_Zeff0 = shi_zeff(h, ion->E);
    build_time_grid(h);

    HIPC(hipGetDevice(&c->device));
    HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    MatBuffers &b = c->mb;
    std::vector<int> atom_of(h.atom_of, h.atom_of + s), shl_off(h.shl_off, h.shl_off + mat->n_atoms + 1),
        osc_off(h.osc_off, h.osc_off + s + 1), Zat(mat->Zat, mat->Zat + mat->n_atoms);
    int rc = 0;
    rc |= upload(c, &b.osc_off, osc_off.data(), osc_off.size());
    rc |= upload(c, &b.atom_of, atom_of.data(), atom_of.size());
    rc |= upload(c, &b.shl_off, shl_off.data(), shl_off.size());
    rc |= upload(c, &b.Zat, Zat.data(), Zat.size());
    rc |= upload(c, &b.oA, mat->osc_A, (size_t)o);
    rc |= upload(c, &b.oE0, mat->osc_E0, (size_t)o);
    rc |= upload(c, &b.oG, mat->osc_Gamma, (size_t)o);
    {
        std::vector<int> kocs((size_t)s);
        for (int sh = 0; sh < s; ++sh) kocs[sh] = (mat->kocs[sh] == 2) ? 2 : 1;
        rc |= upload(c, &b.kocs, kocs.data(), kocs.size());
        rc |= upload(c, &b.Ek, mat->Ek, (size_t)s);
    }
    if (opt->kind_of_DR == 4) {
        // the delta peaks' weights, each oscillator's k-sum above its
        // shell's edge (define_alpha, RF:1385-1388, 1572-1574)
        std::vector<double> alpha((size_t)o);
        for (int sh = 0; sh < s; ++sh)
            for (int k = h.osc_off[sh]; k < h.osc_off[sh + 1]; ++k)
                alpha[k] = trk_delta_alpha(mat->osc_A[k], mat->osc_E0[k], mat->osc_Gamma[k], mat->Ip[sh]);
        rc |= upload(c, &b.oAlpha, alpha.data(), alpha.size());
    }
    rc |= upload(c, &b.phA, mat->ph_A, (size_t)mat->n_phonon);
    rc |= upload(c, &b.phE0, mat->ph_E0, (size_t)mat->n_phonon);
    rc |= upload(c, &b.phG, mat->ph_Gamma, (size_t)mat->n_phonon);
    rc |= upload(c, &b.Ip, mat->Ip, (size_t)s);
    rc |= upload(c, &b.Nel, mat->Nel, (size_t)s);
    rc |= upload(c, &b.Auger, h.m.Auger, (size_t)s);
    rc |= upload(c, &b.Radiat, h.m.Radiat, (size_t)s);
    rc |= upload(c, &b.Mass, mat->Mass, (size_t)mat->n_atoms);
    rc |= upload(c, &b.Pers, mat->Pers, (size_t)mat->n_atoms);
    {
        const int na = mat->n_atoms;
        std::vector<double> mott((size_t)4 * na);
        for (int a = 0; a < na; ++a) {
            const double Z = (double)mat->Zat[a];
            const double Mat = mat->Mass[a] * kMp;
            mott[a] = Z;
            mott[na + a] = trk_pow(Z, 2.0 / 3.0);
            mott[2 * na + a] = ((Mat * kC) * kC) / kE;
            mott[3 * na + a] = mat->Pers[a];
        }
        rc |= upload(c, &b.mott, mott.data(), mott.size());
    }
    rc |= upload(c, &b.dosE, h.dosE.data(), h.dosE.size());
    rc |= upload(c, &b.dos, h.dos.data(), h.dos.size());
    rc |= upload(c, &b.dosI, h.dosI.data(), h.dosI.size());
    rc |= upload(c, &b.dosM, h.dosM.data(), h.dosM.size());
    const bool metal = h.Egap < 0.2;
    if (mat->El_eff_mass == 0.0) {
        rc |= upload(c, &b.k, metal ? h.doskI.data() : h.dosk.data(), (size_t)nd);
        rc |= upload(c, &b.Effm, metal ? h.dosMI.data() : h.dosM.data(), (size_t)nd);
    }
    if (rc) return TRK_ERR_DEVICE;
    MatDev &md = c->md;
    md.nsh = s; md.vb = h.vb; md.n_atoms = mat->n_atoms; md.n_dos = nd;
    md.Egap = h.Egap; md.Mtarget = h.Mtarget; md.Zt = h.Zt; md.sumPers = h.sumPers;
    md.El_eff_mass = mat->El_eff_mass; md.v_f = mat->v_f; md.temp = mat->temp;
    md.N_VB_el = mat->N_VB_el; md.At_Dens = mat->At_Dens; md.cut_off = mat->cut_off;
    md.hole_mass = mat->hole_mass; md.Layer = mat->Layer;
    md.kind_of_DR = opt->kind_of_DR; md.use_k = (mat->El_eff_mass == 0.0); md.n_ph = mat->n_phonon;
    md.kind_of_EMFP = opt->kind_of_EMFP; md.plasmon_Emax = opt->plasmon_Emax;
    md.osc_off = b.osc_off; md.oA = b.oA; md.oE0 = b.oE0; md.oG = b.oG;
    md.phA = b.phA; md.phE0 = b.phE0; md.phG = b.phG;
    md.k = b.k; md.Effm = b.Effm; md.nk = nd;
    md.dosE = b.dosE; md.dos = b.dos; md.dosI = b.dosI; md.dosM = b.dosM;
    md.Ip = b.Ip; md.Nel = b.Nel; md.Auger = b.Auger; md.Radiat = b.Radiat;
    md.atom_of = b.atom_of; md.shl_off = b.shl_off; md.Zat = b.Zat; md.Mass = b.Mass; md.Pers = b.Pers;
    md.mott_Z = b.mott;
    md.mott_Z23 = b.mott + mat->n_atoms;
    md.mott_Mct2 = b.mott + 2 * mat->n_atoms;
    md.mott_Pers = b.mott + 3 * mat->n_atoms;
    md.ion_mass = ion->Mass; md.ion_E = ion->E; md.ion_Zeff0 = h.shi_Zeff0;
    md.ion_Z = ion->Zat; md.ion_kind_zeff = ion->Kind_Zeff; md.ion_fixed_zeff = ion->fixed_Zeff;
    return 0;
}

void trk_ctx_destroy(trk_ctx *c)
{
    if (!c) return;
    free_tables(c->ts);
    mc_free(c->mc);
    MatBuffers &b = c->mb;
    void *ps[] = {b.osc_off, b.atom_of, b.shl_off, b.Zat, b.oA, b.oE0, b.oG, b.phA, b.phE0, b.phG, b.k, b.Effm,
                  b.dosE, b.dos, b.dosI, b.dosM, b.Ip, b.Nel, b.Auger, b.Radiat, b.Mass, b.Pers, b.mott, b.oAlpha,
                  b.kocs, b.Ek};
    for (void *p : ps) if (p) (void)hipFree(p);
    for (hipEvent_t e : c->pe) if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int trk_last_error(const trk_ctx *c, char *buf, size_t n)
{
    if (!c || !buf || n == 0) return TRK_ERR_ARG;
    std::strncpy(buf, c->err.c_str(), n - 1);
    buf[n - 1] = 0;
    return 0;
}

int trk_build_tables(trk_ctx *c)
{
    if (!c) return TRK_ERR_ARG;
    std::string err;
    if (c->base && !c->base->tables_ready)
        return set_error(c, TRK_ERR_STATE, "the material context's tables are not built");
    int rc = build_tables(c->hm, c->md, c->ts, c->stream, err, c->base ? &c->base->ts : nullptr, c->mb.oAlpha);
    if (rc) return set_error(c, rc, "table build failed: %s", err.c_str());
    c->tables_ready = true;
    return 0;
}

int trk_table_size(const trk_ctx *c, int kind, int shell, int32_t *n_out)
{
    if (!c || !n_out || kind < 0 || kind > 6 || !c->tables_ready) return TRK_ERR_ARG;
    const TableDev &t = c->ts.tab[kind];
    if (kind == TRK_TAB_SHI_DIFF) {
        if (shell < 0 || shell >= c->hm.nsh) return TRK_ERR_ARG;
        *n_out = t.dn[shell];
        return 0;
    }
    *n_out = t.n;
    return 0;
}

int trk_table_get(const trk_ctx *c, int kind, int shell, double *E, double *L, double *dEdx, int32_t n)
{
    if (!c || kind < 0 || kind > 6 || !c->tables_ready) return TRK_ERR_ARG;
    const TableDev &t = c->ts.tab[kind];
    trk_ctx *cc = const_cast<trk_ctx *>(c);
    if (kind == TRK_TAB_SHI_DIFF) {
        if (shell < 0 || shell >= c->hm.nsh || n != t.dn[shell]) return TRK_ERR_ARG;
        int64_t o = t.doff[shell];
        if (E && n) { hipError_t e = hipMemcpy(E, t.dE + o, sizeof(double) * n, hipMemcpyDeviceToHost); if (e) return set_error(cc, TRK_ERR_DEVICE, "copy"); }
        if (L && n) { hipError_t e = hipMemcpy(L, t.dL + o, sizeof(double) * n, hipMemcpyDeviceToHost); if (e) return set_error(cc, TRK_ERR_DEVICE, "copy"); }
        return 0;
    }
    if (n != t.n || shell < 0 || shell >= t.nsh) return TRK_ERR_ARG;
    if (E && hipMemcpy(E, t.E, sizeof(double) * n, hipMemcpyDeviceToHost)) return set_error(cc, TRK_ERR_DEVICE, "copy");
    if (L && hipMemcpy(L, t.L + (size_t)shell * n, sizeof(double) * n, hipMemcpyDeviceToHost)) return set_error(cc, TRK_ERR_DEVICE, "copy");
    if (dEdx && hipMemcpy(dEdx, t.dEdx + (size_t)shell * n, sizeof(double) * n, hipMemcpyDeviceToHost)) return set_error(cc, TRK_ERR_DEVICE, "copy");
    return 0;
}

int trk_row_size(const trk_ctx *c, int kind, int shell, int32_t i_E, int32_t *n_out)
{
    if (!c || !n_out || kind < 0 || kind > 3 || !c->tables_ready) return TRK_ERR_ARG;
    const TableDev &t = c->ts.tab[kind];
    if (shell < 0 || shell >= t.nsh || i_E < 0 || i_E >= t.n || !t.has_rows[shell]) return TRK_ERR_ARG;
    size_t slot = (size_t)shell * t.n + i_E;
    *n_out = (int32_t)(t.row_off_h[slot + 1] - t.row_off_h[slot]);
    return 0;
}

int trk_row_get(const trk_ctx *c, int kind, int shell, int32_t i_E, double *hw, double *lam, int32_t n)
{
    int32_t len = 0;
    int rc = trk_row_size(c, kind, shell, i_E, &len);
    if (rc) return rc;
    if (n != len) return TRK_ERR_ARG;
    const TableDev &t = c->ts.tab[kind];
    size_t slot = (size_t)shell * t.n + i_E;
    int64_t o = t.row_off_h[slot];
    trk_ctx *cc = const_cast<trk_ctx *>(c);
    if (hw && hipMemcpy(hw, t.hw + o, sizeof(double) * n, hipMemcpyDeviceToHost)) return set_error(cc, TRK_ERR_DEVICE, "copy");
    if (lam && hipMemcpy(lam, t.lam + o, sizeof(double) * n, hipMemcpyDeviceToHost)) return set_error(cc, TRK_ERR_DEVICE, "copy");
    return 0;
}

// Tables replaced from host arrays: same grid, same row lengths as built.
// DSF_DEMFP / DSF_DEMFP_H (Objects.f90:173-178) as the reference's reader
// leaves them, kept on the host until trk_build_tables places them
int trk_dsf_set(trk_ctx *c, int32_t particle, int32_t nE, int32_t nW, const double *E, const double *dE,
                const double *dL, const double *dL_emit, const double *dL_absorb)
{
    if (!c) return TRK_ERR_ARG;
    HostMat &h = c->hm;
    if (c->base) return set_error(c, TRK_ERR_ARG, "an ion context shares its material's tables; set the DSF there");
    if (h.o.kind_of_EMFP != 2)
        return set_error(c, TRK_ERR_STATE, "the context does not scatter elastically on DSF tables (kind_of_EMFP = 2)");
    if (c->tables_ready) return set_error(c, TRK_ERR_STATE, "DSF tables are set before trk_build_tables");
    if ((particle != 0 && particle != 1) || nE < 2 || nW < 2 || !E || !dE || !dL || !dL_emit || !dL_absorb)
        return set_error(c, TRK_ERR_ARG, "invalid trk_dsf_set arguments");
    for (int i = 0; i < nE; ++i) {
        if (!std::isfinite(E[i]) || !(E[i] > 0.0) || (i > 0 && !(E[i] > E[i - 1])))
            return set_error(c, TRK_ERR_ARG, "DSF energies must be positive, finite and increasing (energy %d)", i);
    }
    const size_t cells = (size_t)nE * nW;
    for (size_t k = 0; k < cells; ++k) {
        if (!std::isfinite(dE[k]) || dL[k] != dL[k] || dL_emit[k] != dL_emit[k] || dL_absorb[k] != dL_absorb[k])
            return set_error(c, TRK_ERR_ARG, "DSF rows hold a NaN or an infinite transfer (entry %zu)", k);
        // cumulative mean free paths: positive (1e30 where nothing scatters)
        if (!(dL[k] > 0.0) || !(dL_emit[k] > 0.0) || !(dL_absorb[k] > 0.0))
            return set_error(c, TRK_ERR_ARG, "DSF mean free paths must be positive (entry %zu)", k);
    }
    h.dsf_nW[particle] = nW;
    h.dsfE[particle].assign(E, E + nE);
    h.dsfdE[particle].assign(dE, dE + cells);
    h.dsfL[particle].assign(dL, dL + cells);
    h.dsfLe[particle].assign(dL_emit, dL_emit + cells);
    h.dsfLa[particle].assign(dL_absorb, dL_absorb + cells);
    return 0;
}

static int check_settable(trk_ctx *c, int kind)
{
    if (!c->tables_ready) return set_error(c, TRK_ERR_STATE, "trk_build_tables must run before tables are set");
    if (kind < 0 || kind > 6) return set_error(c, TRK_ERR_ARG, "no table kind %d", kind);
    if (c->base && kind != TRK_TAB_SHI && kind != TRK_TAB_SHI_DIFF)
        return set_error(c, TRK_ERR_ARG, "an ion context shares its material's tables; set them on the material context");
    return 0;
}

int trk_table_set(trk_ctx *c, int kind, int shell, const double *E, const double *L, const double *dEdx, int32_t n)
{
    if (!c) return TRK_ERR_ARG;
    int rc = check_settable(c, kind);
    if (rc) return rc;
    TableDev &t = c->ts.tab[kind];
    hipStream_t st = c->stream;
    if (kind == TRK_TAB_SHI_DIFF) {
        if (shell < 0 || shell >= c->hm.nsh || n != t.dn[shell])
            return set_error(c, TRK_ERR_ARG, "the ion's transfer row of shell %d has %d nodes", shell,
                             (shell >= 0 && shell < c->hm.nsh) ? t.dn[shell] : -1);
        const int64_t o = t.doff[shell];
        if (E && hipMemcpyAsync(t.dE + o, E, sizeof(double) * n, hipMemcpyHostToDevice, st)) return set_error(c, TRK_ERR_DEVICE, "copy");
        if (L && hipMemcpyAsync(t.dL + o, L, sizeof(double) * n, hipMemcpyHostToDevice, st)) return set_error(c, TRK_ERR_DEVICE, "copy");
    } else {
        if (shell < 0 || shell >= t.nsh || n != t.n)
            return set_error(c, TRK_ERR_ARG, "table kind %d has %d energy nodes and %d shells", kind, t.n, t.nsh);
        if (E) {
            // the grid is the one the tables were built on: it must match bit for bit
            std::vector<double> g((size_t)n);
            if (hipMemcpy(g.data(), t.E, sizeof(double) * n, hipMemcpyDeviceToHost)) return set_error(c, TRK_ERR_DEVICE, "copy");
            for (int i = 0; i < n; ++i)
                if (!(g[i] == E[i])) return set_error(c, TRK_ERR_ARG, "energy node %d differs from the built grid", i);
        }
        if (L && hipMemcpyAsync(t.L + (size_t)shell * n, L, sizeof(double) * n, hipMemcpyHostToDevice, st))
            return set_error(c, TRK_ERR_DEVICE, "copy");
        if (dEdx && hipMemcpyAsync(t.dEdx + (size_t)shell * n, dEdx, sizeof(double) * n, hipMemcpyHostToDevice, st))
            return set_error(c, TRK_ERR_DEVICE, "copy");
    }
    if (hipStreamSynchronize(st)) return set_error(c, TRK_ERR_DEVICE, "synchronize");
    c->tables_dirty = true;
    return 0;
}

int trk_row_set(trk_ctx *c, int kind, int shell, int32_t i_E, const double *hw, const double *lam, int32_t n)
{
    if (!c) return TRK_ERR_ARG;
    int rc = check_settable(c, kind);
    if (rc) return rc;
    int32_t len = 0;
    rc = trk_row_size(c, kind, shell, i_E, &len);
    if (rc) return set_error(c, TRK_ERR_ARG, "no cumulative row (kind %d, shell %d, node %d)", kind, shell, i_E);
    if (n != len) return set_error(c, TRK_ERR_ARG, "the row has %d entries, not %d", len, n);
    const TableDev &t = c->ts.tab[kind];
    const int64_t o = t.row_off_h[(size_t)shell * t.n + i_E];
    if (hw && hipMemcpy(t.hw + o, hw, sizeof(double) * n, hipMemcpyHostToDevice)) return set_error(c, TRK_ERR_DEVICE, "copy");
    if (lam && hipMemcpy(t.lam + o, lam, sizeof(double) * n, hipMemcpyHostToDevice)) return set_error(c, TRK_ERR_DEVICE, "copy");
    c->tables_dirty = true;
    return 0;
}

int trk_tables_finish(trk_ctx *c)
{
    if (!c) return TRK_ERR_ARG;
    if (!c->tables_ready) return set_error(c, TRK_ERR_STATE, "trk_build_tables must run first");
    std::string err;
    const int rc = refresh_derived(c->hm, c->ts, c->stream, err);
    if (rc) return set_error(c, rc, "deriving tables failed: %s", err.c_str());
    c->tables_dirty = false;
    return 0;
}

int trk_time_nodes(const trk_ctx *c, int32_t *Nt_out)
{
    if (!c || !Nt_out) return TRK_ERR_ARG;
    *Nt_out = c->hm.Nt;
    return 0;
}

static int check_run_args(trk_ctx *c, int64_t nmc, int64_t first_iter, const double *Out_R, const double *Out_V,
                          int32_t NR, const trk_tallies *out)
{
    if (!c) return TRK_ERR_ARG;
    if (!c->tables_ready) return set_error(c, TRK_ERR_STATE, "trk_build_tables must run before trk_run_mc");
    if (c->tables_dirty || (c->base && c->base->tables_dirty))
        return set_error(c, TRK_ERR_STATE, "tables were set: trk_tables_finish must run before trk_run_mc");
    if (!out || !Out_R || !Out_V || NR < 1 || nmc < 1 || first_iter < 0 || first_iter + nmc > 0xFFFFFFFFll)
        return set_error(c, TRK_ERR_ARG, "invalid trk_run_mc arguments");
    if (out->Nt != c->hm.Nt) return set_error(c, TRK_ERR_ARG, "tallies have Nt=%d, the run has %d time nodes", out->Nt, c->hm.Nt);
    if (out->NR != NR) return set_error(c, TRK_ERR_ARG, "tallies have NR=%d, the radial grid has %d bins", out->NR, NR);
    if (out->Out_nh || out->Out_Eh || out->Out_Ehkin || out->Out_E_h) {
        int ms = 0;
        for (int a = 0; a < c->hm.m.n_atoms; ++a) ms = c->hm.m.n_shells[a] > ms ? c->hm.m.n_shells[a] : ms;
        if (out->max_shells < ms)
            return set_error(c, TRK_ERR_ARG, "tallies have max_shells=%d, an atom of the material has %d shells",
                             out->max_shells, ms);
    }
    if (out->Out_Eh_vs_E && out->n_dos != c->hm.m.n_dos)
        return set_error(c, TRK_ERR_ARG, "tallies have n_dos=%d, the material's DOS has %d points", out->n_dos, c->hm.m.n_dos);
    for (int j = 1; j < NR; ++j)
        if (!(Out_R[j] > Out_R[j - 1])) return set_error(c, TRK_ERR_ARG, "Out_R must be strictly increasing");
    return 0;
}

int trk_run_mc(trk_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter, const double *Out_R, const double *Out_V,
               int32_t NR, trk_tallies *out)
{
    if (!c) return TRK_ERR_ARG;
    const int rc = check_run_args(c, nmc, first_iter, Out_R, Out_V, NR, out);
    return rc ? rc : run_mc(c, nmc, seed, first_iter, Out_R, Out_V, NR, out);
}

int trk_run_mc_noalloc(trk_ctx *c, int64_t nmc, uint64_t seed, int64_t first_iter, const double *Out_R,
                       const double *Out_V, int32_t NR, trk_tallies *out)
{
    if (!c) return TRK_ERR_ARG;
    const int rc = check_run_args(c, nmc, first_iter, Out_R, Out_V, NR, out);
    return rc ? rc : run_mc_noalloc(c, nmc, seed, first_iter, Out_R, Out_V, NR, out);
}

int trk_last_timing(const trk_ctx *c, double *transport_s, double *total_s, int64_t *launches)
{
    if (!c) return TRK_ERR_ARG;
    if (transport_s) *transport_s = c->last_transport_s;
    if (total_s) *total_s = c->last_total_s;
    if (launches) *launches = c->last_launches;
    return 0;
}

int trk_set_profiling(trk_ctx *c, int on)
{
    if (!c) return TRK_ERR_ARG;
    if (on && !c->pe[0]) {
        if (hipEventCreate(&c->pe[0]) != hipSuccess || hipEventCreate(&c->pe[1]) != hipSuccess)
            return set_error(c, TRK_ERR_DEVICE, "hipEventCreate failed");
    }
    c->prof = on != 0;
    if (on) {
        c->k_ms[0] = c->k_ms[1] = 0.0;
        c->k_launches[0] = c->k_launches[1] = 0;
    }
    return 0;
}

int trk_kernel_time(const trk_ctx *c, int kernel, double *ms_total, int64_t *launches)
{
    if (!c || kernel < 0 || kernel > 1) return TRK_ERR_ARG;
    if (ms_total) *ms_total = c->k_ms[kernel];
    if (launches) *launches = c->k_launches[kernel];
    return 0;
}

int trk_set_trace(trk_ctx *c, double *buf, int64_t cap_records, int64_t *n_out)
{
    if (!c || cap_records < 0) return TRK_ERR_ARG;
    c->trace_host = buf;
    c->trace_cap = buf ? cap_records : 0;
    c->trace_n_out = n_out;
    return 0;
}

}  // extern "C"
