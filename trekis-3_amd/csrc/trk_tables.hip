// trk_tables.hip -- Ritchie-Howie CDF quadrature tables on the MI355X.
//
// Replaces the table builders of the reference:
//   TotIMFP                Cross_sections.f90:881-1050   (electron/hole inelastic + aidCS rows)
//   Tot_EMFP               Cross_sections.f90:2966-3139  (phonon-CDF elastic + rows)
//   Elastic_cross_section  Cross_sections.f90:2892-2963  (Mott branch)
//   Tot_Phot_IMFP          Cross_sections.f90:833-878
//   SHI_TotIMFP            Cross_sections.f90:2452-2597  (table and diff_SHI_MFP rows)
//   drivers                Analytical_IMFPs.f90:1830-2239, 2510-2751
//   How_many_electrons     Monte_Carlo.f90:1902-2057 (fused totals)
//
// Design.  The reference integrates row by row on host threads: for each
// incident energy E_i and shell it walks an adaptive omega grid and, at every
// omega node, a geometric q grid (Simpson with the left-node 1/q weight).
// Here the work is split into four launches so that the expensive part -- the
// q integral behind each omega node, ~10^7 of them per material -- runs one
// per GPU thread:
//   1. plan   (one thread per row)    walk the omega recurrence, count nodes
//   2. nodes  (one thread per row)    write left node E_k and step dE_k
//   3. eval   (one thread per point)  q-integral at the row start, E_k+dE_k/2, E_k+dE_k
//   4. scan   (one thread per row)    Simpson sums in the reference's order,
//                                      cumulative rows, MFP, dE/dx
// Steps 1, 2 and 4 perform exactly the reference's floating-point operations
// in the reference's order, and step 3 is a pure function of its inputs, so
// the tables are bit-identical to a sequential evaluation.  The q loop is
// VALU-bound FP64 work (a correctly rounded division per oscillator per
// node); it has no dense contraction, so MFMA is not used.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "trk_ctx.h"
#include "trk_delta.h"
#include "trk_beb.h"

namespace trk {

#define HIP_OK(x)                                                                 \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            err = std::string(#x) + ": " + hipGetErrorString(e_);                 \
            return TRK_ERR_DEVICE;                                                \
        }                                                                         \
    } while (0)

// ---------------------------------------------------------------- physics
// Loss_func (Cross_sections.f90:446-486) with Extend_E0_to_finite_q (346-393).
// mode 0: electron/hole in the target, 1: photon (q = 0), 2: phonon (atoms).
TRK_DEV double loss(const MatDev &m, int mode, double A, double E, double G, double dE, double dq)
{
    double E0, G1;
    if (mode == 1) {
        G1 = G;
        E0 = E;
    } else {
        double hq2 = ((kH * kH) * dq) * dq;
        if (mode == 2) {
            E0 = E + hq2 / (2.0 * m.Mtarget);
            G1 = G;
        } else {
            double Mass;
            if (m.use_k) {
                double qlim = trk_fabs(dq) * trk_sqrt(kE);
                if (qlim <= m.k[m.nk - 1]) {
                    int j = find_mono_1d(m.k, m.nk, qlim);
                    Mass = m.Effm[j - 1];
                } else {
                    Mass = 1.0;
                }
            } else if (m.El_eff_mass > 0.0) {
                Mass = m.El_eff_mass;
            } else {
                Mass = 1.0;
            }
            double sqq = hq2 / ((2.0 * Mass) * kMe);
            if (m.kind_of_DR == 2) {
                E0 = trk_sqrt(((E * E) + (((m.v_f * m.v_f) * hq2) * 0.3333333333333)) + (sqq * sqq));
                G1 = G;
            } else if (m.kind_of_DR == 3) {
                E0 = trk_pow(trk_pow(E, 0.666666666666) + trk_pow(sqq, 0.666666666666), 1.5);
                G1 = trk_sqrt((G * G) + (sqq * sqq));
            } else {
                E0 = E + sqq;
                G1 = G;
            }
        }
    }
    double dE2 = dE * dE;
    double E02 = E0 * E0;
    return ((A * G1) * dE) / (((dE2 - E02) * (dE2 - E02)) + ((G1 * G1) * dE2));
}

// Imewq, Cross_sections.f90:397-442
TRK_DEV double imewq(const MatDev &m, double hw, double hq, int shell, int photon)
{
    double ImE = 0.0;
    const int o1 = m.osc_off[shell + 1];
    for (int i = m.osc_off[shell]; i < o1; ++i)
        ImE = ImE + loss(m, photon ? 1 : 0, m.oA[i], m.oE0[i], m.oG[i], hw, hq);
    return ImE;
}

// The q quadratures advance by 1% per step, a few thousand steps over any
// physical range; the step limit only ends a loop whose start is 0 or NaN.
constexpr int kMaxQSteps = 1 << 20;
// Nodes of one transfer-energy row (the reference's rows hold up to ~1400).
constexpr int kMaxRowSteps = 1 << 20;

// Diff_cross_section, Cross_sections.f90:2217-2282
TRK_DEV double diff_cs(const MatDev &m, int shell, double Ee, double dE, double Mass)
{
    double qmin, qmax;
    double pre = trk_sqrt((2.0 * Mass) * kMe) / kH;
    if (dE > Ee) {
        qmin = pre * trk_sqrt(Ee);
        qmax = pre * trk_sqrt(Ee);
    } else {
        qmin = pre * (trk_sqrt(Ee) - trk_sqrt(Ee - dE));
        qmax = pre * (trk_sqrt(Ee) + trk_sqrt(Ee - dE));
    }
    double dLs = 0.0, hq = qmin, dLs0 = 0.0;
    for (int g = 0; hq < qmax && g < kMaxQSteps; ++g) {
        double dq = hq / 100.0;
        double a = hq + dq / 2.0;
        double temp1 = imewq(m, dE, a, shell, 0);
        double b = hq + dq;
        double dL = imewq(m, dE, b, shell, 0);
        dLs = dLs + ((dq / 6.0) * ((dLs0 + 4.0 * temp1) + dL)) / hq;
        dLs0 = dL;
        hq = hq + dq;
    }
    double T_fact = (m.temp > 0.0) ? 1.0 / (1.0 - trk_exp(-((dE / m.temp) * kKb))) : 1.0;
    return ((1.0 / ((kPi * kA0) * Ee)) * dLs) * T_fact;
}

// Diff_cross_section_phonon, Cross_sections.f90:3142-3300 (screening 1)
TRK_DEV double diff_cs_phonon(const MatDev &m, double Ee, double dE, double Mass)
{
    const double eps = 1.0e-12;
    double qmin, qmax;
    double pre = trk_sqrt((2.0 * Mass) * kMe) / kH;
    if (trk_fabs(dE) < eps) return 1.31e30;
    else if (dE > (Ee - eps)) qmin = pre * trk_sqrt(Ee);
    else qmin = pre * (trk_sqrt(Ee) - trk_sqrt(trk_fabs(Ee - dE)));
    qmax = ((1.0 * trk_sqrt((2.0 * Mass) * kMe)) / kH) * (trk_sqrt(Ee) + trk_sqrt(trk_fabs(Ee - dE)));
    double dLs = 0.0, hq = qmin, dLs0 = 0.0;
    for (int g = 0; trk_fabs(hq) < trk_fabs(qmax) && g < kMaxQSteps; ++g) {
        double dq = hq / 100.0;
        double a = hq + dq / 2.0;
        double temp1 = 0.0, dL = 0.0;
        for (int i = 0; i < m.n_ph; ++i) temp1 = temp1 + loss(m, 2, m.phA[i], m.phE0[i], m.phG[i], dE, a);
        double b = hq + dq;
        for (int i = 0; i < m.n_ph; ++i) dL = dL + loss(m, 2, m.phA[i], m.phE0[i], m.phG[i], dE, b);
        double Pot = 1.0 / hq;
        dLs = dLs + ((dq / 6.0) * ((dLs0 + 4.0 * temp1) + dL)) * Pot;
        dLs0 = dL;
        hq = hq + dq;
    }
    if (m.temp > 1.0e-6)
        return ((1.0 / ((kPi * kA0) * Ee)) * dLs) / (1.0 - trk_exp(-((dE / m.temp) * kKb)));
    return (1.0 / ((kPi * kA0) * Ee)) * dLs;
}

// SHI_Diff_cross_section, Cross_sections.f90:2683-2744
TRK_DEV double shi_diff_cs(const MatDev &m, int shell, double Ee, double MSHI, double Emax, double hw)
{
    double qmin, qmax;
    if (Ee > 0.0) qmin = (hw / kH) / trk_sqrt((2.0 * Ee) / MSHI);
    else qmin = 0.0;
    if (Emax > 0.0) qmax = (trk_sqrt(2.0 * kMe) / kH) * trk_sqrt(Emax);
    else return 0.0;
    double dLs = 0.0, hq = qmin, dLs0 = 0.0;
    for (int g = 0; hq < qmax && g < kMaxQSteps; ++g) {
        double dq = hq / 100.0;
        double a = hq + dq / 2.0;
        double temp1 = imewq(m, hw, a, shell, 0);
        double b = hq + dq;
        double dL = imewq(m, hw, b, shell, 0);
        dLs = dLs + ((dq / 6.0) * ((dLs0 + 4.0 * temp1) + dL)) / hq;
        dLs0 = dL;
        hq = hq + dq;
    }
    double T_fact = (m.temp > 0.0) ? 1.0 / (1.0 - trk_exp(-((hw / m.temp) * kKb))) : 1.0;
    return dLs * T_fact;
}

// Brand_Kitagawa, Cross_sections.f90:2878-2889: the effective charge of an
// ion of nuclear charge Zp and equilibrium charge Zeff seen at momentum
// transfer hq (reference units: [sqrt(eV/J)/m], times a0 sqrt(e) in metres)
TRK_DEV double bk_charge(double hq, double Zp, double Zeff)
{
    const double a = 0.2400519147;
    const double Z = (Zp - Zeff) / Zp;
    const double Z23 = (Z == 0.0) ? 0.0 : trk_pow(Z, 2.0 / 3.0);   // 0**(2/3) = 0
    const double kl = ((((hq * ((kA0 * 1.0e-10) * trk_sqrt(kE))) * 2.0) * a) * Z23) /
                      (trk_pow(Zp, 2.0 / 3.0) * (1.0 - Z / 7.0));
    const double kl2 = kl * kl;
    return (Zp * ((1.0 - Z) + kl2)) / (1.0 + kl2);
}

// SHI_Diff_cross_section_BK, Cross_sections.f90:2826-2875: the point-charge
// integrand weighted by the squared Brandt-Kitagawa charge at each q
TRK_DEV double shi_diff_cs_bk(const MatDev &m, int shell, double Ee, double MSHI, double Emax, double hw, double Zp,
                              double Zeff)
{
    const double qmin = (hw / kH) / trk_sqrt((2.0 * Ee) / MSHI);
    const double qmax = (trk_sqrt(2.0 * kMe) / kH) * trk_sqrt(Emax);
    double dLs = 0.0, hq = qmin, dLs0 = 0.0;
    for (int g = 0; hq < qmax && g < kMaxQSteps; ++g) {
        double dq = hq / 100.0;
        double a = hq + dq / 2.0;
        double rho = bk_charge(a, Zp, Zeff);
        double temp1 = (imewq(m, hw, a, shell, 0) * rho) * rho;
        double b = hq + dq;
        rho = bk_charge(b, Zp, Zeff);
        double dL = (imewq(m, hw, b, shell, 0) * rho) * rho;
        dLs = dLs + ((dq / 6.0) * ((dLs0 + 4.0 * temp1) + dL)) / hq;
        dLs0 = dL;
        hq = hq + dq;
    }
    // dLs / (1 - exp(-hw/T kB)); at T = 0 the exponential is 0
    if (m.temp > 0.0) return dLs / (1.0 - trk_exp(-((hw / m.temp) * kKb)));
    return dLs;
}

// ---------------------------------------------------------------- row jobs
enum RowType { ROW_INEL = 0, ROW_PHONON = 1, ROW_SHI = 2, ROW_SHI_DIFF = 3 };

struct RowJob {
    double Ele, Estart, Emax, E_low, E_high, dE_min, Mass, Zeff, MSHI;
    double Zp;   // the ion's nuclear charge (ROW_SHI with bk)
    int n, shell, type, with_row;
    int kind, ie;
    int q32;     // FP32 compensated quadrature (ROW_INEL, ROW_SHI, ROW_SHI_DIFF)
    int bk;      // ROW_SHI of a Brandt-Kitagawa ion (SHI_TotIMFP_BK, CS:2748-2823)
};

// ---------------------------------------------------------------- FP32 quadrature
// The same cross sections with the quadrature of the precision form: kQ32W
// transfer-energy intervals, geometric from Estart to Emax, each with its
// midpoint (4001 points), and at each point kQ32Q Simpson intervals in ln q
// (2001 points), evaluated and summed in FP32 with compensated (Kahan) sums.
// Momenta are carried as q/q1 with q1 = sqrt(2 m_e * 1 eV)/hbar, so that
// hbar^2 q^2 / (2 m* m_e) is (q/q1)^2 / m* eV and every FP32 value stays in
// range; int dq/q is int d(ln q), unchanged by the scale.
constexpr int kQ32W = 2000;
constexpr int kQ32Q = 1000;

struct Kahan32 {
    float s = 0.0f, c = 0.0f;
    TRK_DEV void add(float x)
    {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    TRK_DEV double value() const { return (double)s - (double)c; }
};

// Imewq with Loss_func and the dispersion of Extend_E0_to_finite_q (CS:346-486)
// for a constant effective mass, in FP32; q2 = (q/q1)^2.
TRK_DEV float imewq32(const MatDev &m, int shell, float w, float q2, float inv_mstar, float c_vf)
{
    const float sqq = q2 * inv_mstar;
    const float w2 = w * w;
    float ImE = 0.0f;
    const int o1 = m.osc_off[shell + 1];
    for (int i = m.osc_off[shell]; i < o1; ++i) {
        const float E = (float)m.oE0[i], G = (float)m.oG[i], A = (float)m.oA[i];
        float E0, G1;
        if (m.kind_of_DR == 2) {
            E0 = sqrtf(E * E + c_vf * q2 + sqq * sqq);
            G1 = G;
        } else if (m.kind_of_DR == 3) {
            E0 = powf(powf(E, 0.666666666666f) + powf(sqq, 0.666666666666f), 1.5f);
            G1 = sqrtf(G * G + sqq * sqq);
        } else {
            E0 = E + sqq;
            G1 = G;
        }
        const float d = w2 - E0 * E0;
        ImE = ImE + (A * G1 * w) / (d * d + G1 * G1 * w2);
    }
    return ImE;
}

// One point of a row: the q integral of Diff_cross_section (CS:2217-2282) or
// SHI_Diff_cross_section (CS:2683-2744) at transfer energy x, in FP32.
TRK_DEV double eval32(const MatDev &m, const RowJob &j, double x)
{
    // limits in units of q1 (FP64), then the integral in FP32
    double qlo, qhi, pref;
    if (j.type == ROW_INEL) {
        if (!(x < j.Ele)) return 0.0;
        const double sm = trk_sqrt(j.Mass), se = trk_sqrt(j.Ele), sr = trk_sqrt(j.Ele - x);
        qlo = sm * (se - sr);
        qhi = sm * (se + sr);
        pref = 1.0 / ((kPi * kA0) * j.Ele);
    } else {
        if (!(j.Ele > 0.0) || !(j.Emax > 0.0)) return 0.0;
        const double q1 = trk_sqrt(2.0 * kMe) / kH;
        qlo = ((x / kH) / trk_sqrt((2.0 * j.Ele) / j.MSHI)) / q1;
        qhi = trk_sqrt(j.Emax);
        pref = 1.0;
    }
    if (m.temp > 0.0) pref = pref / (1.0 - trk_exp(-((x / m.temp) * kKb)));
    if (!(qlo > 0.0) || !(qhi > qlo)) return 0.0;
    const float inv_mstar = (float)(1.0 / m.El_eff_mass);
    const float c_vf = (float)(((m.v_f * m.v_f) * (2.0 * kMe)) / 3.0);
    const float u0 = (float)trk_log(qlo), h = (float)((trk_log(qhi) - trk_log(qlo)) / (double)kQ32Q);
    const float w = (float)x;
    Kahan32 S;
    for (int k = 0; k <= kQ32Q; ++k) {
        const float q = __expf(u0 + h * (float)k);
        const float f = imewq32(m, j.shell, w, q * q, inv_mstar, c_vf);
        const float wk = (k == 0 || k == kQ32Q) ? 1.0f : ((k & 1) ? 4.0f : 2.0f);
        S.add(wk * f);
    }
    return (double)((float)pref * ((h / 3.0f) * (float)S.value()));
}

__global__ void k_row_plan(const RowJob *__restrict__ jobs, int nrows, int *__restrict__ cnt, int *__restrict__ len)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    RowJob j = jobs[r];
    int c = 0;
    double E = j.Estart;
    if (j.q32) {
        // the fixed geometric intervals of the FP32 quadrature (none when
        // the range is empty); rows hold at least 10 entries, as the
        // reference's allocation does
        c = (j.Emax > j.Estart && j.Estart > 0.0) ? kQ32W : 0;
        cnt[r] = c;
        len[r] = j.with_row ? ((j.type == ROW_SHI_DIFF) ? c : (c > 10 ? c : 10)) : 0;
        return;
    }
    if (j.type == ROW_PHONON) {
        while (trk_fabs(E) <= trk_fabs(j.Emax) && c < kMaxRowSteps) {
            c++;
            E = E + 1.0 * define_dE(j.n, E, j.E_low, j.E_high, j.dE_min);
        }
    } else {
        while (E <= j.Emax && c < kMaxRowSteps) {
            c++;
            E = E + define_dE(j.n, E, j.E_low, j.E_high, j.dE_min);
        }
    }
    if (c >= kMaxRowSteps) { cnt[r] = -1; len[r] = 0; return; }
    int L = 0;
    if (j.with_row) {
        if (j.type == ROW_SHI_DIFF) {
            L = c;
        } else {
            // allocation count of get_diff_CS_grid_size (CS:1278-1343): from E_low
            int nref = 0;
            double Er = j.E_low;
            while (Er <= j.Emax && nref < kMaxRowSteps) {
                nref++;
                Er = Er + define_dE(j.n, Er, j.E_low, j.E_high, j.dE_min);
            }
            L = c;
            if (nref > L) L = nref;
            if (L < 10) L = 10;
        }
    }
    cnt[r] = c;
    len[r] = L;
}

__global__ void k_row_nodes(const RowJob *__restrict__ jobs, int nrows, const int64_t *__restrict__ step_off,
                            double *__restrict__ Ek, double *__restrict__ dEk, int *__restrict__ step_row)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    RowJob j = jobs[r];
    int64_t o = step_off[r];
    int64_t o1 = step_off[r + 1];
    double E = j.Estart;
    if (j.q32) {
        const double lr = trk_log(j.Emax / j.Estart) / (double)kQ32W;
        for (int64_t k = o; k < o1; ++k) {
            const int64_t i = k - o;
            const double a = j.Estart * trk_exp(lr * (double)i);
            const double b = (i + 1 == o1 - o) ? j.Emax : j.Estart * trk_exp(lr * (double)(i + 1));
            Ek[k] = a;
            dEk[k] = b - a;
            step_row[k] = r;
        }
        return;
    }
    for (int64_t k = o; k < o1; ++k) {
        double dE = define_dE(j.n, E, j.E_low, j.E_high, j.dE_min);
        if (j.type == ROW_PHONON) dE = 1.0 * dE;
        Ek[k] = E;
        dEk[k] = dE;
        step_row[k] = r;
        E = E + dE;
    }
}

TRK_DEV double eval_point(const MatDev &m, const RowJob &j, double x)
{
    if (j.q32) return eval32(m, j, x);
    switch (j.type) {
    case ROW_INEL: return diff_cs(m, j.shell, j.Ele, x, j.Mass);
    case ROW_PHONON: return diff_cs_phonon(m, j.Ele, x, j.Mass);
    default:
        if (j.bk) return shi_diff_cs_bk(m, j.shell, j.Ele, j.MSHI, j.Emax, x, j.Zp, j.Zeff);
        return shi_diff_cs(m, j.shell, j.Ele, j.MSHI, j.Emax, x);
    }
}

// Points [0, nrows) are the row starts; points nrows + 2k (+1) are the
// midpoint and right node of step k.
__global__ void k_eval(MatDev m, const RowJob *__restrict__ jobs, int nrows, int64_t nsteps,
                       const double *__restrict__ Ek, const double *__restrict__ dEk,
                       const int *__restrict__ step_row, double *__restrict__ f0,
                       double *__restrict__ fa, double *__restrict__ fb)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = (int64_t)nrows + 2 * nsteps;
    if (p >= total) return;
    if (p < nrows) {
        RowJob j = jobs[p];
        f0[p] = eval_point(m, j, j.Estart);
        return;
    }
    int64_t q = p - nrows;
    int64_t k = q >> 1;
    RowJob j = jobs[step_row[k]];
    double E = Ek[k], dE = dEk[k];
    if ((q & 1) == 0) fa[k] = eval_point(m, j, E + dE / 2.0);
    else fb[k] = eval_point(m, j, E + dE);
}

// inv[k] = 1/L[k] for one row of the ion's transfer table; *nonmono is set
// if the inverse ever decreases (or is NaN) along the row.
__global__ void k_inv_row(const double *__restrict__ L, int n, double *__restrict__ inv, int *nonmono)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double a = 1.0 / L[k];
    inv[k] = a;
    if (k + 1 < n) {
        double b = 1.0 / L[k + 1];
        if (!(a <= b)) atomicOr(nonmono, 1);
    }
}

__global__ void k_log(const double *__restrict__ x, int64_t n, double *__restrict__ out)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = trk_log(x[k]);
}

// mono[slot] = 1 if lam of the row never increases (and holds no NaN): the
// row search can then count entries instead of bisecting
__global__ void k_row_mono(const int64_t *__restrict__ off, int64_t nslots, const double *__restrict__ lam,
                           unsigned char *__restrict__ mono)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nslots) return;
    const int64_t a = off[k], b = off[k + 1];
    unsigned char m = 1;
    for (int64_t i = a; i < b; ++i) {
        if (!(lam[i] == lam[i])) m = 0;
        if (i + 1 < b && !(lam[i] >= lam[i + 1])) m = 0;
    }
    mono[k] = m;
}

struct RowOut {
    double *L, *dEdx;        // table arrays for the row's kind (shell-major)
    int n;                   // table grid size
    double *hw, *lam;        // concatenated rows of the kind
    const int64_t *row_off;  // [nsh*n+1]
};

__global__ void k_row_scan(const RowJob *__restrict__ jobs, int nrows, const int64_t *__restrict__ step_off,
                           const double *__restrict__ Ek, const double *__restrict__ dEk,
                           const double *__restrict__ f0, const double *__restrict__ fa,
                           const double *__restrict__ fb, const RowOut *__restrict__ outs,
                           double *__restrict__ dE_rows, double *__restrict__ dL_rows,
                           const int64_t *__restrict__ drow_off)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    RowJob j = jobs[r];
    int64_t o = step_off[r], o1 = step_off[r + 1];
    double Ltot1 = 0.0, Ltot0 = f0[r], ddEdx = 0.0;
    const RowOut &out = outs[j.kind];
    double *hw = nullptr, *cum = nullptr;
    int64_t rlen = 0;
    if (j.with_row && j.type != ROW_SHI_DIFF) {
        int64_t ro = out.row_off[(int64_t)j.shell * out.n + j.ie];
        rlen = out.row_off[(int64_t)j.shell * out.n + j.ie + 1] - ro;
        hw = out.hw + ro;
        cum = out.lam + ro;
    }
    double *dEr = nullptr, *dLr = nullptr;
    if (j.type == ROW_SHI_DIFF) {
        dEr = dE_rows + drow_off[j.shell];
        dLr = dL_rows + drow_off[j.shell];
    }
    int64_t i = 0;
    if (j.q32) {
        // FP32 Simpson over the transfer-energy intervals with compensated
        // sums; the cumulative values leave in FP64 for the normalisation
        Kahan32 S, D;
        float L0 = (float)Ltot0;
        for (int64_t k = o; k < o1; ++k, ++i) {
            const double E = Ek[k], dE = dEk[k];
            const float La = (float)fa[k], Lb = (float)fb[k];
            const float t2 = (float)(dE / 6.0) * ((L0 + 4.0f * La) + Lb);
            S.add(t2);
            // stopping power: Simpson on w * dsigma/dw over the interval
            const float a = (float)(E + dE / 2.0), b = (float)(E + dE);
            D.add((float)(dE / 6.0) * (((float)E * L0 + (a * 4.0f) * La) + b * Lb));
            L0 = Lb;
            if (hw) {
                hw[i] = E + dE;
                cum[i] = S.value();
            }
            if (dEr) {
                dEr[i] = E;
                dLr[i] = (((kPi * kA0) * j.Ele) * kMe) / (((j.MSHI * j.Zeff) * j.Zeff) * S.value());
            }
        }
        Ltot1 = S.value();
        ddEdx = D.value();
        o1 = o;      // the FP64 loop below has nothing left to do
    }
    for (int64_t k = o; k < o1; ++k, ++i) {
        double E = Ek[k], dE = dEk[k];
        double temp1 = fa[k], dL = fb[k];
        double temp2;
        if (j.type == ROW_PHONON) temp2 = (trk_fabs(dE) / 6.0) * ((Ltot0 + 4.0 * temp1) + dL);
        else temp2 = (dE / 6.0) * ((Ltot0 + 4.0 * temp1) + dL);
        Ltot1 = Ltot1 + temp2;
        if ((j.type == ROW_SHI && !j.bk) || j.type == ROW_SHI_DIFF) {
            double a = E + dE / 2.0;
            double b = E + dE;
            ddEdx = ddEdx + (dE / 6.0) * (((E * Ltot0) + ((a * 4.0) * temp1)) + (b * dL));
        } else {
            ddEdx = ddEdx + E * temp2;
        }
        Ltot0 = dL;
        if (hw) {
            hw[i] = E + dE;
            cum[i] = Ltot1;
        }
        if (dEr) {
            dEr[i] = E;
            dLr[i] = (((kPi * kA0) * j.Ele) * kMe) / (((j.MSHI * j.Zeff) * j.Zeff) * Ltot1);
        }
    }
    if (hw) {
        for (int64_t t = i; t < rlen; ++t) { hw[t] = 0.0; cum[t] = 0.0; }
    }
    int64_t tix = (int64_t)j.shell * out.n + j.ie;
    if (j.type == ROW_INEL) {
        double Sigma, D;
        if (Ltot1 < 1.0e-15) { Sigma = 1.0e20; D = 0.0; }
        else { Sigma = 1.0 / (j.Mass * Ltot1); D = j.Mass * ddEdx; }
        out.L[tix] = Sigma;
        out.dEdx[tix] = D;
        if (hw)
            for (int64_t t = 0; t < rlen; ++t) cum[t] = (cum[t] < 1.0e-15) ? 1.0e15 : 1.0 / (j.Mass * cum[t]);
    } else if (j.type == ROW_PHONON) {
        double Z2M = (j.Zeff * j.Zeff) * j.Mass;
        double Sigma;
        if ((j.Mass < 1.0e-20) || (Ltot1 < 1.0e-15)) Sigma = 1.0e29;
        else Sigma = 1.0 / (Z2M * Ltot1);
        if (Sigma > 1.0e30) Sigma = 1.0e30;
        out.L[j.ie] = Sigma;
        out.dEdx[j.ie] = Z2M * ddEdx;
        if (hw)
            for (int64_t t = 0; t < rlen; ++t) cum[t] = (cum[t] < 1.0e-15) ? 1.0e20 : 1.0 / (Z2M * cum[t]);
    } else if (j.type == ROW_SHI) {
        double pre = (1.0 / ((kPi * kA0) * j.Ele)) * j.MSHI;
        // a Brandt-Kitagawa ion carries its charge inside the integrand
        double Sigma = j.bk ? (pre / kMe) * Ltot1 : (((pre / kMe) * j.Zeff) * j.Zeff) * Ltot1;
        if (Sigma > 1.0e30) Sigma = 1.0e30;
        out.L[tix] = (Sigma > 1.0e-10) ? 1.0 / Sigma : 1.0e28;
        out.dEdx[tix] = j.bk ? (pre / kMe) * ddEdx : (((pre / kMe) * j.Zeff) * j.Zeff) * ddEdx;
    }
}

// Mott elastic MFP (Elastic_cross_section Mott branch, CS:2939-2962 with
// Atomic_elastic_sigma CS:3495-3514); kind_of_EMFP = -1 gives 1e30 (AI:1125).
__global__ void k_mott(MatDev m, const double *__restrict__ E, int n, int hole, double *__restrict__ L,
                       double *__restrict__ dEdx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double Ee = E[i];
    dEdx[i] = 0.0;
    if (m.kind_of_EMFP != 0) { L[i] = 1.0e30; return; }
    double Mass = 1.0;
    if (hole) {
        if (m.hole_mass >= 0.0) Mass = m.hole_mass;
        else Mass = m.dosM[find_mono_1d(m.dosE, m.n_dos, Ee) - 1];
    }
    double S = 0.0;
    for (int a = 0; a < m.n_atoms; ++a) {
        double Zat = (double)m.Zat[a];
        double mec2e = ((kMe * kC) * kC) / kE;
        double Zat137 = Zat / 137.0;
        double RyEe = kRy / Ee;
        double beta2 = (2.0 * Ee) / mec2e;
        double pc = ((1.7e-5 * trk_pow(Zat, 2.0 / 3.0)) * (1.0 - beta2)) / beta2;
        double nc = pc * (1.13 + (((3.76 * (Zat137 * Zat137)) / beta2) * trk_sqrt(Ee / (Ee + mec2e))));
        double sig = (((((((kPi * kA0) * kA0) * Zat) * (Zat + 1.0)) / (nc * (nc + 1.0))) * RyEe) * RyEe) * 1.0e-16;
        S = S + sig * m.Pers[a];
    }
    S = (S / m.sumPers) * Mass;
    L[i] = 1.0e8 / (S * m.At_Dens);
}

// Tot_Phot_IMFP, CS:833-878 (CDF branch)
__global__ void k_phot(MatDev m, const double *__restrict__ E, int n, double *__restrict__ L, double *__restrict__ dEdx)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * m.nsh) return;
    int s = t / n, i = t % n;
    double Ele = E[i];
    double Sigma;
    if (Ele < m.Ip[s]) {
        Sigma = 1.0e30;
    } else {
        double ImE = imewq(m, Ele, 0.0, s, 1);
        Sigma = ((kC * kH) / ((ImE * Ele) * kE)) * 1.0e10;
    }
    L[t] = Sigma;
    dEdx[t] = Ele / Sigma;
}

// How_many_electrons, MC:1990-2055: total MFP = 1/sum(1/lambda_shell)
// TotIMFP's Delta-CDF branch, CS:952-956: per (energy, shell) of an electron
// or hole inelastic table, the mean free path and stopping power of the
// delta peaks (with the electron's mass for holes too, as the reference),
// over the quadrature's values
__global__ void k_delta_tables(MatDev m, const double *__restrict__ alpha, const double *__restrict__ E, int n,
                               int nsh, double *__restrict__ L, double *__restrict__ dEdx)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * nsh) return;
    const int s = (int)(t / n), i = (int)(t % n);
    double Emin = m.Ip[s];
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    const int o0 = m.osc_off[s], o1 = m.osc_off[s + 1];
    const double Sigma = delta_cs(m.oE0, alpha, o0, o1, E[i], Emin, m.At_Dens, 0.0);
    L[(size_t)s * n + i] = delta_mfp(Sigma, m.At_Dens);
    dEdx[(size_t)s * n + i] = delta_eloss(m.oE0, alpha, o0, o1, E[i], Emin, m.At_Dens);
}

__global__ void k_fuse(const double *__restrict__ L, int n, int nsh, int skip_small, double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < nsh; ++k) {
        double l = L[(int64_t)k * n + i];
        if (!skip_small || l > 1.0e-10) s = s + 1.0 / l;
    }
    out[i] = (s < 1.0e-10) ? 1.0e30 : 1.0 / s;
}

// ---------------------------------------------------------------- host
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

static void osc_window(const HostMat &h, int s, double &lo, double &hi)
{
    for (int i = h.osc_off[s]; i < h.osc_off[s + 1]; ++i) {
        double l = h.m.osc_E0[i] - 5.0 * h.m.osc_Gamma[i];
        double u = h.m.osc_E0[i] + 5.0 * h.m.osc_Gamma[i];
        if (i == h.osc_off[s] || l < lo) lo = l;
        if (i == h.osc_off[s] || u > hi) hi = u;
    }
}

static double hole_mass_at(const HostMat &h, double Ele)
{
    if (h.m.hole_mass >= 0.0) return h.m.hole_mass;
    int j = find_mono_1d(h.dosE.data(), h.m.n_dos, Ele);
    return h.dosM[j - 1];
}

// get_E_low, CS:1251-1274
static void get_E_low(double lo_b, double hi_b, double Emin, double Emax, double &lo, double &hi)
{
    if ((Emin > hi_b) || (Emax < lo_b)) {
        lo = Emin;
        hi = Emax;
    } else {
        lo = dmax(lo_b, Emin);
        hi = dmin(hi_b, Emax);
        if (trk_fabs(lo - hi) < 1.0e-3) {
            lo = Emin;
            hi = Emax;
        } else {
            double dE = hi_b - lo_b;
            lo = dmin(lo, hi - dE);
            lo = dmax(lo, Emin);
        }
    }
}

// TotIMFP limits, CS:903-948
static RowJob inel_job(const HostMat &h, double Ele, int s, int hole, int with_row, int kind, int ie)
{
    RowJob j{};
    double Emin = h.m.Ip[s];
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    double Emax, Mass;
    if (hole) {
        Mass = hole_mass_at(h, Ele);
        Emax = ((4.0 * Ele) * Mass) / ((Mass + 1.0) * (Mass + 1.0));
    } else {
        Emax = (Ele + Emin) / 2.0;
        Mass = 1.0;
    }
    if (h.o.plasmon_Emax && Emin == h.Egap) {
        double Ep = trk_sqrt((((((h.m.N_VB_el * h.m.At_Dens) * 1.0e6) * kH) * kH) / (kMe * kE0)) + (h.Egap * h.Egap));
        if (Ep >= Emax) Emax = Ep;
        if (Ele < Emax) Emax = Ele;
    }
    double lo = 0, hi = 0;
    osc_window(h, s, lo, hi);
    j.Ele = Ele;
    j.Estart = Emin;
    j.Emax = Emax;
    j.E_low = dmax(lo, Emin);
    j.E_high = dmin(hi, Emax);
    j.dE_min = 1.0e-3;
    j.Mass = Mass;
    j.n = 100;
    j.shell = s;
    j.type = ROW_INEL;
    j.with_row = with_row;
    j.kind = kind;
    j.ie = ie;
    return j;
}

// Tot_EMFP limits, CS:2992-3041
static RowJob phonon_job(const HostMat &h, double Ele, int hole, int kind, int ie)
{
    RowJob j{};
    double qd = trk_pow(((6.0 * kPi) * kPi) * (h.m.At_Dens * 1.0e6), 0.33333333);
    double Edebay = ((kH * h.m.Vsound) * qd) / kE;
    Edebay = Edebay * 3.0;
    double Emin = 1.0 * 0.1e-8;
    double Mass = hole ? hole_mass_at(h, Ele) : 1.0;
    double Emax = ((((4.0 * Ele) * Mass) * kMe) * h.Mtarget) / ((h.Mtarget + Mass * kMe) * (h.Mtarget + Mass * kMe));
    if (Edebay >= Emax) Emax = Edebay;
    if (Ele < Emax) Emax = Ele;
    Emax = 1.0 * Emax;
    double lo = 0, hi = 0;
    for (int i = 0; i < h.m.n_phonon; ++i) {
        double l = h.m.ph_E0[i] - 5.0 * h.m.ph_Gamma[i];
        double u = h.m.ph_E0[i] + 5.0 * h.m.ph_Gamma[i];
        if (i == 0 || l < lo) lo = l;
        if (i == 0 || u > hi) hi = u;
    }
    double El, Eh;
    get_E_low(lo, hi, Emin, Emax, El, Eh);
    j.Ele = Ele;
    j.Estart = El;
    j.Emax = Emax;
    j.E_low = El;
    j.E_high = Eh;
    j.dE_min = 1.0e-5;
    j.Mass = Mass;
    j.Zeff = zeff_target(h, Ele);
    j.n = 20;
    j.shell = 0;
    j.type = ROW_PHONON;
    j.with_row = 1;
    j.kind = kind;
    j.ie = ie;
    return j;
}

// SHI_TotIMFP limits, CS:2470-2508
static RowJob shi_job(const HostMat &h, double Ele, int s, int diff, int ie)
{
    RowJob j{};
    double MSHI = kMp * h.ion.Mass;
    double Emin = h.m.Ip[s];
    if (Emin <= 1.0e-3) Emin = 1.0e-3;
    double Emax = (((4.0 * Ele) * kMe) * MSHI) / ((MSHI + kMe) * (MSHI + kMe));
    if (h.o.plasmon_Emax && Emin == h.Egap) {
        double Ep = trk_sqrt((((((h.m.N_VB_el * h.m.At_Dens) * 1.0e6) * kH) * kH) / (kMe * kE0)) + (h.Egap * h.Egap));
        if (Ep >= Emax) Emax = Ep;
    }
    double lo = 0, hi = 0;
    osc_window(h, s, lo, hi);
    j.Ele = Ele;
    j.Estart = Emin;
    j.Emax = Emax;
    j.E_low = dmax(lo, Emin);
    j.E_high = dmin(hi, Emax);
    j.dE_min = 1.0e-3;
    j.MSHI = MSHI;
    j.Zeff = shi_zeff(h, Ele);
    j.n = 1000;
    j.shell = s;
    j.type = diff ? ROW_SHI_DIFF : ROW_SHI;
    j.with_row = diff;
    // the ion's mean free paths of a Brandt-Kitagawa ion (SHI_Total_IMFP,
    // CS:2431-2446); its transfer rows stay those of a point charge
    j.bk = (!diff && h.ion.Kind_ion == 1) ? 1 : 0;
    j.Zp = (double)h.ion.Zat;
    j.kind = diff ? TRK_TAB_SHI_DIFF : TRK_TAB_SHI;
    j.ie = ie;
    return j;
}

template <class T>
static int dalloc(T **p, size_t n, std::string &err)
{
    if (n == 0) n = 1;
    HIP_OK(hipMalloc((void **)p, n * sizeof(T)));
    return 0;
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void free_tables(TableSet &ts)
{
    if (ts.shared_material) {
        // forget the borrowed material arrays; free only the ion's own
        for (int k : {TRK_TAB_E_INEL, TRK_TAB_H_INEL, TRK_TAB_E_EL, TRK_TAB_H_EL, TRK_TAB_PHOT}) {
            ts.tab[k] = TableDev();
            ts.ln_E[k] = ts.ln_L[k] = ts.ln_hw[k] = ts.ln_lam[k] = nullptr;
        }
        ts.el_imfp = ts.hole_imfp = ts.phot_imfp = nullptr;
        ts.ln_el_imfp = ts.ln_hole_imfp = ts.ln_Lel = ts.ln_LHel = nullptr;
        ts.dsf[0] = ts.dsf[1] = DsfDev{};
        ts.shared_material = false;
    }
    for (int p = 0; p < 2; ++p) {
        if (ts.dsf_buf[p]) (void)hipFree(ts.dsf_buf[p]);
        ts.dsf_buf[p] = nullptr;
        ts.dsf[p] = DsfDev{};
    }
    for (auto &t : ts.tab) {
        if (t.E) (void)hipFree(t.E);
        if (t.L) (void)hipFree(t.L);
        if (t.dEdx) (void)hipFree(t.dEdx);
        if (t.row_off) (void)hipFree(t.row_off);
        if (t.hw) (void)hipFree(t.hw);
        if (t.lam) (void)hipFree(t.lam);
        if (t.row_mono) (void)hipFree(t.row_mono);
        if (t.dE) (void)hipFree(t.dE);
        if (t.dL) (void)hipFree(t.dL);
        if (t.dInv) (void)hipFree(t.dInv);
        if (t.dNonMono) (void)hipFree(t.dNonMono);
        t = TableDev();
    }
    if (ts.el_imfp) (void)hipFree(ts.el_imfp);
    if (ts.hole_imfp) (void)hipFree(ts.hole_imfp);
    if (ts.phot_imfp) (void)hipFree(ts.phot_imfp);
    if (ts.shi_path) (void)hipFree(ts.shi_path);
    ts.el_imfp = ts.hole_imfp = ts.phot_imfp = ts.shi_path = nullptr;
    for (double *&p : ts.ln_E) { if (p) (void)hipFree(p); p = nullptr; }
    for (int k = 0; k < 7; ++k)
        for (double **p : {&ts.ln_L[k], &ts.ln_hw[k], &ts.ln_lam[k]}) { if (*p) (void)hipFree(*p); *p = nullptr; }
    double **lp[] = {&ts.ln_el_imfp, &ts.ln_hole_imfp, &ts.ln_Lel, &ts.ln_LHel, &ts.ln_shi_path};
    for (double **p : lp) { if (*p) (void)hipFree(*p); *p = nullptr; }
    void *sp[] = {ts.sd_n, ts.sd_off, ts.sd_M, ts.sd_c, ts.sd_bk};
    for (void *p : sp) if (p) (void)hipFree(p);
    ts.sd_n = nullptr; ts.sd_off = nullptr; ts.sd_M = nullptr; ts.sd_c = nullptr; ts.sd_bk = nullptr;
    ts.sd_consts = false;
    ts.grids_increasing = false;
}

static int upload_grid(TableDev &t, const std::vector<double> &g, int nsh, std::string &err)
{
    t.n = (int)g.size();
    t.nsh = nsh;
    if (dalloc(&t.E, g.size(), err)) return TRK_ERR_DEVICE;
    HIP_OK(hipMemcpy(t.E, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice));
    if (dalloc(&t.L, (size_t)nsh * g.size(), err)) return TRK_ERR_DEVICE;
    if (dalloc(&t.dEdx, (size_t)nsh * g.size(), err)) return TRK_ERR_DEVICE;
    t.has_rows.assign(nsh, 0);
    return 0;
}

// The material kinds of a built table set, borrowed by an ion's table set.
static void borrow_material(TableSet &ts, const TableSet &mat)
{
    for (int k : {TRK_TAB_E_INEL, TRK_TAB_H_INEL, TRK_TAB_E_EL, TRK_TAB_H_EL, TRK_TAB_PHOT}) {
        ts.tab[k] = mat.tab[k];
        ts.ln_E[k] = mat.ln_E[k]; ts.ln_L[k] = mat.ln_L[k]; ts.ln_hw[k] = mat.ln_hw[k]; ts.ln_lam[k] = mat.ln_lam[k];
    }
    ts.el_imfp = mat.el_imfp; ts.hole_imfp = mat.hole_imfp; ts.phot_imfp = mat.phot_imfp;
    ts.ln_el_imfp = mat.ln_el_imfp; ts.ln_hole_imfp = mat.ln_hole_imfp; ts.ln_Lel = mat.ln_Lel; ts.ln_LHel = mat.ln_LHel;
    ts.dsf[0] = mat.dsf[0];
    ts.dsf[1] = mat.dsf[1];
    ts.shared_material = true;
}

int build_tables(const HostMat &h, const MatDev &md, TableSet &ts, hipStream_t st, std::string &err,
                 const TableSet *mat, const double *delta_alpha)
{
    auto t_start = std::chrono::steady_clock::now();
    free_tables(ts);
    const bool own = (mat == nullptr);     // build the material kinds too
    const int nsh = h.nsh;
    double dosmax = h.dosE.empty() ? 0.0 : h.dosE[0];
    for (double v : h.dosE) dosmax = (v > dosmax) ? v : dosmax;

    // grids (get_grid_4CS, AI:158, 210, 2270-2273)
    std::vector<double> gE, gH, gEel, gHel, gP, gS;
    grid_4cs(h.m, 0.1, h.e_emax, gE);
    grid_4cs(h.m, 0.1, dosmax + 1.0, gH);
    grid_4cs(h.m, 0.01, h.e_emax, gEel);
    grid_4cs(h.m, 0.01, dosmax + 1.0, gHel);
    if (h.o.kind_of_EMFP == 2) {
        // the elastic tables of DSF scattering live on the DSF's own energies
        // (Elastic_MFP%Total%E = DSF_DEMFP%E, Analytical_IMFPs.f90:905-911)
        if (h.dsfE[0].size() < 2 || h.dsfE[1].size() < 2) {
            err = "DSF elastic scattering (kind_of_EMFP = 2) needs the electron and hole DSF tables (trk_dsf_set)";
            return TRK_ERR_STATE;
        }
        gEel = h.dsfE[0];
        gHel = h.dsfE[1];
    }
    if (h.o.include_photons) grid_4cs(h.m, 0.1, h.e_emax, gP);
    {
        double M = h.ion.Mass * kMp;
        double Emin = std::ceil(((((M + kMe) * (M + kMe)) / (M * kMe)) * h.Egap) / 4.0);
        double Emax = (h.o.shi_grid_Emax > 0.0) ? h.o.shi_grid_Emax : (175.6e6 / 2.0) * h.ion.Mass;
        grid_4cs(h.m, Emin, Emax, gS);
    }
    TableDev *T = ts.tab;
    int rc;
    if (own) {
        if ((rc = upload_grid(T[TRK_TAB_E_INEL], gE, nsh, err))) return rc;
        if ((rc = upload_grid(T[TRK_TAB_H_INEL], gH, nsh, err))) return rc;
        if ((rc = upload_grid(T[TRK_TAB_E_EL], gEel, 1, err))) return rc;
        if ((rc = upload_grid(T[TRK_TAB_H_EL], gHel, 1, err))) return rc;
        if (!gP.empty() && (rc = upload_grid(T[TRK_TAB_PHOT], gP, nsh, err))) return rc;
    } else {
        borrow_material(ts, *mat);
    }
    if ((rc = upload_grid(T[TRK_TAB_SHI], gS, nsh, err))) return rc;
    T[TRK_TAB_SHI_DIFF].nsh = nsh;

    // row jobs
    std::vector<RowJob> jobs;
    if (own) {
        for (int i = 0; i < (int)gE.size(); ++i)
            for (int s = 0; s < nsh; ++s) jobs.push_back(inel_job(h, gE[i], s, 0, 1, TRK_TAB_E_INEL, i));
        for (int i = 0; i < (int)gH.size(); ++i)
            for (int s = 0; s < nsh; ++s) jobs.push_back(inel_job(h, gH[i], s, 1, s == h.vb, TRK_TAB_H_INEL, i));
        if (h.o.kind_of_EMFP == 1) {
            for (int i = 0; i < (int)gEel.size(); ++i) jobs.push_back(phonon_job(h, gEel[i], 0, TRK_TAB_E_EL, i));
            for (int i = 0; i < (int)gHel.size(); ++i) jobs.push_back(phonon_job(h, gHel[i], 1, TRK_TAB_H_EL, i));
        }
    }
    for (int i = 0; i < (int)gS.size(); ++i)
        for (int s = 0; s < nsh; ++s) jobs.push_back(shi_job(h, gS[i], s, 0, i));
    for (int s = 0; s < nsh; ++s) jobs.push_back(shi_job(h, h.ion.E, s, 1, 0));
    // the precision form's quadrature for the inelastic and ion rows (the
    // phonon rows keep the reference's)
    if (h.o.quadrature == 1)
        for (RowJob &j : jobs)
            if (j.type != ROW_PHONON && !j.bk) j.q32 = 1;
    if (own) {
        T[TRK_TAB_E_INEL].has_rows.assign(nsh, 1);
        T[TRK_TAB_H_INEL].has_rows[h.vb] = 1;
        if (h.o.kind_of_EMFP == 1) {
            T[TRK_TAB_E_EL].has_rows[0] = 1;
            T[TRK_TAB_H_EL].has_rows[0] = 1;
        }
    }
    const int nrows = (int)jobs.size();

    RowJob *d_jobs = nullptr;
    int *d_cnt = nullptr, *d_len = nullptr;
    if (dalloc(&d_jobs, nrows, err) || dalloc(&d_cnt, nrows, err) || dalloc(&d_len, nrows, err)) return TRK_ERR_DEVICE;
    HIP_OK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(RowJob) * nrows, hipMemcpyHostToDevice, st));
    k_row_plan<<<nblk(nrows, 64), 64, 0, st>>>(d_jobs, nrows, d_cnt, d_len);
    HIP_OK(hipGetLastError());
    std::vector<int> cnt(nrows), len(nrows);
    HIP_OK(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int) * nrows, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(len.data(), d_len, sizeof(int) * nrows, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int r = 0; r < nrows; ++r)
        if (cnt[r] < 0) {
            err = "transfer-energy grid of row " + std::to_string(r) + " (kind " + std::to_string(jobs[r].kind) +
                  ", shell " + std::to_string(jobs[r].shell) + ", E " + std::to_string(jobs[r].Ele) + " eV) does not close";
            return TRK_ERR_NUMERIC;
        }

    // step offsets
    std::vector<int64_t> step_off(nrows + 1, 0);
    for (int r = 0; r < nrows; ++r) step_off[r + 1] = step_off[r] + cnt[r];
    const int64_t nsteps = step_off[nrows];

    // row storage per kind: row_off over (shell, ie)
    if (own)
        for (int kind : {TRK_TAB_E_INEL, TRK_TAB_H_INEL, TRK_TAB_E_EL, TRK_TAB_H_EL}) {
            TableDev &t = T[kind];
            t.row_off_h.assign((size_t)t.nsh * t.n + 1, 0);
        }
    std::vector<int64_t> rowlen_by_slot[4];
    if (own)
        for (int kind = 0; kind < 4; ++kind) rowlen_by_slot[kind].assign((size_t)T[kind].nsh * T[kind].n, 0);
    std::vector<int> dn(nsh, 0);
    for (int r = 0; r < nrows; ++r) {
        const RowJob &j = jobs[r];
        if (j.type == ROW_SHI_DIFF) { dn[j.shell] = len[r]; continue; }
        if (!j.with_row) continue;
        rowlen_by_slot[j.kind][(size_t)j.shell * T[j.kind].n + j.ie] = len[r];
    }
    for (int kind = 0; kind < 4 && own; ++kind) {
        TableDev &t = T[kind];
        for (size_t i = 0; i < rowlen_by_slot[kind].size(); ++i) t.row_off_h[i + 1] = t.row_off_h[i] + rowlen_by_slot[kind][i];
        int64_t tot = t.row_off_h.back();
        if (dalloc(&t.row_off, t.row_off_h.size(), err)) return TRK_ERR_DEVICE;
        HIP_OK(hipMemcpyAsync(t.row_off, t.row_off_h.data(), sizeof(int64_t) * t.row_off_h.size(), hipMemcpyHostToDevice, st));
        if (dalloc(&t.hw, tot, err) || dalloc(&t.lam, tot, err)) return TRK_ERR_DEVICE;
    }
    TableDev &td = T[TRK_TAB_SHI_DIFF];
    td.dn = dn;
    td.doff.assign(nsh + 1, 0);
    for (int s = 0; s < nsh; ++s) td.doff[s + 1] = td.doff[s] + dn[s];
    int64_t *d_doff = nullptr;
    if (dalloc(&td.dE, td.doff[nsh], err) || dalloc(&td.dL, td.doff[nsh], err) || dalloc(&d_doff, nsh + 1, err))
        return TRK_ERR_DEVICE;
    HIP_OK(hipMemcpyAsync(d_doff, td.doff.data(), sizeof(int64_t) * (nsh + 1), hipMemcpyHostToDevice, st));

    // nodes and evaluations
    int64_t *d_step_off = nullptr;
    double *d_Ek = nullptr, *d_dEk = nullptr, *d_f0 = nullptr, *d_fa = nullptr, *d_fb = nullptr;
    int *d_step_row = nullptr;
    if (dalloc(&d_step_off, nrows + 1, err) || dalloc(&d_Ek, nsteps, err) || dalloc(&d_dEk, nsteps, err) ||
        dalloc(&d_f0, nrows, err) || dalloc(&d_fa, nsteps, err) || dalloc(&d_fb, nsteps, err) ||
        dalloc(&d_step_row, nsteps, err))
        return TRK_ERR_DEVICE;
    HIP_OK(hipMemcpyAsync(d_step_off, step_off.data(), sizeof(int64_t) * (nrows + 1), hipMemcpyHostToDevice, st));
    k_row_nodes<<<nblk(nrows, 64), 64, 0, st>>>(d_jobs, nrows, d_step_off, d_Ek, d_dEk, d_step_row);
    HIP_OK(hipGetLastError());
    int64_t npts = nrows + 2 * nsteps;
    k_eval<<<nblk(npts, 256), 256, 0, st>>>(md, d_jobs, nrows, nsteps, d_Ek, d_dEk, d_step_row, d_f0, d_fa, d_fb);
    HIP_OK(hipGetLastError());
    ts.dcs_evals = npts;

    RowOut outs[7];
    for (int k = 0; k < 7; ++k) {
        outs[k].L = T[k].L;
        outs[k].dEdx = T[k].dEdx;
        outs[k].n = T[k].n;
        outs[k].hw = T[k].hw;
        outs[k].lam = T[k].lam;
        outs[k].row_off = T[k].row_off;
    }
    RowOut *d_outs = nullptr;
    if (dalloc(&d_outs, 7, err)) return TRK_ERR_DEVICE;
    HIP_OK(hipMemcpyAsync(d_outs, outs, sizeof(outs), hipMemcpyHostToDevice, st));
    k_row_scan<<<nblk(nrows, 64), 64, 0, st>>>(d_jobs, nrows, d_step_off, d_Ek, d_dEk, d_f0, d_fa, d_fb, d_outs,
                                                td.dE, td.dL, d_doff);
    HIP_OK(hipGetLastError());

    // inverse rows of the ion's transfer table and their monotonicity
    if (dalloc(&td.dInv, td.doff[nsh] > 0 ? td.doff[nsh] : 1, err) || dalloc(&td.dNonMono, nsh, err))
        return TRK_ERR_DEVICE;
    HIP_OK(hipMemsetAsync(td.dNonMono, 0, sizeof(int) * nsh, st));
    for (int s = 0; s < nsh; ++s)
        if (dn[s] > 0)
            k_inv_row<<<nblk(dn[s], 256), 256, 0, st>>>(td.dL + td.doff[s], dn[s], td.dInv + td.doff[s], td.dNonMono + s);
    HIP_OK(hipGetLastError());

    // the Delta-CDF model's mean free paths and stopping powers
    if (own && h.o.kind_of_DR == 4) {
        if (!delta_alpha) { err = "the Delta-CDF model needs the oscillators' weights"; return TRK_ERR_STATE; }
        for (int k : {TRK_TAB_E_INEL, TRK_TAB_H_INEL})
            k_delta_tables<<<nblk((int64_t)T[k].n * nsh, 128), 128, 0, st>>>(md, delta_alpha, T[k].E, T[k].n, nsh,
                                                                             T[k].L, T[k].dEdx);
        HIP_OK(hipGetLastError());
    }
    // BEB shells (TotIMFP's BEB branch, CS:1035-1041): the closed forms of
    // the mean free path and stopping power, on the host (a few values per
    // shell), over the quadrature's values; the rows are not used
    std::vector<std::vector<double>> beb_vals;
    if (own) {
        for (int k : {TRK_TAB_E_INEL, TRK_TAB_H_INEL}) {
            const std::vector<double> &g = (k == TRK_TAB_E_INEL) ? gE : gH;
            for (int s = 0; s < nsh; ++s) {
                if (h.m.kocs[s] != 2) continue;
                const double B = h.m.Ip[s], U = h.m.Ek[s], N = h.m.Nel[s];
                const double temp1 = ((h.m.At_Dens * 1.0e-24) * h.m.Pers[h.atom_of[s]]) / h.sumPers;
                std::vector<double> L(g.size()), D(g.size());
                for (size_t i = 0; i < g.size(); ++i) {
                    const double Ele = g[i];
                    const double Mass = (k == TRK_TAB_H_INEL) ? hole_mass_at(h, Ele) : 1.0;
                    L[i] = 1.0 / ((Mass * temp1) * beb_sigma(Ele, B, U, N));
                    D[i] = (Mass * temp1) * beb_dsigma_w_int(Ele, (Ele - 1.0) / 2.0, B, U, N);
                }
                beb_vals.push_back(std::move(L));
                HIP_OK(hipMemcpyAsync(T[k].L + (size_t)s * T[k].n, beb_vals.back().data(), sizeof(double) * g.size(),
                                      hipMemcpyHostToDevice, st));
                beb_vals.push_back(std::move(D));
                HIP_OK(hipMemcpyAsync(T[k].dEdx + (size_t)s * T[k].n, beb_vals.back().data(), sizeof(double) * g.size(),
                                      hipMemcpyHostToDevice, st));
            }
        }
        if (!beb_vals.empty()) HIP_OK(hipStreamSynchronize(st));   // the host values are then free to go
    }
    // DSF elastic tables: the total mean free path is each row's last entry
    // (its emission and absorption parts likewise), no stopping power; the
    // rows go to the device for the transfer sampling
    if (own && h.o.kind_of_EMFP == 2) {
        for (int p = 0; p < 2; ++p) {
            const int k = p ? TRK_TAB_H_EL : TRK_TAB_E_EL;
            const int nE = (int)h.dsfE[p].size(), nW = h.dsf_nW[p];
            const size_t cells = (size_t)nE * nW;
            std::vector<double> host(3 * cells + 3 * (size_t)nE);
            double *Lt = host.data() + 3 * cells, *Le = Lt + nE, *La = Le + nE;
            for (int i = 0; i < nE; ++i) {
                const size_t last = (size_t)i * nW + (nW - 1);
                Lt[i] = h.dsfL[p][last];
                Le[i] = h.dsfLe[p][last];
                La[i] = h.dsfLa[p][last];
            }
            std::copy(h.dsfdE[p].begin(), h.dsfdE[p].end(), host.begin());
            std::copy(h.dsfLe[p].begin(), h.dsfLe[p].end(), host.begin() + cells);
            std::copy(h.dsfLa[p].begin(), h.dsfLa[p].end(), host.begin() + 2 * cells);
            if (dalloc(&ts.dsf_buf[p], host.size(), err)) return TRK_ERR_DEVICE;
            HIP_OK(hipMemcpy(ts.dsf_buf[p], host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(T[k].L, Lt, sizeof(double) * nE, hipMemcpyHostToDevice));
            HIP_OK(hipMemset(T[k].dEdx, 0, sizeof(double) * nE));
            DsfDev &dv = ts.dsf[p];
            dv.nE = nE; dv.nW = nW;
            dv.E = T[k].E; dv.Lt = T[k].L;
            dv.dE = ts.dsf_buf[p]; dv.dLe = ts.dsf_buf[p] + cells; dv.dLa = ts.dsf_buf[p] + 2 * cells;
            dv.Le = ts.dsf_buf[p] + 3 * cells + nE; dv.La = dv.Le + nE;
        }
    }
    // Mott / disabled elastic tables
    if (own && h.o.kind_of_EMFP != 1 && h.o.kind_of_EMFP != 2) {
        k_mott<<<nblk(T[TRK_TAB_E_EL].n, 128), 128, 0, st>>>(md, T[TRK_TAB_E_EL].E, T[TRK_TAB_E_EL].n, 0,
                                                              T[TRK_TAB_E_EL].L, T[TRK_TAB_E_EL].dEdx);
        k_mott<<<nblk(T[TRK_TAB_H_EL].n, 128), 128, 0, st>>>(md, T[TRK_TAB_H_EL].E, T[TRK_TAB_H_EL].n, 1,
                                                              T[TRK_TAB_H_EL].L, T[TRK_TAB_H_EL].dEdx);
        HIP_OK(hipGetLastError());
    }
    if (own && !gP.empty()) {
        k_phot<<<nblk((int64_t)T[TRK_TAB_PHOT].n * nsh, 128), 128, 0, st>>>(md, T[TRK_TAB_PHOT].E, T[TRK_TAB_PHOT].n,
                                                                             T[TRK_TAB_PHOT].L, T[TRK_TAB_PHOT].dEdx);
        HIP_OK(hipGetLastError());
    }
    // fused totals
    if (dalloc(&ts.shi_path, T[TRK_TAB_SHI].n, err)) return TRK_ERR_DEVICE;
    k_fuse<<<nblk(T[TRK_TAB_SHI].n, 128), 128, 0, st>>>(T[TRK_TAB_SHI].L, T[TRK_TAB_SHI].n, nsh, 0, ts.shi_path);
    if (own) {
        if (dalloc(&ts.el_imfp, T[TRK_TAB_E_INEL].n, err) || dalloc(&ts.hole_imfp, T[TRK_TAB_H_INEL].n, err))
            return TRK_ERR_DEVICE;
        k_fuse<<<nblk(T[TRK_TAB_E_INEL].n, 128), 128, 0, st>>>(T[TRK_TAB_E_INEL].L, T[TRK_TAB_E_INEL].n, nsh, 1, ts.el_imfp);
        k_fuse<<<nblk(T[TRK_TAB_H_INEL].n, 128), 128, 0, st>>>(T[TRK_TAB_H_INEL].L, T[TRK_TAB_H_INEL].n, nsh, 0, ts.hole_imfp);
    }
    if (own && !gP.empty()) {
        if (dalloc(&ts.phot_imfp, T[TRK_TAB_PHOT].n, err)) return TRK_ERR_DEVICE;
        k_fuse<<<nblk(T[TRK_TAB_PHOT].n, 128), 128, 0, st>>>(T[TRK_TAB_PHOT].L, T[TRK_TAB_PHOT].n, nsh, 0, ts.phot_imfp);
    }
    HIP_OK(hipGetLastError());
    // logs of the grids and of the tables the transport interpolates most
    for (int k = 0; k < 6; ++k) {
        if (!own && k != TRK_TAB_SHI) continue;
        if (T[k].n <= 0 || !T[k].E) continue;
        if (dalloc(&ts.ln_E[k], T[k].n, err)) return TRK_ERR_DEVICE;
        k_log<<<nblk(T[k].n, 128), 128, 0, st>>>(T[k].E, T[k].n, ts.ln_E[k]);
    }
    struct { const double *src; int n; double **dst; bool mat; } lg[] = {
        {ts.el_imfp, T[TRK_TAB_E_INEL].n, &ts.ln_el_imfp, true}, {ts.hole_imfp, T[TRK_TAB_H_INEL].n, &ts.ln_hole_imfp, true},
        {T[TRK_TAB_E_EL].L, T[TRK_TAB_E_EL].n, &ts.ln_Lel, true}, {T[TRK_TAB_H_EL].L, T[TRK_TAB_H_EL].n, &ts.ln_LHel, true},
        {ts.shi_path, T[TRK_TAB_SHI].n, &ts.ln_shi_path, false}};
    for (auto &g : lg) {
        if (!g.src || g.n <= 0 || (g.mat && !own)) continue;
        if (dalloc(g.dst, g.n, err)) return TRK_ERR_DEVICE;
        k_log<<<nblk(g.n, 128), 128, 0, st>>>(g.src, g.n, *g.dst);
    }
    // per-shell mean free paths (shell choice) and cumulative rows (transfer
    // sampling): interpolation then takes one log of the argument
    for (int k : {TRK_TAB_E_INEL, TRK_TAB_H_INEL, TRK_TAB_SHI}) {
        const int64_t m = (int64_t)T[k].nsh * T[k].n;
        if (!T[k].L || m <= 0 || (!own && k != TRK_TAB_SHI)) continue;
        if (dalloc(&ts.ln_L[k], m, err)) return TRK_ERR_DEVICE;
        k_log<<<(unsigned)((m + 127) / 128), 128, 0, st>>>(T[k].L, m, ts.ln_L[k]);
    }
    for (int k : {TRK_TAB_E_INEL, TRK_TAB_H_INEL, TRK_TAB_E_EL, TRK_TAB_H_EL}) {
        const int64_t m = T[k].row_off_h.empty() ? 0 : T[k].row_off_h.back();
        if (!own || !T[k].hw || m <= 0) continue;
        if (dalloc(&ts.ln_hw[k], m, err) || dalloc(&ts.ln_lam[k], m, err)) return TRK_ERR_DEVICE;
        k_log<<<(unsigned)((m + 127) / 128), 128, 0, st>>>(T[k].hw, m, ts.ln_hw[k]);
        k_log<<<(unsigned)((m + 127) / 128), 128, 0, st>>>(T[k].lam, m, ts.ln_lam[k]);
        const int64_t nslots = (int64_t)T[k].nsh * T[k].n;
        if (dalloc(&T[k].row_mono, nslots, err)) return TRK_ERR_DEVICE;
        k_row_mono<<<(unsigned)((nslots + 127) / 128), 128, 0, st>>>(T[k].row_off, nslots, T[k].lam, T[k].row_mono);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));
    // run constants: a particle's last bracket may stand in for a search only
    // on strictly increasing grids; the ion's first free path sizes batches
    {
        bool inc = own || mat->grids_increasing;
        for (const std::vector<double> *g : {&gE, &gEel, &gH, &gHel, &gS})
            for (size_t i = 1; i < g->size(); ++i)
                if (!((*g)[i - 1] < (*g)[i])) inc = false;
        ts.grids_increasing = inc;
        std::vector<double> sp(gS.size());
        HIP_OK(hipMemcpy(sp.data(), ts.shi_path, sizeof(double) * sp.size(), hipMemcpyDeviceToHost));
        ts.shi_path0 = nfp_2d_host(gS.data(), sp.data(), (int)gS.size(), h.ion.E);
        const TableDev &td = T[TRK_TAB_SHI_DIFF];
        if (dalloc(&ts.sd_n, (size_t)nsh, err) || dalloc(&ts.sd_off, (size_t)nsh + 1, err) ||
            dalloc(&ts.sd_M, (size_t)2 * nsh, err) || dalloc(&ts.sd_c, (size_t)2 * nsh, err) ||
            dalloc(&ts.sd_bk, (size_t)nsh * (kShiBuckets + 1), err))
            return TRK_ERR_DEVICE;
        HIP_OK(hipMemcpy(ts.sd_n, td.dn.data(), sizeof(int) * nsh, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(ts.sd_off, td.doff.data(), sizeof(int64_t) * (nsh + 1), hipMemcpyHostToDevice));
        ts.sd_consts = false;
    }
    (void)hipFree(d_jobs); (void)hipFree(d_cnt); (void)hipFree(d_len); (void)hipFree(d_doff);
    (void)hipFree(d_step_off); (void)hipFree(d_Ek); (void)hipFree(d_dEk); (void)hipFree(d_f0);
    (void)hipFree(d_fa); (void)hipFree(d_fb); (void)hipFree(d_step_row); (void)hipFree(d_outs);
    ts.launches = 12;
    ts.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return 0;
}

// After tables were replaced from host arrays (trk_table_set / trk_row_set):
// the fused paths, the logs, the row flags, the ion's inverse transfer rows
// and the run constants derived from them, recomputed in place.
int refresh_derived(const HostMat &h, TableSet &ts, hipStream_t st, std::string &err)
{
    const int nsh = h.nsh;
    TableDev *T = ts.tab;
    const bool own = !ts.shared_material;
    if (own) {
        k_fuse<<<nblk(T[TRK_TAB_E_INEL].n, 128), 128, 0, st>>>(T[TRK_TAB_E_INEL].L, T[TRK_TAB_E_INEL].n, nsh, 1, ts.el_imfp);
        k_fuse<<<nblk(T[TRK_TAB_H_INEL].n, 128), 128, 0, st>>>(T[TRK_TAB_H_INEL].L, T[TRK_TAB_H_INEL].n, nsh, 0, ts.hole_imfp);
        if (ts.phot_imfp)
            k_fuse<<<nblk(T[TRK_TAB_PHOT].n, 128), 128, 0, st>>>(T[TRK_TAB_PHOT].L, T[TRK_TAB_PHOT].n, nsh, 0, ts.phot_imfp);
    }
    k_fuse<<<nblk(T[TRK_TAB_SHI].n, 128), 128, 0, st>>>(T[TRK_TAB_SHI].L, T[TRK_TAB_SHI].n, nsh, 0, ts.shi_path);
    for (int k = 0; k < 6; ++k) {
        if (!own && k != TRK_TAB_SHI) continue;
        if (T[k].n > 0 && T[k].E && ts.ln_E[k]) k_log<<<nblk(T[k].n, 128), 128, 0, st>>>(T[k].E, T[k].n, ts.ln_E[k]);
    }
    struct { const double *src; int n; double *dst; bool mat; } lg[] = {
        {ts.el_imfp, T[TRK_TAB_E_INEL].n, ts.ln_el_imfp, true}, {ts.hole_imfp, T[TRK_TAB_H_INEL].n, ts.ln_hole_imfp, true},
        {T[TRK_TAB_E_EL].L, T[TRK_TAB_E_EL].n, ts.ln_Lel, true}, {T[TRK_TAB_H_EL].L, T[TRK_TAB_H_EL].n, ts.ln_LHel, true},
        {ts.shi_path, T[TRK_TAB_SHI].n, ts.ln_shi_path, false}};
    for (auto &g : lg)
        if (g.src && g.dst && g.n > 0 && (own || !g.mat)) k_log<<<nblk(g.n, 128), 128, 0, st>>>(g.src, g.n, g.dst);
    for (int k : {TRK_TAB_E_INEL, TRK_TAB_H_INEL, TRK_TAB_SHI}) {
        const int64_t m = (int64_t)T[k].nsh * T[k].n;
        if (!T[k].L || !ts.ln_L[k] || m <= 0 || (!own && k != TRK_TAB_SHI)) continue;
        k_log<<<(unsigned)((m + 127) / 128), 128, 0, st>>>(T[k].L, m, ts.ln_L[k]);
    }
    for (int k : {TRK_TAB_E_INEL, TRK_TAB_H_INEL, TRK_TAB_E_EL, TRK_TAB_H_EL}) {
        const int64_t m = T[k].row_off_h.empty() ? 0 : T[k].row_off_h.back();
        if (!own || !T[k].hw || !ts.ln_hw[k] || m <= 0) continue;
        k_log<<<(unsigned)((m + 127) / 128), 128, 0, st>>>(T[k].hw, m, ts.ln_hw[k]);
        k_log<<<(unsigned)((m + 127) / 128), 128, 0, st>>>(T[k].lam, m, ts.ln_lam[k]);
        const int64_t nslots = (int64_t)T[k].nsh * T[k].n;
        k_row_mono<<<(unsigned)((nslots + 127) / 128), 128, 0, st>>>(T[k].row_off, nslots, T[k].lam, T[k].row_mono);
    }
    TableDev &td = T[TRK_TAB_SHI_DIFF];
    HIP_OK(hipMemsetAsync(td.dNonMono, 0, sizeof(int) * nsh, st));
    for (int sh = 0; sh < nsh; ++sh)
        if (td.dn[sh] > 0)
            k_inv_row<<<nblk(td.dn[sh], 256), 256, 0, st>>>(td.dL + td.doff[sh], td.dn[sh], td.dInv + td.doff[sh],
                                                            td.dNonMono + sh);
    HIP_OK(hipGetLastError());
    std::vector<double> sE(T[TRK_TAB_SHI].n), sp(T[TRK_TAB_SHI].n);
    HIP_OK(hipMemcpyAsync(sE.data(), T[TRK_TAB_SHI].E, sizeof(double) * sE.size(), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(sp.data(), ts.shi_path, sizeof(double) * sp.size(), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    ts.shi_path0 = nfp_2d_host(sE.data(), sp.data(), (int)sE.size(), h.ion.E);
    ts.sd_consts = false;
    return 0;
}

}  // namespace trk
