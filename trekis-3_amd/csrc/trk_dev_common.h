// trk_dev_common.h -- device-side definitions shared by the table and
// transport kernels of the MI355X TREKIS-3 path.
//
// Physical constants restate Universal_Constants.f90:28-60 with the same
// defining expressions (so each folds to the same double as the reference's
// parameter statements).  The search helpers reproduce the exact bisection
// paths of Reading_files_and_parameters.f90:3348-3559: the 1D and 2D variants
// differ in their exit test, and a table value equal to a grid point can land
// in a different bracket depending on the path, so the paths are kept.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "trk_detmath.h"
#include "trk_philox.h"
#include "trekis_gpu.h"

#define TRK_DEV __device__ __forceinline__
#define TRK_HD __host__ __device__ __forceinline__

namespace trk {

// Universal_Constants.f90:28-60
constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kE = 1.602176487e-19;
constexpr double kMe = 9.1093821545e-31;
constexpr double kC = 299792458.0;
constexpr double kMp = 1836.1526724780 * kMe;
constexpr double kH = 1.05457162853e-34;
constexpr double kKb = 11604.0;
constexpr double kE0 = 8.854187817620e-12;
constexpr double kRy = 13.6056981;
constexpr double kA0 = 0.5291772085936;
constexpr double kAlpha = kE * kE / (kH * kC * 4.0 * kPi * kE0);

// ---------------------------------------------------------------- searches
// The three searches return a 1-based index in [1, N] for every input: a NaN
// value gives 1 (the reference's searches report an error there; callers here
// carry the NaN on into the particle state, where the kernels detect it and
// the run fails with TRK_ERR_NAN), so no lookup can leave its array.

// Find_in_monotonous_1D_array, RF:3433-3494
TRK_HD int find_mono_1d(const double *A, int N, double v)
{
    int i1 = 1, i2 = N;
    int icur = (i1 + i2) / 2;
    double tv = A[icur - 1];
    if (v != v) return 1;
    if (v < A[0]) {
        icur = 0;
    } else if (v >= A[N - 1]) {
        icur = N - 1;
    } else {
        while (i1 != i2 - 1) {
            if (tv <= v) i1 = icur;
            else i2 = icur;
            icur = (i1 + i2) / 2;
            tv = A[icur - 1];
        }
    }
    return icur + 1;
}

// Find_in_monotonous_2D_array, RF:3496-3559 (exit as soon as the current
// bracket contains the value, inclusive at both ends).  On a sorted array the
// bracket is found in at most log2(N)+1 steps; the step limit only ends the
// loop on an array holding a NaN.
TRK_HD int find_mono_2d(const double *A, int N, double v)
{
    int i1 = 1, i2 = N;
    int icur = (i1 + i2) / 2;
    double tv = A[icur - 1];
    if (v != v) return 1;
    if (v < A[0]) {
        icur = 0;
    } else if (v >= A[N - 1]) {
        icur = N - 1;
    } else {
        for (int guard = 0; guard < 64 && !(v >= A[icur - 1] && v <= A[icur]); ++guard) {
            if (tv <= v) i1 = icur;
            else i2 = icur;
            icur = (i1 + i2) / 2;
            tv = A[icur - 1];
        }
    }
    return icur + 1;
}

// Find_in_monoton_array_decreasing, RF:3376-3429
TRK_HD int find_decr(const double *A, int N, double v)
{
    int i1 = 1, i2 = N;
    int icur = (i1 + i2) / 2;
    double tv = A[icur - 1];
    if (v != v) return 1;
    if (v < A[N - 1]) {
        icur = N;
    } else if (v > A[0]) {
        icur = 1;
    } else {
        while (i2 - i1 > 1) {
            if (tv > v) i1 = icur;
            else i2 = icur;
            icur = (i1 + i2) / 2;
            tv = A[icur - 1];
        }
    }
    return icur;
}

// Interpolate, Cross_sections.f90:4051-4086 (flags 1 linear, 5 log-log)
TRK_HD double interpolate(int flag, double E1, double E2, double S1, double S2, double E)
{
    if (trk_fabs(E2 - E1) < 1.0e-6) return (S1 > S2) ? S1 : S2;
    if (E == E1) return S1;
    if (flag == 5) {
        double E2l = trk_log(E2), E1l = trk_log(E1), El = trk_log(E);
        double S1l = trk_log(S1), S2l = trk_log(S2);
        double o = S1l + ((S2l - S1l) / (E2l - E1l)) * (El - E1l);
        return trk_exp(o);
    }
    return S1 + ((S2 - S1) / (E2 - E1)) * (E - E1);
}

// define_dE, Cross_sections.f90:1374-1412 (new grid, both bounds present)
TRK_HD double define_dE(int n, double E, double E0_min, double E0_max, double dE_min)
{
    double dE;
    if ((E > E0_min) && (E < E0_max)) dE = (E0_max - E0_min) / (double)n;
    else dE = E / (double)n;
    return (dE > dE_min) ? dE : dE_min;
}

// Next_free_path_2d (MC:1874-1899) in the trk namespace, for host code (the
// event code compiles its own copy in each of its namespaces)
TRK_HD double nfp_2d_host(const double *Ea, const double *La, int n, double E)
{
    if (n < 1) return 1.0e30;      // no table (no photons): an infinite path
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

}  // namespace trk
