/*
 * trk_detmath.h -- deterministic FP64 elementary functions for host and device.
 *
 * The TREKIS-3 particle-history path calls log, exp, sin, cos, acos, atan,
 * sqrt and real-exponent powers on every event (Monte_Carlo.f90 and
 * Cross_sections.f90 throughout; e.g. free-path sampling -log(RN) at
 * Monte_Carlo.f90:2332, log-log interpolation at Cross_sections.f90:4078-4083,
 * acos in the kinematics at Monte_Carlo.f90:1155, 1196).  The host libm and
 * the device math library round these differently in the last bit, which
 * would make a CPU history and a GPU history drawn from the same random
 * stream drift apart.  Every function below is built only from IEEE-754
 * +, -, *, / and sqrt (all correctly rounded on x86-64 SSE2 and on CDNA4),
 * evaluated in a fixed order, so the same bits come out on both sides as long
 * as the translation unit is compiled with -ffp-contract=off and without
 * fast-math.
 *
 * Methods (all textbook, no vendor constants):
 *   exp   Cody-Waite reduction x = k ln2 + r with a 2-part ln2, Taylor series
 *         of e^r to degree 18 (|r| <= ln2/2), exact scaling by 2^k.
 *   log   x = 2^e m, m in (sqrt(1/2), sqrt(2)], log m = 2 atanh(s) with
 *         s = (m-1)/(m+1), odd Taylor series to s^27.
 *   sin/cos  Cody-Waite reduction by pi/2 with a 3-part split (exact for
 *         |x| < 2^19 pi/2), Taylor polynomials on |r| <= pi/4.
 *   atan  two half-angle steps atan x = 2 atan(x / (1 + sqrt(1 + x^2)))
 *         bring |x| <= tan(pi/16), then the Gregory series to x^27;
 *         |x| > 1 uses pi/2 - atan(1/x).
 *   acos  acos x = 2 atan(sqrt((1-x)/(1+x))).
 *   pow   exp(y log x) for x > 0.
 * Accuracy is a few ulp (checked against mpmath in the tests); the point is
 * reproducibility, not correct rounding.
 */
#ifndef TRK_DETMATH_H
#define TRK_DETMATH_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define TRK_FN __host__ __device__ static inline
#else
#define TRK_FN static inline
#endif

TRK_FN uint64_t trk_d2u(double x) { union { double d; uint64_t u; } c; c.d = x; return c.u; }
TRK_FN double trk_u2d(uint64_t u) { union { double d; uint64_t u; } c; c.u = u; return c.d; }

#define TRK_INF (trk_u2d(0x7FF0000000000000ull))
#define TRK_NAN (trk_u2d(0x7FF8000000000000ull))

TRK_FN int trk_isnan(double x) { return x != x; }
TRK_FN double trk_fabs(double x) { return trk_u2d(trk_d2u(x) & 0x7FFFFFFFFFFFFFFFull); }
TRK_FN double trk_sqrt(double x) { return __builtin_sqrt(x); }

/* Round to the nearest integer value, ties to even; valid for |x| < 2^51. */
TRK_FN double trk_rint(double x)
{
    const double magic = 0x1.8p+52;
    double t = x + magic;            /* strict IEEE evaluation: not folded without fast-math */
    return t - magic;
}

/* 2^k for -1022 <= k <= 1023 */
TRK_FN double trk_pow2i(int k) { return trk_u2d((uint64_t)(k + 1023) << 52); }

/* --- exp ------------------------------------------------------------------ */
TRK_FN double trk_exp(double x)
{
    const double INV_LN2 = 0x1.71547652b82fep+0;
    const double LN2_HI = 0x1.62e42fee00000p-1;
    const double LN2_LO = 0x1.a39ef35793c76p-33;
    if (x != x) return x;
    if (x > 709.782712893384) return TRK_INF;
    if (x < -745.1332191019412) return 0.0;
    double kd = trk_rint(x * INV_LN2);
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    double p = 0x1.6827863b97d97p-53;          /* 1/18! */
    p = 0x1.952c77030ad4ap-49 + r * p;         /* 1/17! */
    p = 0x1.ae7f3e733b81fp-45 + r * p;         /* 1/16! */
    p = 0x1.ae7f3e733b81fp-41 + r * p;         /* 1/15! */
    p = 0x1.93974a8c07c9dp-37 + r * p;         /* 1/14! */
    p = 0x1.6124613a86d09p-33 + r * p;         /* 1/13! */
    p = 0x1.1eed8eff8d898p-29 + r * p;         /* 1/12! */
    p = 0x1.ae64567f544e4p-26 + r * p;         /* 1/11! */
    p = 0x1.27e4fb7789f5cp-22 + r * p;         /* 1/10! */
    p = 0x1.71de3a556c734p-19 + r * p;         /* 1/9!  */
    p = 0x1.a01a01a01a01ap-16 + r * p;         /* 1/8!  */
    p = 0x1.a01a01a01a01ap-13 + r * p;         /* 1/7!  */
    p = 0x1.6c16c16c16c17p-10 + r * p;         /* 1/6!  */
    p = 0x1.1111111111111p-7 + r * p;          /* 1/5!  */
    p = 0x1.5555555555555p-5 + r * p;          /* 1/4!  */
    p = 0x1.5555555555555p-3 + r * p;          /* 1/3!  */
    p = 0x1.0000000000000p-1 + r * p;          /* 1/2!  */
    p = r + (r * r) * p;
    double e = 1.0 + p;
    int k = (int)kd;
    int k1 = k / 2;
    int k2 = k - k1;
    return (e * trk_pow2i(k1)) * trk_pow2i(k2);
}

/* --- log ------------------------------------------------------------------ */
TRK_FN double trk_log(double x)
{
    const double LN2_HI = 0x1.62e42fee00000p-1;
    const double LN2_LO = 0x1.a39ef35793c76p-33;
    const double SQRT2 = 0x1.6a09e667f3bcdp+0;
    if (x != x) return x;
    if (x < 0.0) return TRK_NAN;
    if (x == 0.0) return -TRK_INF;
    if (x == TRK_INF) return x;
    int e = 0;
    uint64_t u = trk_d2u(x);
    if ((u >> 52) == 0) {                      /* subnormal */
        x = x * 0x1.0p+54;
        u = trk_d2u(x);
        e = -54;
    }
    e += (int)(u >> 52) - 1023;
    double m = trk_u2d((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > SQRT2) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 0x1.2f684bda12f68p-5;          /* 1/27 */
    R = 0x1.47ae147ae147bp-5 + z * R;          /* 1/25 */
    R = 0x1.642c8590b2164p-5 + z * R;          /* 1/23 */
    R = 0x1.8618618618618p-5 + z * R;          /* 1/21 */
    R = 0x1.af286bca1af28p-5 + z * R;          /* 1/19 */
    R = 0x1.e1e1e1e1e1e1ep-5 + z * R;          /* 1/17 */
    R = 0x1.1111111111111p-4 + z * R;          /* 1/15 */
    R = 0x1.3b13b13b13b14p-4 + z * R;          /* 1/13 */
    R = 0x1.745d1745d1746p-4 + z * R;          /* 1/11 */
    R = 0x1.c71c71c71c71cp-4 + z * R;          /* 1/9  */
    R = 0x1.2492492492492p-3 + z * R;          /* 1/7  */
    R = 0x1.999999999999ap-3 + z * R;          /* 1/5  */
    R = 0x1.5555555555555p-2 + z * R;          /* 1/3  */
    double s2 = 2.0 * s;
    double logm = s2 + s2 * (z * R);
    double ed = (double)e;
    return ed * LN2_HI + (ed * LN2_LO + logm);
}

/* --- sin / cos -------------------------------------------------------------- */
TRK_FN double trk_sin_poly(double r)           /* |r| <= pi/4 */
{
    double z = r * r;
    double p = 0x1.952c77030ad4ap-49;          /*  1/17! */
    p = -0x1.ae7f3e733b81fp-41 + z * p;        /* -1/15! */
    p = 0x1.6124613a86d09p-33 + z * p;         /*  1/13! */
    p = -0x1.ae64567f544e4p-26 + z * p;        /* -1/11! */
    p = 0x1.71de3a556c734p-19 + z * p;         /*  1/9!  */
    p = -0x1.a01a01a01a01ap-13 + z * p;        /* -1/7!  */
    p = 0x1.1111111111111p-7 + z * p;          /*  1/5!  */
    p = -0x1.5555555555555p-3 + z * p;         /* -1/3!  */
    return r + (r * z) * p;
}

TRK_FN double trk_cos_poly(double r)           /* |r| <= pi/4 */
{
    double z = r * r;
    double q = -0x1.6827863b97d97p-53;         /* -1/18! */
    q = 0x1.ae7f3e733b81fp-45 + z * q;         /*  1/16! */
    q = -0x1.93974a8c07c9dp-37 + z * q;        /* -1/14! */
    q = 0x1.1eed8eff8d898p-29 + z * q;         /*  1/12! */
    q = -0x1.27e4fb7789f5cp-22 + z * q;        /* -1/10! */
    q = 0x1.a01a01a01a01ap-16 + z * q;         /*  1/8!  */
    q = -0x1.6c16c16c16c17p-10 + z * q;        /* -1/6!  */
    q = 0x1.5555555555555p-5 + z * q;          /*  1/4!  */
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * z) * q);
}

TRK_FN double trk_reduce_pio2(double x, int *q)
{
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double PIO2_1 = 0x1.921fb54400000p+0;
    const double PIO2_2 = 0x1.0b4611a600000p-34;
    const double PIO2_3 = 0x1.3198a2e037073p-69;
    double kd = trk_rint(x * TWO_OVER_PI);
    double r = ((x - kd * PIO2_1) - kd * PIO2_2) - kd * PIO2_3;
    *q = (int)((int64_t)kd & 3);
    return r;
}

/* Both polynomials are always evaluated and the quadrant picks one: on a GPU
 * the lanes of a wave sit in different quadrants, and a branch per quadrant
 * would run every path anyway. */
TRK_FN double trk_sin(double x)
{
    if (x != x || trk_fabs(x) == TRK_INF) return TRK_NAN;
    int q;
    double r = trk_reduce_pio2(x, &q);
    double sp = trk_sin_poly(r), cp = trk_cos_poly(r);
    double v = (q & 1) ? cp : sp;
    return (q & 2) ? -v : v;
}

TRK_FN double trk_cos(double x)
{
    if (x != x || trk_fabs(x) == TRK_INF) return TRK_NAN;
    int q;
    double r = trk_reduce_pio2(x, &q);
    double sp = trk_sin_poly(r), cp = trk_cos_poly(r);
    double v = (q & 1) ? sp : cp;
    return ((q + 1) & 2) ? -v : v;
}

/* sin and cos of one argument from one reduction; each equals trk_sin(x) /
 * trk_cos(x) bit for bit. */
TRK_FN void trk_sincos(double x, double *s, double *c)
{
    if (x != x || trk_fabs(x) == TRK_INF) { *s = TRK_NAN; *c = TRK_NAN; return; }
    int q;
    double r = trk_reduce_pio2(x, &q);
    double sp = trk_sin_poly(r), cp = trk_cos_poly(r);
    double vs = (q & 1) ? cp : sp;
    double vc = (q & 1) ? sp : cp;
    *s = (q & 2) ? -vs : vs;
    *c = ((q + 1) & 2) ? -vc : vc;
}

/* --- atan / acos ------------------------------------------------------------ */
TRK_FN double trk_atan_small(double x)          /* 0 <= x <= 1 */
{
    double y = x / (1.0 + trk_sqrt(1.0 + x * x));
    y = y / (1.0 + trk_sqrt(1.0 + y * y));
    double z = y * y;
    double p = 0x1.2f684bda12f68p-5;           /*  1/27 */
    p = -0x1.47ae147ae147bp-5 + z * p;         /* -1/25 */
    p = 0x1.642c8590b2164p-5 + z * p;          /*  1/23 */
    p = -0x1.8618618618618p-5 + z * p;         /* -1/21 */
    p = 0x1.af286bca1af28p-5 + z * p;          /*  1/19 */
    p = -0x1.e1e1e1e1e1e1ep-5 + z * p;         /* -1/17 */
    p = 0x1.1111111111111p-4 + z * p;          /*  1/15 */
    p = -0x1.3b13b13b13b14p-4 + z * p;         /* -1/13 */
    p = 0x1.745d1745d1746p-4 + z * p;          /*  1/11 */
    p = -0x1.c71c71c71c71cp-4 + z * p;         /* -1/9  */
    p = 0x1.2492492492492p-3 + z * p;          /*  1/7  */
    p = -0x1.999999999999ap-3 + z * p;         /* -1/5  */
    p = 0x1.5555555555555p-2 + z * p;          /*  1/3  */
    double a = y - y * (z * p);
    return 4.0 * a;
}

TRK_FN double trk_atan(double x)
{
    const double PIO2 = 0x1.921fb54442d18p+0;
    const double PIO2_LO = 0x1.1a62633145c07p-54;
    if (x != x) return x;
    double ax = trk_fabs(x);
    const int big = ax > 1.0;
    double a = trk_atan_small(big ? 1.0 / ax : ax);   /* one evaluation for every lane */
    double r = big ? (PIO2 - a) + PIO2_LO : a;
    return (x < 0.0) ? -r : r;
}

TRK_FN double trk_acos(double x)
{
    const double PI = 0x1.921fb54442d18p+1;
    if (x != x) return x;
    if (x > 1.0 || x < -1.0) return TRK_NAN;
    if (x == 1.0) return 0.0;
    if (x == -1.0) return PI;
    double t = trk_sqrt((1.0 - x) / (1.0 + x));
    return 2.0 * trk_atan(t);
}

/* --- pow ------------------------------------------------------------------ */
TRK_FN double trk_pow(double x, double y)
{
    if (x != x || y != y) return TRK_NAN;
    if (y == 0.0) return 1.0;
    if (x == 0.0) return (y > 0.0) ? 0.0 : TRK_INF;
    if (x < 0.0) return TRK_NAN;
    return trk_exp(y * trk_log(x));
}

#endif /* TRK_DETMATH_H */
