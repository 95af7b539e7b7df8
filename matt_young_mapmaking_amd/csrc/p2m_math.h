/*
 * p2m_math.h -- portable double-precision elementary functions.
 *
 * Pixel membership in the halo painter is decided by comparisons and
 * ceil/floor/int() of values built from asin, sin, cos, acos, atan2, tan,
 * log, exp and pow.  A one-ulp difference between two math libraries can
 * move a pixel across a disc edge or a profile-table cell boundary.  This
 * header gives ONE definition of each function, written only with IEEE-754
 * operations that are correctly rounded on every target (+ - * / sqrt,
 * floor/ceil/trunc, integer conversion), so the gfx950 kernels and the CPU
 * restatement of the reference compute bit-identical results.  Both sides
 * must be compiled with floating-point contraction disabled
 * (-ffp-contract=off) and without fast-math.
 *
 * Methods (all coefficients are exact Taylor coefficients rounded once to
 * double; constants were computed from Machin's formula for pi and the
 * atanh series for ln 2):
 *   sin/cos/tan : Cody-Waite reduction by pi/2 (three 33-bit parts) and
 *                 Taylor polynomials on [-pi/4, pi/4] (to x^19 / x^20).
 *   atan/atan2  : reduction to [0,1] by reciprocal, then to |t'|<=1/16 via
 *                 atan(t) = atan(j/8) + atan((t-j/8)/(1+t j/8)), Taylor to t'^17.
 *   asin/acos   : atan2 of (x, sqrt((1-x)(1+x))).
 *   log         : x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)).
 *   exp         : x = k ln2 + r, |r| <= ln2/2, Taylor to r^13, scaled by 2^k.
 *   pow         : exp(y log x) for x > 0.
 * Accuracy against glibc is a few ulp (tests/test_math_portable.py pins it).
 */
#ifndef P2M_MATH_H
#define P2M_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define P2M_HD __host__ __device__ static inline
#else
#define P2M_HD static inline
#endif

/* ---------------------------------------------------------------- bits */
P2M_HD uint64_t p2m_dbits(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
P2M_HD double p2m_dfrom(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }

P2M_HD double p2m_sqrt(double x) { return __builtin_sqrt(x); }
P2M_HD double p2m_floor(double x) { return __builtin_floor(x); }
P2M_HD double p2m_ceil(double x) { return __builtin_ceil(x); }
P2M_HD double p2m_fabs(double x) { return __builtin_fabs(x); }
/* Fortran NINT: nearest, halves away from zero. */
P2M_HD double p2m_nint(double x) { return __builtin_round(x); }

/* ------------------------------------------------------------ constants */
#define P2M_PI_HI    0x1.921fb54442d18p+1
#define P2M_PI_LO    0x1.1a62633145c07p-53
#define P2M_PIO2_HI  0x1.921fb54442d18p+0
#define P2M_PIO2_LO  0x1.1a62633145c07p-54
#define P2M_TWOPI    0x1.921fb54442d18p+2
#define P2M_2OPI     0x1.45f306dc9c883p-1
#define P2M_PIO2_1   0x1.921fb54400000p+0
#define P2M_PIO2_2   0x1.0b4611a600000p-34
#define P2M_PIO2_3   0x1.3198a2e037073p-69
#define P2M_LN2_HI   0x1.62e42ff000000p-1
#define P2M_LN2_LO  -0x1.718432a1b0e26p-35
#define P2M_INVLN2   0x1.71547652b82fep+0
#define P2M_SQRT2    0x1.6a09e667f3bcdp+0

/* ---------------------------------------------------------- sin / cos */
/* Reduce x to r in [-pi/4, pi/4] with x = k*pi/2 + r; returns k mod 4. */
P2M_HD int p2m_rem_pio2(double x, double *r)
{
    double k = p2m_floor(x * P2M_2OPI + 0.5);
    *r = ((x - k * P2M_PIO2_1) - k * P2M_PIO2_2) - k * P2M_PIO2_3;
    int64_t ki = (int64_t)k;
    return (int)(ki & 3);
}

P2M_HD double p2m_sin_poly(double r)
{
    double z = r * r;
    double p = -0x1.2f49b46814157p-57;
    p = 0x1.952c77030ad4ap-49 + z * p;
    p = -0x1.ae7f3e733b81fp-41 + z * p;
    p = 0x1.6124613a86d09p-33 + z * p;
    p = -0x1.ae64567f544e4p-26 + z * p;
    p = 0x1.71de3a556c734p-19 + z * p;
    p = -0x1.a01a01a01a01ap-13 + z * p;
    p = 0x1.1111111111111p-7 + z * p;
    p = -0x1.5555555555555p-3 + z * p;
    return r + (r * z) * p;
}

P2M_HD double p2m_cos_poly(double r)
{
    double z = r * r;
    double p = 0x1.e542ba4020225p-62;
    p = -0x1.6827863b97d97p-53 + z * p;
    p = 0x1.ae7f3e733b81fp-45 + z * p;
    p = -0x1.93974a8c07c9dp-37 + z * p;
    p = 0x1.1eed8eff8d898p-29 + z * p;
    p = -0x1.27e4fb7789f5cp-22 + z * p;
    p = 0x1.a01a01a01a01ap-16 + z * p;
    p = -0x1.6c16c16c16c17p-10 + z * p;
    p = 0x1.5555555555555p-5 + z * p;
    return (1.0 - 0.5 * z) + (z * z) * p;
}

P2M_HD void p2m_sincos(double x, double *s, double *c)
{
    double r;
    int q = p2m_rem_pio2(x, &r);
    double sr = p2m_sin_poly(r), cr = p2m_cos_poly(r);
    switch (q) {
    case 0: *s = sr;  *c = cr;  break;
    case 1: *s = cr;  *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

P2M_HD double p2m_sin(double x) { double s, c; p2m_sincos(x, &s, &c); return s; }
P2M_HD double p2m_cos(double x) { double s, c; p2m_sincos(x, &s, &c); return c; }

P2M_HD double p2m_tan(double x)
{
    double r;
    int q = p2m_rem_pio2(x, &r);
    double sr = p2m_sin_poly(r), cr = p2m_cos_poly(r);
    return (q & 1) ? -cr / sr : sr / cr;
}

/* ------------------------------------------------------------- atan */
/* atan(t) for t in [0, 1]. */
P2M_HD double p2m_atan01(double t)
{
    int j = (int)(t * 8.0 + 0.5);
    double c = (double)j * 0.125;
    double u = (t - c) / (1.0 + t * c);
    double z = u * u;
    double p = 0x1.e1e1e1e1e1e1ep-5;
    p = -0x1.1111111111111p-4 + z * p;
    p = 0x1.3b13b13b13b14p-4 + z * p;
    p = -0x1.745d1745d1746p-4 + z * p;
    p = 0x1.c71c71c71c71cp-4 + z * p;
    p = -0x1.2492492492492p-3 + z * p;
    p = 0x1.999999999999ap-3 + z * p;
    p = -0x1.5555555555555p-2 + z * p;
    double poly = u + (u * z) * p;
    double hi, lo;
    switch (j) {
    case 0: hi = 0.0; lo = 0.0; break;
    case 1: hi = 0x1.fd5ba9aac2f6ep-4; lo = -0x1.cd37686760c17p-59; break;
    case 2: hi = 0x1.f5b75f92c80ddp-3; lo = 0x1.8ab6e3cf7afbdp-57; break;
    case 3: hi = 0x1.6f61941e4def1p-2; lo = -0x1.c63aae6f6e918p-56; break;
    case 4: hi = 0x1.dac670561bb4fp-2; lo = 0x1.a2b7f222f65e2p-56; break;
    case 5: hi = 0x1.1e00babdefeb4p-1; lo = -0x1.928df287a668fp-58; break;
    case 6: hi = 0x1.4978fa3269ee1p-1; lo = 0x1.2419a87f2a458p-56; break;
    case 7: hi = 0x1.700a7c5784634p-1; lo = -0x1.8c34d25aadef6p-56; break;
    default: hi = 0x1.921fb54442d18p-1; lo = 0x1.1a62633145c07p-55; break;
    }
    return hi + (lo + poly);
}

P2M_HD double p2m_atan2(double y, double x)
{
    if (x != x || y != y) return x + y;
    double ax = p2m_fabs(x), ay = p2m_fabs(y);
    int ysign = (int)(p2m_dbits(y) >> 63);
    double a;
    if (ay == 0.0) {
        a = (p2m_dbits(x) >> 63) ? P2M_PI_HI : 0.0;
    } else if (ax == 0.0) {
        a = P2M_PIO2_HI;
    } else if (ay <= ax) {
        a = p2m_atan01(ay / ax);
        if (x < 0.0) a = (P2M_PI_HI - a) + P2M_PI_LO;
    } else {
        a = (P2M_PIO2_HI - p2m_atan01(ax / ay)) + P2M_PIO2_LO;
        if (x < 0.0) a = (P2M_PI_HI - a) + P2M_PI_LO;
    }
    return ysign ? -a : a;
}

P2M_HD double p2m_atan(double x) { return p2m_atan2(x, 1.0); }

P2M_HD double p2m_asin(double x)
{
    if (!(p2m_fabs(x) <= 1.0)) return __builtin_nan("");
    return p2m_atan2(x, p2m_sqrt((1.0 - x) * (1.0 + x)));
}

P2M_HD double p2m_acos(double x)
{
    if (!(p2m_fabs(x) <= 1.0)) return __builtin_nan("");
    return p2m_atan2(p2m_sqrt((1.0 - x) * (1.0 + x)), x);
}

/* -------------------------------------------------------------- log */
P2M_HD double p2m_log(double x)
{
    if (x != x) return x;
    if (x == 0.0) return -__builtin_inf();
    if (x < 0.0) return __builtin_nan("");
    uint64_t u = p2m_dbits(x);
    int e = (int)((u >> 52) & 0x7ff);
    if (e == 0x7ff) return x; /* +inf */
    int eadj = 0;
    if (e == 0) { /* subnormal: scale by 2^54 */
        x = x * 0x1p54;
        u = p2m_dbits(x);
        e = (int)((u >> 52) & 0x7ff);
        eadj = -54;
    }
    double m = p2m_dfrom((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    int k = e - 1023 + eadj;
    if (m > P2M_SQRT2) { m = m * 0.5; k += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 0x1.642c8590b2164p-4;
    p = 0x1.8618618618618p-4 + z * p;
    p = 0x1.af286bca1af28p-4 + z * p;
    p = 0x1.e1e1e1e1e1e1ep-4 + z * p;
    p = 0x1.1111111111111p-3 + z * p;
    p = 0x1.3b13b13b13b14p-3 + z * p;
    p = 0x1.745d1745d1746p-3 + z * p;
    p = 0x1.c71c71c71c71cp-3 + z * p;
    p = 0x1.2492492492492p-2 + z * p;
    p = 0x1.999999999999ap-2 + z * p;
    p = 0x1.5555555555555p-1 + z * p;
    double logm = 2.0 * s + (s * z) * p;
    double dk = (double)k;
    return dk * P2M_LN2_HI + (logm + dk * P2M_LN2_LO);
}

/* -------------------------------------------------------------- exp */
P2M_HD double p2m_ldexp_k(double r, int k)
{
    /* r * 2^k for r in [0.5, 2] and k in [-1100, 1100] */
    int k1 = k / 2, k2 = k - k1;
    double f1 = p2m_dfrom((uint64_t)(int64_t)(k1 + 1023) << 52);
    double f2 = p2m_dfrom((uint64_t)(int64_t)(k2 + 1023) << 52);
    return (r * f1) * f2;
}

P2M_HD double p2m_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -745.2) return 0.0;
    double k = p2m_floor(x * P2M_INVLN2 + 0.5);
    double r = (x - k * P2M_LN2_HI) - k * P2M_LN2_LO;
    double p = 0x1.6124613a86d09p-33;
    p = 0x1.1eed8eff8d898p-29 + r * p;
    p = 0x1.ae64567f544e4p-26 + r * p;
    p = 0x1.27e4fb7789f5cp-22 + r * p;
    p = 0x1.71de3a556c734p-19 + r * p;
    p = 0x1.a01a01a01a01ap-16 + r * p;
    p = 0x1.a01a01a01a01ap-13 + r * p;
    p = 0x1.6c16c16c16c17p-10 + r * p;
    p = 0x1.1111111111111p-7 + r * p;
    p = 0x1.5555555555555p-5 + r * p;
    p = 0x1.5555555555555p-3 + r * p;
    p = 0.5 + r * p;
    double er = 1.0 + (r + (r * r) * p);
    return p2m_ldexp_k(er, (int)k);
}

/* x**y for x > 0 (x == 0 gives 0 for y > 0). */
P2M_HD double p2m_pow(double x, double y)
{
    if (x == 0.0) return (y > 0.0) ? 0.0 : __builtin_inf();
    if (x < 0.0) return __builtin_nan("");
    return p2m_exp(y * p2m_log(x));
}

#endif /* P2M_MATH_H */
