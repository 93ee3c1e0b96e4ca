/*
 * gwh_trig.h -- sine and cosine of FP64 phases for the per-bin kernels.
 *
 * Phases of the hot path reach ~1e5-1e6 rad (2 pi f t_c plus the inspiral phase).  The library sincos of
 * FP64 carries a large-argument (Payne-Hanek) reduction for every call; here the argument is reduced by
 * Cody and Waite's method (W. J. Cody, W. Waite, Software Manual for the Elementary Functions, 1980):
 * q = rint(x 2/pi), r = (x - q p1) - q p2 with pi/2 = p1 + p2 split into two doubles and each step one
 * fused multiply-add (x - q p1 is exact by Sterbenz's lemma, the neglected part of pi/2 is 1e-33, so r is
 * within half an ulp for |x| < 2^30), then sin r and cos r on [-pi/4, pi/4] by their Taylor series and the
 * quadrant by q mod 4.  Arguments beyond 2^30 rad take the library routine.
 *
 * gwh_sincos_table and gwh_exp_table trade the long series for a table lookup (gwh_trig_tables.h, 4.5 kB,
 * correctly rounded entries) and a short series on a tiny remainder: ~20 FP64 operations and one gather for
 * sin and cos together.  Any index they form stays inside the tables, whatever the argument (NaN included).
 */
#ifndef GWH_TRIG_H
#define GWH_TRIG_H

#include <hip/hip_runtime.h>

#include "gwh_trig_tables.h"

/* NT: highest odd power of the sine series (15: truncation below 5e-17 on [-pi/4, pi/4]; 11: below 7e-12) */
template <int NT> __device__ __forceinline__ void gwh_sincos_cw(double x, double &s, double &c)
{
    static_assert(NT == 15 || NT == 11, "series length");
    if (!(fabs(x) < 1073741824.0)) { /* |x| >= 2^30 or NaN */
        sincos(x, &s, &c);
        return;
    }
    const double q = rint(x * 0.63661977236758134308);
    double r = fma(-q, 1.5707963267948966192, x);
    r = fma(-q, 6.123233995736766036e-17, r);
    const double r2 = r * r;
    /* sin r = r + r^3 sum_k (-1)^k r^(2k) / (2k+3)!,  cos r = 1 - r^2/2 + r^4 sum_k (-1)^k r^(2k) / (2k+4)! */
    double ps, pc;
    if (NT == 15) {
        ps = fma(-7.6471637318198164759e-13, r2, 1.6059043836821614599e-10); /* -1/15!, 1/13! */
        ps = fma(ps, r2, -2.5052108385441718775e-08);                        /* -1/11! */
        ps = fma(ps, r2, 2.7557319223985890653e-06);                         /* 1/9! */
        pc = fma(4.7794773323873852974e-14, r2, -1.1470745597729724714e-11); /* 1/16!, -1/14! */
        pc = fma(pc, r2, 2.0876756987868098979e-09);                         /* 1/12! */
        pc = fma(pc, r2, -2.7557319223985890653e-07);                        /* -1/10! */
    } else {
        ps = fma(-2.5052108385441718775e-08, r2, 2.7557319223985890653e-06); /* -1/11!, 1/9! */
        pc = fma(2.0876756987868098979e-09, r2, -2.7557319223985890653e-07); /* 1/12!, -1/10! */
    }
    ps = fma(ps, r2, -1.9841269841269841270e-04); /* -1/7! */
    ps = fma(ps, r2, 8.3333333333333333333e-03);  /* 1/5! */
    ps = fma(ps, r2, -1.6666666666666666667e-01); /* -1/3! */
    pc = fma(pc, r2, 2.4801587301587301587e-05);  /* 1/8! */
    pc = fma(pc, r2, -1.3888888888888888889e-03); /* -1/6! */
    pc = fma(pc, r2, 4.1666666666666666667e-02);  /* 1/4! */
    pc = fma(pc, r2, -0.5);
    const double sr = fma(ps * r2, r, r);
    const double cr = fma(pc, r2, 1.0);
    const int iq = (int)q;
    const double ss = (iq & 1) ? cr : sr, cc = (iq & 1) ? sr : cr;
    s = (iq & 2) ? -ss : ss;
    c = ((iq + 1) & 2) ? -cc : cc;
}

/* sin and cos of an FP64 phase through the table: the phase in turns r = n / 256 + y with n the nearest integer
 * to 256 r (rounded by the 1.5 * 2^52 addition, whose low mantissa bits then hold n mod 256) and y = r - n / 256
 * exact, |y| <= 1/512; sin and cos of t = 2 pi y (|t| <= pi / 256) by their Taylor series to t^5 and t^6
 * (truncation below 1e-17), rotated by the table entry.  Error a few ulp of 1 beyond the rounding of the phase
 * to turns (the same reduction the FP32 path takes); valid for |phase| < 2^44 rad. */
template <class Tab> __device__ __forceinline__ void gwh_sincos_tab(double phase, double &s, double &c, const Tab *tab)
{
    const double r = phase * 0.15915494309189533577;
    const double kRound = 6755399441055744.0; /* 1.5 * 2^52 */
    const double rb = fma(r, 256.0, kRound);
    const double n = rb - kRound;
    const int j = (int)(__double_as_longlong(rb) & 255);
    const double y = fma(n, -0.00390625, r);
    const double t = y * 6.283185307179586477;
    const double t2 = t * t;
    const double sl = t * fma(t2, fma(t2, 1.0 / 120.0, -1.0 / 6.0), 1.0);
    const double cl = fma(t2, fma(t2, fma(t2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0);
    const double2 T = tab[j];
    s = fma(T.y, cl, T.x * sl);
    c = fma(T.x, cl, -T.y * sl);
}
/* the same through the table in global memory */
__device__ __forceinline__ void gwh_sincos_table(double phase, double &s, double &c)
{
    gwh_sincos_tab(phase, s, c, gwh_turn_table);
}

/* e^x for |x| < 700 through the 2^(j/64) table: x = n ln2 / 64 + r with n = rint(64 x / ln2) and the reduction
 * by a two-part ln2 / 64 (r exact to ~1e-19, |r| <= ln2 / 128), e^r by its Taylor series to r^5 (truncation
 * below 4e-17), scaled by 2^(n mod 64 / 64) from the table and 2^(n div 64) exactly.  Error a few ulp. */
__device__ __forceinline__ double gwh_exp_table(double x, const double *tab = gwh_exp2_table)
{
    const double n = rint(x * 92.332482616893656877); /* 64 / ln 2 */
    double r = fma(-n, 0.01083042469326756, x);        /* ln2/64, high 32 bits (n * hi exact for |n| < 2^21) */
    r = fma(-n, 2.9815858269852933e-12, r);            /* ln2/64, the rest */
    const double p = fma(r, fma(r, fma(r, fma(r, fma(r, 1.0 / 120.0, 1.0 / 24.0), 1.0 / 6.0), 0.5), 1.0), 1.0);
    const int ni = (int)n;
    return ldexp(tab[ni & 63] * p, ni >> 6);
}

/* e^x for x <= 0 of moderate size: FP64 split x log2(e) = n + t, FP32 2^t (~6e-8 relative), exact scaling */
__device__ __forceinline__ double exp_fp32(double x)
{
    double y = x * 1.4426950408889634074;
    double n = rint(y);
    float t = (float)(y - n);
    return ldexp((double)__builtin_amdgcn_exp2f(t), (int)n);
}

/* 1/z in FP64: hardware reciprocal and one Newton step (no IEEE division sequence) */
__device__ __forceinline__ double rcp_newton(double z)
{
    double r = __builtin_amdgcn_rcp(z);
    r = fma(r, fma(-z, r, 1.0), r);
    return r;
}

/* z^{-1/2} in FP64 from an FP32 v_rsq_f32 seed (~1e-7) and two Newton steps (~1e-16) */
__device__ __forceinline__ double rsqrt_seeded(double z)
{
    double y = (double)__builtin_amdgcn_rsqf((float)z);
    y = y * (1.5 - 0.5 * z * y * y);
    y = y * (1.5 - 0.5 * z * y * y);
    return y;
}

/* arctangent in FP64 from an FP32 seed: t0 = atanf(x) (within ~1e-7), then the exact correction
 * atan(x) = t0 + atan((x cos t0 - sin t0) / (cos t0 + x sin t0)) with the table sin/cos, where the inner argument
 * is below ~1e-7, so its arctangent is the argument itself to within 1e-21; the denominator is
 * sqrt(1 + x^2) >= 1.  Error a few ulp of pi / 2 for finite x. */
__device__ __forceinline__ double gwh_atan_seeded(double x, const double2 *tab = gwh_turn_table)
{
    const double t0 = (double)atanf((float)x);
    double s, c;
    gwh_sincos_tab(t0, s, c, tab);
    return t0 + fma(x, c, -s) * rcp_newton(fma(x, s, c));
}

#endif
