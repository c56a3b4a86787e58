/*
 * pt_math.h -- the trace's own FP64 transcendental functions, shared verbatim by
 * the gfx950 kernels (device), the C-ABI host code and any CPU replay of the trace.
 *
 * Why own functions: the Monte Carlo trace (calcTransmit, src/RTMCphon.f90:1608-2414)
 * evaluates log(u) for free flights (RT:2172), and cos/sin of 2*pi*u for azimuths
 * (RT:1870, 2203, 2220).  Device libm (ocml) and host libm (glibc) may differ in the
 * last ulp, and a one-ulp difference in a direction can change a later branch.  These
 * functions use only IEEE-754 +, -, *, / and bit manipulation, so with floating-point
 * contraction disabled (-ffp-contract=off) they return identical bits on the CPU and
 * on the GPU.
 *
 * Methods (from the mathematics, not from any vendor library):
 *   log:    x = 2^e * m with m in [sqrt(1/2), sqrt(2)];  log m = 2 atanh(s),
 *           s = (m-1)/(m+1), atanh by its Taylor series (|s| <= 0.1716, 11 odd terms);
 *           log x = e*ln2_hi + (e*ln2_lo + log m)  (ln2 split so e*ln2_hi is exact).
 *   sincos: Cody-Waite reduction r = x - k*pi/2 with a three-part pi/2 (first two parts
 *           with 33 significant bits so k*P1, k*P2 are exact for |k| < 2^20), then
 *           Taylor series of sin and cos on |r| <= pi/4 (terms to r^17 and r^18), and
 *           the quadrant k mod 4 selects signs.
 * Accuracy: within a few ulp of the correctly rounded value on the domains used by the
 * trace (log on (0,1], sincos on [0, 2*pi]); the unit tests bound it against libm.
 *
 * Plain C99 so that it compiles as C (gcc), C++ (g++) and HIP (hipcc, host+device).
 */
#ifndef PT_MATH_H
#define PT_MATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define PT_HD __host__ __device__ static inline
#else
#define PT_HD static inline
#endif

PT_HD uint64_t pt_f64_bits(double x) { uint64_t b; memcpy(&b, &x, 8); return b; }
PT_HD double pt_bits_f64(uint64_t b) { double x; memcpy(&x, &b, 8); return x; }

/* ln 2 = PT_LN2_HI + PT_LN2_LO; PT_LN2_HI keeps 32 significant bits. */
#define PT_LN2_HI 0x1.62e42fee00000p-1
#define PT_LN2_LO 0x1.a39ef35793c76p-33
/* pi/2 = P1 + P2 + P3; P1, P2 keep 33 significant bits. */
#define PT_PIO2_1 0x1.921fb54400000p+0
#define PT_PIO2_2 0x1.0b4611a600000p-34
#define PT_PIO2_3 0x1.3198a2e037073p-69
#define PT_TWO_OVER_PI 0x1.45f306dc9c883p-1
/* 1.5 * 2^52: adding and subtracting it rounds a |t| < 2^51 to the nearest integer. */
#define PT_ROUND_MAGIC 6755399441055744.0

PT_HD double pt_log(double x)
{
    uint64_t b = pt_f64_bits(x);
    int e;
    double m, f, s, z, p, lm, ed;
    if (x != x || x < 0.0) return (x - x) / (x - x);          /* NaN */
    if (x == 0.0) return -1.0 / 0.0;                           /* -inf */
    if ((b >> 52) == 0x7ff) return x;                          /* +inf */
    e = 0;
    if ((b >> 52) == 0) {                                      /* subnormal: scale by 2^54 */
        x = x * 18014398509481984.0;
        b = pt_f64_bits(x);
        e = -54;
    }
    e += (int)((b >> 52) & 0x7ff) - 1023;
    m = pt_bits_f64((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
    if (m > 0x1.6a09e667f3bcdp+0) {                           /* m > sqrt(2) */
        m = m * 0.5;
        e += 1;
    }
    f = m - 1.0;                                               /* exact (Sterbenz) */
    s = f / (2.0 + f);
    z = s * s;
    p = 1.0 / 23.0;
    p = 1.0 / 21.0 + z * p;
    p = 1.0 / 19.0 + z * p;
    p = 1.0 / 17.0 + z * p;
    p = 1.0 / 15.0 + z * p;
    p = 1.0 / 13.0 + z * p;
    p = 1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = 1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = 1.0 / 3.0 + z * p;
    lm = 2.0 * s + (2.0 * s) * (z * p);
    ed = (double)e;
    return ed * PT_LN2_HI + (ed * PT_LN2_LO + lm);
}

/* sin and cos of r, |r| <= pi/4 (Taylor, coefficients are +-1/n!). */
PT_HD void pt_sincos_kernel(double r, double *s, double *c)
{
    double z = r * r, ps, pc;
    ps = 1.0 / 355687428096000.0;              /*  1/17! */
    ps = -1.0 / 1307674368000.0 + z * ps;      /* -1/15! */
    ps = 1.0 / 6227020800.0 + z * ps;          /*  1/13! */
    ps = -1.0 / 39916800.0 + z * ps;           /* -1/11! */
    ps = 1.0 / 362880.0 + z * ps;              /*  1/9!  */
    ps = -1.0 / 5040.0 + z * ps;               /* -1/7!  */
    ps = 1.0 / 120.0 + z * ps;                 /*  1/5!  */
    ps = -1.0 / 6.0 + z * ps;                  /* -1/3!  */
    *s = r + r * (z * ps);
    pc = 1.0 / 6402373705728000.0;             /*  1/18! */
    pc = -1.0 / 20922789888000.0 + z * pc;     /* -1/16! */
    pc = 1.0 / 87178291200.0 + z * pc;         /*  1/14! */
    pc = -1.0 / 479001600.0 + z * pc;          /* -1/12! */
    pc = 1.0 / 3628800.0 + z * pc;             /*  1/10! */
    pc = -1.0 / 40320.0 + z * pc;              /* -1/8!  */
    pc = 1.0 / 720.0 + z * pc;                 /*  1/6!  */
    pc = -1.0 / 24.0 + z * pc;                 /* -1/4!  */
    /* cos r = 1 - z/2 + z^2 * pc */
    *c = (1.0 - 0.5 * z) + (z * z) * (-pc);
}

/* sin and cos of x for |x| < 2^20 (the trace uses x in [0, 2*pi]). */
PT_HD void pt_sincos(double x, double *s, double *c)
{
    double t = x * PT_TWO_OVER_PI;
    double kd = (t + PT_ROUND_MAGIC) - PT_ROUND_MAGIC;
    double r = ((x - kd * PT_PIO2_1) - kd * PT_PIO2_2) - kd * PT_PIO2_3;
    double sr, cr;
    int q = ((int)kd) & 3;
    pt_sincos_kernel(r, &sr, &cr);
    switch (q) {
    case 0: *s = sr;  *c = cr;  break;
    case 1: *s = cr;  *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

PT_HD double pt_sin(double x) { double s, c; pt_sincos(x, &s, &c); return s; }
PT_HD double pt_cos(double x) { double s, c; pt_sincos(x, &s, &c); return c; }

#endif /* PT_MATH_H */
