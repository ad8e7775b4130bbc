/*
 * sun_libm.h -- FP64 log, exp, sine and cosine (of 2*pi*u and of a phase) and atan2,
 * computed by one fixed sequence of IEEE-754 operations, so that host code (gcc) and device code
 * (hipcc for gfx950) produce the same bits from the same input.
 *
 * The heatbath's SU(2) sampler draws from these functions (device_PHB_OVR.h:162-223
 * calls log, cos, sin and exp), the Haar start draws Gaussians through log,
 * sin and cos, and the 8-real storage codec keeps two phases (atan2 to store, sin and
 * cos to load).  With library transcendentals the host and the device differ in the
 * last bit now and then, and a sweep amplifies that (an accept/reject test on a
 * near tie, a chaotic update), so FP64 chains could only be compared within a
 * tolerance.  With these, the FP64 device path and its host restatement agree
 * bit for bit.
 *
 * Method: Cody and Waite (W. J. Cody, W. Waite, "Software Manual for the Elementary
 * Functions", Prentice-Hall 1980): reduce the argument by an exactly computed
 * multiple of a constant that is split into a short high part (so the product is
 * exact) and a low part, then evaluate a polynomial on the small reduced range.
 * The polynomials here are truncated Taylor series, whose coefficients are the
 * exact rationals 1/n! and 2/(2n+1) rounded to double (written as hex literals,
 * so no compiler converts decimal text):
 *   log x    x = 2^k m with m in [sqrt(1/2), sqrt(2)); s = (m-1)/(m+1), |s| <= 0.1716;
 *            log m = 2 atanh s = 2s + sum_{n=1..10} 2 s^(2n+1)/(2n+1) (the first
 *            omitted term is below 2^-61 of the sum); log x = k ln2_hi + (k ln2_lo
 *            + log m), ln2_hi holding 32 significant bits so k ln2_hi is exact.
 *   exp x    k = rint(x/ln 2), r = (x - k ln2_hi) - k ln2_lo, |r| <= 0.347;
 *            e^r = sum_{n=0..14} r^n/n! (omitted term < 2^-57 e^r); exp x = 2^k e^r.
 *   sin, cos (2 pi u)   v = 4u, q = rint(v), t = v - q (exact, |t| <= 1/2); theta =
 *            t pi/2 with pi/2 as a double pair, |theta| <= pi/4; the Taylor series of
 *            sin to theta^19 and of cos to theta^18; the quadrant q mod 4 selects and
 *            negates.  The argument in revolutions needs no large-argument reduction.
 *   sin, cos (theta)    q = rint(2 theta/pi), r = theta - q pi/2 by two fused steps
 *            against the pi/2 pair, then the same series.
 *   atan2    t = min/max of |y|, |x| in [0, 1]; above tan(pi/8), atan t = pi/4 +
 *            atan((t-1)/(t+1)); the Taylor series of atan z to z^41 on |z| <=
 *            tan(pi/8); octant and quadrant by pi/2 - a, pi - a and the signs.
 * Every multiply-add is an explicit fma() (fused identically by both compilers) and
 * the functions forbid contraction of the remaining operations, so the bits do not
 * depend on a compiler's -ffp-contract setting.  frexp, ldexp and rint are exact.
 * Accuracy (tested against glibc): within 2 ulp for log on (0, inf) and for exp on
 * [-700, 700], 3 ulp for atan2; sin and cos within 2 x 2^-53 on u in (0, 1] and on
 * theta in [-pi, pi].
 *
 * Plain C (the host restatement is C) and HIP C++.
 */
#ifndef SUN_LIBM_H
#define SUN_LIBM_H

#include <math.h>

#ifdef __HIPCC__
#define SUN_LIBM_HD __host__ __device__
#else
#define SUN_LIBM_HD
#endif

#if defined(__clang__)
#define SUN_LIBM_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define SUN_LIBM_NOCONTRACT
#endif

#define SUN_LN2_HI 0x1.62e42fee00000p-1     /* ln 2, first 32 significant bits */
#define SUN_LN2_LO 0x1.a39ef35793c76p-33    /* ln 2 - SUN_LN2_HI               */
#define SUN_INV_LN2 0x1.71547652b82fep+0    /* 1 / ln 2                        */
#define SUN_PIO2_HI 0x1.921fb54442d18p+0    /* pi / 2 rounded                  */
#define SUN_PIO2_LO 0x1.1a62633145c07p-54   /* pi / 2 - SUN_PIO2_HI            */
#define SUN_SQRT_HALF 0x1.6a09e667f3bcdp-1  /* sqrt(1/2)                       */
#define SUN_TWO_OVER_PI 0x1.45f306dc9c883p-1 /* 2 / pi                         */
#define SUN_PIO4_HI 0x1.921fb54442d18p-1    /* pi / 4 rounded                  */
#define SUN_PIO4_LO 0x1.1a62633145c07p-55   /* pi / 4 - SUN_PIO4_HI            */
#define SUN_PI_HI 0x1.921fb54442d18p+1      /* pi rounded                      */
#define SUN_PI_LO 0x1.1a62633145c07p-53     /* pi - SUN_PI_HI                  */
#define SUN_TAN_PI8 0x1.a827999fcef32p-2    /* tan(pi/8) = sqrt(2) - 1         */

/* natural logarithm; x <= 0 gives -inf (x = 0) or NaN, as does NaN */
static inline SUN_LIBM_HD double sun_log(double x) {
    SUN_LIBM_NOCONTRACT
    if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    int e;
    double m = frexp(x, &e); /* x = m 2^e, m in [1/2, 1) */
    if (m < SUN_SQRT_HALF) {
        m = 2.0 * m;
        e = e - 1;
    }
    const double k = (double)e;
    const double f = m - 1.0; /* exact */
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 0x1.8618618618618p-4; /* 2/21 */
    p = fma(p, z, 0x1.af286bca1af28p-4); /* 2/19 */
    p = fma(p, z, 0x1.e1e1e1e1e1e1ep-4); /* 2/17 */
    p = fma(p, z, 0x1.1111111111111p-3); /* 2/15 */
    p = fma(p, z, 0x1.3b13b13b13b14p-3); /* 2/13 */
    p = fma(p, z, 0x1.745d1745d1746p-3); /* 2/11 */
    p = fma(p, z, 0x1.c71c71c71c71cp-3); /* 2/9  */
    p = fma(p, z, 0x1.2492492492492p-2); /* 2/7  */
    p = fma(p, z, 0x1.999999999999ap-2); /* 2/5  */
    p = fma(p, z, 0x1.5555555555555p-1); /* 2/3  */
    const double lm = fma(s * z, p, 2.0 * s);
    return k * SUN_LN2_HI + fma(k, SUN_LN2_LO, lm);
}

/* e^x; overflows to +inf above 709.78 and returns 0 below -745.13 */
static inline SUN_LIBM_HD double sun_exp(double x) {
    SUN_LIBM_NOCONTRACT
    if (x != x) return x;
    if (x > 709.78) return INFINITY;
    if (x < -745.13) return 0.0;
    const double k = rint(x * SUN_INV_LN2);
    const double r = fma(-k, SUN_LN2_LO, fma(-k, SUN_LN2_HI, x));
    double p = 0x1.93974a8c07c9dp-37; /* 1/14! */
    p = fma(p, r, 0x1.6124613a86d09p-33); /* 1/13! */
    p = fma(p, r, 0x1.1eed8eff8d898p-29); /* 1/12! */
    p = fma(p, r, 0x1.ae64567f544e4p-26); /* 1/11! */
    p = fma(p, r, 0x1.27e4fb7789f5cp-22); /* 1/10! */
    p = fma(p, r, 0x1.71de3a556c734p-19); /* 1/9!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-16); /* 1/8!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-13); /* 1/7!  */
    p = fma(p, r, 0x1.6c16c16c16c17p-10); /* 1/6!  */
    p = fma(p, r, 0x1.1111111111111p-7);  /* 1/5!  */
    p = fma(p, r, 0x1.5555555555555p-5);  /* 1/4!  */
    p = fma(p, r, 0x1.5555555555555p-3);  /* 1/3!  */
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

/* sin and cos of theta, |theta| <= pi/4, then rotated by quad * pi/2 (quad in 0..3) */
static inline SUN_LIBM_HD void sun_sincos_reduced(double th, int quad, double *sn, double *cs) {
    SUN_LIBM_NOCONTRACT
    const double z = th * th;
    double a = -0x1.2f49b46814157p-57;     /* -1/19! */
    a = fma(a, z, 0x1.952c77030ad4ap-49);  /*  1/17! */
    a = fma(a, z, -0x1.ae7f3e733b81fp-41); /* -1/15! */
    a = fma(a, z, 0x1.6124613a86d09p-33);  /*  1/13! */
    a = fma(a, z, -0x1.ae64567f544e4p-26); /* -1/11! */
    a = fma(a, z, 0x1.71de3a556c734p-19);  /*  1/9!  */
    a = fma(a, z, -0x1.a01a01a01a01ap-13); /* -1/7!  */
    a = fma(a, z, 0x1.1111111111111p-7);   /*  1/5!  */
    a = fma(a, z, -0x1.5555555555555p-3);  /* -1/3!  */
    const double s = fma(th * z, a, th);
    double b = -0x1.6827863b97d97p-53;     /* -1/18! */
    b = fma(b, z, 0x1.ae7f3e733b81fp-45);  /*  1/16! */
    b = fma(b, z, -0x1.93974a8c07c9dp-37); /* -1/14! */
    b = fma(b, z, 0x1.1eed8eff8d898p-29);  /*  1/12! */
    b = fma(b, z, -0x1.27e4fb7789f5cp-22); /* -1/10! */
    b = fma(b, z, 0x1.a01a01a01a01ap-16);  /*  1/8!  */
    b = fma(b, z, -0x1.6c16c16c16c17p-10); /* -1/6!  */
    b = fma(b, z, 0x1.5555555555555p-5);   /*  1/4!  */
    b = fma(b, z, -0.5);                   /* -1/2!  */
    const double c = fma(z, b, 1.0);
    switch (quad) {
        case 0: *sn = s; *cs = c; break;
        case 1: *sn = c; *cs = -s; break;
        case 2: *sn = -s; *cs = -c; break;
        default: *sn = -c; *cs = s; break;
    }
}

/* *sn = sin(2 pi u), *cs = cos(2 pi u); NaN for a non-finite u */
static inline SUN_LIBM_HD void sun_sincos2pi(double u, double *sn, double *cs) {
    SUN_LIBM_NOCONTRACT
    if (!(u - u == 0.0)) {
        *sn = NAN;
        *cs = NAN;
        return;
    }
    const double v = 4.0 * u; /* exact */
    const double q = rint(v);
    const double t = v - q; /* exact, |t| <= 1/2 */
    const double th = fma(t, SUN_PIO2_HI, t * SUN_PIO2_LO);
    /* 2 pi u = q pi/2 + theta: the quadrant q mod 4 rotates (sin, cos) */
    sun_sincos_reduced(th, (int)(q - 4.0 * floor(0.25 * q)), sn, cs);
}

/* *sn = sin(th), *cs = cos(th) for |th| <= 2^20 (8-real phases lie in (-pi, pi]):
 * q = rint(2 th / pi), th - q pi/2 by two fused steps against the pi/2 pair */
static inline SUN_LIBM_HD void sun_sincos(double th, double *sn, double *cs) {
    SUN_LIBM_NOCONTRACT
    if (!(th - th == 0.0)) {
        *sn = NAN;
        *cs = NAN;
        return;
    }
    const double q = rint(th * SUN_TWO_OVER_PI);
    const double r = fma(-q, SUN_PIO2_LO, fma(-q, SUN_PIO2_HI, th));
    sun_sincos_reduced(r, (int)(q - 4.0 * floor(0.25 * q)), sn, cs);
}

/* atan t for t in [0, 1]: above tan(pi/8), atan t = pi/4 + atan((t-1)/(t+1)); the
 * Taylor series of atan z on |z| <= tan(pi/8) to z^41 (omitted term < 2^-58 |z|) */
static inline SUN_LIBM_HD double sun_atan_unit(double t) {
    SUN_LIBM_NOCONTRACT
    double hi = 0.0, lo = 0.0, z = t;
    if (t > SUN_TAN_PI8) {
        z = (t - 1.0) / (t + 1.0);
        hi = SUN_PIO4_HI;
        lo = SUN_PIO4_LO;
    }
    const double w = z * z;
    double p = 0x1.8f9c18f9c18fap-6;       /*  1/41 */
    p = fma(p, w, -0x1.a41a41a41a41ap-6);  /* -1/39 */
    p = fma(p, w, 0x1.bacf914c1bad0p-6);   /*  1/37 */
    p = fma(p, w, -0x1.d41d41d41d41dp-6);  /* -1/35 */
    p = fma(p, w, 0x1.f07c1f07c1f08p-6);   /*  1/33 */
    p = fma(p, w, -0x1.0842108421084p-5);  /* -1/31 */
    p = fma(p, w, 0x1.1a7b9611a7b96p-5);   /*  1/29 */
    p = fma(p, w, -0x1.2f684bda12f68p-5);  /* -1/27 */
    p = fma(p, w, 0x1.47ae147ae147bp-5);   /*  1/25 */
    p = fma(p, w, -0x1.642c8590b2164p-5);  /* -1/23 */
    p = fma(p, w, 0x1.8618618618618p-5);   /*  1/21 */
    p = fma(p, w, -0x1.af286bca1af28p-5);  /* -1/19 */
    p = fma(p, w, 0x1.e1e1e1e1e1e1ep-5);   /*  1/17 */
    p = fma(p, w, -0x1.1111111111111p-4);  /* -1/15 */
    p = fma(p, w, 0x1.3b13b13b13b14p-4);   /*  1/13 */
    p = fma(p, w, -0x1.745d1745d1746p-4);  /* -1/11 */
    p = fma(p, w, 0x1.c71c71c71c71cp-4);   /*  1/9  */
    p = fma(p, w, -0x1.2492492492492p-3);  /* -1/7  */
    p = fma(p, w, 0x1.999999999999ap-3);   /*  1/5  */
    p = fma(p, w, -0x1.5555555555555p-2);  /* -1/3  */
    return hi + (lo + fma(z * w, p, z));
}

/* atan2(y, x) in [-pi, pi], with C's conventions for zeros and NaN */
static inline SUN_LIBM_HD double sun_atan2(double y, double x) {
    SUN_LIBM_NOCONTRACT
    if (x != x || y != y) return x + y;
    const double ax = fabs(x), ay = fabs(y);
    if (ay == 0.0) return signbit(x) ? copysign(SUN_PI_HI, y) : y;
    if (ax == 0.0) return copysign(SUN_PIO2_HI, y);
    double t;
    const int swap = ay > ax;
    if (ax == INFINITY && ay == INFINITY)
        t = 1.0;
    else
        t = swap ? ax / ay : ay / ax;
    double a = sun_atan_unit(t);
    if (swap) a = (SUN_PIO2_HI - a) + SUN_PIO2_LO;
    if (signbit(x)) a = (SUN_PI_HI - a) + SUN_PI_LO;
    return copysign(a, y);
}

/* cos(2 pi u) */
static inline SUN_LIBM_HD double sun_cos2pi(double u) {
    double s, c;
    sun_sincos2pi(u, &s, &c);
    return c;
}

#endif /* SUN_LIBM_H */
