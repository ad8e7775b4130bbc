/* Pins sun_libm.h (the portable FP64 log, exp and sin/cos of 2 pi u shared by the
 * device sampler and the host oracle) against glibc: log and exp against glibc's
 * double routines in ulps, sin and cos of 2 pi u against long-double sinl/cosl of
 * the product (so the reference value carries no rounding of 2 pi u) in units of
 * 2^-53 (also of phases theta in [-pi, pi]), atan2 against glibc in ulps over all
 * quadrants, plus exact special values.  Prints "max_ulp_log max_ulp_exp max_err_sin
 * max_err_cos max_ulp_atan2 nfail" on one line; nfail counts failed exact checks. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sun_libm.h"

static int64_t ordered(double x) {
    int64_t i;
    memcpy(&i, &x, 8);
    return i < 0 ? (int64_t)0x8000000000000000LL - i : i;
}
static double ulps(double a, double b) { return (double)llabs(ordered(a) - ordered(b)); }

static uint64_t state = 0x9E3779B97F4A7C15ull;
static double uniform(void) { /* [0, 1) */
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return (double)(state >> 11) * 0x1p-53;
}

int main(void) {
    const long double two_pi = 6.283185307179586476925286766559005768L;
    double mlog = 0, mexp = 0, msin = 0, mcos = 0, matan = 0;
    int nfail = 0;
    for (int i = 0; i < 1000000; i++) {
        /* the sampler's arguments: u + 1e-10 and Philox uniforms (w+1) 2^-32 */
        const double u = ldexp(floor(uniform() * 4294967296.0) + 1.0, -32);
        double d = ulps(sun_log(u + 1.e-10), log(u + 1.e-10));
        if (d > mlog) mlog = d;
        /* the whole exponent range */
        const double y = ldexp(0.5 + 0.5 * uniform(), (int)(uniform() * 2040.0) - 1020);
        d = ulps(sun_log(y), log(y));
        if (d > mlog) mlog = d;
        const double e1 = -4.0 * uniform(), e2 = 1400.0 * uniform() - 700.0;
        d = ulps(sun_exp(e1), exp(e1));
        if (d > mexp) mexp = d;
        d = ulps(sun_exp(e2), exp(e2));
        if (d > mexp) mexp = d;
        double s, c;
        sun_sincos2pi(u, &s, &c);
        const double rs = (double)sinl(two_pi * (long double)u), rc = (double)cosl(two_pi * (long double)u);
        d = fabs(s - rs) / 0x1p-53;
        if (d > msin) msin = d;
        d = fabs(c - rc) / 0x1p-53;
        if (d > mcos) mcos = d;
        if (sun_cos2pi(u) != c) nfail++;
        /* phases in [-pi, pi] (the 8-real codec) */
        const double th = (2.0 * uniform() - 1.0) * 3.141592653589793;
        sun_sincos(th, &s, &c);
        d = fabs(s - (double)sinl((long double)th)) / 0x1p-53;
        if (d > msin) msin = d;
        d = fabs(c - (double)cosl((long double)th)) / 0x1p-53;
        if (d > mcos) mcos = d;
        /* atan2 over all quadrants and magnitude ratios */
        const double ya = (2.0 * uniform() - 1.0) * ldexp(1.0, (int)(uniform() * 40.0) - 20);
        const double xa = 2.0 * uniform() - 1.0;
        d = ulps(sun_atan2(ya, xa), atan2(ya, xa));
        if (d > matan) matan = d;
        d = ulps(sun_atan2(xa, ya), atan2(xa, ya));
        if (d > matan) matan = d;
    }
    /* exact values */
    double s, c;
    sun_sincos2pi(1.0, &s, &c);
    nfail += !(s == 0.0 && c == 1.0);
    sun_sincos2pi(0.25, &s, &c);
    nfail += !(s == 1.0 && c == 0.0);
    sun_sincos2pi(0.5, &s, &c);
    nfail += !(s == 0.0 && c == -1.0);
    sun_sincos2pi(0.75, &s, &c);
    nfail += !(s == -1.0 && c == 0.0);
    sun_sincos2pi(0.125, &s, &c);
    nfail += !(fabs(s - sqrt(0.5)) <= 0x1p-53 && fabs(c - sqrt(0.5)) <= 0x1p-53);
    nfail += !(sun_log(1.0) == 0.0);
    nfail += !(sun_log(2.0) == log(2.0) && sun_log(0.5) == -log(2.0));
    nfail += !(sun_log(0.0) == -INFINITY && isnan(sun_log(-1.0)) && isnan(sun_log(NAN)));
    nfail += !(sun_log(INFINITY) == INFINITY);
    nfail += !(sun_exp(0.0) == 1.0 && sun_exp(-800.0) == 0.0 && sun_exp(800.0) == INFINITY);
    for (int k = -1000; k <= 1000; k += 7) /* log 2^k = k ln 2 to 1 ulp */
        nfail += !(ulps(sun_log(ldexp(1.0, k)), (double)(k * 0.693147180559945309417232121458176568L)) <= 1);
    sun_sincos2pi(NAN, &s, &c);
    nfail += !(isnan(s) && isnan(c));
    /* atan2 at the axes, zeros (C's sign conventions) and the diagonals */
    const double zx[] = {1.0, -1.0, 0.0, -0.0, 2.5, -2.5};
    const double zy[] = {0.0, -0.0, 1.0, -1.0, 2.5, -2.5};
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            const double a = sun_atan2(zy[i], zx[j]), b = atan2(zy[i], zx[j]);
            nfail += !(ulps(a, b) <= 1 && signbit(a) == signbit(b));
        }
    nfail += !isnan(sun_atan2(NAN, 1.0)) + !isnan(sun_atan2(1.0, NAN));
    sun_sincos(0.0, &s, &c);
    nfail += !(s == 0.0 && c == 1.0);
    printf("%g %g %g %g %g %d\n", mlog, mexp, msin, mcos, matan, nfail);
    return 0;
}
