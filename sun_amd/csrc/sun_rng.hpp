// sun_rng.hpp -- counter-based RNG and the SU(2) subgroup sampler (host + device).
//
// The reference keeps one cuRAND MRG32k3a state per half-volume site
// (random.h:16-22, random.cu:32-136) and streams it through global memory on every
// heatbath launch (monte.cu:82,86).  Here every draw is a pure function of
// (seed, site, sweep, parity, mu, draw index) -- Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, SC'11) -- so there is no RNG state in HBM at all and any lattice can
// be replayed bit-for-bit from (seed, sweep).
#pragma once

#include "sun_libm.h"
#include "sun_math.hpp"

namespace sun {

enum : uint32_t {
    kTagHeatbath = 0x10000000u,
    kTagHotStart = 0x20000000u,
    kTagTest = 0x30000000u,
    kTagMultihit = 0x40000000u,
    kTagMultihitExt = 0x50000000u,
};

struct Philox {
    SUN_HD static void round10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                               uint32_t k0, uint32_t k1) {
#pragma unroll
        for (int r = 0; r < 10; r++) {
            uint64_t p0 = (uint64_t)0xD2511F53u * c0;
            uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
            uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
            uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
#ifdef __HIP_DEVICE_COMPILE__
            // three-input XOR in one instruction (gfx950 v_bitop3_b32, truth table 0x96)
            uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
            uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
#else
            uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
#endif
            c0 = n0;
            c1 = lo1;
            c2 = n2;
            c3 = lo0;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
    }
};

// Uniform in (0,1] from one 32-bit Philox word: (w + 1) * 2^-32 -- 32 bits of
// randomness per draw, as cuRAND's MRG32k3a uniforms (random.h:16-22).  The double
// is exact; the float is the double rounded to 24 bits (at most one ulp apart).
SUN_HD double u01(uint32_t w) { return fma((double)w, 0x1p-32, 0x1p-32); }
SUN_HD float u01f(uint32_t w) { return fmaf((float)w, 0x1p-32f, 0x1p-32f); }

// Stream of uniforms for one (key, c0, c2, c3): draw k is word k%4 of
// Philox4x32-10((c0, k/4, c2, c3), key), so one Philox call serves four draws.
struct Stream {
    uint32_t k0, k1, c0, c2, c3, j;
    uint32_t w[4];
    int nw;

    SUN_HD Stream(uint64_t seed, uint32_t c0_, uint32_t c2_, uint32_t c3_)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), c0(c0_), c2(c2_), c3(c3_), j(0), w{0, 0, 0, 0},
          nw(0) {}

    SUN_HD uint32_t word() {
        if (nw == 0) {
            uint32_t a = c0, b = j, c = c2, d = c3;
            Philox::round10(a, b, c, d, k0, k1);
            w[0] = a;
            w[1] = b;
            w[2] = c;
            w[3] = d;
            j++;
            nw = 4;
        }
        const int i = 4 - nw;
        nw--;
        return i == 0 ? w[0] : i == 1 ? w[1] : i == 2 ? w[2] : w[3];
    }
    SUN_HD double next() { return u01(word()); }
    SUN_HD float next_f() { return u01f(word()); }
};

// Uniforms of the heatbath's SU(2) subgroup updates.  Every subgroup (block b of a
// link update) has counters of its own: attempt k of the sampler (k = 0 the first
// Kennedy-Pendleton attempt; k >= 1 its retries, or the Creutz tries after a rejected
// first attempt) reads the four words of Philox((c0, 32 b + k, c2, c3), seed), and the
// two final draws (a3 and the azimuth) words 0 and 1 of Philox((c0, 32 b + 31, c2, c3)).
// All counters are fixed before the sampler runs, so the first attempt's words and
// the final words of every block are computed with static register indexing (no
// per-draw buffer of a sequential stream to index at run time) and only retries
// compute on demand: the FP32 SU(3) heatbath launch at 32^4 went from 48.0 to
// 42.8 us.  A link update of N colours uses N(N-1)/2 blocks; 20 retries at most.
constexpr uint32_t kBlockCounters = 32;
constexpr uint32_t kFinalCounter = 31;

struct DrawKey {
    uint32_t k0, k1, c0, c2, c3;
    SUN_HD DrawKey(uint64_t seed, uint32_t c0_, uint32_t c2_, uint32_t c3_)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), c0(c0_), c2(c2_), c3(c3_) {}
    // the four words of Philox((c0, c1, c2, c3), key)
    SUN_HD void words(uint32_t c1, uint32_t w[4]) const {
        uint32_t a = c0, b = c1, c = c2, d = c3;
        Philox::round10(a, b, c, d, k0, k1);
        w[0] = a;
        w[1] = b;
        w[2] = c;
        w[3] = d;
    }
};

// the first attempt's words and the two final words of block b
struct BlockDraws {
    uint32_t first[4];
    uint32_t fin[2];
    SUN_HD BlockDraws(const DrawKey& key, int block) {
        key.words((uint32_t)block * kBlockCounters, first);
        uint32_t t[4];
        key.words((uint32_t)block * kBlockCounters + kFinalCounter, t);
        fin[0] = t[0];
        fin[1] = t[1];
    }
};

// Transcendentals of the FP32 sampler.  On the device the native gfx950
// instructions: v_log_f32 (log2), v_cos_f32 / v_sin_f32 (of 2 pi x: the argument is
// in revolutions), v_sqrt_f32, v_rsq_f32, v_rcp_f32.  Measured against libm over the uniforms
// (0, 1]: |cos, sin error| <= 1.3e-7, log2 1.7e-6 absolute at 2^-32 (5e-8 relative),
// sqrt, rsq and rcp <= 1e-7 relative -- the FP32 rounding level of the link they
// update.
SUN_HD float f32_log(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_logf(x) * 0.693147180559945f;
#else
    return logf(x);
#endif
}
SUN_HD float f32_cos2pi(float u) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_cosf(u);
#else
    return cosf((float)kTwoPi * u);
#endif
}
SUN_HD void f32_sincos2pi(float u, float& s, float& c) {
#ifdef __HIP_DEVICE_COMPILE__
    s = __builtin_amdgcn_sinf(u);
    c = __builtin_amdgcn_cosf(u);
#else
    s = sinf((float)kTwoPi * u);
    c = cosf((float)kTwoPi * u);
#endif
}
SUN_HD float f32_sqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
SUN_HD float f32_rsq(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.f / sqrtf(x);
#endif
}
SUN_HD float f32_rcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.f / x;
#endif
}

// SU(2) element a = (a0,a1,a2,a3) with a0 ~ sqrt(1-a0^2) exp(al*a0): the MILC
// Kennedy-Pendleton generator with the Creutz fallback for al <= 2, restating
// device_PHB_OVR.h:162-223 (including the 1e-10 guard and the 20-try caps), with the
// uniforms of block `block` (bd: its first attempt and final words; retries from
// key).  The reference's double literals make this arithmetic double even for
// Real=float; here it is double for every storage precision but the FP32 SU(3) path.
// Its log, exp and sin/cos of 2 pi u are sun_libm.h's, which give the same bits on
// the host and the device.
SUN_HD void generate_su2_milc(double al, const DrawKey& key, int block, const BlockDraws& bd, double a[4]) {
    double xr1, xr2, xr3, xr4, d, r;
    double a0 = 0.0;
    xr1 = u01(bd.first[0]);
    xr1 = sun_log(xr1 + 1.e-10);
    xr2 = u01(bd.first[1]);
    xr2 = sun_log(xr2 + 1.e-10);
    xr3 = u01(bd.first[2]);
    xr4 = u01(bd.first[3]);
    xr3 = sun_cos2pi(xr3);
    d = -(xr2 + xr1 * xr3 * xr3) / al;
    bool accepted = (1.00 - 0.5 * d) > xr4 * xr4;
    const uint32_t cbase = (uint32_t)block * kBlockCounters;
    if (!accepted && al > 2.0) {
        for (int k = 1; k <= 20; k++) {
            uint32_t w[4];
            key.words(cbase + (uint32_t)k, w);
            xr1 = u01(w[0]);
            xr1 = sun_log(xr1 + 1.e-10);
            xr2 = u01(w[1]);
            xr2 = sun_log(xr2 + 1.e-10);
            xr3 = u01(w[2]);
            xr4 = u01(w[3]);
            xr3 = sun_cos2pi(xr3);
            d = -(xr2 + xr1 * xr3 * xr3) / al;
            if ((1.00 - 0.5 * d) > xr4 * xr4) break;
        }
    }
    if (!accepted && al <= 2.0) {
        xr3 = sun_exp(-2.0 * al);
        xr4 = 1.0 - xr3;
        for (int k = 1; k <= 20; k++) {
            uint32_t w[4];
            key.words(cbase + (uint32_t)k, w);
            xr1 = u01(w[0]);
            xr2 = u01(w[1]);
            r = xr3 + xr4 * xr1;
            a0 = 1.00 + sun_log(r) / al;
            if ((1.0 - a0 * a0) > xr2 * xr2) break;
        }
        d = 1.0 - a0;
    }
    a[0] = 1.0 - d;
    xr3 = 1.0 - a[0] * a[0];
    xr3 = fabs(xr3);
    r = sqrt(xr3);
    a[3] = (2.0 * u01(bd.fin[0]) - 1.0) * r;
    xr1 = xr3 - a[3] * a[3];
    xr1 = fabs(xr1);
    xr1 = sqrt(xr1);
    double sn, cs;
    sun_sincos2pi(u01(bd.fin[1]), &sn, &cs);
    a[1] = xr1 * cs;
    a[2] = xr1 * sn;
}

// The same generator evaluated in FP32 (uniforms rounded to float), used by the
// FP32-storage engine: the link it updates is FP32, so the sampler's precision
// matches it; its log/cos/sin/sqrt and the division by al are the native FP32
// instructions above (the library routines cost ~120 more instructions per call).
SUN_HD void generate_su2_milc_f32(float al, const DrawKey& key, int block, const BlockDraws& bd, float a[4]) {
    float xr1, xr2, xr3, xr4, d, r;
    const float inv_al = f32_rcp(al);
    xr1 = u01f(bd.first[0]);
    xr1 = f32_log(xr1 + 1.e-10f);
    xr2 = u01f(bd.first[1]);
    xr2 = f32_log(xr2 + 1.e-10f);
    xr3 = u01f(bd.first[2]);
    xr4 = u01f(bd.first[3]);
    xr3 = f32_cos2pi(xr3);
    d = -(xr2 + xr1 * xr3 * xr3) * inv_al;
    bool accepted = (1.00f - 0.5f * d) > xr4 * xr4;
    const uint32_t cbase = (uint32_t)block * kBlockCounters;
    if (!accepted && al > 2.f) {
        for (int k = 1; k <= 20; k++) {
            uint32_t w[4];
            key.words(cbase + (uint32_t)k, w);
            xr1 = u01f(w[0]);
            xr1 = f32_log(xr1 + 1.e-10f);
            xr2 = u01f(w[1]);
            xr2 = f32_log(xr2 + 1.e-10f);
            xr3 = u01f(w[2]);
            xr4 = u01f(w[3]);
            xr3 = f32_cos2pi(xr3);
            d = -(xr2 + xr1 * xr3 * xr3) * inv_al;
            if ((1.00f - 0.5f * d) > xr4 * xr4) break;
        }
    }
    if (!accepted && al <= 2.f) {
        xr3 = expf(-2.f * al);
        xr4 = 1.f - xr3;
        for (int k = 1; k <= 20; k++) {
            uint32_t w[4];
            key.words(cbase + (uint32_t)k, w);
            xr1 = u01f(w[0]);
            xr2 = u01f(w[1]);
            r = xr3 + xr4 * xr1;
            d = -f32_log(r) * inv_al;  // 1 - a0 for a0 = 1 + log(r) / al
            if (d * (2.f - d) > xr2 * xr2) break;
        }
    }
    // The reference's 1 - a0^2 and 1 - a0^2 - a3^2 (in double there) are formed here
    // without differences of nearly equal FP32 numbers: 1 - a0^2 = d (2 - d) (at
    // a0 ~ 1, 1 - a0*a0 in FP32 loses every significant bit: errors of 6e-5 in a link
    // at 32^4), and with a3 = (2u - 1) r, 1 - a0^2 - a3^2 = r^2 * 4u(1 - u), where u and
    // 1 - u are each converted from the Philox word.
    a[0] = 1.f - d;
    xr3 = fabsf(d * (2.f - d));
    r = f32_sqrt(xr3);
    const uint32_t w = bd.fin[0];
    const float u = u01f(w);                  // (w + 1) 2^-32
    const float v = (float)(~w) * 0x1p-32f;   // 1 - u = (2^32 - 1 - w) 2^-32
    a[3] = (u - v) * r;
    xr1 = f32_sqrt(xr3 * (4.f * u * v));
    float sn, cs;
    f32_sincos2pi(u01f(bd.fin[1]), sn, cs);
    a[1] = xr1 * cs;
    a[2] = xr1 * sn;
}

}  // namespace sun
