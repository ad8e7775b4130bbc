// sun_field.hpp -- lattice geometry and the HBM link layout (host + device).
//
// Geometry restates the reference's even/odd site numbering (index.h:133-202):
//   eo(x) = parity(x)*V/2 + lex(x)/2,  lex(x) = x0 + X0*(x1 + X1*(x2 + X2*x3)),
// and the link of direction mu at site x is link number mu*V + eo(x)
// (monte.cu:58-61), so a link index means the same thing on both sides.
//
// Layout in HBM ("plane" SoA).  The reference stores one complex element per plane
// (element-major SoA, device_load_save.h:76-120).  Here a link's stored reals
// (18/12/8 for SU(3), 2N^2 otherwise) are cut into 16-byte records and record c of
// every link lives in plane c: plane c is an array of 4V records of 16 bytes, so a
// wavefront reads 64 consecutive links of one plane as 1 KiB of contiguous HBM with
// one dwordx4 load per lane.  An FP32 18-real SU(3) link is 72 bytes: four 16-byte
// planes plus one 8-byte tail plane.  Stored reals, in order:
//   full  : e00.re e00.im e01.re e01.im ... (row-major)
//   12    : rows 0 and 1                  (row 2 = conj(r0 x r1) on load)
//   8     : e01 e02 e10 arg(e00) arg(e20) (device_load_save.h:112-120)
#pragma once

#include "sun_math.hpp"

namespace sun {

// n / d for 0 <= n < 2^31 by multiply-high and shift (round-up method,
// Granlund & Montgomery 1994).
struct FastDiv {
    uint32_t d, m, s;
    static FastDiv make(uint32_t d_) {
        FastDiv f;
        f.d = d_;
        uint32_t l = 0;
        while ((1ull << l) < d_) l++;
        f.s = l;
        f.m = (uint32_t)((((1ull << 32) * ((1ull << l) - d_)) / d_) + 1);
        return f;
    }
    SUN_HD uint32_t div(uint32_t n) const {
        uint32_t hi = (uint32_t)(((uint64_t)n * m) >> 32);
        return (hi + n) >> s;
    }
};

struct Geom {
    int X[4];
    int V, half;
    int str[4];          // lexicographic strides 1, X0, X0*X1, X0*X1*X2
    FastDiv dxh, dx1, dx2;  // division by X0/2, X1, X2
    // brick (log2 rows along x1, x2, x3) of the 256-site reduction workgroups, or
    // rb1 = -1 when no brick of 256 sites tiles the lattice (linear order then)
    int rb1, rb2, rb3;
};

// The most cubic brick (least sum of squared log2 extents) of 256 sites of one parity: rows of X0/2 sites, 2^k rows
// split over x1, x2, x3 so that each extent divides the lattice's.
inline void reduction_brick(Geom& g) {
    g.rb1 = -1;
    g.rb2 = g.rb3 = 0;
    const int L = g.X[0] / 2;
    if (L <= 0 || 256 % L) return;
    const int rows = 256 / L;
    if (rows & (rows - 1)) return;
    int k = 0;
    while ((1 << k) < rows) k++;
    int best = 99;
    for (int b1 = k; b1 >= 0; b1--)
        for (int b2 = k - b1; b2 >= 0; b2--) {
            const int b3 = k - b1 - b2;
            if ((g.X[1] % (1 << b1)) || (g.X[2] % (1 << b2)) || (g.X[3] % (1 << b3))) continue;
            const int m = b1 * b1 + b2 * b2 + b3 * b3;
            if (m < best) {
                best = m;
                g.rb1 = b1;
                g.rb2 = b2;
                g.rb3 = b3;
            }
        }
}

inline Geom make_geom(const int dims[4]) {
    Geom g;
    for (int i = 0; i < 4; i++) g.X[i] = dims[i];
    g.V = dims[0] * dims[1] * dims[2] * dims[3];
    g.half = g.V / 2;
    g.str[0] = 1;
    g.str[1] = dims[0];
    g.str[2] = dims[0] * dims[1];
    g.str[3] = dims[0] * dims[1] * dims[2];
    g.dxh = FastDiv::make((uint32_t)(dims[0] / 2));
    g.dx1 = FastDiv::make((uint32_t)dims[1]);
    g.dx2 = FastDiv::make((uint32_t)dims[2]);
    reduction_brick(g);
    return g;
}

// Index_4D_EO (index.h:133-141): coordinates of half-lattice site id of a parity;
// returns lex(x).
SUN_HD int site_coords(const Geom& g, int id, int parity, int x[4]) {
    int za = (int)g.dxh.div((uint32_t)id);
    int zb = (int)g.dx1.div((uint32_t)za);
    x[1] = za - zb * g.X[1];
    x[3] = (int)g.dx2.div((uint32_t)zb);
    x[2] = zb - x[3] * g.X[2];
    int xodd = (x[1] + x[2] + x[3] + parity) & 1;
    x[0] = (2 * id + xodd) - za * g.X[0];
    return 2 * id + xodd;
}

// lex of x + d*e_mu (d = +1 or -1) with periodic wrap; x is updated.
SUN_HD int lex_shift(const Geom& g, int lex, int x[4], int mu, int d) {
    const int s = g.str[mu];
    if (d > 0) {
        if (x[mu] == g.X[mu] - 1) { x[mu] = 0; return lex - (g.X[mu] - 1) * s; }
        x[mu] += 1;
        return lex + s;
    }
    if (x[mu] == 0) { x[mu] = g.X[mu] - 1; return lex + (g.X[mu] - 1) * s; }
    x[mu] -= 1;
    return lex - s;
}

SUN_HD int eo_of_lex(const Geom& g, int lex, int parity) { return (lex >> 1) + parity * g.half; }

// ---- storage -----------------------------------------------------------------
enum : int { kStore18 = 18, kStore12 = 12, kStore8 = 8 };

constexpr int kMaxPlanes = 16;

struct Field {
    char* p[kMaxPlanes];
};

template <typename R, int N, int STORE>
struct Layout {
    static constexpr int NREAL = (STORE == kStore18) ? 2 * N * N : STORE;
    static constexpr int BYTES = NREAL * (int)sizeof(R);
    static constexpr int NFULL = BYTES / 16;
    static constexpr int TAIL = BYTES % 16;  // 0 or 8
    static constexpr int NPLANES = NFULL + (TAIL ? 1 : 0);
    static constexpr int PER16 = 16 / (int)sizeof(R);
    static_assert(STORE == kStore18 || N == 3, "12/8-real storage is SU(3) only");
    static_assert(NPLANES <= kMaxPlanes, "too many planes");
};

// Byte offsets within a plane are 32-bit (a plane holds 4V records of 16 bytes and
// 4V < 2^28 for every supported volume), so each access is a uniform 64-bit plane
// base in SGPRs plus one 32-bit VGPR offset.
template <typename R, int N, int STORE>
SUN_HD void load_reals(const Field& f, uint32_t l, R* r) {
    using L = Layout<R, N, STORE>;
    const uint32_t off = l * 16u;
#pragma unroll
    for (int c = 0; c < L::NFULL; c++) {
        if constexpr (sizeof(R) == 4) {
            float4 v = *reinterpret_cast<const float4*>(f.p[c] + off);
            r[4 * c + 0] = v.x; r[4 * c + 1] = v.y; r[4 * c + 2] = v.z; r[4 * c + 3] = v.w;
        } else {
            double2 v = *reinterpret_cast<const double2*>(f.p[c] + off);
            r[2 * c + 0] = v.x; r[2 * c + 1] = v.y;
        }
    }
    if constexpr (L::TAIL != 0) {
        float2 v = *reinterpret_cast<const float2*>(f.p[L::NFULL] + l * 8u);
        r[L::NREAL - 2] = v.x;
        r[L::NREAL - 1] = v.y;
    }
}

template <typename R, int N, int STORE>
SUN_HD void store_reals(const Field& f, uint32_t l, const R* r) {
    using L = Layout<R, N, STORE>;
    const uint32_t off = l * 16u;
#pragma unroll
    for (int c = 0; c < L::NFULL; c++) {
        if constexpr (sizeof(R) == 4) {
            float4 v;
            v.x = r[4 * c + 0]; v.y = r[4 * c + 1]; v.z = r[4 * c + 2]; v.w = r[4 * c + 3];
            *reinterpret_cast<float4*>(f.p[c] + off) = v;
        } else {
            double2 v;
            v.x = r[2 * c + 0]; v.y = r[2 * c + 1];
            *reinterpret_cast<double2*>(f.p[c] + off) = v;
        }
    }
    if constexpr (L::TAIL != 0) {
        float2 v;
        v.x = r[L::NREAL - 2];
        v.y = r[L::NREAL - 1];
        *reinterpret_cast<float2*>(f.p[L::NFULL] + l * 8u) = v;
    }
}

// sin and cos of a stored phase th in (-pi, pi] (atan2 output): sun_libm.h's, the
// same bits on the host and the device, with a short reduction for this range (the
// general library sin/cos carry a Payne-Hanek path that dominates the registers of a
// kernel decoding 19 links).
SUN_HD void phase(double th, double& s, double& c) { sun_sincos(th, &s, &c); }
// FP32 on the device: the native v_sin_f32 / v_cos_f32 of th / (2 pi) in revolutions
// (within 1.3e-7 of libm on this range, below the FP32 rounding of the stored reals)
SUN_HD void phase(float th, float& s, float& c) {
#ifdef __HIP_DEVICE_COMPILE__
    const float rev = th * 0.159154943091895335769f;
    s = __builtin_amdgcn_sinf(rev);
    c = __builtin_amdgcn_cosf(rev);
#else
    s = sinf(th);
    c = cosf(th);
#endif
}

// the stored phase arg z of the 8-real record (device_load_save.h:112-120): FP64
// through sun_libm.h (host and device agree bitwise), FP32 the library atan2f
SUN_HD double phase_of(cplx<double> z) { return arg_of(z); }
SUN_HD float phase_of(cplx<float> z) { return arg_of(z); }

// square root and reciprocal of the 8-real decode: FP64 exact IEEE; FP32 on the device
// the native v_sqrt_f32 / v_rcp_f32 (1 ulp), as cheap as the FP32 link they decode
SUN_HD double decode_sqrt(double x) { return sqrt(x); }
SUN_HD double decode_rcp(double x) { return 1.0 / x; }
SUN_HD float decode_sqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
SUN_HD float decode_rcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.f / x;
#endif
}

// reconstruct_12p_8p.h:38-82 (8-real SU(3) decode)
template <typename R>
SUN_HD void reconstruct8(Mat<R, 3>& m, R th0, R th1) {
    R p1, p2, diff;
    cplx<R> temp;
    p1 = abs2(m.e[0][1]) + abs2(m.e[0][2]);
    p2 = decode_rcp(p1);
    diff = R(1) - p1;
    p1 = decode_sqrt(diff >= R(0) ? diff : R(0));
    R c0, s0, c1, s1;
    phase(th0, s0, c0);
    phase(th1, s1, c1);
    m.e[0][0].re = p1 * c0;
    m.e[0][0].im = p1 * s0;
    diff = R(1) - abs2(m.e[0][0]) - abs2(m.e[1][0]);
    p1 = decode_sqrt(diff >= R(0) ? diff : R(0));
    m.e[2][0].re = p1 * c1;
    m.e[2][0].im = p1 * s1;
    m.e[1][1] = conj(m.e[2][0] * m.e[0][2]);
    m.e[2][2] = conj(m.e[0][0]) * m.e[1][0];
    m.e[1][1] = m.e[1][1] + m.e[2][2] * m.e[0][1];
    m.e[1][1] = scale(m.e[1][1], -p2);
    m.e[1][2] = conj(m.e[2][0] * m.e[0][1]);
    m.e[1][2] = m.e[1][2] - m.e[2][2] * m.e[0][2];
    m.e[1][2] = scale(m.e[1][2], p2);
    m.e[2][1] = conj(m.e[1][0] * m.e[0][2]);
    temp = conj(m.e[0][0]) * m.e[2][0];
    m.e[2][1] = m.e[2][1] - temp * m.e[0][1];
    m.e[2][1] = scale(m.e[2][1], p2);
    m.e[2][2] = conj(m.e[1][0] * m.e[0][1]);
    m.e[2][2] = m.e[2][2] + temp * m.e[0][2];
    m.e[2][2] = scale(m.e[2][2], -p2);
}

// FP32: the same steps with every multiply-add an explicit fmaf.  Left to the
// compiler, the products and sums above are fused into FMAs differently depending on
// the code the decode is inlined into, and the kernels that decode the same link
// (plain and software-pipelined staples) would round it differently.  The clamps of
// the square-root arguments are fmaxf (one v_max_f32 instead of a compare and a
// select): the same value, as diff is never -0 and a NaN clamps to 0 either way.
namespace r8f {
SUN_HD cplx<float> mul(cplx<float> a, cplx<float> b) {  // a*b
    return {fmaf(-a.im, b.im, a.re * b.re), fmaf(a.im, b.re, a.re * b.im)};
}
SUN_HD cplx<float> conj_mul(cplx<float> a, cplx<float> b) {  // conj(a*b)
    return {fmaf(-a.im, b.im, a.re * b.re), -fmaf(a.im, b.re, a.re * b.im)};
}
SUN_HD cplx<float> mul_ca(cplx<float> a, cplx<float> b) {  // conj(a)*b
    return {fmaf(a.im, b.im, a.re * b.re), fmaf(-a.im, b.re, a.re * b.im)};
}
SUN_HD cplx<float> add_mul(cplx<float> x, cplx<float> a, cplx<float> b) {  // x + a*b
    return {fmaf(-a.im, b.im, fmaf(a.re, b.re, x.re)), fmaf(a.im, b.re, fmaf(a.re, b.im, x.im))};
}
SUN_HD cplx<float> sub_mul(cplx<float> x, cplx<float> a, cplx<float> b) {  // x - a*b
    return {fmaf(a.im, b.im, fmaf(-a.re, b.re, x.re)), fmaf(-a.im, b.re, fmaf(-a.re, b.im, x.im))};
}
SUN_HD cplx<float> scaled(cplx<float> a, float s) { return {a.re * s, a.im * s}; }
}  // namespace r8f

SUN_HD void reconstruct8(Mat<float, 3>& m, float th0, float th1) {
    using namespace r8f;
    float p1 = m.e[0][1].re * m.e[0][1].re;
    p1 = fmaf(m.e[0][1].im, m.e[0][1].im, p1);
    p1 = fmaf(m.e[0][2].re, m.e[0][2].re, p1);
    p1 = fmaf(m.e[0][2].im, m.e[0][2].im, p1);
    const float p2 = decode_rcp(p1);
    float diff = 1.f - p1;
    p1 = decode_sqrt(fmaxf(diff, 0.f));
    float c0, s0, c1, s1;
    phase(th0, s0, c0);
    phase(th1, s1, c1);
    m.e[0][0].re = p1 * c0;
    m.e[0][0].im = p1 * s0;
    diff = fmaf(-m.e[0][0].re, m.e[0][0].re, 1.f);
    diff = fmaf(-m.e[0][0].im, m.e[0][0].im, diff);
    diff = fmaf(-m.e[1][0].re, m.e[1][0].re, diff);
    diff = fmaf(-m.e[1][0].im, m.e[1][0].im, diff);
    p1 = decode_sqrt(fmaxf(diff, 0.f));
    m.e[2][0].re = p1 * c1;
    m.e[2][0].im = p1 * s1;
    const cplx<float> t22 = mul_ca(m.e[0][0], m.e[1][0]);
    m.e[1][1] = scaled(add_mul(conj_mul(m.e[2][0], m.e[0][2]), t22, m.e[0][1]), -p2);
    m.e[1][2] = scaled(sub_mul(conj_mul(m.e[2][0], m.e[0][1]), t22, m.e[0][2]), p2);
    const cplx<float> temp = mul_ca(m.e[0][0], m.e[2][0]);
    m.e[2][1] = scaled(sub_mul(conj_mul(m.e[1][0], m.e[0][2]), temp, m.e[0][1]), p2);
    m.e[2][2] = scaled(add_mul(conj_mul(m.e[1][0], m.e[0][1]), temp, m.e[0][2]), -p2);
}

template <typename R, int N, int STORE>
SUN_HD Mat<R, N> unpack(const R* r) {
    Mat<R, N> m;
    if constexpr (STORE == kStore18) {
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) m.e[i][j] = mk<R>(r[2 * (i * N + j)], r[2 * (i * N + j) + 1]);
    } else if constexpr (STORE == kStore12) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            m.e[0][j] = mk<R>(r[2 * j], r[2 * j + 1]);
            m.e[1][j] = mk<R>(r[6 + 2 * j], r[6 + 2 * j + 1]);
        }
        reconstruct_row2(m);
    } else {
        m.e[0][1] = mk<R>(r[0], r[1]);
        m.e[0][2] = mk<R>(r[2], r[3]);
        m.e[1][0] = mk<R>(r[4], r[5]);
        reconstruct8(m, r[6], r[7]);
    }
    return m;
}

template <typename R, int N, int STORE>
SUN_HD void pack(const Mat<R, N>& m, R* r) {
    if constexpr (STORE == kStore18) {
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) {
                r[2 * (i * N + j)] = m.e[i][j].re;
                r[2 * (i * N + j) + 1] = m.e[i][j].im;
            }
    } else if constexpr (STORE == kStore12) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            r[2 * j] = m.e[0][j].re;
            r[2 * j + 1] = m.e[0][j].im;
            r[6 + 2 * j] = m.e[1][j].re;
            r[6 + 2 * j + 1] = m.e[1][j].im;
        }
    } else {
        r[0] = m.e[0][1].re; r[1] = m.e[0][1].im;
        r[2] = m.e[0][2].re; r[3] = m.e[0][2].im;
        r[4] = m.e[1][0].re; r[5] = m.e[1][0].im;
        r[6] = phase_of(m.e[0][0]);
        r[7] = phase_of(m.e[2][0]);
    }
}

template <typename R, int N, int STORE>
SUN_HD Mat<R, N> load_link(const Field& f, uint32_t l) {
    R r[Layout<R, N, STORE>::NREAL];
    load_reals<R, N, STORE>(f, l, r);
    return unpack<R, N, STORE>(r);
}

template <typename R, int N, int STORE>
SUN_HD void store_link(const Field& f, uint32_t l, const Mat<R, N>& m) {
    R r[Layout<R, N, STORE>::NREAL];
    pack<R, N, STORE>(m, r);
    store_reals<R, N, STORE>(f, l, r);
}

}  // namespace sun
