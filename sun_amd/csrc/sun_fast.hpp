// sun_fast.hpp -- arithmetic-lean SU(3) staple and link updates for FP32 storage.
//
// A (parity, mu) update moves 9 links of 48 bytes (12-real FP32) per site but
// spends ~3k VALU operations on the staple and the SU(2) updates, which on
// gfx950 without packed FP32 is as long as the HBM transfer.  This file removes
// the arithmetic that is not needed:
//   * complex multiply-accumulate as four chained FMAs (the plain form rounds the
//     product and then adds: six operations);
//   * products of SU(3) matrices are computed for rows 0 and 1 only; row 2 of an
//     SU(3) matrix is conj(row0 x row1), so a 12-real link is never expanded unless
//     all three of its rows enter a product;
//   * the overrelaxation applies r^2 once instead of r twice;
//   * the heatbath's SU(2) sampler runs in FP32 (generate_su2_milc_f32).
// Mathematically the same update as heatBathSUN / overrelaxationSUN
// (device_PHB_OVR.h:361-485, 493-608); rounding differs, which the FP32 tolerance
// (5e-5 per link after a sweep) covers.  The FP64 engine keeps the reference's
// operation order (sun_update.hpp).
#pragma once

#include "sun_field.hpp"
#include "sun_update.hpp"

namespace sun {
namespace fast {

using F3 = Mat<float, 3>;
using C = cplx<float>;

// c += a*b
SUN_HD void cmac(C& c, C a, C b) {
    c.re = fmaf(a.re, b.re, c.re);
    c.re = fmaf(-a.im, b.im, c.re);
    c.im = fmaf(a.re, b.im, c.im);
    c.im = fmaf(a.im, b.re, c.im);
}
// c += a*conj(b)
SUN_HD void cmac_cb(C& c, C a, C b) {
    c.re = fmaf(a.re, b.re, c.re);
    c.re = fmaf(a.im, b.im, c.re);
    c.im = fmaf(a.im, b.re, c.im);
    c.im = fmaf(-a.re, b.im, c.im);
}
// c += conj(a)*b
SUN_HD void cmac_ca(C& c, C a, C b) {
    c.re = fmaf(a.re, b.re, c.re);
    c.re = fmaf(a.im, b.im, c.re);
    c.im = fmaf(a.re, b.im, c.im);
    c.im = fmaf(-a.im, b.re, c.im);
}
SUN_HD C cmul(C a, C b) {
    C c;
    c.re = a.re * b.re;
    c.re = fmaf(-a.im, b.im, c.re);
    c.im = a.re * b.im;
    c.im = fmaf(a.im, b.re, c.im);
    return c;
}
SUN_HD C cmul_cb(C a, C b) {   // a*conj(b)
    C c;
    c.re = a.re * b.re;
    c.re = fmaf(a.im, b.im, c.re);
    c.im = a.im * b.re;
    c.im = fmaf(-a.re, b.im, c.im);
    return c;
}
SUN_HD C cmul_ca(C a, C b) {   // conj(a)*b
    C c;
    c.re = a.re * b.re;
    c.re = fmaf(a.im, b.im, c.re);
    c.im = a.re * b.im;
    c.im = fmaf(-a.im, b.re, c.im);
    return c;
}

// conj(a*b - c*d)
SUN_HD C cross_term(C a, C b, C c, C d) {
    C r;
    float re = a.re * b.re;
    re = fmaf(-a.im, b.im, re);
    re = fmaf(-c.re, d.re, re);
    re = fmaf(c.im, d.im, re);
    float im = a.re * b.im;
    im = fmaf(a.im, b.re, im);
    im = fmaf(-c.re, d.im, im);
    im = fmaf(-c.im, d.re, im);
    r.re = re;
    r.im = -im;
    return r;
}

// row 2 = conj(row0 x row1)
SUN_HD void complete_row2(F3& m) {
    m.e[2][0] = cross_term(m.e[0][1], m.e[1][2], m.e[0][2], m.e[1][1]);
    m.e[2][1] = cross_term(m.e[0][2], m.e[1][0], m.e[0][0], m.e[1][2]);
    m.e[2][2] = cross_term(m.e[0][0], m.e[1][1], m.e[0][1], m.e[1][0]);
}

// Links as loaded: rows 0,1 always valid; row 2 valid for 18- and 8-real storage,
// built on demand for 12-real storage.
template <int STORE>
SUN_HD F3 load_rows(const Field& f, uint32_t l) {
    float r[Layout<float, 3, STORE>::NREAL];
    load_reals<float, 3, STORE>(f, l, r);
    F3 m;
    if constexpr (STORE == kStore12) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            m.e[0][j] = mk<float>(r[2 * j], r[2 * j + 1]);
            m.e[1][j] = mk<float>(r[6 + 2 * j], r[6 + 2 * j + 1]);
        }
    } else {
        m = unpack<float, 3, STORE>(r);
    }
    return m;
}
template <int STORE>
SUN_HD void need_row2(F3& m) {
    if constexpr (STORE == kStore12) complete_row2(m);
}

// X(rows 0,1) = A(rows 0,1) * B            (B full)
SUN_HD void mul_2r(const F3& A, const F3& B, F3& X) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            C s = cmul(A.e[i][0], B.e[0][j]);
            cmac(s, A.e[i][1], B.e[1][j]);
            cmac(s, A.e[i][2], B.e[2][j]);
            X.e[i][j] = s;
        }
}
// X(rows 0,1) = A(rows 0,1) * B^+          (B full)
SUN_HD void mul_2r_bdag(const F3& A, const F3& B, F3& X) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            C s = cmul_cb(A.e[i][0], B.e[j][0]);
            cmac_cb(s, A.e[i][1], B.e[j][1]);
            cmac_cb(s, A.e[i][2], B.e[j][2]);
            X.e[i][j] = s;
        }
}
// X(rows 0,1) = A^+ * B                    (A full, B full)
SUN_HD void mul_2r_adag(const F3& A, const F3& B, F3& X) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            C s = cmul_ca(A.e[0][i], B.e[0][j]);
            cmac_ca(s, A.e[1][i], B.e[1][j]);
            cmac_ca(s, A.e[2][i], B.e[2][j]);
            X.e[i][j] = s;
        }
}

SUN_HD void add_full(F3& S, const F3& T) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S.e[i][j] = S.e[i][j] + T.e[i][j];
}

// staple sum (monte.cu:63-80) for a direction known only at run time (used by the
// heatbath, whose sampler needs the registers a specialised kernel would spend).
template <int STORE>
SUN_HD F3 staple_su3_rt(const Field& f, const Geom& g, int id, int parity, int mu, const int x[4], int lex) {
    const uint32_t V = (uint32_t)g.V;
    const int opar = parity ^ 1;
    const uint32_t self = (uint32_t)(id + parity * g.half);
    int xm[4] = {x[0], x[1], x[2], x[3]};
    const uint32_t eo_pmu = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xm, mu, +1), opar);
    F3 S;
    bool first = true;
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        int xn[4] = {x[0], x[1], x[2], x[3]};
        const uint32_t eo_pnu = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xn, nu, +1), opar);
        int xd[4] = {x[0], x[1], x[2], x[3]};
        const int lex_mnu = lex_shift(g, lex, xd, nu, -1);
        const uint32_t eo_mnu = (uint32_t)eo_of_lex(g, lex_mnu, opar);
        const uint32_t eo_mnu_pmu = (uint32_t)eo_of_lex(g, lex_shift(g, lex_mnu, xd, mu, +1), parity);
        F3 A = load_rows<STORE>(f, (uint32_t)nu * V + self);
        F3 B = load_rows<STORE>(f, (uint32_t)mu * V + eo_pnu);
        F3 Cc = load_rows<STORE>(f, (uint32_t)nu * V + eo_pmu);
        need_row2<STORE>(Cc);
        F3 W, T;
        mul_2r_bdag(B, Cc, W);
        complete_row2(W);
        mul_2r(A, W, T);
        complete_row2(T);
        if (first) {
            S = T;
            first = false;
        } else {
            add_full(S, T);
        }
        A = load_rows<STORE>(f, (uint32_t)nu * V + eo_mnu);
        B = load_rows<STORE>(f, (uint32_t)mu * V + eo_mnu);
        Cc = load_rows<STORE>(f, (uint32_t)nu * V + eo_mnu_pmu);
        need_row2<STORE>(Cc);
        need_row2<STORE>(A);
        mul_2r(B, Cc, W);
        complete_row2(W);
        mul_2r_adag(A, W, T);
        complete_row2(T);
        add_full(S, T);
    }
    return S;
}

// staple sum (monte.cu:63-80) for a compile-time direction MU.  With PREFETCH the
// six links U_MU(x +- nu) are loaded first: each of them is needed by six sites
// (y -+ nu), and issuing all six loads at the start of every thread puts those six
// requests at the same point of the program, close in time, so the L2 serves five
// of them.  It costs ~36 VGPRs; the heatbath kernel, whose sampler needs the
// registers, loads them inside the loop instead.
template <int STORE, int MU, bool PREFETCH>
SUN_HD F3 staple_su3(const Field& f, const Geom& g, int id, int parity, const int x[4], int lex) {
    const uint32_t V = (uint32_t)g.V;
    const int opar = parity ^ 1;
    const uint32_t self = (uint32_t)(id + parity * g.half);
    int xm[4] = {x[0], x[1], x[2], x[3]};
    const uint32_t eo_pmu = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xm, MU, +1), opar);
    constexpr int N0 = MU == 0 ? 1 : 0;
    constexpr int N1 = MU <= 1 ? 2 : 1;
    constexpr int N2 = MU <= 2 ? 3 : 2;
    constexpr int NUS[3] = {N0, N1, N2};
    uint32_t eo_pnu[3], eo_mnu[3], eo_mnu_pmu[3];
    F3 Bup[3], Bdn[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int nu = NUS[k];
        int xn[4] = {x[0], x[1], x[2], x[3]};
        eo_pnu[k] = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xn, nu, +1), opar);
        int xd[4] = {x[0], x[1], x[2], x[3]};
        const int lex_mnu = lex_shift(g, lex, xd, nu, -1);
        eo_mnu[k] = (uint32_t)eo_of_lex(g, lex_mnu, opar);
        eo_mnu_pmu[k] = (uint32_t)eo_of_lex(g, lex_shift(g, lex_mnu, xd, MU, +1), parity);
        if constexpr (PREFETCH) {
            Bup[k] = load_rows<STORE>(f, (uint32_t)MU * V + eo_pnu[k]);
            Bdn[k] = load_rows<STORE>(f, (uint32_t)MU * V + eo_mnu[k]);
        }
    }
    F3 S;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int nu = NUS[k];
        if constexpr (!PREFETCH) {
            Bup[k] = load_rows<STORE>(f, (uint32_t)MU * V + eo_pnu[k]);
            Bdn[k] = load_rows<STORE>(f, (uint32_t)MU * V + eo_mnu[k]);
        }
        // up: U_nu(x) [U_mu(x+nu) U_nu(x+mu)^+]
        F3 A = load_rows<STORE>(f, (uint32_t)nu * V + self);
        F3 Cc = load_rows<STORE>(f, (uint32_t)nu * V + eo_pmu);
        need_row2<STORE>(Cc);
        F3 W, T;
        mul_2r_bdag(Bup[k], Cc, W);
        complete_row2(W);
        mul_2r(A, W, T);
        complete_row2(T);
        if (k == 0) S = T; else add_full(S, T);
        // down: U_nu(x-nu)^+ [U_mu(x-nu) U_nu(x-nu+mu)]
        A = load_rows<STORE>(f, (uint32_t)nu * V + eo_mnu[k]);
        Cc = load_rows<STORE>(f, (uint32_t)nu * V + eo_mnu_pmu[k]);
        need_row2<STORE>(Cc);
        need_row2<STORE>(A);
        mul_2r(Bdn[k], Cc, W);
        complete_row2(W);
        mul_2r_adag(A, W, T);
        complete_row2(T);
        add_full(S, T);
    }
    return S;
}

// A link as stored (12, 18 or 8 reals), decoded only when it is multiplied, so a
// register set of in-flight links costs its stored size: 8 reals for 8-real storage.
template <int STORE>
struct RawLink {
    float r[Layout<float, 3, STORE>::NREAL];
};
template <int STORE>
SUN_HD RawLink<STORE> load_raw(const Field& f, uint32_t l) {
    RawLink<STORE> x;
    load_reals<float, 3, STORE>(f, l, x.r);
    return x;
}
// rows 0,1 valid; row 2 valid unless STORE == 12 (need_row2 builds it).  For 8-real
// links the stored reals pass through an empty asm first: the decode (sqrt, sin,
// cos, a division) then cannot be hoisted to the load, where the decoded 18 reals
// of every in-flight link would have to live in registers at once.
template <int STORE>
SUN_HD F3 decode(RawLink<STORE> x) {
    F3 m;
    if constexpr (STORE == kStore8) {
        asm volatile("" : "+v"(x.r[0]), "+v"(x.r[1]), "+v"(x.r[2]), "+v"(x.r[3]), "+v"(x.r[4]), "+v"(x.r[5]),
                     "+v"(x.r[6]), "+v"(x.r[7]));
    }
    if constexpr (STORE == kStore12) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            m.e[0][j] = mk<float>(x.r[2 * j], x.r[2 * j + 1]);
            m.e[1][j] = mk<float>(x.r[6 + 2 * j], x.r[6 + 2 * j + 1]);
        }
    } else {
        m = unpack<float, 3, STORE>(x.r);
    }
    return m;
}

// staple sum (monte.cu:63-80) for a compile-time direction MU, software-pipelined:
// the six links of direction nu_{k+1} are loaded into one register set while the
// other set's direction nu_k is multiplied, so a wavefront keeps ~18 16-byte loads
// in flight instead of ~4.  ~190 VGPRs for 12-real links (less for 8-real, whose
// sets hold the 8 stored reals and decode at use): for kernels run at 2 waves per
// SIMD, the occupancy at which the L2 still holds the neighbour links between their
// uses.  Same products in the same order as staple_su3.  The link being updated,
// U_MU(x), is loaded with the last direction's set into Uown.
template <int STORE, int MU>
SUN_HD F3 staple_su3_piped(const Field& f, const Geom& g, int id, int parity, const int x[4], int lex, F3& Uown) {
    const uint32_t V = (uint32_t)g.V;
    const int opar = parity ^ 1;
    const uint32_t self = (uint32_t)(id + parity * g.half);
    int xm[4] = {x[0], x[1], x[2], x[3]};
    const uint32_t eo_pmu = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xm, MU, +1), opar);
    constexpr int N0 = MU == 0 ? 1 : 0;
    constexpr int N1 = MU <= 1 ? 2 : 1;
    constexpr int N2 = MU <= 2 ? 3 : 2;
    constexpr int NUS[3] = {N0, N1, N2};
    uint32_t eo_pnu[3], eo_mnu[3], eo_mnu_pmu[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int nu = NUS[k];
        int xn[4] = {x[0], x[1], x[2], x[3]};
        eo_pnu[k] = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xn, nu, +1), opar);
        int xd[4] = {x[0], x[1], x[2], x[3]};
        const int lex_mnu = lex_shift(g, lex, xd, nu, -1);
        eo_mnu[k] = (uint32_t)eo_of_lex(g, lex_mnu, opar);
        eo_mnu_pmu[k] = (uint32_t)eo_of_lex(g, lex_shift(g, lex_mnu, xd, MU, +1), parity);
    }
    // register sets [slot]: U_mu(x+nu), U_mu(x-nu), U_nu(x), U_nu(x+mu), U_nu(x-nu), U_nu(x-nu+mu)
    RawLink<STORE> Bu[2], Bd[2], Au[2], Cu[2], Ad[2], Cd[2], Uraw;
#define SUN_PIPE_LOAD(k, slot)                                                          \
    do {                                                                                \
        const uint32_t nuV = (uint32_t)NUS[k] * V;                                      \
        Bu[slot] = load_raw<STORE>(f, (uint32_t)MU * V + eo_pnu[k]);                   \
        Bd[slot] = load_raw<STORE>(f, (uint32_t)MU * V + eo_mnu[k]);                   \
        Au[slot] = load_raw<STORE>(f, nuV + self);                                      \
        Cu[slot] = load_raw<STORE>(f, nuV + eo_pmu);                                    \
        Ad[slot] = load_raw<STORE>(f, nuV + eo_mnu[k]);                                 \
        Cd[slot] = load_raw<STORE>(f, nuV + eo_mnu_pmu[k]);                             \
    } while (0)
    SUN_PIPE_LOAD(0, 0);
    F3 S;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int cur = k & 1;
        if (k + 1 < 3) SUN_PIPE_LOAD(k + 1, cur ^ 1);
        if (k == 1) Uraw = load_raw<STORE>(f, (uint32_t)MU * V + self);
        F3 W, T;
        {
            F3 B = decode(Bu[cur]), Cc = decode(Cu[cur]), A = decode(Au[cur]);
            need_row2<STORE>(Cc);
            mul_2r_bdag(B, Cc, W);
            complete_row2(W);
            mul_2r(A, W, T);
            complete_row2(T);
        }
        if (k == 0) S = T; else add_full(S, T);
        {
            F3 B = decode(Bd[cur]), Cc = decode(Cd[cur]), A = decode(Ad[cur]);
            need_row2<STORE>(Cc);
            need_row2<STORE>(A);
            mul_2r(B, Cc, W);
            complete_row2(W);
            mul_2r_adag(A, W, T);
            complete_row2(T);
        }
        add_full(S, T);
    }
#undef SUN_PIPE_LOAD
    Uown = decode(Uraw);
    return S;
}

// The same staple pipelined by half-directions: stage h = 2k (up term of nu_k) or
// 2k+1 (down term) needs three links; stage h+1's three are in flight while stage h
// is multiplied.  ~150 VGPRs for 12-real links, so 3 waves per SIMD fit (more
// waves, each with half the loads in flight of staple_su3_piped).  Same products in
// the same order as staple_su3.
template <int STORE, int MU>
SUN_HD F3 staple_su3_piped_half(const Field& f, const Geom& g, int id, int parity, const int x[4], int lex,
                                F3& Uown) {
    const uint32_t V = (uint32_t)g.V;
    const int opar = parity ^ 1;
    const uint32_t self = (uint32_t)(id + parity * g.half);
    int xm[4] = {x[0], x[1], x[2], x[3]};
    const uint32_t eo_pmu = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xm, MU, +1), opar);
    constexpr int N0 = MU == 0 ? 1 : 0;
    constexpr int N1 = MU <= 1 ? 2 : 1;
    constexpr int N2 = MU <= 2 ? 3 : 2;
    constexpr int NUS[3] = {N0, N1, N2};
    uint32_t eo_pnu[3], eo_mnu[3], eo_mnu_pmu[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int nu = NUS[k];
        int xn[4] = {x[0], x[1], x[2], x[3]};
        eo_pnu[k] = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xn, nu, +1), opar);
        int xd[4] = {x[0], x[1], x[2], x[3]};
        const int lex_mnu = lex_shift(g, lex, xd, nu, -1);
        eo_mnu[k] = (uint32_t)eo_of_lex(g, lex_mnu, opar);
        eo_mnu_pmu[k] = (uint32_t)eo_of_lex(g, lex_shift(g, lex_mnu, xd, MU, +1), parity);
    }
    // slot: B = U_mu(x +- nu), A = U_nu(x) | U_nu(x-nu), C = U_nu(x+mu) | U_nu(x-nu+mu)
    RawLink<STORE> B[2], A[2], C[2], Uraw;
#define SUN_HALF_LOAD(h, slot)                                                             \
    do {                                                                                   \
        const int k_ = (h) >> 1;                                                           \
        const uint32_t nuV = (uint32_t)NUS[k_] * V;                                        \
        if (((h) & 1) == 0) {                                                              \
            B[slot] = load_raw<STORE>(f, (uint32_t)MU * V + eo_pnu[k_]);                  \
            A[slot] = load_raw<STORE>(f, nuV + self);                                      \
            C[slot] = load_raw<STORE>(f, nuV + eo_pmu);                                    \
        } else {                                                                           \
            B[slot] = load_raw<STORE>(f, (uint32_t)MU * V + eo_mnu[k_]);                  \
            A[slot] = load_raw<STORE>(f, nuV + eo_mnu[k_]);                                \
            C[slot] = load_raw<STORE>(f, nuV + eo_mnu_pmu[k_]);                            \
        }                                                                                  \
    } while (0)
    SUN_HALF_LOAD(0, 0);
    F3 S;
#pragma unroll
    for (int h = 0; h < 6; h++) {
        const int cur = h & 1;
        if (h + 1 < 6) SUN_HALF_LOAD(h + 1, cur ^ 1);
        if (h == 4) Uraw = load_raw<STORE>(f, (uint32_t)MU * V + self);
        F3 W, T;
        F3 Bm = decode(B[cur]), Cm = decode(C[cur]), Am = decode(A[cur]);
        need_row2<STORE>(Cm);
        if ((h & 1) == 0) {
            mul_2r_bdag(Bm, Cm, W);
            complete_row2(W);
            mul_2r(Am, W, T);
            complete_row2(T);
            if (h == 0) S = T; else add_full(S, T);
        } else {
            need_row2<STORE>(Am);
            mul_2r(Bm, Cm, W);
            complete_row2(W);
            mul_2r_adag(Am, W, T);
            complete_row2(T);
            add_full(S, T);
        }
    }
#undef SUN_HALF_LOAD
    Uown = decode(Uraw);
    return S;
}

// rows p,q of U times columns p,q of F = S^+, projected on SU(2):
//   a0=(UF)_pp a1=(UF)_pq a2=(UF)_qp a3=(UF)_qq, (UF)_ab = sum_j U_aj conj(S_bj)
SUN_HD Quat<float> block_proj(const F3& U, const F3& S, int p, int q) {
    C a0 = cmul_cb(U.e[p][0], S.e[p][0]);
    C a1 = cmul_cb(U.e[p][0], S.e[q][0]);
    C a2 = cmul_cb(U.e[q][0], S.e[p][0]);
    C a3 = cmul_cb(U.e[q][0], S.e[q][0]);
#pragma unroll
    for (int j = 1; j < 3; j++) {
        cmac_cb(a0, U.e[p][j], S.e[p][j]);
        cmac_cb(a1, U.e[p][j], S.e[q][j]);
        cmac_cb(a2, U.e[q][j], S.e[p][j]);
        cmac_cb(a3, U.e[q][j], S.e[q][j]);
    }
    Quat<float> r;
    r.a[0] = a0.re + a3.re;
    r.a[1] = a1.im + a2.im;
    r.a[2] = a1.re - a2.re;
    r.a[3] = a0.im - a3.im;
    return r;
}

// rows p,q of U <- [[r0+i r3, r2+i r1], [-r2+i r1, r0-i r3]] * rows p,q
SUN_HD void rotate_rows(F3& U, const Quat<float>& r, int p, int q) {
    const C b0 = mk<float>(r.a[0], r.a[3]), b1 = mk<float>(r.a[2], r.a[1]);
    const C b2 = mk<float>(-r.a[2], r.a[1]), b3 = mk<float>(r.a[0], -r.a[3]);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        C t0 = cmul(b0, U.e[p][j]);
        cmac(t0, b1, U.e[q][j]);
        C t1 = cmul(b2, U.e[p][j]);
        cmac(t1, b3, U.e[q][j]);
        U.e[p][j] = t0;
        U.e[q][j] = t1;
    }
}

// quaternion square: r*r in the product of msu2.h:90-97
SUN_HD Quat<float> qsquare(const Quat<float>& v) {
    Quat<float> r;
    r.a[0] = v.a[0] * v.a[0] - v.a[1] * v.a[1] - v.a[2] * v.a[2] - v.a[3] * v.a[3];
    r.a[1] = 2.f * v.a[0] * v.a[1];
    r.a[2] = 2.f * v.a[0] * v.a[2];
    r.a[3] = 2.f * v.a[0] * v.a[3];
    return r;
}

SUN_HD void overrelax_su3(F3& U, const F3& S) {
#pragma unroll
    for (int block = 0; block < 3; block++) {
        const int p = block_p(3, block), q = block_q(3, block);
        Quat<float> r = block_proj(U, S, p, q);
        conj_normalize(r);
        rotate_rows(U, qsquare(r), p, q);
    }
}

SUN_HD void heatbath_su3(F3& U, const F3& S, double bnc, const DrawKey& key) {
    // the counters of every block's first attempt and final draws are known now
    const BlockDraws bd[3] = {BlockDraws(key, 0), BlockDraws(key, 1), BlockDraws(key, 2)};
#pragma unroll
    for (int block = 0; block < 3; block++) {
        const int p = block_p(3, block), q = block_q(3, block);
        Quat<float> r = block_proj(U, S, p, q);
        // |r| and 1/|r| from one reciprocal square root
        const float n2 = r.a[0] * r.a[0] + r.a[1] * r.a[1] + r.a[2] * r.a[2] + r.a[3] * r.a[3];
        const float k = f32_rsq(n2);
        const float ap = (float)bnc * (n2 * k);
#pragma unroll
        for (int i = 0; i < 4; i++) r.a[i] *= k;
        Quat<float> a;
        generate_su2_milc_f32(ap, key, block, bd[block], a.a);
        rotate_rows(U, mul_uv_dagger(a, r), p, q);
    }
}

// Re Tr(X Y^+) = sum_ij Re(X_ij conj(Y_ij))
SUN_HD float retr_xydag(const F3& X, const F3& Y) {
    float s = X.e[0][0].re * Y.e[0][0].re;
    s = fmaf(X.e[0][0].im, Y.e[0][0].im, s);
#pragma unroll
    for (int k = 1; k < 9; k++) {
        const C a = X.e[k / 3][k % 3], b = Y.e[k / 3][k % 3];
        s = fmaf(a.re, b.re, s);
        s = fmaf(a.im, b.im, s);
    }
    return s;
}

// The six plaquettes at site (idd in [0,V)) as in plaquette_cub.cu:96-149:
// P_munu = Re Tr[U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+] = Re Tr[X Y^+] with
// X = U_mu(x) U_nu(x+mu), Y = U_nu(x) U_mu(x+nu); spatial (nu<3) and temporal sums.
template <int STORE>
SUN_HD void plaquette_site_su3(const Field& f, const Geom& g, int idd, double& sp, double& tp) {
    const int parity = idd >= g.half ? 1 : 0;
    const int id = idd - parity * g.half;
    const uint32_t V = (uint32_t)g.V;
    int x[4];
    const int lex = site_coords(g, id, parity, x);
    F3 L[4];
#pragma unroll
    for (int mu = 0; mu < 4; mu++) L[mu] = load_rows<STORE>(f, (uint32_t)mu * V + (uint32_t)idd);
#pragma unroll
    for (int mu = 0; mu < 3; mu++) {
        int xm[4] = {x[0], x[1], x[2], x[3]};
        const uint32_t eo_pmu = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xm, mu, +1), parity ^ 1);
#pragma unroll
        for (int nu = mu + 1; nu < 4; nu++) {
            int xn[4] = {x[0], x[1], x[2], x[3]};
            const uint32_t eo_pnu = (uint32_t)eo_of_lex(g, lex_shift(g, lex, xn, nu, +1), parity ^ 1);
            F3 B = load_rows<STORE>(f, (uint32_t)nu * V + eo_pmu);
            F3 D = load_rows<STORE>(f, (uint32_t)mu * V + eo_pnu);
            need_row2<STORE>(B);
            need_row2<STORE>(D);
            F3 X, Y;
            mul_2r(L[mu], B, X);
            complete_row2(X);
            mul_2r(L[nu], D, Y);
            complete_row2(Y);
            const double tr = (double)retr_xydag(X, Y);
            if (nu == 3) tp += tr; else sp += tr;
        }
    }
}

}  // namespace fast
}  // namespace sun
