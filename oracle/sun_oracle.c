/*
 * sun_oracle.c -- FP64 CPU restatement of nmrcardoso/sun's quenched SU(N)
 * generation path.  TEST INFRASTRUCTURE ONLY (the checker): it is never part of the
 * product path.  See sun_oracle.h for the layout and the list of functions.
 *
 * Arithmetic follows the reference's operation order (complex.h:150-152,
 * matrixsun.h:298-347) and is built with -ffp-contract=off, so every multiply-add
 * is rounded the way the reference's source writes it.  The sampler's and the Haar
 * start's log, exp and sin/cos of 2 pi u are sun_libm.h's (the engine's own portable
 * implementations, compiled here as on the device and pinned against glibc by the
 * CPU tests), so FP64 chains agree bit for bit with the device; other
 * transcendentals come from the host libm.  The reference uses CUDA libm, a
 * boundary no reference test pins.
 */
#include "sun_oracle.h"
#include "sun_libm.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NMAX 4

#define RNG_TAG_HB   0x10000000u
#define RNG_TAG_HOT  0x20000000u
#define RNG_TAG_TEST 0x30000000u
#define RNG_TAG_MH   0x40000000u
#define RNG_TAG_MHX  0x50000000u

/* ------------------------------------------------------------------------- */
/* complex arithmetic in the reference's operation order (complex.h:82-230)    */
/* ------------------------------------------------------------------------- */
typedef struct { double re, im; } cx;

static inline cx cmk(double r, double i) { cx z; z.re = r; z.im = i; return z; }
static inline cx cadd(cx a, cx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return cmk(a.re - b.re, a.im - b.im); }
/* complex.h:150-152: (x*bx - y*by, y*bx + x*by) */
static inline cx cmul(cx a, cx b) { return cmk(a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im); }
static inline cx cconj(cx a) { return cmk(a.re, -a.im); }
static inline double cabs2(cx a) { return a.re * a.re + a.im * a.im; }
static inline cx cscale(cx a, double s) { return cmk(a.re * s, a.im * s); }

typedef struct { cx e[NMAX][NMAX]; } mat;

static inline void mload(int nc, const double *p, mat *m) {
    if (nc < NMAX) memset(m, 0, sizeof(*m));
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) {
            m->e[i][j].re = p[2 * (i * nc + j)];
            m->e[i][j].im = p[2 * (i * nc + j) + 1];
        }
}
static inline void mstore(int nc, const mat *m, double *p) {
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) {
            p[2 * (i * nc + j)] = m->e[i][j].re;
            p[2 * (i * nc + j) + 1] = m->e[i][j].im;
        }
}
static inline void mzero(int nc, mat *m) {
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) m->e[i][j] = cmk(0.0, 0.0);
}
static inline void munit(int nc, mat *m) {
    mzero(nc, m);
    for (int i = 0; i < nc; i++) m->e[i][i] = cmk(1.0, 0.0);
}
/* matrixsun.h:298-327: res_ij = e_i0*A_0j (+ e_ik*A_kj, k ascending) */
static inline mat mmul(int nc, const mat *a, const mat *b) {
    mat r;
    if (nc < NMAX) memset(&r, 0, sizeof(r));
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) {
            cx s = cmul(a->e[i][0], b->e[0][j]);
            for (int k = 1; k < nc; k++) s = cadd(s, cmul(a->e[i][k], b->e[k][j]));
            r.e[i][j] = s;
        }
    return r;
}
/* matrixsun.h:568-583 */
static inline mat mdag(int nc, const mat *a) {
    mat r;
    if (nc < NMAX) memset(&r, 0, sizeof(r));
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) r.e[i][j] = cconj(a->e[j][i]);
    return r;
}
static inline void madd_to(int nc, mat *acc, const mat *a) {
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) acc->e[i][j] = cadd(acc->e[i][j], a->e[i][j]);
}
/* matrixsun.h:608-630 */
static inline double mrealtrace(int nc, const mat *a) {
    double s = a->e[0][0].re;
    for (int i = 1; i < nc; i++) s += a->e[i][i].re;
    return s;
}
static inline cx mtrace(int nc, const mat *a) {
    cx s = a->e[0][0];
    for (int i = 1; i < nc; i++) s = cadd(s, a->e[i][i]);
    return s;
}

/* matrixsun.h:481-505: SU(3) cofactor expansion; Crout LU (no pivoting) for N!=3 */
static cx mdet(int nc, const mat *a) {
    cx res;
    if (nc == 3) {
        res = cmul(cmul(a->e[0][1], a->e[1][2]), a->e[2][0]);
        res = csub(res, cmul(cmul(a->e[0][2], a->e[1][1]), a->e[2][0]));
        res = cadd(res, cmul(cmul(a->e[0][2], a->e[1][0]), a->e[2][1]));
        res = csub(res, cmul(cmul(a->e[0][0], a->e[1][2]), a->e[2][1]));
        res = csub(res, cmul(cmul(a->e[0][1], a->e[1][0]), a->e[2][2]));
        res = cadd(res, cmul(cmul(a->e[0][0], a->e[1][1]), a->e[2][2]));
        return res;
    }
    mat b = *a;
    for (int j = 0; j < nc; j++) {
        for (int i = 0; i <= j; i++) {
            res = b.e[j][i];
            for (int c = 0; c < i; c++) res = csub(res, cmul(b.e[c][i], b.e[j][c]));
            b.e[j][i] = res;
        }
        for (int i = j + 1; i < nc; i++) {
            res = b.e[j][i];
            for (int c = 0; c < j; c++) res = csub(res, cmul(b.e[c][i], b.e[j][c]));
            /* conj(b_jj)*res / |b_jj|^2 ; complex/real division per complex.h:199 */
            cx t = cmul(cconj(b.e[j][j]), res);
            double d = cabs2(b.e[j][j]);
            b.e[j][i] = cmk(t.re / d, t.im / d);
        }
    }
    res = cmul(b.e[0][0], b.e[1][1]);
    for (int c = 2; c < nc; c++) res = cmul(res, b.e[c][c]);
    return res;
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10                                                               */
/* ------------------------------------------------------------------------- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform double in (0,1] from one 32-bit Philox word: (w + 1) * 2^-32 (32 bits of
 * randomness per draw, as cuRAND's MRG32k3a uniforms have, random.h:16-22). */
double or_u01(uint32_t w) {
    return ((double)w + 1.0) * 0x1p-32;
}

/* sequential uniform stream (the Haar start, or_stream_uniforms): draw k of stream
 * (seed, c0, c2, c3) is word k%4 of Philox4x32-10((c0, k/4, c2, c3), seed). */
typedef struct {
    uint32_t key[2];
    uint32_t c0, c2, c3, j;
    uint32_t w[4];
    int nw;
} usrc;

static void usrc_stream(usrc *s, uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3) {
    s->key[0] = (uint32_t)seed; s->key[1] = (uint32_t)(seed >> 32);
    s->c0 = c0; s->c2 = c2; s->c3 = c3; s->j = 0; s->nw = 0;
}
static inline double unext(usrc *s) {
    if (s->nw == 0) {
        uint32_t ctr[4] = { s->c0, s->j, s->c2, s->c3 };
        or_philox4x32_10(ctr, s->key, s->w);
        s->j++;
        s->nw = 4;
    }
    return or_u01(s->w[4 - s->nw--]);
}

/* Uniforms of one SU(2) subgroup update, block b of a link update with counter words
 * (c0, c2, c3): attempt k (k = 0 the first Kennedy-Pendleton attempt, k >= 1 its
 * retries or the Creutz tries) reads the words of Philox((c0, 32 b + k, c2, c3), seed),
 * the final two draws words 0 and 1 of Philox((c0, 32 b + 31, c2, c3)) -- the engine's
 * layout (sun_rng.hpp), a choice of this build (the reference streams cuRAND
 * sequentially).  An explicit list (tests) is read in order instead: 4 uniforms per
 * Kennedy-Pendleton attempt, 2 per Creutz try, then the 2 final ones. */
typedef struct {
    uint32_t key[2];
    uint32_t c0, c2, c3, cbase;
    const double *list;
    int nlist, pos;
} bsrc;

static void bsrc_block(bsrc *b, uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int block) {
    b->key[0] = (uint32_t)seed; b->key[1] = (uint32_t)(seed >> 32);
    b->c0 = c0; b->c2 = c2; b->c3 = c3; b->cbase = 32u * (uint32_t)block;
    b->list = NULL; b->nlist = 0; b->pos = 0;
}
static double blist(bsrc *b) {
    double v = (b->pos < b->nlist) ? b->list[b->pos] : 1.0;
    b->pos++;
    return v;
}
/* the n (4 or 2) uniforms of attempt k */
static void battempt(bsrc *b, int k, int n, double *u) {
    if (b->list) {
        for (int i = 0; i < n; i++) u[i] = blist(b);
        return;
    }
    uint32_t ctr[4] = { b->c0, b->cbase + (uint32_t)k, b->c2, b->c3 }, w[4];
    or_philox4x32_10(ctr, b->key, w);
    for (int i = 0; i < n; i++) u[i] = or_u01(w[i]);
}
static void bfinal(bsrc *b, double u[2]) {
    if (b->list) {
        u[0] = blist(b);
        u[1] = blist(b);
        return;
    }
    uint32_t ctr[4] = { b->c0, b->cbase + 31u, b->c2, b->c3 }, w[4];
    or_philox4x32_10(ctr, b->key, w);
    u[0] = or_u01(w[0]);
    u[1] = or_u01(w[1]);
}

void or_stream_uniforms(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, double *out) {
    usrc s;
    usrc_stream(&s, seed, c0, c2, c3);
    for (int i = 0; i < n; i++) out[i] = unext(&s);
}

/* ------------------------------------------------------------------------- */
/* SU(2) quaternion helpers (msu2.h:141-155, 211-219)                          */
/* ------------------------------------------------------------------------- */
static inline double q_abs(const double r[4]) {
    return sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
}
static inline void q_conj_normalize(double r[4]) {
    double norm = 1.0 / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
    r[0] *= norm; r[1] *= -norm; r[2] *= -norm; r[3] *= -norm;
}
/* msu2.h:211-219: b = v * u^dagger */
static inline void q_mul_uv_dagger(const double v[4], const double u[4], double b[4]) {
    b[0] = v[0] * u[0] + v[1] * u[1] + v[2] * u[2] + v[3] * u[3];
    b[1] = v[1] * u[0] - v[0] * u[1] + v[2] * u[3] - v[3] * u[2];
    b[2] = v[2] * u[0] - v[0] * u[2] + v[3] * u[1] - v[1] * u[3];
    b[3] = v[3] * u[0] - v[0] * u[3] + v[1] * u[2] - v[2] * u[1];
}

/* device_PHB_OVR.h:162-223 (MILC Kennedy-Pendleton / Creutz generator) */
/* margin (may be NULL): lowered to the smallest |distance from a tie| among the
 * decisions this call takes -- the branch test al > 2, each Kennedy-Pendleton test
 * (1 - d/2) - xr4^2 and each Creutz test (1 - a0^2) - xr2^2.  Arithmetic of another
 * precision can only decide differently where that distance is at its rounding
 * level; the FP32 parity tests use it to attribute links that differ by O(1). */
#define TIE(m) do { if (margin) { double t_ = fabs(m); if (t_ < *margin) *margin = t_; } } while (0)
static void gen_su2_milc(double al, bsrc *st, double a[4], double *margin) {
    double xr1, xr2, xr3, xr4, d, r;
    int k;
    double a0 = 0.0, u[4];
    battempt(st, 0, 4, u);
    xr1 = u[0];
    xr1 = sun_log(xr1 + 1.e-10);
    xr2 = u[1];
    xr2 = sun_log(xr2 + 1.e-10);
    xr3 = u[2];
    xr4 = u[3];
    xr3 = sun_cos2pi(xr3);
    d = -(xr2 + xr1 * xr3 * xr3) / al;
    int nacd = 0;
    TIE(al - 2.0);
    TIE((1.00 - 0.5 * d) - xr4 * xr4);
    if ((1.00 - 0.5 * d) > xr4 * xr4) nacd = 1;
    if (nacd == 0 && al > 2.0) {
        for (k = 1; k <= 20; k++) {
            battempt(st, k, 4, u);
            xr1 = u[0];
            xr1 = sun_log(xr1 + 1.e-10);
            xr2 = u[1];
            xr2 = sun_log(xr2 + 1.e-10);
            xr3 = u[2];
            xr4 = u[3];
            xr3 = sun_cos2pi(xr3);
            d = -(xr2 + xr1 * xr3 * xr3) / al;
            TIE((1.00 - 0.5 * d) - xr4 * xr4);
            if ((1.00 - 0.5 * d) > xr4 * xr4) break;
        }
    }
    if (nacd == 0 && al <= 2.0) {
        xr3 = sun_exp(-2.0 * al);
        xr4 = 1.0 - xr3;
        for (k = 1; k <= 20; k++) {
            battempt(st, k, 2, u);
            xr1 = u[0];
            xr2 = u[1];
            r = xr3 + xr4 * xr1;
            a0 = 1.00 + sun_log(r) / al;
            TIE((1.0 - a0 * a0) - xr2 * xr2);
            if ((1.0 - a0 * a0) > xr2 * xr2) break;
        }
        d = 1.0 - a0;
    }
    a[0] = 1.0 - d;
    xr3 = 1.0 - a[0] * a[0];
    xr3 = fabs(xr3);
    r = sqrt(xr3);
    double fin[2];
    bfinal(st, fin);
    a[3] = (2.0 * fin[0] - 1.0) * r;
    xr1 = xr3 - a[3] * a[3];
    xr1 = fabs(xr1);
    xr1 = sqrt(xr1);
    /* sin/cos of PII * u (device_PHB_OVR.h:217-221) with the angle in revolutions */
    double sn, cs;
    sun_sincos2pi(fin[1], &sn, &cs);
    a[1] = xr1 * cs;
    a[2] = xr1 * sn;
}

int or_generate_su2_milc(double al, const double *u, int nu, double a[4]) {
    bsrc s;
    memset(&s, 0, sizeof(s));
    s.list = u; s.nlist = nu; s.pos = 0;
    gen_su2_milc(al, &s, a, NULL);
    return s.pos;
}

void or_sample_a0(double al, uint64_t seed, int n, double *a0_out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        bsrc s;
        double a[4];
        bsrc_block(&s, seed, (uint32_t)i, 0u, RNG_TAG_TEST, 0);
        gen_su2_milc(al, &s, a, NULL);
        a0_out[i] = a[0];
    }
}

/* ------------------------------------------------------------------------- */
/* geometry (include/index.h:133-202)                                          */
/* ------------------------------------------------------------------------- */
void or_eo_to_coords(int id, int parity, const int X[4], int x[4]) {
    int za = id / (X[0] / 2);
    int zb = za / X[1];
    x[1] = za - zb * X[1];
    x[3] = zb / X[2];
    x[2] = zb - x[3] * X[2];
    int xodd = (x[1] + x[2] + x[3] + parity) & 1;
    x[0] = (2 * id + xodd) - za * X[0];
}

int or_coords_to_eo(const int x[4], const int X[4]) {
    int lex = ((x[3] * X[2] + x[2]) * X[1] + x[1]) * X[0] + x[0];
    int par = (x[0] + x[1] + x[2] + x[3]) & 1;
    int vol = X[0] * X[1] * X[2] * X[3];
    return (lex >> 1) + par * (vol / 2);
}

int or_neighbor_eo(int id, int parity, const int X[4], int mu, int lmu, int nu, int lnu) {
    int x[4];
    or_eo_to_coords(id, parity, X, x);
    x[mu] = (x[mu] + lmu + X[mu]) % X[mu];
    if (nu >= 0) x[nu] = (x[nu] + lnu + X[nu]) % X[nu];
    return or_coords_to_eo(x, X);
}

static inline long volume_of(const int X[4]) { return (long)X[0] * X[1] * X[2] * X[3]; }
static inline const double *link_at(int nc, const double *U, long V, int mu, long s) {
    return U + ((long)mu * V + s) * nc * nc * 2;
}
static inline double *link_at_w(int nc, double *U, long V, int mu, long s) {
    return U + ((long)mu * V + s) * nc * nc * 2;
}

/* ------------------------------------------------------------------------- */
/* storage codecs                                                              */
/* ------------------------------------------------------------------------- */
/* reconstruct_12p_8p.h:17-21: third row = conj(r0 x r1) */
static void reconstruct12(mat *A) {
    A->e[2][0] = cconj(csub(cmul(A->e[0][1], A->e[1][2]), cmul(A->e[0][2], A->e[1][1])));
    A->e[2][1] = cconj(csub(cmul(A->e[0][2], A->e[1][0]), cmul(A->e[0][0], A->e[1][2])));
    A->e[2][2] = cconj(csub(cmul(A->e[0][0], A->e[1][1]), cmul(A->e[0][1], A->e[1][0])));
}

/* reconstruct_12p_8p.h:38-82 */
static void reconstruct8(mat *m, cx theta) {
    double p1, p2, diff;
    cx temp;
    p1 = cabs2(m->e[0][1]) + cabs2(m->e[0][2]);
    p2 = 1.0 / p1;
    diff = 1.0 - p1;
    p1 = sqrt(diff >= 0 ? diff : 0.0);
    double s0, c0, s1, c1;
    sun_sincos(theta.re, &s0, &c0);
    sun_sincos(theta.im, &s1, &c1);
    m->e[0][0].re = p1 * c0;
    m->e[0][0].im = p1 * s0;
    diff = 1.0 - cabs2(m->e[0][0]) - cabs2(m->e[1][0]);
    p1 = sqrt(diff >= 0 ? diff : 0.0);
    m->e[2][0].re = p1 * c1;
    m->e[2][0].im = p1 * s1;
    m->e[1][1] = cconj(cmul(m->e[2][0], m->e[0][2]));
    m->e[2][2] = cmul(cconj(m->e[0][0]), m->e[1][0]);
    m->e[1][1] = cadd(m->e[1][1], cmul(m->e[2][2], m->e[0][1]));
    m->e[1][1] = cscale(m->e[1][1], -p2);
    m->e[1][2] = cconj(cmul(m->e[2][0], m->e[0][1]));
    m->e[1][2] = csub(m->e[1][2], cmul(m->e[2][2], m->e[0][2]));
    m->e[1][2] = cscale(m->e[1][2], p2);
    m->e[2][1] = cconj(cmul(m->e[1][0], m->e[0][2]));
    temp = cmul(cconj(m->e[0][0]), m->e[2][0]);
    m->e[2][1] = csub(m->e[2][1], cmul(temp, m->e[0][1]));
    m->e[2][1] = cscale(m->e[2][1], p2);
    m->e[2][2] = cconj(cmul(m->e[1][0], m->e[0][1]));
    m->e[2][2] = cadd(m->e[2][2], cmul(temp, m->e[0][2]));
    m->e[2][2] = cscale(m->e[2][2], -p2);
}

void or_encode12(const double *link, double *p12) { memcpy(p12, link, 12 * sizeof(double)); }
void or_decode12(const double *p12, double *link) {
    mat m;
    memset(&m, 0, sizeof(m));
    for (int j = 0; j < 3; j++) {
        m.e[0][j] = cmk(p12[2 * j], p12[2 * j + 1]);
        m.e[1][j] = cmk(p12[6 + 2 * j], p12[6 + 2 * j + 1]);
    }
    reconstruct12(&m);
    mstore(3, &m, link);
}
/* device_load_save.h:112-120: (e01, e02, e10, (arg e00, arg e20)) */
void or_encode8(const double *link, double *p8) {
    mat m;
    mload(3, link, &m);
    p8[0] = m.e[0][1].re; p8[1] = m.e[0][1].im;
    p8[2] = m.e[0][2].re; p8[3] = m.e[0][2].im;
    p8[4] = m.e[1][0].re; p8[5] = m.e[1][0].im;
    p8[6] = sun_atan2(m.e[0][0].im, m.e[0][0].re);
    p8[7] = sun_atan2(m.e[2][0].im, m.e[2][0].re);
}
void or_decode8(const double *p8, double *link) {
    mat m;
    memset(&m, 0, sizeof(m));
    m.e[0][1] = cmk(p8[0], p8[1]);
    m.e[0][2] = cmk(p8[2], p8[3]);
    m.e[1][0] = cmk(p8[4], p8[5]);
    reconstruct8(&m, cmk(p8[6], p8[7]));
    mstore(3, &m, link);
}

static inline void project_link(int nc, int storage, double *l) {
    if (nc != 3 || storage == OR_STORE_18) return;
    double p[12];
    if (storage == OR_STORE_12) { or_encode12(l, p); or_decode12(p, l); }
    else { or_encode8(l, p); or_decode8(p, l); }
}

void or_project_storage(int nc, int storage, long nlinks, double *U) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nlinks; i++) project_link(nc, storage, U + i * nc * nc * 2);
}

/* ------------------------------------------------------------------------- */
/* reunitarisation                                                             */
/* ------------------------------------------------------------------------- */
/* reunitlink.h:16-85 (N=3: row Gram-Schmidt + cross product; N>3: column G-S +
 * determinant phase).  N=2 is not compilable in the reference (reunitlink.h:86);
 * here the textbook SU(2) projection a=(U00+U11*)/2, b=(U01-U10*)/2 normalised. */
static void reunit_mat(int nc, mat *U) {
    if (nc == 3) {
        double t1 = 0.0;
        cx t2 = cmk(0.0, 0.0);
        for (int c = 0; c < 3; c++) t1 += cabs2(U->e[0][c]);
        t1 = 1.0 / sqrt(t1);
        for (int c = 0; c < 3; c++) U->e[0][c] = cscale(U->e[0][c], t1);
        for (int c = 0; c < 3; c++) t2 = cadd(t2, cmul(cconj(U->e[0][c]), U->e[1][c]));
        for (int c = 0; c < 3; c++) U->e[1][c] = csub(U->e[1][c], cmul(t2, U->e[0][c]));
        t1 = 0.0;
        for (int c = 0; c < 3; c++) t1 += cabs2(U->e[1][c]);
        t1 = 1.0 / sqrt(t1);
        for (int c = 0; c < 3; c++) U->e[1][c] = cscale(U->e[1][c], t1);
        reconstruct12(U);
        return;
    }
    if (nc == 2) {
        cx a = cmk((U->e[0][0].re + U->e[1][1].re) * 0.5, (U->e[0][0].im - U->e[1][1].im) * 0.5);
        cx b = cmk((U->e[0][1].re - U->e[1][0].re) * 0.5, (U->e[0][1].im + U->e[1][0].im) * 0.5);
        double n = 1.0 / sqrt(cabs2(a) + cabs2(b));
        a = cscale(a, n);
        b = cscale(b, n);
        U->e[0][0] = a;
        U->e[0][1] = b;
        U->e[1][0] = cmk(-b.re, b.im);
        U->e[1][1] = cconj(a);
        return;
    }
    double t1 = 0.0;
    cx t2;
    for (int c = 0; c < nc; c++) t1 += cabs2(U->e[c][0]);
    t1 = 1.0 / sqrt(t1);
    for (int c = 0; c < nc; c++) U->e[c][0] = cscale(U->e[c][0], t1);
    for (int j = 1; j < nc; j++) {
        for (int i = 0; i < j; i++) {
            t2 = cmul(cconj(U->e[0][i]), U->e[0][j]);
            for (int c = 1; c < nc; ++c) t2 = cadd(t2, cmul(cconj(U->e[c][i]), U->e[c][j]));
            for (int c = 0; c < nc; ++c) U->e[c][j] = csub(U->e[c][j], cmul(t2, U->e[c][i]));
        }
        t1 = 0.0;
        for (int c = 0; c < nc; ++c) t1 += cabs2(U->e[c][j]);
        t1 = 1.0 / sqrt(t1);
        for (int c = 0; c < nc; ++c) U->e[c][j] = cscale(U->e[c][j], t1);
    }
    t2 = mdet(nc, U);
    double ai = sun_atan2(t2.im, t2.re), sn, cs;
    sun_sincos(ai, &sn, &cs);
    t2 = cmk(cs, -sn);
    for (int c = 0; c < nc; ++c) U->e[c][nc - 1] = cmul(U->e[c][nc - 1], t2);
}

void or_reunit_link(int nc, double *link) {
    mat m;
    mload(nc, link, &m);
    reunit_mat(nc, &m);
    mstore(nc, &m, link);
}

/* reunitarize.cu:23-30: load (decode), reunit, store (encode) for every link */
void or_reunitarize(int nc, long nlinks, int storage, double *U) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nlinks; i++) {
        double *l = U + i * nc * nc * 2;
        or_reunit_link(nc, l);
        project_link(nc, storage, l);
    }
}

/* ------------------------------------------------------------------------- */
/* initialisation                                                              */
/* ------------------------------------------------------------------------- */
void or_cold_start(int nc, const int X[4], double *U) {
    long V = volume_of(X);
    mat one;
    munit(nc, &one);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 4 * V; i++) mstore(nc, &one, U + i * nc * nc * 2);
}

/* Haar hot start (replaces the reference's non-Haar randomize+reunit,
 * device_PHB_OVR.h:645-663 / gaugearray.cu:353-374): complex Gaussian entries by
 * Box-Muller from the stream (seed, link, 0, TAG_HOT|mu), one Philox call per
 * entry, row-major; then the reunitarisation above (Gram-Schmidt of a Gaussian
 * matrix is Haar on U(N); the det-phase / cross-product step makes it SU(N)). */
void or_hot_start(int nc, const int X[4], uint64_t seed, int storage, double *U) {
    long V = volume_of(X);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < 4 * V; l++) {
        int mu = (int)(l / V);
        long s = l - (long)mu * V;
        usrc st;
        usrc_stream(&st, seed, (uint32_t)s, 0u, RNG_TAG_HOT | (uint32_t)mu);
        mat m;
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nc; j++) {
                double u1 = unext(&st), u2 = unext(&st);
                double rr = sqrt(-2.0 * sun_log(u1));
                double sn, cs;
                sun_sincos2pi(u2, &sn, &cs);
                m.e[i][j] = cmk(rr * cs, rr * sn);
            }
        reunit_mat(nc, &m);
        double *dst = link_at_w(nc, U, V, mu, s);
        mstore(nc, &m, dst);
        project_link(nc, storage, dst);
    }
}

/* ------------------------------------------------------------------------- */
/* staple (monte.cu:63-80)                                                     */
/* ------------------------------------------------------------------------- */
/* nu runs over 0..nu_end-1 without mu (nu_end 3: the spatial staples of APE smearing) */
static mat staple_range(int nc, const int X[4], const double *U, int parity, int mu, int id, int nu_end) {
    long V = volume_of(X);
    long half = V / 2;
    long idxoddbit = id + parity * half;
    mat staple, link, t;
    mzero(nc, &staple);
    int newidmu1 = or_neighbor_eo(id, parity, X, mu, 1, -1, 0);
    for (int nu = 0; nu < nu_end; nu++) {
        if (nu == mu) continue;
        mat A, B, C;
        /* up: U_nu(x) U_mu(x+nu) U_nu(x+mu)^dagger */
        mload(nc, link_at(nc, U, V, nu, idxoddbit), &A);
        mload(nc, link_at(nc, U, V, mu, or_neighbor_eo(id, parity, X, nu, 1, -1, 0)), &B);
        mload(nc, link_at(nc, U, V, nu, newidmu1), &C);
        link = mmul(nc, &A, &B);
        t = mdag(nc, &C);
        link = mmul(nc, &link, &t);
        madd_to(nc, &staple, &link);
        /* down: U_nu(x-nu)^dagger U_mu(x-nu) U_nu(x-nu+mu) */
        int newidnum1 = or_neighbor_eo(id, parity, X, nu, -1, -1, 0);
        mload(nc, link_at(nc, U, V, nu, newidnum1), &A);
        mload(nc, link_at(nc, U, V, mu, newidnum1), &B);
        mload(nc, link_at(nc, U, V, nu, or_neighbor_eo(id, parity, X, mu, 1, nu, -1)), &C);
        link = mdag(nc, &A);
        link = mmul(nc, &link, &B);
        link = mmul(nc, &link, &C);
        madd_to(nc, &staple, &link);
    }
    return staple;
}

static mat staple_at(int nc, const int X[4], const double *U, int parity, int mu, int id) {
    return staple_range(nc, X, U, parity, mu, id, 4);
}

void or_staple(int nc, const int X[4], const double *U, int parity, int mu, int id, double *S) {
    mat s = staple_at(nc, X, U, parity, mu, id);
    mstore(nc, &s, S);
}

/* ------------------------------------------------------------------------- */
/* IndexBlock (constants.h:143-192)                                            */
/* ------------------------------------------------------------------------- */
static void index_block(int nc, int block, int *p, int *q) {
    if (nc == 3) {
        if (block == 0) { *p = 0; *q = 1; }
        else if (block == 1) { *p = 1; *q = 2; }
        else { *p = 0; *q = 2; }
        return;
    }
    int i1 = 0, found = 0, del_i = 0, index = -1;
    while (del_i < (nc - 1) && found == 0) {
        del_i++;
        for (i1 = 0; i1 < (nc - del_i); i1++) {
            index++;
            if (index == block) { found = 1; break; }
        }
    }
    *q = i1 + del_i;
    *p = i1;
}

/* device_PHB_OVR.h:296-303: link <- u * link on rows (x,y) */
static void mul_block_sun(int nc, const double u[4], mat *link, int x, int y) {
    cx a00 = cmk(u[0], u[3]), a01 = cmk(u[2], u[1]);
    cx a10 = cmk(-u[2], u[1]), a11 = cmk(u[0], -u[3]);
    for (int j = 0; j < nc; j++) {
        cx tmp = cadd(cmul(a00, link->e[x][j]), cmul(a01, link->e[y][j]));
        link->e[y][j] = cadd(cmul(a10, link->e[x][j]), cmul(a11, link->e[y][j]));
        link->e[x][j] = tmp;
    }
}

/* device_PHB_OVR.h:265-272 */
static void get_block_su2(const mat *m, int x, int y, double r[4]) {
    r[0] = m->e[x][x].re + m->e[y][y].re;
    r[1] = m->e[x][y].im + m->e[y][x].im;
    r[2] = m->e[x][y].re - m->e[y][x].re;
    r[3] = m->e[x][x].im - m->e[y][y].im;
}

/* rows p,q of U times columns p,q of F (device_PHB_OVR.h:385-400) */
static void block_proj_uf(const mat *U, const mat *F, int p, int q, double r[4]) {
    cx a0 = cmk(0, 0), a1 = cmk(0, 0), a2 = cmk(0, 0), a3 = cmk(0, 0);
    for (int j = 0; j < 3; j++) {
        a0 = cadd(a0, cmul(U->e[p][j], F->e[j][p]));
        a1 = cadd(a1, cmul(U->e[p][j], F->e[j][q]));
        a2 = cadd(a2, cmul(U->e[q][j], F->e[j][p]));
        a3 = cadd(a3, cmul(U->e[q][j], F->e[j][q]));
    }
    r[0] = a0.re + a3.re;
    r[1] = a1.im + a2.im;
    r[2] = a1.re - a2.re;
    r[3] = a0.im - a3.im;
}

/* device_PHB_OVR.h:361-485 */
/* counters (seed, c0, c2, c3) of the link; block b draws as bsrc_block describes */
static void heatbath_mat(int nc, double beta_over_nc, mat *U, const mat *F, uint64_t seed, uint32_t c0, uint32_t c2,
                         uint32_t c3, double *margin) {
    if (nc == 3) {
        for (int block = 0; block < 3; block++) {
            int p, q;
            index_block(3, block, &p, &q);
            double r[4], a[4], rr[4];
            block_proj_uf(U, F, p, q, r);
            double k = q_abs(r);
            double ap = beta_over_nc * k;
            k = 1.0 / k;
            for (int i = 0; i < 4; i++) r[i] *= k;
            bsrc st;
            bsrc_block(&st, seed, c0, c2, c3, block);
            gen_su2_milc(ap, &st, a, margin);
            q_mul_uv_dagger(a, r, rr);
            cx a0 = cmk(rr[0], rr[3]), a1 = cmk(rr[2], rr[1]);
            cx a2 = cmk(-rr[2], rr[1]), a3 = cmk(rr[0], -rr[3]);
            for (int j = 0; j < 3; j++) {
                cx tmp0 = cadd(cmul(a0, U->e[p][j]), cmul(a1, U->e[q][j]));
                U->e[q][j] = cadd(cmul(a2, U->e[p][j]), cmul(a3, U->e[q][j]));
                U->e[p][j] = tmp0;
            }
        }
        return;
    }
    mat M = mmul(nc, U, F);
    for (int block = 0; block < nc * (nc - 1) / 2; block++) {
        int x, y;
        index_block(nc, block, &x, &y);
        double r[4], a[4], rr[4];
        get_block_su2(&M, x, y, r);
        double k = q_abs(r);
        double ap = beta_over_nc * k;
        k = 1.0 / k;
        for (int i = 0; i < 4; i++) r[i] *= k;
        bsrc st;
        bsrc_block(&st, seed, c0, c2, c3, block);
        gen_su2_milc(ap, &st, a, margin);
        q_mul_uv_dagger(a, r, rr);
        mul_block_sun(nc, rr, U, x, y);
        mul_block_sun(nc, rr, &M, x, y);
    }
}

/* device_PHB_OVR.h:493-608 */
static void overrelax_mat(int nc, mat *U, const mat *F) {
    if (nc == 3) {
        for (int block = 0; block < 3; block++) {
            int p, q;
            index_block(3, block, &p, &q);
            double r[4];
            block_proj_uf(U, F, p, q, r);
            q_conj_normalize(r);
            cx a0 = cmk(r[0], r[3]), a1 = cmk(r[2], r[1]);
            cx a2 = cmk(-r[2], r[1]), a3 = cmk(r[0], -r[3]);
            for (int j = 0; j < 3; j++) {
                cx tmp0 = cadd(cmul(a0, U->e[p][j]), cmul(a1, U->e[q][j]));
                cx tmp1 = cadd(cmul(a2, U->e[p][j]), cmul(a3, U->e[q][j]));
                U->e[p][j] = cadd(cmul(a0, tmp0), cmul(a1, tmp1));
                U->e[q][j] = cadd(cmul(a2, tmp0), cmul(a3, tmp1));
            }
        }
        return;
    }
    mat M = mmul(nc, U, F);
    for (int block = 0; block < nc * (nc - 1) / 2; block++) {
        int x, y;
        index_block(nc, block, &x, &y);
        double r[4];
        get_block_su2(&M, x, y, r);
        q_conj_normalize(r);
        mul_block_sun(nc, r, U, x, y);
        mul_block_sun(nc, r, U, x, y);
        mul_block_sun(nc, r, &M, x, y);
        mul_block_sun(nc, r, &M, x, y);
    }
}

void or_overrelax_link(int nc, double *U, const double *F) {
    mat u, f;
    mload(nc, U, &u);
    mload(nc, F, &f);
    overrelax_mat(nc, &u, &f);
    mstore(nc, &u, U);
}

void or_heatbath_link(int nc, double beta, double *U, const double *F,
                      uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3) {
    mat u, f;
    mload(nc, U, &u);
    mload(nc, F, &f);
    heatbath_mat(nc, beta / (double)nc, &u, &f, seed, c0, c2, c3, NULL);
    mstore(nc, &u, U);
}

static int check_args(int nc, const int X[4], int storage) {
    if (nc < 2 || nc > NMAX) return -1;
    for (int i = 0; i < 4; i++)
        if (X[i] < 2 || (X[i] & 1)) return -2;
    if (storage != OR_STORE_18 && storage != OR_STORE_12 && storage != OR_STORE_8) return -3;
    if (nc != 3 && storage != OR_STORE_18) return -3;
    return 0;
}

/* monte.cu:44-88: one (parity, mu) heatbath launch.  RNG stream of the site:
 * key = seed, counter = (id, draw, sweep, TAG_HB | parity<<4 | mu). */
static int heatbath_sub(int nc, const int X[4], double beta, uint64_t seed, uint32_t sweep,
                        int storage, double *U, int parity, int mu, double *margins) {
    int rc = check_args(nc, X, storage);
    if (rc) return rc;
    long V = volume_of(X);
    long half = V / 2;
    double bnc = beta / (double)nc;   /* constants.cu:443 */
#pragma omp parallel for schedule(static)
    for (long id = 0; id < half; id++) {
        mat S = staple_at(nc, X, U, parity, mu, (int)id);
        mat F = mdag(nc, &S);
        const uint32_t tag = RNG_TAG_HB | ((uint32_t)parity << 4) | (uint32_t)mu;
        double *l = link_at_w(nc, U, V, mu, id + parity * half);
        mat u;
        mload(nc, l, &u);
        double *m = margins ? &margins[(long)mu * V + id + parity * half] : NULL;
        if (m) *m = INFINITY;
        heatbath_mat(nc, bnc, &u, &F, seed, (uint32_t)id, sweep, tag, m);
        mstore(nc, &u, l);
        project_link(nc, storage, l);
    }
    return 0;
}

int or_heatbath_sub(int nc, const int X[4], double beta, uint64_t seed, uint32_t sweep,
                    int storage, double *U, int parity, int mu) {
    return heatbath_sub(nc, X, beta, seed, sweep, storage, U, parity, mu, NULL);
}

/* ovr.cu:39-80 */
int or_overrelax_sub(int nc, const int X[4], int storage, double *U, int parity, int mu) {
    int rc = check_args(nc, X, storage);
    if (rc) return rc;
    long V = volume_of(X);
    long half = V / 2;
#pragma omp parallel for schedule(static)
    for (long id = 0; id < half; id++) {
        mat S = staple_at(nc, X, U, parity, mu, (int)id);
        mat F = mdag(nc, &S);
        double *l = link_at_w(nc, U, V, mu, id + parity * half);
        mat u;
        mload(nc, l, &u);
        overrelax_mat(nc, &u, &F);
        mstore(nc, &u, l);
        project_link(nc, storage, l);
    }
    return 0;
}

int or_heatbath_sweep(int nc, const int X[4], double beta, uint64_t seed, uint32_t sweep,
                      int storage, double *U) {
    for (int parity = 0; parity < 2; parity++)
        for (int mu = 0; mu < 4; mu++) {
            int rc = or_heatbath_sub(nc, X, beta, seed, sweep, storage, U, parity, mu);
            if (rc) return rc;
        }
    return 0;
}

int or_heatbath_sweep_margins(int nc, const int X[4], double beta, uint64_t seed, uint32_t sweep,
                              int storage, double *U, double *margins) {
    for (int parity = 0; parity < 2; parity++)
        for (int mu = 0; mu < 4; mu++) {
            int rc = heatbath_sub(nc, X, beta, seed, sweep, storage, U, parity, mu, margins);
            if (rc) return rc;
        }
    return 0;
}

int or_overrelax_sweep(int nc, const int X[4], int storage, double *U) {
    for (int parity = 0; parity < 2; parity++)
        for (int mu = 0; mu < 4; mu++) {
            int rc = or_overrelax_sub(nc, X, storage, U, parity, mu);
            if (rc) return rc;
        }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* observables; partial sums over fixed chunks of sites, then summed in order,  */
/* so the result does not depend on the thread count                          */
/* ------------------------------------------------------------------------- */
#define CHUNK 4096

void or_plaquette(int nc, const int X[4], const double *U, double out[2]) {
    long V = volume_of(X);
    long half = V / 2;
    long nchunk = (V + CHUNK - 1) / CHUNK;
    double *part = (double *)calloc((size_t)nchunk * 2, sizeof(double));
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nchunk; c++) {
        double sp = 0.0, tp = 0.0;
        long end = (c + 1) * CHUNK < V ? (c + 1) * CHUNK : V;
        for (long idd = c * CHUNK; idd < end; idd++) {
            int parity = idd >= half ? 1 : 0;
            int id = (int)(parity ? idd - half : idd);
            long idxoddbit = id + parity * half;
            for (int mu = 0; mu < 3; mu++) {
                mat link1, A, B, C, t, link;
                mload(nc, link_at(nc, U, V, mu, idxoddbit), &link1);
                int newidmu1 = or_neighbor_eo(id, parity, X, mu, 1, -1, 0);
                for (int nu = mu + 1; nu < 4; nu++) {
                    mload(nc, link_at(nc, U, V, nu, newidmu1), &A);
                    mload(nc, link_at(nc, U, V, mu, or_neighbor_eo(id, parity, X, nu, 1, -1, 0)), &B);
                    mload(nc, link_at(nc, U, V, nu, idxoddbit), &C);
                    t = mdag(nc, &B);
                    link = mmul(nc, &A, &t);
                    t = mdag(nc, &C);
                    link = mmul(nc, &link, &t);
                    t = mmul(nc, &link1, &link);
                    double tr = mrealtrace(nc, &t);
                    if (nu == 3) tp += tr; else sp += tr;
                }
            }
        }
        part[2 * c] = sp;
        part[2 * c + 1] = tp;
    }
    double s = 0.0, t = 0.0;
    for (long c = 0; c < nchunk; c++) { s += part[2 * c]; t += part[2 * c + 1]; }
    free(part);
    double norm = (double)(3 * nc) * (double)V;
    out[0] = s / norm;
    out[1] = t / norm;
}

/* Link monitors: linktrsum.cu:26-55,130-150 (sum of Tr U over all links / (4 N V))
 * and linkdetsum.cu:31-60,84-101 (sum of det U / (4 V)); out[4] = max over links of
 * max_ij |(U U^+)_ij - delta_ij|, out[5] = max over links of |det U - 1|. */
void or_link_monitor(int nc, long V, const double *U, double out[6]) {
    long nl = 4 * V;
    long nchunk = (nl + CHUNK - 1) / CHUNK;
    double *part = (double *)calloc((size_t)nchunk * 6, sizeof(double));
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nchunk; c++) {
        double *p = part + 6 * c;
        long end = (c + 1) * CHUNK < nl ? (c + 1) * CHUNK : nl;
        for (long l = c * CHUNK; l < end; l++) {
            mat a;
            mload(nc, U + (size_t)l * nc * nc * 2, &a);
            for (int i = 0; i < nc; i++) {
                p[0] += a.e[i][i].re;
                p[1] += a.e[i][i].im;
            }
            cx d = mdet(nc, &a);
            p[2] += d.re;
            p[3] += d.im;
            double dd = cabs2(cmk(d.re - 1.0, d.im));
            if (dd > p[5] || dd != dd) p[5] = dd;  /* a NaN stays */
            for (int i = 0; i < nc; i++)
                for (int j = 0; j < nc; j++) {
                    cx e = cmk(i == j ? -1.0 : 0.0, 0.0);
                    for (int k = 0; k < nc; k++) e = cadd(e, cmul(a.e[i][k], cconj(a.e[j][k])));
                    double ee = cabs2(e);
                    if (ee > p[4] || ee != ee) p[4] = ee;
                }
        }
    }
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (long c = 0; c < nchunk; c++) {
        for (int k = 0; k < 4; k++) acc[k] += part[6 * c + k];
        for (int k = 4; k < 6; k++)
            if (part[6 * c + k] > acc[k] || part[6 * c + k] != part[6 * c + k]) acc[k] = part[6 * c + k];
    }
    free(part);
    out[0] = acc[0] / (4.0 * nc * (double)V);
    out[1] = acc[1] / (4.0 * nc * (double)V);
    out[2] = acc[2] / (4.0 * (double)V);
    out[3] = acc[3] / (4.0 * (double)V);
    out[4] = sqrt(acc[4]);
    out[5] = sqrt(acc[5]);
}

void or_polyakov(int nc, const int X[4], const double *U, double out[2]) {
    long V = volume_of(X);
    long half = V / 2;
    long ts = (long)X[0] * X[1] * X[2];
    long nchunk = (ts + CHUNK - 1) / CHUNK;
    double *part = (double *)calloc((size_t)nchunk * 2, sizeof(double));
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nchunk; c++) {
        double sr = 0.0, si = 0.0;
        long end = (c + 1) * CHUNK < ts ? (c + 1) * CHUNK : ts;
        for (long idd = c * CHUNK; idd < end; idd++) {
            int parity = idd >= ts / 2 ? 1 : 0;
            int id = (int)(parity ? idd - ts / 2 : idd);
            int x[4];
            or_eo_to_coords(id, parity, X, x);
            long idx = ((long)x[2] * X[1] + x[1]) * X[0] + x[0];
            int idl = x[0] + x[1] + x[2];
            mat L, A;
            munit(nc, &L);
            for (int t = 0; t < X[3]; t++) {
                long id0 = (idx + (long)t * ts) >> 1;
                id0 += ((idl + t) & 1) * half;
                mload(nc, link_at(nc, U, V, 3, id0), &A);
                L = mmul(nc, &L, &A);
            }
            cx tr = mtrace(nc, &L);
            sr += tr.re;
            si += tr.im;
        }
        part[2 * c] = sr;
        part[2 * c + 1] = si;
    }
    double s = 0.0, t = 0.0;
    for (long c = 0; c < nchunk; c++) { s += part[2 * c]; t += part[2 * c + 1]; }
    free(part);
    double norm = (double)nc * (double)ts;
    out[0] = s / norm;
    out[1] = t / norm;
}

/* ------------------------------------------------------------------------- */
/* ------------------------------------------------------------------------- */
/* measurement chain: SU(N) projection, APE smearing, multihit, Wilson loops   */
/* ------------------------------------------------------------------------- */
/* Re Tr(a b^+) (matrixsun.h:934-950, realtraceUVdagger) */
static double retrace_uvdag(int nc, const mat *a, const mat *b) {
    double s = 0.0;
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) s += a->e[i][j].re * b->e[i][j].re + a->e[i][j].im * b->e[i][j].im;
    return s;
}

/* projectlink.h:16-64 (project_link): SU(2)-subgroup hits maximising Re Tr(w q^+);
 * with tol > 0 stop once the relative change of realtraceUVdagger(w, q^+) / N is
 * at most tol (the reference measures with the already daggered q). */
static void su_project(int nc, mat *w, const mat *q_in, int nhit, double tol) {
    mat q = mdag(nc, q_in);
    double conver = 1.0, old_tr = 0.0;
    if (tol > 0) old_tr = retrace_uvdag(nc, w, &q) / nc;
    int nblocks = nc * (nc - 1) / 2;
    for (int hit = 0; hit < nhit && conver > tol; hit++) {
        for (int block = 0; block < nblocks; block++) {
            int x, y;
            index_block(nc, block, &x, &y);
            mat action = mmul(nc, w, &q);
            double r[4];
            get_block_su2(&action, x, y, r);
            double k = q_abs(r);
            if (k == 0.0) {
                k = 1.0;
                r[0] = 1.0;
            }
            k = 1.0 / k;
            r[0] *= k;
            r[1] *= -k;
            r[2] *= -k;
            r[3] *= -k;
            mul_block_sun(nc, r, w, x, y);
        }
        if (tol > 0) {
            double new_tr = retrace_uvdag(nc, w, &q) / nc;
            conver = (new_tr - old_tr) / old_tr;
            old_tr = new_tr;
        }
    }
}

void or_project_link(int nc, double *w, const double *q, int nhit, double tol) {
    mat a, b;
    mload(nc, w, &a);
    mload(nc, q, &b);
    su_project(nc, &a, &b, nhit, tol);
    mstore(nc, &a, w);
}

/* SU3project (projectlink.h:119-139): w from the identity, the blocks of the
 * product V = w F^+ projected by SU2project (projectlink.h:68-94) and w <- V_b^+ w. */
void or_su3_project(double *F, int nhit) {
    mat f, w;
    mload(3, F, &f);
    munit(3, &w);
    su_project(3, &w, &f, nhit, 0.0);
    mstore(3, &w, F);
}

/* ape.cu:41-94 (kernel_APE_Space), one direction: link = (U + w S) / (1 + 6w) with S
 * the spatial staples, projected (N = 3: SU3project, else project_link from U). */
static void ape_dir(int nc, const int X[4], const double *Uin, double *Uout, int mu, double w, int nhit,
                    double tol) {
    long V = volume_of(X), half = V / 2;
#pragma omp parallel for schedule(static)
    for (long idd = 0; idd < V; idd++) {
        int parity = idd >= half ? 1 : 0;
        int id = (int)(idd - parity * half);
        mat S = staple_range(nc, X, Uin, parity, mu, id, 3);
        mat U, link, res;
        mload(nc, link_at(nc, Uin, V, mu, idd), &U);
        double d = 1.0 + 6.0 * w;
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nc; j++)
                link.e[i][j] = cmk((U.e[i][j].re + S.e[i][j].re * w) / d, (U.e[i][j].im + S.e[i][j].im * w) / d);
        if (nc == 3) {
            munit(3, &res);
            su_project(3, &res, &link, nhit, 0.0);
        } else {
            res = U;
            su_project(nc, &res, &link, nhit, tol);
        }
        mstore(nc, &res, link_at_w(nc, Uout, V, mu, idd));
    }
}

/* ape.cu:198-264: ApplyAPEinSpace (mu = 0, 1, 2 per step, every direction from the
 * field of the previous step) or ApplyAPEinTime (mu = 3) */
int or_ape_smear(int nc, const int X[4], double *U, double w, int steps, int nhit, double tol, int time_dir) {
    int rc = check_args(nc, X, OR_STORE_18);
    if (rc) return rc;
    long n = 4 * volume_of(X) * nc * nc * 2;
    double *out = (double *)malloc(sizeof(double) * (size_t)n);
    if (!out) return -4;
    memcpy(out, U, sizeof(double) * (size_t)n);
    for (int st = 0; st < steps; st++) {
        if (time_dir) {
            ape_dir(nc, X, U, out, 3, w, nhit, tol);
        } else {
            for (int mu = 0; mu < 3; mu++) ape_dir(nc, X, U, out, mu, w, nhit, tol);
        }
        memcpy(U, out, sizeof(double) * (size_t)n);
    }
    free(out);
    return 0;
}

/* exp(iQ) for traceless Hermitian 3x3 Q: Morningstar & Peardon, Phys. Rev. D 69 (2004)
 * 054501, eqs. (19)-(34): f0 + f1 Q + f2 Q^2, the f_j from c0 = Tr Q^3 / 3 and
 * c1 = Tr Q^2 / 2; for c1 < 4e-3 the exponential series reduced with Q^3 = c1 Q + c0
 * (order 16).  (stout.cu:94-466 computes the same function.) */
static mat exp_iq(const mat *Q) {
    mat QQ = mmul(3, Q, Q), QQQ = mmul(3, Q, &QQ);
    double c1 = 0.5 * mrealtrace(3, &QQ), c0 = mrealtrace(3, &QQQ) / 3.0;
    cx f0, f1, f2;
    if (c1 < 4.0e-3) {
        double a = 1.0, b = 0.0, c = 0.0;
        cx t = cmk(1.0, 0.0);
        f0 = f1 = f2 = cmk(0.0, 0.0);
        for (int k = 0; k <= 16; k++) {
            f0 = cadd(f0, cscale(t, a));
            f1 = cadd(f1, cscale(t, b));
            f2 = cadd(f2, cscale(t, c));
            double na = c * c0, nb = a + c * c1, nc = b;
            a = na;
            b = nb;
            c = nc;
            t = cmk(-t.im / (k + 1), t.re / (k + 1));
        }
    } else {
        double s3 = sqrt(c1 / 3.0), c0max = 2.0 * (c1 / 3.0) * s3;
        double ratio = fabs(c0) / c0max;
        double theta = acos(ratio < 1.0 ? ratio : 1.0);
        double u = s3 * cos(theta / 3.0), w = sqrt(c1) * sin(theta / 3.0);
        double u2 = u * u, w2 = w * w;
        double xi0 = fabs(w) < 0.05 ? 1.0 - w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0)) : sin(w) / w;
        double cw = cos(w);
        cx e2iu = cmk(cos(2.0 * u), sin(2.0 * u)), emiu = cmk(cos(u), -sin(u));
        cx h0 = cadd(cscale(e2iu, u2 - w2), cmul(emiu, cmk(8.0 * u2 * cw, 2.0 * u * (3.0 * u2 + w2) * xi0)));
        cx h1 = csub(cscale(e2iu, 2.0 * u), cmul(emiu, cmk(2.0 * u * cw, -(3.0 * u2 - w2) * xi0)));
        cx h2 = csub(e2iu, cmul(emiu, cmk(cw, 3.0 * u * xi0)));
        double den = 1.0 / (9.0 * u2 - w2);
        f0 = cscale(h0, den);
        f1 = cscale(h1, den);
        f2 = cscale(h2, den);
        if (c0 < 0) {
            f0 = cconj(f0);
            f1 = cmk(-f1.re, f1.im);
            f2 = cconj(f2);
        }
    }
    mat E;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            cx e = cadd(cmul(f1, Q->e[i][j]), cmul(f2, QQ.e[i][j]));
            E.e[i][j] = i == j ? cadd(e, f0) : e;
        }
    return E;
}

void or_exp_iq(const double *Q, double *E) {
    mat q, e;
    mload(3, Q, &q);
    e = exp_iq(&q);
    mstore(3, &e, E);
}

/* stout.cu:86-92,519-535,556-601: W = exp(iQ) U, Q the traceless part of
 * (i/2)(Omega^+ - Omega), Omega = rho S U^+, S the spatial staples */
static void stout_dir(const int X[4], const double *Uin, double *Uout, int mu, double rho) {
    long V = volume_of(X), half = V / 2;
#pragma omp parallel for schedule(static)
    for (long idd = 0; idd < V; idd++) {
        int parity = idd >= half ? 1 : 0;
        int id = (int)(idd - parity * half);
        mat S = staple_range(3, X, Uin, parity, mu, id, 3);
        mat U, Ud, Om, Q, E, W;
        mload(3, link_at(3, Uin, V, mu, idd), &U);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) S.e[i][j] = cscale(S.e[i][j], rho);
        Ud = mdag(3, &U);
        Om = mmul(3, &S, &Ud);
        cx tr = cmk(0.0, 0.0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                cx d = csub(cconj(Om.e[j][i]), Om.e[i][j]);
                Q.e[i][j] = cmk(-0.5 * d.im, 0.5 * d.re);
            }
        for (int i = 0; i < 3; i++) tr = cadd(tr, Q.e[i][i]);
        for (int i = 0; i < 3; i++) Q.e[i][i] = csub(Q.e[i][i], cscale(tr, 1.0 / 3.0));
        E = exp_iq(&Q);
        W = mmul(3, &E, &U);
        mstore(3, &W, link_at_w(3, Uout, V, mu, idd));
    }
}

/* stout.cu:687-766: ApplyStoutinSpace (mu = 0, 1, 2) / ApplyStoutinTime (mu = 3) */
int or_stout_smear(const int X[4], double *U, double rho, int steps, int time_dir) {
    int rc = check_args(3, X, OR_STORE_18);
    if (rc) return rc;
    long n = 4 * volume_of(X) * 18;
    double *out = (double *)malloc(sizeof(double) * (size_t)n);
    if (!out) return -4;
    memcpy(out, U, sizeof(double) * (size_t)n);
    for (int st = 0; st < steps; st++) {
        if (time_dir) {
            stout_dir(X, U, out, 3, rho);
        } else {
            for (int mu = 0; mu < 3; mu++) stout_dir(X, U, out, mu, rho);
        }
        memcpy(U, out, sizeof(double) * (size_t)n);
    }
    free(out);
    return 0;
}

/* HYP smearing, the reference's recursion restated (hyp.cu:51-187): decorated links
 * recomputed for every use.  Geometry through coordinates, periodic. */
typedef struct {
    int nc;
    const int *X;
    const double *U;
    long V;
    double a1, a2, a3, tol;
    int nhit;
} hyp_ctx;

static long hyp_eo(const hyp_ctx *h, const int x[4], int d1, int s1, int d2, int s2) {
    int y[4] = {x[0], x[1], x[2], x[3]};
    y[d1] = (y[d1] + s1 + h->X[d1]) % h->X[d1];
    if (d2 >= 0) y[d2] = (y[d2] + s2 + h->X[d2]) % h->X[d2];
    return or_coords_to_eo(y, h->X);
}

static mat hyp_link(const hyp_ctx *h, long eo, int mu) {
    mat m;
    mload(h->nc, link_at(h->nc, h->U, h->V, mu, eo), &m);
    return m;
}

/* hyp.cu:52-79 (CalcStaple) + 82-103 (HYPSmearVbar) */
static mat hyp_vbar(const hyp_ctx *h, const int x[4], int mu, int nu, int rho) {
    int nc = h->nc, eta = 6 - mu - nu - rho;
    mat a = hyp_link(h, hyp_eo(h, x, 0, 0, -1, 0), eta), b = hyp_link(h, hyp_eo(h, x, eta, 1, -1, 0), mu);
    mat c = hyp_link(h, hyp_eo(h, x, mu, 1, -1, 0), eta), t, st, dn;
    st = mmul(nc, &a, &b);
    t = mdag(nc, &c);
    st = mmul(nc, &st, &t);
    a = hyp_link(h, hyp_eo(h, x, eta, -1, -1, 0), eta);
    b = hyp_link(h, hyp_eo(h, x, eta, -1, -1, 0), mu);
    c = hyp_link(h, hyp_eo(h, x, eta, -1, mu, 1), eta);
    t = mdag(nc, &a);
    dn = mmul(nc, &t, &b);
    dn = mmul(nc, &dn, &c);
    madd_to(nc, &st, &dn);
    mat w = hyp_link(h, hyp_eo(h, x, 0, 0, -1, 0), mu), v;
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) v.e[i][j] = cadd(cscale(st.e[i][j], 0.5 * h->a3), cscale(w.e[i][j], 1.0 - h->a3));
    su_project(nc, &w, &v, h->nhit, h->tol);
    return w;
}

static void hyp_shift(const hyp_ctx *h, const int x[4], int d1, int s1, int d2, int s2, int y[4]) {
    for (int i = 0; i < 4; i++) y[i] = x[i];
    y[d1] = (y[d1] + s1 + h->X[d1]) % h->X[d1];
    if (d2 >= 0) y[d2] = (y[d2] + s2 + h->X[d2]) % h->X[d2];
}

/* hyp.cu:105-134 (HYPSmearVtil) */
static mat hyp_vtil(const hyp_ctx *h, const int x[4], int mu, int nu) {
    int nc = h->nc, y[4];
    mat vt, l, t;
    mzero(nc, &vt);
    for (int rho = 0; rho < 4; rho++) {
        if (rho == mu || rho == nu) continue;
        l = hyp_vbar(h, x, rho, nu, mu);
        hyp_shift(h, x, rho, 1, -1, 0, y);
        t = hyp_vbar(h, y, mu, rho, nu);
        l = mmul(nc, &l, &t);
        hyp_shift(h, x, mu, 1, -1, 0, y);
        t = hyp_vbar(h, y, rho, nu, mu);
        t = mdag(nc, &t);
        l = mmul(nc, &l, &t);
        madd_to(nc, &vt, &l);
        hyp_shift(h, x, rho, -1, -1, 0, y);
        t = hyp_vbar(h, y, rho, nu, mu);
        l = mdag(nc, &t);
        t = hyp_vbar(h, y, mu, rho, nu);
        l = mmul(nc, &l, &t);
        hyp_shift(h, x, rho, -1, mu, 1, y);
        t = hyp_vbar(h, y, rho, nu, mu);
        l = mmul(nc, &l, &t);
        madd_to(nc, &vt, &l);
    }
    mat w = hyp_link(h, hyp_eo(h, x, 0, 0, -1, 0), mu), v;
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) v.e[i][j] = cadd(cscale(vt.e[i][j], 0.25 * h->a2), cscale(w.e[i][j], 1.0 - h->a2));
    su_project(nc, &w, &v, h->nhit, h->tol);
    return w;
}

/* hyp.cu:136-167 (HYPSmearV) */
static mat hyp_v(const hyp_ctx *h, const int x[4], int mu) {
    int nc = h->nc, y[4];
    mat v, l, t;
    mzero(nc, &v);
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        l = hyp_vtil(h, x, nu, mu);
        hyp_shift(h, x, nu, 1, -1, 0, y);
        t = hyp_vtil(h, y, mu, nu);
        l = mmul(nc, &l, &t);
        hyp_shift(h, x, mu, 1, -1, 0, y);
        t = hyp_vtil(h, y, nu, mu);
        t = mdag(nc, &t);
        l = mmul(nc, &l, &t);
        madd_to(nc, &v, &l);
        hyp_shift(h, x, nu, -1, -1, 0, y);
        t = hyp_vtil(h, y, nu, mu);
        l = mdag(nc, &t);
        t = hyp_vtil(h, y, mu, nu);
        l = mmul(nc, &l, &t);
        hyp_shift(h, x, nu, -1, mu, 1, y);
        t = hyp_vtil(h, y, nu, mu);
        l = mmul(nc, &l, &t);
        madd_to(nc, &v, &l);
    }
    mat w = hyp_link(h, hyp_eo(h, x, 0, 0, -1, 0), mu), vv;
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) vv.e[i][j] = cadd(cscale(v.e[i][j], h->a1 / 6.0), cscale(w.e[i][j], 1.0 - h->a1));
    su_project(nc, &w, &vv, h->nhit, h->tol);
    return w;
}

/* hyp.cu:277-368: ApplyHYPinSpace (mu = 0, 1, 2) / ApplyHYPinTime (mu = 3) */
int or_hyp_smear(int nc, const int X[4], double *U, double a1, double a2, double a3, int steps, int nhit, double tol,
                 int time_dir) {
    int rc = check_args(nc, X, OR_STORE_18);
    if (rc) return rc;
    long V = volume_of(X), half = V / 2;
    long n = 4 * V * nc * nc * 2;
    double *out = (double *)malloc(sizeof(double) * (size_t)n);
    if (!out) return -4;
    memcpy(out, U, sizeof(double) * (size_t)n);
    hyp_ctx h = {nc, X, U, V, a1, a2, a3, tol, nhit};
    for (int st = 0; st < steps; st++) {
        for (int mu = time_dir ? 3 : 0; mu < (time_dir ? 4 : 3); mu++) {
#pragma omp parallel for schedule(dynamic, 16)
            for (long idd = 0; idd < V; idd++) {
                int parity = idd >= half ? 1 : 0, x[4];
                or_eo_to_coords((int)(idd - parity * half), parity, X, x);
                mat w = hyp_v(&h, x, mu);
                mstore(nc, &w, link_at_w(nc, out, V, mu, idd));
            }
        }
        memcpy(U, out, sizeof(double) * (size_t)n);
    }
    free(out);
    return 0;
}

/* multihit.cu:112-147 (kernel_evenodd): every temporal link replaced by the mean of
 * itself and nhit successive heatbath updates in its (fixed) staple, all from the
 * input field; spatial links copied.  Hit k of the link (parity, id) draws from the
 * stream (seed, id, k, TAG_MH | parity<<4 | 3). */
static int multihit(int nc, const int X[4], double beta, uint64_t seed, int nhit, const double *Uin, double *Uout,
                    double *margins) {
    int rc = check_args(nc, X, OR_STORE_18);
    if (rc) return rc;
    long V = volume_of(X), half = V / 2;
    long n = 4 * V * nc * nc * 2;
    memcpy(Uout, Uin, sizeof(double) * (size_t)n);
    double bnc = beta / (double)nc;
#pragma omp parallel for schedule(static)
    for (long idd = 0; idd < V; idd++) {
        int parity = idd >= half ? 1 : 0;
        int id = (int)(idd - parity * half);
        mat S = staple_at(nc, X, Uin, parity, 3, id);
        mat F = mdag(nc, &S);
        mat u, acc;
        mload(nc, link_at(nc, Uin, V, 3, idd), &u);
        acc = u;
        double *m = margins ? &margins[3 * V + idd] : NULL;
        if (m) *m = INFINITY;
        for (int hit = 0; hit < nhit; hit++) {
            heatbath_mat(nc, bnc, &u, &F, seed, (uint32_t)id, (uint32_t)hit, RNG_TAG_MH | ((uint32_t)parity << 4) | 3u, m);
            madd_to(nc, &acc, &u);
        }
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nc; j++)
                acc.e[i][j] = cmk(acc.e[i][j].re / (double)(nhit + 1), acc.e[i][j].im / (double)(nhit + 1));
        mstore(nc, &acc, link_at_w(nc, Uout, V, 3, idd));
    }
    return 0;
}

int or_multihit(int nc, const int X[4], double beta, uint64_t seed, int nhit, const double *Uin, double *Uout) {
    return multihit(nc, X, beta, seed, nhit, Uin, Uout, NULL);
}

int or_multihit_margins(int nc, const int X[4], double beta, uint64_t seed, int nhit, const double *Uin,
                        double *Uout, double *margins) {
    return multihit(nc, X, beta, seed, nhit, Uin, Uout, margins);
}

/* multihitsp.cu:115-161: spatial links -> mean of nhit successive overrelaxation
 * steps in the spatial staple; temporal links copied */
int or_multihit_space(int nc, const int X[4], int nhit, const double *Uin, double *Uout) {
    int rc = check_args(nc, X, OR_STORE_18);
    if (rc) return rc;
    if (nhit < 1) return -5;
    long V = volume_of(X), half = V / 2;
    memcpy(Uout, Uin, sizeof(double) * (size_t)(4 * V * nc * nc * 2));
#pragma omp parallel for schedule(static)
    for (long idd = 0; idd < V; idd++) {
        int parity = idd >= half ? 1 : 0;
        int id = (int)(idd - parity * half);
        for (int mu = 0; mu < 3; mu++) {
            mat S = staple_range(nc, X, Uin, parity, mu, id, 3);
            mat F = mdag(nc, &S);
            mat u, acc;
            mload(nc, link_at(nc, Uin, V, mu, idd), &u);
            mzero(nc, &acc);
            for (int hit = 0; hit < nhit; hit++) {
                overrelax_mat(nc, &u, &F);
                madd_to(nc, &acc, &u);
            }
            for (int i = 0; i < nc; i++)
                for (int j = 0; j < nc; j++) acc.e[i][j] = cmk(acc.e[i][j].re / nhit, acc.e[i][j].im / nhit);
            mstore(nc, &acc, link_at_w(nc, Uout, V, mu, idd));
        }
    }
    return 0;
}

/* multihitext.cu:241-330 (kernel_multihitext) with :395-459 (ApplyMultiHitExt: mu = 3,
 * LEN_SPACE 2, LEN_TIME 1).  For every site x, on a copy of the input: the 27 temporal
 * links of the spatial box x + (i, j, k, 0), |i|, |j|, |k| <= 1 (time slice of x),
 * receive nhit checkerboard heatbath sweeps -- parity (i + j + k) & 1 first 0 then 1,
 * box order i fastest, as the reference's loops -- in staples whose other links keep
 * their input values; the output temporal link at x is the mean of the centre link
 * after each sweep.  Spatial links copied.  Box link b = (i+1) + 3 (j+1) + 9 (k+1)
 * of site idd (even/odd index) in hit h draws from stream (seed, idd, h, TAG_MHX | b).
 * The reference runs the sites in 128 launches over a halo-extended copy refreshed
 * before each: a site's result depends only on the input and its own random numbers. */
static int mhx_in_box(const int o[3]) {
    return o[0] >= -1 && o[0] <= 1 && o[1] >= -1 && o[1] <= 1 && o[2] >= -1 && o[2] <= 1;
}
static int mhx_index(const int o[3]) { return (o[0] + 1) + 3 * (o[1] + 1) + 9 * (o[2] + 1); }

static int multihit_ext(int nc, const int X[4], double beta, uint64_t seed, int nhit, const double *Uin,
                        double *Uout, double *margins) {
    int rc = check_args(nc, X, OR_STORE_18);
    if (rc) return rc;
    if (nhit < 1) return -5;
    long V = volume_of(X), half = V / 2;
    memcpy(Uout, Uin, sizeof(double) * (size_t)(4 * V * nc * nc * 2));
    double bnc = beta / (double)nc;
#pragma omp parallel for schedule(static)
    for (long idd = 0; idd < V; idd++) {
        int parity = idd >= half ? 1 : 0;
        int x[4];
        or_eo_to_coords((int)(idd - parity * half), parity, X, x);
        mat box[27];
        for (int b = 0; b < 27; b++) {
            int y[4] = {(x[0] + b % 3 - 1 + X[0]) % X[0], (x[1] + (b / 3) % 3 - 1 + X[1]) % X[1],
                        (x[2] + b / 9 - 1 + X[2]) % X[2], x[3]};
            mload(nc, link_at(nc, Uin, V, 3, or_coords_to_eo(y, X)), &box[b]);
        }
        mat acc;
        mzero(nc, &acc);
        double *m = margins ? &margins[3 * V + idd] : NULL;
        if (m) *m = INFINITY;
        for (int hit = 0; hit < nhit; hit++) {
            for (int par = 0; par < 2; par++)
                for (int b = 0; b < 27; b++) {
                    int o[3] = {b % 3 - 1, (b / 3) % 3 - 1, b / 9 - 1};
                    if (((o[0] + o[1] + o[2]) & 1) != par) continue;
                    int y[4] = {x[0] + o[0], x[1] + o[1], x[2] + o[2], x[3]};
                    mat staple, link, t, A, B, C;
                    mzero(nc, &staple);
                    for (int nu = 0; nu < 3; nu++) {
                        int yp[4], yt[4], ym[4], ymt[4];
                        for (int d = 0; d < 4; d++) yp[d] = yt[d] = ym[d] = ymt[d] = y[d];
                        yp[nu] += 1;
                        yt[3] += 1;
                        ym[nu] -= 1;
                        ymt[nu] -= 1;
                        ymt[3] += 1;
                        int op[3] = {o[0], o[1], o[2]}, om[3] = {o[0], o[1], o[2]};
                        op[nu] += 1;
                        om[nu] -= 1;
                        int *pts[4] = {yp, yt, ym, ymt};
                        for (int q = 0; q < 4; q++)
                            for (int d = 0; d < 4; d++) pts[q][d] = (pts[q][d] + X[d]) % X[d];
                        int yy[4] = {(y[0] + X[0]) % X[0], (y[1] + X[1]) % X[1], (y[2] + X[2]) % X[2], y[3]};
                        /* up: U_nu(y) U_3(y+nu) U_nu(y+3)^dagger */
                        mload(nc, link_at(nc, Uin, V, nu, or_coords_to_eo(yy, X)), &A);
                        if (mhx_in_box(op)) B = box[mhx_index(op)];
                        else mload(nc, link_at(nc, Uin, V, 3, or_coords_to_eo(yp, X)), &B);
                        mload(nc, link_at(nc, Uin, V, nu, or_coords_to_eo(yt, X)), &C);
                        link = mmul(nc, &A, &B);
                        t = mdag(nc, &C);
                        link = mmul(nc, &link, &t);
                        madd_to(nc, &staple, &link);
                        /* down: U_nu(y-nu)^dagger U_3(y-nu) U_nu(y-nu+3) */
                        mload(nc, link_at(nc, Uin, V, nu, or_coords_to_eo(ym, X)), &A);
                        if (mhx_in_box(om)) B = box[mhx_index(om)];
                        else mload(nc, link_at(nc, Uin, V, 3, or_coords_to_eo(ym, X)), &B);
                        mload(nc, link_at(nc, Uin, V, nu, or_coords_to_eo(ymt, X)), &C);
                        link = mdag(nc, &A);
                        link = mmul(nc, &link, &B);
                        link = mmul(nc, &link, &C);
                        madd_to(nc, &staple, &link);
                    }
                    mat F = mdag(nc, &staple);
                    heatbath_mat(nc, bnc, &box[b], &F, seed, (uint32_t)idd, (uint32_t)hit, RNG_TAG_MHX | (uint32_t)b, m);
                }
            madd_to(nc, &acc, &box[13]);
        }
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nc; j++)
                acc.e[i][j] = cmk(acc.e[i][j].re / (double)nhit, acc.e[i][j].im / (double)nhit);
        mstore(nc, &acc, link_at_w(nc, Uout, V, 3, idd));
    }
    return 0;
}

int or_multihit_ext(int nc, const int X[4], double beta, uint64_t seed, int nhit, const double *Uin, double *Uout) {
    return multihit_ext(nc, X, beta, seed, nhit, Uin, Uout, NULL);
}

int or_multihit_ext_margins(int nc, const int X[4], double beta, uint64_t seed, int nhit, const double *Uin,
                            double *Uout, double *margins) {
    return multihit_ext(nc, X, beta, seed, nhit, Uin, Uout, margins);
}

/* wilsonloop.cu:44-121,494-516 (WilsonLineSP + WilsonLoopSP): W(r, T) = sum over x and
 * mu < 3 of Tr(L_r(x) t1 L_r(x + T t)^+ t0^+) / (3 N V), L_r(x) the product of r links
 * along mu from x, t0 (t1) the product of T temporal links from x (x + r mu).
 * out[2 (T + r (tmax + 1)) + {0, 1}]. */
int or_wilson_loops(int nc, const int X[4], const double *U, int rmax, int tmax, double *out) {
    int rc = check_args(nc, X, OR_STORE_18);
    if (rc) return rc;
    long V = volume_of(X);
    int nrt = (rmax + 1) * (tmax + 1);
    long nchunk = (V + CHUNK - 1) / CHUNK;
    double *part = (double *)calloc((size_t)nchunk * nrt * 2, sizeof(double));
    if (!part) return -4;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nchunk; c++) {
        double *p = part + (size_t)c * nrt * 2;
        long end = (c + 1) * CHUNK < V ? (c + 1) * CHUNK : V;
        for (long s = c * CHUNK; s < end; s++) {
            int parity = s >= V / 2 ? 1 : 0;
            int x[4];
            or_eo_to_coords((int)(s - parity * (V / 2)), parity, X, x);
            for (int mu = 0; mu < 3; mu++) {
                for (int r = 0; r <= rmax; r++) {
                    mat t0, t1;
                    munit(nc, &t0);
                    munit(nc, &t1);
                    for (int T = 0; T <= tmax; T++) {
                        mat lb, lt, a, m;
                        munit(nc, &lb);
                        munit(nc, &lt);
                        for (int k = 0; k < r; k++) {
                            int y[4] = {x[0], x[1], x[2], x[3]};
                            y[mu] = (y[mu] + k) % X[mu];
                            mload(nc, link_at(nc, U, V, mu, or_coords_to_eo(y, X)), &a);
                            lb = mmul(nc, &lb, &a);
                            y[3] = (y[3] + T) % X[3];
                            mload(nc, link_at(nc, U, V, mu, or_coords_to_eo(y, X)), &a);
                            lt = mmul(nc, &lt, &a);
                        }
                        m = mmul(nc, &lb, &t1);
                        a = mdag(nc, &lt);
                        m = mmul(nc, &m, &a);
                        a = mdag(nc, &t0);
                        m = mmul(nc, &m, &a);
                        cx tr = mtrace(nc, &m);
                        p[2 * (T + r * (tmax + 1))] += tr.re;
                        p[2 * (T + r * (tmax + 1)) + 1] += tr.im;
                        if (T < tmax) {
                            int y[4] = {x[0], x[1], x[2], (x[3] + T) % X[3]};
                            mload(nc, link_at(nc, U, V, 3, or_coords_to_eo(y, X)), &a);
                            t0 = mmul(nc, &t0, &a);
                            y[mu] = (y[mu] + r) % X[mu];
                            mload(nc, link_at(nc, U, V, 3, or_coords_to_eo(y, X)), &a);
                            t1 = mmul(nc, &t1, &a);
                        }
                    }
                }
            }
        }
    }
    double norm = 3.0 * nc * (double)V;
    for (int k = 0; k < 2 * nrt; k++) {
        double acc = 0.0;
        for (long c = 0; c < nchunk; c++) acc += part[(size_t)c * nrt * 2 + k];
        out[k] = acc / norm;
    }
    free(part);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* operator-basis Wilson loops: the operator fields of src/wl/calcop_dg_A0.cu   */
/* (with include/mop/mo0.cuh) and src/wl/calcop_dg_33.cu, and the correlator of  */
/* src/wl/wilsonloop_dg_A0.cu / wilsonloop_dg.cu                                */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nc;
    const int *X;
    long V;
    const double *U;
} lctx;

/* one step of a path at y: code = +(d+1) takes U_d(y) and moves to y + d; code =
 * -(d+1) moves to y - d and takes U_d(y - d)^+ */
static mat path_step(const lctx *c, int y[4], int code) {
    mat a;
    int d = (code > 0 ? code : -code) - 1;
    if (code < 0) y[d] = (y[d] - 1 + c->X[d]) % c->X[d];
    mload(c->nc, link_at(c->nc, c->U, c->V, d, or_coords_to_eo(y, c->X)), &a);
    if (code > 0) {
        y[d] = (y[d] + 1) % c->X[d];
        return a;
    }
    return mdag(c->nc, &a);
}

/* ordered product of the steps st[0..n) from x, left to right; it starts from the
 * identity when ident (mo0.cuh:22, calcop_dg_A0.cu:60), from the first link
 * otherwise (calcop_dg_33.cu: link0 = GAUGE_LOAD(...) then *=) */
static mat path_walk(const lctx *c, const int x[4], const int *st, int n, int ident) {
    int y[4] = {x[0], x[1], x[2], x[3]};
    mat m, a;
    if (ident) munit(c->nc, &m);
    for (int k = 0; k < n; k++) {
        a = path_step(c, y, st[k]);
        if (k == 0 && !ident)
            m = a;
        else
            m = mmul(c->nc, &m, &a);
    }
    return m;
}

/* append `count` copies of `code` */
static int path_seg(int *st, int n, int code, int count) {
    for (int k = 0; k < count; k++) st[n++] = code;
    return n;
}

/* acc += i^q L (q = 0..3), the reference's  mop += link / += timesI(link) / -= link
 * / -= timesI(link)  (matrixsun.h:968-975: timesI(a) = (-Im a, Re a)) */
static void acc_phase(int nc, mat *acc, const mat *L, int q) {
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) {
            cx a = acc->e[i][j], l = L->e[i][j];
            if (q == 0) a = cadd(a, l);
            else if (q == 1) a = cmk(a.re + (-l.im), a.im + l.re);
            else if (q == 2) a = csub(a, l);
            else a = cmk(a.re - (-l.im), a.im - l.re);
            acc->e[i][j] = a;
        }
}

static void mat_div(int nc, mat *m, double s) {
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) m->e[i][j] = cmk(m->e[i][j].re / s, m->e[i][j].im / s);
}
static void mat_mulr(int nc, mat *m, double s) {
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) m->e[i][j] = cmk(m->e[i][j].re * s, m->e[i][j].im * s);
}

/* Phase tables of the 33 family: q[p][k] is the power of i with which path p enters
 * output k (the first path of a family initialises every output).  Paths of family
 * I and III: p = 4 part + 2 (down) + il; family II: p = 8 part + 2 type + il, type 0
 * up-forward, 1 up-backward, 2 down-forward, 3 down-backward; il = 0 steps first
 * along mu+1, il = 1 along mu+2 (mod 3).  calcop_dg_33.cu:150-260 (I), :319-672
 * (II), :734-852 (III). */
static const signed char WLOP_Q1[8][6] = {
    {0, 0, 0, 0, 0, 0}, {0, 1, 2, 0, 1, 2}, {0, 2, 0, 0, 2, 0}, {0, 3, 2, 0, 3, 2},
    {0, 0, 0, 2, 2, 2}, {0, 1, 2, 2, 3, 0}, {0, 2, 0, 2, 0, 2}, {0, 3, 2, 2, 1, 0}};
static const signed char WLOP_Q2[16][12] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 2, 0, 2, 1, 3, 1, 3, 2, 0, 2, 0},
    {0, 2, 0, 2, 0, 2, 0, 2, 0, 2, 0, 2}, {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2},
    {0, 2, 0, 2, 2, 0, 2, 0, 0, 2, 0, 2}, {0, 0, 0, 0, 3, 3, 3, 3, 2, 2, 2, 2},
    {0, 0, 0, 0, 2, 2, 2, 2, 0, 0, 0, 0}, {0, 2, 0, 2, 3, 1, 3, 1, 2, 0, 2, 0},
    {0, 0, 2, 2, 0, 0, 2, 2, 0, 0, 2, 2}, {0, 2, 2, 0, 1, 3, 3, 1, 0, 0, 0, 2},
    {0, 2, 2, 0, 0, 2, 2, 0, 2, 2, 2, 0}, {0, 0, 2, 2, 1, 1, 3, 3, 2, 2, 0, 0},
    {0, 2, 2, 0, 2, 0, 0, 2, 0, 2, 2, 0}, {0, 0, 2, 2, 3, 3, 1, 1, 2, 2, 0, 0},
    {0, 0, 2, 2, 2, 2, 0, 0, 0, 0, 2, 2}, {0, 2, 2, 0, 3, 1, 1, 3, 2, 0, 0, 2}};
static const signed char WLOP_Q3[8][6] = {
    {0, 0, 0, 0, 0, 0}, {0, 0, 1, 1, 2, 2}, {0, 0, 2, 2, 0, 0}, {0, 0, 3, 3, 2, 2},
    {0, 2, 0, 2, 0, 2}, {0, 2, 1, 3, 2, 0}, {0, 2, 2, 0, 0, 2}, {0, 2, 3, 1, 2, 0}};
/* output k of each family -> operator slot (calcop_dg_33.cu:263-268, :680-690, :853-858) */
static const int WLOP_SLOT1[6] = {1, 9, 17, 5, 13, 21};
static const int WLOP_SLOT2[12] = {2, 4, 6, 8, 10, 11, 14, 15, 18, 19, 22, 23};
static const int WLOP_SLOT3[6] = {3, 7, 12, 16, 20, 24};

static void wlop_accumulate(int nc, mat *acc, int nout, const signed char *q, int p, const mat *L) {
    for (int k = 0; k < nout; k++) {
        if (p == 0)
            acc[k] = *L;
        else
            acc_phase(nc, &acc[k], L, q[k]);
    }
}

/* the 25 operators of calcop_dg_33.cu:95-860 at site x (radius r >= 1, direction mu) */
static void wlop33_site(const lctx *c, const int x[4], int r, int mu, mat *out) {
    const int nc = c->nc, M = mu + 1;
    const int h = r / 2, h2 = (r + 1) / 2;
    const int dd[2] = {(mu + 1) % 3, (mu + 2) % 3};
    int xh[4] = {x[0], x[1], x[2], x[3]}, xs2[4] = {x[0], x[1], x[2], x[3]};
    xh[mu] = (xh[mu] + h) % c->X[mu];
    xs2[mu] = (xs2[mu] + h2) % c->X[mu];
    int st[256];
    int n;
    mat acc[12], link0, link, a;
    const double rt8 = 2.0 * sqrt(2.0);

    /* slot 0: the straight line from the identity (calcop_dg_33.cu:103-113) */
    n = path_seg(st, 0, M, r);
    out[0] = path_walk(c, x, st, n, 1);

    /* lines: line1 = links h..r-1 along mu (from x + h mu), line2 = links 0..h2-1 */
    n = path_seg(st, 0, M, r - h);
    const mat line1 = path_walk(c, xh, st, n, 0);
    n = path_seg(st, 0, M, h2);
    const mat line2 = path_walk(c, x, st, n, 0);

    /* family I: a unit staple over the first (last) half of the line */
    for (int part = 0; part < 2; part++)
        for (int down = 0; down < 2; down++)
            for (int il = 0; il < 2; il++) {
                const int D = dd[il] + 1, s = down ? -1 : 1;
                n = path_seg(st, 0, s * D, 1);
                n = path_seg(st, n, M, part ? r - h2 : h);
                n = path_seg(st, n, -s * D, 1);
                link0 = path_walk(c, part ? xs2 : x, st, n, 0);
                link = part ? mmul(nc, &line2, &link0) : mmul(nc, &link0, &line1);
                int p = 4 * part + 2 * down + il;
                wlop_accumulate(nc, acc, 6, WLOP_Q1[p], p, &link);
            }
    for (int k = 0; k < 6; k++) {
        mat_div(nc, &acc[k], rt8);
        out[WLOP_SLOT1[k]] = acc[k];
    }

    /* family II: the staple's top leg displaced one step along the other transverse
     * direction (forward or backward) */
    for (int part = 0; part < 2; part++)
        for (int type = 0; type < 4; type++)
            for (int il = 0; il < 2; il++) {
                const int D = dd[il] + 1, E = dd[1 - il] + 1;
                const int sd = (type < 2) ? 1 : -1, se = (type & 1) ? -1 : 1;
                n = path_seg(st, 0, sd * D, 1);
                n = path_seg(st, n, se * E, 1);
                n = path_seg(st, n, M, part ? r - h2 : h);
                n = path_seg(st, n, -se * E, 1);
                n = path_seg(st, n, -sd * D, 1);
                link0 = path_walk(c, part ? xs2 : x, st, n, 0);
                link = part ? mmul(nc, &line2, &link0) : mmul(nc, &link0, &line1);
                int p = 8 * part + 2 * type + il;
                wlop_accumulate(nc, acc, 12, WLOP_Q2[p], p, &link);
            }
    for (int k = 0; k < 12; k++) {
        mat_mulr(nc, &acc[k], 0.25);
        out[WLOP_SLOT2[k]] = acc[k];
    }

    /* family III: the staple run backwards and forwards over half the line */
    n = path_seg(st, 0, M, h > 0 ? h : 1);
    const mat line0a = path_walk(c, x, st, n, 0);
    n = path_seg(st, 0, M, r - h2 > 0 ? r - h2 : 1);
    const mat line0b = path_walk(c, xs2, st, n, 0);
    for (int part = 0; part < 2; part++)
        for (int down = 0; down < 2; down++)
            for (int il = 0; il < 2; il++) {
                const int D = dd[il] + 1, s = down ? -1 : 1;
                n = path_seg(st, 0, s * D, 1);
                n = path_seg(st, n, M, part ? r - h2 : h);
                n = path_seg(st, n, -s * D, 1);
                link0 = path_walk(c, part ? xs2 : x, st, n, 0);
                mat l0d = mdag(nc, &link0);
                if (!part) {
                    a = mmul(nc, &line0a, &l0d);
                    a = mmul(nc, &a, &line0a);
                    link = mmul(nc, &a, &line1);
                } else {
                    a = mmul(nc, &line2, &line0b);
                    a = mmul(nc, &a, &l0d);
                    link = mmul(nc, &a, &line0b);
                }
                int p = 4 * part + 2 * down + il;
                wlop_accumulate(nc, acc, 6, WLOP_Q3[p], p, &link);
            }
    for (int k = 0; k < 6; k++) {
        mat_div(nc, &acc[k], rt8);
        out[WLOP_SLOT3[k]] = acc[k];
    }
}

/* the 10 operators of calcop_dg_A0.cu:46-135 at site x: slot 0 the straight line,
 * slots 1..8 MO0(l = 1..8) (mo0.cuh:12-60: the four staples of height l over the line
 * in the two transverse spatial directions, / 2), slot 9 zero (the field is cleared
 * before the kernel, calcop_dg_A0.cu:241, and nothing writes that slot) */
static void wlopA0_site(const lctx *c, const int x[4], int r, int mu, mat *out) {
    const int nc = c->nc, M = mu + 1;
    const int dd[2] = {(mu + 1) % 3, (mu + 2) % 3};
    int st[256];
    int n = path_seg(st, 0, M, r);
    out[0] = path_walk(c, x, st, n, 1);
    for (int l = 1; l <= 8; l++) {
        mat acc, link0;
        mzero(nc, &acc);
        for (int down = 0; down < 2; down++)
            for (int il = 0; il < 2; il++) {
                const int D = dd[il] + 1, s = down ? -1 : 1;
                n = path_seg(st, 0, s * D, l);
                n = path_seg(st, n, M, r);
                n = path_seg(st, n, -s * D, l);
                link0 = path_walk(c, x, st, n, 1);
                madd_to(nc, &acc, &link0);
            }
        mat_div(nc, &acc, 2.0);
        out[l] = acc;
    }
    mzero(nc, &out[9]);
}

int or_wlop_fields(int nc, const int X[4], const double *U, int family, int r, int mu, double *ops) {
    int rc = check_args(nc, X, OR_STORE_18);
    if (rc) return rc;
    if (family != OR_WLOP_A0 && family != OR_WLOP_33) return -3;
    if (mu < 0 || mu > 2 || r < 1 || r > X[mu] || r > 64) return -2;
    const long V = volume_of(X);
    const int nop = family == OR_WLOP_A0 ? 10 : 25;
    const lctx c = {nc, X, V, U};
#pragma omp parallel for schedule(static)
    for (long s = 0; s < V; s++) {
        int parity = s >= V / 2 ? 1 : 0;
        int x[4];
        mat out[25];
        or_eo_to_coords((int)(s - parity * (V / 2)), parity, X, x);
        if (family == OR_WLOP_A0)
            wlopA0_site(&c, x, r, mu, out);
        else
            wlop33_site(&c, x, r, mu, out);
        for (int k = 0; k < nop; k++) mstore(nc, &out[k], ops + ((size_t)k * V + s) * nc * nc * 2);
    }
    return 0;
}

int or_wlop_correlate(int nc, const int X[4], const double *U, const double *ops, int nop, int r, int mu, int tmax,
                      double *out) {
    int rc = check_args(nc, X, OR_STORE_18);
    if (rc) return rc;
    if (nop < 1 || mu < 0 || mu > 2 || r < 0 || r > X[mu] || tmax < 0 || tmax > X[3]) return -2;
    const long V = volume_of(X);
    const int npair = nop * nop;
    const long nchunk = (V + CHUNK - 1) / CHUNK;
    double *part = (double *)calloc((size_t)nchunk * npair * (tmax + 1), sizeof(double));
    if (!part) return -4;
#pragma omp parallel for schedule(static)
    for (long ch = 0; ch < nchunk; ch++) {
        double *p = part + (size_t)ch * npair * (tmax + 1);
        long end = (ch + 1) * CHUNK < V ? (ch + 1) * CHUNK : V;
        for (long s = ch * CHUNK; s < end; s++) {
            int parity = s >= V / 2 ? 1 : 0;
            int x[4];
            or_eo_to_coords((int)(s - parity * (V / 2)), parity, X, x);
            int xr[4] = {x[0], x[1], x[2], x[3]};
            xr[mu] = (xr[mu] + r) % X[mu];
            mat t0, t1, lb, lt, m, a;
            munit(nc, &t0);
            munit(nc, &t1);
            for (int T = 0; T <= tmax; T++) {
                int y[4] = {x[0], x[1], x[2], (x[3] + T) % X[3]};
                const long top = or_coords_to_eo(y, X);
                for (int i = 0; i < nop; i++) {
                    mload(nc, ops + ((size_t)i * V + s) * nc * nc * 2, &lb);
                    for (int j = 0; j < nop; j++) {
                        mload(nc, ops + ((size_t)j * V + top) * nc * nc * 2, &a);
                        lt = mdag(nc, &a);
                        m = mmul(nc, &lb, &t1);
                        m = mmul(nc, &m, &lt);
                        a = mdag(nc, &t0);
                        m = mmul(nc, &m, &a);
                        p[(size_t)(i * nop + j) * (tmax + 1) + T] += mrealtrace(nc, &m);
                    }
                }
                if (T < tmax) {
                    mload(nc, link_at(nc, U, V, 3, top), &a);
                    t0 = mmul(nc, &t0, &a);
                    int z[4] = {xr[0], xr[1], xr[2], y[3]};
                    mload(nc, link_at(nc, U, V, 3, or_coords_to_eo(z, X)), &a);
                    t1 = mmul(nc, &t1, &a);
                }
            }
        }
    }
    for (long k = 0; k < (long)npair * (tmax + 1); k++) {
        double acc = 0.0;
        for (long ch = 0; ch < nchunk; ch++) acc += part[(size_t)ch * npair * (tmax + 1) + k];
        out[k] += acc;
    }
    free(part);
    return 0;
}

void or_matmul(int nc, const double *A, const double *B, double *C) {
    mat a, b, c;
    mload(nc, A, &a);
    mload(nc, B, &b);
    c = mmul(nc, &a, &b);
    mstore(nc, &c, C);
}

void or_det(int nc, const double *A, double out[2]) {
    mat a;
    mload(nc, A, &a);
    cx d = mdet(nc, &a);
    out[0] = d.re;
    out[1] = d.im;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
