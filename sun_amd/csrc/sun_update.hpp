// sun_update.hpp -- single-link Cabibbo-Marinari updates (host + device).
//
// heatbath_link restates heatBathSUN (device_PHB_OVR.h:361-485) and overrelax_link
// restates overrelaxationSUN (device_PHB_OVR.h:493-608): N=3 uses the reference's
// block-projection fast path, other N the M = U*F path.  The link, the staple and
// the SU(2) algebra run in the storage precision R; the sampler runs in double for
// FP64 links and in FP32 (generate_su2_milc_f32) for FP32 links.
#pragma once

#include <type_traits>
#include <utility>

#include "sun_rng.hpp"

namespace sun {

template <typename R>
struct Quat {
    R a[4];
};

// v * u^dagger (msu2.h:211-219)
template <typename R>
SUN_HD Quat<R> mul_uv_dagger(const Quat<R>& v, const Quat<R>& u) {
    Quat<R> b;
    b.a[0] = v.a[0] * u.a[0] + v.a[1] * u.a[1] + v.a[2] * u.a[2] + v.a[3] * u.a[3];
    b.a[1] = v.a[1] * u.a[0] - v.a[0] * u.a[1] + v.a[2] * u.a[3] - v.a[3] * u.a[2];
    b.a[2] = v.a[2] * u.a[0] - v.a[0] * u.a[2] + v.a[3] * u.a[1] - v.a[1] * u.a[3];
    b.a[3] = v.a[3] * u.a[0] - v.a[0] * u.a[3] + v.a[1] * u.a[2] - v.a[2] * u.a[1];
    return b;
}

template <typename R>
SUN_HD R qabs(const Quat<R>& r) {
    return sqrt(r.a[0] * r.a[0] + r.a[1] * r.a[1] + r.a[2] * r.a[2] + r.a[3] * r.a[3]);
}

// msu2.h:148-155 (FP32: the reciprocal square root is one v_rsq_f32)
template <typename R>
SUN_HD void conj_normalize(Quat<R>& r) {
    R norm;
    if constexpr (std::is_same<R, float>::value)
        norm = f32_rsq(r.a[0] * r.a[0] + r.a[1] * r.a[1] + r.a[2] * r.a[2] + r.a[3] * r.a[3]);
    else
        norm = R(1) / sqrt(r.a[0] * r.a[0] + r.a[1] * r.a[1] + r.a[2] * r.a[2] + r.a[3] * r.a[3]);
    r.a[0] *= norm;
    r.a[1] *= -norm;
    r.a[2] *= -norm;
    r.a[3] *= -norm;
}

// r r for a unit quaternion in the SU(2) product of msu2.h:90-97:
// (a0^2 - |a|^2, 2 a0 a1, 2 a0 a2, 2 a0 a3)
template <typename R>
SUN_HD Quat<R> quat_square(const Quat<R>& v) {
    Quat<R> r;
    r.a[0] = v.a[0] * v.a[0] - v.a[1] * v.a[1] - v.a[2] * v.a[2] - v.a[3] * v.a[3];
    r.a[1] = R(2) * v.a[0] * v.a[1];
    r.a[2] = R(2) * v.a[0] * v.a[2];
    r.a[3] = R(2) * v.a[0] * v.a[3];
    return r;
}

// rows p,q of U times columns p,q of F, projected on SU(2) (device_PHB_OVR.h:385-400)
template <typename R, int N>
SUN_HD Quat<R> block_projection(const Mat<R, N>& U, const Mat<R, N>& F, int p, int q) {
    cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0), a2 = mk<R>(0, 0), a3 = mk<R>(0, 0);
#pragma unroll
    for (int j = 0; j < N; j++) {
        a0 = a0 + U.e[p][j] * F.e[j][p];
        a1 = a1 + U.e[p][j] * F.e[j][q];
        a2 = a2 + U.e[q][j] * F.e[j][p];
        a3 = a3 + U.e[q][j] * F.e[j][q];
    }
    Quat<R> r;
    r.a[0] = a0.re + a3.re;
    r.a[1] = a1.im + a2.im;
    r.a[2] = a1.re - a2.re;
    r.a[3] = a0.im - a3.im;
    return r;
}

// device_PHB_OVR.h:265-272
template <typename R, int N>
SUN_HD Quat<R> get_block_su2(const Mat<R, N>& m, int x, int y) {
    Quat<R> r;
    r.a[0] = m.e[x][x].re + m.e[y][y].re;
    r.a[1] = m.e[x][y].im + m.e[y][x].im;
    r.a[2] = m.e[x][y].re - m.e[y][x].re;
    r.a[3] = m.e[x][x].im - m.e[y][y].im;
    return r;
}

// link <- u * link on rows (x,y) (device_PHB_OVR.h:296-303)
template <typename R, int N>
SUN_HD void mul_block(const Quat<R>& u, Mat<R, N>& link, int x, int y) {
    cplx<R> a00 = mk<R>(u.a[0], u.a[3]), a01 = mk<R>(u.a[2], u.a[1]);
    cplx<R> a10 = mk<R>(-u.a[2], u.a[1]), a11 = mk<R>(u.a[0], -u.a[3]);
#pragma unroll
    for (int j = 0; j < N; j++) {
        cplx<R> tmp = a00 * link.e[x][j] + a01 * link.e[y][j];
        link.e[y][j] = a10 * link.e[x][j] + a11 * link.e[y][j];
        link.e[x][j] = tmp;
    }
}

// one SU(2)-subgroup heatbath step of the N != 3 path on rows (x, y) of block B:
// U <- r U, M <- r M with r = a (M-block)^+ normalised (device_PHB_OVR.h:430-482)
template <typename R, int N, int B>
SUN_HD void heatbath_block(Mat<R, N>& U, Mat<R, N>& M, double beta_over_nc, const DrawKey& key) {
    constexpr int x = block_p(N, B), y = block_q(N, B);
    Quat<R> r = get_block_su2<R, N>(M, x, y);
    R k = qabs(r);
    R ap = R(beta_over_nc) * k;
    k = R(1) / k;
#pragma unroll
    for (int i = 0; i < 4; i++) r.a[i] *= k;
    Quat<R> a;
    const BlockDraws bd(key, B);
    if constexpr (std::is_same<R, float>::value) {
        // FP32 links: the FP32 sampler (as the SU(3) FP32 path)
        generate_su2_milc_f32(ap, key, B, bd, a.a);
    } else {
        double ad[4];
        generate_su2_milc((double)ap, key, B, bd, ad);
#pragma unroll
        for (int i = 0; i < 4; i++) a.a[i] = R(ad[i]);
    }
    const Quat<R> rr = mul_uv_dagger(a, r);
    mul_block<R, N>(rr, U, x, y);
    mul_block<R, N>(rr, M, x, y);
}

template <typename R, int N, int... B>
SUN_HD void heatbath_blocks(Mat<R, N>& U, Mat<R, N>& M, double beta_over_nc, const DrawKey& key,
                            std::integer_sequence<int, B...>) {
    (heatbath_block<R, N, B>(U, M, beta_over_nc, key), ...);
}

template <typename R, int N>
SUN_HD void heatbath_link(Mat<R, N>& U, const Mat<R, N>& F, double beta_over_nc, const DrawKey& key) {
    if constexpr (N == 3) {
#pragma unroll
        for (int block = 0; block < 3; block++) {
            const int p = block_p(3, block), q = block_q(3, block);
            Quat<R> r = block_projection<R, 3>(U, F, p, q);
            R k = qabs(r);
            R ap = R(beta_over_nc) * k;
            k = R(1) / k;
#pragma unroll
            for (int i = 0; i < 4; i++) r.a[i] *= k;
            double ad[4];
            generate_su2_milc((double)ap, key, block, BlockDraws(key, block), ad);
            Quat<R> a;
#pragma unroll
            for (int i = 0; i < 4; i++) a.a[i] = R(ad[i]);
            r = mul_uv_dagger(a, r);
            cplx<R> b0 = mk<R>(r.a[0], r.a[3]), b1 = mk<R>(r.a[2], r.a[1]);
            cplx<R> b2 = mk<R>(-r.a[2], r.a[1]), b3 = mk<R>(r.a[0], -r.a[3]);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                cplx<R> tmp0 = b0 * U.e[p][j] + b1 * U.e[q][j];
                U.e[q][j] = b2 * U.e[p][j] + b3 * U.e[q][j];
                U.e[p][j] = tmp0;
            }
        }
    } else {
        Mat<R, N> M = U * F;
        // one instance per block: the block's rows are compile-time constants (a
        // rolled loop would index U and M at run time, i.e. from scratch memory)
        heatbath_blocks<R, N>(U, M, beta_over_nc, key, std::make_integer_sequence<int, N * (N - 1) / 2>{});
    }
}

template <typename R, int N>
SUN_HD void overrelax_link(Mat<R, N>& U, const Mat<R, N>& F) {
    if constexpr (N == 3) {
#pragma unroll
        for (int block = 0; block < 3; block++) {
            const int p = block_p(3, block), q = block_q(3, block);
            Quat<R> r = block_projection<R, 3>(U, F, p, q);
            conj_normalize(r);
            cplx<R> b0 = mk<R>(r.a[0], r.a[3]), b1 = mk<R>(r.a[2], r.a[1]);
            cplx<R> b2 = mk<R>(-r.a[2], r.a[1]), b3 = mk<R>(r.a[0], -r.a[3]);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                cplx<R> tmp0 = b0 * U.e[p][j] + b1 * U.e[q][j];
                cplx<R> tmp1 = b2 * U.e[p][j] + b3 * U.e[q][j];
                U.e[p][j] = b0 * tmp0 + b1 * tmp1;
                U.e[q][j] = b2 * tmp0 + b3 * tmp1;
            }
        }
    } else {
        Mat<R, N> M = U * F;
#pragma unroll
        for (int block = 0; block < N * (N - 1) / 2; block++) {
            const int x = block_p(N, block), y = block_q(N, block);
            Quat<R> r = get_block_su2<R, N>(M, x, y);
            conj_normalize(r);
            if constexpr (std::is_same<R, float>::value) {
                // FP32 links: the rotation applied once as r^2 (half the row updates;
                // equal to applying r twice up to FP32 rounding)
                const Quat<R> r2 = quat_square(r);
                mul_block<R, N>(r2, U, x, y);
                mul_block<R, N>(r2, M, x, y);
            } else {
                mul_block<R, N>(r, U, x, y);
                mul_block<R, N>(r, U, x, y);
                mul_block<R, N>(r, M, x, y);
                mul_block<R, N>(r, M, x, y);
            }
        }
    }
}

// Projection onto SU(N) by SU(2)-subgroup maximisation of Re Tr(w q^+)
// (projectlink.h:16-64, project_link): per hit, for every subgroup block, the block
// of w q^+ is normalised, conjugated and multiplied into w from the left.  With
// tol > 0 the hits stop once the relative change of Re Tr(w q) / N (the reference's
// measure, realtraceUVdagger of w and the daggered q) falls to tol.  SU3project
// (projectlink.h:119-139), the reference's N = 3 choice, is this with w = 1,
// tol = 0 and the blocks (0,1), (1,2), (0,2); it has no zero-block guard (a block
// of exactly zero norm would give NaN there and the identity block here).
template <typename R, int N>
SUN_HD R retrace_conj(const Mat<R, N>& a, const Mat<R, N>& b) {  // Re Tr(a b) with b stored daggered
    R s = R(0);
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) s += a.e[i][j].re * b.e[i][j].re + a.e[i][j].im * b.e[i][j].im;
    return s;
}

template <typename R, int N>
SUN_HD void project_link(Mat<R, N>& w, const Mat<R, N>& q_in, int nhit, R tol) {
    const Mat<R, N> q = dagger(q_in);
    R conver = R(1), old_tr = R(0);
    if (tol > R(0)) old_tr = retrace_conj<R, N>(w, q) / R(N);
#pragma unroll 1
    for (int hit = 0; hit < nhit && conver > tol; hit++) {
        // unrolled: compile-time rows keep w in registers
#pragma unroll
        for (int block = 0; block < N * (N - 1) / 2; block++) {
            const int x = block_p(N, block), y = block_q(N, block);
            const Mat<R, N> action = w * q;
            Quat<R> r = get_block_su2<R, N>(action, x, y);
            R k = qabs(r);
            if (k == R(0)) {
                k = R(1);
                r.a[0] = R(1);
            }
            k = R(1) / k;
            r.a[0] *= k;
            r.a[1] *= -k;
            r.a[2] *= -k;
            r.a[3] *= -k;
            mul_block<R, N>(r, w, x, y);
        }
        if (tol > R(0)) {
            const R new_tr = retrace_conj<R, N>(w, q) / R(N);
            conver = (new_tr - old_tr) / old_tr;
            old_tr = new_tr;
        }
    }
}

// Reunitarisation (reunitlink.h:16-85).  N=3: row Gram-Schmidt + cross product;
// N>3: column Gram-Schmidt + determinant phase on the last column.  N=2 is not
// compilable in the reference (reunitlink.h:86); here the textbook projection
// a=(U00+conj U11)/2, b=(U01-conj U10)/2, normalised.
template <typename R, int N>
SUN_HD void reunit_link(Mat<R, N>& U) {
    if constexpr (N == 3) {
        R t1 = R(0);
        cplx<R> t2 = mk<R>(0, 0);
#pragma unroll
        for (int c = 0; c < 3; c++) t1 += abs2(U.e[0][c]);
        t1 = R(1) / sqrt(t1);
#pragma unroll
        for (int c = 0; c < 3; c++) U.e[0][c] = scale(U.e[0][c], t1);
#pragma unroll
        for (int c = 0; c < 3; c++) t2 = t2 + conj(U.e[0][c]) * U.e[1][c];
#pragma unroll
        for (int c = 0; c < 3; c++) U.e[1][c] = U.e[1][c] - t2 * U.e[0][c];
        t1 = R(0);
#pragma unroll
        for (int c = 0; c < 3; c++) t1 += abs2(U.e[1][c]);
        t1 = R(1) / sqrt(t1);
#pragma unroll
        for (int c = 0; c < 3; c++) U.e[1][c] = scale(U.e[1][c], t1);
        reconstruct_row2(U);
    } else if constexpr (N == 2) {
        cplx<R> a = mk<R>((U.e[0][0].re + U.e[1][1].re) * R(0.5), (U.e[0][0].im - U.e[1][1].im) * R(0.5));
        cplx<R> b = mk<R>((U.e[0][1].re - U.e[1][0].re) * R(0.5), (U.e[0][1].im + U.e[1][0].im) * R(0.5));
        R n = R(1) / sqrt(abs2(a) + abs2(b));
        a = scale(a, n);
        b = scale(b, n);
        U.e[0][0] = a;
        U.e[0][1] = b;
        U.e[1][0] = mk<R>(-b.re, b.im);
        U.e[1][1] = conj(a);
    } else {
        R t1 = R(0);
        cplx<R> t2;
#pragma unroll
        for (int c = 0; c < N; c++) t1 += abs2(U.e[c][0]);
        t1 = R(1) / sqrt(t1);
#pragma unroll
        for (int c = 0; c < N; c++) U.e[c][0] = scale(U.e[c][0], t1);
#pragma unroll
        for (int j = 1; j < N; j++) {
#pragma unroll
            for (int i = 0; i < j; i++) {
                t2 = conj(U.e[0][i]) * U.e[0][j];
#pragma unroll
                for (int c = 1; c < N; ++c) t2 = t2 + conj(U.e[c][i]) * U.e[c][j];
#pragma unroll
                for (int c = 0; c < N; ++c) U.e[c][j] = U.e[c][j] - t2 * U.e[c][i];
            }
            t1 = R(0);
#pragma unroll
            for (int c = 0; c < N; ++c) t1 += abs2(U.e[c][j]);
            t1 = R(1) / sqrt(t1);
#pragma unroll
            for (int c = 0; c < N; ++c) U.e[c][j] = scale(U.e[c][j], t1);
        }
        t2 = det<R, N>(U);
        const R ai = arg_of(t2);
        R sn, cs;
        sincos_of(ai, sn, cs);
        t2 = mk<R>(cs, -sn);
#pragma unroll
        for (int c = 0; c < N; ++c) U.e[c][N - 1] = U.e[c][N - 1] * t2;
    }
}


// exp(iQ) for a traceless Hermitian 3x3 Q (Morningstar and Peardon, Phys. Rev. D 69
// (2004) 054501, eqs. (19)-(34)): exp(iQ) = f0 + f1 Q + f2 Q^2 with the f_j from
// c0 = det Q = Tr Q^3 / 3 and c1 = Tr Q^2 / 2.  For c1 below kStoutSeriesC1 (|Q| <
// 0.08) the f_j are the exponential series reduced by Cayley-Hamilton,
// Q^3 = c1 Q + c0, summed to order 16 (the closed form divides by ~c1 there).
constexpr double kStoutSeriesC1 = 4.0e-3;

SUN_HD void exp_iq_su3(const Mat<double, 3>& Q, Mat<double, 3>& E) {
    const Mat<double, 3> QQ = Q * Q;
    const double c1 = 0.5 * realtrace(QQ);
    const double c0 = realtrace(Q * QQ) / 3.0;
    cplx<double> f0, f1, f2;
    if (c1 < kStoutSeriesC1) {
        // Q^k = a + b Q + c Q^2 -> Q^(k+1) = c c0 + (a + c c1) Q + b Q^2
        double a = 1.0, b = 0.0, c = 0.0;
        cplx<double> t = mk<double>(1.0, 0.0);  // i^k / k!
        f0 = f1 = f2 = mk<double>(0.0, 0.0);
#pragma unroll 1
        for (int k = 0; k <= 16; k++) {
            f0 = f0 + scale(t, a);
            f1 = f1 + scale(t, b);
            f2 = f2 + scale(t, c);
            const double na = c * c0, nb = a + c * c1, nc = b;
            a = na;
            b = nb;
            c = nc;
            t = mk<double>(-t.im / (k + 1), t.re / (k + 1));
        }
    } else {
        const double c0abs = fabs(c0);
        const double s3 = sqrt(c1 / 3.0);
        const double c0max = 2.0 * (c1 / 3.0) * s3;
        const double theta = acos(fmin(c0abs / c0max, 1.0));
        const double u = s3 * cos(theta / 3.0), w = sqrt(c1) * sin(theta / 3.0);
        const double u2 = u * u, w2 = w * w;
        const double xi0 = fabs(w) < 0.05 ? 1.0 - w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0)) : sin(w) / w;
        const double cw = cos(w);
        const cplx<double> e2iu = mk<double>(cos(2.0 * u), sin(2.0 * u));
        const cplx<double> emiu = mk<double>(cos(u), -sin(u));
        const cplx<double> h0 = scale(e2iu, u2 - w2) + emiu * mk<double>(8.0 * u2 * cw, 2.0 * u * (3.0 * u2 + w2) * xi0);
        const cplx<double> h1 = scale(e2iu, 2.0 * u) - emiu * mk<double>(2.0 * u * cw, -(3.0 * u2 - w2) * xi0);
        const cplx<double> h2 = e2iu - emiu * mk<double>(cw, 3.0 * u * xi0);
        const double den = 1.0 / (9.0 * u2 - w2);
        f0 = scale(h0, den);
        f1 = scale(h1, den);
        f2 = scale(h2, den);
        if (c0 < 0) {  // f_j(-c0) = (-1)^j conj f_j(c0)
            f0 = conj(f0);
            f1 = mk<double>(-f1.re, f1.im);
            f2 = conj(f2);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<double> e = f1 * Q.e[i][j] + f2 * QQ.e[i][j];
            if (i == j) e = e + f0;
            E.e[i][j] = e;
        }
}

// One stout step of a link (stout.cu:86-92, 519-535): Omega = V U^+ with V = rho x
// the staple sum, Q = the traceless part of (i/2)(Omega^+ - Omega), W = exp(iQ) U.
SUN_HD Mat<double, 3> stout_link(const Mat<double, 3>& V, const Mat<double, 3>& U) {
    const Mat<double, 3> Om = V * dagger(U);
    Mat<double, 3> Q;
    cplx<double> tr = mk<double>(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            // (i/2)(conj(Om_ji) - Om_ij)
            const cplx<double> d = conj(Om.e[j][i]) - Om.e[i][j];
            Q.e[i][j] = mk<double>(-0.5 * d.im, 0.5 * d.re);
        }
#pragma unroll
    for (int i = 0; i < 3; i++) tr = tr + Q.e[i][i];
#pragma unroll
    for (int i = 0; i < 3; i++) Q.e[i][i] = Q.e[i][i] - scale(tr, 1.0 / 3.0);
    Mat<double, 3> E;
    exp_iq_su3(Q, E);
    return E * U;
}

}  // namespace sun
