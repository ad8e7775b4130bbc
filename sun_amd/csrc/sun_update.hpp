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

// msu2.h:148-155
template <typename R>
SUN_HD void conj_normalize(Quat<R>& r) {
    R norm = R(1) / sqrt(r.a[0] * r.a[0] + r.a[1] * r.a[1] + r.a[2] * r.a[2] + r.a[3] * r.a[3]);
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
SUN_HD void heatbath_block(Mat<R, N>& U, Mat<R, N>& M, double beta_over_nc, Stream& st) {
    constexpr int x = block_p(N, B), y = block_q(N, B);
    Quat<R> r = get_block_su2<R, N>(M, x, y);
    R k = qabs(r);
    R ap = R(beta_over_nc) * k;
    k = R(1) / k;
#pragma unroll
    for (int i = 0; i < 4; i++) r.a[i] *= k;
    Quat<R> a;
    if constexpr (std::is_same<R, float>::value) {
        // FP32 links: the FP32 sampler (as the SU(3) FP32 path)
        generate_su2_milc_f32(ap, st, a.a);
    } else {
        double ad[4];
        generate_su2_milc((double)ap, st, ad);
#pragma unroll
        for (int i = 0; i < 4; i++) a.a[i] = R(ad[i]);
    }
    const Quat<R> rr = mul_uv_dagger(a, r);
    mul_block<R, N>(rr, U, x, y);
    mul_block<R, N>(rr, M, x, y);
}

template <typename R, int N, int... B>
SUN_HD void heatbath_blocks(Mat<R, N>& U, Mat<R, N>& M, double beta_over_nc, Stream& st,
                            std::integer_sequence<int, B...>) {
    (heatbath_block<R, N, B>(U, M, beta_over_nc, st), ...);
}

template <typename R, int N>
SUN_HD void heatbath_link(Mat<R, N>& U, const Mat<R, N>& F, double beta_over_nc, Stream& st) {
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
            generate_su2_milc((double)ap, st, ad);
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
        heatbath_blocks<R, N>(U, M, beta_over_nc, st, std::make_integer_sequence<int, N * (N - 1) / 2>{});
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
        R ai = atan2(t2.im, t2.re);
        t2 = mk<R>(cos(ai), -sin(ai));
#pragma unroll
        for (int c = 0; c < N; ++c) U.e[c][N - 1] = U.e[c][N - 1] * t2;
    }
}

}  // namespace sun
