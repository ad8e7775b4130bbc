// sun_math.hpp -- complex numbers, N x N colour matrices and SU(2) quaternions for
// the HIP engine (host + device).  Operation order follows the reference's
// complex.h:150-152 (complex product) and matrixsun.h:298-347 (matrix product), so
// the FP64 engine tracks the FP64 oracle to rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sun_libm.h"

#define SUN_HD __host__ __device__ __forceinline__

namespace sun {

constexpr double kTwoPi = 6.2831853071795864769252867665590;  // PII, constants.h:27

template <typename R>
struct cplx {
    R re, im;
};

template <typename R>
SUN_HD cplx<R> mk(R r, R i) { cplx<R> z; z.re = r; z.im = i; return z; }
template <typename R>
SUN_HD cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk<R>(a.re + b.re, a.im + b.im); }
template <typename R>
SUN_HD cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk<R>(a.re - b.re, a.im - b.im); }
// (x*bx - y*by, y*bx + x*by)
template <typename R>
SUN_HD cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return mk<R>(a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im);
}
template <typename R>
SUN_HD cplx<R> conj(cplx<R> a) { return mk<R>(a.re, -a.im); }
template <typename R>
SUN_HD R abs2(cplx<R> a) { return a.re * a.re + a.im * a.im; }
template <typename R>
SUN_HD cplx<R> scale(cplx<R> a, R s) { return mk<R>(a.re * s, a.im * s); }

// arg z and (sin, cos) of a phase: FP64 through sun_libm.h (the host restatement
// computes the same bits), FP32 the library routines
SUN_HD double arg_of(cplx<double> z) { return sun_atan2(z.im, z.re); }
SUN_HD float arg_of(cplx<float> z) { return atan2f(z.im, z.re); }
SUN_HD void sincos_of(double th, double& s, double& c) { sun_sincos(th, &s, &c); }
SUN_HD void sincos_of(float th, float& s, float& c) {
    s = sinf(th);
    c = cosf(th);
}

template <typename R, int N>
struct Mat {
    cplx<R> e[N][N];
};

template <typename R, int N>
SUN_HD Mat<R, N> mzero() {
    Mat<R, N> m;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) m.e[i][j] = mk<R>(R(0), R(0));
    return m;
}

template <typename R, int N>
SUN_HD Mat<R, N> munit() {
    Mat<R, N> m = mzero<R, N>();
#pragma unroll
    for (int i = 0; i < N; i++) m.e[i][i] = mk<R>(R(1), R(0));
    return m;
}

// res_ij = a_i0*b_0j (+ a_ik*b_kj, k ascending)
template <typename R, int N>
SUN_HD Mat<R, N> operator*(const Mat<R, N>& a, const Mat<R, N>& b) {
    Mat<R, N> r;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            cplx<R> s = a.e[i][0] * b.e[0][j];
#pragma unroll
            for (int k = 1; k < N; k++) s = s + a.e[i][k] * b.e[k][j];
            r.e[i][j] = s;
        }
    return r;
}

template <typename R, int N>
SUN_HD Mat<R, N> dagger(const Mat<R, N>& a) {
    Mat<R, N> r;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) r.e[i][j] = conj(a.e[j][i]);
    return r;
}

template <typename R, int N>
SUN_HD void add_to(Mat<R, N>& acc, const Mat<R, N>& a) {
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) acc.e[i][j] = acc.e[i][j] + a.e[i][j];
}

template <typename R, int N>
SUN_HD R realtrace(const Mat<R, N>& a) {
    R s = a.e[0][0].re;
#pragma unroll
    for (int i = 1; i < N; i++) s += a.e[i][i].re;
    return s;
}

template <typename R, int N>
SUN_HD cplx<R> trace(const Mat<R, N>& a) {
    cplx<R> s = a.e[0][0];
#pragma unroll
    for (int i = 1; i < N; i++) s = s + a.e[i][i];
    return s;
}

template <typename Rout, typename Rin, int N>
SUN_HD Mat<Rout, N> mconvert(const Mat<Rin, N>& a) {
    Mat<Rout, N> r;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) r.e[i][j] = mk<Rout>(Rout(a.e[i][j].re), Rout(a.e[i][j].im));
    return r;
}

// Third row of an SU(3) matrix from the first two: conj(r0 x r1)
// (reconstruct_12p_8p.h:17-21 of the reference).
template <typename R>
SUN_HD void reconstruct_row2(Mat<R, 3>& A) {
    A.e[2][0] = conj(A.e[0][1] * A.e[1][2] - A.e[0][2] * A.e[1][1]);
    A.e[2][1] = conj(A.e[0][2] * A.e[1][0] - A.e[0][0] * A.e[1][2]);
    A.e[2][2] = conj(A.e[0][0] * A.e[1][1] - A.e[0][1] * A.e[1][0]);
}

// Determinant: cofactor expansion for N=3, Crout LU without pivoting otherwise
// (matrixsun.h:481-505).
template <typename R, int N>
SUN_HD cplx<R> det(const Mat<R, N>& a) {
    cplx<R> res;
    if constexpr (N == 3) {
        res = a.e[0][1] * a.e[1][2] * a.e[2][0];
        res = res - a.e[0][2] * a.e[1][1] * a.e[2][0];
        res = res + a.e[0][2] * a.e[1][0] * a.e[2][1];
        res = res - a.e[0][0] * a.e[1][2] * a.e[2][1];
        res = res - a.e[0][1] * a.e[1][0] * a.e[2][2];
        res = res + a.e[0][0] * a.e[1][1] * a.e[2][2];
        return res;
    } else {
        Mat<R, N> b = a;
#pragma unroll
        for (int j = 0; j < N; j++) {
#pragma unroll
            for (int i = 0; i <= j; i++) {
                res = b.e[j][i];
#pragma unroll
                for (int c = 0; c < i; c++) res = res - b.e[c][i] * b.e[j][c];
                b.e[j][i] = res;
            }
#pragma unroll
            for (int i = j + 1; i < N; i++) {
                res = b.e[j][i];
#pragma unroll
                for (int c = 0; c < j; c++) res = res - b.e[c][i] * b.e[j][c];
                cplx<R> t = conj(b.e[j][j]) * res;
                R d = abs2(b.e[j][j]);
                b.e[j][i] = mk<R>(t.re / d, t.im / d);
            }
        }
        res = b.e[0][0] * b.e[1][1];
#pragma unroll
        for (int c = 2; c < N; c++) res = res * b.e[c][c];
        return res;
    }
}

// ---- SU(2) subgroup enumeration (constants.h:143-192) ----------------------------
// N=3 uses the reference's fixed order (0,1),(1,2),(0,2); other N the diagonal sweep
// (0,1),(1,2),...,(0,2),(1,3),...  Both are the same enumeration for N=3.
SUN_HD constexpr int block_p(int nc, int block) {
    int index = -1;
    for (int del = 1; del < nc; del++)
        for (int i1 = 0; i1 < nc - del; i1++) {
            index++;
            if (index == block) return i1;
        }
    return 0;
}
SUN_HD constexpr int block_q(int nc, int block) {
    int index = -1;
    for (int del = 1; del < nc; del++)
        for (int i1 = 0; i1 < nc - del; i1++) {
            index++;
            if (index == block) return i1 + del;
        }
    return 1;
}

}  // namespace sun
