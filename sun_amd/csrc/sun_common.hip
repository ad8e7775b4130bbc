// sun_common.hip -- variant-independent kernels: the fixed-order final reduction
// of per-block partial sums (single lattice and batched).
#include <hip/hip_runtime.h>

#include "sun_internal.hpp"

namespace sun {

namespace {

__device__ __forceinline__ double2 block_sum_all(double2 v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_down(v.x, off, 64);
        v.y += __shfl_down(v.y, off, 64);
    }
    __shared__ double2 ws[kReduceBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) ws[w] = v;
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < kReduceBlock / 64; i++) {
            r.x += ws[i].x;
            r.y += ws[i].y;
        }
    }
    return r;
}

// thread t sums partials t, t+256, t+512, ... in order; then the block tree.  The
// loads are issued kSumBatch at a time before the additions (which keep their order),
// so a thread waits for one memory latency per batch instead of one per partial.
constexpr int kSumBatch = 8;

__device__ __forceinline__ double2 ordered_sum(const double2* p, int n) {
    double2 acc = make_double2(0.0, 0.0);
    int i = threadIdx.x;
    for (; i + (kSumBatch - 1) * kReduceBlock < n; i += kSumBatch * kReduceBlock) {
        double2 v[kSumBatch];
#pragma unroll
        for (int k = 0; k < kSumBatch; k++) v[k] = p[i + k * kReduceBlock];
#pragma unroll
        for (int k = 0; k < kSumBatch; k++) {
            acc.x += v[k].x;
            acc.y += v[k].y;
        }
    }
    for (; i < n; i += kReduceBlock) {
        acc.x += p[i].x;
        acc.y += p[i].y;
    }
    return block_sum_all(acc);
}

__global__ void __launch_bounds__(kReduceBlock) k_finalize(const double2* partials, int n, double s0, double s1,
                                                           double2* out) {
    double2 v = ordered_sum(partials, n);
    if (threadIdx.x == 0) out[0] = make_double2(v.x * s0, v.y * s1);
}

// trace and determinant segments summed in the order of ordered_sum, maxima of the
// deviation segment; square roots of the maxima at the end
__global__ void __launch_bounds__(kReduceBlock) k_monitor_finalize(const double2* partials, int n, double s_tr,
                                                                   double s_det, double2* out) {
    const double2 tr = ordered_sum(partials, n);
    __syncthreads();
    const double2 dt = ordered_sum(partials + n, n);
    double2 m = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < n; i += kReduceBlock) {
        m.x = nan_max(m.x, partials[2 * n + i].x);
        m.y = nan_max(m.y, partials[2 * n + i].y);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        m.x = nan_max(m.x, __shfl_down(m.x, off, 64));
        m.y = nan_max(m.y, __shfl_down(m.y, off, 64));
    }
    __shared__ double2 wm[kReduceBlock / 64];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 1; i < kReduceBlock / 64; i++) {
            m.x = nan_max(m.x, wm[i].x);
            m.y = nan_max(m.y, wm[i].y);
        }
        out[0] = make_double2(tr.x * s_tr, tr.y * s_tr);
        out[1] = make_double2(dt.x * s_det, dt.y * s_det);
        out[2] = make_double2(sqrt(m.x), sqrt(m.y));
    }
}

__global__ void __launch_bounds__(kReduceBlock) k_finalize_iteration(const double2* partials, int pb, int qb,
                                                                     double s_plaq, double s_poly, double2* out,
                                                                     double2* history, uint32_t* ctr, uint32_t first) {
    const double2 p = ordered_sum(partials, pb);
    __syncthreads();
    const double2 q = ordered_sum(partials + pb, qb);
    if (threadIdx.x == 0) {
        const double2 a = make_double2(p.x * s_plaq, p.y * s_plaq), b = make_double2(q.x * s_poly, q.y * s_poly);
        out[0] = a;
        out[1] = b;
        const size_t row = (size_t)(ctr[0] - first);
        history[2 * row] = a;
        history[2 * row + 1] = b;
        ctr[0] += 1u;
    }
}

__global__ void __launch_bounds__(kReduceBlock) k_wl_finalize(const double2* partials, int nb, int tmax, double scale,
                                                              double2* out) {
    const int T = blockIdx.x;
    double2 acc = make_double2(0.0, 0.0);
    for (int mu = 0; mu < 3; mu++) {
        const double2 v = ordered_sum(partials + (size_t)(mu * (tmax + 1) + T) * nb, nb);
        __syncthreads();
        acc.x += v.x;
        acc.y += v.y;
    }
    if (threadIdx.x == 0) out[T] = make_double2(acc.x * scale, acc.y * scale);
}

__global__ void __launch_bounds__(kReduceBlock) k_wlop_finalize(const double* partials, int nb, double* accum) {
    const double* p = partials + (size_t)blockIdx.x * nb;
    double acc = 0.0;
    int i = threadIdx.x;
    for (; i + (kSumBatch - 1) * kReduceBlock < nb; i += kSumBatch * kReduceBlock) {
        double v[kSumBatch];
#pragma unroll
        for (int k = 0; k < kSumBatch; k++) v[k] = p[i + k * kReduceBlock];
#pragma unroll
        for (int k = 0; k < kSumBatch; k++) acc += v[k];
    }
    for (; i < nb; i += kReduceBlock) acc += p[i];
    const double2 v = block_sum_all(make_double2(acc, 0.0));
    if (threadIdx.x == 0) accum[blockIdx.x] += v.x;
}

__global__ void __launch_bounds__(kReduceBlock) k_batch_finalize(const BatchEntry* table, double2* out) {
    const BatchEntry& E = table[blockIdx.x];
    double2 v = ordered_sum(E.partials, E.nblocks);
    if (threadIdx.x == 0) out[blockIdx.x] = make_double2(v.x * E.scale, v.y * E.scale);
}

__global__ void k_batch_record(BatchEntry* table, int n, const double2* plaq, const double2* poly) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    BatchEntry& E = table[i];
    if (E.hist) {
        E.hist[2 * (size_t)E.hist_iter] = plaq[i];
        E.hist[2 * (size_t)E.hist_iter + 1] = poly ? poly[i] : make_double2(0.0, 0.0);
    }
    E.hist_iter += 1u;
}

// the end of a large-volume group's iteration in one launch (one workgroup per
// member): the plaquette and Polyakov sums of k_batch_finalize, the history record of
// k_batch_record and the sweep advance of k_batch_advance, with the same values
__global__ void __launch_bounds__(kReduceBlock) k_group_close(BatchEntry* t_plaq, const BatchEntry* t_poly,
                                                              BatchEntry* t_update, double2* plaq, double2* poly) {
    const int i = blockIdx.x;
    BatchEntry& P = t_plaq[i];
    const BatchEntry& Q = t_poly[i];
    double2 a = ordered_sum(P.partials, P.nblocks);
    __syncthreads();
    double2 b = ordered_sum(Q.partials, Q.nblocks);
    if (threadIdx.x == 0) {
        a = make_double2(a.x * P.scale, a.y * P.scale);
        b = make_double2(b.x * Q.scale, b.y * Q.scale);
        plaq[i] = a;
        poly[i] = b;
        if (P.hist) {
            P.hist[2 * (size_t)P.hist_iter] = a;
            P.hist[2 * (size_t)P.hist_iter + 1] = b;
        }
        P.hist_iter += 1u;
        t_update[i].sweep += 1u;
    }
}

__global__ void k_advance_counter(uint32_t* ctr) {
    if (threadIdx.x == 0) ctr[0] += 1u;
}

__global__ void k_record_obs(const double2* src, int n, double2* history, const uint32_t* ctr, uint32_t first) {
    const int i = threadIdx.x;
    if (i < n) history[(size_t)(ctr[0] - first) * n + i] = src[i];
}

__global__ void k_batch_advance(BatchEntry* table, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) table[i].sweep += 1u;
}

}  // namespace

hipError_t launch_advance_counter(uint32_t* ctr, hipStream_t s) {
    hipLaunchKernelGGL(k_advance_counter, dim3(1), dim3(64), 0, s, ctr);
    return hipGetLastError();
}

hipError_t launch_record_obs(const double2* src, int n, double2* history, const uint32_t* ctr, uint32_t first,
                             hipStream_t s) {
    if (n > 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_record_obs, dim3(1), dim3(((n + 63) / 64) * 64), 0, s, src, n, history, ctr, first);
    return hipGetLastError();
}

hipError_t launch_batch_advance(BatchEntry* table, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_advance, dim3((n + 255) / 256), dim3(256), 0, s, table, n);
    return hipGetLastError();
}

hipError_t launch_finalize(const double2* partials, int n, double s0, double s1, double2* out, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kReduceBlock), 0, s, partials, n, s0, s1, out);
    return hipGetLastError();
}

hipError_t launch_monitor_finalize(const double2* partials, int n, double s_tr, double s_det, double2* out,
                                   hipStream_t s) {
    hipLaunchKernelGGL(k_monitor_finalize, dim3(1), dim3(kReduceBlock), 0, s, partials, n, s_tr, s_det, out);
    return hipGetLastError();
}

hipError_t launch_finalize_iteration(const double2* partials, int pb, int qb, double s_plaq, double s_poly,
                                     double2* out, double2* history, uint32_t* ctr, uint32_t first, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize_iteration, dim3(1), dim3(kReduceBlock), 0, s, partials, pb, qb, s_plaq, s_poly, out,
                       history, ctr, first);
    return hipGetLastError();
}

hipError_t launch_wl_finalize(const double2* partials, int nb, int tmax, double scale, double2* out, hipStream_t s) {
    hipLaunchKernelGGL(k_wl_finalize, dim3(tmax + 1), dim3(kReduceBlock), 0, s, partials, nb, tmax, scale, out);
    return hipGetLastError();
}

hipError_t launch_wlop_finalize(const double* partials, int nb, int nout, double* accum, hipStream_t s) {
    if (nout <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_wlop_finalize, dim3(nout), dim3(kReduceBlock), 0, s, partials, nb, accum);
    return hipGetLastError();
}

hipError_t launch_batch_finalize(const BatchEntry* table, int nlat, double2* out, hipStream_t s) {
    if (nlat <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_finalize, dim3(nlat), dim3(kReduceBlock), 0, s, table, out);
    return hipGetLastError();
}

hipError_t launch_group_close(BatchEntry* t_plaq, const BatchEntry* t_poly, BatchEntry* t_update, int nlat,
                              double2* plaq, double2* poly, hipStream_t s) {
    if (nlat <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_group_close, dim3(nlat), dim3(kReduceBlock), 0, s, t_plaq, t_poly, t_update, plaq, poly);
    return hipGetLastError();
}

hipError_t launch_batch_record(BatchEntry* table, int nlat, const double2* plaq, const double2* poly,
                               hipStream_t s) {
    if (nlat <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_record, dim3((nlat + 255) / 256), dim3(256), 0, s, table, nlat, plaq, poly);
    return hipGetLastError();
}

}  // namespace sun
