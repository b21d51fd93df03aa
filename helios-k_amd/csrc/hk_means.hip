// hk_means.hip -- consumers of the cross sections: Planck / Rosseland mean
// sums over the grid and the per-bin transmission function.
//
// Means (Mean_kernel + IntegrateMean_kernel, resample.h:635-771, called at
// heliosk.cu:1953-2000): with x_j = fma(j, dnu, numin),
//   B_j    = 2 h c^2 x^3 / (exp(h c x / (kB T)) - 1)
//   dB_j   = t1 t2 / (T (e - 2 + 1/e)),  t1 = 2 h c^2 x^3, t2 = h c x/(kB T),
//            e = exp(t2)  (dB/dT, in the form that stays finite for large t2)
//   sums   = [sum B K, sum_{K != 0} dB / K, sum B, sum dB], B = dB = 0 at x = 0.
// The reference adds a Simpson end correction inside IntegrateMean_kernel,
// but its weights (constant memory) are only set later, before the
// transmission (heliosk.cu:2392), so the means it writes are the plain sums;
// those are what is computed here.
//
// Transmission (Integrate_kernel, resample.h:541-614, heliosk.cu:2400-2440):
// for bin b of Nxb sorted values and m_t = exp((t - nTr/2) dTr),
//   a = sum_k exp(-K_k m) + sum_{i<3} (w_i exp(-K_i m) + w_{2-i} exp(-K_{Nxb-1-i} m)),
//   w = (-5/8, 1/6, -1/24),
//   Tr = a / (Nxb - 1) (Nxb - Nxmin_b) / Nxb + exp(-kmin m) Nxmin_b / Nxb.
// (The end weights are the reference's, including their order at the upper
// end.)
//
// Both are one pass over K with fixed-order reductions: deterministic.
#include <math.h>

#include "hk_device.h"
#include "hk_internal.h"

namespace {

constexpr int kMeanThreads = 256;
constexpr int kMeanBlocks = 1024;     // partial sums per launch (fixed order)

__device__ inline void planck_terms(double x, double T, double &B, double &dB) {
    using namespace hk;
    const double t1 = 2.0 * kH * x * x * x * kC * kC;
    const double t2 = kH * x * kC / (kKB * T);
    const double e = exp(t2);
    B = t1 / (e - 1.0);
    dB = t1 * t2 / (T * (e - 2.0 + 1.0 / e));
    if (x == 0.0) { B = 0.0; dB = 0.0; }
}

// Per block: the four sums over points j = block*256 + tid + k*stride.
__global__ __launch_bounds__(kMeanThreads) void means_partial_kernel(
    const double *__restrict__ K, double numin, double dnu, int64_t nx, double T,
    double *__restrict__ part) {
    __shared__ double s[4][kMeanThreads];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kMeanThreads;
    for (int64_t j = (int64_t)blockIdx.x * kMeanThreads + threadIdx.x; j < nx; j += stride) {
        double B, dB;
        planck_terms(fma((double)j, dnu, numin), T, B, dB);
        const double k = K[j];
        a0 += B * k;
        if (k != 0.0) a1 += dB / k;
        a2 += B;
        a3 += dB;
    }
    const int t = threadIdx.x;
    s[0][t] = a0; s[1][t] = a1; s[2][t] = a2; s[3][t] = a3;
    __syncthreads();
    for (int h = kMeanThreads / 2; h > 0; h >>= 1) {
        if (t < h)
            for (int q = 0; q < 4; ++q) s[q][t] += s[q][t + h];
        __syncthreads();
    }
    if (t < 4) part[(size_t)blockIdx.x * 4 + t] = s[t][0];
}

// Trapezoid form on explicit points x (IntegrateMean_kernel with
// useIndividualX, resample.h:700-716): per interval [x_j, x_j+1] the mean of
// the end values times the width; the Rosseland term needs both K != 0.
__global__ __launch_bounds__(kMeanThreads) void means_x_partial_kernel(
    const double *__restrict__ K, const double *__restrict__ x, int64_t nx, double T,
    double *__restrict__ part) {
    __shared__ double s[4][kMeanThreads];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kMeanThreads;
    for (int64_t j = (int64_t)blockIdx.x * kMeanThreads + threadIdx.x; j + 1 < nx; j += stride) {
        double B0, dB0, B1, dB1;
        const double x0 = x[j], x1 = x[j + 1];
        planck_terms(x0, T, B0, dB0);
        planck_terms(x1, T, B1, dB1);
        const double k0 = K[j], k1 = K[j + 1], w = x1 - x0;
        a0 += 0.5 * (B0 * k0 + B1 * k1) * w;
        if (k0 != 0.0 && k1 != 0.0) a1 += 0.5 * (dB0 / k0 + dB1 / k1) * w;
        a2 += 0.5 * (B0 + B1) * w;
        a3 += 0.5 * (dB0 + dB1) * w;
    }
    const int t = threadIdx.x;
    s[0][t] = a0; s[1][t] = a1; s[2][t] = a2; s[3][t] = a3;
    __syncthreads();
    for (int h = kMeanThreads / 2; h > 0; h >>= 1) {
        if (t < h)
            for (int q = 0; q < 4; ++q) s[q][t] += s[q][t + h];
        __syncthreads();
    }
    if (t < 4) part[(size_t)blockIdx.x * 4 + t] = s[t][0];
}

__global__ __launch_bounds__(kMeanThreads) void means_final_kernel(
    const double *__restrict__ part, int nblocks, double *__restrict__ sums) {
    __shared__ double s[4][kMeanThreads];
    const int t = threadIdx.x;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = t; b < nblocks; b += kMeanThreads)
        for (int q = 0; q < 4; ++q) a[q] += part[(size_t)b * 4 + q];
    for (int q = 0; q < 4; ++q) s[q][t] = a[q];
    __syncthreads();
    for (int h = kMeanThreads / 2; h > 0; h >>= 1) {
        if (t < h)
            for (int q = 0; q < 4; ++q) s[q][t] += s[q][t + h];
        __syncthreads();
    }
    if (t < 4) sums[t] = s[t][0];
}

constexpr int kTrThreads = 256;
constexpr int kTrM = 8;               // m values per workgroup

// grid (ceil(nTr / kTrM), nbins): one workgroup sums kTrM transmission
// points of one bin; K of the bin is read once per workgroup (coalesced).
__global__ __launch_bounds__(kTrThreads) void transmission_kernel(
    const double *__restrict__ Ks, int Nxb, const int32_t *__restrict__ nxmin, double kmin,
    int nTr, double dTr, double *__restrict__ Tr) {
    __shared__ double s[kTrM][kTrThreads];
    const int b = blockIdx.y;
    const int t0 = blockIdx.x * kTrM;
    const int tid = threadIdx.x;
    const double *Kb = Ks + (size_t)b * Nxb;
    double m[kTrM], a[kTrM];
#pragma unroll
    for (int q = 0; q < kTrM; ++q) {
        m[q] = exp((double)(t0 + q - nTr / 2) * dTr);
        a[q] = 0.0;
    }
    for (int k = tid; k < Nxb; k += kTrThreads) {
        const double kv = Kb[k];
#pragma unroll
        for (int q = 0; q < kTrM; ++q) a[q] += exp(-kv * m[q]);
    }
    if (tid < 3) {
        const double w[3] = {-5.0 / 8.0, 1.0 / 6.0, -1.0 / 24.0};
        const double klo = Kb[tid], khi = Kb[Nxb - 1 - tid];
#pragma unroll
        for (int q = 0; q < kTrM; ++q) {
            a[q] += w[tid] * exp(-klo * m[q]);
            a[q] += w[2 - tid] * exp(-khi * m[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < kTrM; ++q) s[q][tid] = a[q];
    __syncthreads();
    for (int h = kTrThreads / 2; h > 0; h >>= 1) {
        if (tid < h)
#pragma unroll
            for (int q = 0; q < kTrM; ++q) s[q][tid] += s[q][tid + h];
        __syncthreads();
    }
    if (tid < kTrM && t0 + tid < nTr) {
        const double nmin = nxmin ? (double)nxmin[b] : 0.0;
        const double nb = (double)Nxb;
        Tr[(size_t)b * nTr + t0 + tid] =
            s[tid][0] / (nb - 1.0) * (nb - nmin) / nb + exp(-kmin * m[tid]) * nmin / nb;
    }
}

}  // namespace

extern "C" size_t hk_means_workspace_bytes(void) {
    return (size_t)kMeanBlocks * 4 * sizeof(double) + 256;
}

extern "C" int hk_means(const double *K, const hk_grid *grid, double T, double *sums,
                        void *workspace, size_t workspace_bytes, hk_stream_t stream) {
    HK_CHECK_ARG(K && grid && sums && workspace, "hk_means: null argument");
    HK_CHECK_ARG(grid->nx > 0 && grid->dnu > 0.0, "hk_means: empty grid or dnu <= 0");
    HK_CHECK_ARG(T > 0.0, "hk_means: T must be > 0");
    if (workspace_bytes < hk_means_workspace_bytes()) {
        hk::set_error("hk_means: workspace %zu bytes, %zu needed", workspace_bytes,
                      hk_means_workspace_bytes());
        return HK_ERR_WORKSPACE;
    }
    double *part = (double *)hk::align_up((size_t)workspace, 256);
    const int64_t want = (grid->nx + kMeanThreads - 1) / kMeanThreads;
    const int nblocks = (int)(want < kMeanBlocks ? want : kMeanBlocks);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(means_partial_kernel, dim3(nblocks), dim3(kMeanThreads), 0, s, K,
                       grid->numin, grid->dnu, grid->nx, T, part);
    HK_LAUNCH_CHECK();
    hipLaunchKernelGGL(means_final_kernel, dim3(1), dim3(kMeanThreads), 0, s, part, nblocks,
                       sums);
    HK_LAUNCH_CHECK();
    return HK_OK;
}

extern "C" int hk_means_x(const double *K, const double *x, int64_t nx, double T, double *sums,
                          void *workspace, size_t workspace_bytes, hk_stream_t stream) {
    HK_CHECK_ARG(K && x && sums && workspace, "hk_means_x: null argument");
    HK_CHECK_ARG(nx > 0, "hk_means_x: no points");
    HK_CHECK_ARG(T > 0.0, "hk_means_x: T must be > 0");
    if (workspace_bytes < hk_means_workspace_bytes()) {
        hk::set_error("hk_means_x: workspace %zu bytes, %zu needed", workspace_bytes,
                      hk_means_workspace_bytes());
        return HK_ERR_WORKSPACE;
    }
    double *part = (double *)hk::align_up((size_t)workspace, 256);
    const int64_t want = (nx + kMeanThreads - 1) / kMeanThreads;
    const int nblocks = (int)(want < kMeanBlocks ? want : kMeanBlocks);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(means_x_partial_kernel, dim3(nblocks), dim3(kMeanThreads), 0, s, K, x, nx, T,
                       part);
    HK_LAUNCH_CHECK();
    hipLaunchKernelGGL(means_final_kernel, dim3(1), dim3(kMeanThreads), 0, s, part, nblocks,
                       sums);
    HK_LAUNCH_CHECK();
    return HK_OK;
}

extern "C" int hk_transmission(const double *Ksorted, int nbins, int Nxb, const int32_t *nxmin,
                               double kmin, int nTr, double dTr, double *Tr,
                               hk_stream_t stream) {
    HK_CHECK_ARG(Ksorted && Tr, "hk_transmission: null argument");
    HK_CHECK_ARG(nbins > 0 && nbins <= 65535, "hk_transmission: nbins %d out of 1..65535", nbins);
    HK_CHECK_ARG(Nxb >= 3, "hk_transmission: %d points per bin (at least 3)", Nxb);
    HK_CHECK_ARG(nTr > 0, "hk_transmission: nTr must be > 0");
    hipStream_t s = (hipStream_t)stream;
    dim3 g((unsigned)((nTr + kTrM - 1) / kTrM), (unsigned)nbins);
    hipLaunchKernelGGL(transmission_kernel, g, dim3(kTrThreads), 0, s, Ksorted, Nxb, nxmin, kmin,
                       nTr, dTr, Tr);
    HK_LAUNCH_CHECK();
    return HK_OK;
}
