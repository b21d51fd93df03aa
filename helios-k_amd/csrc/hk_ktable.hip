// hk_ktable.hip -- k-distribution tables: per-bin sort, Nxmin, rescale, ln K,
// Chebyshev least squares, and the K(y) reconstruction.
//
// Reference sequence (heliosk.cu:2046-2193, resample.h):
//   thrust sort_by_key + stable_sort_by_key     -> ascending K within each bin
//   findCut_kernel (resample.h:362)             -> Nxmin per bin
//   rescale_kernel f=+1, copyK2 (resample.h:398, 445)
//   lnK_kernel (resample.h:472)
//   Vandermonde + QR_kernel (resample.h:14, 73) -> once per (Nxb, nC)
//   leastSquare_kernel (resample.h:243)         -> NC-1 reflections + back substitution
//
// Here:
//   * bins of <= 16384 points are sorted by one workgroup each with a
//     bitonic network in LDS; longer bins are cut into LDS-sorted chunks and
//     merged pass by pass (one merge-path split per workgroup, slices staged
//     through LDS, coalesced reads and writes);
//   * one workgroup per (spectrum, bin) applies the findCut rules, the
//     rescale, ln, and forms the nC Chebyshev moments m_i = sum_k T_i(x_k) b_k
//     plus the first nC values b0 of b = ln K2;
//   * the fit is c = P1 m + P2 b0 with an nC x 2nC matrix P computed once on
//     the host.  The reference's solve c = G b (Householder QR, NC-1
//     reflections applied to b, back substitution) is linear in b and every
//     row of G lies in span{Chebyshev columns, e_0..e_{nC-1}}; P is found by
//     running that exact solve on 2nC probe vectors.  One pass over the bin
//     replaces nC-1 sequential reflection passes.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "hk_device.h"
#include "hk_internal.h"

namespace hk {

constexpr int kSortLDS = 16384;          // doubles per workgroup (128 KiB)
constexpr int kSortThreads = 1024;
constexpr int kMaxNC = 32;
constexpr int kFitThreads = 256;
constexpr int kMergeE = 8;               // outputs per thread in a merge pass
// The moment form c = P1 m + P2 b0 is exact algebra but inherits the
// conditioning of the Chebyshev normal equations; for short bins
// (Nxb < kPlanMinRatio * nC) the fit applies the Householder reflections
// literally instead (b fits in LDS there).
constexpr int kPlanMinRatio = 25;
constexpr int kHHMaxNxb = kPlanMinRatio * kMaxNC;

struct Seg {
    int64_t start, len;
};

// ---------------------------------------------------------------------------
// Bitonic sort of one segment (or chunk) of length <= N in LDS.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(kSortThreads) void bitonic_kernel(const double *__restrict__ in,
                                                               double *__restrict__ out,
                                                               const Seg *__restrict__ segs,
                                                               int64_t chunk) {
    __shared__ double s[N];
    const Seg sg = segs[blockIdx.y];
    const int64_t off = (int64_t)blockIdx.x * chunk;
    if (off >= sg.len && !(off == 0 && sg.len == 0)) return;
    const int len = (int)min((int64_t)chunk, sg.len - off);
    if (len <= 0) return;
    int n = 1;
    while (n < len) n <<= 1;
    const double *src = in + sg.start + off;
    for (int i = threadIdx.x; i < n; i += kSortThreads) s[i] = i < len ? src[i] : INFINITY;
    __syncthreads();
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < (n >> 1); i += kSortThreads) {
                const int lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                const int hi = lo + j;
                const bool asc = (lo & k) == 0;
                const double a = s[lo], b = s[hi];
                if ((a > b) == asc) { s[lo] = b; s[hi] = a; }
            }
            __syncthreads();
        }
    }
    double *dst = out + sg.start + off;
    for (int i = threadIdx.x; i < len; i += kSortThreads) dst[i] = s[i];
}

// Merge pass with LDS staging: each workgroup emits kMB consecutive outputs
// of one merged pair; the two input slices it needs are found by one merge
// path search per workgroup, loaded coalesced into LDS, merged there (each
// thread kMergeE outputs) and written back coalesced.
constexpr int kMB = 256 * kMergeE;

__device__ __forceinline__ int64_t merge_split(const double *A, int64_t n0, const double *B,
                                               int64_t n1, int64_t diag) {
    int64_t lo = max((int64_t)0, diag - n1), hi = min(diag, n0);
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (A[mid] <= B[diag - mid - 1]) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void merge_lds_kernel(const double *__restrict__ in,
                                                        double *__restrict__ out,
                                                        const Seg *__restrict__ segs, int64_t w) {
    __shared__ double sA[kMB], sB[kMB], sO[kMB];
    __shared__ int64_t sp[4];
    const Seg sg = segs[blockIdx.y];
    const int64_t o0 = (int64_t)blockIdx.x * kMB;
    if (o0 >= sg.len) return;
    const int t = threadIdx.x;
    const int64_t p = o0 / (2 * w), ps = p * 2 * w;
    const int64_t n0 = min(w, sg.len - ps), n1 = min(w, sg.len - ps - n0);
    const double *A = in + sg.start + ps;
    const double *B = A + n0;
    const int64_t d0 = o0 - ps, d1 = min(d0 + kMB, n0 + n1);
    if (t == 0) sp[0] = merge_split(A, n0, B, n1, d0);
    if (t == 64) sp[1] = merge_split(A, n0, B, n1, d1);
    __syncthreads();
    const int64_t a0 = sp[0], a1 = sp[1], b0 = d0 - a0, b1 = d1 - a1;
    const int na = (int)(a1 - a0), nb = (int)(b1 - b0), tot = (int)(d1 - d0);
    for (int i = t; i < na; i += 256) sA[i] = A[a0 + i];
    for (int i = t; i < nb; i += 256) sB[i] = B[b0 + i];
    __syncthreads();
    const int d = t * kMergeE;
    if (d < tot) {
        int lo = max(0, d - nb), hi = min(d, na);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sA[mid] <= sB[d - mid - 1]) lo = mid + 1; else hi = mid;
        }
        int i = lo, j = d - lo;
        const int cnt = min(kMergeE, tot - d);
        for (int q = 0; q < cnt; ++q) {
            const bool takeA = j >= nb || (i < na && sA[i] <= sB[j]);
            sO[d + q] = takeA ? sA[i] : sB[j];
            if (takeA) ++i; else ++j;
        }
    }
    __syncthreads();
    double *dst = out + sg.start + o0;
    for (int i = t; i < tot; i += 256) dst[i] = sO[i];
}

// ---------------------------------------------------------------------------
// Fit, fast form for bins that are sorted in place (bin segments == the
// fixed Nxb chunks).  findCut (resample.h:370-386) on a sorted bin reduces
// to c = #{K <= kmin}: Nxmin = c, except in the last bin where
// K[Nx-3] <= kmin (c >= Nxb-2) gives Nxb.  Each workgroup takes kFitChunk
// points of one bin and writes their nC partial moments; kfit_combine sums
// the partials in chunk order and applies P.
// ---------------------------------------------------------------------------
constexpr int kFitChunk = 4096;

__device__ __forceinline__ int nxmin_sorted(const double *Kb, int Nxb, bool last, double kmin) {
    int lo = 0, hi = Nxb;                  // first index with K > kmin
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (Kb[mid] <= kmin) lo = mid + 1; else hi = mid;
    }
    return (last && lo >= Nxb - 2) ? Nxb : lo;
}

__global__ __launch_bounds__(256) void kfit_partial_kernel(const double *__restrict__ Ksorted,
                                                           int64_t nx, int Nxb, int nbins, int nC,
                                                           double kmin, double unit_scale,
                                                           double *__restrict__ part,
                                                           double *__restrict__ b0out,
                                                           int32_t *__restrict__ nxmin_out) {
    __shared__ double red[kMaxNC][4];
    __shared__ int s_nxm;
    const int ch = blockIdx.x, b = blockIdx.y, sp = blockIdx.z, t = threadIdx.x;
    const int nch = gridDim.x;
    const double *Kb = Ksorted + (int64_t)sp * nx + (int64_t)b * Nxb;
    if (t == 0) s_nxm = nxmin_sorted(Kb, Nxb, b == nbins - 1, kmin);
    __syncthreads();
    const int nxm = s_nxm;
    const int64_t bid = (int64_t)sp * nbins + b;
    if (ch == 0 && t == 0) nxmin_out[bid] = nxm;
    double m[kMaxNC];
#pragma unroll
    for (int i = 0; i < kMaxNC; ++i) m[i] = 0.0;
    const double sres = 1.0 - nxm / ((double)(Nxb - 1));
    const double nlm1 = (double)Nxb - 1.0;
    const int k0 = ch * kFitChunk, k1 = min(Nxb, k0 + kFitChunk);
    for (int k = k0 + t; k < k1; k += 256) {
        double Kv;
        if (nxm > 0) {
            double ii = fma(sres, (double)k, (double)nxm);
            if (ii >= Nxb - 1) ii = 0.999999 * (Nxb - 1);
            const int il = (int)(ii / ((double)Nxb) * Nxb);
            const double Kl = Kb[il], Kr = Kb[il + 1];
            Kv = fma(Kr - Kl, ii - il, Kl);
        } else {
            Kv = Kb[k];
        }
        const double bv = log(Kv * unit_scale);
        if (k < nC) b0out[bid * nC + k] = bv;
        const double x = k * 2.0 / nlm1 - 1.0;
        double T0 = 1.0, T1 = x;
        m[0] += bv;
        if (nC > 1) m[1] = fma(x, bv, m[1]);
#pragma unroll
        for (int i = 2; i < kMaxNC; ++i) {
            if (i < nC) {
                const double T2 = fma(2.0 * x, T1, -T0);
                m[i] = fma(T2, bv, m[i]);
                T0 = T1; T1 = T2;
            }
        }
    }
    // wave sums (fixed butterfly), then the 4 waves in order
    const int lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int i = 0; i < kMaxNC; ++i) {
        if (i < nC) {
            double v = m[i];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) red[i][wv] = v;
        }
    }
    __syncthreads();
    if (t < nC) part[(bid * nch + ch) * nC + t] = ((red[t][0] + red[t][1]) + red[t][2]) + red[t][3];
}

__global__ __launch_bounds__(64) void kfit_combine_kernel(const double *__restrict__ part,
                                                          const double *__restrict__ b0in, int nch,
                                                          int nC, const double *__restrict__ P,
                                                          double *__restrict__ coeff) {
    __shared__ double mom[kMaxNC];
    const int64_t bid = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    const int t = threadIdx.x;
    if (t < nC) {
        double s = 0.0;
        for (int c = 0; c < nch; ++c) s += part[(bid * nch + c) * nC + t];
        mom[t] = s;
    }
    __syncthreads();
    if (t < nC) {
        const double *Pr = P + (int64_t)t * 2 * nC;
        const double *b0 = b0in + bid * nC;
        double c = 0.0;
        for (int i = 0; i < nC; ++i) c = fma(Pr[i], mom[i], c);
        for (int i = 0; i < nC; ++i) c = fma(Pr[nC + i], b0[i], c);
        coeff[bid * nC + t] = c;
    }
}

// ---------------------------------------------------------------------------
// Fit of one (spectrum, bin).  Ks: per-bin sorted K of the spectrum.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kFitThreads) void kfit_kernel(const double *__restrict__ Ksorted,
                                                          int64_t nx, int Nxb, int nbins, int nC,
                                                          double kmin, double unit_scale,
                                                          const double *__restrict__ P,
                                                          const double *__restrict__ HV,
                                                          double *__restrict__ coeff,
                                                          int32_t *__restrict__ nxmin_out) {
    __shared__ double red[kFitThreads];
    __shared__ double bl[kHHMaxNxb];
    __shared__ double mom[kMaxNC];
    __shared__ double b0[kMaxNC];
    __shared__ int s_r1[kFitThreads], s_r2[kFitThreads], s_r3[kFitThreads];
    const int b = blockIdx.x, sp = blockIdx.y, t = threadIdx.x;
    const double *Ks = Ksorted + (int64_t)sp * nx;
    const double *Kb = Ks + (int64_t)b * Nxb;
    // findCut rules (resample.h:370-386), global index id = b*Nxb + k
    int r1 = -1, r2 = 0, r3 = 0;
    for (int k = t; k < Nxb; k += kFitThreads) {
        const int64_t id = (int64_t)b * Nxb + k;
        if (id >= nx - 1) continue;
        const double K0 = Ks[id], K1 = Ks[id + 1];
        if (K0 <= kmin && K1 > kmin) r1 = max(r1, k + 1);
        if (K0 <= kmin && k == Nxb - 1) r2 = 1;
        if (K0 <= kmin && id == nx - 3) r3 = 1;
    }
    s_r1[t] = r1; s_r2[t] = r2; s_r3[t] = r3;
    __syncthreads();
    for (int s = kFitThreads / 2; s > 0; s >>= 1) {
        if (t < s) {
            s_r1[t] = max(s_r1[t], s_r1[t + s]);
            s_r2[t] |= s_r2[t + s];
            s_r3[t] |= s_r3[t + s];
        }
        __syncthreads();
    }
    const int nxm = s_r3[0] ? Nxb : (s_r2[0] ? Nxb : (s_r1[0] >= 0 ? s_r1[0] : 0));
    // rescale (f = +1), ln, moments
    double m[kMaxNC];
#pragma unroll
    for (int i = 0; i < kMaxNC; ++i) m[i] = 0.0;
    const double sres = 1.0 - nxm / ((double)(Nxb - 1));
    const double nlm1 = (double)Nxb - 1.0;
    for (int k = t; k < Nxb; k += kFitThreads) {
        double Kv;
        if (nxm > 0) {
            double ii = fma(sres, (double)k, (double)nxm);
            if (ii >= Nxb - 1) ii = 0.999999 * (Nxb - 1);
            const int il = (int)(ii / ((double)Nxb) * Nxb);
            const double Kl = Kb[il], Kr = Kb[il + 1];
            Kv = fma(Kr - Kl, ii - il, Kl);
        } else {
            Kv = Kb[k];
        }
        const double bv = log(Kv * unit_scale);
        if (k < nC) b0[k] = bv;
        if (HV) { bl[k] = bv; continue; }
        const double x = k * 2.0 / nlm1 - 1.0;
        double T0 = 1.0, T1 = x;
        m[0] += bv;
        if (nC > 1) m[1] = fma(x, bv, m[1]);
#pragma unroll
        for (int i = 2; i < kMaxNC; ++i) {
            if (i < nC) {
                const double T2 = fma(2.0 * x, T1, -T0);
                m[i] = fma(T2, bv, m[i]);
                T0 = T1; T1 = T2;
            }
        }
    }
    if (HV) {
        // literal leastSquare_kernel (resample.h:243-353) on the LDS copy of b;
        // HV = [V (Nxb x nC, Householder vectors) | C (nC) | D (nC)]
        __syncthreads();
        const double *Cc = HV + (int64_t)Nxb * nC, *Dd = Cc + nC;
        for (int j = 0; j < nC - 1; ++j) {
            const double *vj = HV + (int64_t)Nxb * j;
            double dot = 0.0;
            for (int k = j + t; k < Nxb; k += kFitThreads) dot = fma(vj[k], bl[k], dot);
            red[t] = dot;
            __syncthreads();
            for (int s2 = kFitThreads / 2; s2 > 0; s2 >>= 1) {
                if (t < s2) red[t] += red[t + s2];
                __syncthreads();
            }
            const double tq = red[0] / Cc[j];
            __syncthreads();
            for (int k = j + t; k < Nxb; k += kFitThreads) bl[k] -= tq * vj[k];
            __syncthreads();
        }
        if (t == 0) {
            bl[nC - 1] /= Dd[nC - 1];
            for (int i = nC - 2; i >= 0; --i) {
                double a = 0.0;
                for (int j = i + 1; j < nC; ++j) a += HV[i + (int64_t)Nxb * j] * bl[j];
                bl[i] = (bl[i] - a) / Dd[i];
            }
        }
        __syncthreads();
        if (t < nC) coeff[((int64_t)sp * nbins + b) * nC + t] = bl[t];
        if (t == 0) nxmin_out[(int64_t)sp * nbins + b] = nxm;
        return;
    }
    // deterministic tree reduction of each moment
    for (int i = 0; i < nC; ++i) {
        red[t] = m[i];
        __syncthreads();
        for (int s = kFitThreads / 2; s > 0; s >>= 1) {
            if (t < s) red[t] += red[t + s];
            __syncthreads();
        }
        if (t == 0) mom[i] = red[0];
        __syncthreads();
    }
    if (t < nC) {
        const double *Pr = P + (int64_t)t * 2 * nC;
        double c = 0.0;
        for (int i = 0; i < nC; ++i) c = fma(Pr[i], mom[i], c);
        for (int i = 0; i < nC; ++i) c = fma(Pr[nC + i], b0[i], c);
        coeff[((int64_t)sp * nbins + b) * nC + t] = c;
    }
    if (t == 0) nxmin_out[(int64_t)sp * nbins + b] = nxm;
}

// ---------------------------------------------------------------------------
// Reconstruction: expfx_kernel (resample.h:497-524) then rescale f=-1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void expfx_kernel(const double *__restrict__ coeff, int nC,
                                                    int Nxb, double unit_scale,
                                                    double *__restrict__ F) {
    __shared__ double cs[kMaxNC];
    const int64_t bin = blockIdx.x;
    if ((int)threadIdx.x < nC) cs[threadIdx.x] = coeff[bin * nC + threadIdx.x];
    __syncthreads();
    double *out = F + bin * Nxb;
    for (int k = threadIdx.x; k < Nxb; k += 256) {
        const double x = -1.0 + k * 2.0 / ((double)Nxb);
        double d1 = 0.0, d2 = 0.0;
        for (int i = nC - 1; i >= 1; --i) {
            const double tt = d1;
            d1 = fma(2.0 * x, d1, -d2) + cs[i];
            d2 = tt;
        }
        const double f = fma(x, d1, -d2) + 0.5 * cs[0] + 0.5 * cs[0];
        out[k] = exp(f) / unit_scale;
    }
}

__global__ __launch_bounds__(256) void rescale_inv_kernel(const double *__restrict__ F,
                                                          const int32_t *__restrict__ nxmin,
                                                          int Nxb, double kmin,
                                                          double *__restrict__ Ky) {
    const int64_t bin = blockIdx.x;
    const int n = nxmin[bin];
    const double *Fb = F + bin * Nxb;
    double *out = Ky + bin * Nxb;
    for (int k = threadIdx.x; k < Nxb; k += 256) {
        double Ki;
        if (n > 0) {
            Ki = kmin;
            if (k >= n) {
                double ii = (k - n) / ((double)(Nxb - n - 1)) * (Nxb - 1);
                if (ii >= Nxb - 1) ii = 0.999999 * (Nxb - 1);
                const int il = ii == ii ? (int)(ii / ((double)Nxb) * Nxb) : 0;
                const double Kl = Fb[il], Kr = Fb[il + 1];
                Ki = fma(Kr - Kl, ii - il, Kl);
            }
        } else {
            Ki = Fb[k];
        }
        out[k] = Ki;
    }
}

// ---------------------------------------------------------------------------
// Host: the least-squares map P (nC x 2nC).
// ---------------------------------------------------------------------------
struct QRPlan {
    std::vector<double> V, C, D;   // QR_kernel output, V column-major Nxb x nC
    std::vector<double> Van;       // plain Vandermonde (Chebyshev columns)
};

// Vandermonde_kernel + QR_kernel (resample.h:14-37, 73-227), sequential sums.
static void host_qr(int NL, int NC, QRPlan &q) {
    q.V.assign((size_t)NL * NC, 0.0);
    q.C.assign(NC, 0.0);
    q.D.assign(NC, 0.0);
    const double nl = (double)NL;
    for (int id = 0; id < NL; ++id) {
        const double x = id * 2.0 / (nl - 1.0) - 1.0;
        q.V[id] = 1.0;
        if (NC > 1) q.V[id + (size_t)NL] = x;
        double t = x, t0 = 1.0;
        for (int i = 2; i < NC; ++i) {
            const double tt = t;
            t = fma(2.0 * x, t, -t0);
            t0 = tt;
            q.V[id + (size_t)NL * i] = t;
        }
    }
    q.Van = q.V;
    for (int i = 0; i < NC; ++i) {
        double *vi = &q.V[(size_t)NL * i];
        double scale = 0.0;
        for (int k = i; k < NL; ++k) scale = fmax(scale, fabs(vi[k]));
        double ss = 0.0;
        for (int k = i; k < NL; ++k) { vi[k] /= scale; ss += vi[k] * vi[k]; }
        const double s = sqrt(ss), v = vi[i];
        const double sigma = v < 0.0 ? -s : s;
        vi[i] += sigma;
        q.C[i] = sigma * (v + sigma);
        q.D[i] = -scale * sigma;
        for (int j = i + 1; j < NC; ++j) {
            double *vj = &q.V[(size_t)NL * j];
            double dot = 0.0;
            for (int k = i; k < NL; ++k) dot += vi[k] * vj[k];
            const double tq = dot / q.C[i];
            for (int k = i; k < NL; ++k) vj[k] -= tq * vi[k];
        }
    }
}

// leastSquare_kernel (resample.h:243-353): NC-1 reflections, back substitution.
static void host_ls(const QRPlan &q, int NL, int NC, std::vector<double> &b, double *c) {
    for (int j = 0; j < NC - 1; ++j) {
        const double *vj = &q.V[(size_t)NL * j];
        double dot = 0.0;
        for (int k = j; k < NL; ++k) dot += vj[k] * b[k];
        const double t = dot / q.C[j];
        for (int k = j; k < NL; ++k) b[k] -= t * vj[k];
    }
    b[NC - 1] /= q.D[NC - 1];
    for (int i = NC - 2; i >= 0; --i) {
        double a = 0.0;
        for (int j = i + 1; j < NC; ++j) a += q.V[i + (size_t)NL * j] * b[j];
        b[i] = (b[i] - a) / q.D[i];
    }
    for (int i = 0; i < NC; ++i) c[i] = b[i];
}

static int compute_plan(int NL, int NC, std::vector<double> &P) {
    QRPlan q;
    host_qr(NL, NC, q);
    const int M = 2 * NC;
    // probe vectors: Chebyshev columns, then unit vectors e_0..e_{NC-1}
    std::vector<long double> Z((size_t)M * M), Cp((size_t)NC * M);
    std::vector<double> b(NL), c(NC);
    for (int p = 0; p < M; ++p) {
        if (p < NC) for (int k = 0; k < NL; ++k) b[k] = q.Van[(size_t)NL * p + k];
        else { std::fill(b.begin(), b.end(), 0.0); b[p - NC] = 1.0; }
        // features z = [m ; b0] of the probe
        for (int i = 0; i < NC; ++i) {
            long double m = 0.0L;
            for (int k = 0; k < NL; ++k) m += (long double)q.Van[(size_t)NL * i + k] * b[k];
            Z[(size_t)i * M + p] = m;
            Z[(size_t)(NC + i) * M + p] = b[i];
        }
        std::vector<double> bb = b;
        host_ls(q, NL, NC, bb, c.data());
        for (int i = 0; i < NC; ++i) Cp[(size_t)i * M + p] = c[i];
    }
    // P Z = Cp  ->  Z^T P^T = Cp^T ; LU with partial pivoting on Z^T.
    std::vector<long double> A((size_t)M * M);
    for (int r = 0; r < M; ++r)
        for (int cc = 0; cc < M; ++cc) A[(size_t)r * M + cc] = Z[(size_t)cc * M + r];
    std::vector<int> piv(M);
    for (int k = 0; k < M; ++k) {
        int pmax = k;
        for (int r = k + 1; r < M; ++r)
            if (fabsl(A[(size_t)r * M + k]) > fabsl(A[(size_t)pmax * M + k])) pmax = r;
        if (A[(size_t)pmax * M + k] == 0.0L) return HK_ERR_ARG;
        piv[k] = pmax;
        if (pmax != k)
            for (int cc = 0; cc < M; ++cc) std::swap(A[(size_t)k * M + cc], A[(size_t)pmax * M + cc]);
        for (int r = k + 1; r < M; ++r) {
            const long double f = A[(size_t)r * M + k] / A[(size_t)k * M + k];
            A[(size_t)r * M + k] = f;
            for (int cc = k + 1; cc < M; ++cc) A[(size_t)r * M + cc] -= f * A[(size_t)k * M + cc];
        }
    }
    P.assign((size_t)NC * M, 0.0);
    std::vector<long double> rhs(M);
    for (int i = 0; i < NC; ++i) {
        for (int r = 0; r < M; ++r) rhs[r] = Cp[(size_t)i * M + r];
        for (int k = 0; k < M; ++k) std::swap(rhs[k], rhs[piv[k]]);
        for (int r = 1; r < M; ++r)
            for (int k = 0; k < r; ++k) rhs[r] -= A[(size_t)r * M + k] * rhs[k];
        for (int r = M - 1; r >= 0; --r) {
            for (int k = r + 1; k < M; ++k) rhs[r] -= A[(size_t)r * M + k] * rhs[k];
            rhs[r] /= A[(size_t)r * M + r];
        }
        for (int r = 0; r < M; ++r) P[(size_t)i * M + r] = (double)rhs[r];
    }
    return HK_OK;
}

static std::mutex g_plan_mu;
static std::map<std::pair<int, int>, std::vector<double>> g_plans;

static int get_plan(int Nxb, int nC, std::vector<double> &P) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto key = std::make_pair(Nxb, nC);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) { P = it->second; return HK_OK; }
    int e = compute_plan(Nxb, nC, P);
    if (e == HK_OK) g_plans[key] = P;
    return e;
}

// Workspace layout.
struct KtLayout {
    size_t off_seg, off_plan, off_hv, off_segs, off_part, off_b0, off_sorted, off_tmp, total;
    bool householder;
    int nch;
};

static bool use_householder(int64_t Nxb, int nC) { return Nxb < (int64_t)kPlanMinRatio * nC; }

static KtLayout kt_layout(int64_t nx, int nbins, int nC, int nspec) {
    KtLayout L;
    size_t o = 0;
    const int64_t Nxb = nx / nbins;
    L.householder = use_householder(Nxb, nC);
    L.off_seg = o; o = align_up(o + sizeof(int64_t) * (nbins + 2), 256);
    L.off_plan = o; o = align_up(o + sizeof(double) * 2 * nC * nC, 256);
    L.off_hv = o;
    if (L.householder) o = align_up(o + sizeof(double) * (size_t)(Nxb * nC + 2 * nC), 256);
    L.off_segs = o; o = align_up(o + sizeof(Seg) * (size_t)nspec * (nbins + 1), 256);
    L.nch = (int)((Nxb + kFitChunk - 1) / kFitChunk);
    L.off_part = o; o = align_up(o + sizeof(double) * (size_t)nspec * nbins * L.nch * nC, 256);
    L.off_b0 = o; o = align_up(o + sizeof(double) * (size_t)nspec * nbins * nC, 256);
    L.off_sorted = o; o = align_up(o + sizeof(double) * (size_t)nspec * nx, 256);
    L.off_tmp = o; o = align_up(o + sizeof(double) * (size_t)nspec * nx, 256);
    L.total = o + 256;
    return L;
}

}  // namespace hk

using namespace hk;

// binKey semantics (voigt.h:779-806): key(j) = first bin b with
// bnd[b] <= x_j < bnd[b+1], nbins if x_j >= numax; bnd[b] computed as in
// heliosk.cu:438-441.
extern "C" int hk_bin_segments(const hk_grid *g, double numax, int nbins, int64_t *seg) {
    HK_CHECK_ARG(g && seg && nbins > 0, "hk_bin_segments: bad arguments");
    std::vector<double> bnd(nbins + 1);
    for (int i = 0; i < nbins; ++i) bnd[i] = g->numin + i * (numax - g->numin) / ((double)nbins);
    bnd[nbins] = numax;
    auto key = [&](int64_t j) {
        const double xj = fma((double)j, g->dnu, g->numin);
        int k = 0;
        for (int i = 0; i < nbins; ++i)
            if (bnd[i] <= xj && xj < bnd[i + 1]) { k = i; break; }
        if (xj >= numax) k = nbins;
        return k;
    };
    for (int b = 0; b <= nbins + 1; ++b) {
        int64_t lo = 0, hi = g->nx;
        while (lo < hi) {
            const int64_t m = (lo + hi) / 2;
            if (key(m) < b) lo = m + 1; else hi = m;
        }
        seg[b] = lo;
    }
    return HK_OK;
}

extern "C" int hk_ktable_plan_host(int Nxb, int nC, double *P) {
    HK_CHECK_ARG(P && Nxb > nC && nC >= 2 && nC <= kMaxNC,
                 "hk_ktable_plan_host: need 2 <= nC <= %d and Nxb > nC", kMaxNC);
    std::vector<double> v;
    int e = get_plan(Nxb, nC, v);
    if (e) { set_error("hk_ktable_plan_host: singular probe system"); return e; }
    memcpy(P, v.data(), v.size() * sizeof(double));
    return HK_OK;
}

extern "C" size_t hk_ktable_workspace_bytes(int64_t nx, int nbins, int nC, int nspec) {
    if (nx <= 0 || nbins <= 0 || nC <= 0 || nspec <= 0) return 0;
    return kt_layout(nx, nbins, nC, nspec).total;
}

extern "C" int hk_ktable_prepare(const hk_grid *g, double numax, int nbins, int nC, void *ws,
                                 size_t ws_bytes) {
    HK_CHECK_ARG(g && ws && nbins > 0, "hk_ktable_prepare: bad arguments");
    HK_CHECK_ARG(g->nx % nbins == 0, "hk_ktable_prepare: Nx %lld not divisible by nbins %d",
                 (long long)g->nx, nbins);
    const int Nxb = (int)(g->nx / nbins);
    HK_CHECK_ARG(nC >= 2 && nC <= kMaxNC && Nxb > nC,
                 "hk_ktable_prepare: need 2 <= nC <= %d < Nxb", kMaxNC);
    KtLayout L = kt_layout(g->nx, nbins, nC, 1);
    HK_CHECK_ARG(ws_bytes >= L.off_segs + 256, "hk_ktable_prepare: workspace too small");
    std::vector<int64_t> seg(nbins + 2);
    if (int e = hk_bin_segments(g, numax, nbins, seg.data())) return e;
    std::vector<double> P;
    if (int e = get_plan(Nxb, nC, P)) { set_error("hk_ktable_prepare: singular probe system"); return e; }
    char *base = (char *)align_up((size_t)ws, 256);
    HK_HIP(hipMemcpy(base + L.off_seg, seg.data(), seg.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HK_HIP(hipMemcpy(base + L.off_plan, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice));
    if (L.householder) {
        QRPlan q;
        host_qr(Nxb, nC, q);
        std::vector<double> hv(q.V);
        hv.insert(hv.end(), q.C.begin(), q.C.end());
        hv.insert(hv.end(), q.D.begin(), q.D.end());
        HK_HIP(hipMemcpy(base + L.off_hv, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return HK_OK;
}

namespace {
__global__ void build_segs_kernel(const int64_t *__restrict__ seg, int nbins, int64_t nx, int nspec,
                                  Seg *__restrict__ segs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nper = nbins + 1;
    if (i >= nspec * nper) return;
    const int s = i / nper, b = i % nper;
    Seg sg;
    sg.start = (int64_t)s * nx + seg[b];
    sg.len = seg[b + 1] - seg[b];
    segs[i] = sg;
}

}  // namespace

extern "C" int hk_ktable(const double *K, int nspec, const hk_grid *g, double numax, int nbins,
                         int nC, double kmin, double unit_scale, double *coeff, int32_t *nxmin,
                         double *Ksorted, void *ws, size_t ws_bytes, hk_stream_t stream) {
    HK_CHECK_ARG(K && g && coeff && nxmin && ws && nspec > 0 && nbins > 0,
                 "hk_ktable: bad arguments");
    HK_CHECK_ARG(g->nx % nbins == 0, "hk_ktable: Nx %lld not divisible by nbins %d",
                 (long long)g->nx, nbins);
    const int Nxb = (int)(g->nx / nbins);
    HK_CHECK_ARG(nC >= 2 && nC <= kMaxNC && Nxb > nC, "hk_ktable: need 2 <= nC <= %d < Nxb", kMaxNC);
    KtLayout L = kt_layout(g->nx, nbins, nC, nspec);
    if (ws_bytes < L.total) {
        set_error("hk_ktable: workspace %zu bytes, %zu needed", ws_bytes, L.total);
        return HK_ERR_WORKSPACE;
    }
    (void)numax;
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)align_up((size_t)ws, 256);
    const int64_t *seg = (const int64_t *)(base + L.off_seg);
    const double *P = (const double *)(base + L.off_plan);
    Seg *segs = (Seg *)(base + L.off_segs);
    double *sorted = Ksorted ? Ksorted : (double *)(base + L.off_sorted);
    double *tmp = (double *)(base + L.off_tmp);
    const int nsegs = nspec * (nbins + 1);
    hipLaunchKernelGGL(build_segs_kernel, dim3((nsegs + 255) / 256), dim3(256), 0, s, seg, nbins,
                       g->nx, nspec, segs);
    HK_LAUNCH_CHECK();
    // longest segment (host copy of seg is cheap: recompute)
    std::vector<int64_t> hseg(nbins + 2);
    if (int e = hk_bin_segments(g, numax, nbins, hseg.data())) return e;
    int64_t maxlen = 0;
    for (int b = 0; b <= nbins; ++b) maxlen = std::max(maxlen, hseg[b + 1] - hseg[b]);
    // segments equal to the fixed Nxb chunks (the usual case): sorted bins can
    // use the chunked fit; otherwise the literal per-bin rules
    bool aligned = true;
    for (int b = 0; b <= nbins; ++b)
        if (hseg[b] != (int64_t)b * Nxb) aligned = false;
    if (maxlen <= kSortLDS) {
        hipLaunchKernelGGL((bitonic_kernel<kSortLDS>), dim3(1, nsegs), dim3(kSortThreads), 0, s, K,
                           sorted, segs, (int64_t)kSortLDS);
        HK_LAUNCH_CHECK();
    } else {
        const int64_t chunk = kSortLDS;
        const int64_t nchunks = (maxlen + chunk - 1) / chunk;
        int npass = 0;
        for (int64_t w = chunk; w < maxlen; w *= 2) ++npass;
        // arrange the ping-pong so the last pass writes into `sorted`
        double *first = (npass % 2 == 0) ? sorted : tmp;
        double *other = (npass % 2 == 0) ? tmp : sorted;
        hipLaunchKernelGGL((bitonic_kernel<kSortLDS>), dim3((unsigned)nchunks, nsegs),
                           dim3(kSortThreads), 0, s, K, first, segs, chunk);
        HK_LAUNCH_CHECK();
        double *src = first, *dst = other;
        for (int64_t w = chunk; w < maxlen; w *= 2) {
            hipLaunchKernelGGL(merge_lds_kernel, dim3((unsigned)((maxlen + kMB - 1) / kMB), nsegs),
                               dim3(256), 0, s, src, dst, segs, w);
            HK_LAUNCH_CHECK();
            std::swap(src, dst);
        }
    }
    const double *HV = L.householder ? (const double *)(base + L.off_hv) : nullptr;
    if (aligned && !L.householder && Nxb >= 4) {
        double *part = (double *)(base + L.off_part);
        double *b0 = (double *)(base + L.off_b0);
        hipLaunchKernelGGL(kfit_partial_kernel, dim3(L.nch, nbins, nspec), dim3(256), 0, s, sorted,
                           g->nx, Nxb, nbins, nC, kmin, unit_scale, part, b0, nxmin);
        HK_LAUNCH_CHECK();
        hipLaunchKernelGGL(kfit_combine_kernel, dim3(nbins, nspec), dim3(64), 0, s, part, b0, L.nch,
                           nC, P, coeff);
        HK_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(kfit_kernel, dim3(nbins, nspec), dim3(kFitThreads), 0, s, sorted, g->nx,
                           Nxb, nbins, nC, kmin, unit_scale, P, HV, coeff, nxmin);
        HK_LAUNCH_CHECK();
    }
    return HK_OK;
}

// Ascending sort of arbitrary bins [seg[b], seg[b+1]) of one spectrum (the
// per-bin thrust sorts of heliosk.cu:2046-2051 for bins-file and per-bin
// grids).  Workspace: the segment table and one spectrum for the merge
// ping-pong.
extern "C" size_t hk_sort_bins_workspace_bytes(int64_t nx, int nseg) {
    if (nx <= 0 || nseg <= 0) return 0;
    return align_up(sizeof(Seg) * (size_t)nseg, 256) + align_up(sizeof(double) * (size_t)nx, 256) +
           256;
}

extern "C" int hk_sort_bins(const double *K, int64_t nx, const int64_t *seg, int nseg,
                            double *Ksorted, void *ws, size_t ws_bytes, hk_stream_t stream) {
    HK_CHECK_ARG(K && seg && Ksorted && ws && nx > 0 && nseg > 0 && nseg < 65536,
                 "hk_sort_bins: bad arguments");
    HK_CHECK_ARG(K != Ksorted, "hk_sort_bins: in-place sort is not supported");
    HK_CHECK_ARG(seg[0] >= 0 && seg[nseg] <= nx, "hk_sort_bins: segments outside [0, nx)");
    int64_t maxlen = 0;
    for (int b = 0; b < nseg; ++b) {
        HK_CHECK_ARG(seg[b + 1] >= seg[b], "hk_sort_bins: segment %d has negative length", b);
        maxlen = std::max(maxlen, seg[b + 1] - seg[b]);
    }
    const size_t need = hk_sort_bins_workspace_bytes(nx, nseg);
    if (ws_bytes < need) {
        set_error("hk_sort_bins: workspace %zu bytes, %zu needed", ws_bytes, need);
        return HK_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)align_up((size_t)ws, 256);
    Seg *segs = (Seg *)base;
    double *tmp = (double *)(base + align_up(sizeof(Seg) * (size_t)nseg, 256));
    std::vector<Seg> hs(nseg);
    for (int b = 0; b < nseg; ++b) { hs[b].start = seg[b]; hs[b].len = seg[b + 1] - seg[b]; }
    HK_HIP(hipMemcpyAsync(segs, hs.data(), hs.size() * sizeof(Seg), hipMemcpyHostToDevice, s));
    HK_HIP(hipStreamSynchronize(s));         // hs is host stack memory
    if (maxlen == 0) return HK_OK;
    if (maxlen <= kSortLDS) {
        hipLaunchKernelGGL((bitonic_kernel<kSortLDS>), dim3(1, nseg), dim3(kSortThreads), 0, s, K,
                           Ksorted, segs, (int64_t)kSortLDS);
        HK_LAUNCH_CHECK();
        return HK_OK;
    }
    const int64_t chunk = kSortLDS;
    const int64_t nchunks = (maxlen + chunk - 1) / chunk;
    int npass = 0;
    for (int64_t w = chunk; w < maxlen; w *= 2) ++npass;
    double *first = (npass % 2 == 0) ? Ksorted : tmp;
    double *other = (npass % 2 == 0) ? tmp : Ksorted;
    hipLaunchKernelGGL((bitonic_kernel<kSortLDS>), dim3((unsigned)nchunks, nseg),
                       dim3(kSortThreads), 0, s, K, first, segs, chunk);
    HK_LAUNCH_CHECK();
    double *src = first, *dst = other;
    for (int64_t w = chunk; w < maxlen; w *= 2) {
        hipLaunchKernelGGL(merge_lds_kernel, dim3((unsigned)((maxlen + kMB - 1) / kMB), nseg),
                           dim3(256), 0, s, src, dst, segs, w);
        HK_LAUNCH_CHECK();
        std::swap(src, dst);
    }
    return HK_OK;
}

extern "C" int hk_ktable_reconstruct(const double *coeff, const int32_t *nxmin, int nspec,
                                     int nbins, int Nxb, int nC, double kmin, double unit_scale,
                                     double *Ky, double *scratch, hk_stream_t stream) {
    HK_CHECK_ARG(coeff && nxmin && Ky && scratch && nspec > 0 && nbins > 0 && Nxb > 1,
                 "hk_ktable_reconstruct: bad arguments");
    HK_CHECK_ARG(nC >= 1 && nC <= kMaxNC, "hk_ktable_reconstruct: nC %d", nC);
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = (unsigned)(nspec * nbins);
    hipLaunchKernelGGL(expfx_kernel, dim3(nb), dim3(256), 0, s, coeff, nC, Nxb, unit_scale, scratch);
    HK_LAUNCH_CHECK();
    hipLaunchKernelGGL(rescale_inv_kernel, dim3(nb), dim3(256), 0, s, scratch, nxmin, Nxb, kmin, Ky);
    HK_LAUNCH_CHECK();
    return HK_OK;
}
