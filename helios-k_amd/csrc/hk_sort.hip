// hk_sort.hip -- order a raw line list by centre on the device.
//
// The reference sorts each chunk of lines by nu with thrust::sort_by_key and
// then gathers every field through the permutation (Sort_kernel /
// Copy_kernel, heliosk.cu:1147-1171, voigt.h:843-866).  Here the (nu, line
// index) pairs are merge-sorted in lexicographic order, so lines with equal
// centres keep their input order (a stable sort by nu):
//   * pair_block_sort_kernel: runs of 8192 pairs sorted by one workgroup in
//     LDS -- 8 pairs per thread sorted in registers by a bitonic network,
//     then merge rounds in LDS, each thread placing its 8 outputs after a
//     merge-path search;
//   * pair_merge_kernel: runs of w pairs merged into runs of 2w; each
//     workgroup emits 2048 outputs of one merged pair after one merge-path
//     split, with both input slices staged in LDS;
//   * the last pass writes the sorted centres into the output list and the
//     permutation, and gather_kernel moves the other seven fields.
#include <climits>

#include <algorithm>

#include "hk_internal.h"

namespace {

constexpr int kNT = 1024;                 // threads of the block sort
constexpr int kIPT = 8;                   // pairs per thread
constexpr int kRun = kNT * kIPT;          // pairs per LDS-sorted run (8192)
constexpr int kMergeE = 8;                // outputs per thread in a merge pass
constexpr int kMB = 256 * kMergeE;        // outputs per merge workgroup (2048)
constexpr int64_t kMaxSortLines = 0x7fffffff;

// (key, index) order: by key, then by index.
__device__ __forceinline__ bool pless(double ka, int32_t ia, double kb, int32_t ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// LDS slot of pair i: one spare slot after every kIPT pairs, so the pairs
// one thread owns start kIPT + 1 slots after its neighbour's.
__device__ __forceinline__ int slot(int i) { return i + i / kIPT; }
constexpr int kSlots = kRun + kRun / kIPT;

__global__ __launch_bounds__(kNT) void pair_block_sort_kernel(const double *__restrict__ key,
                                                              int64_t n,
                                                              double *__restrict__ okey,
                                                              int32_t *__restrict__ oidx) {
    __shared__ double sk[kSlots];
    __shared__ int32_t si[kSlots];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kRun;
    const int len = (int)std::min<int64_t>(kRun, n - base);
    for (int i = t; i < len; i += kNT) {
        sk[slot(i)] = key[base + i];
        si[slot(i)] = (int32_t)(base + i);
    }
    __syncthreads();
    const int np = (len + kIPT - 1) / kIPT * kIPT;
    const int d0 = t * kIPT;
    const bool act = d0 < np;
    double vk[kIPT];
    int32_t vi[kIPT];
    if (act) {
#pragma unroll
        for (int q = 0; q < kIPT; ++q) {
            const bool in = d0 + q < len;
            vk[q] = in ? sk[slot(d0 + q)] : INFINITY;
            vi[q] = in ? si[slot(d0 + q)] : INT_MAX;
        }
#pragma unroll
        for (int k = 2; k <= kIPT; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
                for (int i = 0; i < kIPT; ++i) {
                    const int l = i ^ j;
                    if (l > i) {
                        const bool sw = ((i & k) == 0) == pless(vk[l], vi[l], vk[i], vi[i]);
                        if (sw) {
                            const double tk = vk[i]; vk[i] = vk[l]; vk[l] = tk;
                            const int32_t ti = vi[i]; vi[i] = vi[l]; vi[l] = ti;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int q = 0; q < kIPT; ++q) { sk[slot(d0 + q)] = vk[q]; si[slot(d0 + q)] = vi[q]; }
    }
    for (int L = kIPT; L < np; L <<= 1) {
        __syncthreads();
        if (act) {
            const int ps = d0 / (2 * L) * (2 * L);
            const int na = min(L, np - ps), nb = min(L, np - ps - na);
            const int diag = d0 - ps, a0 = ps, b0 = ps + na;
            int lo = max(0, diag - nb), hi = min(diag, na);
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                const int bm = b0 + diag - mid - 1;
                if (!pless(sk[slot(bm)], si[slot(bm)], sk[slot(a0 + mid)], si[slot(a0 + mid)]))
                    lo = mid + 1;
                else
                    hi = mid;
            }
            int a = lo, b = diag - lo;
            double ka = a < na ? sk[slot(a0 + a)] : INFINITY, kb = b < nb ? sk[slot(b0 + b)] : INFINITY;
            int32_t ia = a < na ? si[slot(a0 + a)] : INT_MAX, ib = b < nb ? si[slot(b0 + b)] : INT_MAX;
#pragma unroll
            for (int q = 0; q < kIPT; ++q) {
                const bool ta = b >= nb || (a < na && !pless(kb, ib, ka, ia));
                vk[q] = ta ? ka : kb;
                vi[q] = ta ? ia : ib;
                if (ta) {
                    ++a;
                    ka = a < na ? sk[slot(a0 + a)] : INFINITY;
                    ia = a < na ? si[slot(a0 + a)] : INT_MAX;
                } else {
                    ++b;
                    kb = b < nb ? sk[slot(b0 + b)] : INFINITY;
                    ib = b < nb ? si[slot(b0 + b)] : INT_MAX;
                }
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < kIPT; ++q) { sk[slot(d0 + q)] = vk[q]; si[slot(d0 + q)] = vi[q]; }
        }
    }
    __syncthreads();
    for (int i = t; i < len; i += kNT) {
        okey[base + i] = sk[slot(i)];
        oidx[base + i] = si[slot(i)];
    }
}

// First a in [max(0, diag-nb), min(diag, na)] with A[a] after B[diag-a-1].
__device__ __forceinline__ int64_t pair_split(const double *Ak, const int32_t *Ai, int64_t na,
                                              const double *Bk, const int32_t *Bi, int64_t nb,
                                              int64_t diag) {
    int64_t lo = std::max<int64_t>(0, diag - nb), hi = std::min<int64_t>(diag, na);
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int64_t bm = diag - mid - 1;
        if (!pless(Bk[bm], Bi[bm], Ak[mid], Ai[mid])) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void pair_merge_kernel(const double *__restrict__ ik,
                                                         const int32_t *__restrict__ ii,
                                                         double *__restrict__ ok,
                                                         int32_t *__restrict__ oi, int64_t n,
                                                         int64_t w) {
    __shared__ double sAk[kMB], sBk[kMB];
    __shared__ int32_t sAi[kMB], sBi[kMB];
    __shared__ int64_t sp[2];
    const int t = threadIdx.x;
    const int64_t o0 = (int64_t)blockIdx.x * kMB;
    if (o0 >= n) return;
    const int64_t ps = o0 / (2 * w) * (2 * w);
    const int64_t na = std::min(w, n - ps), nb = std::min(w, n - ps - na);
    const double *Ak = ik + ps, *Bk = Ak + na;
    const int32_t *Ai = ii + ps, *Bi = Ai + na;
    const int64_t d0 = o0 - ps, d1 = std::min(d0 + kMB, na + nb);
    if (t == 0) sp[0] = pair_split(Ak, Ai, na, Bk, Bi, nb, d0);
    if (t == 64) sp[1] = pair_split(Ak, Ai, na, Bk, Bi, nb, d1);
    __syncthreads();
    const int64_t a0 = sp[0], a1 = sp[1], b0 = d0 - a0, b1 = d1 - a1;
    const int la = (int)(a1 - a0), lb = (int)(b1 - b0), tot = (int)(d1 - d0);
    for (int i = t; i < la; i += 256) { sAk[i] = Ak[a0 + i]; sAi[i] = Ai[a0 + i]; }
    for (int i = t; i < lb; i += 256) { sBk[i] = Bk[b0 + i]; sBi[i] = Bi[b0 + i]; }
    __syncthreads();
    const int d = t * kMergeE;
    double vk[kMergeE];
    int32_t vi[kMergeE];
    int cnt = 0;
    if (d < tot) {
        int lo = max(0, d - lb), hi = min(d, la);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const int bm = d - mid - 1;
            if (!pless(sBk[bm], sBi[bm], sAk[mid], sAi[mid])) lo = mid + 1; else hi = mid;
        }
        int a = lo, b = d - lo;
        cnt = min(kMergeE, tot - d);
        for (int q = 0; q < cnt; ++q) {
            const bool ta = b >= lb || (a < la && !pless(sBk[b], sBi[b], sAk[a], sAi[a]));
            vk[q] = ta ? sAk[a] : sBk[b];
            vi[q] = ta ? sAi[a] : sBi[b];
            if (ta) ++a; else ++b;
        }
    }
    // outputs are written from registers: the slices in LDS are not reused
    for (int q = 0; q < cnt; ++q) {
        ok[o0 + d + q] = vk[q];
        oi[o0 + d + q] = vi[q];
    }
}

struct Fields {
    const double *src[6];
    double *dst[6];
    const int32_t *iso_src;
    int32_t *iso_dst;
};

// out_f[i] = in_f[perm[i]] for S, EL, A, delta, vy, n and iso.
__global__ __launch_bounds__(256) void gather_kernel(const int32_t *__restrict__ perm, int64_t n,
                                                     Fields f) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        const int64_t j = perm[i];
#pragma unroll
        for (int k = 0; k < 6; ++k) f.dst[k][i] = f.src[k][j];
        if (f.iso_src) f.iso_dst[i] = f.iso_src[j];
    }
}

int grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

int merge_passes(int64_t n) {
    int p = 0;
    for (int64_t w = kRun; w < n; w *= 2) ++p;
    return p;
}

}  // namespace

extern "C" size_t hk_lines_sort_workspace_bytes(int64_t nlines) {
    if (nlines < 0 || nlines > kMaxSortLines) return 0;
    // two ping-pong pairs of (key, index) arrays
    const size_t kb = hk::align_up((size_t)nlines * sizeof(double), 256);
    const size_t ib = hk::align_up((size_t)nlines * sizeof(int32_t), 256);
    return 2 * (kb + ib) + 256;
}

extern "C" int hk_lines_sort(const hk_lines *in, hk_lines *out, void *workspace, size_t bytes,
                             hk_stream_t stream) {
    HK_CHECK_ARG(in && out, "hk_lines_sort: null argument");
    const int64_t n = in->nlines;
    HK_CHECK_ARG(n >= 0 && n <= kMaxSortLines, "hk_lines_sort: %lld lines (at most %lld per call)",
                 (long long)n, (long long)kMaxSortLines);
    HK_CHECK_ARG(in->form >= HK_FORM_HITRAN && in->form <= HK_FORM_SUPERLINE,
                 "hk_lines_sort: unknown line form %d", in->form);
    const double *src[7] = {in->nu, in->S, in->EL, in->A, in->delta, in->vy, in->n};
    const double *dst[7] = {out->nu, out->S, out->EL, out->A, out->delta, out->vy, out->n};
    if (n > 0) {
        for (int k = 0; k < 7; ++k) {
            HK_CHECK_ARG(src[k] && dst[k], "hk_lines_sort: field %d is NULL", k);
            HK_CHECK_ARG(src[k] != dst[k], "hk_lines_sort: field %d sorted in place", k);
        }
        HK_CHECK_ARG(!in->iso || (out->iso && out->iso != in->iso),
                     "hk_lines_sort: iso needs a distinct output array");
    }
    const size_t need = hk_lines_sort_workspace_bytes(n);
    HK_CHECK_ARG(need > 0, "hk_lines_sort: workspace size query failed");
    if (bytes < need) {
        hk::set_error("hk_lines_sort: workspace of %zu bytes, %zu needed", bytes, need);
        return HK_ERR_WORKSPACE;
    }
    HK_CHECK_ARG(workspace, "hk_lines_sort: null workspace");
    // scalar fields travel with the lines
    const int niso = in->niso, form = in->form;
    double mass[HK_MAX_ISO], ab[HK_MAX_ISO], ss[HK_MAX_ISO];
    for (int k = 0; k < HK_MAX_ISO; ++k) {
        mass[k] = in->iso_mass[k]; ab[k] = in->iso_ab[k]; ss[k] = in->iso_sscale[k];
    }
    const double mad = in->max_abs_delta;
    const int32_t *iso_src = in->iso;
    out->nlines = n;
    out->niso = niso;
    out->form = form;
    for (int k = 0; k < HK_MAX_ISO; ++k) {
        out->iso_mass[k] = mass[k]; out->iso_ab[k] = ab[k]; out->iso_sscale[k] = ss[k];
    }
    out->max_abs_delta = mad;
    if (n == 0) return HK_OK;

    hipStream_t s = (hipStream_t)stream;
    char *w = (char *)hk::align_up((size_t)workspace, 256);
    const size_t kb = hk::align_up((size_t)n * sizeof(double), 256);
    const size_t ib = hk::align_up((size_t)n * sizeof(int32_t), 256);
    double *k0 = (double *)w;
    int32_t *i0 = (int32_t *)(w + kb);
    double *k1 = (double *)(w + kb + ib);
    int32_t *i1 = (int32_t *)(w + 2 * kb + ib);
    // the last pass writes the sorted centres into out->nu
    const int npass = merge_passes(n);
    double *kfirst = npass % 2 == 0 ? (double *)dst[0] : k0;
    const unsigned nrun = (unsigned)((n + kRun - 1) / kRun);
    hipLaunchKernelGGL(pair_block_sort_kernel, dim3(nrun), dim3(kNT), 0, s, src[0], n, kfirst, i0);
    HK_LAUNCH_CHECK();
    double *ck = kfirst, *nk = npass % 2 == 0 ? k0 : (double *)dst[0];
    int32_t *ci = i0, *ni = i1;
    const unsigned nblk = (unsigned)((n + kMB - 1) / kMB);
    for (int64_t wl = kRun; wl < n; wl *= 2) {
        hipLaunchKernelGGL(pair_merge_kernel, dim3(nblk), dim3(256), 0, s, ck, ci, nk, ni, n, wl);
        HK_LAUNCH_CHECK();
        std::swap(ck, nk);
        std::swap(ci, ni);
    }
    Fields f;
    for (int k = 0; k < 6; ++k) {
        f.src[k] = src[k + 1];
        f.dst[k] = (double *)dst[k + 1];
    }
    f.iso_src = iso_src;
    f.iso_dst = iso_src ? (int32_t *)out->iso : nullptr;
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n)), dim3(256), 0, s, ci, n, f);
    HK_LAUNCH_CHECK();
    return HK_OK;
}
