// hk_sort.hip -- order a raw line list by centre on the device.
//
// The reference sorts each chunk of lines by nu with thrust::sort_by_key and
// then gathers every field through the permutation (Sort_kernel /
// Copy_kernel, heliosk.cu:1147-1171, voigt.h:843-866).  Here: one stable LSD
// radix sort of (nu, line index) pairs (rocPRIM through hipCUB) writes the
// sorted centres straight into the output list, and one gather kernel moves
// the other seven fields.  Stability makes the order identical to a stable
// comparison sort: lines with equal centres keep their input order.
#include <hipcub/hipcub.hpp>

#include "hk_internal.h"

namespace {

__global__ __launch_bounds__(256) void iota_kernel(int32_t *__restrict__ v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        v[i] = (int32_t)i;
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

constexpr int64_t kMaxSortLines = 0x7fffffff;

int grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

// Radix-sort scratch of n pairs (0 when the query fails).
size_t radix_bytes(int64_t n) {
    size_t t = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, t, (const double *)nullptr, (double *)nullptr,
                                           (const int32_t *)nullptr, (int32_t *)nullptr, (int)n,
                                           0, 64, (hipStream_t)0) != hipSuccess)
        return 0;
    return t;
}

}  // namespace

extern "C" size_t hk_lines_sort_workspace_bytes(int64_t nlines) {
    if (nlines < 0 || nlines > kMaxSortLines) return 0;
    const size_t rb = radix_bytes(nlines);
    if (nlines > 0 && rb == 0) return 0;
    // permutation in and out, then the radix scratch; each 256-byte aligned
    return 2 * hk::align_up((size_t)nlines * sizeof(int32_t), 256) + rb + 3 * 256;
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
    const int64_t nl = n;
    const int niso = in->niso, form = in->form;
    double mass[HK_MAX_ISO], ab[HK_MAX_ISO], ss[HK_MAX_ISO];
    for (int k = 0; k < HK_MAX_ISO; ++k) {
        mass[k] = in->iso_mass[k]; ab[k] = in->iso_ab[k]; ss[k] = in->iso_sscale[k];
    }
    const double mad = in->max_abs_delta;
    const int32_t *iso_src = in->iso;
    out->nlines = nl;
    out->niso = niso;
    out->form = form;
    for (int k = 0; k < HK_MAX_ISO; ++k) {
        out->iso_mass[k] = mass[k]; out->iso_ab[k] = ab[k]; out->iso_sscale[k] = ss[k];
    }
    out->max_abs_delta = mad;
    if (n == 0) return HK_OK;

    hipStream_t s = (hipStream_t)stream;
    char *w = (char *)hk::align_up((size_t)workspace, 256);
    int32_t *perm_in = (int32_t *)w;
    w += hk::align_up((size_t)n * sizeof(int32_t), 256);
    int32_t *perm = (int32_t *)w;
    w += hk::align_up((size_t)n * sizeof(int32_t), 256);
    size_t rb = radix_bytes(n);
    iota_kernel<<<grid_for(n), 256, 0, s>>>(perm_in, n);
    HK_LAUNCH_CHECK();
    HK_HIP(hipcub::DeviceRadixSort::SortPairs((void *)w, rb, src[0], (double *)dst[0], perm_in,
                                              perm, (int)n, 0, 64, s));
    Fields f;
    for (int k = 0; k < 6; ++k) {
        f.src[k] = src[k + 1];
        f.dst[k] = (double *)dst[k + 1];
    }
    f.iso_src = iso_src;
    f.iso_dst = iso_src ? (int32_t *)out->iso : nullptr;
    gather_kernel<<<grid_for(n), 256, 0, s>>>(perm, n, f);
    HK_LAUNCH_CHECK();
    return HK_OK;
}
