// hk_tiles.h -- line ranges of grid tiles, shared by the kernels that sum
// lines onto the grid (Voigt passes, the other profiles, the plinth).
//
// Lines are sorted by their search key (the raw centre).  A tile of points
// [j0, j1] can only receive lines with key in [x_j0 - R, x_j1 + R], R the
// largest reach (cut half width plus centre shift) over the list; one
// wavefront per tile finds that range by binary search and, for the FP32
// form, a power-of-two scale from the per-block strength maxima.
#pragma once
#include "hk_device.h"

namespace hk {

constexpr int kScaleTop = 60;        // FP32 form: max strength -> 2^60

struct TileInfo {
    int64_t lo, hi;     // line range [lo, hi)
    int32_t e;          // FP32 form: strengths scaled by 2^-e
    int32_t pad;
    int64_t pad2;
};

__device__ __forceinline__ int64_t lower_bound_dev(const double *v, int64_t n, double key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t m = (lo + hi) >> 1;
        if (v[m] < key) lo = m + 1; else hi = m;
    }
    return lo;
}

__device__ __forceinline__ int64_t upper_bound_dev(const double *v, int64_t n, double key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t m = (lo + hi) >> 1;
        if (v[m] <= key) lo = m + 1; else hi = m;
    }
    return lo;
}

// Line range (and FP32 scale) of one tile, computed by one wavefront.
__device__ __forceinline__ void tile_range(const double *__restrict__ key, int64_t nl,
                                           const double *__restrict__ radius,
                                           const double *__restrict__ blk_max, double numin,
                                           double dnu, int64_t nx, int tile_pts, int64_t tile,
                                           int scale, TileInfo *__restrict__ tiles) {
    const int lane = threadIdx.x & 63;
    const double R = radius[0];
    const int64_t j0 = tile * tile_pts;
    const int64_t j1 = min(j0 + (int64_t)tile_pts, nx) - 1;
    const double xlo = fma((double)j0, dnu, numin) - R;
    const double xhi = fma((double)j1, dnu, numin) + R;
    int64_t lo = 0, hi = 0;
    if (R > 0.0 && nl > 0) {
        lo = lower_bound_dev(key, nl, xlo);
        hi = upper_bound_dev(key, nl, xhi);
    }
    double m = 0.0;
    if (scale && lo < hi) {
        for (int64_t b = lo / kBlockLines + lane; b <= (hi - 1) / kBlockLines; b += 64)
            m = fmax(m, blk_max[b]);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (lane == 0) {
        TileInfo ti;
        ti.lo = lo; ti.hi = hi;
        ti.e = (scale && m > 0.0) ? ilogb(m) - kScaleTop : 0;
        ti.pad = 0; ti.pad2 = 0;
        tiles[tile] = ti;
    }
}

namespace {
__global__ __launch_bounds__(256) __attribute__((unused)) void tile_ranges_kernel(
    const double *__restrict__ key, int64_t nl, const double *__restrict__ radius,
    const double *__restrict__ blk_max, double numin, double dnu, int64_t nx, int tile_pts,
    int64_t ntiles, int scale, TileInfo *__restrict__ tiles) {
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
    tile_range(key, nl, radius, blk_max, numin, dnu, nx, tile_pts, tile, scale, tiles);
}
}  // namespace

}  // namespace hk
