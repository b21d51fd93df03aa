// hk_profiles.hip -- the line profiles other than Voigt, and plinth removal.
//
// Profiles (Line6fA_kernel<profile>, voigt.h:1235-1262, with the S1 of
// S3_kernel, voigt.h:458-470), per (line, point) pair inside the cut, in
// Doppler units x = (x_j - nu) / alpha_D and with S the line state's strength
// (S(T) / alpha_D):
//   2 Lorentz          (S y / pi) / (x^2 + y^2)
//   3 Doppler          (S / sqrt(pi)) exp(-x^2)
//   4 binned Gaussian  S / vb * (erf(x + vb/2) - erf(x - vb/2)) / 2,
//                      vb = dnu / alpha_D: the Gaussian averaged over the
//                      point's grid cell (Hill et al. 2012); the difference
//                      is formed from erfc on the side away from the centre,
//                      which keeps it accurate in the wings.
// Plinth removal (Plinth_kernel, voigt.h:1091-1164; removePlinth): every
// pair also subtracts the line's profile value at its cut, x_c = hw /
// alpha_D, so a truncated line falls to zero at the cut.  For the Voigt
// profile the plain Voigt sum runs first (hk_voigt_accumulate, either form)
// and a second pass subtracts the plinth of every line over its window.
// Sub-Lorentzian wings (sLChi, voigt.h:47-72, applied at voigt.h:1345-1350 and
// in every profile branch): inside the given x range each pair's profile
// value is multiplied by chi(|x_j - nu|) x_j / nu (not the plinth); with the
// factor on, every profile -- Voigt included, through the FP64 Faddeeva
// function -- runs in this file's FP64 gather.
//
// Kernels: one workgroup of 256 threads per tile of 1024 points (4 per
// thread, coalesced); the tile's line range comes from the shared
// binary-search kernel; lines are staged 256 at a time in LDS with their
// cut window clipped to the tile; sums stay in registers in line order, so
// results are deterministic.  All arithmetic is FP64.
#include <math.h>

#include "hk_device.h"
#include "hk_internal.h"
#include "hk_tiles.h"

namespace {

using namespace hk;

constexpr int kPT = 256;             // threads per workgroup
constexpr int kPPT = 4;              // points per thread
constexpr int kPW = kPT * kPPT;      // points per tile

struct PRec {
    double nu, ia, S1, y, pl, vb;
    int32_t lo, hi;                  // tile-relative window [lo, hi)
};

// Binned Gaussian cell average of exp(-t^2)/sqrt(pi) times vb:
// (erf(x + h) - erf(x - h)) / 2 with h = vb/2.
__device__ __forceinline__ double binned_gauss(double x, double h) {
    return x >= 0.0 ? 0.5 * (erfc(x - h) - erfc(x + h)) : 0.5 * (erfc(-x - h) - erfc(-x + h));
}

// Profile value of one pair; S1 is the per-line prefactor formed at staging.
template <int PROFILE>
__device__ __forceinline__ double profile_value(double S1, double x, double y, double vb) {
    if (PROFILE == 1) return S1 * rew64(x, y);
    if (PROFILE == 2) return S1 / (x * x + y * y);
    if (PROFILE == 3) return S1 * exp(-x * x);
    return S1 / vb * binned_gauss(x, 0.5 * vb);
}

__device__ __forceinline__ double prefactor(int profile, double S, double y) {
    if (profile == 2) return S * y / kPi;
    if (profile == 3 || profile == 1) return S * kISqrtPi;
    return S;
}

struct Chi {
    double s1, s2, s3, B1, B2, B3, lo, hi;
};

// Perrin & Hartmann (1989) chi of the distance d = |x - nu| [cm^-1].
__device__ __forceinline__ double chi_of(const Chi &c, double d) {
    if (d < c.s1) return 1.0;
    if (d < c.s2) return exp(-c.B1 * (d - c.s1));
    if (d < c.s3) return exp(-c.B1 * (c.s2 - c.s1) - c.B2 * (d - c.s2));
    return exp(-c.B1 * (c.s2 - c.s1) - c.B2 * (c.s3 - c.s2) - c.B3 * (d - c.s3));
}

// Plinth of each line: its profile at the cut (0 for empty lines).
__global__ __launch_bounds__(256) void plinth_kernel(
    const double *__restrict__ S, const double *__restrict__ ia, const double *__restrict__ y,
    const double *__restrict__ hw, int64_t n, int profile, double *__restrict__ pl) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    const double h = hw[i];
    if (h > 0.0) {
        const double xc = h * ia[i], yy = y[i];
        if (profile == 1) v = S[i] * rew64(xc, yy) * kISqrtPi;
        else if (profile == 2) v = prefactor(2, S[i], yy) / (xc * xc + yy * yy);
        else v = prefactor(3, S[i], yy) * exp(-xc * xc);
    }
    pl[i] = v;
}

struct PArgs {
    const double *nu, *S, *ia, *y, *hw, *pl;
    int64_t nl;
    const TileInfo *tiles;
    double numin, dnu;
    int64_t nx;
    double kmin;
    int accumulate;
    double *K;
    Chi chi;
};

// PROFILE 1..4: profile sums (minus plinths when PLINTH); PROFILE 0: the
// plinth subtraction alone (after a Voigt sum), always accumulating.
template <int PROFILE, bool PLINTH, bool CHI>
__global__ __launch_bounds__(kPT) void profile_tiles_kernel(PArgs a) {
    __shared__ PRec rec[kPT];
    const int tid = threadIdx.x;
    const int64_t j0 = (int64_t)blockIdx.x * kPW;
    const TileInfo ti = a.tiles[blockIdx.x];
    double acc[kPPT];
#pragma unroll
    for (int k = 0; k < kPPT; ++k) acc[k] = 0.0;
    for (int64_t base = ti.lo; base < ti.hi; base += kPT) {
        const int64_t i = base + tid;
        PRec r;
        r.lo = 0; r.hi = 0;
        r.nu = 0.0; r.ia = 0.0; r.S1 = 0.0; r.y = 0.0; r.pl = 0.0; r.vb = 0.0;
        if (i < ti.hi) {
            int64_t lo, hi;
            cut_window(a.nu[i], a.hw[i], a.numin, a.dnu, a.nx, lo, hi);
            lo = max(lo, j0) - j0;
            hi = min(hi, j0 + kPW) - j0;
            if (lo < hi) {
                r.lo = (int32_t)lo; r.hi = (int32_t)hi;
                r.nu = a.nu[i]; r.ia = a.ia[i]; r.y = a.y[i];
                r.S1 = PROFILE ? prefactor(PROFILE, a.S[i], r.y) : 0.0;
                r.pl = PLINTH ? a.pl[i] : 0.0;
                r.vb = a.dnu * r.ia;
            }
        }
        rec[tid] = r;
        __syncthreads();
        const int nr = (int)min((int64_t)kPT, ti.hi - base);
        for (int q = 0; q < nr; ++q) {
            const PRec &L = rec[q];
            if (L.lo >= L.hi) continue;
#pragma unroll
            for (int k = 0; k < kPPT; ++k) {
                const int jj = tid + k * kPT;
                if (jj >= L.lo && jj < L.hi) {
                    double v = 0.0;
                    if (PROFILE) {
                        const double xj = fma((double)(j0 + jj), a.dnu, a.numin);
                        const double d = xj - L.nu;
                        v = profile_value<PROFILE>(L.S1, d * L.ia, L.y, L.vb);
                        if (CHI && xj > a.chi.lo && xj < a.chi.hi)
                            v *= chi_of(a.chi, fabs(d)) * xj / L.nu;
                    }
                    acc[k] += v - L.pl;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kPPT; ++k) {
        const int64_t j = j0 + tid + k * kPT;
        if (j < a.nx) a.K[j] = (a.accumulate || PROFILE == 0) ? a.K[j] + acc[k] : a.kmin + acc[k];
    }
}

template <int PROFILE, bool PLINTH, bool CHI = false>
void launch_profile(int64_t ntiles, hipStream_t s, const PArgs &a) {
    hipLaunchKernelGGL((profile_tiles_kernel<PROFILE, PLINTH, CHI>), dim3((unsigned)ntiles),
                       dim3(kPT), 0, s, a);
}

template <int PROFILE>
void launch_chi(bool plinth, int64_t ntiles, hipStream_t s, const PArgs &a) {
    if (plinth) launch_profile<PROFILE, true, true>(ntiles, s, a);
    else launch_profile<PROFILE, false, true>(ntiles, s, a);
}

size_t tiles_bytes(const hk_grid *g) {
    const int64_t nt = (g->nx + kPW - 1) / kPW;
    return align_up((size_t)nt * sizeof(TileInfo), 256);
}

size_t region_bytes(const hk_grid *g) {
    hk_tune t = {0, 0};
    const size_t v = hk_voigt_workspace_bytes(g, &t);
    const size_t p = tiles_bytes(g);
    return align_up(v > p ? v : p, 256);
}

}  // namespace

extern "C" size_t hk_profile_workspace_bytes(const hk_grid *grid, int64_t nlines) {
    if (!grid || grid->nx <= 0 || nlines < 0) return 0;
    return region_bytes(grid) + align_up((size_t)nlines * sizeof(double), 256) + 512;
}

extern "C" int hk_profile_accumulate(const hk_line_state *st, const hk_grid *grid, double kmin,
                                     int accumulate, int profile, int remove_plinth,
                                     int voigt_mode, const hk_chi *chi, double *K, void *ws,
                                     size_t ws_bytes, hk_stream_t stream) {
    HK_CHECK_ARG(st && grid && K && ws, "hk_profile_accumulate: null argument");
    HK_CHECK_ARG(profile >= HK_PROFILE_VOIGT && profile <= HK_PROFILE_BINNED_GAUSS,
                 "hk_profile_accumulate: profile %d (1 Voigt, 2 Lorentz, 3 Doppler, 4 binned "
                 "Gaussian)", profile);
    HK_CHECK_ARG(!(remove_plinth && profile == HK_PROFILE_BINNED_GAUSS),
                 "hk_profile_accumulate: plinth removal is not supported for profile 4");
    HK_CHECK_ARG(grid->nx > 0 && grid->nx < ((int64_t)1 << 31) && grid->dnu > 0.0,
                 "hk_profile_accumulate: grid of %lld points, dnu %g", (long long)grid->nx,
                 grid->dnu);
    HK_CHECK_ARG(st->nlines == 0 || (st->nu && st->S && st->ialphaD && st->y && st->hw &&
                                     st->radius && st->key),
                 "hk_profile_accumulate: line state not initialised by hk_line_params");
    const size_t need = hk_profile_workspace_bytes(grid, st->nlines);
    if (ws_bytes < need) {
        set_error("hk_profile_accumulate: workspace %zu bytes, %zu needed", ws_bytes, need);
        return HK_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    char *w = (char *)align_up((size_t)ws, 256);
    const size_t rb = region_bytes(grid);
    double *pl = (double *)(w + rb);
    if (chi) {
        HK_CHECK_ARG(chi->sigma[0] <= chi->sigma[1] && chi->sigma[1] <= chi->sigma[2],
                     "hk_profile_accumulate: chi breaks must be ascending");
    }
    if (profile == HK_PROFILE_VOIGT && !chi) {
        hk_tune t = {0, 0};
        int rc = hk_voigt_accumulate(st, grid, kmin, accumulate, voigt_mode, &t, K, w, rb, stream);
        if (rc != HK_OK || !remove_plinth) return rc;
    }
    const int64_t n = st->nlines;
    if (remove_plinth && n > 0) {
        hipLaunchKernelGGL(plinth_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           st->S, st->ialphaD, st->y, st->hw, n, profile, pl);
        HK_LAUNCH_CHECK();
    }
    const int64_t ntiles = (grid->nx + kPW - 1) / kPW;
    TileInfo *tiles = (TileInfo *)w;
    hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, s,
                       st->key, n, st->radius, st->blk_max, grid->numin, grid->dnu, grid->nx,
                       kPW, ntiles, 0, tiles);
    HK_LAUNCH_CHECK();
    PArgs a;
    a.nu = st->nu; a.S = st->S; a.ia = st->ialphaD; a.y = st->y; a.hw = st->hw; a.pl = pl;
    a.nl = n; a.tiles = tiles; a.numin = grid->numin; a.dnu = grid->dnu; a.nx = grid->nx;
    a.kmin = kmin; a.accumulate = accumulate; a.K = K;
    a.chi = Chi{0, 0, 0, 0, 0, 0, 0, 0};
    const bool P = remove_plinth != 0;
    if (chi) {
        a.chi = Chi{chi->sigma[0], chi->sigma[1], chi->sigma[2], chi->B[0], chi->B[1], chi->B[2],
                    chi->range[0], chi->range[1]};
        switch (profile) {
        case HK_PROFILE_VOIGT: launch_chi<1>(P, ntiles, s, a); break;
        case HK_PROFILE_LORENTZ: launch_chi<2>(P, ntiles, s, a); break;
        case HK_PROFILE_DOPPLER: launch_chi<3>(P, ntiles, s, a); break;
        default: launch_profile<4, false, true>(ntiles, s, a); break;
        }
        HK_LAUNCH_CHECK();
        return HK_OK;
    }
    switch (profile) {
    case HK_PROFILE_VOIGT: launch_profile<0, true>(ntiles, s, a); break;
    case HK_PROFILE_LORENTZ:
        P ? launch_profile<2, true>(ntiles, s, a) : launch_profile<2, false>(ntiles, s, a);
        break;
    case HK_PROFILE_DOPPLER:
        P ? launch_profile<3, true>(ntiles, s, a) : launch_profile<3, false>(ntiles, s, a);
        break;
    default: launch_profile<4, false>(ntiles, s, a); break;
    }
    HK_LAUNCH_CHECK();
    return HK_OK;
}
