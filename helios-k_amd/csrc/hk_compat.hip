// hk_compat.hip -- HK_VOIGT_COMPAT: the reference's own single-precision
// Voigt arithmetic, pair by pair, with FP64 accumulation.
//
// The reference evaluates every (line, point) pair in FP32 from per-line
// FP32 parameters that S3_kernel (voigt.h:436-482) derives from the FP64
// line state:
//   S1 = (float)(S y / pi), Sf = (float)S, yf = (float)y,
//   va = (float)((numin - nu) / alpha_D), vb = (float)(dnu / alpha_D),
//   vcut2 = (float)(cut^2 / alpha_D^2)            (cut mode 0),
// and, per pair, x = va + j vb (the grid index j converted to float, so
// beyond 2^24 points the positions round to even indices, as there),
// t1 = x^2, r2 = t1 + yf^2, kept when t1 < vcut2, with
//   region A, r2 >= 1e6 (Line6fA_kernel, voigt.h:1177-1281): S1 / r2;
//   region B, 100 <= r2 < 1e6 (Line6fB_kernel, voigt.h:1437-1515):
//     S1 t4, t4 = (pi t2 (2 t2 + r2) - 12 x^2) / (3 r2 (pi t2^2 - 6 x^2)),
//     t2 = r2 + 1.5;
//   region C, r2 < 100 (Line6fC_kernel, voigt.h:1616-1706, Sigmabf,
//     voigt.h:217-268): Algorithm 916 in FP32 with the per-term series
//     summed until all three sums change by less than def_TOLF, constants
//     a = pi / sqrt(-ln(TOLF/2)), b = 1/sqrt(pi), c = 2a/pi
//     (heliosk.cu:1656-1658), times Sf b.
// The values are added in FP64, as the reference's K_s/K_d do.  The
// arithmetic here is written out operation by operation with contraction
// switched off, so it rounds as written (the reference's compiler may fuse
// some of these products; that changes last bits only).  erfcxf(yf) is
// taken as the FP64 erfcx of yf rounded to float.  The cut: vcut2 =
// (float)(hw^2 / alpha_D^2) with hw the line state's half width, which is
// the reference's expression for cut mode 0; for modes 1 and 2 the
// reference squares cut y or cut instead (the same value up to the last
// bit).
//
// Layout: tiles of 1024 points, one workgroup each, lanes owning points
// 64k + lane of their wave's 256; lines staged 256 at a time through LDS in
// line order; each point sums its lines in line order.  Deterministic, no
// atomics.  The pass has no far field and no line-core split: it exists to
// reproduce the reference's numbers, not for speed.
#include <math.h>

#include "hk_device.h"
#include "hk_internal.h"
#include "hk_tiles.h"

namespace hk {
namespace {

constexpr int kCThreads = 256;
constexpr int kCPPT = 4;
constexpr int kCW = kCThreads * kCPPT;   // points per tile
constexpr float kTolF = 2.48e-12f;       // def_TOLF, define.h:48
constexpr float kPiF = 3.14159265358979323846f;

struct CRec {
    float S1, S, y, va, vb, vcut2, ey;
    int32_t lo, hi;                      // tile-relative point range to test
    int32_t pad[3];
};

struct CArgs {
    const double *nu, *S, *ia, *y, *hw;
    int64_t nl;
    const TileInfo *tiles;
    double numin, dnu;
    int64_t nx;
    double kmin;
    int accumulate;
    float a, b, c;
    double *K;
};

// Sigmabf (voigt.h:217-268): the three series of Algorithm 916 in FP32,
// summed term by term until every sum changes by less than def_TOLF.
__device__ void sigma_f32(float x, float y, float a, float &s1, float &s2, float &s3) {
#pragma clang fp contract(off)
    s1 = 0.0f; s2 = 0.0f; s3 = 0.0f;
    float o1 = 0.0f, o2 = 0.0f, o3 = 0.0f;
    const float yy = y * y;
    x = fabsf(x);
    const int n0 = (int)ceilf(x / a);
    bool st1 = false, st2 = false, st3 = false;
    for (int n = 1; n < 100; ++n) {
        const int n3p = n0 + n - 1, n3n = n0 - n;
        const float an = a * (float)n, an3p = a * (float)n3p, an3n = a * (float)n3n;
        const float f = 1.0f / (an * an + yy);
        const float f3p = 1.0f / (an3p * an3p + yy);
        const float f3n = 1.0f / (an3n * an3n + yy);
        s1 += f * expf(-(an * an + x * x));
        s2 += f * expf(-(an + x) * (an + x));
        s3 += f3p * expf(-(an3p - x) * (an3p - x));
        if (n3n >= 1) s3 += f3n * expf(-(an3n - x) * (an3n - x));
        if (fabsf(s1 - o1) < kTolF) st1 = true;
        if (fabsf(s2 - o2) < kTolF) st2 = true;
        if (fabsf(s3 - o3) < kTolF) st3 = true;
        if (st1 && st2 && st3) break;
        o1 = s1; o2 = s2; o3 = s3;
    }
}

// One pair in the reference's arithmetic; 0 outside the cut.
__device__ float compat_pair(const CRec &q, int64_t j, float a, float b, float c) {
#pragma clang fp contract(off)
    const float x = fmaf((float)j, q.vb, q.va);
    const float t1 = x * x;
    const float xxyy = t1 + q.y * q.y;
    if (!(t1 < q.vcut2)) return 0.0f;
    if (xxyy >= 1.0e6f) return q.S1 / xxyy;
    if (xxyy >= 100.0f) {
        const float t6 = t1 * 6.0f;
        const float t2 = xxyy + 1.5f;
        const float t3 = kPiF * t2;
        const float t4 = (t3 * (2.0f * t2 + xxyy) - 2.0f * t6) / (3.0f * xxyy * (t3 * t2 - t6));
        return q.S1 * t4;
    }
    const float y = q.y;
    float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
    const float ex2 = expf(-x * x);
    if (x != 0.0f && y != 0.0f) sigma_f32(x, y, a, s1, s2, s3);
    const float xy = x * y;
    const float cos2xy = cosf(2.0f * xy);
    const float sinxy = sinf(xy);
    float u1 = ex2 * q.ey * cos2xy;
    const float u2 = sinxy * ex2 * sinxy / y;
    const float u3 = y * (-cos2xy * s1 + 0.5f * (s2 + s3));
    u1 += c * (u2 + u3);
    if (x == 0.0f) u1 = q.ey;
    if (y == 0.0f) u1 = ex2;
    return q.S * u1 * b;
}

// Line range of each tile: the main pass's search widened by pad [cm^-1]
// (the FP32 cut test can admit a point just outside the FP64 window).
__global__ __launch_bounds__(256) void compat_ranges_kernel(const double *__restrict__ key,
                                                            int64_t nl,
                                                            const double *__restrict__ radius,
                                                            double pad_rel, double numin,
                                                            double dnu, int64_t nx,
                                                            int64_t ntiles,
                                                            TileInfo *__restrict__ tiles) {
    const int64_t tile = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (tile >= ntiles) return;
    const double R0 = radius[0];
    const double R = R0 + 2.0 * dnu + pad_rel * R0;
    const int64_t j0 = tile * kCW;
    const int64_t j1 = min(j0 + (int64_t)kCW, nx) - 1;
    int64_t lo = 0, hi = 0;
    if (R0 > 0.0 && nl > 0) {
        lo = lower_bound_dev(key, nl, fma((double)j0, dnu, numin) - R);
        hi = upper_bound_dev(key, nl, fma((double)j1, dnu, numin) + R);
    }
    TileInfo ti;
    ti.lo = lo; ti.hi = hi; ti.e = 0; ti.pad = 0; ti.pad2 = 0;
    tiles[tile] = ti;
}

constexpr double kPadRel = 1.0e-6;   // relative widening of each window

__global__ __launch_bounds__(kCThreads) void compat_tiles_kernel(CArgs a) {
    __shared__ CRec rec[kCThreads];
    __shared__ int cnt[kCThreads / 64];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int jr0 = wave * 64 * kCPPT + lane;
    const int64_t tile = blockIdx.x;
    const TileInfo ti = a.tiles[tile];
    const int64_t j0 = tile * kCW;
    double acc[kCPPT];
#pragma unroll
    for (int k = 0; k < kCPPT; ++k) acc[k] = 0.0;
    for (int64_t base = ti.lo; base < ti.hi; base += kCThreads) {
        const int64_t i = base + t;
        CRec r;
        bool valid = false;
        if (i < ti.hi) {
            const double nu = a.nu[i], hw = a.hw[i];
            if (hw > 0.0) {
                int64_t wl, wh;
                cut_window(nu, hw, a.numin, a.dnu, a.nx, wl, wh);
                // widen by the FP32 rounding of x (relative ~1e-7) plus 2 points
                const int64_t m = 2 + (int64_t)(kPadRel * hw / a.dnu);
                wl = max(wl - m, (int64_t)0);
                wh = min(wh + m, a.nx);
                const int64_t clo = max(wl, j0), chi = min(wh, j0 + kCW);
                if (clo < chi) {
                    const double S = a.S[i], ia = a.ia[i], y = a.y[i];
                    r.S1 = (float)(S * y / kPi);
                    r.S = (float)S;
                    r.y = (float)y;
                    r.va = (float)((a.numin - nu) * ia);
                    r.vb = (float)(a.dnu * ia);
                    r.vcut2 = (float)(hw * hw * ia * ia);
                    r.ey = (float)erfcx((double)r.y);
                    r.lo = (int32_t)(clo - j0);
                    r.hi = (int32_t)(chi - j0);
                    valid = true;
                }
            }
        }
        __syncthreads();   // the previous stage's readers of rec[] are done
        const unsigned long long bal = __ballot(valid);
        if (lane == 0) cnt[wave] = __popcll(bal);
        __syncthreads();
        int off = 0, n = 0;
#pragma unroll
        for (int w = 0; w < kCThreads / 64; ++w) {
            off += w < wave ? cnt[w] : 0;
            n += cnt[w];
        }
        if (valid) rec[off + __popcll(bal & ((1ull << lane) - 1ull))] = r;
        __syncthreads();
        for (int l = 0; l < n; ++l) {
            const CRec &q = rec[l];
#pragma unroll
            for (int k = 0; k < kCPPT; ++k) {
                const int jr = jr0 + 64 * k;
                if (jr >= q.lo && jr < q.hi)
                    acc[k] += (double)compat_pair(q, j0 + jr, a.a, a.b, a.c);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kCPPT; ++k) {
        const int64_t j = j0 + jr0 + 64 * k;
        if (j < a.nx) a.K[j] = a.accumulate ? a.K[j] + acc[k] : a.kmin + acc[k];
    }
}

}  // namespace

// Workspace: one tile table (fits in hk_voigt_workspace_bytes, which sizes
// tables of 256-point tiles).
int compat_accumulate(const hk_line_state *st, const hk_grid *grid, double kmin, int accumulate,
                      double *K, void *ws, hipStream_t s) {
    const int64_t ntiles = (grid->nx + kCW - 1) / kCW;
    TileInfo *tiles = (TileInfo *)align_up((size_t)ws, 256);
    hipLaunchKernelGGL(compat_ranges_kernel, dim3((unsigned)((ntiles + 255) / 256)), dim3(256), 0,
                       s, st->key, st->nlines, st->radius, kPadRel, grid->numin, grid->dnu,
                       grid->nx, ntiles, tiles);
    HK_LAUNCH_CHECK();
    CArgs a;
    a.nu = st->nu; a.S = st->S; a.ia = st->ialphaD; a.y = st->y; a.hw = st->hw;
    a.nl = st->nlines; a.tiles = tiles; a.numin = grid->numin; a.dnu = grid->dnu;
    a.nx = grid->nx; a.kmin = kmin; a.accumulate = accumulate; a.K = K;
    // heliosk.cu:1656-1658, with def_TOLF the float 2.48e-12f
    const double av = (double)(float)(kPi * sqrt(-1.0 / log((double)kTolF * 0.5)));
    a.a = (float)av;
    a.b = (float)(1.0 / sqrt(kPi));
    a.c = (float)(2.0 * av / kPi);
    hipLaunchKernelGGL(compat_tiles_kernel, dim3((unsigned)ntiles), dim3(kCThreads), 0, s, a);
    HK_LAUNCH_CHECK();
    return HK_OK;
}

}  // namespace hk
