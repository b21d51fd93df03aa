// hk_voigt.hip -- gather-style Voigt summation onto the wavenumber grid.
//
// Replaces the reference's scatter design (Line6fA/B/C_kernel, voigt.h:1177,
// 1437, 1616; 32 partial spectra summed by AddKStreams_kernel, voigt.h:728)
// with an owner-computes gather:
//   * the grid is cut into tiles of W points, one workgroup per tile;
//   * tile_ranges_kernel finds, by binary search on the sorted line
//     centres, the lines whose cut window can reach the tile, and (FP32
//     form) a power-of-two strength scale for the tile;
//   * each workgroup stages 256 lines at a time through LDS; every thread
//     owns W/256 grid points and keeps their sums in registers, adding the
//     lines in index order -- no atomics, bit-identical reruns;
//   * a line contributes to point j iff |x_j - nu| < hw in FP64 (the
//     integer window is computed once per line and tile, cut_window).
//
// Forms:
//   HK_VOIGT_FP64  x = (x_j - nu)/alpha_D in FP64, FP64 profile (Algorithm
//                  916 core, 8- and 4-node Gauss-Hermite, 3-term asymptotic)
//                  and FP64 accumulation.
//   HK_VOIGT_FP32C the line offset is formed in FP64 once per line and tile
//                  (nearest grid point jc, remainder r), then
//                  x = (j - jc) vb + r in FP32 with vb split hi+lo; FP32
//                  profile (916 core, 4-node Gauss-Hermite, 2-term
//                  asymptotic); per-point FP32 partial sums over 32 lines
//                  folded into a two-sum (hi, lo) pair; strengths scaled by
//                  2^-e per tile so the largest maps to 2^60.
#include <math.h>

#include <algorithm>
#include <vector>

#include "hk_device.h"
#include "hk_internal.h"

namespace hk {

constexpr int kThreads = 256;
constexpr int kStage = 256;          // lines per LDS stage (one per thread)
constexpr int kScaleTop = 60;        // FP32 form: max strength -> 2^60

struct TileInfo {
    int64_t lo, hi;     // line range [lo, hi)
    int32_t e;          // FP32 form: strengths scaled by 2^-e
    int32_t pad;
    int64_t pad2;
};

struct VArgs {
    const double *nu, *S, *ia, *y, *hw;
    const TileInfo *tiles;
    double numin, dnu;
    int64_t nx;
    double kmin;
    int accumulate;
    double *K;
};

// LDS record of one line for one tile.  lo/hi: the line's cut window;
// [r1lo, r1hi), [r2lo, r2hi), [r3lo, r3hi): conservative tile-relative index
// intervals outside which x^2 + y^2 is >= R1, R2, R3 (nested; R3 bounds the
// core).  A wave whose points miss an interval takes the branch-free path of
// the next region out.
struct Rec64 {
    double nu, ia, y, y2, A, B, G, p2, p3, p4, ey, q1;
    int32_t lo, hi, r1lo, r1hi, r2lo, r2hi, r3lo, r3hi;
};

struct Rec32 {
    float r, vbh, vbl, y, y2, A, B, G, cy, ey, q1, pad2;
    int32_t d0, lo, hi, r1lo, r1hi, r2lo, r2hi, r3lo, r3hi, dop, pad0, pad1;
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

// One wavefront per tile.
__global__ __launch_bounds__(256) void tile_ranges_kernel(
    const double *__restrict__ key, int64_t nl, const double *__restrict__ radius,
    const double *__restrict__ blk_max, double numin, double dnu, int64_t nx, int tile_pts,
    int64_t ntiles, int scale, TileInfo *__restrict__ tiles) {
    const int lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
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

// Conservative tile-relative interval of the points with x^2 + y^2 < R, for
// x_j ~ (j - jc) vb + rr and tile start j0 = jc + d0 ... as index j - j0.
__device__ __forceinline__ void region_interval(double R, double y2, double rr, double vb,
                                                double d0, int W, int32_t &lo, int32_t &hi) {
    if (!(y2 < R)) { lo = W + 1; hi = -1; return; }
    const double X = sqrt(R - y2);
    const double mlo = floor((-X - rr) / vb) - 1.0 - d0;
    const double mhi = ceil((X - rr) / vb) + 2.0 - d0;
    lo = (int32_t)fmin(fmax(mlo, -1.0), (double)(W + 1));
    hi = (int32_t)fmin(fmax(mhi, -1.0), (double)(W + 1));
}

__device__ __forceinline__ bool misses(int32_t lo, int32_t hi, int wlo, int whi) {
    return hi <= wlo || lo >= whi;
}

__device__ __noinline__ void core_consts64(double y, double &ey, double &q1) {
    ey = erfcx(y);
    q1 = core_q1(y, kA64, kN64);
}
__device__ __noinline__ void core_consts32(float y, float &ey, float &q1) {
    ey = (float)erfcx((double)y);
    q1 = (float)core_q1((double)y, kA32, kN32);
}

// Paths of the per-(line, wave) dispatch.
enum { kPathFar = 0, kPathMid = 1, kPathMid2 = 2, kPathAny = 3 };

// ---------------------------------------------------------------------------
// FP64 form.  Wave w owns the contiguous points [w*64*PPT, (w+1)*64*PPT) of
// the tile; lane l holds points w*64*PPT + 64k + l.  FULL: the line's cut
// window covers the wave's points, so no per-point mask.
// ---------------------------------------------------------------------------
template <int PATH, bool FULL, int PPT>
__device__ __forceinline__ void eval64(const Rec64 &qr, const double *xj, double *acc, int jr0) {
    const double nu = qr.nu, ia = qr.ia, y2 = qr.y2, B = qr.B, G = qr.G;
    const double p2 = qr.p2, p3 = qr.p3, p4 = qr.p4;
    const int32_t lo = qr.lo, hi = qr.hi;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const double x = (xj[k] - nu) * ia;
        const double r2 = fma(x, x, y2);
        double v;
        if (PATH == kPathFar) {
            const double inv = rcp64(r2);
            v = B * inv * fma(fma(fma(fma(p4, inv, p3), inv, p2), inv, 1.5), inv, 1.0);
        } else if (PATH == kPathMid) {
            v = G * gh4_sum64(x * x, r2);
        } else if (PATH == kPathMid2) {
            v = G * gh8_sum64(x * x, r2);
        } else {
            if (r2 >= kRF64) {
                const double inv = rcp64(r2);
                v = B * inv * fma(fma(fma(fma(p4, inv, p3), inv, p2), inv, 1.5), inv, 1.0);
            } else if (r2 >= kRM64) {
                v = G * gh4_sum64(x * x, r2);
            } else if (r2 >= kRB64) {
                v = G * gh8_sum64(x * x, r2);
            } else {
                v = qr.A * rew_core64(x, qr.y, qr.ey, qr.q1);
            }
        }
        if (!FULL) {
            const int jr = jr0 + 64 * k;
            v = (jr >= lo && jr < hi) ? v : 0.0;
        }
        acc[k] += v;
    }
}

template <bool FULL, int PPT>
__device__ __forceinline__ void dispatch64(const Rec64 &q, const double *xj, double *acc, int jr0,
                                           int wlo, int whi) {
    if (misses(q.r1lo, q.r1hi, wlo, whi)) eval64<kPathFar, FULL, PPT>(q, xj, acc, jr0);
    else if (misses(q.r2lo, q.r2hi, wlo, whi)) eval64<kPathMid, FULL, PPT>(q, xj, acc, jr0);
    else if (misses(q.r3lo, q.r3hi, wlo, whi)) eval64<kPathMid2, FULL, PPT>(q, xj, acc, jr0);
    else eval64<kPathAny, FULL, PPT>(q, xj, acc, jr0);
}

template <int PPT>
__device__ __forceinline__ void tile_f64(const VArgs &a, int64_t tile, Rec64 *rec) {
    constexpr int W = kThreads * PPT;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wlo = wave * 64 * PPT, whi = wlo + 64 * PPT;
    const int jr0 = wlo + lane;
    const TileInfo ti = a.tiles[tile];
    const int64_t j0 = tile * W;
    double xj[PPT], acc[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        xj[k] = fma((double)(j0 + jr0 + 64 * k), a.dnu, a.numin);
        acc[k] = 0.0;
    }
    for (int64_t base = ti.lo; base < ti.hi; base += kStage) {
        const int64_t i = base + t;
        Rec64 r;
        r.lo = 0; r.hi = 0;
        if (i < ti.hi) {
            const double nu = a.nu[i];
            int64_t wl, wh;
            cut_window(nu, a.hw[i], a.numin, a.dnu, a.nx, wl, wh);
            const int64_t clo = max(wl, j0), chi = min(wh, j0 + W);
            if (clo < chi) {
                const double ia = a.ia[i], y = a.y[i], y2 = y * y;
                const double A = a.S[i] * kISqrtPi;
                r.nu = nu; r.ia = ia; r.y = y; r.y2 = y2;
                r.A = A; r.B = A * y * kISqrtPi; r.G = A * y * (2.0 / kPi);
                r.p2 = 3.75 - 2.0 * y2; r.p3 = -15.0 * y2; r.p4 = 12.0 * y2 * y2;
                r.lo = (int32_t)(clo - j0); r.hi = (int32_t)(chi - j0);
                const double jcd = rint((nu - a.numin) / a.dnu);
                const double rr = (fma(jcd, a.dnu, a.numin) - nu) * ia;
                const double vb = a.dnu * ia, d0 = (double)j0 - jcd;
                region_interval(kRF64, y2, rr, vb, d0, W, r.r1lo, r.r1hi);
                region_interval(kRM64, y2, rr, vb, d0, W, r.r2lo, r.r2hi);
                region_interval(kRB64, y2, rr, vb, d0, W, r.r3lo, r.r3hi);
                if (r.r3lo < r.r3hi && r.r3hi > 0 && r.r3lo < W) {
                    core_consts64(y, r.ey, r.q1);
                } else {
                    r.ey = 0.0; r.q1 = 0.0;
                }
            }
        }
        __syncthreads();
        rec[t] = r;
        __syncthreads();
        const int nst = (int)min((int64_t)kStage, ti.hi - base);
        for (int l = 0; l < nst; ++l) {
            const Rec64 &q = rec[l];
            const int32_t qlo = q.lo, qhi = q.hi;
            if (qhi <= wlo || qlo >= whi) continue;          // window misses this wave
            if (qlo <= wlo && qhi >= whi) dispatch64<true, PPT>(q, xj, acc, jr0, wlo, whi);
            else dispatch64<false, PPT>(q, xj, acc, jr0, wlo, whi);
        }
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int64_t j = j0 + jr0 + 64 * k;
        if (j < a.nx) a.K[j] = a.accumulate ? a.K[j] + acc[k] : a.kmin + acc[k];
    }
}

// ---------------------------------------------------------------------------
// FP32 compensated form (same point ownership as the FP64 form)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void two_sum_acc(float &hi, float &lo, float p) {
    const float s = hi + p;
    const float bp = s - hi;
    const float ap = s - bp;
    lo += (hi - ap) + (p - bp);
    hi = s;
}

constexpr int kDT = 16;      // per-line core table stride (entries 1..kNT32 used)

template <int PATH, bool FULL, int PPT>
__device__ __forceinline__ void eval32(const Rec32 &qr, const float *dtab, float *part, int jr0) {
    const float r0 = qr.r, vbh = qr.vbh, vbl = qr.vbl, y2 = qr.y2, B = qr.B, G = qr.G, cy = qr.cy;
    const int32_t lo = qr.lo, hi = qr.hi;
    const float m0 = (float)(jr0 + qr.d0);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const float m = m0 + (float)(64 * k);
        const float x = fmaf(m, vbh, fmaf(m, vbl, r0));
        const float r2 = fmaf(x, x, y2);
        float v;
        if (PATH == kPathFar) {
            const float inv = rcp32(r2);
            v = (B * inv) * fmaf(fmaf(cy, inv, 1.5f), inv, 1.0f);
        } else if (PATH == kPathMid) {
            v = G * gh4_sum32(x * x, r2);
        } else {
            if (r2 >= kRF32) {
                const float inv = rcp32(r2);
                v = (B * inv) * fmaf(fmaf(cy, inv, 1.5f), inv, 1.0f);
            } else if (r2 >= kRM32) {
                v = G * gh4_sum32(x * x, r2);
            } else if (PATH == kPathMid2 || r2 >= kRB32) {
                v = G * gh8_sum32(x * x, r2);
                if (qr.dop) v = fmaf(qr.A, __expf(-x * x), v);
            } else {
                v = qr.A * rew_core32(x, qr.y, qr.ey, qr.q1, dtab);
            }
        }
        if (!FULL) {
            const int jr = jr0 + 64 * k;
            v = (jr >= lo && jr < hi) ? v : 0.0f;
        }
        part[k] += v;
    }
}

template <bool FULL, int PPT>
__device__ __forceinline__ void dispatch32(const Rec32 &q, const float *dtab, float *part, int jr0,
                                           int wlo, int whi) {
    if (misses(q.r1lo, q.r1hi, wlo, whi)) {
        eval32<kPathFar, FULL, PPT>(q, dtab, part, jr0);
    } else if (misses(q.r2lo, q.r2hi, wlo, whi)) {
        // 4-node rule for the whole wave unless x is so large that its
        // single-reciprocal form could overflow FP32
        const float m0 = (float)(wlo + q.d0), m1 = (float)(whi - 1 + q.d0);
        const float x0 = fmaf(m0, q.vbh, q.r), x1 = fmaf(m1, q.vbh, q.r);
        if (fmaxf(x0 * x0, x1 * x1) < 1.0e9f) eval32<kPathMid, FULL, PPT>(q, dtab, part, jr0);
        else eval32<kPathMid2, FULL, PPT>(q, dtab, part, jr0);
    } else if (misses(q.r3lo, q.r3hi, wlo, whi)) {
        eval32<kPathMid2, FULL, PPT>(q, dtab, part, jr0);
    } else {
        eval32<kPathAny, FULL, PPT>(q, dtab, part, jr0);
    }
}

template <int PPT>
__device__ __forceinline__ void tile_f32c(const VArgs &a, int64_t tile, Rec32 *rec, float *dtab) {
    constexpr int W = kThreads * PPT;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wlo = wave * 64 * PPT, whi = wlo + 64 * PPT;
    const int jr0 = wlo + lane;
    const TileInfo ti = a.tiles[tile];
    const int64_t j0 = tile * W;
    float hi[PPT], lo[PPT], part[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) { hi[k] = 0.0f; lo[k] = 0.0f; part[k] = 0.0f; }
    int since = 0;
    for (int64_t base = ti.lo; base < ti.hi; base += kStage) {
        const int64_t i = base + t;
        Rec32 r;
        r.lo = 0; r.hi = 0;
        bool core = false;
        if (i < ti.hi) {
            const double nu = a.nu[i];
            int64_t wl, wh;
            cut_window(nu, a.hw[i], a.numin, a.dnu, a.nx, wl, wh);
            const int64_t clo = max(wl, j0), chi = min(wh, j0 + W);
            if (clo < chi) {
                const double ia = a.ia[i], y = a.y[i];
                // nearest grid point to the centre (may lie off the grid)
                const double jcd = rint((nu - a.numin) / a.dnu);
                const double rr = (fma(jcd, a.dnu, a.numin) - nu) * ia;
                const double vb = a.dnu * ia, d0 = (double)j0 - jcd;
                const float vbh = (float)vb;
                const double A = ldexp(a.S[i], -ti.e) * kISqrtPi;
                const float yf = (float)y;
                r.r = (float)rr;
                r.vbh = vbh;
                r.vbl = (float)(vb - (double)vbh);
                r.y = yf;
                r.y2 = yf * yf;
                r.A = (float)A;
                r.B = (float)(A * y * kISqrtPi);
                r.G = (float)(A * y * (2.0 / kPi));
                r.cy = -2.0f * yf * yf;
                r.d0 = (int32_t)d0;
                r.lo = (int32_t)(clo - j0);
                r.hi = (int32_t)(chi - j0);
                r.dop = yf < kY0 ? 1 : 0;
                const double y2 = (double)yf * (double)yf;
                region_interval((double)kRF32, y2, rr, vb, d0, W, r.r1lo, r.r1hi);
                region_interval((double)kRM32, y2, rr, vb, d0, W, r.r2lo, r.r2hi);
                region_interval((double)kRB32, y2, rr, vb, d0, W, r.r3lo, r.r3hi);
                core = r.r3lo < r.r3hi && r.r3hi > 0 && r.r3lo < W;
                if (core) core_consts32(yf, r.ey, r.q1);
                else { r.ey = 0.0f; r.q1 = 0.0f; }
            }
        }
        // the previous stage's readers of rec[] and dtab[] are done after this
        __syncthreads();
        rec[t] = r;
        if (core) {
            const double y2 = (double)r.y2;
            for (int n = 1; n <= kNT32; ++n) {
                const double an = kA32 * n;
                dtab[t * kDT + n] = (float)(1.0 / fma(an, an, y2));
            }
        }
        __syncthreads();
        const int nst = (int)min((int64_t)kStage, ti.hi - base);
        for (int l = 0; l < nst; ++l) {
            const Rec32 &q = rec[l];
            const int32_t qlo = q.lo, qhi = q.hi;
            if (qlo >= qhi) continue;
            if (!(qhi <= wlo || qlo >= whi)) {
                const float *dt = dtab + l * kDT;
                if (qlo <= wlo && qhi >= whi) dispatch32<true, PPT>(q, dt, part, jr0, wlo, whi);
                else dispatch32<false, PPT>(q, dt, part, jr0, wlo, whi);
            }
            if (++since == 32) {
#pragma unroll
                for (int k = 0; k < PPT; ++k) { two_sum_acc(hi[k], lo[k], part[k]); part[k] = 0.0f; }
                since = 0;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        two_sum_acc(hi[k], lo[k], part[k]);
        const int64_t j = j0 + jr0 + 64 * k;
        if (j < a.nx) {
            const double sum = ldexp((double)hi[k] + (double)lo[k], ti.e);
            a.K[j] = a.accumulate ? a.K[j] + sum : a.kmin + sum;
        }
    }
}

template <int MODE, int PPT>
__global__ __launch_bounds__(256) void voigt_tiles_kernel(VArgs a) {
    if (MODE == HK_VOIGT_FP64) {
        __shared__ Rec64 rec[kStage];
        tile_f64<PPT>(a, blockIdx.x, rec);
    } else {
        __shared__ Rec32 rec[kStage];
        __shared__ float dtab[kStage * kDT];
        tile_f32c<PPT>(a, blockIdx.x, rec, dtab);
    }
}

// Persistent form: workgroups take (spectrum, tile) items from a queue.
// item_start[s] = first item of spectrum s (nspec+1 entries).
template <int MODE, int PPT>
__global__ __launch_bounds__(256) void voigt_persistent_kernel(const VArgs *__restrict__ specs,
                                                               const int64_t *__restrict__ item_start,
                                                               int nspec, unsigned long long *counter) {
    __shared__ int64_t s_item;
    __shared__ float dtab[MODE == HK_VOIGT_FP64 ? 1 : kStage * kDT];
    __shared__ __attribute__((aligned(16))) char smem[(MODE == HK_VOIGT_FP64 ? sizeof(Rec64) : sizeof(Rec32)) * kStage];
    const int64_t total = item_start[nspec];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t item = s_item;
        if (item >= total) break;
        int s = 0;
        while (s + 1 < nspec && item_start[s + 1] <= item) ++s;
        const VArgs a = specs[s];
        const int64_t tile = item - item_start[s];
        if (MODE == HK_VOIGT_FP64) tile_f64<PPT>(a, tile, (Rec64 *)smem);
        else tile_f32c<PPT>(a, tile, (Rec32 *)smem, dtab);
    }
}

__global__ void reset_counter_kernel(unsigned long long *c) { *c = 0ull; }

// Algorithmic evaluation count: per line, the points of its cut window, split
// by the reference's regions (voigt.h:1223-1230, 1490, 1665): A r2 >= 1e6,
// B 100 <= r2 < 1e6, C r2 < 100 with r2 = x^2 + y^2.
__global__ __launch_bounds__(256) void pair_count_kernel(const double *__restrict__ nu,
                                                         const double *__restrict__ hw,
                                                         const double *__restrict__ ia,
                                                         const double *__restrict__ y, int64_t nl,
                                                         double numin, double dnu, int64_t nx,
                                                         unsigned long long *__restrict__ out) {
    __shared__ unsigned long long red[3][256];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * 256 + t;
    unsigned long long tot = 0, bc = 0, c = 0;
    if (i < nl) {
        int64_t lo, hi;
        cut_window(nu[i], hw[i], numin, dnu, nx, lo, hi);
        tot = (unsigned long long)(hi - lo);
        const double y2 = y[i] * y[i], a = ia[i];
        if (tot && a > 0.0) {
            if (y2 < 1.0e6) {
                cut_window(nu[i], fmin(hw[i], sqrt(1.0e6 - y2) / a), numin, dnu, nx, lo, hi);
                bc = (unsigned long long)(hi - lo);
            }
            if (y2 < 100.0) {
                cut_window(nu[i], fmin(hw[i], sqrt(100.0 - y2) / a), numin, dnu, nx, lo, hi);
                c = (unsigned long long)(hi - lo);
            }
        }
    }
    red[0][t] = tot; red[1][t] = bc; red[2][t] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) { red[0][t] += red[0][t + s]; red[1][t] += red[1][t + s]; red[2][t] += red[2][t + s]; }
        __syncthreads();
    }
    if (t == 0) {
        atomicAdd(&out[0], red[0][0]);
        atomicAdd(&out[1], red[0][0] - red[1][0]);
        atomicAdd(&out[2], red[1][0] - red[2][0]);
        atomicAdd(&out[3], red[2][0]);
    }
}

__global__ void zero4_kernel(unsigned long long *c) {
    if (threadIdx.x < 4) c[threadIdx.x] = 0ull;
}

}  // namespace hk

using namespace hk;

namespace {

int pick_ppt(int mode, const hk_tune *tune) {
    int tp = tune && tune->tile_points > 0 ? tune->tile_points : (mode == HK_VOIGT_FP64 ? 1024 : 2048);
    return tp / kThreads;
}

bool ppt_ok(int ppt) { return ppt == 2 || ppt == 4 || ppt == 8 || ppt == 16; }

template <int MODE>
void launch_tiles(int ppt, int64_t ntiles, hipStream_t s, const VArgs &a) {
    dim3 g((unsigned)ntiles), b(kThreads);
    switch (ppt) {
        case 2: hipLaunchKernelGGL((voigt_tiles_kernel<MODE, 2>), g, b, 0, s, a); break;
        case 4: hipLaunchKernelGGL((voigt_tiles_kernel<MODE, 4>), g, b, 0, s, a); break;
        case 8: hipLaunchKernelGGL((voigt_tiles_kernel<MODE, 8>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((voigt_tiles_kernel<MODE, 16>), g, b, 0, s, a); break;
    }
}

template <int MODE>
void launch_persistent(int ppt, unsigned grid, hipStream_t s, const VArgs *specs,
                       const int64_t *item_start, int nspec, unsigned long long *counter) {
    dim3 g(grid), b(kThreads);
    switch (ppt) {
        case 2: hipLaunchKernelGGL((voigt_persistent_kernel<MODE, 2>), g, b, 0, s, specs, item_start, nspec, counter); break;
        case 4: hipLaunchKernelGGL((voigt_persistent_kernel<MODE, 4>), g, b, 0, s, specs, item_start, nspec, counter); break;
        case 8: hipLaunchKernelGGL((voigt_persistent_kernel<MODE, 8>), g, b, 0, s, specs, item_start, nspec, counter); break;
        default: hipLaunchKernelGGL((voigt_persistent_kernel<MODE, 16>), g, b, 0, s, specs, item_start, nspec, counter); break;
    }
}

int check_state(const hk_line_state *st, const char *fn) {
    HK_CHECK_ARG(st->nlines == 0 || (st->nu && st->S && st->ialphaD && st->y && st->hw &&
                                     st->blk_max && st->radius && st->key),
                 "%s: line state not initialised by hk_line_params", fn);
    return HK_OK;
}

int check_grid(const hk_grid *g, const char *fn) {
    HK_CHECK_ARG(g->nx > 0 && g->nx < ((int64_t)1 << 31), "%s: nx %lld out of range", fn,
                 (long long)g->nx);
    HK_CHECK_ARG(g->dnu > 0.0, "%s: dnu must be > 0", fn);
    return HK_OK;
}

}  // namespace

extern "C" size_t hk_voigt_workspace_bytes(const hk_grid *grid, const hk_tune *tune) {
    if (!grid || grid->nx <= 0) return 0;
    int64_t w = 256;   // smallest tile
    (void)tune;
    int64_t ntiles = (grid->nx + w - 1) / w;
    return (size_t)ntiles * sizeof(TileInfo) + 256;
}

extern "C" int hk_voigt_accumulate(const hk_line_state *st, const hk_grid *grid, double kmin,
                                   int accumulate, int mode, const hk_tune *tune, double *K,
                                   void *ws, size_t ws_bytes, hk_stream_t stream) {
    HK_CHECK_ARG(st && grid && K && ws, "hk_voigt_accumulate: null argument");
    HK_CHECK_ARG(mode == HK_VOIGT_FP64 || mode == HK_VOIGT_FP32C, "hk_voigt_accumulate: mode %d",
                 mode);
    if (int e = check_state(st, "hk_voigt_accumulate")) return e;
    if (int e = check_grid(grid, "hk_voigt_accumulate")) return e;
    const int ppt = pick_ppt(mode, tune);
    HK_CHECK_ARG(ppt_ok(ppt), "hk_voigt_accumulate: tile_points must be 512, 1024, 2048 or 4096");
    if (ws_bytes < hk_voigt_workspace_bytes(grid, tune)) {
        set_error("hk_voigt_accumulate: workspace %zu bytes, %zu needed", ws_bytes,
                  hk_voigt_workspace_bytes(grid, tune));
        return HK_ERR_WORKSPACE;
    }
    const int W = kThreads * ppt;
    const int64_t ntiles = (grid->nx + W - 1) / W;
    TileInfo *tiles = (TileInfo *)align_up((size_t)ws, 256);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, s,
                       st->key, st->nlines, st->radius, st->blk_max, grid->numin, grid->dnu,
                       grid->nx, W, ntiles, mode == HK_VOIGT_FP32C ? 1 : 0, tiles);
    HK_LAUNCH_CHECK();
    VArgs a;
    a.nu = st->nu; a.S = st->S; a.ia = st->ialphaD; a.y = st->y; a.hw = st->hw;
    a.tiles = tiles; a.numin = grid->numin; a.dnu = grid->dnu; a.nx = grid->nx;
    a.kmin = kmin; a.accumulate = accumulate; a.K = K;
    if (mode == HK_VOIGT_FP64) launch_tiles<HK_VOIGT_FP64>(ppt, ntiles, s, a);
    else launch_tiles<HK_VOIGT_FP32C>(ppt, ntiles, s, a);
    HK_LAUNCH_CHECK();
    return HK_OK;
}

// ---------------------------------------------------------------------------
// Batched: workspace = [VArgs x nspec | item_start x (nspec+1) | counter |
//                       tile tables of every spectrum]
// ---------------------------------------------------------------------------
namespace {
struct BatchLayout {
    size_t off_args, off_items, off_counter, off_tiles, total;
    std::vector<size_t> tile_off;
    std::vector<int64_t> ntiles;
};

BatchLayout batch_layout(int nspec, const hk_spectrum *spec, int W) {
    BatchLayout L;
    size_t o = 0;
    L.off_args = o; o = align_up(o + sizeof(VArgs) * nspec, 256);
    L.off_items = o; o = align_up(o + sizeof(int64_t) * (nspec + 1), 256);
    L.off_counter = o; o = align_up(o + 64, 256);
    L.off_tiles = o;
    for (int i = 0; i < nspec; ++i) {
        int64_t nt = (spec[i].grid.nx + W - 1) / W;
        L.ntiles.push_back(nt);
        L.tile_off.push_back(o);
        o = align_up(o + sizeof(TileInfo) * nt, 256);
    }
    L.total = o + 256;
    return L;
}
}  // namespace

extern "C" size_t hk_voigt_batched_workspace_bytes(int nspec, const hk_spectrum *spec,
                                                   const hk_tune *tune) {
    if (nspec <= 0 || !spec) return 0;
    (void)tune;
    return batch_layout(nspec, spec, 256).total;   // smallest tile: largest tables
}

extern "C" int hk_voigt_batched(int nspec, const hk_spectrum *spec, int mode, const hk_tune *tune,
                                void *ws, size_t ws_bytes, hk_stream_t stream) {
    HK_CHECK_ARG(nspec > 0 && spec && ws, "hk_voigt_batched: bad arguments");
    HK_CHECK_ARG(mode == HK_VOIGT_FP64 || mode == HK_VOIGT_FP32C, "hk_voigt_batched: mode %d", mode);
    const int ppt = pick_ppt(mode, tune);
    HK_CHECK_ARG(ppt_ok(ppt), "hk_voigt_batched: tile_points must be 512, 1024, 2048 or 4096");
    for (int i = 0; i < nspec; ++i) {
        if (int e = check_state(&spec[i].st, "hk_voigt_batched")) return e;
        if (int e = check_grid(&spec[i].grid, "hk_voigt_batched")) return e;
        HK_CHECK_ARG(spec[i].K, "hk_voigt_batched: spectrum %d has no K", i);
    }
    if (ws_bytes < hk_voigt_batched_workspace_bytes(nspec, spec, tune)) {
        set_error("hk_voigt_batched: workspace %zu bytes, %zu needed", ws_bytes,
                  hk_voigt_batched_workspace_bytes(nspec, spec, tune));
        return HK_ERR_WORKSPACE;
    }
    const int W = kThreads * ppt;
    char *base = (char *)align_up((size_t)ws, 256);
    BatchLayout L = batch_layout(nspec, spec, W);
    // Host tables: spectra ordered by line count (largest first) so the
    // heaviest tiles are dequeued first.
    std::vector<int> order(nspec);
    for (int i = 0; i < nspec; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return spec[x].st.nlines > spec[y].st.nlines;
    });
    std::vector<VArgs> args(nspec);
    std::vector<int64_t> items(nspec + 1, 0);
    hipStream_t s = (hipStream_t)stream;
    for (int q = 0; q < nspec; ++q) {
        const int i = order[q];
        const hk_spectrum &sp = spec[i];
        TileInfo *tiles = (TileInfo *)(base + L.tile_off[i]);
        hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((L.ntiles[i] + 3) / 4)), dim3(256),
                           0, s, sp.st.key, sp.st.nlines, sp.st.radius, sp.st.blk_max,
                           sp.grid.numin, sp.grid.dnu, sp.grid.nx, W, L.ntiles[i],
                           mode == HK_VOIGT_FP32C ? 1 : 0, tiles);
        HK_LAUNCH_CHECK();
        VArgs &a = args[q];
        a.nu = sp.st.nu; a.S = sp.st.S; a.ia = sp.st.ialphaD; a.y = sp.st.y; a.hw = sp.st.hw;
        a.tiles = tiles; a.numin = sp.grid.numin; a.dnu = sp.grid.dnu; a.nx = sp.grid.nx;
        a.kmin = sp.kmin; a.accumulate = sp.accumulate; a.K = sp.K;
        items[q + 1] = items[q] + L.ntiles[i];
    }
    HK_HIP(hipMemcpyAsync(base + L.off_args, args.data(), sizeof(VArgs) * nspec,
                          hipMemcpyHostToDevice, s));
    HK_HIP(hipMemcpyAsync(base + L.off_items, items.data(), sizeof(int64_t) * (nspec + 1),
                          hipMemcpyHostToDevice, s));
    unsigned long long *counter = (unsigned long long *)(base + L.off_counter);
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, counter);
    HK_LAUNCH_CHECK();
    int dev = 0, ncu = 0;
    HK_HIP(hipGetDevice(&dev));
    HK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t total = items[nspec];
    unsigned grid = (unsigned)std::min<int64_t>(total, (int64_t)ncu * 8);
    if (grid == 0) return HK_OK;
    if (mode == HK_VOIGT_FP64)
        launch_persistent<HK_VOIGT_FP64>(ppt, grid, s, (const VArgs *)(base + L.off_args),
                                         (const int64_t *)(base + L.off_items), nspec, counter);
    else
        launch_persistent<HK_VOIGT_FP32C>(ppt, grid, s, (const VArgs *)(base + L.off_args),
                                          (const int64_t *)(base + L.off_items), nspec, counter);
    HK_LAUNCH_CHECK();
    // the host tables must outlive the asynchronous copies
    HK_HIP(hipStreamSynchronize(s));
    return HK_OK;
}

extern "C" int hk_voigt_pair_count(const hk_line_state *st, const hk_grid *grid,
                                   unsigned long long *counts, hk_stream_t stream) {
    HK_CHECK_ARG(st && grid && counts, "hk_voigt_pair_count: null argument");
    if (int e = check_state(st, "hk_voigt_pair_count")) return e;
    if (int e = check_grid(grid, "hk_voigt_pair_count")) return e;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(zero4_kernel, dim3(1), dim3(64), 0, s, counts);
    HK_LAUNCH_CHECK();
    if (st->nlines > 0) {
        hipLaunchKernelGGL(pair_count_kernel, dim3((unsigned)((st->nlines + 255) / 256)), dim3(256),
                           0, s, st->nu, st->hw, st->ialphaD, st->y, st->nlines, grid->numin,
                           grid->dnu, grid->nx, counts);
        HK_LAUNCH_CHECK();
    }
    return HK_OK;
}
