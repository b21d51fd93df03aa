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
//     owns W/256 grid points and keeps their sums in registers (the FP64
//     form adds the lines in index order, the FP32 form path by path within
//     a stage) -- no atomics, bit-identical reruns;
//   * a line contributes to point j iff |x_j - nu| < hw in FP64 (the
//     integer window is computed once per line and tile, cut_window).
//
// Two passes per spectrum, in stream order:
//   * the main pass above evaluates every (line, point) pair of the cut
//     windows except the line cores, x^2 + y^2 < R_core, which it skips;
//   * the line-core pass (voigt_core_kernel) adds the core pairs: tiles of
//     256 points, one wave per 64 points, over only the lines whose core can
//     reach the tile (a second binary search with the core reach), compacted
//     in LDS in line order.  Both passes form x and r2 with the same
//     operations, so each pair is evaluated by exactly one of them.
//   Keeping the core series out of the main kernel keeps its register
//   footprint, and so its occupancy, set by the wing formulas.
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
//                  Far lines (see the far-field section) are summed in FP32
//                  at the Chebyshev nodes of the tile (9, 15 or 19 by
//                  distance), or of a wave's quarter of it, and interpolated
//                  to the points in FP64.
#include <math.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "hk_device.h"
#include "hk_internal.h"
#include "hk_tiles.h"

namespace hk {

constexpr int kThreads = 256;
constexpr int kStage = 256;          // lines per LDS stage (one per thread)

struct VArgs {
    const double *nu, *S, *ia, *y, *hw, *ey;
    const double *key, *blk_max, *radius;   // line-range search (batched form)
    int64_t nl;
    TileInfo *tiles;
    double numin, dnu;
    int64_t nx;
    double kmin;
    int accumulate;
    int far;                                // FP32 form: far-field interpolation on
    double *K;
};

// LDS record of one line for one tile.  lo/hi: the line's cut window
// (tile-relative); code: 4 bits per wave, the evaluation path that wave takes
// for this line (wave_code), decided once when the record is built.
struct Rec64 {
    double nu, ia, y2, B, G, p2, p3, p4;
    int32_t lo, hi, code, pad;
};

// FP32 record, grouped by the path that reads it: x and r2; the asymptotic
// series of the wings with the strength folded into the coefficients
// (far: B, 1.5 B, BC; mid: q2..q6, see far4_32); the 8-node rule and the
// Doppler term; window and dispatch word.  72 bytes: with the far field's
// sums, a 256-line stage fits six workgroups in one CU's LDS.
struct Rec32 {
    float r, vbh, vbl, y2;
    float B, BC;
    int32_t d0;
    float q2, q3, q4, q5;
    float q6, G, A;
    int32_t dop;
    int32_t code, lo, hi;
};

// Line-core pass records: lo/hi bound the points of the tile that lie both
// in the line's cut window and in its (conservative) core interval.
struct Core64 {
    double nu, ia, y, y2, A, ey, q1;
    int32_t lo, hi;
};

struct Core32 {
    float r, vbh, vbl, y, y2, A, ey, q1;
    int32_t lo, hi, d0, pad;
};

// Main-pass record loops: lane l of a wave reads the dispatch nibble of
// record c0 + l, one ballot marks the records this wave must take, and the
// wave walks the set bits in ascending order with scalar instructions, so
// records that miss the wave cost nothing and no record waits on an LDS read
// of its own code.  (The line-core passes keep a plain loop: the ballot form
// costs them a register block and a wave of occupancy.)
__device__ __forceinline__ int lowest_bit(unsigned long long &m) {
    const int b = __builtin_ctzll(m);
    m &= m - 1ull;
    return b;
}

// Spectrum of a queue item: item_start[s] <= item < item_start[s + 1].
__device__ __forceinline__ int item_spectrum(const int64_t *__restrict__ item_start, int nspec,
                                             int64_t item) {
    int lo = 0, hi = nspec - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (item_start[m] <= item) lo = m; else hi = m - 1;
    }
    return lo;
}

// Batched form: tile ranges of every (spectrum, tile) item, one wave each.
__global__ __launch_bounds__(256) void batched_ranges_kernel(const VArgs *__restrict__ specs,
                                                             const int64_t *__restrict__ item_start,
                                                             int nspec, int tile_pts, int scale) {
    const int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= item_start[nspec]) return;
    const int s = item_spectrum(item_start, nspec, item);
    const VArgs &a = specs[s];
    tile_range(a.key, a.nl, a.radius, a.blk_max, a.numin, a.dnu, a.nx, tile_pts,
               item - item_start[s], scale, a.tiles);
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

// Per-wave path code of one line: 0 = the window misses the wave's points;
// otherwise 1 + 2 * path + (window covers the wave ? 0 : 1).  [r1lo, r1hi),
// [r2lo, r2hi), [r3lo, r3hi) are conservative tile-relative intervals
// outside which x^2 + y^2 >= R1 > R2 > R3; a wave that misses an interval
// takes the branch-free path of the next region out.  The whole-wave
// Gauss-Hermite paths form products of order (x^2+y^2)^3 (4 nodes, with the
// strength folded in) and (x^2+y^2)^7 (8 nodes); r2max(w) bounds x^2 + y^2
// over wave w, and a wave beyond mid_max / mid2_max takes the per-point path,
// which uses those rules only below the far-wing threshold.
enum { kPathFar = 0, kPathMid = 1, kPathMid2 = 2, kPathAny = 3 };

template <int WPTS, class R2Max>
__device__ __forceinline__ int32_t wave_codes(int32_t lo, int32_t hi, const int32_t (&ri)[6],
                                              R2Max r2max, double mid_max, double mid2_max) {
    int32_t code = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        const int wlo = w * WPTS, whi = wlo + WPTS;
        if (misses(lo, hi, wlo, whi)) continue;
        int path;
        if (misses(ri[0], ri[1], wlo, whi)) path = kPathFar;
        else if (misses(ri[2], ri[3], wlo, whi)) path = r2max(w) < mid_max ? kPathMid : kPathAny;
        else if (misses(ri[4], ri[5], wlo, whi)) path = r2max(w) < mid2_max ? kPathMid2 : kPathAny;
        else path = kPathAny;
        const int full = (lo <= wlo && hi >= whi) ? 0 : 1;
        code |= (1 + 2 * path + full) << (4 * w);
    }
    return code;
}

// FP64 per-line core constant q1 = sum_{n=1..13} e^{-a^2n^2} / (a^2 n^2 +
// y^2), a = kA64, with e^{-a^2 n^2} and a^2 n^2 as literals.
__device__ __forceinline__ double core_q1_64(double y) {
    constexpr double c[13] = {0.755683658862919, 0.32610700211508303, 0.08036376573837029,
                              0.011309424249254312, 0.0009088679210029174, 4.171008895272509e-05,
                              1.0931040894428156e-06, 1.6359197067600175e-08, 1.3981137635663895e-10,
                              6.823435648847615e-13, 1.9017076395301824e-15, 3.0266657632272252e-18,
                              2.7508394197950315e-21};
    constexpr double an2[13] = {0.2801324310308401, 1.1205297241233605, 2.5211918792775614,
                                4.482118896493442, 7.003310775771004, 10.084767517110246,
                                13.726489120511166, 17.928475585973768, 22.690726913498057,
                                28.013243103084015, 33.89602415473165, 40.33907006844098,
                                47.34238084421199};
    const double yy = y * y;
    double s = 0.0;
#pragma unroll
    for (int n = 0; n < 13; ++n) s = fma(c[n], rcp64(an2[n] + yy), s);
    return s;
}

// FP32 per-line core constant q1 = sum_{n=1..6} e^{-a^2n^2} / (a^2 n^2 + y^2),
// a = kA32, with e^{-a^2 n^2} and a^2 n^2 as literals.
__device__ __forceinline__ float core_q1_32(float y) {
    constexpr float c[6] = {0.5559463030549833f, 0.09552815420260209f, 0.005073359286059877f,
                            8.327709087244148e-05f, 4.2249460229375276e-07f,
                            6.624948055035271e-10f};
    constexpr float an2[6] = {0.5870835666346752f, 2.348334266538701f, 5.283752099712076f,
                              9.393337066154803f, 14.67708916586688f, 21.135008398848303f};
    const float yy = y * y;
    float s = 0.0f;
#pragma unroll
    for (int n = 0; n < 6; ++n) s = fmaf(c[n], rcp32(an2[n] + yy), s);
    return s;
}


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
                v = 0.0;            // line core: added by the line-core pass
            }
        }
        if (!FULL) {
            const int jr = jr0 + 64 * k;
            v = (jr >= lo && jr < hi) ? v : 0.0;
        }
        acc[k] += v;
    }
}

template <int PPT>
__device__ __forceinline__ void dispatch64(const Rec64 &q, int code, const double *xj, double *acc,
                                           int jr0) {
    switch (code) {
        case 1: eval64<kPathFar, true, PPT>(q, xj, acc, jr0); break;
        case 2: eval64<kPathFar, false, PPT>(q, xj, acc, jr0); break;
        case 3: eval64<kPathMid, true, PPT>(q, xj, acc, jr0); break;
        case 4: eval64<kPathMid, false, PPT>(q, xj, acc, jr0); break;
        case 5: eval64<kPathMid2, true, PPT>(q, xj, acc, jr0); break;
        case 6: eval64<kPathMid2, false, PPT>(q, xj, acc, jr0); break;
        case 7: eval64<kPathAny, true, PPT>(q, xj, acc, jr0); break;
        default: eval64<kPathAny, false, PPT>(q, xj, acc, jr0); break;
    }
}

// LM: stages of kThreads * LM lines (each thread stages LM lines per stage);
// results do not depend on it (lines are added in index order either way).
template <int PPT, int LM>
__device__ __forceinline__ void tile_f64(const VArgs &a, int64_t tile, Rec64 *rec) {
    constexpr int W = kThreads * PPT;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wlo = wave * 64 * PPT;
    const int jr0 = wlo + lane;
    const TileInfo ti = a.tiles[tile];
    const int64_t j0 = tile * W;
    double xj[PPT], acc[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        xj[k] = fma((double)(j0 + jr0 + 64 * k), a.dnu, a.numin);
        acc[k] = 0.0;
    }
    for (int64_t base = ti.lo; base < ti.hi; base += kThreads * LM) {
#pragma unroll 1
        for (int q = 0; q < LM; ++q) {
            const int64_t i = base + q * kThreads + t;
            Rec64 r;
            r.lo = 0; r.hi = 0; r.code = 0;
            if (i < ti.hi) {
                const double nu = a.nu[i];
                int64_t wl, wh;
                cut_window(nu, a.hw[i], a.numin, a.dnu, a.nx, wl, wh);
                const int64_t clo = max(wl, j0), chi = min(wh, j0 + W);
                if (clo < chi) {
                    const double ia = a.ia[i], y = a.y[i], y2 = y * y;
                    const double A = a.S[i] * kISqrtPi;
                    r.nu = nu; r.ia = ia; r.y2 = y2;
                    r.B = A * y * kISqrtPi; r.G = A * y * (2.0 / kPi);
                    r.p2 = 3.75 - 2.0 * y2; r.p3 = -15.0 * y2; r.p4 = 12.0 * y2 * y2;
                    r.lo = (int32_t)(clo - j0); r.hi = (int32_t)(chi - j0);
                    const double jcd = rint((nu - a.numin) / a.dnu);
                    const double rr = (fma(jcd, a.dnu, a.numin) - nu) * ia;
                    const double vb = a.dnu * ia, d0 = (double)j0 - jcd;
                    int32_t ri[6];
                    region_interval(kRF64, y2, rr, vb, d0, W, ri[0], ri[1]);
                    region_interval(kRM64, y2, rr, vb, d0, W, ri[2], ri[3]);
                    region_interval(kRB64, y2, rr, vb, d0, W, ri[4], ri[5]);
                    r.code = wave_codes<64 * PPT>(r.lo, r.hi, ri, [](int) { return 0.0; }, 1.0, 1.0);
                }
            }
            // the previous stage's readers of rec[] must be done before it is
            // written (the first record is built while they finish)
            if (q == 0) __syncthreads();
            rec[q * kThreads + t] = r;
        }
        __syncthreads();
        const int nst = (int)min((int64_t)kThreads * LM, ti.hi - base);
        for (int c0 = 0; c0 < nst; c0 += 64) {
            const int nib = c0 + lane < nst ? (rec[c0 + lane].code >> (4 * wave)) & 15 : 0;
            unsigned long long m = __ballot(nib != 0);       // 0: window misses this wave
            while (m) {
                const int b = lowest_bit(m);
                const int code = __builtin_amdgcn_readlane(nib, b);
                dispatch64<PPT>(rec[c0 + b], code, xj, acc, jr0);
            }
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

// Per-line core table stride (entries 1..kNT32 used).  Odd, so that lanes
// reading different lines' tables in one round of the core walk fall in
// different LDS banks (a stride of 16 put every fourth line on one bank:
// the c2 line-core pass 20 ms slower).
constexpr int kDT = 17;
// Largest x^2 + y^2 over a wave for the whole-wave 8-node FP32 path: keeps
// D12 D34 ~ r2^8 below the FP32 range.  The whole-wave series path has no
// such limit (its coefficients are bounded by y^2 < 4096).
constexpr double kMidMax32 = 1.0e300, kMid2Max32 = 3.0e4;

// Wings from the asymptotic series w(z) ~ (i/sqrt(pi)) sum_n c_n z^-(2n+1),
// c_n = (2n-1)!!/2^n.  With inv = 1/r2 and t = y^2 inv its real part is
//   Re w sqrt(pi) = y inv [1 + inv (3/2 - 2t) + inv^2 (15/4 - 15t + 12t^2)
//                   + inv^3 (105/8 - 105t + 210t^2 - 120t^3) + ...],
// regrouped in powers of inv with per-line coefficients in y^2 and the
// strength B = A y / sqrt(pi) folded in.
//  far32 (r2 >= 4096): terms n <= 1, truncation ~0.75/r2^2 (< 5e-8);
//  far4_32 (100 <= r2 < 4096, y^2 < 4096): terms n <= 3, truncation
//  ~59/r2^4 relative (< 1e-6 at r2 = 100).
__device__ __forceinline__ float far32(const Rec32 &q, float r2) {
    const float inv = rcp32(r2);
    return inv * fmaf(fmaf(q.BC, inv, 1.5f * q.B), inv, q.B);
}

__device__ __forceinline__ float far4_32(const Rec32 &q, float r2) {
    const float inv = rcp32(r2);
    float p = fmaf(q.q6, inv, q.q5);
    p = fmaf(p, inv, q.q4);
    p = fmaf(p, inv, q.q3);
    p = fmaf(p, inv, q.q2);
    p = fmaf(p, inv, 1.5f * q.B);
    p = fmaf(p, inv, q.B);
    return inv * p;
}

template <int PATH, bool FULL, int PPT>
__device__ __forceinline__ void eval32(const Rec32 &qr, float *part, int jr0) {
    const float r0 = qr.r, vbh = qr.vbh, vbl = qr.vbl, y2 = qr.y2, G = qr.G;
    const int32_t lo = qr.lo, hi = qr.hi;
    const float m0 = (float)(jr0 + qr.d0);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const float m = m0 + (float)(64 * k);
        float v;
        if (PATH == kPathFar || PATH == kPathMid) {
            // r2 >= 100 on these paths: |d ln v / dx| <= 2|x|/r2, so the
            // relative error 6e-8 |x| of x = m fl(vb) + r moves v by < 1.2e-7
            const float x = fmaf(m, vbh, r0);
            v = PATH == kPathFar ? far32(qr, fmaf(x, x, y2)) : far4_32(qr, fmaf(x, x, y2));
        } else {
            // these paths come with y^2 < kRF32 (else x^2 + y^2 >= kRF32 at
            // every point, the far path), where the four-term series holds
            // for every r2 >= kRM32: one wing formula, one branch fewer to
            // diverge on
            const float x = fmaf(m, vbh, fmaf(m, vbl, r0));
            const float r2 = fmaf(x, x, y2);
            if (r2 >= kRM32) {
                v = far4_32(qr, r2);
            } else if (PATH == kPathMid2 || r2 >= kRB32) {
                v = G * gh8_sum32(x * x, r2);
                if (qr.dop) v = fmaf(qr.A, __expf(-x * x), v);
            } else {
                v = 0.0f;           // line core: added by the line-core pass
            }
        }
        if (!FULL) {
            const int jr = jr0 + 64 * k;
            v = (jr >= lo && jr < hi) ? v : 0.0f;
        }
        part[k] += v;
    }
}

// The records of one path code that this wave takes (mask m over records
// c0 .. c0 + 63), in ascending order, with the next record's LDS fields
// loaded while the current one is evaluated (the loads are the latency of a
// record; an eval of the cheap paths is a few dozen instructions).  Two
// register sets take the records in turn, so no record is copied from one
// set to the other (a copy per record was 6-8 VALU instructions, a fifth of
// the far path's).
template <int PPT>
__device__ __forceinline__ void fold_parts(float *part, float *hi, float *lo, int &since) {
    if (++since == 32) {
#pragma unroll
        for (int k = 0; k < PPT; ++k) { two_sum_acc(hi[k], lo[k], part[k]); part[k] = 0.0f; }
        since = 0;
    }
}

template <int PATH, bool FULL, int PPT>
__device__ __forceinline__ void run_path(const Rec32 *rec, int c0, unsigned long long m,
                                         float *part, float *hi, float *lo, int &since,
                                         int jr0) {
    if (!m) return;
    Rec32 a = rec[c0 + lowest_bit(m)];
    // the first record's loads complete here, so that inside the loop only
    // the next record's loads are in flight during an evaluation
    __builtin_amdgcn_s_waitcnt(0xc07f);      // lgkmcnt(0)
    for (;;) {
        bool more = m != 0ull;
        Rec32 b;
        if (more) b = rec[c0 + lowest_bit(m)];
        eval32<PATH, FULL, PPT>(a, part, jr0);
        fold_parts<PPT>(part, hi, lo, since);
        if (!more) break;
        more = m != 0ull;
        if (more) a = rec[c0 + lowest_bit(m)];
        eval32<PATH, FULL, PPT>(b, part, jr0);
        fold_parts<PPT>(part, hi, lo, since);
        if (!more) break;
    }
}

// Position of this thread's valid record among the workgroup's valid
// records (line order) and their total; cnt[] is 4 ints of LDS.
__device__ __forceinline__ int compact_slot(bool valid, int *cnt, int &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long bal = __ballot(valid);
    const int below = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) cnt[wave] = __popcll(bal);
    __syncthreads();
    int off = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        const int c = cnt[w];
        off += w < wave ? c : 0;
        total += c;
    }
    return off + below;
}

// ---------------------------------------------------------------------------
// Far field of the FP32 form.  A line whose cut window covers the whole tile
// and whose centre lies well away from the tile contributes a function that
// is analytic over the tile with its nearest singularity (the line centre)
// outside it.  Such lines are not evaluated point by point: each adds its
// FP64 values at the tile's Chebyshev-Gauss nodes, and the sum over them is
// interpolated to the tile's points once.  Three bands, by the distance d
// of the centre from the tile's centre in tile half widths:
//   band 1, d >= 4:       9 nodes;
//   band 2, 2 <= d < 4:   15 nodes;
//   band 3, 1.5 <= d < 2: 19 nodes.
// The worst line of a band has its centre at the band's inner edge and the
// narrowest profile, 1/(u - d)^2 over u in [-1, 1] (a double pole at the
// centre); the pointwise relative interpolation error there is 2.2e-7
// (band 1), 1.4e-7 (band 2) and 9.9e-7 (band 3), largest at the tile end
// nearest the line.  Far lines contribute positive values, so the far
// field's share of K carries at most 1e-6 relative.
// (tests/test_far_field_bounds.py measures these on the constants below.)
// The node values use the asymptotic series w(z) sqrt(pi) = (i/z) sum_n c_n
// z^-2n, c_n = (2n-1)!!/2^n, n <= 3, which requires x^2 + y^2 >= kFarR2 over
// the tile (first omitted term of the real part (105/16) 9 y / x^10 against
// y / x^2: < 3e-9 relative).
// ---------------------------------------------------------------------------
constexpr double kFarR2 = 400.0;
constexpr int kFarMinPts = 256;       // smaller (edge) tiles sum every pair
constexpr int kFN1 = 9, kFN2 = 15, kFN3 = 19;
constexpr double kFarRatio1 = 4.0, kFarRatio2 = 2.0, kFarRatio3 = 1.5;
// Chebyshev-Gauss nodes cos(pi (k + 1/2) / n)
__constant__ double kFarU1[kFN1] = {
    0.98480775301220802032, 0.86602540378443870761, 0.64278760968653936292,
    0.34202014332566882393, 0, -0.34202014332566849086,
    -0.64278760968653936292, -0.86602540378443848557, -0.98480775301220802032};
__constant__ double kFarU2[kFN2] = {
    0.99452189536827328986, 0.95105651629515353118, 0.86602540378443870761,
    0.74314482547739424412, 0.5877852522924731371, 0.4067366430758003748,
    0.20791169081775923155, 0, -0.20791169081775934258,
    -0.40673664307580004174, -0.58778525229247302608, -0.74314482547739402207,
    -0.86602540378443870761, -0.95105651629515353118, -0.99452189536827340088};

__constant__ double kFarU3[kFN3] = {
    0.99658449300666984705, 0.96940026593933037447, 0.91577332665505739584,
    0.83716647826252865716, 0.73572391067313169799, 0.61421271268966792789,
    0.47594739303707345179, 0.32469946920468356621, 0.16459459028073375464,
    0, -0.16459459028073383791, -0.32469946920468323315,
    -0.47594739303707356282, -0.61421271268966781687, -0.73572391067313158697,
    -0.83716647826252832409, -0.91577332665505750686, -0.96940026593933037447,
    -0.99658449300666984705};
template <int N>
__device__ __forceinline__ double far_u(int k) {
    return N == kFN1 ? kFarU1[k] : (N == kFN2 ? kFarU2[k] : kFarU3[k]);
}

// One band's node sums (FP32, compensated: hi + lo), their totals and the
// Chebyshev coefficients.  Each wave keeps its own column of sums, so no two
// waves write one word.  Lines that reach every point of the tile add to the
// common columns (hi, lo); a line whose window ends inside the tile adds only
// to the targets (the waves whose points it covers) in the target columns
// (xhi, xlo).  Totals and coefficients are per target wave: each wave
// interpolates its own sums at its own points.
constexpr int kWaves = kThreads / 64;
template <int N>
struct FarBand {
    float hi[N][kWaves];              // common node sums, one column per wave
    float lo[N][kWaves];
    // [node][target][wave]; band 1 only: a line whose window ends inside the
    // tile lies a cut from its centre, far from the tile, and the other
    // bands' lines with such windows stay with the pair-by-pair path
    float xhi[N == kFN1 ? N : 1][kWaves][kWaves];
    float xlo[N == kFN1 ? N : 1][kWaves][kWaves];
};

// Node totals and Chebyshev coefficients per target wave.  They exist only
// after the line loop, so they live in the record array's LDS (FarCoefs).
template <int N>
struct FarCoef {
    double F[kWaves][N];
    double c[kWaves][N];
};

// Wave-level far field (wave_far_codes): per wave, node sums of its three
// bands one after the other (9, 15, 19 nodes) and their coefficients.
constexpr int kWN = kFN1 + kFN2 + kFN3;
constexpr int kWOff2 = kFN1, kWOff3 = kFN1 + kFN2;

struct FarCoefs {
    FarCoef<kFN1> b1;
    FarCoef<kFN2> b2;
    FarCoef<kFN3> b3;
    double wc[kWaves][kWN];
};

struct FarLDS {
    FarBand<kFN1> b1;
    FarBand<kFN2> b2;
    FarBand<kFN3> b3;
    float ws[kWaves][kWN][2];         // wave-level node sums (hi, lo)
    int cnt[2][kWaves];               // compaction counts, alternate stages
    int any1, any2, any3;             // band used (set to 1 by any wave)
};

// Re w(x, y) / sqrt(pi) from the asymptotic series in FP32 (x^2 + y^2 >=
// kFarR2): w sqrt(pi) = (i/z) S, S = sum_n c_n z^-2n, n <= 3.  Every term
// is below the previous by >= 1/r2 <= 2.5e-3, so the sum carries a few
// FP32 roundings (< 4e-7 relative).
__device__ __forceinline__ float far_node32(float x, float y) {
    const float r2 = fmaf(x, x, y * y);
    const float ir = rcp32(r2);
    const float zr = x * ir, zi = -y * ir;                 // 1/z
    const float ur = (zr - zi) * (zr + zi), ui = 2.0f * zr * zi;   // 1/z^2
    float sr = 1.875f, si = 0.0f;
#pragma unroll
    for (int n = 2; n >= 0; --n) {
        const float cn = n == 2 ? 0.75f : n == 1 ? 0.5f : 1.0f;
        const float tr = fmaf(sr, ur, fmaf(-si, ui, cn));
        si = fmaf(sr, ui, si * ur);
        sr = tr;
    }
    // Re(i S / z) = -(Im S Re(1/z) + Re S Im(1/z))
    return -(si * zr + sr * zi) * (float)(1.0 / kPi);
}

// DPP move of a double: each lane takes the value of the lane that the
// (compile-time) DPP control selects; rows outside row_mask keep 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp32(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

// Sum of v over the wave, complete in lane 63 (other lanes hold partial
// sums), in a fixed order: pairs and quads (quad_perm), 8 and 16 lanes (row
// half mirror, row mirror), then rows 0+1 and 2+3 (row_bcast:15 into rows 1
// and 3) and all four (row_bcast:31 into rows 2 and 3).  DPP moves are VALU
// operations; a butterfly of __shfl_xor goes through ds_bpermute instead and
// waits on the LDS path at every step.  For v >= 0 (node values) the six
// levels of the tree leave the total within 6 roundings (3.6e-7) of exact.
__device__ __forceinline__ float wave_total32(float v) {
    v += dpp32<0xb1, 0xf>(v);     // quad_perm [1,0,3,2]
    v += dpp32<0x4e, 0xf>(v);     // quad_perm [2,3,0,1]
    v += dpp32<0x141, 0xf>(v);    // row_half_mirror
    v += dpp32<0x140, 0xf>(v);    // row_mirror
    v += dpp32<0x142, 0xa>(v);    // row_bcast:15
    v += dpp32<0x143, 0xc>(v);    // row_bcast:31
    return v;
}

// Maximum of v over the wave, complete in lane 63 (the DPP tree of
// wave_total32).
__device__ __forceinline__ int wave_max_i32(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));
    return v;
}

// Far-field band (0: evaluated point by point) of a line (centre nu, cut
// half width hw, 1/alpha_D ia, y) for the tile [j0, j0 + Wv) with half
// width hW (points): its centre is at least kFarRatio3 half widths from the
// tile's centre (band 2 from kFarRatio2 on, band 1 from kFarRatio1 on) and
// x^2 + y^2 >= kFarR2 at the nearest point.  full: the window covers the
// tile (the cut predicate |x_j - nu| < hw holds at both end points, so at
// all of them).  No division, so it is cheap enough to run for every staged
// line before its window is computed.
__device__ __forceinline__ int far_band(double nu, double hw, double ia, double y, int64_t j0,
                                        int64_t Wv, double hW, double numin, double dnu,
                                        bool &full) {
    full = false;
    if (Wv < kFarMinPts || !(hw > 0.0)) return 0;
    const double half = hW * dnu;
    const double d = fabs(nu - fma((double)j0 + hW, dnu, numin));
    if (d < kFarRatio3 * half) return 0;
    const double xnear = (d - half) * ia;
    if (!(fma(xnear, xnear, y * y) >= kFarR2)) return 0;
    const double dlo = fma((double)j0, dnu, numin) - nu;
    const double dhi = fma((double)(j0 + Wv - 1), dnu, numin) - nu;
    full = dlo > -hw && dlo < hw && dhi > -hw && dhi < hw;
    return d >= kFarRatio1 * half ? 1 : (d >= kFarRatio2 * half ? 2 : 3);
}

// Line offsets of the FP32 records: jc = the grid point nearest the centre
// (may lie off the grid), rr = (x_jc - nu) / alpha_D, vb = dnu / alpha_D and
// d0 = j0 - jc, so that x at tile-relative point j is (j + d0) vb + rr.
__device__ __forceinline__ void line_offsets(double nu, double ia, double numin, double dnu,
                                             int64_t j0, double &rr, double &vb, double &d0) {
    const double jcd = rint((nu - numin) / dnu);
    rr = (fma(jcd, dnu, numin) - nu) * ia;
    vb = dnu * ia;
    d0 = (double)j0 - jcd;
}

// A far line whose window ends inside the tile: the window's tile-relative
// part [lo, hi) (valid points, Wv of them) is split into the waves it covers
// (wave w owns [w wpts, (w + 1) wpts)), which take the line from the far
// field, and the rest, which stays with the pair-by-pair path: lo and hi are
// moved past the covered waves.  Returns the covered-wave mask with the waves
// that own no valid point set too (mask 0xF: the far field takes every
// point); 0 when no wave with valid points is covered.  The centre lies
// outside the tile, so the window reaches one end of it and the covered waves
// are a run at that end: [lo, hi) stays one interval.
__device__ __forceinline__ int far_cover(int32_t &lo, int32_t &hi, int Wv, int wpts) {
    int mask = 0, vmask = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const int a = w * wpts, b = min(a + wpts, Wv);
        if (a >= Wv) mask |= 1 << w;
        else {
            vmask |= 1 << w;
            if (lo <= a && hi >= b) mask |= 1 << w;
        }
    }
    if (!(mask & vmask)) return 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        if (lo < hi && ((mask >> (lo / wpts)) & 1)) lo = min((lo / wpts + 1) * wpts, hi);
        if (lo < hi && ((mask >> ((hi - 1) / wpts)) & 1)) hi = max(((hi - 1) / wpts) * wpts, lo);
    }
    return mask;
}

// The tile bands' distance rules for an interval of half width `half`
// whose centre lies dc from the line centre, both in Doppler units, for
// the wave-level far field: also x^2 + y^2 >= kRF32 at the interval's
// nearest point, so the node values take the two-term series of far32.
__device__ __forceinline__ int wave_band(double dc, double half, double y2) {
    if (dc < kFarRatio3 * half) return 0;
    const double xn = dc - half;
    if (!(fma(xn, xn, y2) >= (double)kRF32)) return 0;
    return dc >= kFarRatio1 * half ? 1 : (dc >= kFarRatio2 * half ? 2 : 3);
}

// Wave-level far field.  A line that the tile takes pair by pair (its centre
// too near the tile for the tile's far field) can lie far from some of the
// tile's waves: wave w (tile-relative points [w wpts, (w + 1) wpts), all
// valid and all inside the line's window) takes the line from a far field of
// its own -- Chebyshev nodes over the wave's interval, the tile bands' rules
// with the wave's half width -- instead of pair by pair.  x at tile-relative
// point j is (j + d0) vb + rr.  Returns per wave (4 bits each) 8 + band, or 0.
template <int WPTS>
__device__ __forceinline__ int32_t wave_far_codes(int32_t lo, int32_t hi, int64_t Wv, double d0,
                                                  double vb, double rr, double y2) {
    constexpr double hwp = 0.5 * (double)(WPTS - 1);
    const double half = hwp * vb;
    int32_t code = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const int a = w * WPTS, b = a + WPTS;
        if ((int64_t)b > Wv || lo > a || hi < b) continue;
        const int band = wave_band(fabs(((double)a + hwp + d0) * vb + rr), half, y2);
        if (band) code |= (8 + band) << (4 * w);
    }
    return code;
}

template <int N>
__device__ __forceinline__ void band_init(FarBand<N> *b, int t) {
    for (int e = t; e < N * kWaves; e += kThreads) {
        (&b->hi[0][0])[e] = 0.0f;
        (&b->lo[0][0])[e] = 0.0f;
    }
    if constexpr (N == kFN1) {
        for (int e = t; e < N * kWaves * kWaves; e += kThreads) {
            (&b->xhi[0][0][0])[e] = 0.0f;
            (&b->xlo[0][0][0])[e] = 0.0f;
        }
    }
}

// The wave's lines of this band (mine) add their node values to the wave's
// columns of node sums (all lanes take part in the wave sums).  A line's
// Doppler-unit offset from the tile centre D and the tile half width h come
// in FP64 (one per line and tile); each node's x = u_k h + D is rounded to
// FP32 once, and the series, the strength and the wave sum are FP32.  The
// column sums over the stages are compensated (two-sum, hi + lo).  When a
// line of the wave covers only some waves (cover != 0xF), the wave sums go
// per target wave into the target columns instead (four wave sums a node).
template <int N>
__device__ __forceinline__ void band_nodes(FarBand<N> *b, bool mine, int cover, double D, double h,
                                           float fy, float fB, int lane, int wave) {
    const bool split = __ballot(mine && cover != 0xF) != 0ull;
    // x^2 + y^2 >= kRF32 at every node of every line of the wave: the
    // two-term series of the pair-by-pair far path (truncation < 5e-8)
    // instead of the four-term complex one
    const float fD = (float)D, fh = (float)h;
    const double xn = fabs(D) - h;
    const bool series2 = __ballot(mine && !(fma(xn, xn, (double)fy * fy) >= (double)kRF32)) == 0ull;
    const float y2 = fy * fy, B = fB * fy * (float)(1.0 / kPi);
    const float B15 = 1.5f * B, BC = fmaxf(-2.0f * y2 * B, -1.0e37f);
    // the wave totals are gathered into lanes (node k, or target tw and node
    // k, in lane tw N + k) and added to the columns in one pass at the end
    float tot = 0.0f;
#pragma unroll 1
    for (int k = 0; k < N; ++k) {
        // far lines: |x| >= |D| - h >= |D| / 3, so x in FP32 from FP32 D and
        // h keeps ~1e-7 of x
        const float x = fmaf((float)far_u<N>(k), fh, fD);
        float v = 0.0f;
        if (series2) {
            const float inv = rcp32(fmaf(x, x, y2));
            v = mine ? inv * fmaf(fmaf(BC, inv, B15), inv, B) : 0.0f;
        } else {
            v = mine ? fB * far_node32(x, fy) : 0.0f;
        }
        if (N != kFN1 || !split) {
            const float sw = __int_as_float(
                __builtin_amdgcn_readlane(__float_as_int(wave_total32(v)), 63));
            tot = lane == k ? sw : tot;
        } else {
#pragma unroll
            for (int tw = 0; tw < kWaves; ++tw) {
                const float sw = __int_as_float(__builtin_amdgcn_readlane(
                    __float_as_int(wave_total32((cover >> tw) & 1 ? v : 0.0f)), 63));
                tot = lane == tw * N + k ? sw : tot;
            }
        }
    }
    if (N != kFN1 || !split) {
        if (lane < N) {
            float hi = b->hi[lane][wave], lo = b->lo[lane][wave];
            two_sum_acc(hi, lo, tot);
            b->hi[lane][wave] = hi;
            b->lo[lane][wave] = lo;
        }
    } else if (lane < kWaves * N) {
        const int tw = lane / N, k = lane - tw * N;
        float hi = b->xhi[k][tw][wave], lo = b->xlo[k][tw][wave];
        two_sum_acc(hi, lo, tot);
        b->xhi[k][tw][wave] = hi;
        b->xlo[k][tw][wave] = lo;
    }
}

// Node totals -> Chebyshev coefficients, per target wave, by threads t0 ..
// t0 + 4N - 1 (two barriers are taken by the caller between the phases).
// The interpolation -- totals, coefficients and the per-point Clenshaw sums
// -- runs in FP64: a band-3 tile's far field varies by up to 25x across it,
// and an FP32 Clenshaw sum would carry up to ~3e-5 of the smallest value.
template <int N>
__device__ __forceinline__ void band_sum(const FarBand<N> *b, FarCoef<N> *fc, int t) {
    if (t >= 0 && t < N * kWaves) {
        const int tw = t / N, k = t % N;
        double F = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) F += (double)b->hi[k][w] + (double)b->lo[k][w];
        if constexpr (N == kFN1) {
#pragma unroll
            for (int w = 0; w < kWaves; ++w) F += (double)b->xhi[k][tw][w] + (double)b->xlo[k][tw][w];
        }
        fc->F[tw][k] = F;
    }
}

template <int N>
__device__ __forceinline__ void band_coeffs(FarCoef<N> *b, int t) {
    if (t >= 0 && t < N * kWaves) {
        const int tw = t / N, m = t % N;
        // c_m = (2 - [m == 0]) / N sum_k F_k T_m(u_k), T_m by its recurrence
        // (not unrolled: unrolled, the node constants are hoisted out of the
        // persistent kernel's item loop and held in registers throughout)
        double c = 0.0;
#pragma unroll 1
        for (int k = 0; k < N; ++k) {
            const double u = far_u<N>(k);
            double tm = 1.0, tc = u;
            for (int q = 2; q <= m; ++q) {
                const double tn = fma(2.0 * u, tc, -tm);
                tm = tc;
                tc = tn;
            }
            c += b->F[tw][k] * (m == 0 ? 1.0 : tc);
        }
        b->c[tw][m] = (m == 0 ? 1.0 : 2.0) / (double)N * c;
    }
}

// Clenshaw sum of an N-term Chebyshev series at u in [-1, 1].
template <int N>
__device__ __forceinline__ double cheb_eval(const double *c, double u) {
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int m = N - 1; m >= 1; --m) {
        const double b0 = fma(2.0 * u, b1, c[m] - b2);
        b2 = b1;
        b1 = b0;
    }
    return fma(u, b1, c[0] - b2);
}

// Window ends inside a wave (cut edges).  A far line whose cut window ends
// inside the tile (about one cut from it) leaves, after the far field takes
// the waves its window covers, a remainder in one wave: [lo, wave end) for a
// line right of the tile (code 14), [wave start, hi) for one left of it
// (code 15).  Taken this way only when its centre lies >= kEdgeMin wave
// half-widths from the wave's centre, where it is smooth over the wave: each
// lane of a 64-record chunk expands its line over
// the wave's interval in kEN Chebyshev terms (far32 at the nodes), and the
// terms are summed over the lines whose windows reach each point: with the
// lines' window ends in lane order, those lines are a prefix (right) or a
// suffix (left) of the chunk, so an inclusive lane scan of the terms, read
// at the point's count of open windows, gives the sum.  Chunks whose window
// ends are not in lane order, or that hold both kinds, are taken record by
// record on the far path.  kEN = 5 nodes keep the interpolation within
// 3.8e-7 of a line's value at kEdgeMin = 16 half-widths (a window end
// inside a band-1 tile can lie as near as 11, where 5 nodes would allow
// 2.5e-6: those lines keep their records; at a cut of many tiles, as at
// 100 cm^-1 and 1024 points, every cut edge is far beyond 16).
constexpr int kEN = 5;
constexpr double kEdgeMin = 16.0;
__device__ __forceinline__ float edge_u(int k) {
    return k == 0 ? 9.5105651630e-01f : k == 1 ? 5.8778525229e-01f : k == 2 ? 0.0f
         : k == 3 ? -5.8778525229e-01f : -9.5105651630e-01f;
}
// (2 - [m == 0]) / kEN T_m(u_k)
__device__ __forceinline__ float edge_w(int m, int k) {
    constexpr float w[kEN][kEN] = {
        {2.0e-01f, 2.0e-01f, 2.0e-01f, 2.0e-01f, 2.0e-01f},
        {3.8042260652e-01f, 2.3511410092e-01f, 0.0f, -2.3511410092e-01f, -3.8042260652e-01f},
        {3.2360679775e-01f, -1.2360679775e-01f, -4.0e-01f, -1.2360679775e-01f, 3.2360679775e-01f},
        {2.3511410092e-01f, -3.8042260652e-01f, 0.0f, 3.8042260652e-01f, -2.3511410092e-01f},
        {1.2360679775e-01f, -3.2360679775e-01f, 4.0e-01f, -3.2360679775e-01f, 1.2360679775e-01f}};
    return w[m][k];
}

// Inclusive scans over the wave's lanes in lane order (DPP: rows of 16, then
// rows 0+1 and 2+3, then all four); fixed order, so deterministic.  Every
// lane must be active.
__device__ __forceinline__ float wave_scan32(float v) {
    v += dpp32<0x111, 0xf>(v);    // row_shr:1
    v += dpp32<0x112, 0xf>(v);    // row_shr:2
    v += dpp32<0x114, 0xf>(v);    // row_shr:4
    v += dpp32<0x118, 0xf>(v);    // row_shr:8
    v += dpp32<0x142, 0xa>(v);    // row_bcast:15
    v += dpp32<0x143, 0xc>(v);    // row_bcast:31
    return v;
}

__device__ __forceinline__ int wave_scan_max_i32(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x111, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x112, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x114, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x118, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x143, 0xc, 0xf, false));
    return v;
}

// ds_bpermute reads other lanes' values; a lane outside the exec mask reads
// as 0, so these are called with every lane active (never under a branch).
__device__ __forceinline__ float lane_get(float v, int l) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(4 * l, __float_as_int(v)));
}
__device__ __forceinline__ int lane_get_i(int v, int l) {
    return __builtin_amdgcn_ds_bpermute(4 * l, v);
}

// One chunk's edge records (nib 14 / 15 for this wave).  Returns false when
// the chunk must be taken record by record.
template <int PPT>
__device__ __forceinline__ bool edge_chunk(const Rec32 *rec, int c0, int nst, int nib,
                                           float *part, int jr0, float wcen, float whp) {
    const int lane = threadIdx.x & 63;
    const bool er = nib == 14, el = nib == 15;
    const unsigned long long mr = __ballot(er), ml = __ballot(el);
    if (mr && ml) return false;
    const bool right = mr != 0ull;
    const bool mine = right ? er : el;
    // lane order of the scan: as loaded (right) or reversed (left), so that
    // the lines reaching a point are a prefix
    const int src = right ? lane : 63 - lane;
    const Rec32 q = rec[min(c0 + src, nst - 1)];
    const bool act = right ? ((mr >> src) & 1ull) : ((ml >> src) & 1ull);
    (void)mine;
    // keys: window start (right, non-decreasing) or end (left, taken as its
    // negative so that the reversed order is non-decreasing too); lanes
    // without a line take the running maximum
    const int key = act ? (right ? q.lo : -q.hi) : INT_MIN;
    const int run = wave_scan_max_i32(key);
    if (__ballot(act && run != key)) return false;    // window ends out of order
    // Chebyshev terms of this lane's line over the wave
    float c[kEN];
#pragma unroll
    for (int m = 0; m < kEN; ++m) c[m] = 0.0f;
    if (act) {
        const float m0 = wcen + (float)q.d0;
#pragma unroll
        for (int k = 0; k < kEN; ++k) {
            const float mm = fmaf(edge_u(k), whp, m0);
            const float x = fmaf(mm, q.vbh, fmaf(mm, q.vbl, q.r));
            const float v = far32(q, fmaf(x, x, q.y2));
#pragma unroll
            for (int m = 0; m < kEN; ++m) c[m] = fmaf(edge_w(m, k), v, c[m]);
        }
    }
#pragma unroll
    for (int m = 0; m < kEN; ++m) c[m] = wave_scan32(c[m]);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int j = jr0 + 64 * k;
        // lines reaching point j: key <= j (right: lo <= j; left: -hi <= -j - 1)
        const int kj = right ? j : -j - 1;
        int pos = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if (lane_get_i(run, pos + step - 1) <= kj) pos += step;
        // the halving steps reach at most 63: the 64th lane is one more test
        const int r63 = lane_get_i(run, 63);
        if (pos == 63 && r63 <= kj) pos = 64;
        // the prefix sum at lane pos - 1 (none when pos == 0); the reads are
        // unconditional, with every lane active
        const int ls = max(pos - 1, 0);
        float cs[kEN];
#pragma unroll
        for (int m = 0; m < kEN; ++m) {
            const float t = lane_get(c[m], ls);
            cs[m] = pos > 0 ? t : 0.0f;
        }
        const float u = ((float)j - wcen) / whp;
        float b1 = 0.0f, b2 = 0.0f;
#pragma unroll
        for (int m = kEN - 1; m >= 1; --m) {
            const float b0 = fmaf(2.0f * u, b1, cs[m] - b2);
            b2 = b1;
            b1 = b0;
        }
        part[k] += fmaf(u, b1, cs[0] - b2);
    }
    return true;
}

// Wave-level far field, node sums: lane l holds the record of its own line
// (record c0 + l of the stage); the lanes whose line this wave takes in
// band N (mine) add its values at the N nodes of the wave's interval
// (centre cw, half width hwp in points) to the wave's compensated sums s.
// x = m vb + rr in FP32 as in the pair-by-pair path, m = cw + d0 + u_k hwp
// (|x| >= |D| / 3 for these lines); the value is far32's two-term series.
template <int N>
__device__ __forceinline__ void wave_far_nodes(const Rec32 &q, bool mine, float cw, float hwp,
                                               float (*s)[2], int lane) {
    const float m0 = cw + (float)q.d0;
    float tot = 0.0f;           // lane k: node k's wave total
#pragma unroll 1
    for (int k = 0; k < N; ++k) {
        const float m = fmaf((float)far_u<N>(k), hwp, m0);
        const float x = fmaf(m, q.vbh, fmaf(m, q.vbl, q.r));
        const float v = mine ? far32(q, fmaf(x, x, q.y2)) : 0.0f;
        const float sw = __int_as_float(
            __builtin_amdgcn_readlane(__float_as_int(wave_total32(v)), 63));
        tot = lane == k ? sw : tot;
    }
    if (lane < N) {
        float h = s[lane][0], l = s[lane][1];
        two_sum_acc(h, l, tot);
        s[lane][0] = h;
        s[lane][1] = l;
    }
}

// Wave-level far field, coefficients of one band from the wave's node sums
// (FP64, as the tile bands), by lane m < N.
template <int N>
__device__ __forceinline__ void wave_far_coeffs(const float (*s)[2], double *c, int m) {
    if (m < 0 || m >= N) return;
    double acc = 0.0;
#pragma unroll 1
    for (int k = 0; k < N; ++k) {
        const double u = far_u<N>(k);
        double tm = 1.0, tc = u;
        for (int q = 2; q <= m; ++q) {
            const double tn = fma(2.0 * u, tc, -tm);
            tm = tc;
            tc = tn;
        }
        acc += ((double)s[k][0] + (double)s[k][1]) * (m == 0 ? 1.0 : tc);
    }
    c[m] = (m == 0 ? 1.0 : 2.0) / (double)N * acc;
}

template <int PPT, int LM>
__device__ __forceinline__ void tile_f32c(const VArgs &a, int64_t tile, Rec32 *rec, FarLDS *fl) {
    constexpr int W = kThreads * PPT;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wlo = wave * 64 * PPT;
    const int jr0 = wlo + lane;
    const TileInfo ti = a.tiles[tile];
    const int64_t j0 = tile * W;
    float hi[PPT], lo[PPT], part[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) { hi[k] = 0.0f; lo[k] = 0.0f; part[k] = 0.0f; }
    int since = 0;
    // far field: valid points of the tile, half width, nodes
    const int64_t Wv = min((int64_t)W, a.nx - j0);
    const double hW = 0.5 * (double)(Wv - 1);
    const bool ffok = a.far && Wv >= kFarMinPts;
    band_init<kFN1>(&fl->b1, t);
    band_init<kFN2>(&fl->b2, t);
    band_init<kFN3>(&fl->b3, t);
    // wave-level far field: this wave's node sums and the bands it used
    constexpr float whp = 0.5f * (float)(64 * PPT - 1);
    const float wcen = (float)wlo + whp;
    float (*ws)[2] = fl->ws[wave];
    for (int e = lane; e < 2 * kWN; e += 64) (&ws[0][0])[e] = 0.0f;
    int wused = 0;
    const double xc = fma((double)j0 + hW, a.dnu, a.numin);     // tile centre [cm^-1]
    if (t == 0) { fl->any1 = 0; fl->any2 = 0; fl->any3 = 0; }
    __syncthreads();
    int nprev = 0, par = 0;
    for (int64_t base = ti.lo; base < ti.hi; base += kThreads * LM) {
      int nst = 0;
#pragma unroll 1
      for (int q = 0; q < LM; ++q) {
        const int64_t i = base + q * kThreads + t;
        Rec32 r;
        r.lo = 0; r.hi = 0; r.code = 0;
        // the line's window on the tile, [clo, chi) tile-relative, less the
        // waves the far field takes
        int32_t clo = 0, chi = 0;
        bool edge1 = false;             // band-1 far line whose window ends in the tile
        // far lines first: FP32 values at the nodes, summed per wave (their
        // registers are free again before the direct records are built)
        {
            double fD = 0.0, fh = 0.0;
            float fy = 0.0f, fB = 0.0f;
            int band = 0, cover = 0;
            if (i < ti.hi) {
                const double fnu = a.nu[i], fhw = a.hw[i];
                bool full = false;
                if (ffok) band = far_band(fnu, fhw, a.ia[i], a.y[i], j0, Wv, hW, a.numin, a.dnu, full);
                if (band && full) {
                    cover = 0xF;
                } else {
                    int64_t wl, wh;
                    cut_window(fnu, fhw, a.numin, a.dnu, a.nx, wl, wh);
                    clo = (int32_t)(max(wl, j0) - j0);
                    chi = (int32_t)(min(wh, j0 + W) - j0);
                    if (band) {
                        edge1 = band == 1;
                        cover = band == 1 && clo < chi ? far_cover(clo, chi, (int)Wv, 64 * PPT) : 0;
                        if (!cover) band = 0;
                    }
                }
                if (band) {
                    const double fia = a.ia[i];
                    fD = (xc - fnu) * fia;
                    fh = hW * a.dnu * fia;
                    fy = (float)a.y[i];
                    fB = (float)ldexp(a.S[i], -ti.e);
                }
            }
            if (__ballot(band == 1) != 0ull) {
                band_nodes<kFN1>(&fl->b1, band == 1, cover, fD, fh, fy, fB, lane, wave);
                if (lane == 0) fl->any1 = 1;
            }
            if (__ballot(band == 2) != 0ull) {
                band_nodes<kFN2>(&fl->b2, band == 2, cover, fD, fh, fy, fB, lane, wave);
                if (lane == 0) fl->any2 = 1;
            }
            if (__ballot(band == 3) != 0ull) {
                band_nodes<kFN3>(&fl->b3, band == 3, cover, fD, fh, fy, fB, lane, wave);
                if (lane == 0) fl->any3 = 1;
            }
        }
        if (clo < chi) {
            const double nu = a.nu[i];
            {
                const double ia = a.ia[i], y = a.y[i];
                // nearest grid point to the centre (may lie off the grid)
                double rr, vb, d0;
                line_offsets(nu, ia, a.numin, a.dnu, j0, rr, vb, d0);
                const float vbh = (float)vb;
                const double A = ldexp(a.S[i], -ti.e) * kISqrtPi;
                const float yf = (float)y;
                r.r = (float)rr;
                r.vbh = vbh;
                r.vbl = (float)(vb - (double)vbh);
                r.y2 = yf * yf;
                r.A = (float)A;
                const double B = A * y * kISqrtPi, G = A * y * (2.0 / kPi);
                const double yy = (double)r.y2;
                r.B = (float)B;
                // -2 y^2 B may exceed FP32 only for y > ~1e6, where the term
                // it multiplies (-2 y^2 / r2^2 <= 2 / r2) is negligible
                r.BC = (float)fmax(-2.0 * yy * B, -1.0e37);
                r.G = (float)G;
                // series coefficients of far4_32, used only where y^2 < 4096
                const double Bm = yy < (double)kRF32 ? B : 0.0;
                r.q2 = (float)((3.75 - 2.0 * yy) * Bm);
                r.q3 = (float)((13.125 - 15.0 * yy) * Bm);
                r.q4 = (float)((12.0 * yy - 105.0) * yy * Bm);
                r.q5 = (float)(210.0 * yy * yy * Bm);
                r.q6 = (float)(-120.0 * yy * yy * yy * Bm);
                r.d0 = (int32_t)d0;
                r.lo = clo;
                r.hi = chi;
                r.dop = yf < kY0 ? 1 : 0;
                const double y2 = (double)yf * (double)yf;
                int32_t ri[6];
                region_interval((double)kRF32, y2, rr, vb, d0, W, ri[0], ri[1]);
                region_interval((double)kRM32, y2, rr, vb, d0, W, ri[2], ri[3]);
                region_interval((double)kRB32, y2, rr, vb, d0, W, ri[4], ri[5]);
                // x is monotonic over a wave: its largest x^2 is at an end
                auto r2max = [&](int w) {
                    const double xa = ((double)(w * 64 * PPT) + d0) * vb + rr;
                    const double xb = ((double)((w + 1) * 64 * PPT - 1) + d0) * vb + rr;
                    return fmax(xa * xa, xb * xb) + y2;
                };
                r.code = wave_codes<64 * PPT>(r.lo, r.hi, ri, r2max, kMidMax32, kMid2Max32);
                if (ffok) {
                    // waves far from the line take it from their own far field
                    const int32_t wf = wave_far_codes<64 * PPT>(r.lo, r.hi, Wv, d0, vb, rr, y2);
                    r.code = (r.code & ~(((wf & 0x88888888) >> 3) * 0xF)) | wf;
                }
                if (edge1) {
                    // the window's remainder in one wave, on the far path,
                    // from a centre >= kEdgeMin wave half-widths away: an
                    // edge record of that wave (14: right, 15: left)
                    constexpr int wpts = 64 * PPT;
                    constexpr double hwp = 0.5 * (wpts - 1);
                    const int ws = r.lo / wpts;
                    if ((r.hi - 1) / wpts == ws && ((r.code >> (4 * ws)) & 15) == 2 &&
                        fabs((double)ws * wpts + hwp + d0) >= kEdgeMin * hwp) {
                        const int ea = ws * wpts, eb = (int)min((int64_t)(ea + wpts), Wv);
                        const int ty = r.hi >= eb ? 14 : (r.lo <= ea ? 15 : 0);
                        if (ty) r.code = (r.code & ~(15 << (4 * ws))) | (ty << (4 * ws));
                    }
                }
            }
        }
        // Barriers: the previous stage's readers of rec[] must be done before
        // it is written, and the writes before this stage reads it; a stage
        // of far lines only (nst == 0) needs neither, only the one inside
        // compact_slot.  The counts alternate between two buffers, so a wave
        // that skips ahead cannot overwrite counts another wave has not read
        // (it cannot pass the next compact_slot barrier before that wave).
        if (q == 0 && nprev > 0) __syncthreads();
        int nq;
        const int slot = compact_slot(r.code != 0, fl->cnt[par], nq);
        par ^= 1;
        if (r.code != 0) rec[nst + slot] = r;
        nst += nq;
      }
        if (nst > 0) __syncthreads();
        nprev = nst;
        for (int c0 = 0; c0 < nst; c0 += 64) {
            const int nib = c0 + lane < nst ? (rec[c0 + lane].code >> (4 * wave)) & 15 : 0;
            // 0: the window misses this wave; otherwise wave_codes' path code
            // (1 + 2 path + partial), taken path by path
            run_path<kPathFar, true, PPT>(rec, c0, __ballot(nib == 1), part, hi, lo, since, jr0);
            run_path<kPathFar, false, PPT>(rec, c0, __ballot(nib == 2), part, hi, lo, since, jr0);
            run_path<kPathMid, true, PPT>(rec, c0, __ballot(nib == 3), part, hi, lo, since, jr0);
            run_path<kPathMid, false, PPT>(rec, c0, __ballot(nib == 4), part, hi, lo, since, jr0);
            run_path<kPathMid2, true, PPT>(rec, c0, __ballot(nib == 5), part, hi, lo, since, jr0);
            run_path<kPathMid2, false, PPT>(rec, c0, __ballot(nib == 6), part, hi, lo, since, jr0);
            run_path<kPathAny, true, PPT>(rec, c0, __ballot(nib == 7), part, hi, lo, since, jr0);
            run_path<kPathAny, false, PPT>(rec, c0, __ballot(nib == 8), part, hi, lo, since, jr0);
            if (__ballot(nib >= 14)) {
                if (edge_chunk<PPT>(rec, c0, nst, nib, part, jr0, wcen, whp))
                    fold_parts<PPT>(part, hi, lo, since);
                else
                    run_path<kPathFar, false, PPT>(rec, c0, __ballot(nib >= 14), part, hi, lo, since, jr0);
            }
            // wave-level far field: each lane evaluates its own record's line
            const unsigned long long wfm = __ballot(nib > 8 && nib < 14);
            if (wfm) {
                const Rec32 q = rec[min(c0 + lane, nst - 1)];
                if (__ballot(nib == 9)) { wave_far_nodes<kFN1>(q, nib == 9, wcen, whp, ws, lane); wused |= 1; }
                if (__ballot(nib == 10)) { wave_far_nodes<kFN2>(q, nib == 10, wcen, whp, ws + kWOff2, lane); wused |= 2; }
                if (__ballot(nib == 11)) { wave_far_nodes<kFN3>(q, nib == 11, wcen, whp, ws + kWOff3, lane); wused |= 4; }
            }
        }
    }
    // far field: node sums -> Chebyshev coefficients -> the tile's points
    // (totals and coefficients in the records' LDS, free after this barrier;
    // the next tile writes records only after its own barriers)
    static_assert(sizeof(FarCoefs) <= sizeof(Rec32) * kThreads, "far coefficients overlay");
    FarCoefs *fc = reinterpret_cast<FarCoefs *>(rec);
    __syncthreads();
    const int any = fl->any1 | (fl->any2 << 1) | (fl->any3 << 2);
    if (any) {
        // band 1 on threads 0..35, band 2 on 40..99, band 3 on 100..175
        band_sum<kFN1>(&fl->b1, &fc->b1, t);
        band_sum<kFN2>(&fl->b2, &fc->b2, t - 40);
        band_sum<kFN3>(&fl->b3, &fc->b3, t - 100);
        __syncthreads();
        band_coeffs<kFN1>(&fc->b1, t);
        band_coeffs<kFN2>(&fc->b2, t - 40);
        band_coeffs<kFN3>(&fc->b3, t - 100);
    }
    // wave-level coefficients, each wave its own (lanes 0..42), read by the
    // wave's lanes after the barrier
    if (ffok) {
        double *wc = fc->wc[wave];
        if (wused & 1) wave_far_coeffs<kFN1>(ws, wc, lane);
        if (wused & 2) wave_far_coeffs<kFN2>(ws + kWOff2, wc + kWOff2, lane - kWOff2);
        if (wused & 4) wave_far_coeffs<kFN3>(ws + kWOff3, wc + kWOff3, lane - kWOff3);
    }
    if (any || ffok) __syncthreads();
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        two_sum_acc(hi[k], lo[k], part[k]);
        const int64_t j = j0 + jr0 + 64 * k;
        double ff = 0.0;
        if (any) {
            // the interpolants at u = (jr - hW) / hW
            const double u = ((double)(jr0 + 64 * k) - hW) / hW;
            if (any & 1) ff += cheb_eval<kFN1>(fc->b1.c[wave], u);
            if (any & 2) ff += cheb_eval<kFN2>(fc->b2.c[wave], u);
            if (any & 4) ff += cheb_eval<kFN3>(fc->b3.c[wave], u);
        }
        if (wused) {
            // the wave's own interpolants, u over the wave's interval
            const double u = ((double)(jr0 + 64 * k) - (double)wcen) / (double)whp;
            const double *wc = fc->wc[wave];
            if (wused & 1) ff += cheb_eval<kFN1>(wc, u);
            if (wused & 2) ff += cheb_eval<kFN2>(wc + kWOff2, u);
            if (wused & 4) ff += cheb_eval<kFN3>(wc + kWOff3, u);
        }
        if (j < a.nx) {
            const double sum = ldexp(((double)hi[k] + (double)lo[k]) + ff, ti.e);
            a.K[j] = a.accumulate ? a.K[j] + sum : a.kmin + sum;
        }
    }
}

// ---------------------------------------------------------------------------
// Line-core pass.  Tile of kCoreW = 256 points; wave w owns the 64 points
// [64w, 64w + 64), lane l point 64w + l.  Per stage, each thread builds the
// record of one line of the tile's core range; the lines whose core meets
// the tile are compacted into LDS in line order (wave ballots), then each
// wave adds the records that overlap its points.  x and r2 are formed
// exactly as in the main pass, so r2 < R_core selects the complementary
// pairs.
// ---------------------------------------------------------------------------
constexpr int kCoreW = 256;

__device__ __forceinline__ void core_f64(const VArgs &a, int64_t tile, Core64 *rec, int *cnt) {
    constexpr int W = kCoreW;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wlo = wave * 64, whi = wlo + 64;
    const int jr = wlo + lane;
    const TileInfo ti = a.tiles[tile];
    const int64_t j0 = tile * W;
    const double xj = fma((double)(j0 + jr), a.dnu, a.numin);
    double acc = 0.0;
    bool touched = false;
    for (int64_t base = ti.lo; base < ti.hi; base += kStage) {
        const int64_t i = base + t;
        Core64 r;
        bool valid = false;
        if (i < ti.hi) {
            const double nu = a.nu[i];
            int64_t wl, wh;
            cut_window(nu, a.hw[i], a.numin, a.dnu, a.nx, wl, wh);
            const int64_t clo = max(wl, j0), chi = min(wh, j0 + W);
            const double y = a.y[i], y2 = y * y;
            if (clo < chi && y2 < kRB64) {
                const double ia = a.ia[i];
                const double jcd = rint((nu - a.numin) / a.dnu);
                const double rr = (fma(jcd, a.dnu, a.numin) - nu) * ia;
                int32_t clo3, chi3;
                region_interval(kRB64, y2, rr, a.dnu * ia, (double)j0 - jcd, W, clo3, chi3);
                r.lo = max((int32_t)(clo - j0), max(clo3, 0));
                r.hi = min((int32_t)(chi - j0), min(chi3, W));
                valid = r.lo < r.hi;
                if (valid) {
                    r.nu = nu; r.ia = ia; r.y = y; r.y2 = y2;
                    r.A = a.S[i] * kISqrtPi;
                    r.ey = a.ey[i];
                    r.q1 = core_q1_64(y);
                }
            }
        }
        // the previous stage's readers of rec[] and cnt[] are done after this
        __syncthreads();
        int n;
        const int slot = compact_slot(valid, cnt, n);
        if (valid) rec[slot] = r;
        __syncthreads();
        for (int l = 0; l < n; ++l) {
            const Core64 &q = rec[l];
            if ((q.hi <= wlo) | (q.lo >= whi)) continue;      // one LDS read for both
            touched = true;
            const double x = (xj - q.nu) * q.ia;
            const double r2 = fma(x, x, q.y2);
            if (jr >= q.lo && jr < q.hi && r2 < kRB64) acc += q.A * rew_core64(x, q.y, q.ey, q.q1);
        }
    }
    const int64_t j = j0 + jr;
    if (touched && j < a.nx) a.K[j] += acc;
}

__device__ __forceinline__ void core_f32c(const VArgs &a, int64_t tile, Core32 *rec, float *dtab,
                                          int *cnt) {
    constexpr int W = kCoreW;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wlo = wave * 64, whi = wlo + 64;
    const int jr = wlo + lane;
    const TileInfo ti = a.tiles[tile];
    const int64_t j0 = tile * W;
    float hi = 0.0f, lo = 0.0f, part = 0.0f;
    int since = 0;
    bool touched = false;
    for (int64_t base = ti.lo; base < ti.hi; base += kStage) {
        const int64_t i = base + t;
        Core32 r;
        bool valid = false;
        if (i < ti.hi) {
            const double nu = a.nu[i];
            int64_t wl, wh;
            cut_window(nu, a.hw[i], a.numin, a.dnu, a.nx, wl, wh);
            const int64_t clo = max(wl, j0), chi = min(wh, j0 + W);
            const double y = a.y[i];
            const float yf = (float)y;
            const double y2 = (double)yf * (double)yf;
            if (clo < chi && y2 < (double)kRB32) {
                const double ia = a.ia[i];
                const double jcd = rint((nu - a.numin) / a.dnu);
                const double rr = (fma(jcd, a.dnu, a.numin) - nu) * ia;
                const double vb = a.dnu * ia, d0 = (double)j0 - jcd;
                int32_t clo3, chi3;
                region_interval((double)kRB32, y2, rr, vb, d0, W, clo3, chi3);
                r.lo = max((int32_t)(clo - j0), max(clo3, 0));
                r.hi = min((int32_t)(chi - j0), min(chi3, W));
                valid = r.lo < r.hi;
                if (valid) {
                    const float vbh = (float)vb;
                    r.r = (float)rr;
                    r.vbh = vbh;
                    r.vbl = (float)(vb - (double)vbh);
                    r.y = yf;
                    r.y2 = yf * yf;
                    r.A = (float)(ldexp(a.S[i], -ti.e) * kISqrtPi);
                    r.ey = (float)a.ey[i];
                    r.q1 = core_q1_32(yf);
                    r.d0 = (int32_t)d0;
                }
            }
        }
        // the previous stage's readers of rec[], dtab[] and cnt[] are done after this
        __syncthreads();
        int n;
        const int slot = compact_slot(valid, cnt, n);
        if (valid) {
            rec[slot] = r;
            dtab[slot * kDT] = 0.0f;
            for (int k = 1; k <= kNT32; ++k) {
                const double an = kA32 * k;
                dtab[slot * kDT + k] = (float)(1.0 / fma(an, an, (double)r.y2));
            }
        }
        __syncthreads();
        // The stage's records (up to kStage, in four words of 64) that
        // reach this wave.  Unless one of its points lies in every such
        // record's interval (then each record needs a round anyway: wide
        // cores, or a stage of dense narrow ones), each lane marks the
        // records whose core interval holds its own point, and if the
        // busiest lane then holds under three quarters of them, the lanes
        // take their records in ascending order, one per round, so every
        // round evaluates a core for each lane that still has one (a
        // 64-point wave meets far more narrow cores than cover any one of
        // its points).  Otherwise the wave walks the records with uniform
        // loads.  A point sees its records in line order either way.
        static_assert(kStage == 256, "four mask words");
        unsigned long long mw[4], cov[4];
        int rlo[4], rhi[4], nrec = 0, maxlo = 0, minhi = 1 << 20;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int c0 = 64 * w, cn = min(64, n - c0);
            rlo[w] = 0;
            rhi[w] = 0;
            if (lane < cn) { rlo[w] = rec[c0 + lane].lo; rhi[w] = rec[c0 + lane].hi; }
            mw[w] = __ballot(lane < cn && rhi[w] > wlo && rlo[w] < whi);
            cov[w] = 0ull;
            nrec += __popcll(mw[w]);
            const bool in = (mw[w] >> lane) & 1ull;
            maxlo = max(maxlo, in ? max(rlo[w], wlo) : 0);
            minhi = min(minhi, in ? min(rhi[w], whi) : (1 << 20));
        }
        if (!nrec) continue;
        touched = true;
        maxlo = __builtin_amdgcn_readlane(wave_max_i32(maxlo), 63);
        minhi = (1 << 20) - __builtin_amdgcn_readlane(wave_max_i32((1 << 20) - minhi), 63);
        int rounds = nrec;
        if (maxlo >= minhi) {
            int mine = 0;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                unsigned long long m = mw[w];
                while (m) {
                    const int b = lowest_bit(m);
                    const int blo = __builtin_amdgcn_readlane(rlo[w], b);
                    const int bhi = __builtin_amdgcn_readlane(rhi[w], b);
                    if (jr >= blo && jr < bhi) cov[w] |= 1ull << b;
                }
                mine += __popcll(cov[w]);
            }
            rounds = __builtin_amdgcn_readlane(wave_max_i32(mine), 63);
        }
        auto eval_core = [&](int l) {
            const Core32 &q = rec[l];
            const float mm = (float)(jr + q.d0);
            const float x = fmaf(mm, q.vbh, fmaf(mm, q.vbl, q.r));
            const float r2 = fmaf(x, x, q.y2);
            if (jr >= q.lo && jr < q.hi && r2 < kRB32)
                part += q.A * rew_core32(x, q.y, q.ey, q.q1, dtab + l * kDT);
            if (++since == 32) { two_sum_acc(hi, lo, part); part = 0.0f; since = 0; }
        };
        if (4 * rounds >= 3 * nrec) {
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                unsigned long long m = mw[w];
                while (m) eval_core(64 * w + lowest_bit(m));
            }
            continue;
        }
        while (__ballot((cov[0] | cov[1] | cov[2] | cov[3]) != 0ull)) {
            int l = -1;
            if (cov[0]) { l = __builtin_ctzll(cov[0]); cov[0] &= cov[0] - 1ull; }
            else if (cov[1]) { l = 64 + __builtin_ctzll(cov[1]); cov[1] &= cov[1] - 1ull; }
            else if (cov[2]) { l = 128 + __builtin_ctzll(cov[2]); cov[2] &= cov[2] - 1ull; }
            else if (cov[3]) { l = 192 + __builtin_ctzll(cov[3]); cov[3] &= cov[3] - 1ull; }
            if (l >= 0) {
                const Core32 &q = rec[l];
                const float mm = (float)(jr + q.d0);
                const float x = fmaf(mm, q.vbh, fmaf(mm, q.vbl, q.r));
                const float r2 = fmaf(x, x, q.y2);
                if (r2 < kRB32) part += q.A * rew_core32(x, q.y, q.ey, q.q1, dtab + l * kDT);
            }
            if (++since == 32) { two_sum_acc(hi, lo, part); part = 0.0f; since = 0; }
        }
    }
    two_sum_acc(hi, lo, part);
    const int64_t j = j0 + jr;
    if (touched && j < a.nx) a.K[j] += ldexp((double)hi + (double)lo, ti.e);
}

// Tile b is workgroup b, so the round-robin dealing of workgroups to the 8
// XCDs interleaves the grid over them.  The work per tile grows with the
// wavenumber (Doppler widths scale with nu), and this interleaving keeps the
// XCDs balanced; giving each XCD a contiguous range instead (better L2 reuse
// of line data) measured 1.1x slower for the main pass and 2x for the
// line-core pass.
constexpr int main_waves(int mode, int ppt, int lm);

template <int MODE, int PPT, int LM>
__global__ __launch_bounds__(256, main_waves(MODE, PPT, LM)) void voigt_tiles_kernel(VArgs a) {
    const int64_t tile = blockIdx.x;
    if (MODE == HK_VOIGT_FP64) {
        __shared__ Rec64 rec[kThreads * LM];
        tile_f64<PPT, LM>(a, tile, rec);
    } else {
        __shared__ Rec32 rec[kThreads * LM];
        __shared__ FarLDS fl;
        tile_f32c<PPT, LM>(a, tile, rec, &fl);
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void voigt_core_kernel(VArgs a) {
    __shared__ int cnt[kThreads / 64];
    const int64_t tile = blockIdx.x;
    if (MODE == HK_VOIGT_FP64) {
        __shared__ Core64 rec[kStage];
        core_f64(a, tile, rec, cnt);
    } else {
        __shared__ Core32 rec[kStage];
        __shared__ float dtab[kStage * kDT];
        core_f32c(a, tile, rec, dtab, cnt);
    }
}

// Persistent form: workgroups take (spectrum, tile) items from a queue.
// item_start[s] = first item of spectrum s (nspec+1 entries).  PPT == 0
// selects the line-core pass.
// Occupancy floors (waves per SIMD): the FP32C main pass is latency-bound
// (LDS record reads, staging loads, a barrier per stage), so occupancy beats
// registers: 6 for tiles up to 1024 points with 256-line stages (25 KB of
// LDS per workgroup; the values that then spill are reloaded once per stage
// or per tile, outside the record loops: c2's batched sum 1070 ms at 4, 964
// at 5, 891 at 6), 3 with 512-line stages (LDS allows 3 workgroups per CU) and
// for 2048-point tiles (without a floor the far-field code raises the
// register count of the persistent form enough to halve its occupancy); the
// FP64 main pass 4; the FP32 line-core pass 5 (as its one-spectrum kernel has
// without a floor).
constexpr int main_waves(int mode, int ppt, int lm) {
    return ppt <= 4 ? (lm == 2 ? 3 : (mode == HK_VOIGT_FP32C ? 6 : 4)) : (ppt == 8 ? 3 : 1);
}

template <int MODE, int PPT, int LM>
__global__ __launch_bounds__(256, PPT == 0 ? (MODE == HK_VOIGT_FP32C ? 5 : 1) : main_waves(MODE, PPT, LM))
void voigt_persistent_kernel(const VArgs *__restrict__ specs,
                                                               const int64_t *__restrict__ item_start,
                                                               int nspec, unsigned long long *counter) {
    constexpr size_t kRecBytes =
        PPT == 0 ? (MODE == HK_VOIGT_FP64 ? sizeof(Core64) : sizeof(Core32))
                 : (MODE == HK_VOIGT_FP64 ? sizeof(Rec64) : sizeof(Rec32));
    __shared__ int64_t s_item;
    __shared__ int cnt[kThreads / 64];
    __shared__ float dtab[(PPT == 0 && MODE == HK_VOIGT_FP32C) ? kStage * kDT : 1];
    __shared__ __attribute__((aligned(16))) char smem[kRecBytes * kThreads * LM];
    __shared__ FarLDS fl[1];
    const int64_t total = item_start[nspec];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(counter, 1ull);
        __syncthreads();
        // uniform across the workgroup: keep the item, spectrum and its
        // arguments in scalar registers
        const int64_t iv = s_item;
        const int64_t item =
            (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(iv >> 32)) << 32) |
                      (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)iv));
        if (item >= total) break;
        const int s = __builtin_amdgcn_readfirstlane(item_spectrum(item_start, nspec, item));
        const VArgs a = specs[s];
        const int64_t tile = item - item_start[s];
        if constexpr (PPT == 0) {
            if constexpr (MODE == HK_VOIGT_FP64) core_f64(a, tile, (Core64 *)smem, cnt);
            else core_f32c(a, tile, (Core32 *)smem, dtab, cnt);
        } else {
            if constexpr (MODE == HK_VOIGT_FP64) tile_f64<PPT, LM>(a, tile, (Rec64 *)smem);
            else tile_f32c<PPT, LM>(a, tile, (Rec32 *)smem, fl);
        }
    }
}

__global__ void reset_counters_kernel(unsigned long long *c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0ull;
}

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

// Work of one FP32 Voigt sum with the far field at tile width W: per line,
// the tiles its window touches, split into far tiles (pairs summed through
// the far field, node evaluations) and direct pairs by region.  The band
// decision is the kernel's (far_band on the same FP64 values).  out: direct
// A, B, C pairs, far-field pairs, far (line, tile) counts of band 1 (9
// nodes), band 2 (15 nodes) and band 3 (19 nodes), wave-level far (line,
// wave) counts of bands 1-3, and the cut-edge remainders the kernel may take
// through its per-wave expansion: (line, wave) count and their pairs (the
// kernel takes them record by record in chunks whose window ends are out of
// lane order or hold both kinds; counted here as expanded).
__global__ __launch_bounds__(256) void work_count_kernel(const double *__restrict__ nu,
                                                         const double *__restrict__ hw,
                                                         const double *__restrict__ ia,
                                                         const double *__restrict__ y, int64_t nl,
                                                         double numin, double dnu, int64_t nx,
                                                         int ppt, int far,
                                                         unsigned long long *__restrict__ out) {
    constexpr int kOut = 12;
    __shared__ unsigned long long red[kOut][256];
    const int t = threadIdx.x;
    const int W = kThreads * ppt, WP = 64 * ppt;
    const int64_t i = (int64_t)blockIdx.x * 256 + t;
    // direct A, B, C; far-field pairs; far (line, tile) pairs of bands 1-3;
    // far (line, wave) pairs of the wave-level bands 1-3
    unsigned long long c[kOut] = {};
    if (i < nl) {
        int64_t wl, wh, bl = 0, bh = 0, cl = 0, ch = 0;
        cut_window(nu[i], hw[i], numin, dnu, nx, wl, wh);
        const double yy = y[i], y2 = yy * yy, a = ia[i];
        if (wh > wl && a > 0.0) {
            if (y2 < 1.0e6) cut_window(nu[i], fmin(hw[i], sqrt(1.0e6 - y2) / a), numin, dnu, nx, bl, bh);
            if (y2 < 100.0) cut_window(nu[i], fmin(hw[i], sqrt(100.0 - y2) / a), numin, dnu, nx, cl, ch);
            auto ov = [](int64_t l0, int64_t h0, int64_t l1, int64_t h1) {
                const int64_t l = max(l0, l1), h = min(h0, h1);
                return (unsigned long long)(h > l ? h - l : 0);
            };
            for (int64_t tl = wl / W; tl <= (wh - 1) / W; ++tl) {
                const int64_t j0 = tl * W, Wv = min((int64_t)W, nx - j0);
                const double hW = 0.5 * (double)(Wv - 1);
                bool full = false;
                int band = far ? far_band(nu[i], hw[i], a, yy, j0, Wv, hW, numin, dnu, full) : 0;
                const bool edge1 = band == 1 && !full;
                // the pairs left to the pair-by-pair path: [dlo, dhi)
                int32_t dlo = (int32_t)(max(wl, j0) - j0), dhi = (int32_t)(min(wh, j0 + Wv) - j0);
                const int32_t n0 = dhi - dlo;
                if (band && !full && (band != 1 || !far_cover(dlo, dhi, (int)Wv, WP))) band = 0;
                if (band) {
                    c[3] += (unsigned long long)(full ? Wv : n0 - (dhi - dlo));
                    c[3 + band] += 1;
                    if (full) continue;
                }
                // waves that take the line from their own far field (the
                // kernel's test, on the record's FP32 y)
                int32_t wf = 0, fr0 = 0, fr1 = 0;
                double fd0 = 0.0;
                if (far && Wv >= kFarMinPts && dlo < dhi) {
                    double rr, vb, d0;
                    line_offsets(nu[i], a, numin, dnu, j0, rr, vb, d0);
                    fd0 = d0;
                    const double yf = (double)(float)yy, y2f = yf * yf;
                    // points with x^2 + y^2 < kRF32 (the far path's bound)
                    region_interval((double)kRF32, y2f, rr, vb, d0, W, fr0, fr1);
                    switch (ppt) {
                        case 2: wf = wave_far_codes<128>(dlo, dhi, Wv, d0, vb, rr, y2f); break;
                        case 4: wf = wave_far_codes<256>(dlo, dhi, Wv, d0, vb, rr, y2f); break;
                        case 8: wf = wave_far_codes<512>(dlo, dhi, Wv, d0, vb, rr, y2f); break;
                        default: wf = wave_far_codes<1024>(dlo, dhi, Wv, d0, vb, rr, y2f); break;
                    }
                }
                for (int w = 0; w < kWaves; ++w) {
                    const int64_t p0 = j0 + max(dlo, w * WP), p1 = j0 + min(dhi, (w + 1) * WP);
                    if (p1 <= p0) continue;
                    const int wb = (wf >> (4 * w)) & 15;
                    if (wb) {
                        c[3] += (unsigned long long)(p1 - p0);
                        c[7 + wb - 9] += 1;
                        continue;
                    }
                    // a cut-edge remainder in one wave, all on the far path,
                    // reaching the wave's end (right) or start (left)
                    if (edge1 && (dhi - 1) / WP == dlo / WP && misses(fr0, fr1, w * WP, (w + 1) * WP) &&
                        fabs((double)w * WP + 0.5 * (WP - 1) + fd0) >= kEdgeMin * 0.5 * (WP - 1)) {
                        const int ea = w * WP, eb = (int)min((int64_t)(ea + WP), Wv);
                        if (dhi >= eb || dlo <= ea) {
                            c[10] += 1;
                            c[11] += (unsigned long long)(p1 - p0);
                            continue;
                        }
                    }
                    const unsigned long long tot = ov(wl, wh, p0, p1);
                    const unsigned long long bc = ov(bl, bh, p0, p1);
                    const unsigned long long cc = ov(cl, ch, p0, p1);
                    c[0] += tot - bc;
                    c[1] += bc - cc;
                    c[2] += cc;
                }
            }
        }
    }
    for (int q = 0; q < kOut; ++q) red[q][t] = c[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < kOut; ++q) red[q][t] += red[q][t + s];
        __syncthreads();
    }
    if (t < kOut) atomicAdd(&out[t], red[t][0]);
}

__global__ void zero12_kernel(unsigned long long *c) {
    if (threadIdx.x < 12) c[threadIdx.x] = 0ull;
}

}  // namespace hk

using namespace hk;

namespace {

int pick_ppt(int mode, const hk_tune *tune) {
    // defaults: 1024-point tiles for both forms (with the FP32 far field,
    // smaller tiles trade direct pairs for node evaluations; 1024 was
    // fastest at 1e7 lines x 3e7 points)
    int tp = tune && tune->tile_points > 0 ? tune->tile_points : 1024;
    return tp / kThreads;
}

bool ppt_ok(int ppt) { return ppt == 2 || ppt == 4 || ppt == 8 || ppt == 16; }

// Lines per LDS stage of the main pass: 256 (default) or 512; returns the
// multiple of kThreads, 0 when the request is not one of those.
int pick_lm(const hk_tune *tune) {
    const int sl = tune && tune->stage_lines > 0 ? tune->stage_lines : kThreads;
    return sl == kThreads ? 1 : (sl == 2 * kThreads ? 2 : 0);
}

template <int MODE, int LM>
void launch_tiles_lm(int ppt, int64_t ntiles, hipStream_t s, const VArgs &a) {
    dim3 g((unsigned)ntiles), b(kThreads);
    switch (ppt) {
        case 2: hipLaunchKernelGGL((voigt_tiles_kernel<MODE, 2, LM>), g, b, 0, s, a); break;
        case 4: hipLaunchKernelGGL((voigt_tiles_kernel<MODE, 4, LM>), g, b, 0, s, a); break;
        case 8: hipLaunchKernelGGL((voigt_tiles_kernel<MODE, 8, LM>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((voigt_tiles_kernel<MODE, 16, LM>), g, b, 0, s, a); break;
    }
}

template <int MODE>
void launch_tiles(int ppt, int lm, int64_t ntiles, hipStream_t s, const VArgs &a) {
    if (lm == 2) launch_tiles_lm<MODE, 2>(ppt, ntiles, s, a);
    else launch_tiles_lm<MODE, 1>(ppt, ntiles, s, a);
}

template <int MODE, int LM>
void launch_persistent_lm(int ppt, unsigned grid, hipStream_t s, const VArgs *specs,
                          const int64_t *item_start, int nspec, unsigned long long *counter) {
    dim3 g(grid), b(kThreads);
    switch (ppt) {
        case 2: hipLaunchKernelGGL((voigt_persistent_kernel<MODE, 2, LM>), g, b, 0, s, specs, item_start, nspec, counter); break;
        case 4: hipLaunchKernelGGL((voigt_persistent_kernel<MODE, 4, LM>), g, b, 0, s, specs, item_start, nspec, counter); break;
        case 8: hipLaunchKernelGGL((voigt_persistent_kernel<MODE, 8, LM>), g, b, 0, s, specs, item_start, nspec, counter); break;
        default: hipLaunchKernelGGL((voigt_persistent_kernel<MODE, 16, LM>), g, b, 0, s, specs, item_start, nspec, counter); break;
    }
}

// ppt == 0: the line-core pass (its stages are kThreads lines)
template <int MODE>
void launch_persistent(int ppt, int lm, unsigned grid, hipStream_t s, const VArgs *specs,
                       const int64_t *item_start, int nspec, unsigned long long *counter) {
    if (ppt == 0)
        hipLaunchKernelGGL((voigt_persistent_kernel<MODE, 0, 1>), dim3(grid), dim3(kThreads), 0, s,
                           specs, item_start, nspec, counter);
    else if (lm == 2)
        launch_persistent_lm<MODE, 2>(ppt, grid, s, specs, item_start, nspec, counter);
    else
        launch_persistent_lm<MODE, 1>(ppt, grid, s, specs, item_start, nspec, counter);
}

int check_state(const hk_line_state *st, const char *fn) {
    HK_CHECK_ARG(st->nlines == 0 || (st->nu && st->S && st->ialphaD && st->y && st->hw && st->ey &&
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

// Workspace: the main pass's tile table (tiles of >= 512 points) followed
// by the line-core pass's (tiles of kCoreW points).
extern "C" size_t hk_voigt_workspace_bytes(const hk_grid *grid, const hk_tune *tune) {
    if (!grid || grid->nx <= 0) return 0;
    (void)tune;
    const int64_t ncore = (grid->nx + kCoreW - 1) / kCoreW;
    return align_up((size_t)ncore * sizeof(TileInfo), 256) * 2 + 256;
}

extern "C" int hk_voigt_accumulate(const hk_line_state *st, const hk_grid *grid, double kmin,
                                   int accumulate, int mode, const hk_tune *tune, double *K,
                                   void *ws, size_t ws_bytes, hk_stream_t stream) {
    HK_CHECK_ARG(st && grid && K && ws, "hk_voigt_accumulate: null argument");
    HK_CHECK_ARG(mode == HK_VOIGT_FP64 || mode == HK_VOIGT_FP32C || mode == HK_VOIGT_COMPAT,
                 "hk_voigt_accumulate: mode %d", mode);
    if (int e = check_state(st, "hk_voigt_accumulate")) return e;
    if (int e = check_grid(grid, "hk_voigt_accumulate")) return e;
    const int ppt = pick_ppt(mode, tune);
    HK_CHECK_ARG(ppt_ok(ppt), "hk_voigt_accumulate: tile_points must be 512, 1024, 2048 or 4096");
    const int lm = pick_lm(tune);
    HK_CHECK_ARG(lm > 0, "hk_voigt_accumulate: stage_lines must be 256 or 512");
    if (ws_bytes < hk_voigt_workspace_bytes(grid, tune)) {
        set_error("hk_voigt_accumulate: workspace %zu bytes, %zu needed", ws_bytes,
                  hk_voigt_workspace_bytes(grid, tune));
        return HK_ERR_WORKSPACE;
    }
    if (mode == HK_VOIGT_COMPAT)
        return compat_accumulate(st, grid, kmin, accumulate, K, ws, (hipStream_t)stream);
    const int W = kThreads * ppt;
    const int64_t ntiles = (grid->nx + W - 1) / W;
    const int64_t ncore = (grid->nx + kCoreW - 1) / kCoreW;
    TileInfo *tiles = (TileInfo *)align_up((size_t)ws, 256);
    TileInfo *ctiles = (TileInfo *)((char *)tiles + align_up((size_t)ncore * sizeof(TileInfo), 256));
    const int scale = mode == HK_VOIGT_FP32C ? 1 : 0;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, s,
                       st->key, st->nlines, st->radius, st->blk_max, grid->numin, grid->dnu,
                       grid->nx, W, ntiles, scale, tiles);
    HK_LAUNCH_CHECK();
    hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((ncore + 3) / 4)), dim3(256), 0, s,
                       st->key, st->nlines, st->radius + 1, st->blk_max, grid->numin, grid->dnu,
                       grid->nx, kCoreW, ncore, scale, ctiles);
    HK_LAUNCH_CHECK();
    VArgs a;
    a.nu = st->nu; a.S = st->S; a.ia = st->ialphaD; a.y = st->y; a.hw = st->hw; a.ey = st->ey;
    a.key = st->key; a.blk_max = st->blk_max; a.radius = st->radius; a.nl = st->nlines;
    a.tiles = tiles; a.numin = grid->numin; a.dnu = grid->dnu; a.nx = grid->nx;
    a.kmin = kmin; a.accumulate = accumulate; a.K = K;
    a.far = (mode == HK_VOIGT_FP32C && !(tune && tune->exact_pairs)) ? 1 : 0;
    if (mode == HK_VOIGT_FP64) launch_tiles<HK_VOIGT_FP64>(ppt, lm, ntiles, s, a);
    else launch_tiles<HK_VOIGT_FP32C>(ppt, lm, ntiles, s, a);
    HK_LAUNCH_CHECK();
    a.tiles = ctiles;
    if (mode == HK_VOIGT_FP64)
        hipLaunchKernelGGL(voigt_core_kernel<HK_VOIGT_FP64>, dim3((unsigned)ncore), dim3(kThreads),
                           0, s, a);
    else
        hipLaunchKernelGGL(voigt_core_kernel<HK_VOIGT_FP32C>, dim3((unsigned)ncore), dim3(kThreads),
                           0, s, a);
    HK_LAUNCH_CHECK();
    return HK_OK;
}

// ---------------------------------------------------------------------------
// No-allocation form: the tile tables in a static device array, one per GPU.
// The array's name is a host-side handle only; its device address is looked
// up once per device (hipGetSymbolAddress, with that device current) and that
// address is what the kernels receive.  The tile tables are rewritten in full
// by every call, so nothing in the array needs clearing between calls.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[HK_VOIGT_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(void **out) {
    int dev = 0;
    HK_HIP(hipGetDevice(&dev));
    HK_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "hk_voigt_accumulate_noalloc: device %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HK_CHECK_ARG(p != nullptr, "hk_voigt_accumulate_noalloc: no device address for scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return HK_OK;
}
}  // namespace

extern "C" int hk_voigt_accumulate_noalloc(const hk_line_state *st, const hk_grid *grid,
                                           double kmin, int accumulate, int mode,
                                           const hk_tune *tune, double *K, hk_stream_t stream) {
    HK_CHECK_ARG(st && grid && K, "hk_voigt_accumulate_noalloc: null argument");
    HK_CHECK_ARG(grid->nx <= HK_VOIGT_NOALLOC_MAX_NX,
                 "hk_voigt_accumulate_noalloc: nx %lld exceeds the static scratch limit of %lld "
                 "points (use hk_voigt_accumulate with a workspace)",
                 (long long)grid->nx, (long long)HK_VOIGT_NOALLOC_MAX_NX);
    const size_t need = hk_voigt_workspace_bytes(grid, tune);
    HK_CHECK_ARG(need <= HK_VOIGT_NOALLOC_SCRATCH_BYTES,
                 "hk_voigt_accumulate_noalloc: %zu scratch bytes needed, %zu available", need,
                 (size_t)HK_VOIGT_NOALLOC_SCRATCH_BYTES);
    void *ws = nullptr;
    if (int e = scratch_for_current_device(&ws)) return e;
    return hk_voigt_accumulate(st, grid, kmin, accumulate, mode, tune, K, ws,
                               HK_VOIGT_NOALLOC_SCRATCH_BYTES, stream);
}

// ---------------------------------------------------------------------------
// Batched: workspace = [VArgs x nspec, main pass | VArgs x nspec, core pass |
//   item_start x (nspec+1), main | item_start x (nspec+1), core | 2 queue
//   counters | per spectrum: main tile table, core tile table]
// hk_voigt_batched_prepare writes the tables (one synchronous upload);
// hk_voigt_batched_launch then issues kernels only, so a caller can capture
// it in a graph and replay it.
// ---------------------------------------------------------------------------
namespace {
struct BatchLayout {
    size_t off_args, off_cargs, off_items, off_citems, off_counter, total;
    std::vector<size_t> tile_off, ctile_off;
    std::vector<int64_t> ntiles, ncore;
};

BatchLayout batch_layout(int nspec, const hk_spectrum *spec, int W) {
    BatchLayout L;
    size_t o = 0;
    L.off_args = o; o = align_up(o + sizeof(VArgs) * nspec, 256);
    L.off_cargs = o; o = align_up(o + sizeof(VArgs) * nspec, 256);
    L.off_items = o; o = align_up(o + sizeof(int64_t) * (nspec + 1), 256);
    L.off_citems = o; o = align_up(o + sizeof(int64_t) * (nspec + 1), 256);
    L.off_counter = o; o = align_up(o + 64, 256);
    for (int i = 0; i < nspec; ++i) {
        const int64_t nt = (spec[i].grid.nx + W - 1) / W;
        const int64_t nc = (spec[i].grid.nx + kCoreW - 1) / kCoreW;
        L.ntiles.push_back(nt);
        L.ncore.push_back(nc);
        L.tile_off.push_back(o);
        o = align_up(o + sizeof(TileInfo) * nt, 256);
        L.ctile_off.push_back(o);
        o = align_up(o + sizeof(TileInfo) * nc, 256);
    }
    L.total = o + 256;
    return L;
}

unsigned persistent_grid(int64_t items) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    return (unsigned)std::min<int64_t>(items, (int64_t)ncu * 8);
}
}  // namespace

extern "C" size_t hk_voigt_batched_workspace_bytes(int nspec, const hk_spectrum *spec,
                                                   const hk_tune *tune) {
    if (nspec <= 0 || !spec) return 0;
    (void)tune;
    return batch_layout(nspec, spec, 512).total;   // smallest main tile: largest tables
}

extern "C" int hk_voigt_batched_prepare(int nspec, const hk_spectrum *spec, int mode,
                                        const hk_tune *tune, void *ws, size_t ws_bytes,
                                        hk_voigt_batch *batch, hk_stream_t stream) {
    HK_CHECK_ARG(nspec > 0 && spec && ws && batch, "hk_voigt_batched_prepare: bad arguments");
    HK_CHECK_ARG(mode == HK_VOIGT_FP64 || mode == HK_VOIGT_FP32C,
                 "hk_voigt_batched_prepare: mode %d", mode);
    const int ppt = pick_ppt(mode, tune);
    HK_CHECK_ARG(ppt_ok(ppt), "hk_voigt_batched_prepare: tile_points must be 512, 1024, 2048 or 4096");
    const int lm = pick_lm(tune);
    HK_CHECK_ARG(lm > 0, "hk_voigt_batched_prepare: stage_lines must be 256 or 512");
    for (int i = 0; i < nspec; ++i) {
        if (int e = check_state(&spec[i].st, "hk_voigt_batched_prepare")) return e;
        if (int e = check_grid(&spec[i].grid, "hk_voigt_batched_prepare")) return e;
        HK_CHECK_ARG(spec[i].K, "hk_voigt_batched_prepare: spectrum %d has no K", i);
    }
    if (ws_bytes < hk_voigt_batched_workspace_bytes(nspec, spec, tune)) {
        set_error("hk_voigt_batched_prepare: workspace %zu bytes, %zu needed", ws_bytes,
                  hk_voigt_batched_workspace_bytes(nspec, spec, tune));
        return HK_ERR_WORKSPACE;
    }
    const int W = kThreads * ppt;
    char *base = (char *)align_up((size_t)ws, 256);
    BatchLayout L = batch_layout(nspec, spec, W);
    // Spectra ordered by line count (largest first) so the heaviest tiles are
    // dequeued first.
    std::vector<int> order(nspec);
    for (int i = 0; i < nspec; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return spec[x].st.nlines > spec[y].st.nlines;
    });
    std::vector<VArgs> args(nspec), cargs(nspec);
    std::vector<int64_t> items(nspec + 1, 0), citems(nspec + 1, 0);
    for (int q = 0; q < nspec; ++q) {
        const int i = order[q];
        const hk_spectrum &sp = spec[i];
        VArgs &a = args[q];
        a.nu = sp.st.nu; a.S = sp.st.S; a.ia = sp.st.ialphaD; a.y = sp.st.y; a.hw = sp.st.hw;
        a.ey = sp.st.ey;
        a.key = sp.st.key; a.blk_max = sp.st.blk_max; a.radius = sp.st.radius;
        a.nl = sp.st.nlines;
        a.tiles = (TileInfo *)(base + L.tile_off[i]);
        a.numin = sp.grid.numin; a.dnu = sp.grid.dnu; a.nx = sp.grid.nx;
        a.kmin = sp.kmin; a.accumulate = sp.accumulate; a.K = sp.K;
        a.far = (mode == HK_VOIGT_FP32C && !(tune && tune->exact_pairs)) ? 1 : 0;
        cargs[q] = a;
        cargs[q].radius = sp.st.radius + 1;
        cargs[q].tiles = (TileInfo *)(base + L.ctile_off[i]);
        items[q + 1] = items[q] + L.ntiles[i];
        citems[q + 1] = citems[q] + L.ncore[i];
    }
    hipStream_t s = (hipStream_t)stream;
    HK_HIP(hipMemcpyAsync(base + L.off_args, args.data(), sizeof(VArgs) * nspec,
                          hipMemcpyHostToDevice, s));
    HK_HIP(hipMemcpyAsync(base + L.off_cargs, cargs.data(), sizeof(VArgs) * nspec,
                          hipMemcpyHostToDevice, s));
    HK_HIP(hipMemcpyAsync(base + L.off_items, items.data(), sizeof(int64_t) * (nspec + 1),
                          hipMemcpyHostToDevice, s));
    HK_HIP(hipMemcpyAsync(base + L.off_citems, citems.data(), sizeof(int64_t) * (nspec + 1),
                          hipMemcpyHostToDevice, s));
    // the host tables must outlive the asynchronous copies
    HK_HIP(hipStreamSynchronize(s));
    batch->workspace = base;
    batch->nspec = nspec;
    batch->mode = mode;
    batch->tile_points = W;
    batch->stage_lines = kThreads * lm;
    batch->items_main = items[nspec];
    batch->items_core = citems[nspec];
    batch->off_args = L.off_args; batch->off_cargs = L.off_cargs;
    batch->off_items = L.off_items; batch->off_citems = L.off_citems;
    batch->off_counter = L.off_counter;
    return HK_OK;
}

extern "C" int hk_voigt_batched_launch(const hk_voigt_batch *batch, hk_stream_t stream) {
    HK_CHECK_ARG(batch && batch->workspace && batch->nspec > 0,
                 "hk_voigt_batched_launch: batch not prepared");
    const int ppt = batch->tile_points / kThreads;
    const int lm = batch->stage_lines / kThreads;
    HK_CHECK_ARG(ppt_ok(ppt) && (lm == 1 || lm == 2) &&
                     (batch->mode == HK_VOIGT_FP64 || batch->mode == HK_VOIGT_FP32C),
                 "hk_voigt_batched_launch: batch not prepared");
    char *base = (char *)batch->workspace;
    const VArgs *args = (const VArgs *)(base + batch->off_args);
    const VArgs *cargs = (const VArgs *)(base + batch->off_cargs);
    const int64_t *items = (const int64_t *)(base + batch->off_items);
    const int64_t *citems = (const int64_t *)(base + batch->off_citems);
    unsigned long long *counter = (unsigned long long *)(base + batch->off_counter);
    const int nspec = batch->nspec;
    const int scale = batch->mode == HK_VOIGT_FP32C ? 1 : 0;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, counter);
    HK_LAUNCH_CHECK();
    if (batch->items_main > 0) {
        hipLaunchKernelGGL(batched_ranges_kernel, dim3((unsigned)((batch->items_main + 3) / 4)),
                           dim3(256), 0, s, args, items, nspec, batch->tile_points, scale);
        HK_LAUNCH_CHECK();
        hipLaunchKernelGGL(batched_ranges_kernel, dim3((unsigned)((batch->items_core + 3) / 4)),
                           dim3(256), 0, s, cargs, citems, nspec, kCoreW, scale);
        HK_LAUNCH_CHECK();
        const unsigned gm = persistent_grid(batch->items_main);
        const unsigned gc = persistent_grid(batch->items_core);
        if (batch->mode == HK_VOIGT_FP64) {
            launch_persistent<HK_VOIGT_FP64>(ppt, lm, gm, s, args, items, nspec, counter);
            HK_LAUNCH_CHECK();
            launch_persistent<HK_VOIGT_FP64>(0, 1, gc, s, cargs, citems, nspec, counter + 1);
        } else {
            launch_persistent<HK_VOIGT_FP32C>(ppt, lm, gm, s, args, items, nspec, counter);
            HK_LAUNCH_CHECK();
            launch_persistent<HK_VOIGT_FP32C>(0, 1, gc, s, cargs, citems, nspec, counter + 1);
        }
        HK_LAUNCH_CHECK();
    }
    return HK_OK;
}

extern "C" int hk_voigt_batched(int nspec, const hk_spectrum *spec, int mode, const hk_tune *tune,
                                void *ws, size_t ws_bytes, hk_stream_t stream) {
    hk_voigt_batch b;
    if (int e = hk_voigt_batched_prepare(nspec, spec, mode, tune, ws, ws_bytes, &b, stream))
        return e;
    return hk_voigt_batched_launch(&b, stream);
}

// Tuning cache: the choice of the first tuning of a problem is kept for the
// process, so reruns of the same problem take the same tile width (and give
// the same bits).  The signature: form, far-field setting, grid, line count,
// and the line population's cut reach and median y (y of 64 lines spread
// over the list), both rounded to 3 significant digits.
namespace {
struct TuneKey {
    int mode, exact;
    int64_t nx, nl;
    double numin, dnu, reach, ymed;
    bool operator<(const TuneKey &o) const {
        return std::tie(mode, exact, nx, nl, numin, dnu, reach, ymed) <
               std::tie(o.mode, o.exact, o.nx, o.nl, o.numin, o.dnu, o.reach, o.ymed);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tuned;     // -> (tile points, stage lines)

double round3(double v) {
    if (!(v != 0.0) || !std::isfinite(v)) return v;
    const double e = std::pow(10.0, std::floor(std::log10(std::fabs(v))) - 2.0);
    return std::round(v / e) * e;
}

constexpr int kTuneSamples = 64;

__global__ void tune_sample_kernel(const double *__restrict__ y, int64_t nl,
                                   const double *__restrict__ radius, double *__restrict__ out) {
    const int t = threadIdx.x;
    if (t < kTuneSamples) out[t] = nl > 0 ? y[(nl - 1) * t / (kTuneSamples - 1)] : 0.0;
    if (t == 0) out[kTuneSamples] = radius[0];
}
}  // namespace

extern "C" void hk_voigt_autotune_reset(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
}

extern "C" int hk_voigt_autotune(const hk_line_state *st, const hk_grid *grid, int mode,
                                 double *K, void *ws, size_t ws_bytes, hk_tune *best, float *ms,
                                 hk_stream_t stream) {
    HK_CHECK_ARG(st && grid && K && ws && best, "hk_voigt_autotune: null argument");
    if (int e = check_state(st, "hk_voigt_autotune")) return e;
    if (int e = check_grid(grid, "hk_voigt_autotune")) return e;
    constexpr int kCand[4] = {512, 1024, 2048, 4096};
    constexpr int kStageCand[2] = {256, 512};
    hipStream_t s = (hipStream_t)stream;
    const int exact = best->exact_pairs;     // tuned for the caller's setting
    // signature (K serves as the scratch of the samples: it is overwritten anyway)
    TuneKey key = {mode, exact, grid->nx, st->nlines, grid->numin, grid->dnu, 0.0, 0.0};
    const bool cacheable = grid->nx > kTuneSamples;
    if (cacheable) {
        hipLaunchKernelGGL(tune_sample_kernel, dim3(1), dim3(kTuneSamples), 0, s, st->y,
                           st->nlines, st->radius, K);
        HK_LAUNCH_CHECK();
        double h[kTuneSamples + 1];
        HK_HIP(hipMemcpyAsync(h, K, sizeof(h), hipMemcpyDeviceToHost, s));
        HK_HIP(hipStreamSynchronize(s));
        std::sort(h, h + kTuneSamples);
        key.ymed = round3(0.5 * (h[kTuneSamples / 2 - 1] + h[kTuneSamples / 2]));
        key.reach = round3(h[kTuneSamples]);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            best->tile_points = it->second.first;
            best->stage_lines = it->second.second;
            if (ms) for (int c = 0; c < 8; ++c) ms[c] = 0.0f;
            return HK_OK;
        }
    }
    hipEvent_t e0, e1;
    HK_HIP(hipEventCreate(&e0));
    HK_HIP(hipEventCreate(&e1));
    float bestms = 0.0f;
    int bestw = 0, bests = 0, rc = HK_OK;
    for (int c = 0; c < 8 && rc == HK_OK; ++c) {
        hk_tune t = {kCand[c % 4], exact, kStageCand[c / 4]};
        rc = hk_voigt_accumulate(st, grid, 0.0, 0, mode, &t, K, ws, ws_bytes, stream);
        if (rc != HK_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = HK_ERR_HIP; break; }
        rc = hk_voigt_accumulate(st, grid, 0.0, 0, mode, &t, K, ws, ws_bytes, stream);
        if (rc != HK_OK) break;
        float dt = 0.0f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&dt, e0, e1) != hipSuccess) {
            set_error("hk_voigt_autotune: event timing failed");
            rc = HK_ERR_HIP;
            break;
        }
        if (ms) ms[c] = dt;
        if (bestw == 0 || dt < bestms) { bestms = dt; bestw = t.tile_points; bests = t.stage_lines; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == HK_OK) {
        best->tile_points = bestw;
        best->stage_lines = bests;
        if (cacheable) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned.emplace(key, std::make_pair(bestw, bests));
        }
    }
    return rc;
}

extern "C" int hk_voigt_work_count(const hk_line_state *st, const hk_grid *grid, int mode,
                                   const hk_tune *tune, unsigned long long *counts,
                                   hk_stream_t stream) {
    HK_CHECK_ARG(st && grid && counts, "hk_voigt_work_count: null argument");
    HK_CHECK_ARG(mode == HK_VOIGT_FP64 || mode == HK_VOIGT_FP32C, "hk_voigt_work_count: mode %d",
                 mode);
    if (int e = check_state(st, "hk_voigt_work_count")) return e;
    if (int e = check_grid(grid, "hk_voigt_work_count")) return e;
    const int ppt = pick_ppt(mode, tune);
    HK_CHECK_ARG(ppt_ok(ppt), "hk_voigt_work_count: tile_points must be 512, 1024, 2048 or 4096");
    const int far = (mode == HK_VOIGT_FP32C && !(tune && tune->exact_pairs)) ? 1 : 0;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(zero12_kernel, dim3(1), dim3(64), 0, s, counts);
    HK_LAUNCH_CHECK();
    if (st->nlines > 0) {
        hipLaunchKernelGGL(work_count_kernel, dim3((unsigned)((st->nlines + 255) / 256)), dim3(256),
                           0, s, st->nu, st->hw, st->ialphaD, st->y, st->nlines, grid->numin,
                           grid->dnu, grid->nx, ppt, far, counts);
        HK_LAUNCH_CHECK();
    }
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
