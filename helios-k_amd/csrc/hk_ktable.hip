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
// Here the sort and the fit read K once:
//   * a bin of at most 16384 points is loaded into LDS by one workgroup,
//     sorted there (block merge sort, hk_lds_sort.h) and fitted from LDS in
//     the same kernel: Nxmin (a binary search: findCut's rules on a sorted
//     bin), the rescale, ln(K unit_scale), the nC Chebyshev moments
//     m_i = sum_k T_i(x_k) b_k and the first nC values b0 of b = ln K2;
//   * a longer bin is cut into LDS-sorted runs of 16384 points (written
//     once), a splitter kernel per bin picks value splitters from a regular
//     sample of the runs (every m-th value of each run), and one workgroup
//     per bucket [s_w, s_w+1) gathers the bucket's values from every run,
//     sorts them in LDS and fits the output points whose rescale nodes fall
//     in the bucket's ranks; per-bucket partial moments are summed in
//     bucket order.  With regular sampling a bucket holds at most
//     m (g - 1 + runs) values above its lower splitter (g samples per
//     bucket), so it always fits in LDS; the values equal to the lower
//     splitter are counted, not loaded.  Bins of more than 16 runs are
//     first merged pairwise (merge_lds_kernel) down to 16 runs;
//   * the fit is c = P1 m + P2 b0 with an nC x 2nC matrix P computed once on
//     the host.  The reference's solve c = G b (Householder QR, NC-1
//     reflections applied to b, back substitution) is linear in b and every
//     row of G lies in span{Chebyshev columns, e_0..e_{nC-1}}; P is found by
//     running that exact solve on 2nC probe vectors.  Short bins
//     (Nxb < 25 nC), where the moment form loses accuracy, apply the
//     reflections literally to b in LDS.
// HBM traffic: 8 B per point read (+ 8 B written when the sorted bins are
// requested); longer bins add one written and read-back copy of the runs.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "hk_device.h"
#include "hk_internal.h"
#include "hk_lds_sort.h"

namespace hk {

constexpr int kSortNT = 1024;            // threads of the bin / run sorts
constexpr int kIPT = 16;                 // values per thread in an LDS sort
constexpr int kSortCap = kSortNT * kIPT; // values one workgroup sorts (16384)
constexpr int kBucketNT = 256;           // threads of a bucket workgroup
constexpr int kBucketCap = kBucketNT * kIPT;   // values a bucket may hold (4096)
constexpr int kMaxRuns = 16;             // runs per bin the bucket pass takes
constexpr int kMaxNC = 100;              // largest nC (the driver's def_NmaxSample)
constexpr int kMomGroup = 32;            // moments accumulated per pass over a bin
constexpr int kFitGroup = 20;            // the same in the bin and bucket fits (nC = 20
                                         // in one pass within their register budgets)
// The moment form c = P1 m + P2 b0 is exact algebra but inherits the
// conditioning of the Chebyshev normal equations; for short bins
// (Nxb < kPlanMinRatio * nC) the fit applies the Householder reflections
// literally instead (b fits in LDS there).
constexpr int kPlanMinRatio = 25;
constexpr int kHHMaxNxb = kPlanMinRatio * kMaxNC;

struct Seg {
    int64_t start, len;
};

// What a segment is: segments are numbered spec * per_spec + b; those with
// b < nfit are k-table bins (fit index spec * nfit + b), the others are only
// sorted.
struct SegMap {
    int per_spec, nfit;
};

// Fit constants shared by the kernels.
struct FitArgs {
    int Nxb, nC;
    double kmin, unit_scale;
    const double *P;        // nC x 2nC least-squares map
    const double *HV;       // Householder vectors | C | D (short bins), or null
    double *coeff;          // (fit index) * nC
    int32_t *nxmin;         // per fit index
    double *part;           // bucket pass: per (fit index, bucket) partial moments
    double *b0;             // bucket pass: per fit index, first nC values of b
    int bmax;               // buckets per bin (bucket pass)
};

// findCut (resample.h:370-386) on a sorted bin: cnt = #{K <= kmin}; in the
// last bin K[Nx-3] <= kmin (cnt >= Nxb-2) gives Nxb.
__device__ __forceinline__ int nxmin_rule(int cnt, int Nxb, bool last) {
    return (last && cnt >= Nxb - 2) ? Nxb : cnt;
}

// rescale_kernel f=+1 (resample.h:398-437): output point k of a bin with
// Nxmin = nxm > 0 interpolates the sorted values at il and il + 1.
__device__ __forceinline__ void rescale_node(int k, int nxm, int Nxb, double sres, double &ii,
                                             int &il) {
    ii = fma(sres, (double)k, (double)nxm);
    if (ii >= Nxb - 1) ii = 0.999999 * (Nxb - 1);
    il = (int)(ii / ((double)Nxb) * Nxb);
}

// Lower rescale node of output point k (k itself when nxm == 0).
__device__ __forceinline__ int rescale_lower(int k, int nxm, int Nxb, double sres) {
    if (nxm == 0) return k;
    double ii;
    int il;
    rescale_node(k, nxm, Nxb, sres, ii, il);
    return il;
}

// b_k = ln(K2_k unit_scale), K2 the rescaled bin; val(r) = sorted value of rank r.
template <class V>
__device__ __forceinline__ double rescaled_log(const V &val, int k, int nxm, int Nxb,
                                               double sres, double unit_scale) {
    double Kv;
    if (nxm > 0) {
        double ii;
        int il;
        rescale_node(k, nxm, Nxb, sres, ii, il);
        const double Kl = val(il), Kr = val(il + 1);
        Kv = fma(Kr - Kl, ii - il, Kl);
    } else {
        Kv = val(k);
    }
    return log(Kv * unit_scale);
}

// Chebyshev moments i in [g0, g0 + kMomGroup) of one value b at x, T_i by
// the recurrence T_{i+1} = 2x T_i - T_{i-1} started from T_{-1} = T_1 = x,
// T_0 = 1 (so T_1 = 2x - x = x and T_2 = fma(2x, x, -1) exactly as written
// from T_0, T_1).
template <int G>
__device__ __forceinline__ void add_moments(double (&m)[G], int g0, int nC, double x, double bv) {
    double prev = x, cur = 1.0;
    for (int i = 0; i < g0; ++i) {
        const double nx = fma(2.0 * x, cur, -prev);
        prev = cur;
        cur = nx;
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
        if (g0 + q < nC) m[q] = fma(cur, bv, m[q]);
        const double nx = fma(2.0 * x, cur, -prev);
        prev = cur;
        cur = nx;
    }
}

// Sum of m[0..n) over the workgroup (NW waves), deterministic: a fixed
// butterfly per wave, then the waves in order.  out[i] for i < n.
template <int NW, int G>
__device__ __forceinline__ void block_sum(double (&m)[G], int n, double (*red)[kMomGroup],
                                          double *out) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int i = 0; i < G; ++i) {
        if (i < n) {
            double v = m[i];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) red[wv][i] = v;
        }
    }
    __syncthreads();
    if (t < n) {
        double s = 0.0;
        for (int w = 0; w < NW; ++w) s += red[w][t];
        out[t] = s;
    }
    __syncthreads();
}

// c = P1 m + P2 b0 for the threads t < nC.
__device__ __forceinline__ void apply_plan(const double *P, int nC, const double *mom,
                                           const double *b0, double *coeff) {
    const int t = threadIdx.x;
    if (t < nC) {
        const double *Pr = P + (int64_t)t * 2 * nC;
        double c = 0.0;
        for (int i = 0; i < nC; ++i) c = fma(Pr[i], mom[i], c);
        for (int i = 0; i < nC; ++i) c = fma(Pr[nC + i], b0[i], c);
        coeff[t] = c;
    }
}

// Flat segment index of a launch: blockIdx.x, folded over blockIdx.y for
// launches of more than 2^31 - 1 workgroups (never at the sizes here, but
// the fold keeps every grid dimension within its limit).
__device__ __forceinline__ int64_t flat_block() {
    return (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
}

// ---------------------------------------------------------------------------
// One workgroup per segment of at most kSortCap values: load, sort in LDS,
// write the sorted values (when `out` is given) and, for k-table bins, fit.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortNT) void bin_sort_fit_kernel(const double *__restrict__ K,
                                                               const Seg *__restrict__ segs,
                                                               int64_t nseg, SegMap map,
                                                               FitArgs fa,
                                                               double *__restrict__ out) {
    __shared__ double s[lds_sort_slots(kSortCap)];
    __shared__ double red[kSortNT / 64][kMomGroup];
    __shared__ double mom[kMaxNC], b0s[kMaxNC];
    __shared__ int s_nxm;
    const int64_t f = flat_block();
    if (f >= nseg) return;
    const Seg sg = segs[f];
    const int n = (int)sg.len;
    const int t = threadIdx.x;
    const double *src = K + sg.start;
    for (int i = t; i < n; i += kSortNT) s[pad_index(i)] = src[i];
    lds_sort<kSortNT, kIPT>(s, n);
    if (out) {
        double *dst = out + sg.start;
        for (int i = t; i < n; i += kSortNT) dst[i] = s[pad_index(i)];
    }
    const int spec = (int)(f / map.per_spec), b = (int)(f % map.per_spec);
    if (b >= map.nfit) return;
    const int64_t bid = (int64_t)spec * map.nfit + b;
    const int Nxb = fa.Nxb, nC = fa.nC;
    if (t == 0) {
        int lo = 0, hi = Nxb;                  // #{K <= kmin}
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s[pad_index(mid)] <= fa.kmin) lo = mid + 1; else hi = mid;
        }
        s_nxm = nxmin_rule(lo, Nxb, b == map.nfit - 1);
        fa.nxmin[bid] = s_nxm;
    }
    __syncthreads();
    const int nxm = s_nxm;
    const double sres = 1.0 - nxm / ((double)(Nxb - 1));
    const double nlm1 = (double)Nxb - 1.0;
    auto val = [&](int r) { return s[pad_index(r)]; };
    // b_k = ln(K2_k unit_scale) replaces the sorted value k in LDS.  Point k
    // reads the sorted values at its rescale nodes il, il + 1 with
    // il >= k - 1 (ii >= k except at the clamped last point, where
    // il = Nxb - 2), so the points are taken in chunks of kSortNT and a
    // chunk's b is stored only after the next chunk has read its nodes.
    double prev = 0.0;
    int kprev = -1;
    for (int c = 0; c * kSortNT < Nxb; ++c) {
        const int k = c * kSortNT + t;
        const double cur = k < Nxb ? rescaled_log(val, k, nxm, Nxb, sres, fa.unit_scale) : 0.0;
        __syncthreads();
        if (kprev >= 0 && kprev < Nxb) s[pad_index(kprev)] = prev;
        prev = cur;
        kprev = k;
    }
    __syncthreads();
    if (kprev >= 0 && kprev < Nxb) s[pad_index(kprev)] = prev;
    __syncthreads();
    auto bl = [&](int k) -> double & { return s[pad_index(k)]; };
    if (fa.HV) {
        // literal leastSquare_kernel (resample.h:243-353) on b in LDS;
        // HV = [V (Nxb x nC, Householder vectors) | C (nC) | D (nC)]
        const double *Cc = fa.HV + (int64_t)Nxb * nC, *Dd = Cc + nC;
        double *rsum = &red[0][0];
        for (int j = 0; j < nC - 1; ++j) {
            const double *vj = fa.HV + (int64_t)Nxb * j;
            double dot = 0.0;
            for (int k = j + t; k < Nxb; k += kSortNT) dot = fma(vj[k], bl(k), dot);
            for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
            if ((t & 63) == 0) rsum[t >> 6] = dot;
            __syncthreads();
            double tot = 0.0;
            for (int w = 0; w < kSortNT / 64; ++w) tot += rsum[w];
            const double tq = tot / Cc[j];
            __syncthreads();
            for (int k = j + t; k < Nxb; k += kSortNT) bl(k) -= tq * vj[k];
            __syncthreads();
        }
        if (t == 0) {
            bl(nC - 1) /= Dd[nC - 1];
            for (int i = nC - 2; i >= 0; --i) {
                double a = 0.0;
                for (int j = i + 1; j < nC; ++j) a += fa.HV[i + (int64_t)Nxb * j] * bl(j);
                bl(i) = (bl(i) - a) / Dd[i];
            }
        }
        __syncthreads();
        if (t < nC) fa.coeff[bid * nC + t] = bl(t);
        return;
    }
    // kFitGroup moments per pass over b keep the sums within the register
    // budget of a 1024-thread workgroup (32 per pass spill)
    constexpr int G = kFitGroup;
    static_assert(G <= kMomGroup, "block_sum rows");
    for (int g0 = 0; g0 < nC; g0 += G) {
        double m[G];
#pragma unroll
        for (int i = 0; i < G; ++i) m[i] = 0.0;
        for (int k = t; k < Nxb; k += kSortNT) {
            const double bk = bl(k);
            if (g0 == 0 && k < nC) b0s[k] = bk;
            add_moments(m, g0, nC, k * 2.0 / nlm1 - 1.0, bk);
        }
        block_sum<kSortNT / 64>(m, min(G, nC - g0), red, mom + g0);
    }
    apply_plan(fa.P, nC, mom, b0s, fa.coeff + bid * nC);
}

// ---------------------------------------------------------------------------
// Longer segments, step 1: LDS-sorted runs of kSortCap values.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortNT) void run_sort_kernel(const double *__restrict__ K,
                                                           const Seg *__restrict__ segs,
                                                           int64_t nseg, int nrun_max,
                                                           double *__restrict__ runs) {
    __shared__ double s[lds_sort_slots(kSortCap)];
    const int64_t f = flat_block();
    if (f >= nseg * nrun_max) return;
    const Seg sg = segs[f / nrun_max];
    const int64_t off = (int64_t)(f % nrun_max) * kSortCap;
    if (off >= sg.len) return;
    const int n = (int)min((int64_t)kSortCap, sg.len - off);
    const int t = threadIdx.x;
    const double *src = K + sg.start + off;
    for (int i = t; i < n; i += kSortNT) s[pad_index(i)] = src[i];
    lds_sort<kSortNT, kIPT>(s, n);
    double *dst = runs + sg.start + off;
    for (int i = t; i < n; i += kSortNT) dst[i] = s[pad_index(i)];
}

// Pairwise merge of runs of length w into runs of 2w (bins of more than
// kMaxRuns runs): each workgroup emits kMB consecutive outputs of one merged
// pair after one merge-path split, both input slices staged in LDS.
constexpr int kMergeE = 8;
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
                                                        const Seg *__restrict__ segs,
                                                        int64_t nseg, int64_t nblk, int64_t w) {
    __shared__ double sA[kMB], sB[kMB], sO[kMB];
    __shared__ int64_t sp[2];
    const int64_t f = flat_block();
    if (f >= nseg * nblk) return;
    const Seg sg = segs[f / nblk];
    const int64_t o0 = (f % nblk) * kMB;
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

// Bucket pass parameters of one launch (the same for every segment of it).
struct BucketArgs {
    int64_t runlen;     // sorted run length
    int m;              // sample spacing: values m-1, 2m-1, ... of each run
    int g;              // samples per bucket
    int bmax;           // buckets per segment
    double *spl;        // per segment: bmax - 1 splitters
    int32_t *nbk;       // per segment: buckets used
};

// ---------------------------------------------------------------------------
// Longer segments, step 2: value splitters from the regular sample of the
// runs (and, for k-table bins, Nxmin = #{K <= kmin} over all runs).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortNT) void splitter_kernel(const double *__restrict__ runs,
                                                           const Seg *__restrict__ segs,
                                                           int64_t nseg, SegMap map,
                                                           BucketArgs ba, FitArgs fa) {
    __shared__ double s[lds_sort_slots(kSortCap)];
    __shared__ int s_cnt[kSortNT];
    const int64_t f = flat_block();
    if (f >= nseg) return;
    const Seg sg = segs[f];
    const int t = threadIdx.x;
    const int64_t R = ba.runlen;
    const int nrun = (int)((sg.len + R - 1) / R);
    const int per_full = (int)(R / ba.m);
    const int64_t last_len = sg.len - (int64_t)(nrun - 1) * R;
    const int nsamp = nrun > 0 ? (nrun - 1) * per_full + (int)(last_len / ba.m) : 0;
    const double *base = runs + sg.start;
    for (int i = t; i < nsamp; i += kSortNT) {
        const int c = i / per_full, j = i % per_full;
        s[pad_index(i)] = base[(int64_t)c * R + (int64_t)(j + 1) * ba.m - 1];
    }
    // each run's samples are in order: with a power-of-two count per run
    // the sort starts from those runs
    const bool pow2 = per_full >= kIPT && (per_full & (per_full - 1)) == 0;
    lds_sort<kSortNT, kIPT>(s, nsamp, pow2 ? per_full : kIPT);
    const int nb = max(1, (nsamp + ba.g - 1) / ba.g);
    double *spl = ba.spl + f * (ba.bmax - 1);
    for (int w = 1 + t; w < nb; w += kSortNT) spl[w - 1] = s[pad_index(w * ba.g)];
    if (t == 0) ba.nbk[f] = nb;
    const int spec = (int)(f / map.per_spec), b = (int)(f % map.per_spec);
    if (b >= map.nfit) return;
    int cnt = 0;
    for (int c = t; c < nrun; c += kSortNT) {
        const int64_t len = c < nrun - 1 ? R : last_len;
        cnt += (int)run_upper_bound(base + (int64_t)c * R, len, fa.kmin);
    }
    s_cnt[t] = cnt;
    __syncthreads();
    if (t == 0) {
        int tot = 0;
        for (int i = 0; i < min(nrun, kSortNT); ++i) tot += s_cnt[i];
        fa.nxmin[(int64_t)spec * map.nfit + b] = nxmin_rule(tot, (int)sg.len, b == map.nfit - 1);
    }
}

// ---------------------------------------------------------------------------
// Longer segments, step 3: one workgroup per bucket [s_lo, s_hi).  Its
// values above s_lo are gathered from every run into LDS and sorted; the
// run of values equal to s_lo is counted only.  Ranks o0 .. o1-1 of the
// sorted segment are then: E copies of s_lo, then the sorted values; rank
// o1 (the next bucket's first) is the smallest run value >= s_hi.
// Four workgroups per CU: 128 VGPRs with a few spilled, faster than three
// per CU without spills.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBucketNT, 4) void bucket_kernel(const double *__restrict__ runs,
                                                           const Seg *__restrict__ segs,
                                                           int64_t nseg, SegMap map,
                                                           BucketArgs ba, FitArgs fa,
                                                           double *__restrict__ out) {
    __shared__ double s[lds_sort_slots(kBucketCap)];
    __shared__ double red[kBucketNT / 64][kMomGroup];
    __shared__ double mom[kMaxNC];
    __shared__ int64_t r_lo[kMaxRuns], r_eq[kMaxRuns], r_hi[kMaxRuns];
    __shared__ int s_off[kMaxRuns + 1];
    __shared__ double s_halo;
    __shared__ int s_k[2];
    const int64_t f = flat_block();
    if (f >= nseg * ba.bmax) return;
    const int64_t sgi = f / ba.bmax;
    const int w = (int)(f % ba.bmax);
    const Seg sg = segs[sgi];
    const int t = threadIdx.x;
    const int spec = (int)(sgi / map.per_spec), b = (int)(sgi % map.per_spec);
    const bool fit = b < map.nfit;
    const int64_t bid = (int64_t)spec * map.nfit + b;
    const int nbk = ba.nbk[sgi];
    double *part = fit ? fa.part + (bid * ba.bmax + w) * fa.nC : nullptr;
    if (w >= nbk) {                            // unused bucket: zero partials
        if (fit)
            for (int i = t; i < fa.nC; i += kBucketNT) part[i] = 0.0;
        return;
    }
    const double *spl = ba.spl + sgi * (ba.bmax - 1);
    const bool has_lo = w > 0, has_hi = w < nbk - 1;
    const double slo = has_lo ? spl[w - 1] : -INFINITY;
    const double shi = has_hi ? spl[w] : INFINITY;
    const int64_t R = ba.runlen;
    const int nrun = (int)((sg.len + R - 1) / R);
    const double *base = runs + sg.start;
    const bool empty = has_lo && has_hi && !(slo < shi);
    if (t < nrun) {
        const double *rv = base + (int64_t)t * R;
        const int64_t len = min(R, sg.len - (int64_t)t * R);
        int64_t lo = 0, eq = 0, hi = len;
        if (has_lo) {
            lo = run_lower_bound(rv, len, slo);
            eq = run_upper_bound(rv, len, slo);
        }
        if (has_hi) hi = run_lower_bound(rv, len, shi);
        if (empty) { eq = lo; hi = lo; }
        r_lo[t] = lo; r_eq[t] = eq; r_hi[t] = max(hi, eq);
    }
    __syncthreads();
    if (t == 0) {
        int off = 0;
        double halo = INFINITY;
        for (int c = 0; c < nrun; ++c) {
            s_off[c] = off;
            off += (int)min((int64_t)kBucketCap - off, r_hi[c] - r_eq[c]);
            const int64_t len = min(R, sg.len - (int64_t)c * R);
            if (r_hi[c] < len) halo = fmin(halo, base[(int64_t)c * R + r_hi[c]]);
        }
        s_off[nrun] = off;
        s_halo = halo;
    }
    __syncthreads();
    int64_t o0 = 0, E = 0;
    for (int c = 0; c < nrun; ++c) { o0 += r_lo[c]; E += r_eq[c] - r_lo[c]; }
    const int M = s_off[nrun];
    for (int c = 0; c < nrun; ++c) {
        const int cnt = s_off[c + 1] - s_off[c];
        const double *rv = base + (int64_t)c * R + r_eq[c];
        for (int i = t; i < cnt; i += kBucketNT) s[pad_index(s_off[c] + i)] = rv[i];
    }
    lds_sort<kBucketNT, kIPT>(s, M);
    const int64_t o1 = o0 + E + M;
    const double halo = s_halo;
    auto val = [&](int64_t r) {
        return r < o0 + E ? slo : (r < o1 ? s[pad_index((int)(r - o0 - E))] : halo);
    };
    if (out) {
        double *dst = out + sg.start;
        for (int64_t r = o0 + t; r < o1; r += kBucketNT) dst[r] = val(r);
    }
    if (!fit) return;
    const int Nxb = fa.Nxb, nC = fa.nC;
    const int nxm = fa.nxmin[bid];
    const double sres = 1.0 - nxm / ((double)(Nxb - 1));
    const double nlm1 = (double)Nxb - 1.0;
    // output points whose lower rescale node lies in [o0, o1): the rescale
    // node is non-decreasing in k
    if (t < 2) {
        const int64_t target = t == 0 ? o0 : o1;
        int lo = 0, hi = Nxb;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (rescale_lower(mid, nxm, Nxb, sres) < target) lo = mid + 1; else hi = mid;
        }
        s_k[t] = lo;
    }
    __syncthreads();
    const int k0 = s_k[0], k1 = s_k[1];
    auto vali = [&](int r) { return val((int64_t)r); };
    for (int g0 = 0; g0 < nC; g0 += kFitGroup) {
        double m[kFitGroup];
#pragma unroll
        for (int i = 0; i < kFitGroup; ++i) m[i] = 0.0;
        for (int k = k0 + t; k < k1; k += kBucketNT) {
            const double bv = rescaled_log(vali, k, nxm, Nxb, sres, fa.unit_scale);
            if (g0 == 0 && k < nC) fa.b0[bid * nC + k] = bv;
            add_moments(m, g0, nC, k * 2.0 / nlm1 - 1.0, bv);
        }
        block_sum<kBucketNT / 64>(m, min(kFitGroup, nC - g0), red, mom + g0);
    }
    for (int i = t; i < nC; i += kBucketNT) part[i] = mom[i];
}

// Partial moments of the buckets summed in bucket order, then c = P1 m + P2 b0.
__global__ __launch_bounds__(128) void bucket_combine_kernel(FitArgs fa, int64_t nfit) {
    __shared__ double mom[kMaxNC];
    const int64_t bid = flat_block();
    if (bid >= nfit) return;
    const int t = threadIdx.x, nC = fa.nC;
    if (t < nC) {
        double s = 0.0;
        for (int w = 0; w < fa.bmax; ++w) s += fa.part[(bid * fa.bmax + w) * nC + t];
        mom[t] = s;
    }
    __syncthreads();
    apply_plan(fa.P, nC, mom, fa.b0 + bid * nC, fa.coeff + bid * nC);
}

// ---------------------------------------------------------------------------
// Fit with findCut's literal rules (resample.h:370-386) on the global index
// id = b*Nxb + k, for bins whose bin-key segments differ from the fixed Nxb
// chunks (each sorted by segment first).  Ks: the per-segment sorted K of
// one spectrum.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kfit_literal_kernel(const double *__restrict__ Ksorted,
                                                           int64_t nx, int nbins, int nspec,
                                                           FitArgs fa) {
    __shared__ double red[4][kMomGroup];
    __shared__ double bl[kHHMaxNxb];
    __shared__ double mom[kMaxNC];
    __shared__ double b0[kMaxNC];
    __shared__ int s_r1[256], s_r2[256], s_r3[256];
    const int64_t f = flat_block();
    if (f >= (int64_t)nbins * nspec) return;
    const int b = (int)(f % nbins), sp = (int)(f / nbins), t = threadIdx.x;
    const int Nxb = fa.Nxb, nC = fa.nC;
    const double *Ks = Ksorted + (int64_t)sp * nx;
    const double *Kb = Ks + (int64_t)b * Nxb;
    const double kmin = fa.kmin;
    int r1 = -1, r2 = 0, r3 = 0;
    for (int k = t; k < Nxb; k += 256) {
        const int64_t id = (int64_t)b * Nxb + k;
        if (id >= nx - 1) continue;
        const double K0 = Ks[id], K1 = Ks[id + 1];
        if (K0 <= kmin && K1 > kmin) r1 = max(r1, k + 1);
        if (K0 <= kmin && k == Nxb - 1) r2 = 1;
        if (K0 <= kmin && id == nx - 3) r3 = 1;
    }
    s_r1[t] = r1; s_r2[t] = r2; s_r3[t] = r3;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            s_r1[t] = max(s_r1[t], s_r1[t + s]);
            s_r2[t] |= s_r2[t + s];
            s_r3[t] |= s_r3[t + s];
        }
        __syncthreads();
    }
    const int nxm = s_r3[0] ? Nxb : (s_r2[0] ? Nxb : (s_r1[0] >= 0 ? s_r1[0] : 0));
    const int64_t bid = (int64_t)sp * nbins + b;
    const double sres = 1.0 - nxm / ((double)(Nxb - 1));
    const double nlm1 = (double)Nxb - 1.0;
    auto val = [&](int r) { return Kb[r]; };
    if (fa.HV) {
        for (int k = t; k < Nxb; k += 256) bl[k] = rescaled_log(val, k, nxm, Nxb, sres, fa.unit_scale);
        __syncthreads();
        const double *Cc = fa.HV + (int64_t)Nxb * nC, *Dd = Cc + nC;
        for (int j = 0; j < nC - 1; ++j) {
            const double *vj = fa.HV + (int64_t)Nxb * j;
            double dot = 0.0;
            for (int k = j + t; k < Nxb; k += 256) dot = fma(vj[k], bl[k], dot);
            for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
            if ((t & 63) == 0) red[t >> 6][0] = dot;
            __syncthreads();
            const double tq = (((red[0][0] + red[1][0]) + red[2][0]) + red[3][0]) / Cc[j];
            __syncthreads();
            for (int k = j + t; k < Nxb; k += 256) bl[k] -= tq * vj[k];
            __syncthreads();
        }
        if (t == 0) {
            bl[nC - 1] /= Dd[nC - 1];
            for (int i = nC - 2; i >= 0; --i) {
                double a = 0.0;
                for (int j = i + 1; j < nC; ++j) a += fa.HV[i + (int64_t)Nxb * j] * bl[j];
                bl[i] = (bl[i] - a) / Dd[i];
            }
        }
        __syncthreads();
        if (t < nC) fa.coeff[bid * nC + t] = bl[t];
        if (t == 0) fa.nxmin[bid] = nxm;
        return;
    }
    for (int g0 = 0; g0 < nC; g0 += kMomGroup) {
        double m[kMomGroup];
#pragma unroll
        for (int i = 0; i < kMomGroup; ++i) m[i] = 0.0;
        for (int k = t; k < Nxb; k += 256) {
            const double bv = rescaled_log(val, k, nxm, Nxb, sres, fa.unit_scale);
            if (g0 == 0 && k < nC) b0[k] = bv;
            add_moments(m, g0, nC, k * 2.0 / nlm1 - 1.0, bv);
        }
        block_sum<4>(m, min(kMomGroup, nC - g0), red, mom + g0);
    }
    apply_plan(fa.P, nC, mom, b0, fa.coeff + bid * nC);
    if (t == 0) fa.nxmin[bid] = nxm;
}

// ---------------------------------------------------------------------------
// Backward of the fit: dL/dc -> dL/dK.  The forward per bin is
//   K -> sorted Ks -> K2_k = Ks[il] + (Ks[il+1] - Ks[il]) f_k (rescale, or
//   K2 = Ks when Nxmin = 0) -> b = ln(K2 unit_scale) -> c = G b,
// with Nxmin and the sort order piecewise constant in K (held fixed).
// Transposed: gb = G^T gc (moment form: gb_k = sum_i u_i T_i(x_k) + v_k for
// k < nC with u = P1^T gc, v = P2^T gc; short bins: the transposed
// back substitution and the reflections in reverse order), gK2 = gb / K2,
// then each sorted rank gathers its interpolation weights (deterministic),
// and each point takes the gradient of its rank (values equal within a bin
// share the mean gradient of their ranks).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kt_bwd_b_kernel(const double *__restrict__ Ksorted,
                                                       int64_t nx, int nbins, int64_t nfit,
                                                       FitArgs fa,
                                                       const double *__restrict__ gcoeff,
                                                       double *__restrict__ gK2) {
    __shared__ double u[kMaxNC], v[kMaxNC], gz[kMaxNC], gc[kMaxNC];
    __shared__ double gb[kHHMaxNxb];
    __shared__ double red[4];
    const int64_t bid = flat_block();
    if (bid >= nfit) return;
    const int t = threadIdx.x, Nxb = fa.Nxb, nC = fa.nC;
    const int64_t sp = bid / nbins, b = bid % nbins;
    const double *Kb = Ksorted + sp * nx + b * Nxb;
    double *g2 = gK2 + sp * nx + b * Nxb;
    const int nxm = fa.nxmin[bid];
    const double sres = 1.0 - nxm / ((double)(Nxb - 1));
    const double nlm1 = (double)Nxb - 1.0;
    if (t < nC) gc[t] = gcoeff[bid * nC + t];
    __syncthreads();
    auto val = [&](int r) { return Kb[r]; };
    auto K2 = [&](int k) {
        if (nxm > 0) {
            double ii;
            int il;
            rescale_node(k, nxm, Nxb, sres, ii, il);
            const double Kl = Kb[il], Kr = Kb[il + 1];
            return fma(Kr - Kl, ii - il, Kl);
        }
        return val(k);
    };
    if (fa.HV) {
        // c = R^{-1} (H_{nC-2} ... H_0 b)[0:nC], R_ii = D_i, R_ij = V[i + Nxb j]
        const double *Cc = fa.HV + (int64_t)Nxb * nC, *Dd = Cc + nC;
        if (t == 0) {
            for (int i = 0; i < nC; ++i) {            // R^T gz = gc
                double a = 0.0;
                for (int j = 0; j < i; ++j) a += fa.HV[j + (int64_t)Nxb * i] * gz[j];
                gz[i] = (gc[i] - a) / Dd[i];
            }
        }
        __syncthreads();
        for (int k = t; k < Nxb; k += 256) gb[k] = k < nC ? gz[k] : 0.0;
        __syncthreads();
        for (int j = nC - 2; j >= 0; --j) {
            const double *vj = fa.HV + (int64_t)Nxb * j;
            double dot = 0.0;
            for (int k = j + t; k < Nxb; k += 256) dot = fma(vj[k], gb[k], dot);
            for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
            if ((t & 63) == 0) red[t >> 6] = dot;
            __syncthreads();
            const double tq = (((red[0] + red[1]) + red[2]) + red[3]) / Cc[j];
            __syncthreads();
            for (int k = j + t; k < Nxb; k += 256) gb[k] -= tq * vj[k];
            __syncthreads();
        }
        for (int k = t; k < Nxb; k += 256) g2[k] = gb[k] / K2(k);
        return;
    }
    if (t < nC) {
        double a = 0.0, c = 0.0;
        for (int r = 0; r < nC; ++r) {
            a = fma(fa.P[(int64_t)r * 2 * nC + t], gc[r], a);
            c = fma(fa.P[(int64_t)r * 2 * nC + nC + t], gc[r], c);
        }
        u[t] = a;
        v[t] = c;
    }
    __syncthreads();
    for (int k = t; k < Nxb; k += 256) {
        const double x = k * 2.0 / nlm1 - 1.0;
        double prev = x, cur = 1.0, g = 0.0;
        for (int i = 0; i < nC; ++i) {
            g = fma(u[i], cur, g);
            const double nxt = fma(2.0 * x, cur, -prev);
            prev = cur;
            cur = nxt;
        }
        if (k < nC) g += v[k];
        g2[k] = g / K2(k);
    }
}

// dL/dKs at each sorted rank r of a bin: sum over the output points whose
// rescale nodes are r (weight 1 - f) or r - 1 (weight f).
__global__ __launch_bounds__(256) void kt_bwd_rescale_kernel(const double *__restrict__ gK2,
                                                             int64_t nx, int nbins, FitArgs fa,
                                                             double *__restrict__ gKs) {
    const int64_t bid = (int64_t)blockIdx.y * nbins + blockIdx.x;
    const int t = threadIdx.x, Nxb = fa.Nxb;
    const int64_t sp = bid / nbins, b = bid % nbins;
    const double *g2 = gK2 + sp * nx + b * Nxb;
    double *gs = gKs + sp * nx + b * Nxb;
    const int nxm = fa.nxmin[bid];
    const double sres = 1.0 - nxm / ((double)(Nxb - 1));
    auto first = [&](int64_t target) {        // first k with il(k) >= target
        int lo = 0, hi = Nxb;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (rescale_lower(mid, nxm, Nxb, sres) < target) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    for (int r = blockIdx.z * 256 + t; r < Nxb; r += 256 * gridDim.z) {
        if (nxm == 0) { gs[r] = g2[r]; continue; }
        const int ka = first(r - 1), kb = first(r), kc = first(r + 1);
        double g = 0.0;
        for (int k = kb; k < kc; ++k) {        // lower node r
            double ii;
            int il;
            rescale_node(k, nxm, Nxb, sres, ii, il);
            g = fma(g2[k], 1.0 - (ii - il), g);
        }
        for (int k = ka; k < kb; ++k) {        // upper node r (lower node r - 1)
            double ii;
            int il;
            rescale_node(k, nxm, Nxb, sres, ii, il);
            g = fma(g2[k], ii - il, g);
        }
        gs[r] = g;
    }
}

// dL/dK_j = dL/dKs at the rank of K_j in its sorted bin (mean over the ranks
// of equal values).
__global__ __launch_bounds__(256) void kt_bwd_unsort_kernel(const double *__restrict__ K,
                                                            const double *__restrict__ Ksorted,
                                                            const double *__restrict__ gKs,
                                                            int64_t nx, int nbins, int Nxb,
                                                            int64_t n, double *__restrict__ gK) {
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * 256) {
        const int64_t sp = j / nx, q = j % nx, b = q / Nxb;
        if (b >= nbins) { gK[j] = 0.0; continue; }
        const int64_t o = sp * nx + b * Nxb;
        const double kv = K[j];
        const int64_t lo = run_lower_bound(Ksorted + o, Nxb, kv);
        const int64_t hi = run_upper_bound(Ksorted + o, Nxb, kv);
        double g = 0.0;
        for (int64_t r = lo; r < hi; ++r) g += gKs[o + r];
        gK[j] = hi > lo ? g / (double)(hi - lo) : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Reconstruction: expfx_kernel (resample.h:497-524) then rescale f=-1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void expfx_kernel(const double *__restrict__ coeff, int nC,
                                                    int Nxb, double unit_scale,
                                                    double *__restrict__ F, int64_t nbin) {
    __shared__ double cs[kMaxNC];
    const int64_t bin = flat_block();
    if (bin >= nbin) return;
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
                                                          double *__restrict__ Ky, int64_t nbin) {
    const int64_t bin = flat_block();
    if (bin >= nbin) return;
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

// ---------------------------------------------------------------------------
// Sorting plan for segments longer than one LDS sort.
// ---------------------------------------------------------------------------
struct SortPlan {
    bool chunked = false;   // some segment holds more than kSortCap values
    int nrun = 0;           // LDS-sorted runs of the longest segment
    int merges = 0;         // pairwise merge passes before the bucket pass
    int64_t runlen = 0;     // run length after the merges
    int m = 0, g = 0, bmax = 0;
};

static int plan_sort(int64_t maxlen, SortPlan &sp) {
    sp = SortPlan();
    if (maxlen <= kSortCap) return HK_OK;
    sp.chunked = true;
    sp.nrun = (int)((maxlen + kSortCap - 1) / kSortCap);
    sp.runlen = kSortCap;
    // sample spacing: at most kSortCap - kMaxRuns samples per segment
    sp.m = (int)std::max<int64_t>(8, (maxlen + (kSortCap - kMaxRuns) - 1) / (kSortCap - kMaxRuns));
    if (sp.m > kBucketCap / 2) {
        set_error("k-table: a bin of %lld points is longer than the per-bin sort supports",
                  (long long)maxlen);
        return HK_ERR_UNSUPPORTED;
    }
    int64_t k = (maxlen + sp.runlen - 1) / sp.runlen;
    // at most kMaxRuns runs, and room for a bucket: m (g + runs) <= kBucketCap
    while (k > kMaxRuns || sp.m * (k + 1) > kBucketCap) {
        sp.runlen *= 2;
        ++sp.merges;
        k = (maxlen + sp.runlen - 1) / sp.runlen;
    }
    sp.g = kBucketCap / sp.m - (int)k;
    const int64_t nsamp = maxlen / sp.m;
    sp.bmax = (int)std::max<int64_t>(1, (nsamp + sp.g - 1) / sp.g);
    return HK_OK;
}

// Workspace of the sort (and bucket fit) of nseg segments of nvals values in
// all, nfit of them k-table bins with nC coefficients.
struct SortLayout {
    size_t off_runs, off_runs2, off_spl, off_nbk, off_part, off_b0, total;
};

static SortLayout sort_layout(const SortPlan &sp, int64_t nvals, int64_t nseg, int64_t nfit,
                              int nC, size_t o) {
    SortLayout L;
    L.off_runs = L.off_runs2 = L.off_spl = L.off_nbk = L.off_part = L.off_b0 = o;
    if (sp.chunked) {
        L.off_runs = o; o = align_up(o + sizeof(double) * (size_t)nvals, 256);
        L.off_runs2 = o;
        if (sp.merges > 0) o = align_up(o + sizeof(double) * (size_t)nvals, 256);
        L.off_spl = o; o = align_up(o + sizeof(double) * (size_t)nseg * std::max(1, sp.bmax - 1), 256);
        L.off_nbk = o; o = align_up(o + sizeof(int32_t) * (size_t)nseg, 256);
        L.off_part = o; o = align_up(o + sizeof(double) * (size_t)nfit * sp.bmax * nC, 256);
        L.off_b0 = o; o = align_up(o + sizeof(double) * (size_t)nfit * nC, 256);
    }
    L.total = o;
    return L;
}

// Grid of n workgroups numbered by flat_block().
static dim3 flat_grid(int64_t n) {
    const int64_t X = (int64_t)1 << 30;
    if (n <= X) return dim3((unsigned)std::max<int64_t>(n, 1));
    return dim3((unsigned)X, (unsigned)((n + X - 1) / X));
}

// Sorts every segment of segs (device) ascending into `out` (may be null
// when only the fit is wanted) and, for the segments map marks as k-table
// bins, fits them (fa).  base: the workspace past the caller's tables.
static int sort_fit_segments(const double *K, const Seg *segs, int64_t nseg, SegMap map,
                             const SortPlan &sp, const SortLayout &L, char *base, FitArgs fa,
                             double *out, hipStream_t s) {
    const int64_t nfit = (nseg / map.per_spec) * map.nfit;
    if (!sp.chunked) {
        hipLaunchKernelGGL(bin_sort_fit_kernel, flat_grid(nseg), dim3(kSortNT), 0, s, K, segs,
                           nseg, map, fa, out);
        HK_LAUNCH_CHECK();
        return HK_OK;
    }
    double *runs = (double *)(base + L.off_runs);
    double *runs2 = (double *)(base + L.off_runs2);
    hipLaunchKernelGGL(run_sort_kernel, flat_grid(nseg * sp.nrun), dim3(kSortNT), 0, s, K, segs,
                       nseg, sp.nrun, runs);
    HK_LAUNCH_CHECK();
    // runs and K are addressed alike: run values of a segment start at its
    // own offset in the runs buffer
    const int64_t maxlen = (int64_t)sp.nrun * kSortCap;
    const int64_t nblk = (maxlen + kMB - 1) / kMB;
    int64_t w = kSortCap;
    for (int p = 0; p < sp.merges; ++p, w *= 2) {
        hipLaunchKernelGGL(merge_lds_kernel, flat_grid(nseg * nblk), dim3(256), 0, s, runs, runs2,
                           segs, nseg, nblk, w);
        HK_LAUNCH_CHECK();
        std::swap(runs, runs2);
    }
    BucketArgs ba;
    ba.runlen = sp.runlen;
    ba.m = sp.m;
    ba.g = sp.g;
    ba.bmax = sp.bmax;
    ba.spl = (double *)(base + L.off_spl);
    ba.nbk = (int32_t *)(base + L.off_nbk);
    fa.part = (double *)(base + L.off_part);
    fa.b0 = (double *)(base + L.off_b0);
    fa.bmax = sp.bmax;
    hipLaunchKernelGGL(splitter_kernel, flat_grid(nseg), dim3(kSortNT), 0, s, runs, segs, nseg, map,
                       ba, fa);
    HK_LAUNCH_CHECK();
    hipLaunchKernelGGL(bucket_kernel, flat_grid(nseg * sp.bmax), dim3(kBucketNT), 0, s, runs, segs,
                       nseg, map, ba, fa, out);
    HK_LAUNCH_CHECK();
    if (nfit > 0) {
        hipLaunchKernelGGL(bucket_combine_kernel, flat_grid(nfit), dim3(128), 0, s, fa, nfit);
        HK_LAUNCH_CHECK();
    }
    return HK_OK;
}

// Workspace layout of hk_ktable.
struct KtLayout {
    size_t off_seg, off_plan, off_hv, off_segs, off_sorted, total;
    bool householder, aligned;
    SortPlan sp;
    SortLayout sl;
};

static bool use_householder(int64_t Nxb, int nC) { return Nxb < (int64_t)kPlanMinRatio * nC; }

// Bin-key segments of the grid (host) and whether they equal the fixed Nxb
// chunks; maxlen = longest segment.
static int bin_layout(const hk_grid *g, double numax, int nbins, std::vector<int64_t> &seg,
                      bool &aligned, int64_t &maxlen) {
    seg.assign(nbins + 2, 0);
    if (int e = hk_bin_segments(g, numax, nbins, seg.data())) return e;
    const int64_t Nxb = g->nx / nbins;
    aligned = true;
    maxlen = 0;
    for (int b = 0; b <= nbins; ++b) {
        maxlen = std::max(maxlen, seg[b + 1] - seg[b]);
        if (seg[b] != (int64_t)b * Nxb) aligned = false;
    }
    return HK_OK;
}

static int kt_layout(const hk_grid *g, double numax, int nbins, int nC, int nspec, KtLayout &L) {
    std::vector<int64_t> seg;
    int64_t maxlen = 0;
    if (int e = bin_layout(g, numax, nbins, seg, L.aligned, maxlen)) return e;
    const int64_t nx = g->nx, Nxb = nx / nbins;
    if (int e = plan_sort(maxlen, L.sp)) return e;
    size_t o = 0;
    L.householder = use_householder(Nxb, nC);
    L.off_seg = o; o = align_up(o + sizeof(int64_t) * (nbins + 2), 256);
    L.off_plan = o; o = align_up(o + sizeof(double) * 2 * nC * nC, 256);
    L.off_hv = o;
    if (L.householder) o = align_up(o + sizeof(double) * (size_t)(Nxb * nC + 2 * nC), 256);
    const int64_t nseg = (int64_t)nspec * (nbins + 1);
    L.off_segs = o; o = align_up(o + sizeof(Seg) * (size_t)nseg, 256);
    L.off_sorted = o;
    if (!L.aligned) o = align_up(o + sizeof(double) * (size_t)nspec * nx, 256);
    L.sl = sort_layout(L.sp, (int64_t)nspec * nx, nseg, L.aligned ? (int64_t)nspec * nbins : 0,
                       nC, o);
    L.total = L.sl.total + 256;
    return HK_OK;
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
        if (xj >= numax) return nbins;
        // first bin whose interval holds x_j (bounds ascending)
        int lo = 0, hi = nbins;
        while (lo < hi) {
            const int m = (lo + hi) / 2;
            if (bnd[m + 1] <= xj) lo = m + 1; else hi = m;
        }
        return (lo < nbins && bnd[lo] <= xj && xj < bnd[lo + 1]) ? lo : 0;
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
    HK_CHECK_ARG(P && Nxb >= 2 && nC >= 1 && nC <= kMaxNC && nC <= Nxb,
                 "hk_ktable_plan_host: need 1 <= nC <= %d, nC <= Nxb and Nxb >= 2", kMaxNC);
    std::vector<double> v;
    int e = get_plan(Nxb, nC, v);
    if (e) { set_error("hk_ktable_plan_host: singular probe system"); return e; }
    memcpy(P, v.data(), v.size() * sizeof(double));
    return HK_OK;
}

static int check_kt_args(const hk_grid *g, int nbins, int nC, const char *fn) {
    HK_CHECK_ARG(g->nx % nbins == 0, "%s: Nx %lld not divisible by nbins %d", fn,
                 (long long)g->nx, nbins);
    const int64_t Nxb = g->nx / nbins;
    HK_CHECK_ARG(Nxb >= 2 && Nxb < ((int64_t)1 << 31), "%s: %lld points per bin (need 2 .. 2^31-1)",
                 fn, (long long)Nxb);
    HK_CHECK_ARG(nC >= 1 && nC <= kMaxNC && nC <= Nxb,
                 "%s: nC %d (need 1 <= nC <= %d and nC <= %lld points per bin)", fn, nC, kMaxNC,
                 (long long)Nxb);
    return HK_OK;
}

extern "C" size_t hk_ktable_workspace_bytes(int64_t nx, int nbins, int nC, int nspec) {
    if (nx <= 0 || nbins <= 0 || nC <= 0 || nspec <= 0 || nx % nbins) return 0;
    // the segments of the byte count: fixed Nxb chunks (the longest any grid
    // of nx points and nbins bins can have is Nxb + 1 by the bin-key rule)
    const hk_grid g = {0.0, 1.0, nx};
    KtLayout L;
    if (kt_layout(&g, (double)nx, nbins, std::min(nC, kMaxNC), nspec, L)) return 0;
    // room for unaligned segments too (bin keys of other numin/dnu)
    SortPlan sp;
    if (plan_sort(nx / nbins + 1, sp)) return 0;
    const int64_t nseg = (int64_t)nspec * (nbins + 1);
    const size_t base = align_up(L.off_segs + sizeof(Seg) * (size_t)nseg, 256) +
                        align_up(sizeof(double) * (size_t)nspec * nx, 256);
    const SortLayout sl = sort_layout(sp, (int64_t)nspec * nx, nseg, (int64_t)nspec * nbins, nC,
                                      base);
    return std::max(L.total, sl.total + 256);
}

extern "C" int hk_ktable_prepare(const hk_grid *g, double numax, int nbins, int nC, void *ws,
                                 size_t ws_bytes) {
    HK_CHECK_ARG(g && ws && nbins > 0, "hk_ktable_prepare: bad arguments");
    if (int e = check_kt_args(g, nbins, nC, "hk_ktable_prepare")) return e;
    const int Nxb = (int)(g->nx / nbins);
    KtLayout L;
    if (int e = kt_layout(g, numax, nbins, nC, 1, L)) return e;
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
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nper = nbins + 1;
    if (i >= (int64_t)nspec * nper) return;
    const int s = (int)(i / nper), b = (int)(i % nper);
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
    if (int e = check_kt_args(g, nbins, nC, "hk_ktable")) return e;
    const int Nxb = (int)(g->nx / nbins);
    KtLayout L;
    if (int e = kt_layout(g, numax, nbins, nC, nspec, L)) return e;
    if (ws_bytes < L.total) {
        set_error("hk_ktable: workspace %zu bytes, %zu needed", ws_bytes, L.total);
        return HK_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)align_up((size_t)ws, 256);
    const int64_t *seg = (const int64_t *)(base + L.off_seg);
    Seg *segs = (Seg *)(base + L.off_segs);
    const int64_t nseg = (int64_t)nspec * (nbins + 1);
    hipLaunchKernelGGL(build_segs_kernel, dim3((unsigned)((nseg + 255) / 256)), dim3(256), 0, s,
                       seg, nbins, g->nx, nspec, segs);
    HK_LAUNCH_CHECK();
    FitArgs fa = {};
    fa.Nxb = Nxb;
    fa.nC = nC;
    fa.kmin = kmin;
    fa.unit_scale = unit_scale;
    fa.P = (const double *)(base + L.off_plan);
    fa.HV = L.householder ? (const double *)(base + L.off_hv) : nullptr;
    fa.coeff = coeff;
    fa.nxmin = nxmin;
    if (L.aligned) {
        // sorted bins are the fixed Nxb chunks: sort and fit in one pass
        return sort_fit_segments(K, segs, nseg, SegMap{nbins + 1, nbins}, L.sp, L.sl, base, fa,
                                 Ksorted, s);
    }
    // bin-key segments differ from the fixed chunks: sort the segments, then
    // apply findCut's literal rules on the fixed chunks of the sorted spectrum
    double *sorted = Ksorted ? Ksorted : (double *)(base + L.off_sorted);
    if (int e = sort_fit_segments(K, segs, nseg, SegMap{nbins + 1, 0}, L.sp, L.sl, base, fa, sorted,
                                  s))
        return e;
    hipLaunchKernelGGL(kfit_literal_kernel, flat_grid((int64_t)nbins * nspec), dim3(256), 0, s,
                       sorted, g->nx, nbins, nspec, fa);
    HK_LAUNCH_CHECK();
    return HK_OK;
}

extern "C" size_t hk_ktable_backward_workspace_bytes(int64_t nx, int nbins, int nC, int nspec) {
    const size_t kt = hk_ktable_workspace_bytes(nx, nbins, nC, nspec);
    if (kt == 0) return 0;
    return align_up(kt, 256) + 2 * align_up(sizeof(double) * (size_t)nspec * nx, 256) + 256;
}

extern "C" int hk_ktable_backward(const double *K, const double *Ksorted, const int32_t *nxmin,
                                  int nspec, const hk_grid *g, double numax, int nbins, int nC,
                                  const double *gcoeff, double *gK, void *ws, size_t ws_bytes,
                                  hk_stream_t stream) {
    HK_CHECK_ARG(K && Ksorted && nxmin && g && gcoeff && gK && ws && nspec > 0 && nbins > 0,
                 "hk_ktable_backward: bad arguments");
    if (int e = check_kt_args(g, nbins, nC, "hk_ktable_backward")) return e;
    const int Nxb = (int)(g->nx / nbins);
    KtLayout L;
    if (int e = kt_layout(g, numax, nbins, nC, nspec, L)) return e;
    if (!L.aligned) {
        set_error("hk_ktable_backward: bins whose bin-key segments differ from the fixed "
                  "%d-point chunks are not supported", Nxb);
        return HK_ERR_UNSUPPORTED;
    }
    const size_t need = hk_ktable_backward_workspace_bytes(g->nx, nbins, nC, nspec);
    if (ws_bytes < need) {
        set_error("hk_ktable_backward: workspace %zu bytes, %zu needed", ws_bytes, need);
        return HK_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)align_up((size_t)ws, 256);
    const size_t kt = align_up(hk_ktable_workspace_bytes(g->nx, nbins, nC, nspec), 256);
    double *gK2 = (double *)(base + kt);
    double *gKs = (double *)(base + kt + align_up(sizeof(double) * (size_t)nspec * g->nx, 256));
    FitArgs fa = {};
    fa.Nxb = Nxb;
    fa.nC = nC;
    fa.P = (const double *)(base + L.off_plan);
    fa.HV = L.householder ? (const double *)(base + L.off_hv) : nullptr;
    fa.nxmin = const_cast<int32_t *>(nxmin);
    const int64_t nfit = (int64_t)nspec * nbins;
    hipLaunchKernelGGL(kt_bwd_b_kernel, flat_grid(nfit), dim3(256), 0, s, Ksorted, g->nx, nbins,
                       nfit, fa, gcoeff, gK2);
    HK_LAUNCH_CHECK();
    const unsigned zs = (unsigned)std::max(1, std::min(64, Nxb / 2048));
    HK_CHECK_ARG(nspec <= 65535, "hk_ktable_backward: at most 65535 spectra per call");
    hipLaunchKernelGGL(kt_bwd_rescale_kernel, dim3(nbins, nspec, zs), dim3(256), 0, s, gK2, g->nx,
                       nbins, fa, gKs);
    HK_LAUNCH_CHECK();
    const int64_t n = (int64_t)nspec * g->nx;
    const unsigned nb = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(kt_bwd_unsort_kernel, dim3(nb), dim3(256), 0, s, K, Ksorted, gKs, g->nx,
                       nbins, Nxb, n, gK);
    HK_LAUNCH_CHECK();
    return HK_OK;
}

// Ascending sort of arbitrary bins [seg[b], seg[b+1]) of one spectrum (the
// per-bin thrust sorts of heliosk.cu:2046-2051 for bins-file and per-bin
// grids).  Workspace: the segment table and the run buffers of the longest
// segment.
static int sort_bins_plan(int64_t nx, const int64_t *seg, int nseg, SortPlan &sp, SortLayout &sl,
                          size_t &total) {
    int64_t maxlen = 0;
    for (int b = 0; b < nseg; ++b) maxlen = std::max(maxlen, seg[b + 1] - seg[b]);
    if (int e = plan_sort(maxlen, sp)) return e;
    sl = sort_layout(sp, nx, nseg, 0, 1, align_up(sizeof(Seg) * (size_t)nseg, 256));
    total = sl.total + 256;
    return HK_OK;
}

extern "C" size_t hk_sort_bins_workspace_bytes(int64_t nx, int nseg) {
    if (nx <= 0 || nseg <= 0) return 0;
    // the longest possible segment is the whole spectrum
    SortPlan sp;
    if (plan_sort(nx, sp)) return 0;
    return sort_layout(sp, nx, nseg, 0, 1, align_up(sizeof(Seg) * (size_t)nseg, 256)).total + 256;
}

extern "C" int hk_sort_bins(const double *K, int64_t nx, const int64_t *seg, int nseg,
                            double *Ksorted, void *ws, size_t ws_bytes, hk_stream_t stream) {
    HK_CHECK_ARG(K && seg && Ksorted && ws && nx > 0 && nseg > 0, "hk_sort_bins: bad arguments");
    HK_CHECK_ARG(K != Ksorted, "hk_sort_bins: in-place sort is not supported");
    HK_CHECK_ARG(seg[0] >= 0 && seg[nseg] <= nx, "hk_sort_bins: segments outside [0, nx)");
    for (int b = 0; b < nseg; ++b)
        HK_CHECK_ARG(seg[b + 1] >= seg[b], "hk_sort_bins: segment %d has negative length", b);
    SortPlan sp;
    SortLayout sl;
    size_t need = 0;
    if (int e = sort_bins_plan(nx, seg, nseg, sp, sl, need)) return e;
    if (ws_bytes < need) {
        set_error("hk_sort_bins: workspace %zu bytes, %zu needed", ws_bytes, need);
        return HK_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)align_up((size_t)ws, 256);
    Seg *segs = (Seg *)base;
    std::vector<Seg> hs(nseg);
    for (int b = 0; b < nseg; ++b) { hs[b].start = seg[b]; hs[b].len = seg[b + 1] - seg[b]; }
    HK_HIP(hipMemcpyAsync(segs, hs.data(), hs.size() * sizeof(Seg), hipMemcpyHostToDevice, s));
    HK_HIP(hipStreamSynchronize(s));         // hs is host memory of this call
    FitArgs fa = {};
    fa.nC = 1;
    return sort_fit_segments(K, segs, nseg, SegMap{nseg, 0}, sp, sl, base, fa, Ksorted, s);
}

extern "C" int hk_ktable_reconstruct(const double *coeff, const int32_t *nxmin, int nspec,
                                     int nbins, int Nxb, int nC, double kmin, double unit_scale,
                                     double *Ky, double *scratch, hk_stream_t stream) {
    HK_CHECK_ARG(coeff && nxmin && Ky && scratch && nspec > 0 && nbins > 0 && Nxb > 1,
                 "hk_ktable_reconstruct: bad arguments");
    HK_CHECK_ARG(nC >= 1 && nC <= kMaxNC, "hk_ktable_reconstruct: nC %d", nC);
    hipStream_t s = (hipStream_t)stream;
    const int64_t nb = (int64_t)nspec * nbins;
    hipLaunchKernelGGL(expfx_kernel, flat_grid(nb), dim3(256), 0, s, coeff, nC, Nxb, unit_scale,
                       scratch, nb);
    HK_LAUNCH_CHECK();
    hipLaunchKernelGGL(rescale_inv_kernel, flat_grid(nb), dim3(256), 0, s, scratch, nxmin, Nxb, kmin,
                       Ky, nb);
    HK_LAUNCH_CHECK();
    return HK_OK;
}
