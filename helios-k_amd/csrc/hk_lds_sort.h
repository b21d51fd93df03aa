// hk_lds_sort.h -- ascending sort of up to NT*IPT doubles held in LDS by one
// workgroup (block merge sort), shared by the per-bin sorts.
//
//   * values live at s[pad_index(i)]: one spare slot after every 16 values,
//     so the IPT = 16 values a thread owns start 17 slots after its
//     neighbour's and the blocked reads and writes spread over the banks;
//   * each thread loads IPT consecutive values into registers and sorts them
//     with a bitonic network (fully unrolled, registers only);
//   * then log2(n/IPT) merge rounds: a thread finds where its IPT outputs of
//     one merged pair of runs start by a merge-path binary search in LDS,
//     forms them in registers, and writes them back after a barrier.  Its
//     outputs are the IPT smallest of the next IPT values of each run: the
//     two windows are loaded at once (no load waits on a comparison) and a
//     bitonic half-cleaner plus a 4-level network orders them.
//
// Only comparisons move values, so the sorted values are the input values;
// equal values are interchangeable (ties need no order).  Positions past n
// are filled with +inf, which sort last.  A bin holding negative zeros
// (for which min/max instructions may pick either operand) is sorted with
// compares and selects and a value-by-value merge instead, so -0 and +0
// keep their bits.  NaNs have no place in the order the merge-path splits
// rely on: a bin holding them is not sorted, and its values may repeat or
// go missing.
#pragma once
#include <hip/hip_runtime.h>

namespace hk {

__device__ __forceinline__ int pad_index(int i) { return i + (i >> 4); }

// LDS doubles a sort of capacity n needs.
// (+ 32: the merge windows read up to 17 slots past a run's last value)
constexpr int lds_sort_slots(int n) { return n + n / 16 + 32; }

template <int N>
__device__ __forceinline__ void sort_regs(double (&v)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const double a = v[i], b = v[l];
                    const double lo = a < b ? a : b, hi = a < b ? b : a;
                    if ((i & k) == 0) { v[i] = lo; v[l] = hi; }
                    else { v[i] = hi; v[l] = lo; }
                }
            }
        }
    }
}

// sort_regs with min/max instructions (no NaNs, no negative zeros).
template <int N>
__device__ __forceinline__ void sort_regs_minmax(double (&v)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const double lo = fmin(v[i], v[l]), hi = fmax(v[i], v[l]);
                    if ((i & k) == 0) { v[i] = lo; v[l] = hi; }
                    else { v[i] = hi; v[l] = lo; }
                }
            }
        }
    }
}

// Merge-path split of one merged pair: the number of outputs before
// output diag that come from run A (na values at a0) rather than run B (nb
// values at b0); ties go to A first.
__device__ __forceinline__ int merge_path(const double *s, int a0, int na, int b0, int nb,
                                          int diag) {
    int lo = max(0, diag - nb), hi = min(diag, na);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s[pad_index(a0 + mid)] <= s[pad_index(b0 + diag - mid - 1)]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// IPT consecutive values of a run from logical index i0 on, +inf from
// index i0 + m on (m = values left in the run).  The padded positions of
// 16 consecutive indices cross at most one pad slot, after 16 - (i0 & 15)
// values, so each load is the window's base plus a constant offset.
template <int IPT>
__device__ __forceinline__ void load_window(const double *s, int i0, int m, double (&w)[IPT]) {
    const double *p = s + pad_index(i0);
    const int r = 16 - (i0 & 15);
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
        const double x = q < r ? p[q] : p[q + 1];
        w[q] = q < m ? x : INFINITY;
    }
}

template <int NT, int IPT, bool FAST>
__device__ __forceinline__ void lds_sort_run(double *s, int np, double (&v)[IPT], bool act,
                                             int d0, int L0) {
    if (act) {
        if (L0 <= IPT) {
            if (FAST) sort_regs_minmax<IPT>(v);
            else sort_regs<IPT>(v);
        }
#pragma unroll
        for (int q = 0; q < IPT; ++q) s[pad_index(d0 + q)] = v[q];
    }
    for (int L = max(L0, IPT); L < np; L <<= 1) {
        __syncthreads();
        if (act) {
            const int ps = d0 / (2 * L) * (2 * L);
            const int na = min(L, np - ps);
            const int nb = min(L, np - ps - na);
            const int diag = d0 - ps;
            const int a0 = ps, b0 = ps + na;
            const int a = merge_path(s, a0, na, b0, nb, diag), b = diag - a;
            if (FAST) {
                // this thread's IPT outputs are the IPT smallest of the next
                // IPT values of each run: the half-cleaner of a bitonic merge
                // (A ascending against B descending) picks them, and its
                // sorting network orders them
                double wa[IPT], wb[IPT];
                load_window<IPT>(s, a0 + a, na - a, wa);
                load_window<IPT>(s, b0 + b, nb - b, wb);
#pragma unroll
                for (int q = 0; q < IPT; ++q) v[q] = fmin(wa[q], wb[IPT - 1 - q]);
#pragma unroll
                for (int j = IPT / 2; j > 0; j >>= 1) {
#pragma unroll
                    for (int i = 0; i < IPT; ++i) {
                        if ((i & j) == 0) {
                            const double lo = fmin(v[i], v[i + j]), hi = fmax(v[i], v[i + j]);
                            v[i] = lo;
                            v[i + j] = hi;
                        }
                    }
                }
            } else {
                int ia = a, ib = b;
                double va = ia < na ? s[pad_index(a0 + ia)] : INFINITY;
                double vb = ib < nb ? s[pad_index(b0 + ib)] : INFINITY;
#pragma unroll
                for (int q = 0; q < IPT; ++q) {
                    const bool ta = ib >= nb || (ia < na && va <= vb);
                    v[q] = ta ? va : vb;
                    if (ta) { ++ia; va = ia < na ? s[pad_index(a0 + ia)] : INFINITY; }
                    else { ++ib; vb = ib < nb ? s[pad_index(b0 + ib)] : INFINITY; }
                }
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < IPT; ++q) s[pad_index(d0 + q)] = v[q];
        }
    }
}

// Sorts the n values at s[pad_index(i)], i < n <= NT*IPT, in place.  Every
// thread of the workgroup (blockDim.x == NT) must call it; it begins and
// ends with a barrier.  L0 (a power of two >= IPT): the values already form
// sorted runs [r L0, (r + 1) L0), so the merge rounds start there.  Without NaNs or negative zeros (checked here, once
// per call), compare-exchanges are min/max instructions and each merge
// output block comes from a bitonic half-cleaner; otherwise every step
// compares and selects and the merge walks the runs one value at a time.
template <int NT, int IPT>
__device__ __forceinline__ void lds_sort(double *s, int n, int L0 = IPT) {
    static_assert(IPT == 16, "pad_index assumes 16 values per thread");
    const int t = threadIdx.x;
    const int np = (n + IPT - 1) / IPT * IPT;
    const int d0 = t * IPT;
    const bool act = d0 < np;
    double v[IPT];
    __syncthreads();
    int odd = 0;
    if (act) {
#pragma unroll
        for (int q = 0; q < IPT; ++q) {
            v[q] = d0 + q < n ? s[pad_index(d0 + q)] : INFINITY;
            odd |= (v[q] != v[q]) | (v[q] == 0.0 && signbit(v[q]));
        }
    }
    if (__syncthreads_or(odd)) lds_sort_run<NT, IPT, false>(s, np, v, act, d0, L0);
    else lds_sort_run<NT, IPT, true>(s, np, v, act, d0, L0);
    __syncthreads();
}

// First index in the sorted run v[0..n) whose value is >= key (lower) or
// > key (upper); global memory.
__device__ __forceinline__ int64_t run_lower_bound(const double *v, int64_t n, double key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (v[m] < key) lo = m + 1; else hi = m;
    }
    return lo;
}

__device__ __forceinline__ int64_t run_upper_bound(const double *v, int64_t n, double key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (v[m] <= key) lo = m + 1; else hi = m;
    }
    return lo;
}

}  // namespace hk
