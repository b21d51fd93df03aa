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
//     merges them into registers, and writes them back after a barrier.
//
// Only comparisons move values, so the sorted values are the input values;
// equal values are interchangeable (ties need no order).  Positions past n
// are filled with +inf, which sort last.
#pragma once
#include <hip/hip_runtime.h>

namespace hk {

__device__ __forceinline__ int pad_index(int i) { return i + (i >> 4); }

// LDS doubles a sort of capacity n needs.
constexpr int lds_sort_slots(int n) { return n + n / 16; }

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

// Sorts the n values at s[pad_index(i)], i < n <= NT*IPT, in place.  Every
// thread of the workgroup (blockDim.x == NT) must call it; it begins and
// ends with a barrier.
template <int NT, int IPT>
__device__ void lds_sort(double *s, int n) {
    static_assert(IPT == 16, "pad_index assumes 16 values per thread");
    const int t = threadIdx.x;
    const int np = (n + IPT - 1) / IPT * IPT;
    const int d0 = t * IPT;
    const bool act = d0 < np;
    double v[IPT];
    __syncthreads();
    if (act) {
#pragma unroll
        for (int q = 0; q < IPT; ++q) v[q] = d0 + q < n ? s[pad_index(d0 + q)] : INFINITY;
        sort_regs<IPT>(v);
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int q = 0; q < IPT; ++q) s[pad_index(d0 + q)] = v[q];
    }
    for (int L = IPT; L < np; L <<= 1) {
        __syncthreads();
        if (act) {
            const int ps = d0 / (2 * L) * (2 * L);
            const int na = min(L, np - ps);
            const int nb = min(L, np - ps - na);
            const int diag = d0 - ps;
            const int a0 = ps, b0 = ps + na;
            int lo = max(0, diag - nb), hi = min(diag, na);
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s[pad_index(a0 + mid)] <= s[pad_index(b0 + diag - mid - 1)]) lo = mid + 1;
                else hi = mid;
            }
            int a = lo, b = diag - lo;
            double va = a < na ? s[pad_index(a0 + a)] : INFINITY;
            double vb = b < nb ? s[pad_index(b0 + b)] : INFINITY;
#pragma unroll
            for (int q = 0; q < IPT; ++q) {
                const bool ta = b >= nb || (a < na && va <= vb);
                v[q] = ta ? va : vb;
                if (ta) { ++a; va = a < na ? s[pad_index(a0 + a)] : INFINITY; }
                else { ++b; vb = b < nb ? s[pad_index(b0 + b)] : INFINITY; }
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < IPT; ++q) s[pad_index(d0 + q)] = v[q];
        }
    }
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
