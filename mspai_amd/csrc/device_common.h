/*
 * Shared device definitions of the MSPAI column kernels (gfx950, wave64).
 *
 * Column status bits, the argument block every column kernel receives, and
 * the wave-level primitives: the fixed 64-lane dot-product order, sorting of
 * small index sets, and binary search.
 */
#ifndef MSPAI_DEVICE_COMMON_H
#define MSPAI_DEVICE_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "introsort.h"

namespace mspai {

enum ColStatus : int {
    COL_SINGULAR = 1,
    COL_UNDERDET = 2,
    COL_NAN_RHO = 4,
    COL_FALLBACK = 8,
    COL_K_NOT_IN_I = 16,
    COL_OVERFLOW = 32,  // internal: column did not fit its bin, retried in a larger one
};

constexpr int kWave = 64;
constexpr int kNumWork = 8;

/* Device view of a CSC matrix (Compressed_Lines, src/Compressed_Lines.h:52-101). */
struct DevCsc {
    int nrows, ncols;
    const int* colptr;
    const int* rowind;
    const double* val;
};

/* Per-wave scratch layout of the generic kernel (byte offsets). */
struct ScratchLayout {
    size_t J, I, sbuf, ubuf, ahat, aqr, b, b0, r, tau, jt, rho, pairs, sel;
    size_t per_wave;
    int capJ, capI, capU, capSort;
};

/* Everything a column kernel needs; passed by value. */
struct BuildArgs {
    DevCsc A;               // C (A, or A with probing rows)
    const int* rp_ptr;      // row pattern of C (CSR)
    const int* rp_ind;
    const int* p_ptr;       // start pattern (CSC)
    const int* p_ind;
    DevCsc B;               // target; B.colptr == nullptr means identity
    const double* ajn;      // ||a_j|| as Euclidean_Norm computes it
    double epsilon;
    int maxnew, ns, use_mean;
    int cap;                // capacity of one output column
    const int* queue;       // columns to build in this launch
    int nqueue;
    int* counter;           // persistent work-queue head
    int* overflow;          // overflow list (columns that did not fit)
    int* n_overflow;
    int* m_len;
    int* m_idx;
    double* m_val;
    double* res;
    int* nsolve;
    int* status;
    double* work;
    char* scratch;
    ScratchLayout L;
};

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

/* The 64-lane reduction: partial sums are combined by the butterfly
 * p_l <- p_l + p_(l xor s), s = 32 .. 1; every lane ends with p_0. */
__device__ __forceinline__ double wave_sum_fixed(double p)
{
#pragma unroll
    for (int s = kWave / 2; s >= 1; s >>= 1) p = p + __shfl_xor(p, s, kWave);
    return p;
}

__device__ __forceinline__ double wave_min(double p)
{
#pragma unroll
    for (int s = kWave / 2; s >= 1; s >>= 1) p = fmin(p, __shfl_xor(p, s, kWave));
    return p;
}

__device__ __forceinline__ int wave_sum_int(int p)
{
#pragma unroll
    for (int s = kWave / 2; s >= 1; s >>= 1) p += __shfl_xor(p, s, kWave);
    return p;
}

__device__ __forceinline__ int lower_bound_i(const int* a, int n, int key)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* Sort buf[0..count) ascending (bitonic network over the next power of two,
 * padded with INT_MAX; buf must hold that many ints) and write the distinct
 * values to out.  Whole-wave call; returns the number of distinct values. */
__device__ inline int wave_sort_unique(int* buf, int count, int* out)
{
    const int lane = lane_id();
    int N = 1;
    while (N < count) N <<= 1;
    for (int i = count + lane; i < N; i += kWave) buf[i] = 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= N; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < N; i += kWave) {
                int l = i ^ j;
                if (l > i) {
                    int a = buf[i], b = buf[l];
                    bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        buf[i] = b;
                        buf[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    int total = 0;
    for (int base = 0; base < count; base += kWave) {
        int i = base + lane;
        int v = i < count ? buf[i] : 0;
        bool keep = i < count && (i == 0 || buf[i - 1] != v);
        unsigned long long mask = __ballot(keep);
        int pos = total + __popcll(mask & ((1ull << lane) - 1ull));
        if (keep) out[pos] = v;
        total += __popcll(mask);
    }
    __syncthreads();
    return total;
}

}  // namespace mspai

#endif
