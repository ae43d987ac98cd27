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
/* Leading dimension of the Gram matrix / R of the mixed 64-lane LDS bins
 * (|J| <= 32; padded against bank conflicts). */
constexpr int kGramLd = 33;
constexpr int kNumWork = 8;
constexpr int kLaunchStat = 13;  // per-launch totals: kNumWork counters, columns, 3 near-margin counts, nnz
/* A selection decision of the FP64 path whose margin is at most kNearMargin
 * relative (a score against the mean, the maxnew cut, ||r|| against eps) is
 * counted as near: another operation order (MKL's) could decide it the
 * other way. */
constexpr double kNearMargin = 1e-12;

/* Unrolled |J| width of the register-resident QR for a lane-group width. */
__host__ __device__ constexpr int register_width(int group) { return group == 16 ? 8 : (group == 32 ? 16 : 32); }
/* A size bin runs the register-resident QR when one group holds every row
 * (|I| <= G) and |J| fits the unrolled width. */
__host__ __device__ constexpr bool register_bin(int group, int capI, int capJ)
{
    return group > 0 && capI <= group && capJ <= register_width(group);
}

/* Device view of a CSC matrix (Compressed_Lines, src/Compressed_Lines.h:52-101). */
struct DevCsc {
    int nrows, ncols;
    const int* colptr;
    const int* rowind;
    const double* val;
};

/* Per-wave working-set layout (byte offsets) and the sizes it holds.  One
 * layout per size bin; the generic layout is sized for the worst case. */
struct ScratchLayout {
    size_t J, I, sbuf, aqr, b, b0, r, tau, jt, rho, pairs, sel, stack, bq32, g, y, z;
    size_t srow, spos, iold, sposold, jold, colid, reflrows, bperm, yq;
    size_t segs, segoff, sid, wsum, aq32, rq, a0;
    size_t per_wave;
    int keep_a0;   // a0: an untouched copy of A(I,J) for the residual (the generic kernel's HBM QR overwrites aqr)
    int capJ;      // |J|
    int capI;      // |I| (also the leading dimension of A(I,J))
    int capU;      // distinct candidates |U|
    int capSort;   // raw index lists before sorting (a power of two)
    int capSid;    // flat entries of the columns of J (Get_I_Set's raw list, the A(I,J) gather)
    int capPairs;  // candidates on the std::sort tie path
    int ldq;       // leading dimension of aq32: capI (FP32 copy of A(I,J)), or
                   // capJ in register bins, where aq32 holds only the FP32 R, or
                   // kGramLd with `gram`
    int gram;      // mixed 64-lane LDS bin: R from the Gram matrix on the matrix
                   // cores (aq32: kGramLd x kGramLd floats, the Gram rows, then R)
    int tie_in_main;        // pairs/stack in the main region (else in BuildArgs::tie_scratch)
    int max_rounds;         // most columns per group in one chunk of the work queue
    size_t tie_per_group;   // bytes of a group's tie region when not in the main one
};

/* Routing capacities of the size bins (|J| and raw row-list length). */
struct BinEdges {
    int capJ[8];
    int capSort[8];
    int capI[8];
};

/* Everything a column kernel needs; passed by value. */
struct BuildArgs {
    DevCsc A;               // C (A, or A with probing rows)
    const int* rp_ptr;      // row pattern of C (CSR)
    const int* rp_ind;
    const double* rp_val;   // C's values in row-pattern order: candidate scores over the rows of I
                            // (nullptr: over the candidates' columns)
    int rp_maxlen;          // longest row of the pattern (groups of fewer lanes score by columns)
    const int* p_ptr;       // start pattern (CSC)
    const int* p_ind;
    const int* up_ptr;      // upper pattern (CSC, -up): candidates of column k in UP(:,k); nullptr: none
    const int* up_ind;
    DevCsc B;               // target; B.colptr == nullptr means identity
    const double* ajn;      // ||a_j|| as Euclidean_Norm computes it
    double epsilon;
    int maxnew, ns, use_mean;
    int cap;                // capacity of one output column
    const int* queue;       // columns to build in this launch
    int nqueue;
    int* counter;           // persistent work-queue head
    int* overflow;          // columns that did not fit this launch's layout
    int* n_overflow;        // their count
    int* fallback;          // mixed path: columns to rebuild in FP64
    int* n_fallback;
    int status_or;          // status bits added to every column of this launch
    int save_state;         // a column that outgrows this launch's layout leaves its state (J, step,
                            // status, counters) for the next launch of the same precision to resume;
                            // 0: the next launch starts it afresh (another precision follows)
    const double* m_in;     // adjoint: values of M (pattern in p_ptr/p_ind)
    const double* gM;       // adjoint: dL/dM, indexed like m_in
    double* gC;             // adjoint: dL/dC, indexed like A.val (accumulated)
    double* gB;             // adjoint: dL/dB, indexed like B.val (may be null)
    int cache_slots;        // pattern->R cache: slots (power of two), 0 = off
    int cache_capI, cache_capJ;
    int* cache_state;       // 0 empty, 1 being written, 2 ready
    uint32_t* cache_key;
    int* cache_m;
    int* cache_n;
    int* cache_status;
    double* cache_data;
    int* cache_hits;        // number of solves served from the cache
    int* m_len;
    int* m_idx;
    double* m_val;
    double* res;
    int* nsolve;
    int* status;
    double* work;
    double* lwork;          // builds with counters: totals of this launch's finished columns
                            // (kLaunchStat doubles: the 8 work counters, columns, near-margin
                            // decisions at the mean, at the maxnew cut, at epsilon, nnz of M)
    char* scratch;          // per-wave HBM working sets of the generic kernel
    char* tie_scratch;      // per-group HBM tie-path buffers of the LDS bins
};

/* Block SPAI (-bs != 1): the block matrix, its patterns and the outputs.
 * Block e of block column j is stored column-major (bsize[rowind[e]] x
 * bsize[j]) at val[voff[e]]. */
struct BlockArgs {
    int nb;                   // block columns (= block rows)
    const int* bsize;         // nb
    const int* bstart;        // nb + 1 scalar offsets
    const int* colptr;        // block CSC
    const int* rowind;
    const long long* voff;
    const double* val;
    const int* rp_ptr;        // block row pattern (CSR)
    const int* rp_ind;
    const int* p_ptr;         // start pattern over block columns (CSC)
    const int* p_ind;
    const double* sqinv;      // (A_j^T A_j)^-1, column-major, at sqoff[j]
    const long long* sqoff;
    double epsilon;
    int maxnew, ns, use_mean;
    int capJ;                 // blocks per output column
    long long capX;           // values per output column
    int* m_len;               // blocks of M's block column
    int* m_idx;               // their block rows, nb x capJ
    double* m_val;            // their values, nb x capX (Get_Solution_Vals_Block order)
    double* res;
    int* nsolve;
    int* status;
    int* head;                // work counter of the persistent kernel
    int* error;               // set when a column outgrew the working set
    char* scratch;            // per-wave working sets
};

/* Per-wave working set of the block kernel (byte offsets) and its sizes. */
struct BlockLayout {
    size_t J, I, sbuf, jt, fr, fc, sel, stack, pairs, rho, ahat, aqr, b, b0, r, rb, x, nbuf;
    size_t per_wave;
    int capJ;     // blocks in J
    int capI;     // blocks in I
    int capIs;    // scalar rows of A(I,J) (its leading dimension)
    int capJs;    // scalar columns of A(I,J)
    int capU;     // candidates
    int capSort;  // raw index lists before sorting (a power of two)
    int maxbs;    // largest block size
};

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

/*
 * Lane groups.  A wave runs 64 / G columns at once, one per group of G lanes
 * (G = 16, 32 or 64).  Row r of a column lives on lane r % G of its group.
 * For |I| <= G this reproduces the 64-lane order exactly: the 64-lane
 * butterfly's first steps (offsets >= G) only add +0 from lanes that hold no
 * row, which leaves every nonzero partial unchanged.  Larger columns run with
 * G = 64.
 */
template <int G>
struct Group {
    static_assert(G == 16 || G == 32 || G == 64, "group width");
    static __device__ __forceinline__ int lane() { return threadIdx.x & (G - 1); }
    static __device__ __forceinline__ int index() { return (threadIdx.x & (kWave - 1)) / G; }
    /* value of lane (l xor S) of the same group.  DPP within a row of 16:
     * quad_perm [1,0,3,2] (S = 1), [2,3,0,1] (S = 2), row_half_mirror then
     * quad_perm [3,2,1,0] (S = 4: 7 - (l xor 3) = l xor 4 within 8 lanes),
     * row_ror:8 (S = 8); ds_swizzle xor mode within 32 lanes (S = 16);
     * ds_bpermute across the halves (S = 32). */
    template <int S>
    static __device__ __forceinline__ int xor_lane(int v)
    {
        if (S == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);
        if (S == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);
        if (S == 4) return __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false), 0x1B, 0xF,
                                                    0xF, false);
        if (S == 8) return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, false);
        if (S == 16) return __builtin_amdgcn_ds_swizzle(v, 0x401F);
        return __shfl_xor(v, S, kWave);
    }
    template <int S>
    static __device__ __forceinline__ double xor_lane(double v)
    {
        long long b = __double_as_longlong(v);
        int lo = xor_lane<S>((int)(b & 0xffffffffll));
        int hi = xor_lane<S>((int)(b >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
    /* p_l + p_(l xor S) for S = 32 and 16 with the gfx950 lane swaps:
     * v_permlane32_swap of two copies of p leaves p_l in one and p_(l xor 32)
     * in the other (lanes below 32), or the same pair the other way round
     * (lanes above); v_permlane16_swap does the same for the 16-lane rows.
     * Both are VALU moves, not LDS-pipe permutes. */
    template <int S>
    static __device__ __forceinline__ double swap_add(double p)
    {
        const long long bits = __double_as_longlong(p);
        const unsigned lo = (unsigned)bits, hi = (unsigned)(bits >> 32);
        unsigned x0, x1, y0, y1;
        if (S == 32) {
            auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
            auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
            x0 = l[0], x1 = h[0], y0 = l[1], y1 = h[1];
        } else {
            auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
            auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
            x0 = l[0], x1 = h[0], y0 = l[1], y1 = h[1];
        }
        const double x = __longlong_as_double((long long)(((unsigned long long)x1 << 32) | x0));
        const double y = __longlong_as_double((long long)(((unsigned long long)y1 << 32) | y0));
        return x + y;
    }
    template <int S>
    static __device__ __forceinline__ double sum_step(double p)
    {
        if (S >= G) return p;
        if (S >= 16) return swap_add<S>(p);
        return p + xor_lane<S>(p);
    }
    /* the fixed-order sum: p_l <- p_l + p_(l xor s), s = G/2 .. 1 */
    static __device__ __forceinline__ double sum_fixed(double p)
    {
        p = sum_step<32>(p);
        p = sum_step<16>(p);
        p = sum_step<8>(p);
        p = sum_step<4>(p);
        p = sum_step<2>(p);
        p = sum_step<1>(p);
        return p;
    }
    /* K independent fixed-order sums, step by step so that their moves and
     * adds interleave */
    template <int S, int K>
    static __device__ __forceinline__ void sum_step_n(double (&p)[K])
    {
#pragma unroll
        for (int k = 0; k < K; ++k) p[k] = sum_step<S>(p[k]);
    }
    template <int K>
    static __device__ __forceinline__ void sum_fixed_n(double (&p)[K])
    {
        sum_step_n<32>(p);
        sum_step_n<16>(p);
        sum_step_n<8>(p);
        sum_step_n<4>(p);
        sum_step_n<2>(p);
        sum_step_n<1>(p);
    }
    /* one v_permlane{16,32}_swap on the two halves of a pair of doubles */
    template <int S>
    static __device__ __forceinline__ void swap_pair(double& a, double& b)
    {
        const long long ab = __double_as_longlong(a), bb = __double_as_longlong(b);
        unsigned al = (unsigned)ab, ah = (unsigned)(ab >> 32), bl = (unsigned)bb, bh = (unsigned)(bb >> 32);
        if (S == 32) {
            auto l = __builtin_amdgcn_permlane32_swap(al, bl, false, false);
            auto h = __builtin_amdgcn_permlane32_swap(ah, bh, false, false);
            al = l[0], bl = l[1], ah = h[0], bh = h[1];
        } else {
            auto l = __builtin_amdgcn_permlane16_swap(al, bl, false, false);
            auto h = __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
            al = l[0], bl = l[1], ah = h[0], bh = h[1];
        }
        a = __longlong_as_double((long long)(((unsigned long long)ah << 32) | al));
        b = __longlong_as_double((long long)(((unsigned long long)bh << 32) | bl));
    }
    /* Four fixed-order sums over a 64-lane group, every lane ending with all
     * four.  The xor-32 and xor-16 steps of the four butterflies share their
     * lane swaps: swapping p0 with p2 leaves lanes 0-31 the pair (p0_l,
     * p0_(l+32)) and lanes 32-63 the pair (p2_(l-32), p2_l), so one add gives
     * the xor-32 step of p0 in the lower and of p2 in the upper half; the
     * same for p1/p3, then a 16-lane swap of the two halves leaves row r of
     * 16 lanes with the xor-16 step of p_r.  The last four steps run on that
     * one value; two rounds of swaps give every lane all four results.  Each
     * sum is formed from the same pairs as sum_fixed (additions commute). */
    static __device__ __forceinline__ void sum4_bcast(double (&p)[4])
    {
        static_assert(G == 64, "sum4_bcast needs the whole wave");
        double a = p[0], c = p[2], b = p[1], d = p[3];
        swap_pair<32>(a, c);
        swap_pair<32>(b, d);
        double qa = a + c, qb = b + d;
        swap_pair<16>(qa, qb);
        double q = qa + qb;
        q = sum_step<8>(q);
        q = sum_step<4>(q);
        q = sum_step<2>(q);
        q = sum_step<1>(q);
        double u = q, v = q;
        swap_pair<16>(u, v);  // u: rows 0/2 results, v: rows 1/3
        double u2 = u, v2 = v;
        swap_pair<32>(u, u2);  // u: p0 everywhere, u2: p2 everywhere
        swap_pair<32>(v, v2);  // v: p1, v2: p3
        p[0] = u, p[1] = v, p[2] = u2, p[3] = v2;
    }
    /* Two fixed-order sums over a 32-lane group (xor-16 step by one shared
     * 16-lane swap), every lane ending with both. */
    static __device__ __forceinline__ void sum2_bcast(double& p0, double& p1)
    {
        static_assert(G == 32, "sum2_bcast is for 32-lane groups");
        double a = p0, b = p1;
        swap_pair<16>(a, b);
        double q = a + b;
        q = sum_step<8>(q);
        q = sum_step<4>(q);
        q = sum_step<2>(q);
        q = sum_step<1>(q);
        double u = q, v = q;
        swap_pair<16>(u, v);
        p0 = u, p1 = v;
    }
    /* K sums (K a multiple of 4), every lane ending with all of them */
    template <int K>
    static __device__ __forceinline__ void sum_bcast_n(double (&p)[K])
    {
        if constexpr (G == 64) {
#pragma unroll
            for (int k = 0; k < K; k += 4) {
                double t[4] = {p[k], p[k + 1], p[k + 2], p[k + 3]};
                sum4_bcast(t);
                p[k] = t[0], p[k + 1] = t[1], p[k + 2] = t[2], p[k + 3] = t[3];
            }
        } else if constexpr (G == 32) {
#pragma unroll
            for (int k = 0; k < K; k += 2) sum2_bcast(p[k], p[k + 1]);
        } else {
            sum_fixed_n(p);
        }
    }
    /* value of lane `src` of the group (src the same in every active lane
     * of the wave) */
    static __device__ __forceinline__ double lane_value(double v, int src)
    {
        const long long bits = __double_as_longlong(v);
        const unsigned lo = (unsigned)lane_value_i((int)bits, src);
        const unsigned hi = (unsigned)lane_value_i((int)(bits >> 32), src);
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    }
    /* 32-bit version: for whole waves and 32-lane groups one readlane per
     * group, then each lane keeps its own group's value -- no LDS-pipe
     * permute.  src must be the same in every active lane of the wave (a
     * loop counter that all groups step together), not only within the
     * group. */
    static __device__ __forceinline__ int lane_value_i(int v, int src)
    {
        if (G == kWave) return __builtin_amdgcn_readlane(v, src);
        if (G == 32) {
            const int a = __builtin_amdgcn_readlane(v, src), b = __builtin_amdgcn_readlane(v, src + 32);
            return (threadIdx.x & 32) ? b : a;
        }
        /* 16-lane groups: four readlanes and a select cost more than the
         * permute (measured) */
        return __shfl(v, src, G);
    }
    template <int S>
    static __device__ __forceinline__ float sum_step_f(float p)
    {
        if (S >= G) return p;
        const int bits = __float_as_int(p);
        if (S >= 16) {
            auto r = S == 32 ? __builtin_amdgcn_permlane32_swap(bits, bits, false, false)
                             : __builtin_amdgcn_permlane16_swap(bits, bits, false, false);
            return __int_as_float((int)r[0]) + __int_as_float((int)r[1]);
        }
        return p + __int_as_float(xor_lane<S>(bits));
    }
    static __device__ __forceinline__ float sum_f(float p)
    {
        p = sum_step_f<32>(p);
        p = sum_step_f<16>(p);
        p = sum_step_f<8>(p);
        p = sum_step_f<4>(p);
        p = sum_step_f<2>(p);
        p = sum_step_f<1>(p);
        return p;
    }
    /* one v_permlane{16,32}_swap on a pair of floats (see swap_pair) */
    template <int S>
    static __device__ __forceinline__ void swap_pair_f(float& a, float& b)
    {
        const int ai = __float_as_int(a), bi = __float_as_int(b);
        auto r = S == 32 ? __builtin_amdgcn_permlane32_swap(ai, bi, false, false)
                         : __builtin_amdgcn_permlane16_swap(ai, bi, false, false);
        a = __int_as_float((int)r[0]);
        b = __int_as_float((int)r[1]);
    }
    /* Four FP32 sums with every lane ending with all four, the lane swaps of
     * the xor-32 and xor-16 steps shared as in sum4_bcast (64-lane groups),
     * or two pairs sharing the xor-16 swap (32-lane groups); 16-lane groups
     * run four butterflies.  The summation order differs from sum_f's (mixed
     * path only, where it is free). */
    static __device__ __forceinline__ void sum4_bcast_f(float (&p)[4])
    {
        if constexpr (G == 64) {
            float a = p[0], c = p[2], b = p[1], d = p[3];
            swap_pair_f<32>(a, c);
            swap_pair_f<32>(b, d);
            float qa = a + c, qb = b + d;
            swap_pair_f<16>(qa, qb);
            float q = qa + qb;
            q = sum_step_f<8>(q);
            q = sum_step_f<4>(q);
            q = sum_step_f<2>(q);
            q = sum_step_f<1>(q);
            float u = q, v = q;
            swap_pair_f<16>(u, v);  // u: rows 0/2 results, v: rows 1/3
            float u2 = u, v2 = v;
            swap_pair_f<32>(u, u2);  // u: p0 everywhere, u2: p2 everywhere
            swap_pair_f<32>(v, v2);  // v: p1, v2: p3
            p[0] = u, p[1] = v, p[2] = u2, p[3] = v2;
        } else if constexpr (G == 32) {
#pragma unroll
            for (int k = 0; k < 4; k += 2) {
                float a = p[k], b = p[k + 1];
                swap_pair_f<16>(a, b);
                float q = a + b;
                q = sum_step_f<8>(q);
                q = sum_step_f<4>(q);
                q = sum_step_f<2>(q);
                q = sum_step_f<1>(q);
                float u = q, v = q;
                swap_pair_f<16>(u, v);
                p[k] = u, p[k + 1] = v;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) p[k] = sum_f(p[k]);
        }
    }
    /* 32-bit value of lane (l xor S), S < G, without LDS-pipe permutes:
     * DPP for S < 16, a lane swap for 16 and 32 (of the two swap results the
     * partner's value is the one from the other row / half) */
    template <int S>
    static __device__ __forceinline__ int partner(int v)
    {
        if (S < 16) return xor_lane<S>(v);
        if (S == 16) {
            auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
            return (threadIdx.x & 16) ? (int)r[0] : (int)r[1];
        }
        auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (threadIdx.x & 32) ? (int)r[0] : (int)r[1];
    }
    template <int S>
    static __device__ __forceinline__ double partner(double v)
    {
        const long long b = __double_as_longlong(v);
        const int lo = partner<S>((int)b), hi = partner<S>((int)(b >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
    template <int S>
    static __device__ __forceinline__ float partner(float v)
    {
        return __int_as_float(partner<S>(__float_as_int(v)));
    }
    /* all-reduce with an exact, order-free operation (min, max, integer add) */
    template <class T, class Op>
    static __device__ __forceinline__ T allreduce(T p, Op op)
    {
        if (G > 32) p = op(p, partner<32>(p));
        if (G > 16) p = op(p, partner<16>(p));
        p = op(p, partner<8>(p));
        p = op(p, partner<4>(p));
        p = op(p, partner<2>(p));
        p = op(p, partner<1>(p));
        return p;
    }
    static __device__ __forceinline__ double min(double p)
    {
        return allreduce(p, [](double x, double y) { return fmin(x, y); });
    }
    static __device__ __forceinline__ double max(double p)
    {
        return allreduce(p, [](double x, double y) { return fmax(x, y); });
    }
    static __device__ __forceinline__ float minf(float p)
    {
        return allreduce(p, [](float x, float y) { return fminf(x, y); });
    }
    static __device__ __forceinline__ float maxf(float p)
    {
        return allreduce(p, [](float x, float y) { return fmaxf(x, y); });
    }
    static __device__ __forceinline__ int sum_int(int p)
    {
        return allreduce(p, [](int x, int y) { return x + y; });
    }
    static __device__ __forceinline__ int min_int(int p)
    {
        return allreduce(p, [](int x, int y) { return x < y ? x : y; });
    }
    static __device__ __forceinline__ int max_int(int p)
    {
        return allreduce(p, [](int x, int y) { return x > y ? x : y; });
    }
    /* 32-bit value of group lane `src` (a compile-time lane): readlane for
     * a whole wave, DPP row_newbcast (lane SRC of each row of 16, a VALU
     * move) for 16-lane groups, ds_swizzle's broadcast mode for 32-lane
     * groups */
    template <int SRC>
    static __device__ __forceinline__ int lane_bcast(int v)
    {
        static_assert(SRC >= 0 && SRC < G, "source lane");
        if constexpr (G == kWave) {
            return __builtin_amdgcn_readlane(v, SRC);
        } else if constexpr (G == 16) {
            return __builtin_amdgcn_mov_dpp(v, 0x150 + SRC, 0xF, 0xF, false);
        } else {
            /* lane l reads ((l & and) | or) within its 32-lane half */
            return __builtin_amdgcn_ds_swizzle(v, SRC << 5);
        }
    }
    template <int SRC>
    static __device__ __forceinline__ double lane_bcast(double v)
    {
        const long long b = __double_as_longlong(v);
        const int lo = lane_bcast<SRC>((int)b), hi = lane_bcast<SRC>((int)(b >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
    /* value of lane 0 of the group */
    static __device__ __forceinline__ int bcast0(int v) { return lane_bcast<0>(v); }
    static __device__ __forceinline__ double bcast0(double v) { return lane_bcast<0>(v); }
    /* value of the group's last lane */
    static __device__ __forceinline__ int bcast_last(int v) { return lane_bcast<G - 1>(v); }
    /* inclusive prefix sum over the group: DPP row shifts inside each row of
     * 16 lanes, then row_bcast:15 / row_bcast:31 carry the row totals on */
    static __device__ __forceinline__ int scan_incl(int v)
    {
        v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
        v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
        v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
        v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
        if (G > 16) v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // rows 1,3 += lane 15, 47
        if (G > 32) v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // rows 2,3 += lane 31
        return v;
    }
    /* ballot restricted to the group, bit i = group lane i */
    static __device__ __forceinline__ unsigned long long ballot(bool pred)
    {
        unsigned long long m = __ballot(pred);
        if (G == 64) return m;
        return (m >> (index() * G)) & ((1ull << G) - 1ull);
    }
    static __device__ __forceinline__ unsigned long long lanes_below() { return (1ull << lane()) - 1ull; }
};

/* Backwards-compatible whole-wave helpers. */
__device__ __forceinline__ double wave_sum_fixed(double p) { return Group<64>::sum_fixed(p); }
__device__ __forceinline__ double wave_min(double p) { return Group<64>::min(p); }
__device__ __forceinline__ int wave_sum_int(int p) { return Group<64>::sum_int(p); }

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
 * values to out (out may equal buf).  Whole-group call; the wave must be
 * converged at the barriers.  Returns the number of distinct values. */
/* Sorted distinct values of buf[0..count) into out (out may be buf).
 * For a whole wave (G = 64) and up to 8 values per lane the bitonic network
 * runs in registers: element i = k*64 + lane sits in slot k of its lane;
 * partners at distance j < 64 are reached with DPP / lane swaps, larger
 * distances pair slots of the same lane.  Otherwise (16/32-lane groups, where
 * the LDS network measured faster, or longer lists) the same network runs
 * through buf in LDS.  Either way the result is written back to buf and
 * compacted by ballot. */
template <int G, int J>
__device__ __forceinline__ void bitonic_lanes(int (&x)[8], int kslots, int size)
{
    const int l = Group<G>::lane();
    const bool low = (l & J) == 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (k < kslots) {
            const int i = k * G + l;
            const bool up = (i & size) == 0;
            const int p = Group<G>::template partner<J>(x[k]);
            x[k] = (low == up) ? min(x[k], p) : max(x[k], p);
        }
    }
}

template <int G>
__device__ __forceinline__ void bitonic_slots(int (&x)[8], int kslots, int size, int js)
{
    const int l = Group<G>::lane();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
#pragma unroll
        for (int d = 1; d < 8; d <<= 1) {
            if (d == js && (k & d) == 0 && (k | d) < kslots) {
                const int i = k * G + l;
                const bool up = (i & size) == 0;
                const int a = x[k], b = x[k | d];
                x[k] = up ? min(a, b) : max(a, b);
                x[k | d] = up ? max(a, b) : min(a, b);
            }
        }
    }
}

template <int G>
__device__ __forceinline__ int group_sort_unique(int* buf, int count, int* out)
{
    using Gr = Group<G>;
    const int l = Gr::lane();
    int N = 1;
    while (N < count) N <<= 1;
    if (G == kWave && N <= 8 * G) {
        const int kslots = (N + G - 1) / G;
        int x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = (k < kslots && k * G + l < count) ? buf[k * G + l] : 0x7fffffff;
        for (int size = 2; size <= N; size <<= 1) {
            for (int j = size >> 1; j > 0; j >>= 1) {
                if (j >= G) {
                    bitonic_slots<G>(x, kslots, size, j / G);
                } else {
                    switch (j) {
                    case 1: bitonic_lanes<G, 1>(x, kslots, size); break;
                    case 2: bitonic_lanes<G, 2>(x, kslots, size); break;
                    case 4: bitonic_lanes<G, 4>(x, kslots, size); break;
                    case 8: bitonic_lanes<G, 8>(x, kslots, size); break;
                    case 16: if (G > 16) bitonic_lanes<G, (G > 16 ? 16 : 1)>(x, kslots, size); break;
                    default: if (G > 32) bitonic_lanes<G, (G > 32 ? 32 : 1)>(x, kslots, size); break;
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < kslots && k * G + l < count) buf[k * G + l] = x[k];
    } else {
        for (int i = count + l; i < N; i += G) buf[i] = 0x7fffffff;
        __syncthreads();
        for (int k = 2; k <= N; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = l; i < N; i += G) {
                    int p = i ^ j;
                    if (p > i) {
                        int a = buf[i], b = buf[p];
                        bool up = (i & k) == 0;
                        if ((a > b) == up) {
                            buf[i] = b;
                            buf[p] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    int total = 0;
    for (int base = 0; base < count; base += G) {
        int i = base + l;
        int v = i < count ? buf[i] : 0;
        bool keep = i < count && (i == 0 || buf[i - 1] != v);
        unsigned long long mask = Gr::ballot(keep);
        int pos = total + __popcll(mask & Gr::lanes_below());
        if (keep) out[pos] = v;
        total += __popcll(mask);
    }
    __syncthreads();
    return total;
}

__device__ inline int wave_sort_unique(int* buf, int count, int* out) { return group_sort_unique<64>(buf, count, out); }

/* c <- (I - tau v v^T) c over rows i..m-1, v_i = 1 (dlarf('L')). */
template <int G>
__device__ __forceinline__ void apply_reflector(const double* v, double* c, int i, int m, double tau)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    double p = 0.0;
    for (int r = lane; r < m; r += G) {
        if (r < i) continue;
        double vr = (r == i) ? 1.0 : v[r];
        p = __fma_rn(vr, c[r], p);
    }
    double w = Gr::sum_fixed(p);
    double t = (-tau) * w;
    for (int r = lane; r < m; r += G) {
        if (r < i) continue;
        double vr = (r == i) ? 1.0 : v[r];
        c[r] = __fma_rn(vr, t, c[r]);
    }
    __syncthreads();
}

/* Householder vector of storage column c over storage rows [c, rows)
 * (dlarfg in the fixed order).  Returns tau. */
template <int G>
__device__ __forceinline__ double make_reflector(double* ai, int c, int rows)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    double tau = 0.0;
    if (c + 1 < rows) {
        double alpha = ai[c];
        double p = 0.0;
        for (int r = lane; r < rows; r += G)
            if (r > c) p = __fma_rn(ai[r], ai[r], p);
        double xn2 = Gr::sum_fixed(p);
        if (xn2 != 0.0) {
            double nrm = sqrt(__fma_rn(alpha, alpha, xn2));
            double beta = -copysign(nrm, alpha);
            tau = (beta - alpha) / beta;
            double scal = 1.0 / (alpha - beta);
            __syncthreads();
            for (int r = lane; r < rows; r += G)
                if (r > c) ai[r] = ai[r] * scal;
            if (lane == 0) ai[c] = beta;
            __syncthreads();
        }
    }
    return tau;
}


}  // namespace mspai

#endif
