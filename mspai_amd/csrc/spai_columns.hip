/*
 * FP64 column kernels: one wave64 per column of M.
 *
 * The same code runs in two memory configurations.  In the binned kernel the
 * per-wave working set (J, I, A(I,J), right-hand side, residual, candidates)
 * lives in LDS, sized by the host for the common column sizes; a column that
 * outgrows those sizes is appended to an overflow queue and rebuilt by the
 * generic kernel, whose working set lives in a per-wave region of HBM sized
 * for the worst case.  Both give bitwise identical results.
 *
 * One wave runs Spai_Unrestrained<T>::SPAI_Column
 * (src/Spai_Unrestrained.imp:33-133) for one column: I-set, gather of A(I,J),
 * Householder QR with Q^T b carried as an extra column, back substitution,
 * FP64 residual, and the augmentation step (candidate union, rho scores, mean
 * threshold, top-maxnew selection).  Floating-point operations follow the
 * fixed 64-lane order (rows r on lane r % 64, partials combined by the xor
 * butterfly), which makes the result bitwise reproducible; arithmetic the
 * reference compiles itself (Sqrt_Sum, Compute_Numerator, the rho formula and
 * the mean) is sequential and unfused, as there.
 *
 * Columns come from a persistent work queue (one atomic per column), the
 * device form of Load_Balance::Grab_M_Col (src/Load_Balance.imp:34-69).
 */
#include <map>
#include <mutex>
#include <tuple>

#include "device_common.h"
#include "spai_host.h"

namespace mspai {

namespace {

struct WaveScratch {
    int* J;
    int* I;
    int* sbuf;
    double* aqr;
    double* b;
    double* b0;
    double* r;
    double* tau;
    int* jt;
    double* rho;
    RhoPair* pairs;
    int* sel;
    int* stack;
    float* aq32;  // mixed path: FP32 copy of A(I,J) for the LDS QR, or the register QR's FP32 R
    float* bq32;  // mixed path: FP32 right-hand side
    double* g;    // mixed path: A(I,J)^T r
    double* y;    // mixed path: triangular-solve work vector; backward: m
    double* z;    // backward: copy of dL/dm
    /* QR-update path: storage order of rows / columns of the kept factor */
    int* srow;     // storage row -> row index of A
    int* spos;     // position in sorted I -> storage row
    int* iold;     // previous sorted I
    int* sposold;  // previous spos
    int* jold;     // previous J
    int* colid;    // storage column -> column index of A
    int* reflrows; // rows a reflector spans (rows present when it was made)
    double* bperm; // right-hand side in storage row order
    double* yq;    // Q^T b work vector in storage row order
    int* segs;     // colptr[J[t]] of the current J
    int* segoff;   // first flat entry of column J[t]; segoff[|J|] = total
    int* sid;      // flat entry -> position t in J
    double* wsum;  // the column's work counters
    double* rq;    // adjoint: R of the register QR (capJ x capJ)
    double* a0;    // untouched copy of A(I,J), or nullptr
    unsigned* htab;  // row -> position hash of the candidate scores: always in LDS
    int hslots;      // its slots (a power of two)
};

__device__ __forceinline__ WaveScratch wave_scratch(char* ws, char* tie, const ScratchLayout& L)
{
    WaveScratch s;
    s.J = (int*)(ws + L.J);
    s.I = (int*)(ws + L.I);
    s.sbuf = (int*)(ws + L.sbuf);
    s.aqr = (double*)(ws + L.aqr);
    s.b = (double*)(ws + L.b);
    s.b0 = (double*)(ws + L.b0);
    s.r = (double*)(ws + L.r);
    s.tau = (double*)(ws + L.tau);
    s.jt = (int*)(ws + L.jt);
    s.rho = (double*)(ws + L.rho);
    s.pairs = (RhoPair*)(tie + L.pairs);
    s.sel = (int*)(ws + L.sel);
    s.stack = (int*)(tie + L.stack);
    s.aq32 = (float*)(ws + L.aq32);
    s.bq32 = (float*)(ws + L.bq32);
    s.g = (double*)(ws + L.g);
    s.y = (double*)(ws + L.y);
    s.z = (double*)(ws + L.z);
    s.srow = (int*)(ws + L.srow);
    s.spos = (int*)(ws + L.spos);
    s.iold = (int*)(ws + L.iold);
    s.sposold = (int*)(ws + L.sposold);
    s.jold = (int*)(ws + L.jold);
    s.colid = (int*)(ws + L.colid);
    s.reflrows = (int*)(ws + L.reflrows);
    s.bperm = (double*)(ws + L.bperm);
    s.yq = (double*)(ws + L.yq);
    s.segs = (int*)(ws + L.segs);
    s.segoff = (int*)(ws + L.segoff);
    s.sid = (int*)(ws + L.sid);
    s.wsum = (double*)(ws + L.wsum);
    s.rq = (double*)(ws + L.rq);
    s.a0 = L.keep_a0 ? (double*)(ws + L.a0) : nullptr;
    s.htab = nullptr;  // set by the kernel (LDS)
    s.hslots = 0;
    return s;
}

constexpr int kOverflow = -2;
constexpr int kGenericHashSlots = 2048;  // LDS of the generic kernel's score hash (8 KB per wave)
constexpr int kXcds = 8;  // MI355X: 8 XCDs, 32 CUs each
constexpr int kFallback = -3;

/* Outcome of one column in one launch. */
enum ColumnOutcome { kDone = 0, kMoveOn = 1, kRedoFp64 = 2 };

/* Mixed path guards.  A selection decision whose margin is below kGuard (on
 * the scale of ||r||^2 for the scores, of epsilon for the stop test) is not
 * trusted to the refined FP32 solution: the column is rebuilt by the FP64
 * kernels, which reproduce the oracle bit for bit.  kCondGuard bounds
 * min |R_ii| / max |R_ii| of the FP32 factor (below it, one or two
 * refinement steps are not certain to reach FP64 accuracy). */
constexpr double kGuard = 1e-10;
constexpr float kCondGuard = 1e-4f;
constexpr int kRefineSteps = 3;

/* Column segments of the current J: every lane t < |J| reads colptr[J[t]]
 * and colptr[J[t] + 1] at once, an exclusive scan gives each column's first
 * flat entry, and sid maps a flat entry back to its column.  Gathers over
 * the flat entries then issue all their loads together instead of one
 * column after another.  Returns the number of entries; sid is written only
 * when it fits (<= cap). */
template <int G>
__device__ __forceinline__ int load_segments(const BuildArgs& a, const WaveScratch& s, int n, int cap)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    int carry = 0;
    for (int base = 0; base < n; base += G) {
        const int t = base + lane;
        int start = 0, cnt = 0;
        if (t < n) {
            const int j = s.J[t];
            start = a.A.colptr[j];
            cnt = a.A.colptr[j + 1] - start;
        }
        const int incl = Gr::scan_incl(cnt);
        if (t < n) {
            s.segs[t] = start;
            s.segoff[t] = carry + incl - cnt;
        }
        carry += Gr::bcast_last(incl);
    }
    if (lane == 0) s.segoff[n] = carry;
    __syncthreads();
    if (carry <= cap) {
        for (int t = lane; t < n; t += G) {
            const int o = s.segoff[t], e = s.segoff[t + 1];
            for (int k = o; k < e; ++k) s.sid[k] = t;
        }
    }
    __syncthreads();
    return carry;
}

/* Sorted distinct values of src[0..total) that are not in excl[0..nexcl),
 * written to out, without a sort: every value lies in [lo, hi] and sets its
 * bit of a bitmap bm of at least (hi - lo) / 32 + 1 words of LDS; the
 * excluded values clear theirs; the set bits are read out in increasing
 * order.  Returns the count, or kOverflow (before any write past out[cap))
 * when it exceeds cap. */
template <int G>
__device__ __forceinline__ int bitmap_unique(const int* src, int total, int lo, int hi, const int* excl, int nexcl,
                                             int* out, int cap, unsigned* bm)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int words = ((hi - lo) >> 5) + 1;
    for (int w = lane; w < words; w += G) bm[w] = 0u;
    __syncthreads();
    for (int e = lane; e < total; e += G) {
        const int d = src[e] - lo;
        atomicOr(&bm[d >> 5], 1u << (d & 31));
    }
    __syncthreads();
    for (int t = lane; t < nexcl; t += G) {
        const int v = excl[t];
        if (v >= lo && v <= hi) atomicAnd(&bm[(v - lo) >> 5], ~(1u << ((v - lo) & 31)));
    }
    __syncthreads();
    int c = 0;
    for (int w0 = 0; w0 < words; w0 += G) {
        const int w = w0 + lane;
        unsigned bits = w < words ? bm[w] : 0u;
        const int cnt = __popc(bits);
        const int incl = Gr::scan_incl(cnt);
        const int chunk = Gr::bcast_last(incl);
        if (c + chunk > cap) return kOverflow;
        int o = c + incl - cnt;
        while (bits) {
            const int k = __ffs(bits) - 1;
            bits &= bits - 1u;
            out[o++] = lo + (w << 5) + k;
        }
        c += chunk;
    }
    __syncthreads();
    return c;
}

/* Word ranks of a bitmap: pre[w] = set bits in the words before w, stored
 * after the bitmap (words more words).  The rank of a set bit d is then
 * pre[d >> 5] + popc(bits below d in its word), its position in the sorted
 * list bitmap_unique read out. */
template <int G>
__device__ __forceinline__ void bitmap_ranks(unsigned* bm, int words)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    unsigned* const pre = bm + words;
    int run = 0;
    for (int w0 = 0; w0 < words; w0 += G) {
        const int w = w0 + lane;
        const int cnt = w < words ? __popc(bm[w]) : 0;
        const int incl = Gr::scan_incl(cnt);
        if (w < words) pre[w] = (unsigned)(run + incl - cnt);
        run += Gr::bcast_last(incl);
    }
    __syncthreads();
}

__device__ __forceinline__ int bitmap_rank(const unsigned* bm, int words, int d)
{
    const unsigned bits = bm[d >> 5], bit = 1u << (d & 31);
    return (bits & bit) ? (int)bm[words + (d >> 5)] + __popc(bits & (bit - 1u)) : -1;
}

/* LDS words for bitmap_unique: in a size bin the score region (free outside
 * the scoring, and aliased with A(I,J) only while A(I,J) is dead), in the
 * generic kernel the LDS hash region (its working set is in HBM, where
 * atomics in L2 and reads through the vector L1 would not meet). */
template <bool kHbm>
__device__ __forceinline__ unsigned* bitmap_words(const ScratchLayout& L, const WaveScratch& s, int* avail)
{
    if (kHbm) {
        *avail = s.htab ? s.hslots : 0;
        return s.htab;
    }
    *avail = 2 * L.capU;
    return reinterpret_cast<unsigned*>(s.rho);
}

/* Get_I_Set (src/Spai_Sub.imp:65-122): sorted distinct rows of the columns in
 * J (by a bitmap over their range when it fits, else a sort).  Returns |I|,
 * or kOverflow when the raw list or I exceeds the layout.  Leaves the column
 * segments of J for gather_ahat.  kHbm: the working set is in HBM. */
template <int G, bool kHbm>
__device__ __forceinline__ int build_I(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s, int n,
                                       bool* packed = nullptr)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int total = load_segments<G>(a, s, n, L.capSid);
    if (total > L.capSid) return kOverflow;
    int vmin = 0x7fffffff, vmax = -1;
    for (int e = lane; e < total; e += G) {
        const int t = s.sid[e];
        const int v = a.A.rowind[s.segs[t] + e - s.segoff[t]];
        s.sbuf[e] = v;
        vmin = min(vmin, v);
        vmax = max(vmax, v);
    }
    __syncthreads();
    vmin = Gr::min_int(vmin);
    vmax = Gr::max_int(vmax);
    int avail = 0;
    unsigned* const bm = bitmap_words<kHbm>(L, s, &avail);
    if (total > 0 && ((vmax - vmin) >> 5) < avail) {
        const int m = bitmap_unique<G>(s.sbuf, total, vmin, vmax, nullptr, 0, s.I, L.capI, bm);
        const int words = ((vmax - vmin) >> 5) + 1;
        if (packed && m != kOverflow && 2 * words <= avail && L.capI < 0x10000 && L.capJ < 0x8000) {
            /* every entry's position in I while the bitmap is at hand,
             * packed below its column slot: the gather then needs neither
             * the row index nor a search of I */
            bitmap_ranks<G>(bm, words);
            for (int e = lane; e < total; e += G)
                s.sid[e] = (s.sid[e] << 16) | bitmap_rank(bm, words, s.sbuf[e] - vmin);
            __syncthreads();
            *packed = true;
        }
        return m;
    }
    int m = group_sort_unique<G>(s.sbuf, total, s.sbuf);
    if (m > L.capI) return kOverflow;
    for (int i = lane; i < m; i += G) s.I[i] = s.sbuf[i];
    __syncthreads();
    return m;
}

/* Row -> position in I for the candidate scores (and, in the generic kernel,
 * for the gather and the residual): open addressing (linear probing,
 * Fibonacci hashing) over s.hslots slots of LDS -- in a size bin the raw
 * candidate list, which is dead from the union to the next I-set; in the
 * generic kernel (working set in HBM, where a binary search of I is a chain
 * of HBM reads) a region of LDS of its own, so that the table is never
 * written by atomics in L2 and read back through the vector L1.  Entries pack the row above the
 * position; at most half of the slots are used.  When I does not fit (or
 * rows need more bits than the packing leaves) the lookups binary-search
 * the sorted I; a probe sequence is also cut off after every slot. */
struct RowHash {
    unsigned* tab;  // nullptr: binary search
    int bits, pbits;
};

template <int G>
__device__ __forceinline__ RowHash row_hash_build(const WaveScratch& s, const ScratchLayout& L, int m, int nrows)
{
    const int lane = Group<G>::lane();
    RowHash h{nullptr, 31 - __clz(s.hslots > 1 ? s.hslots : 1), 32 - __clz(L.capI > 2 ? L.capI - 1 : 1)};
    const int nslots = 1 << h.bits;
    if (!s.htab || 2 * m > nslots || (unsigned)nrows > (0xffffffffu >> h.pbits)) return h;
    h.tab = s.htab;
    for (int i = lane; i < nslots; i += G) h.tab[i] = 0xffffffffu;
    __syncthreads();
    for (int i = lane; i < m; i += G) {
        const unsigned row = (unsigned)s.I[i];
        unsigned slot = (row * 0x9e3779b1u) >> (32 - h.bits);
        /* at least half of the slots stay empty: a free one is found */
        for (int probe = 0; probe < nslots; ++probe) {
            if (atomicCAS(&h.tab[slot], 0xffffffffu, (row << h.pbits) | (unsigned)i) == 0xffffffffu) break;
            slot = (slot + 1) & (unsigned)(nslots - 1);
        }
    }
    __syncthreads();
    return h;
}

/* Position of `row` in I, or -1. */
__device__ __forceinline__ int row_position(const RowHash& h, const int* I, int m, int row)
{
    if (h.tab) {
        unsigned slot = ((unsigned)row * 0x9e3779b1u) >> (32 - h.bits);
        const unsigned mask = (1u << h.bits) - 1u;
        for (int probe = 0; probe <= (int)mask; ++probe) {
            const unsigned e = h.tab[slot];
            if (e == 0xffffffffu) return -1;
            if ((e >> h.pbits) == (unsigned)row) return (int)(e & ((1u << h.pbits) - 1u));
            slot = (slot + 1) & mask;
        }
    }
    const int p = lower_bound_i(I, m, row);
    return (p < m && I[p] == row) ? p : -1;
}

/* Create_Submatrix_AHat (src/Spai_Sub.imp:221-282): A(I,J) column-major with
 * leading dimension capI, straight into the QR working array (and an FP32
 * copy for the mixed path's LDS QR; the register QR converts as it loads).
 * Uses the column segments build_I left for this J. */
template <int G, bool kMixed, bool kCopy = false>
__device__ __forceinline__ void gather_ahat(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                            int m, int n, const RowHash& h = RowHash{nullptr, 0, 0},
                                            bool packed = false)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int ld = L.capI;
    double* const a0 = kCopy ? s.a0 : nullptr;  // the generic kernel's untouched copy
    for (int jj = 0; jj < n; ++jj)
        for (int i = lane; i < m; i += G) {
            s.aqr[jj * ld + i] = 0.0;
            if (a0) a0[jj * ld + i] = 0.0;
            if (kMixed) s.aq32[jj * ld + i] = 0.f;
        }
    __syncthreads();
    const int total = s.segoff[n];
    for (int e = lane; e < total; e += G) {
        const int pk = s.sid[e];
        const int t = packed ? pk >> 16 : pk;
        const int src = s.segs[t] + e - s.segoff[t];
        /* every row of J's columns is in I */
        const int pos = packed ? (pk & 0xffff) : row_position(h, s.I, m, a.A.rowind[src]);
        const double v = a.A.val[src];
        s.aqr[t * ld + pos] = v;
        if (a0) a0[t * ld + pos] = v;
        if (kMixed) s.aq32[t * ld + pos] = (float)v;
    }
    __syncthreads();
}

/* Create_ek_Hat / Create_bk_Hat (src/Spai_Sub.imp:629-658, 735-775).
 * Returns COL_K_NOT_IN_I when no entry of the target column lies in I. */
template <int G>
__device__ __forceinline__ int make_rhs(const BuildArgs& a, const WaveScratch& s, int m, int col)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    for (int i = lane; i < m; i += G) s.b[i] = 0.0;
    __syncthreads();
    int hit = 0;
    if (!a.B.colptr) {
        int pos = lower_bound_i(s.I, m, col);
        hit = pos < m && s.I[pos] == col;
        if (hit && lane == 0) s.b[pos] = 1.0;
    } else {
        int e0 = a.B.colptr[col], cnt = a.B.colptr[col + 1] - e0;
        for (int e = lane; e < cnt; e += G) {
            int row = a.B.rowind[e0 + e];
            int pos = lower_bound_i(s.I, m, row);
            if (pos < m && s.I[pos] == row) {
                s.b[pos] = a.B.val[e0 + e];
                hit = 1;
            }
        }
        hit = Gr::ballot(hit) != 0ull;
    }
    __syncthreads();
    for (int i = lane; i < m; i += G) s.b0[i] = s.b[i];
    __syncthreads();
    return hit ? 0 : COL_K_NOT_IN_I;
}

/* Approximative_Solution (src/Spai_Unrestrained.imp:135-195): QR, Q^T b,
 * dtrtrs.  On exit b[0..n) holds the solution.  Returns status bits. */
template <int G, int kC = 4>
__device__ __forceinline__ int ls_solve(const ScratchLayout& L, const WaveScratch& s, int m, int n)
{
    using Gr = Group<G>;
    static_assert(kC % 4 == 0, "columns per group");
    const int lane = Gr::lane();
    const int ld = L.capI;
    int k = m < n ? m : n;
    for (int i = 0; i < k; ++i) {
        double* ai = s.aqr + i * ld;
        double tau = 0.0;
        if (i + 1 < m) {
            double alpha = ai[i];
            double p = 0.0;
            for (int r = lane; r < m; r += G)
                if (r > i) p = __fma_rn(ai[r], ai[r], p);
            double xn2 = Gr::sum_fixed(p);
            if (xn2 != 0.0) {
                double nrm = sqrt(__fma_rn(alpha, alpha, xn2));
                double beta = -copysign(nrm, alpha);
                tau = (beta - alpha) / beta;
                double scal = 1.0 / (alpha - beta);
                __syncthreads();
                for (int r = lane; r < m; r += G)
                    if (r > i) ai[r] = ai[r] * scal;
                if (lane == 0) ai[i] = beta;
                __syncthreads();
            }
        }
        if (lane == 0) s.tau[i] = tau;
        if (tau != 0.0) {
            /* H_i on columns i+1..n-1 and b, kC at a time: each column's
             * dot product and update are apply_reflector's (rows of a lane in
             * order, fused; the 64-lane tree), but the reductions share
             * their lane moves and the loads of kC columns are in flight
             * together; one barrier per reflector */
            const int nc = n - i;  // columns i+1..n-1, then b
            const double mt = -tau;
            for (int c0 = 0; c0 < nc; c0 += kC) {
                double* cp[kC];
                bool live[kC];
#pragma unroll
                for (int t = 0; t < kC; ++t) {
                    const int c = c0 + t;
                    live[t] = c < nc;
                    cp[t] = c < nc - 1 ? s.aqr + (i + 1 + c) * ld : s.b;
                }
                /* a row's kC + 1 loads are issued together, unconditionally
                 * (a dead slot reads b, in bounds, and adds or stores
                 * nothing): behind per-slot branches each load waited for
                 * the one before it */
                double p[kC];
#pragma unroll
                for (int t = 0; t < kC; ++t) p[t] = 0.0;
                for (int r = lane; r < m; r += G) {
                    if (r < i) continue;
                    const double air = ai[r];
                    double x[kC];
#pragma unroll
                    for (int t = 0; t < kC; ++t) x[t] = cp[t][r];
                    const double vr = (r == i) ? 1.0 : air;
#pragma unroll
                    for (int t = 0; t < kC; ++t) p[t] = live[t] ? __fma_rn(vr, x[t], p[t]) : p[t];
                }
                Gr::sum_bcast_n(p);
#pragma unroll
                for (int t = 0; t < kC; ++t) p[t] = mt * p[t];
                for (int r = lane; r < m; r += G) {
                    if (r < i) continue;
                    const double air = ai[r];
                    double x[kC];
#pragma unroll
                    for (int t = 0; t < kC; ++t) x[t] = cp[t][r];
                    const double vr = (r == i) ? 1.0 : air;
#pragma unroll
                    for (int t = 0; t < kC; ++t)
                        if (live[t]) cp[t][r] = __fma_rn(vr, p[t], x[t]);
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (m < n) {
        for (int i = lane; i < n; i += G) s.b[i] = 0.0;
        __syncthreads();
        return COL_UNDERDET;
    }
    bool zero = false;
    for (int i = lane; i < n; i += G) zero |= (s.aqr[i * ld + i] == 0.0);
    if (Gr::ballot(zero) != 0ull) return COL_SINGULAR;
    for (int K = n - 1; K >= 0; --K) {
        double xk = s.b[K];
        if (xk != 0.0) {
            xk = xk / s.aqr[K * ld + K];
            double nx = -xk;
            __syncthreads();
            if (lane == 0) s.b[K] = xk;
            for (int i = lane; i < K; i += G) s.b[i] = __fma_rn(nx, s.aqr[K * ld + i], s.b[i]);
        }
        __syncthreads();
    }
    return 0;
}

/* v <- (R^T R)^-1 v in FP64 for the upper-triangular R stored column-major
 * in q (leading dimension ld; FP64 or FP32 entries promoted to FP64). */
template <int G, typename RT>
__device__ __forceinline__ void normal_solve(const RT* q, int ld, double* v, int n)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    for (int K = 0; K < n; ++K) {  // R^T y = v
        double yk = v[K] / (double)q[K * ld + K];
        __syncthreads();
        if (lane == 0) v[K] = yk;
        for (int i = K + 1 + lane; i < n; i += G) v[i] = __fma_rn(-yk, (double)q[i * ld + K], v[i]);
        __syncthreads();
    }
    for (int K = n - 1; K >= 0; --K) {  // R d = y
        double dk = v[K] / (double)q[K * ld + K];
        __syncthreads();
        if (lane == 0) v[K] = dk;
        for (int i = lane; i < K; i += G) v[i] = __fma_rn(-dk, (double)q[K * ld + i], v[i]);
        __syncthreads();
    }
}

/* (R^T R)^-1 v with v_i on group lane i (n <= G): normal_solve's steps with
 * the pivots broadcast from their lane instead of going through LDS with a
 * barrier per step, and each pivot division a product with dinv = 1/R_ii
 * held by lane i (mixed refinement only: its corrections need no last bit). */
template <int G, typename RT>
__device__ __forceinline__ double normal_solve_lanes(const RT* q, int ld, double v, int n, double dinv)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    for (int K = 0; K < n; ++K) {  // R^T y = v
        const double yk = Gr::lane_value(v * dinv, K);
        if (lane == K) v = yk;
        else if (lane > K && lane < n) v = __fma_rn(-yk, (double)q[lane * ld + K], v);
    }
    /* K counts down from G - 1 for every group of the wave alike: the
     * broadcast reads lane K of each group with one readlane per group,
     * which needs the same K in all of them */
    for (int K = G - 1; K >= 0; --K) {  // R d = y
        if (K >= n) continue;
        const double dk = Gr::lane_value(v * dinv, K);
        if (lane == K) v = dk;
        else if (lane < K) v = __fma_rn(-dk, (double)q[K * ld + lane], v);
    }
    return v;
}

/* Reduce-scatter over a group: v[j] (j < G/2) summed over the G lanes; the
 * sum of v[j] ends on lanes 2j and 2j+1.  Each step hands the partner (lane
 * xor S) the half it keeps and adds the half it receives: G/2 lane moves in
 * all, against G/2 full butterflies of log2(G) moves each.  The summation
 * order is fixed but not the 64-lane order (mixed path only). */
template <int G, int S, int C, int N>
__device__ __forceinline__ void reduce_scatter_step(double (&v)[N])
{
    if constexpr (C >= 2) {
        const bool lower = (Group<G>::lane() & S) == 0;
#pragma unroll
        for (int k = 0; k < C / 2; ++k) {
            const double send = lower ? v[k + C / 2] : v[k];
            const double keep = lower ? v[k] : v[k + C / 2];
            v[k] = keep + Group<G>::template partner<S>(send);
        }
        reduce_scatter_step<G, S / 2, C / 2, N>(v);
    } else if constexpr (S >= 1) {
        v[0] = v[0] + Group<G>::template partner<S>(v[0]);
        reduce_scatter_step<G, S / 2, 1, N>(v);
    }
}

/* Group lane k's value, k a constant after unrolling (folds to one
 * broadcast). */
template <int G, int K = 0>
__device__ __forceinline__ double bcast_k(double v, int k)
{
    if constexpr (K >= G) {
        return 0.0;
    } else {
        if (k == K) return Group<G>::template lane_bcast<K>(v);
        return bcast_k<G, K + 1>(v, k);
    }
}

/* Sum over the group of v[j], j < NV (NV a power of two, NV <= G/2): the
 * halving steps hand the partner (lane xor S) the half it keeps and add the
 * half received, then plain butterflies finish; sum j ends on the G/NV
 * lanes j*G/NV .. (j+1)*G/NV - 1.  G/2 + log2(G/NV) - 1 lane moves in all,
 * against NV butterflies of log2(G) each.  The order is fixed but is not the
 * 64-lane order (mixed path only). */
template <int G, int NV>
__device__ __forceinline__ double reduce_scatter(double (&v)[NV])
{
    static_assert(NV >= 1 && 2 * NV <= G, "values per group");
    reduce_scatter_step<G, G / 2, NV, NV>(v);
    return v[0];
}

/* (A^T r)_j on lane j (0 on lanes >= n), with row i of A in a[] and r_i on
 * lane i: reduce-scatters of at most 16 sums (fewer live registers than
 * one of NV).  Mixed path only (not the 64-lane order). */
template <int G, int NV>
__device__ __forceinline__ double rows_t_times(const double (&a)[NV], double r, int n)
{
    constexpr int CH = NV < 16 ? NV : 16;
    const int lane = Group<G>::lane();
    double g = 0.0;
#pragma unroll
    for (int c0 = 0; c0 < NV; c0 += CH) {
        double v[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) v[j] = a[c0 + j] * r;
        const double t = __shfl(reduce_scatter<G, CH>(v), (lane - c0) * (G / CH), G);
        if (lane >= c0 && lane < c0 + CH) g = t;
    }
    return lane < n ? g : 0.0;
}

/* CSNE refinement of the FP32 register QR's solution for a group of G lanes
 * and |J| <= NV (16- and 64-lane groups, and 32-lane groups with |J| <= 8;
 * with |J| <= 16 in 32-lane groups it measured slower than the LDS-fed
 * steps): row i of A(I,J) in FP64 and column i
 * of R on lane i stay in registers across the steps; r = b0 - A x per row, A^T r by one
 * reduce-scatter, R^T R d = A^T r with the reciprocal diagonal, x += d.
 * Stops when max|d| <= 1e-10 max|x|; kFallback after kRefineSteps. */
template <int G, int NV>
__device__ __forceinline__ int refine_regs(const ScratchLayout& L, const WaveScratch& s, int m, int n)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int ld = L.capI, ldq = L.ldq;
    const float* q = s.aq32;
    const bool row = lane < m;
    /* column `lane` of R in registers for 16-lane groups; 32-lane groups read
     * it from LDS (registers would cost a wave per SIMD) */
    constexpr bool kRcRegs = G == 16;
    double a[NV], rc[kRcRegs ? NV : 1];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        a[j] = (row && j < n) ? s.aqr[j * ld + lane] : 0.0;
        if constexpr (kRcRegs) rc[j] = (lane < n && j < lane) ? (double)q[lane * ldq + j] : 0.0;  // R(j, lane)
    }
    const double dinv = lane < n ? 1.0 / (double)q[lane * ldq + lane] : 0.0;
    const double b0 = row ? s.b0[lane] : 0.0;
    double x = lane < n ? s.b[lane] : 0.0;
    for (int it = 0; it < kRefineSteps; ++it) {
        double r = b0;  // r = b0 - A x (FP64)
#pragma unroll
        for (int j = 0; j < NV; ++j)
            if (j < n) r = __fma_rn(-bcast_k<G>(x, j), a[j], r);
        double g = rows_t_times<G, NV>(a, r, n);  // (A^T r)_j on lane j
#pragma unroll
        for (int K = 0; K < NV; ++K) {  // R^T y = g
            if (K < n) {
                const double yk = bcast_k<G>(g * dinv, K);
                if (lane == K)
                    g = yk;
                else if (lane > K && lane < n)
                    g = __fma_rn(-yk, kRcRegs ? rc[kRcRegs ? K : 0] : (double)q[lane * ldq + K], g);
            }
        }
#pragma unroll
        for (int K = NV - 1; K >= 0; --K) {  // R d = y
            if (K < n) {
                const double dk = bcast_k<G>(g * dinv, K);
                if (lane == K)
                    g = dk;
                else if (lane < K)
                    g = __fma_rn(-dk, (double)q[K * ldq + lane], g);
            }
        }
        x += g;
        const double dmx = Gr::max(lane < n ? fabs(g) : 0.0);
        const double xmx = Gr::max(lane < n ? fabs(x) : 0.0);
        if (dmx <= 1e-10 * xmx) {
            if (lane < n) s.b[lane] = x;
            __syncthreads();
            return 0;
        }
    }
    if (lane < n) s.b[lane] = x;
    __syncthreads();
    return kFallback;
}

/* The mixed adjoint's solve R^T R lambda = gbar with the FP32 register QR's
 * R, refined in FP64 as the forward solve is: rows of A(I,J) in registers,
 * u = A lambda per row, A^T u by reduce-scatter, lambda += (R^T R)^-1
 * (gbar - A^T u) until max|d| <= 1e-10 max|lambda|.  gbar in s.g on entry,
 * lambda there on return; kFallback when it does not converge. */
template <int G, int NV>
__device__ __forceinline__ int refine_adjoint_regs(const ScratchLayout& L, const WaveScratch& s, int m, int n)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int ld = L.capI, ldq = L.ldq;
    const float* q = s.aq32;
    const bool row = lane < m;
    double a[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) a[j] = (row && j < n) ? s.aqr[j * ld + lane] : 0.0;
    const double dinv = lane < n ? 1.0 / (double)q[lane * ldq + lane] : 0.0;
    const double gbar = lane < n ? s.g[lane] : 0.0;
    double lam = normal_solve_lanes<G>(q, ldq, gbar, n, dinv);
    for (int it = 0; it < kRefineSteps; ++it) {
        double u = 0.0;  // (A lambda)_i on lane i
#pragma unroll
        for (int j = 0; j < NV; ++j)
            if (j < n) u = __fma_rn(bcast_k<G>(lam, j), a[j], u);
        const double h = rows_t_times<G, NV>(a, row ? u : 0.0, n);
        const double d = normal_solve_lanes<G>(q, ldq, lane < n ? gbar - h : 0.0, n, dinv);
        lam += d;
        const double dmx = Gr::max(lane < n ? fabs(d) : 0.0);
        const double lmx = Gr::max(lane < n ? fabs(lam) : 0.0);
        if (dmx <= 1e-10 * lmx) {
            if (lane < n) s.g[lane] = lam;
            __syncthreads();
            return 0;
        }
    }
    return kFallback;
}

/* out[i] = sum_j C(I_i, J_j) w_j over I in FP64, from the CSC columns or
 * (dense) from A(I,J) in the working set; with `base`, out = base - C w. */
template <int G>
__device__ __forceinline__ void csc_apply(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s, int m,
                                          int n, const double* w, const double* base, double* out, bool dense)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const double sign = base ? -1.0 : 1.0;
    if (dense) {
        /* from A(I,J) in the working set (FP64, untouched by the FP32 QR) */
        const int ld = L.capI;
        for (int i = lane; i < m; i += G) {
            double o = base ? base[i] : 0.0;
            for (int jj = 0; jj < n; ++jj) o = __fma_rn(sign * w[jj], s.aqr[jj * ld + i], o);
            out[i] = o;
        }
        __syncthreads();
        return;
    }
    for (int i = lane; i < m; i += G) out[i] = base ? base[i] : 0.0;
    __syncthreads();
    for (int jj = 0; jj < n; ++jj) {
        double wj = sign * w[jj];
        int j = s.J[jj];
        int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
        for (int e = lane; e < cnt; e += G) {
            int pos = lower_bound_i(s.I, m, a.A.rowind[e0 + e]);
            out[pos] = __fma_rn(wj, a.A.val[e0 + e], out[pos]);
        }
        __syncthreads();
    }
}

/* out[jj] = (A(I,J)^T v)_jj, from the sparse columns or (dense) from A(I,J)
 * in the working set. */
template <int G>
__device__ __forceinline__ void csc_apply_t(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s, int m,
                                            int n, const double* v, double* out, bool dense)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    if (dense) {
        const int ld = L.capI;
        for (int jj = 0; jj < n; ++jj) {
            double p = 0.0;
            for (int i = lane; i < m; i += G) p = __fma_rn(s.aqr[jj * ld + i], v[i], p);
            p = Gr::sum_fixed(p);
            if (lane == 0) out[jj] = p;
        }
        __syncthreads();
        return;
    }
    for (int jj = 0; jj < n; ++jj) {
        int j = s.J[jj];
        int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
        double p = 0.0;
        for (int e = lane; e < cnt; e += G) {
            int pos = lower_bound_i(s.I, m, a.A.rowind[e0 + e]);
            p = __fma_rn(a.A.val[e0 + e], v[pos], p);
        }
        p = Gr::sum_fixed(p);
        if (lane == 0) out[jj] = p;
    }
    __syncthreads();
}

/* c <- (I - tau v v^T) c in FP32 (any summation order). */
template <int G>
__device__ __forceinline__ void apply_reflector_f(const float* v, float* c, int i, int m, float tau)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    float p = 0.f;
    for (int r = lane; r < m; r += G) {
        if (r < i) continue;
        float vr = (r == i) ? 1.f : v[r];
        p = __fmaf_rn(vr, c[r], p);
    }
    float w = Gr::sum_f(p);
    float t = (-tau) * w;
    for (int r = lane; r < m; r += G) {
        if (r < i) continue;
        float vr = (r == i) ? 1.f : v[r];
        c[r] = __fmaf_rn(vr, t, c[r]);
    }
    __syncthreads();
}

/* Mixed-precision least squares: Householder QR of A(I,J) in FP32 (Q^T b
 * carried as an extra column), x0 = R^-1 Q^T b in FP32, then corrected
 * semi-normal-equation steps in FP64: r = b - A x, R^T R d = A^T r, x += d,
 * with A(I,J) read in FP64 from the CSC columns and R promoted to FP64.
 * Returns 0, or kFallback when the FP32 factor is too ill-conditioned or the
 * corrections have not settled after kRefineSteps steps. */
/* Householder QR of the FP32 copy of A(I,J) with the FP32 right-hand side
 * carried along; kFallback when min |R_ii| <= kCondGuard max |R_ii|. */
template <int G>
__device__ __forceinline__ int qr_f32(const ScratchLayout& L, const WaveScratch& s, int m, int n)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int ld = L.capI;
    if (m < n || n == 0) return kFallback;
    float* q = s.aq32;
    for (int i = lane; i < m; i += G) s.bq32[i] = (float)s.b[i];
    __syncthreads();
    for (int i = 0; i < n; ++i) {
        float* ai = q + i * ld;
        float tau = 0.f;
        if (i + 1 < m) {
            float alpha = ai[i];
            float p = 0.f;
            for (int r = lane; r < m; r += G)
                if (r > i) p = __fmaf_rn(ai[r], ai[r], p);
            float xn2 = Gr::sum_f(p);
            if (xn2 != 0.f) {
                float nrm = sqrtf(__fmaf_rn(alpha, alpha, xn2));
                float beta = -copysignf(nrm, alpha);
                tau = (beta - alpha) / beta;
                float scal = 1.f / (alpha - beta);
                __syncthreads();
                for (int r = lane; r < m; r += G)
                    if (r > i) ai[r] = ai[r] * scal;
                if (lane == 0) ai[i] = beta;
                __syncthreads();
            }
        }
        if (tau != 0.f) {
            for (int j = i + 1; j < n; ++j) apply_reflector_f<G>(ai, q + j * ld, i, m, tau);
            apply_reflector_f<G>(ai, s.bq32, i, m, tau);
        }
    }
    float dmax = 0.f, dmin = INFINITY;
    for (int i = lane; i < n; i += G) {
        float d = fabsf(q[i * ld + i]);
        dmax = fmaxf(dmax, d);
        dmin = fminf(dmin, d);
    }
    dmax = Gr::maxf(dmax);
    dmin = Gr::minf(dmin);
    return (dmin > kCondGuard * dmax) ? 0 : kFallback;
}

/* FP32 Householder QR of A(I,J) with row r in the registers of group lane r
 * (|I| <= G, |J| <= NMAX), the FP32 right-hand side in slot NMAX; x0 =
 * R^-1 Q^T b in FP32.  Rows come from the FP64 A(I,J) in s.aqr; leaves R
 * (upper triangle, column-major, leading dimension L.ldq) in s.aq32 for the
 * refinement and x0 in s.b.  kFallback when min |R_ii| <= kCondGuard
 * max |R_ii| (summation order is free here: the FP64 refinement decides). */
template <int G, int NMAX>
__device__ __forceinline__ int qr_regs_f32(const ScratchLayout& L, const WaveScratch& s, int m, int n)
{
    using Gr = Group<G>;
    constexpr int kB = 4;
    const int lane = Gr::lane();
    const int ld = L.capI;
    if (m < n || n == 0) return kFallback;
    const bool row = lane < m;
    float a[NMAX + 1];
#pragma unroll
    for (int j = 0; j < NMAX; ++j) a[j] = (row && j < n) ? (float)s.aqr[j * ld + lane] : 0.f;
    a[NMAX] = row ? (float)s.b[lane] : 0.f;
    float dmax = 0.f, dmin = INFINITY;
    const auto plus = [](float x, float y) { return x + y; };
#pragma unroll
    for (int ib = 0; ib < NMAX; ib += kB) {
        if (ib >= n) break;
#pragma unroll 1
        for (int r = 0; r < kB; ++r) {
            const int i = ib + r;
            if (i >= n) break;
            float ai = a[ib];
#pragma unroll
            for (int t = 1; t < kB; ++t)
                if (t == r) ai = a[ib + t];
            /* whole waves: i is wave-uniform, so a readlane; narrower groups
             * may stand at different steps (their n differ) */
            const float alpha = G == kWave ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ai), i))
                                           : __shfl(ai, i, G);
            const float xn2 = Gr::allreduce(lane > i ? ai * ai : 0.f, plus);
            float rii = ai;
            if (i + 1 < m && xn2 != 0.f) {
                const float nrm = sqrtf(__fmaf_rn(alpha, alpha, xn2));
                const float beta = -copysignf(nrm, alpha);
                const float tau = (beta - alpha) / beta;
                const float scal = 1.f / (alpha - beta);
                const float vi = lane > i ? ai * scal : (lane == i ? 1.f : 0.f);
                const float ni = lane > i ? vi : (lane == i ? beta : ai);
                rii = beta;
#pragma unroll
                for (int t = 0; t < kB; ++t)
                    if (t == r) a[ib + t] = ni;
                const bool act = lane >= i;
#pragma unroll
                for (int jb = ib; jb <= NMAX; jb += kB) {
                    if ((jb + kB - 1 > i && jb < n) || jb + kB > NMAX) {
                        float w[kB];
#pragma unroll
                        for (int t = 0; t < kB; ++t)
                            w[t] = (jb + t <= NMAX && act) ? vi * a[jb + t <= NMAX ? jb + t : NMAX] : 0.f;
                        if (jb + kB - 1 <= NMAX) {
                            Gr::sum4_bcast_f(w);  // four sums sharing their lane swaps
                        } else {
#pragma unroll
                            for (int t = 0; t < kB; ++t)
                                if (jb + t <= NMAX) w[t] = Gr::allreduce(w[t], plus);
                        }
#pragma unroll
                        for (int t = 0; t < kB; ++t) {
                            const int j = jb + t;
                            if (j <= NMAX) {
                                const bool live = j == NMAX || (j > i && j < n);
                                if (live && act) a[j] = __fmaf_rn(vi, -tau * w[t], a[j]);
                            }
                        }
                    }
                }
            }
            if (lane == i) {
                dmax = fmaxf(dmax, fabsf(rii));
                dmin = fminf(dmin, fabsf(rii));
            }
        }
    }
    dmax = Gr::maxf(dmax);
    dmin = Gr::minf(dmin);
    if (!(dmin > kCondGuard * dmax)) return kFallback;
    float x = a[NMAX];
#pragma unroll
    for (int K = NMAX - 1; K >= 0; --K) {
        if (K < n) {
            /* K is a constant: one readlane per group, the same in every group */
            const float xk = __int_as_float(Gr::lane_value_i(__float_as_int(x), K)) /
                             __int_as_float(Gr::lane_value_i(__float_as_int(a[K]), K));
            if (lane == K) x = xk;
            if (lane < K) x = __fmaf_rn(-xk, a[K], x);
        }
    }
#pragma unroll
    for (int K = 0; K < NMAX; ++K)
        if (K < n && lane <= K) s.aq32[K * L.ldq + lane] = a[K];
    if (lane < n) s.b[lane] = (double)x;
    __syncthreads();
    return 0;
}

/* The mixed factor of the 64-lane LDS bins: G = A^T A of the FP32-rounded
 * A(I,J) on the matrix cores (v_mfma_f32_32x32x2f32: lane l supplies
 * A(r0 + l/32, l%32), which is both the A^T and the A operand; two rows of
 * A per instruction, two accumulators), written to LDS by its accumulator
 * layout (lane l, register v: row 8(v/4) + 4(l/32) + v%4, column l%32),
 * then an FP32 Cholesky G = R^T R with row p of G in the registers of lane
 * p (rows of R broadcast through LDS, one pivot at a time).  R (upper,
 * column-major, leading dimension kGramLd) is left in aq32 for the CSNE
 * refinement.  kFallback when |J| > 32 (a 33rd column would need a second
 * accumulator set and 33 registers per row: the kernel then spills), when
 * a pivot is not positive, or when R's diagonal ratio is below kGramGuard
 * (the refinement's rate is about kappa^2 u_32). */
constexpr float kGramGuard = 3e-3f;
constexpr int kGramSteps = 5;
__device__ __forceinline__ int gram_chol(const ScratchLayout& L, const WaveScratch& s, int m, int n)
{
    typedef float f16v __attribute__((ext_vector_type(16)));
    using Gr = Group<kWave>;
    const int lane = Gr::lane();
    if (n > 32 || n == 0 || m < n) return kFallback;
    const int ci = lane & 31, kr = lane >> 5;
    const bool live = ci < n;
    const double* const col = s.aqr + (size_t)ci * L.capI;
    f16v acc0 = {}, acc1 = {};
    int r0 = 0;
    for (; r0 + 4 <= m; r0 += 4) {
        const float a0 = live ? (float)col[r0 + kr] : 0.f;
        const float a1 = live ? (float)col[r0 + 2 + kr] : 0.f;
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, a0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, a1, acc1, 0, 0, 0);
    }
    for (; r0 < m; r0 += 2) {
        const int r = r0 + kr;
        const float a0 = (live && r < m) ? (float)col[r] : 0.f;
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, a0, acc0, 0, 0, 0);
    }
    float* const q = s.aq32;  // kGramLd x kGramLd
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int i = 8 * (v >> 2) + 4 * kr + (v & 3);
        q[i * kGramLd + ci] = acc0[v] + acc1[v];  // row i, column ci of G
    }
    __syncthreads();
    float g[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) g[c] = lane < n ? q[lane * kGramLd + c] : 0.f;
    __syncthreads();
    bool bad = false;
    float dmax = 0.f, dmin = INFINITY;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        if (j >= n) break;
        if (lane == j) {
            /* row j of R: R_jj = sqrt(pivot), R_jc = G_jc / R_jj, to LDS
             * (R_jc at q[c * kGramLd + j]) */
            const float piv = g[j];
            bad |= !(piv > 0.f);
            const float d = sqrtf(piv > 0.f ? piv : 1.f);
            const float inv = 1.f / d;
#pragma unroll
            for (int c = 0; c < 32; ++c)
                if (c >= j && c < n) q[c * kGramLd + j] = c == j ? d : g[c] * inv;
            dmax = fmaxf(dmax, d);
            dmin = fminf(dmin, d);
        }
        __syncthreads();
        if (lane > j && lane < n) {
            /* G_pc -= R_jp R_jc for c >= p (the upper part of row p) */
            const float rjp = q[lane * kGramLd + j];
#pragma unroll
            for (int c = 0; c < 32; ++c)
                if (c > j && c >= lane && c < n) g[c] = __fmaf_rn(-rjp, q[c * kGramLd + j], g[c]);
        }
    }
    __syncthreads();
    bad = Gr::ballot(bad) != 0ull;
    dmax = Gr::maxf(dmax);
    dmin = Gr::minf(dmin);
    if (bad || !(dmin > kGramGuard * dmax)) return kFallback;
    return 0;
}

template <int G, int NMAX>
__device__ __forceinline__ int ls_solve_mixed(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                              int m, int n)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int ld = L.capI;
    const float* q = s.aq32;
    if (NMAX > 0 && m <= G && n <= NMAX) {
        if (qr_regs_f32<G, (NMAX > 0 ? NMAX : 4)>(L, s, m, n) == kFallback) return kFallback;
        /* the register-resident refinement in 16- and 64-lane groups (the
         * 64-lane bins are held to 8 waves per CU by LDS, so the registers
         * cost no wave there) and in 32-lane groups with |J| <= 8; wider
         * 32-lane columns keep the LDS-fed steps below (their rows in
         * registers would cost a wave per SIMD) */
        if constexpr (G == 16 && NMAX == 8) return refine_regs<G, 8>(L, s, m, n);
        if constexpr (G == 32) {
            if (n <= 8) return refine_regs<G, 8>(L, s, m, n);
        }
        if constexpr (G == 64 && (NMAX == 32 || NMAX == 16)) return refine_regs<G, NMAX>(L, s, m, n);
        /* the same steps with one row (r_i) or one unknown (x_i, g_i) per lane:
         * no barrier except after publishing x for the next residual */
        const bool row = lane < m;
        double x = lane < n ? s.b[lane] : 0.0;
        const int ldq = L.ldq;
        const double dinv = lane < n ? 1.0 / (double)q[lane * ldq + lane] : 0.0;  // R's reciprocal diagonal
        for (int it = 0; it < kRefineSteps; ++it) {
            double r = row ? s.b0[lane] : 0.0;  // r = b0 - A x (FP64)
            for (int jj = 0; jj < n; ++jj)
                if (row) r = __fma_rn(-s.b[jj], s.aqr[jj * ld + lane], r);
            double g = 0.0;  // g = A^T r, g_j on lane j; four sums at a time share their lane swaps
            for (int jb = 0; jb < n; jb += 4) {
                double p[4];
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    p[t] = (row && jb + t < n) ? __fma_rn(s.aqr[(jb + t) * ld + lane], r, 0.0) : 0.0;
                Gr::sum_bcast_n(p);
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (lane == jb + t) g = p[t];
            }
            g = normal_solve_lanes<G>(q, ldq, g, n, dinv);  // R^T R d = g, R from the FP32 factor
            x += g;
            const double dmx = Gr::max(lane < n ? fabs(g) : 0.0);
            const double xmx = Gr::max(lane < n ? fabs(x) : 0.0);
            if (lane < n) s.b[lane] = x;
            __syncthreads();
            if (dmx <= 1e-10 * xmx) return 0;
        }
        return kFallback;
    } else if (G == kWave && L.gram) {
        /* R from the Gram matrix; x0 = 0, so the first step is the
         * semi-normal equations' solution */
        if (NMAX > 0 || gram_chol(L, s, m, n) == kFallback) return kFallback;
        for (int i = lane; i < n; i += G) s.b[i] = 0.0;
        __syncthreads();
    } else {
        /* register bins hold register columns only (their aq32 has room for R alone) */
        if (NMAX > 0 || qr_f32<G>(L, s, m, n) == kFallback) return kFallback;
        /* x0 in FP32, column-oriented back substitution */
        for (int K = n - 1; K >= 0; --K) {
            float xk = s.bq32[K] / q[K * ld + K];
            __syncthreads();
            if (lane == 0) s.bq32[K] = xk;
            for (int i = lane; i < K; i += G) s.bq32[i] = __fmaf_rn(-xk, q[K * ld + i], s.bq32[i]);
            __syncthreads();
        }
        for (int i = lane; i < n; i += G) s.b[i] = (double)s.bq32[i];
        __syncthreads();
    }
    const bool gram = G == kWave && L.gram;
    const int ldr = gram ? L.ldq : ld, steps = gram ? kGramSteps : kRefineSteps;
    for (int it = 0; it < steps; ++it) {
        csc_apply<G>(a, L, s, m, n, s.b, s.b0, s.r, true);  // r = b0 - A x (FP64)
        csc_apply_t<G>(a, L, s, m, n, s.r, s.g, true);      // g = A^T r
        normal_solve<G>(q, ldr, s.g, n);                     // R^T R d = g, R from the FP32 factor
        double dmx = 0.0, xmx = 0.0;
        for (int i = lane; i < n; i += G) {
            double x = s.b[i] + s.g[i];
            s.b[i] = x;
            dmx = fmax(dmx, fabs(s.g[i]));
            xmx = fmax(xmx, fabs(x));
        }
        dmx = Gr::max(dmx);
        xmx = Gr::max(xmx);
        __syncthreads();
        /* each step shrinks the error by about kappa^2 u_32; a correction
         * below 1e-10 |x| leaves the next one below FP64 rounding */
        if (dmx <= 1e-10 * xmx) return 0;
    }
    return kFallback;  // not converged in kRefineSteps: let FP64 decide
}

/* Element j of a register row (j uniform in the group, any value). */
template <int NMAX>
__device__ __forceinline__ double row_pick(const double (&a)[NMAX], int j)
{
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < NMAX; ++t)
        if (t == j) v = a[t];
    return v;
}

/* Approximative_Solution (src/Spai_Unrestrained.imp:135-195) with row r of
 * A(I,J) held in the registers of group lane r (|I| <= G, |J| <= NMAX).  The
 * operations and their order are those of ls_solve (a single row per lane
 * is the 64-lane order's only slot); the matrix never goes through LDS and
 * the group needs no barrier between reflectors.  Slot NMAX of the row
 * holds the right-hand side; a reflector is applied to four slots at a time
 * so that their reductions interleave (slots outside (i, n) and b keep their
 * values). */
template <int G, int NMAX, bool kROut = false>
__device__ __forceinline__ int ls_solve_regs(const ScratchLayout& L, const WaveScratch& s, int m, int n)
{
    using Gr = Group<G>;
    constexpr int kB = 4;  // slots per block
    static_assert(NMAX % kB == 0, "register width");
    const int lane = Gr::lane();
    const int ld = L.capI;
    const bool row = lane < m;
    double a[NMAX + 1];
#pragma unroll
    for (int j = 0; j < NMAX; ++j) a[j] = (row && j < n) ? s.aqr[j * ld + lane] : 0.0;
    a[NMAX] = row ? s.b[lane] : 0.0;
    const int k = m < n ? m : n;
    bool zero_diag = false;  // R_ii == 0 on this lane's row (final once step i is done)
    /* step i = ib + r: the block of i is known at compile time, so column i is
     * picked from four registers, and only the blocks from ib on are visited */
#pragma unroll
    for (int ib = 0; ib < NMAX; ib += kB) {
        if (ib >= k) break;
#pragma unroll 1
        for (int r = 0; r < kB; ++r) {
            const int i = ib + r;
            if (i >= k) break;
            double ai = a[ib];
#pragma unroll
            for (int t = 1; t < kB; ++t)
                if (t == r) ai = a[ib + t];
            const double alpha = Gr::lane_value(ai, i);
            const double xn2 = Gr::sum_fixed(lane > i ? __fma_rn(ai, ai, 0.0) : 0.0);
            double rii = ai;
            if (i + 1 < m && xn2 != 0.0) {
                const double nrm = sqrt(__fma_rn(alpha, alpha, xn2));
                const double beta = -copysign(nrm, alpha);
                const double tau = (beta - alpha) / beta;
                const double scal = 1.0 / (alpha - beta);
                const double vi = lane > i ? ai * scal : (lane == i ? 1.0 : 0.0);
                const double ni = lane > i ? vi : (lane == i ? beta : ai);
                rii = beta;
#pragma unroll
                for (int t = 0; t < kB; ++t)
                    if (t == r) a[ib + t] = ni;
                const double mt = -tau;
                const bool act = lane >= i;
#pragma unroll
                for (int jb = ib; jb <= NMAX; jb += kB) {
                    /* the block holds a slot in (i, n) or the right-hand side */
                    if ((jb + kB - 1 > i && jb < n) || jb + kB > NMAX) {
                        double p[kB];
#pragma unroll
                        for (int t = 0; t < kB; ++t)
                            p[t] = (jb + t <= NMAX && act) ? __fma_rn(vi, a[jb + t <= NMAX ? jb + t : NMAX], 0.0)
                                                           : 0.0;
                        Gr::sum_bcast_n(p);
#pragma unroll
                        for (int t = 0; t < kB; ++t) {
                            const int j = jb + t;
                            if (j <= NMAX) {
                                const bool live = j == NMAX || (j > i && j < n);
                                if (live && act) a[j] = __fma_rn(vi, mt * p[t], a[j]);
                            }
                        }
                    }
                }
            }
            zero_diag |= lane == i && rii == 0.0;
        }
    }
    if (kROut) {
        /* the adjoint's normal-equation solves need R, not x: R_iK (i <= K)
         * to s.rq, column-major with leading dimension capJ */
#pragma unroll
        for (int K = 0; K < NMAX; ++K)
            if (K < n && lane <= K) s.rq[K * L.capJ + lane] = a[K];
        __syncthreads();
        if (m < n) return COL_UNDERDET;
        return Gr::ballot(zero_diag) != 0ull ? COL_SINGULAR : 0;
    }
    double b = a[NMAX];
    int st = 0;
    if (m < n) {
        st = COL_UNDERDET;
        b = 0.0;
    } else if (Gr::ballot(zero_diag) != 0ull) {
        st = COL_SINGULAR;
    } else {
        /* dtrsm order; K runs over compile-time slots so R_iK is a fixed
         * register of lane i */
#pragma unroll
        for (int K = NMAX - 1; K >= 0; --K) {
            if (K < n) {
                double xk = Gr::lane_value(b, K);
                const double rkk = Gr::lane_value(a[K], K);
                if (xk != 0.0) {
                    xk = xk / rkk;
                    if (lane == K) b = xk;
                    if (lane < K) b = __fma_rn(-xk, a[K], b);
                }
            }
        }
    }
    if (lane < n) s.b[lane] = b;
    __syncthreads();
    return st;
}

/* Row-group layout for 64-lane groups: lane l = 16 c + q holds rows q,
 * q + 16, q + 32, q + 48 (t = 0..3) of the columns j = c + 4 s (slot s =
 * 0..7); the right-hand side is column 32 (c = 0, slot 8).  The dot product
 * of a column adds its four rows in the lane first -- (p_t0 + p_t2) +
 * (p_t1 + p_t3) are the xor-32 and xor-16 steps of the 64-lane order -- then
 * the xor-8..1 steps run as DPP inside the 16-lane row, so every row group
 * reduces its own column and no broadcast of the sums is needed.  The
 * Householder vector of step i lives in row group i mod 4 and is copied to
 * the other three with lane swaps. */

/* value X of row group `src` (0..3) in every row group (same q) */
__device__ __forceinline__ unsigned rowgroup_bcast(unsigned x, int src)
{
    auto a = __builtin_amdgcn_permlane16_swap(x, x, false, false);  // a[0]: rows 0,0,2,2; a[1]: 1,1,3,3
    const unsigned z = (src & 1) ? a[1] : a[0];
    auto b = __builtin_amdgcn_permlane32_swap(z, z, false, false);  // b[0]: z rows 0,1,0,1; b[1]: 2,3,2,3
    return (src & 2) ? b[1] : b[0];
}
__device__ __forceinline__ double rowgroup_bcast(double x, int src)
{
    const long long v = __double_as_longlong(x);
    const unsigned lo = rowgroup_bcast((unsigned)v, src), hi = rowgroup_bcast((unsigned)(v >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
/* value X of row group SRC (compile time) in row group 0 */
template <int SRC>
__device__ __forceinline__ unsigned rowgroup_to0(unsigned x)
{
    if (SRC == 0) return x;
    if (SRC == 1) return __builtin_amdgcn_permlane16_swap(x, x, false, false)[1];
    if (SRC == 2) return __builtin_amdgcn_permlane32_swap(x, x, false, false)[1];
    const unsigned y = __builtin_amdgcn_permlane16_swap(x, x, false, false)[1];
    return __builtin_amdgcn_permlane32_swap(y, y, false, false)[1];
}
template <int SRC>
__device__ __forceinline__ double rowgroup_to0(double x)
{
    const long long v = __double_as_longlong(x);
    const unsigned lo = rowgroup_to0<SRC>((unsigned)v), hi = rowgroup_to0<SRC>((unsigned)(v >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
/* the four in-lane rows, then xor 8..1 inside the 16-lane row */
__device__ __forceinline__ double rowgroup_sum(const double (&p)[4])
{
    using Gr = Group<64>;
    double v = (p[0] + p[2]) + (p[1] + p[3]);
    v = Gr::sum_step<8>(v);
    v = Gr::sum_step<4>(v);
    v = Gr::sum_step<2>(v);
    v = Gr::sum_step<1>(v);
    return v;
}

/* The in-lane part of a 64-lane-order sum for the row-group layout with RT
 * rows per lane: virtual lane v = q + 16u (u = 0..3) owns rows v and, when
 * RT = 8, v + 64 (slot u + 4), which the order adds to the lane's chain
 * first; f(t) gives slot t's term and `fma2(t, acc)` continues the chain. */
template <int RT, class F, class F2>
__device__ __forceinline__ double rowgroup_sum_rt(F first, F2 next)
{
    double p[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        p[u] = first(u);
        if constexpr (RT == 8) p[u] = next(u + 4, p[u]);
    }
    return rowgroup_sum(p);
}

/* NS column slots per lane (|J| <= 4 NS); slot NS holds b.  RT rows per
 * lane and slot: 4 (|I| <= 64) or 8 (|I| <= 128: rows q + 16 t, t = 0..7;
 * row r + 64 continues the chain of row r, as in the 64-lane order). */
template <bool kROut, int NS = 8, int RT = 4>
__device__ __forceinline__ int ls_solve_rows64(const ScratchLayout& L, const WaveScratch& s, int m, int n)
{
    static_assert(RT == 4 || RT == 8, "rows per lane");
    const int lane = lane_id();
    const int c = lane >> 4, q = lane & 15;
    const int ld = L.capI;
    double a[NS + 1][RT];
#pragma unroll
    for (int sl = 0; sl < NS; ++sl)
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const int j = c + 4 * sl, r = q + 16 * t;
            a[sl][t] = (r < m && j < n) ? s.aqr[j * ld + r] : 0.0;
        }
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int r = q + 16 * t;
        a[NS][t] = (c == 0 && r < m) ? s.b[r] : 0.0;
    }
    const int k = m < n ? m : n;
    bool zero_diag = false;
#pragma unroll
    for (int si = 0; si < NS; ++si) {
        if (4 * si >= k) break;
#pragma unroll 1
        for (int ci = 0; ci < 4; ++ci) {
            const int i = 4 * si + ci;
            if (i >= k) break;
            const int ti = i >> 4, qi = i & 15;
            /* column i sits in row group ci, slot si */
            const double xn2 = Group<64>::lane_value(
                rowgroup_sum_rt<RT>(
                    [&](int t) {
                        const int r = q + 16 * t;
                        return r > i ? __fma_rn(a[si][t], a[si][t], 0.0) : 0.0;
                    },
                    [&](int t, double acc) {
                        const int r = q + 16 * t;
                        return (r > i && r < m) ? __fma_rn(a[si][t], a[si][t], acc) : acc;
                    }),
                16 * ci);
            double aii_l = a[si][0];
#pragma unroll
            for (int t = 1; t < RT; ++t)
                if (t == ti) aii_l = a[si][t];
            const double alpha = Group<64>::lane_value(aii_l, 16 * ci + qi);
            double rii = alpha;
            if (i + 1 < m && xn2 != 0.0) {
                const double nrm = sqrt(__fma_rn(alpha, alpha, xn2));
                const double beta = -copysign(nrm, alpha);
                const double tau = (beta - alpha) / beta;
                const double scal = 1.0 / (alpha - beta);
                rii = beta;
                /* v over this lane's rows, formed in row group ci, copied to all */
                double v[RT];
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    const int r = q + 16 * t;
                    const double vl = r > i ? a[si][t] * scal : (r == i ? 1.0 : 0.0);
                    v[t] = rowgroup_bcast(vl, ci);
                    /* R_ii into column i's own lane; rows below keep v (dead) */
                    if (c == ci && r == i) a[si][t] = beta;
                    else if (c == ci && r > i) a[si][t] = vl;
                }
                const double mt = -tau;
#pragma unroll
                for (int sl = si; sl <= NS; ++sl) {
                    if ((4 * sl + 3 > i && 4 * sl < n) || sl == NS) {
                        const int j = c + 4 * sl;
                        const bool live = sl == NS ? c == 0 : (j > i && j < n);
                        const double w = rowgroup_sum_rt<RT>(
                            [&](int t) { return q + 16 * t >= i ? __fma_rn(v[t], a[sl][t], 0.0) : 0.0; },
                            [&](int t, double acc) {
                                const int r = q + 16 * t;
                                return (r >= i && r < m) ? __fma_rn(v[t], a[sl][t], acc) : acc;
                            });
                        const double tw = mt * w;
#pragma unroll
                        for (int t = 0; t < RT; ++t)
                            if (live && q + 16 * t >= i) a[sl][t] = __fma_rn(v[t], tw, a[sl][t]);
                    }
                }
            }
            zero_diag |= c == ci && q == qi && rii == 0.0;
        }
    }
    if (kROut) {
#pragma unroll
        for (int sl = 0; sl < NS; ++sl)
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const int j = c + 4 * sl, r = q + 16 * t;
                if (j < n && r <= j) s.rq[j * L.capJ + r] = a[sl][t];
            }
        __syncthreads();
        if (m < n) return COL_UNDERDET;
        return __ballot(zero_diag) != 0ull ? COL_SINGULAR : 0;
    }
    int st = 0;
    if (m < n) {
        st = COL_UNDERDET;
#pragma unroll
        for (int t = 0; t < RT; ++t) a[NS][t] = 0.0;
    } else if (__ballot(zero_diag) != 0ull) {
        st = COL_SINGULAR;
    } else {
        /* dtrsm order: K over compile-time columns, x_K to all, column K of R
         * moved from its row group to row group 0 where b lives */
#pragma unroll
        for (int K = 4 * NS - 1; K >= 0; --K) {
            if (K < n) {
                const int cK = K & 3, sK = K >> 2, tK = K >> 4, qK = K & 15;
                double xk = Group<64>::lane_value(a[NS][tK], qK);
                const double rkk = Group<64>::lane_value(a[sK][tK], 16 * cK + qK);
                if (xk != 0.0) {
                    xk = xk / rkk;
                    double rcol[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (t <= tK) {
                            switch (cK) {
                            case 0: rcol[t] = rowgroup_to0<0>(a[sK][t]); break;
                            case 1: rcol[t] = rowgroup_to0<1>(a[sK][t]); break;
                            case 2: rcol[t] = rowgroup_to0<2>(a[sK][t]); break;
                            default: rcol[t] = rowgroup_to0<3>(a[sK][t]); break;
                            }
                        }
                    }
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (t <= tK) {
                            const int r = q + 16 * t;
                            if (c == 0 && r == K) a[NS][t] = xk;
                            else if (c == 0 && r < K) a[NS][t] = __fma_rn(-xk, rcol[t], a[NS][t]);
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int r = q + 16 * t;
        if (c == 0 && r < n) s.b[r] = a[NS][t];
    }
    __syncthreads();
    return st;
}

/* Residual_Norm (src/Spai_Sub.imp:915-941): r = A(I,J) x - b in dgemv's column
 * order with beta = -1, then sqrt of the sequential sum of squares.  With
 * `dense`, from A(I,J) in the working set; else from the sparse columns (a
 * stored entry of column j updates its row in order of j, as the dense
 * product does; the dense product's zero terms leave every nonzero residual
 * unchanged). */
template <int G>
__device__ __forceinline__ double residual_norm(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                                int m, int n, bool dense,
                                                const RowHash& h = RowHash{nullptr, 0, 0},
                                                const double* Ad = nullptr)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    if (dense) {
        /* A(I,J) is intact in Ad (default aqr: the register QR leaves it; the
         * generic kernel keeps a copy): dgemv's own column order,
         * r_i = fma(x_j, A_ij, r_i) for each x_j != 0 */
        const int ld = L.capI;
        const double* const A = Ad ? Ad : s.aqr;
        for (int i = lane; i < m; i += G) {
            double r = -s.b0[i];
            for (int jj = 0; jj < n; ++jj) {
                const double xj = s.b[jj];
                if (xj != 0.0) r = __fma_rn(xj, A[jj * ld + i], r);
            }
            s.r[i] = r;
        }
        __syncthreads();
    } else {
        for (int i = lane; i < m; i += G) s.r[i] = -s.b0[i];
        __syncthreads();
        /* the columns' bounds and x_j one per lane; then kD columns' entries
         * loaded and their rows found together, the updates applied column
         * by column (a row takes its terms in order of j, as before) */
        constexpr int kD = 4;
        for (int j0 = 0; j0 < n; j0 += G) {
            const int jj = j0 + lane;
            int e0 = 0, cnt = 0;
            double xj = 0.0;
            if (jj < n) {
                xj = s.b[jj];
                const int j = s.J[jj];
                e0 = a.A.colptr[j];
                cnt = xj != 0.0 ? a.A.colptr[j + 1] - e0 : 0;
            }
            const int cols = n - j0 < G ? n - j0 : G;
            if (Gr::max_int(cnt) > G) {
                /* a column longer than the group: one column at a time */
                for (int q = 0; q < cols; ++q) {
                    const int c = Gr::lane_value_i(cnt, q), b = Gr::lane_value_i(e0, q);
                    const double x = Gr::lane_value(xj, q);
                    for (int e = lane; e < c; e += G) {
                        const int pos = row_position(h, s.I, m, a.A.rowind[b + e]);
                        s.r[pos] = __fma_rn(x, a.A.val[b + e], s.r[pos]);
                    }
                    __syncthreads();
                }
                continue;
            }
            for (int q0 = 0; q0 < cols; q0 += kD) {
                int pos[kD];
                double v[kD];
#pragma unroll
                for (int d = 0; d < kD; ++d) {
                    pos[d] = -1;
                    v[d] = 0.0;
                    const int q = q0 + d;
                    if (q < cols) {
                        const int c = Gr::lane_value_i(cnt, q), b = Gr::lane_value_i(e0, q);
                        if (lane < c) {
                            pos[d] = a.A.rowind[b + lane];
                            v[d] = a.A.val[b + lane];
                        }
                    }
                }
#pragma unroll
                for (int d = 0; d < kD; ++d)
                    if (pos[d] >= 0) pos[d] = row_position(h, s.I, m, pos[d]);
#pragma unroll
                for (int d = 0; d < kD; ++d) {
                    const int q = q0 + d;
                    if (q < cols) {
                        const double x = Gr::lane_value(xj, q);
                        if (pos[d] >= 0) s.r[pos[d]] = __fma_rn(x, v[d], s.r[pos[d]]);
                    }
                }
            }
            __syncthreads();
        }
    }
    double sum = 0.0;
    if constexpr (G == kWave) {
        /* the same sequential sum: each square formed on its row's lane,
         * the terms taken in row order by readlane (no LDS round trip per
         * row on one lane) */
        for (int i0 = 0; i0 < m; i0 += kWave) {
            const int i = i0 + lane;
            const double v = i < m ? s.r[i] : 0.0;
            const double sq = __dmul_rn(v, v);
            const int cnt = m - i0 < kWave ? m - i0 : kWave;
            for (int l = 0; l < cnt; ++l) sum = __dadd_rn(sum, Gr::lane_value(sq, l));
        }
    } else {
        if (lane == 0) {
            for (int i = 0; i < m; ++i) {
                double v = s.r[i];
                sum = __dadd_rn(sum, __dmul_rn(v, v));
            }
        }
        sum = Gr::bcast0(sum);
    }
    return sqrt(sum);
}

/* The candidate scores of augment with the numerators r^T a_j summed over
 * the rows of I instead of over the columns of the candidates: row i of I
 * adds r_i a_ij to the numerator of every candidate j of its row, the rows
 * in increasing order, so each numerator takes its terms in the order of
 * its column's stored entries restricted to I -- the column-wise sum, the
 * same bits -- while only the entries of I's rows are read (a candidate
 * column's entries outside I are never touched).
 *   1. The raw list (the row patterns of I then of col, from the union) is
 *      turned in place into candidate slots: rank of j in the bitmap of
 *      candidates (word ranks in the words after the bitmap), -1 for j in J.
 *   2. Numerators (in the score slots, zeroed): rows of I in order, eight
 *      rows' slots and values loaded together, then each row's entries
 *      (distinct slots) updated in turn.
 *   3. rho_j from the numerator and ||a_j||, in its slot.
 * `rp_val` holds C's values in row-pattern order; rows hold at most G
 * entries. */
template <int G, bool kMixed, int kA>
__device__ __forceinline__ void scores_by_rows(const BuildArgs& a, const WaveScratch& s, int m, int c, int off,
                                               int vmin, int words, unsigned* bm, double rn, double rn2,
                                               bool* rho_near0, double* acc4)
{
    using Gr = Group<G>;
    constexpr int kD = 8;  // rows whose loads are in flight together
    const int lane = Gr::lane();
    bitmap_ranks<G>(bm, words);
    for (int k = lane; k < off; k += G) s.sbuf[k] = bitmap_rank(bm, words, s.sbuf[k] - vmin);
    __syncthreads();
    int run = 0;
    double* const num = s.rho;  // the word ranks are dead now
    for (int t = lane; t < c; t += G) num[t] = 0.0;
    __syncthreads();
    for (int ip0 = 0; ip0 < m; ip0 += G) {
        const int ip = ip0 + lane;
        int len = 0, base = 0;
        double rr = 0.0;
        if (ip < m) {
            const int i = s.I[ip];
            base = a.rp_ptr[i];
            len = a.rp_ptr[i + 1] - base;
            rr = s.r[ip];
        }
        const int incl = Gr::scan_incl(len);
        const int roff = run + incl - len;  // the row's entries in the raw list
        run += Gr::bcast_last(incl);
        const int rows = m - ip0 < G ? m - ip0 : G;
        for (int q0 = 0; q0 < rows; q0 += kD) {
            int slot[kD];
            double val[kD];
#pragma unroll
            for (int d = 0; d < kD; ++d) {
                slot[d] = -1;
                val[d] = 0.0;
                const int q = q0 + d;
                if (q < rows) {
                    /* the row's numbers from its lane while every lane of the
                     * group is active (16-lane groups read them by permute,
                     * which sees only active lanes) */
                    const int ln = Gr::lane_value_i(len, q), ro = Gr::lane_value_i(roff, q);
                    const int bs = Gr::lane_value_i(base, q);
                    if (lane < ln) {
                        slot[d] = s.sbuf[ro + lane];
                        val[d] = a.rp_val[bs + lane];
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < kD; ++d) {
                const int q = q0 + d;
                if (q < rows) {
                    const double rq = Gr::lane_value(rr, q);
                    if (slot[d] >= 0) num[slot[d]] = __dadd_rn(num[slot[d]], __dmul_rn(rq, val[d]));
                }
            }
        }
    }
    __syncthreads();
    for (int t0 = lane; t0 < c; t0 += kA * G) {
        double ajn[kA];
#pragma unroll
        for (int k = 0; k < kA; ++k) {
            const int t = t0 + k * G;
            ajn[k] = t < c ? a.ajn[s.jt[t]] : 1.0;
        }
#pragma unroll
        for (int k = 0; k < kA; ++k) {
            const int t = t0 + k * G;
            if (t < c) {
                const double nm = num[t];
                const double n2 = __dmul_rn(nm, nm);
                const double rho = __dsub_rn(rn2, __ddiv_rn(n2, __dmul_rn(ajn[k], ajn[k])));
                if (kMixed) *rho_near0 |= rho <= kGuard * rn * rn;
                s.rho[t] = sqrt(rho);
                if (acc4) {
                    const int j = s.jt[t];
                    *acc4 += 16.0 + 12.0 * (a.A.colptr[j + 1] - a.A.colptr[j]);
                }
            }
        }
    }
}

/* Augment_Sparsity (src/Spai_Sub.imp:1010-1116) with Union_Nl_Sets and
 * Compute_Rhos.  Returns the new |J|, -1 when J~ is empty, or kOverflow.
 * Without kTies a tie of scores across the cut also returns kOverflow: the
 * std::sort emulation that orders such ties lives only in the generic
 * kernel, which keeps the size bins' register budget small (ties at the cut
 * are rare: none in the five workloads, a handful in the CSparse tests). */
template <int G, bool kMixed, bool kTies, int kB, int kR, int kA>
__device__ __forceinline__ int augment(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s, int col,
                                       int m, int n, double rn, int* st, double* wk, int* near = nullptr)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    /* L = I u {col}: append col when absent (the union does not need order) */
    int pos = lower_bound_i(s.I, m, col);
    bool col_in = pos < m && s.I[pos] == col;
    int nl = col_in ? m : m + 1;
    double acc3 = 0.0, acc4 = 0.0;  // work counters: exact integers, any order
    /* size of the raw candidate list */
    int total = 0;
    for (int li = lane; li < nl; li += G) {
        int l = li < m ? s.I[li] : col;
        int len = a.rp_ptr[l + 1] - a.rp_ptr[l];
        total += len;
        acc3 += 8.0 + 4.0 * len;
    }
    total = Gr::sum_int(total);
    acc3 = Gr::sum_fixed(acc3);
    if (total > L.capSort) return kOverflow;
    /* gather the row patterns of L (and the range of their indices) */
    int off = 0, vmin = 0x7fffffff, vmax = -1;
    for (int base = 0; base < nl; base += G) {
        int li = base + lane;
        int l = li < m ? s.I[li] : col;
        int len = 0, e0 = 0;
        if (li < nl) {
            e0 = a.rp_ptr[l];
            len = a.rp_ptr[l + 1] - e0;
        }
        int incl = Gr::scan_incl(len);
        int excl = incl - len;
        /* four loads in flight per round instead of one */
        for (int e = 0; e < len; e += 4) {
            int v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = e + q < len ? a.rp_ind[e0 + e + q] : 0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (e + q < len) {
                    s.sbuf[off + excl + e + q] = v[q];
                    vmin = min(vmin, v[q]);
                    vmax = max(vmax, v[q]);
                }
        }
        off += Gr::bcast_last(incl);
    }
    __syncthreads();
    vmin = Gr::min_int(vmin);
    vmax = Gr::max_int(vmax);
    /* J~ = U \ J in increasing order: through a bitmap over [vmin, vmax] in
     * LDS when the range fits (the score region, free until the scores, in a
     * size bin; the hash region in the generic kernel), else by sorting the
     * raw list */
    int avail = 0;
    unsigned* const bm = bitmap_words<kTies>(L, s, &avail);
    int c = 0;
    bool by_rows = false;
    if (off > 0 && ((vmax - vmin) >> 5) < avail) {
        c = bitmap_unique<G>(s.sbuf, off, vmin, vmax, s.J, n, s.jt, L.capU, bm);
        if (c == kOverflow) return kOverflow;
        /* the raw list is intact and the bitmap leaves room for its word
         * ranks: the numerators can run over the rows of I */
        by_rows = !kTies && a.rp_val != nullptr && a.rp_maxlen <= G && 2 * (((vmax - vmin) >> 5) + 1) <= avail;
    } else {
        int nu = group_sort_unique<G>(s.sbuf, off, s.sbuf);
        if (nu > L.capU) return kOverflow;
        for (int base = 0; base < nu; base += G) {
            int t = base + lane;
            int u = t < nu ? s.sbuf[t] : 0;
            bool keep = false;
            if (t < nu) {
                int p = lower_bound_i(s.J, n, u);
                keep = !(p < n && s.J[p] == u);
            }
            unsigned long long mask = Gr::ballot(keep);
            int q = c + __popcll(mask & Gr::lanes_below());
            if (keep) s.jt[q] = u;
            c += __popcll(mask);
        }
        __syncthreads();
    }
    if (kTies && a.up_ptr && c > 0) {
        /* J~ n UP(:,col) (Set_Intersection, src/Spai_Sub.imp:1072-1094), in
         * the generic kernel only (builds with an upper pattern run
         * unbinned; the size-bin kernels keep their register budgets): each
         * candidate is searched in the column of the upper pattern, the kept
         * ones compacted in order (a lane's write never passes a slot not yet
         * read).  The scores then run over the candidates' columns. */
        const int u0 = a.up_ptr[col], ulen = a.up_ptr[col + 1] - u0;
        int kept = 0;
        for (int base = 0; base < c; base += G) {
            const int t = base + lane;
            const int j = t < c ? s.jt[t] : 0;
            bool keep = false;
            if (t < c && ulen > 0) {
                const int p = lower_bound_i(a.up_ind + u0, ulen, j);
                keep = p < ulen && a.up_ind[u0 + p] == j;
            }
            const unsigned long long mask = Gr::ballot(keep);
            if (keep) s.jt[kept + __popcll(mask & Gr::lanes_below())] = j;
            kept += __popcll(mask);
        }
        __syncthreads();
        c = kept;
        by_rows = false;
    }
    if (c == 0) {
        if (wk && lane == 0) wk[3] += acc3;
        return -1;
    }
    /* rho_j = sqrt(||r||^2 - (r^T a_j)^2 / ||a_j||^2), numerator over I only.
     * First the column bounds of every candidate of the lane, kA columns'
     * loads in flight at a time, parked in the candidate's score slot (the
     * lane that parks a slot reads it back and then overwrites it with the
     * score).  Then kB candidates at a time (t, t + G, ...): up to kR row
     * indices of each per round, all in flight together; only the entries
     * whose row is in I (a small share) load their value.  Every numerator
     * still adds its column's entries in order. */
    bool rho_near0 = false;  // mixed path: rho^2 within rounding of 0 (its sign decides NaN or not)
    const double rn2 = __dmul_rn(rn, rn);
    if (by_rows) {
        scores_by_rows<G, kMixed, kA>(a, s, m, c, off, vmin, ((vmax - vmin) >> 5) + 1, bm, rn, rn2, &rho_near0,
                                   wk ? &acc4 : nullptr);
    } else {
    const RowHash h = row_hash_build<G>(s, L, m, a.A.nrows);
    int2* const bounds = reinterpret_cast<int2*>(s.rho);
    for (int t0 = lane; t0 < c; t0 += kA * G) {
        int lo[kA], hi[kA];
#pragma unroll
        for (int k = 0; k < kA; ++k) {
            const int t = t0 + k * G;
            lo[k] = hi[k] = 0;
            if (t < c) {
                const int j = s.jt[t];
                lo[k] = a.A.colptr[j];
                hi[k] = a.A.colptr[j + 1];
            }
        }
#pragma unroll
        for (int k = 0; k < kA; ++k)
            if (t0 + k * G < c) bounds[t0 + k * G] = make_int2(lo[k], hi[k] - lo[k]);
    }
    for (int t0 = lane; t0 < c; t0 += kB * G) {
        int e0[kB], len[kB];
        double ajn[kB], num[kB];
        int lmax = 0;
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            const int t = t0 + k * G;
            e0[k] = 0;
            len[k] = 0;
            ajn[k] = 1.0;
            num[k] = 0.0;
            if (t < c) {
                const int2 bd = bounds[t];
                e0[k] = bd.x;
                len[k] = bd.y;
                ajn[k] = a.ajn[s.jt[t]];
            }
            lmax = max(lmax, len[k]);
        }
        for (int base = 0; base < lmax; base += kR) {
            /* with several candidates per lane (the wide LDS kernels, a few
             * waves per CU, latency-bound) the values load with the row
             * indices instead of waiting for the lookups, one memory round
             * trip instead of two; elsewhere only the hits load a value
             * (fewer bytes for the many resident waves) */
            constexpr bool kEager = kB > 1;
            int pos[kB][kR];
            double v[kB][kR];
#pragma unroll
            for (int k = 0; k < kB; ++k)
#pragma unroll
                for (int q = 0; q < kR; ++q) {
                    const bool in = base + q < len[k];
                    pos[k][q] = in ? a.A.rowind[e0[k] + base + q] : -1;
                    v[k][q] = (kEager && in) ? a.A.val[e0[k] + base + q] : 0.0;
                }
#pragma unroll
            for (int k = 0; k < kB; ++k)
#pragma unroll
                for (int q = 0; q < kR; ++q)
                    if (pos[k][q] >= 0) pos[k][q] = row_position(h, s.I, m, pos[k][q]);
            if constexpr (!kEager) {
#pragma unroll
                for (int k = 0; k < kB; ++k)
#pragma unroll
                    for (int q = 0; q < kR; ++q)
                        if (pos[k][q] >= 0) v[k][q] = a.A.val[e0[k] + base + q];
            }
#pragma unroll
            for (int k = 0; k < kB; ++k)
#pragma unroll
                for (int q = 0; q < kR; ++q)
                    if (pos[k][q] >= 0) num[k] = __dadd_rn(num[k], __dmul_rn(s.r[pos[k][q]], v[k][q]));
        }
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            const int t = t0 + k * G;
            if (t < c) {
                const double n2 = __dmul_rn(num[k], num[k]);
                const double rho = __dsub_rn(rn2, __ddiv_rn(n2, __dmul_rn(ajn[k], ajn[k])));
                if (kMixed) rho_near0 |= rho <= kGuard * rn * rn;
                s.rho[t] = sqrt(rho);
                acc4 += 16.0 + 12.0 * len[k];
            }
        }
    }
    }
    __syncthreads();
    if (kMixed && Gr::ballot(rho_near0) != 0ull) return kFallback;
    acc4 = Gr::sum_fixed(acc4);
    if (wk && lane == 0) {
        wk[3] += acc3;
        wk[4] += acc4;
        wk[5] += (double)c;
    }
    /* The threshold is the reference's sequential sum of the scores (over
     * the candidates in increasing j), divided by c with -um 1.  A parallel
     * sum differs from it by at most 2 gamma_{c-1} sum|rho| (both orders are
     * within gamma_{c-1} sum|rho| of the exact sum; rho >= 0), and the
     * division adds at most one rounding to each: with
     * band = 4 (c + 1) u threshold, no score outside the band around the
     * parallel threshold compares differently with the sequential one.  The
     * sequential sum runs only when some score lies inside the band (or in
     * builds with counters, whose near-margin count is exact). */
    double part = 0.0;
    for (int t = lane; t < c; t += G) part += s.rho[t];
    double mean = Gr::sum_fixed(part);
    if (a.use_mean) mean = __ddiv_rn(mean, (double)c);
    bool exact = wk != nullptr || !isfinite(mean);
    if (!exact) {
        const double band = 4.0 * (double)(c + 1) * 0x1p-53 * mean;
        bool inside = false;
        for (int t = lane; t < c; t += G) inside |= fabs(s.rho[t] - mean) <= band;
        exact = Gr::ballot(inside) != 0ull;
    }
    if (exact) {
        double seq = 0.0;
        if (lane == 0) {
            for (int t = 0; t < c; ++t) seq = __dadd_rn(seq, s.rho[t]);
            if (a.use_mean) seq = __ddiv_rn(seq, (double)c);
        }
        mean = Gr::bcast0(seq);
    }
    if (isnan(mean)) {
        if (kMixed) return kFallback;
        *st |= COL_NAN_RHO;
        return n;
    }
    if (!kMixed && near) {
        /* scores within kNearMargin of the mean threshold */
        int cnt = 0;
        for (int t = lane; t < c; t += G) cnt += fabs(s.rho[t] - mean) <= kNearMargin * fabs(mean);
        near[0] += Gr::sum_int(cnt);
    }
    if (kMixed) {
        /* no score within the guard band of the mean threshold */
        bool close = false;
        const double band = kGuard * rn * rn;
        for (int t = lane; t < c; t += G) close |= fabs(s.rho[t] - mean) * (s.rho[t] + mean) <= band;
        if (Gr::ballot(close) != 0ull) return kFallback;
    }
    /* q = #{rho <= mean}; the first min(maxnew, q) after std::sort are taken */
    int q = 0;
    for (int t = lane; t < c; t += G) q += (s.rho[t] <= mean);
    q = Gr::sum_int(q);
    int sneed = a.maxnew < q ? a.maxnew : q;
    if (sneed <= 0) return n;
    if (n + sneed > L.capJ) return kOverflow;
    /* the sneed-th smallest score v* and the number of scores below it */
    double prev = -1.0, vstar = 0.0;
    int below = 0, ties = 0;
    for (int it = 0; it < c; ++it) {
        double v = INFINITY;
        for (int t = lane; t < c; t += G)
            if (s.rho[t] > prev) v = fmin(v, s.rho[t]);
        v = Gr::min(v);
        int cnt = 0;
        for (int t = lane; t < c; t += G) cnt += (s.rho[t] == v);
        cnt = Gr::sum_int(cnt);
        if (below + cnt >= sneed || cnt == 0) {
            vstar = v;
            ties = cnt;
            break;
        }
        below += cnt;
        prev = v;
    }
    int take_ties = sneed - below;
    if (!kMixed && near && q > sneed) {
        /* the cut after the sneed-th score: a tie across it, or the next
         * score within kNearMargin */
        double vnext = INFINITY;
        for (int t = lane; t < c; t += G)
            if (s.rho[t] > vstar) vnext = fmin(vnext, s.rho[t]);
        vnext = Gr::min(vnext);
        near[1] += (take_ties != ties || vnext - vstar <= kNearMargin * vstar) ? 1 : 0;
    }
    if (kMixed && q > sneed) {
        /* the cut after the sneed-th score must be clear of the next score */
        if (take_ties != ties) return kFallback;
        double vnext = INFINITY;
        for (int t = lane; t < c; t += G)
            if (s.rho[t] > vstar) vnext = fmin(vnext, s.rho[t]);
        vnext = Gr::min(vnext);
        if ((vnext - vstar) * (vnext + vstar) <= kGuard * rn * rn) return kFallback;
    }
    if (take_ties == ties) {
        int ns = 0;
        for (int base = 0; base < c; base += G) {
            int t = base + lane;
            bool keep = t < c && s.rho[t] <= vstar;
            unsigned long long mask = Gr::ballot(keep);
            int q2 = ns + __popcll(mask & Gr::lanes_below());
            if (keep) s.sel[q2] = s.jt[t];
            ns += __popcll(mask);
        }
    } else {
        /* a tie straddles the cut: reproduce std::sort's order */
        if (!kTies || c > L.capPairs) return kOverflow;
        if (lane == 0) {
            for (int t = 0; t < c; ++t) {
                s.pairs[t].idx = s.jt[t];
                s.pairs[t].rho = s.rho[t];
            }
            std_sort_rho(s.pairs, s.pairs + c, s.stack);
            for (int t = 0; t < sneed; ++t) s.sel[t] = s.pairs[t].idx;
            for (int t = 1; t < sneed; ++t) {
                int v = s.sel[t], u = t - 1;
                while (u >= 0 && s.sel[u] > v) {
                    s.sel[u + 1] = s.sel[u];
                    --u;
                }
                s.sel[u + 1] = v;
            }
        }
    }
    __syncthreads();
    /* J <- sort(J u selected): merge two sorted lists (lane 0, |J| is small) */
    if (lane == 0) {
        int i = n - 1, t = sneed - 1, w = n + sneed - 1;
        while (t >= 0) {
            if (i >= 0 && s.J[i] > s.sel[t])
                s.J[w--] = s.J[i--];
            else
                s.J[w--] = s.sel[t--];
        }
    }
    __syncthreads();
    return n + sneed;
}

/* Builds with counters: add a finished column's work counters, near-margin
 * counts and |J| to its launch's totals (one atomic per counter). */
__device__ __forceinline__ void launch_totals(const BuildArgs& a, const double* wsum, const int* near, int lane,
                                              int nnz)
{
    if (lane < kNumWork) atomicAdd(&a.lwork[lane], wsum[lane]);
    if (lane == 0) {
        atomicAdd(&a.lwork[kNumWork], 1.0);
        atomicAdd(&a.lwork[kNumWork + 4], (double)nnz);
        if (near)
            for (int q = 0; q < 3; ++q)
                if (near[q]) atomicAdd(&a.lwork[kNumWork + 1 + q], (double)near[q]);
    }
}

template <int G>
__device__ __forceinline__ void count_solve(const BuildArgs& a, const WaveScratch& s, int m, int n, double* wk)
{
    (void)a;
    if (!wk || Group<G>::lane() != 0) return;
    double dm = (double)m, dn = (double)n;
    double kk = dm < dn ? dm : dn;
    wk[0] = __dadd_rn(wk[0], __dsub_rn(2.0 * dm * dn * dn, (2.0 / 3.0) * kk * kk * kk));
    wk[1] = __dadd_rn(wk[1], (4.0 * dm * dn - 2.0 * dn * dn) + dn * dn + 2.0 * dm * dn);
    /* sum over J of (8 + 12 nnz(a_j)): exact integers */
    wk[2] += 8.0 * dn + 12.0 * (double)s.segoff[n];
    wk[6] += dm;
    wk[7] += dn;
}

template <int G, bool kMixed, int NMAX, bool kHbm>
__device__ __forceinline__ int backward_column(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                               int col)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int ld = L.capI;
    int p0 = a.p_ptr[col], n = a.p_ptr[col + 1] - p0;
    if (n > L.capJ) return kMoveOn;
    if (n == 0) return kDone;
    for (int i = lane; i < n; i += G) {
        s.J[i] = a.p_ind[p0 + i];
        s.y[i] = a.m_in[p0 + i];
        s.g[i] = a.gM[p0 + i];
    }
    __syncthreads();
    int m = build_I<G, kHbm>(a, L, s, n);
    if (m == kOverflow) return kMoveOn;
    if (m < n) return kDone;  // degenerate column: no gradient
    gather_ahat<G, kMixed && NMAX == 0>(a, L, s, m, n);
    make_rhs<G>(a, s, m, col);
    /* the register QRs leave A(I,J) intact in LDS: products from it */
    const bool regs = NMAX > 0 && m <= G && n <= NMAX;
    const double* R = s.aqr;
    int ldr = ld;
    if (kMixed) {
        if (regs) {
            if (qr_regs_f32<G, (NMAX > 0 ? NMAX : 4)>(L, s, m, n) == kFallback) return kRedoFp64;
        } else if (NMAX > 0 || qr_f32<G>(L, s, m, n) == kFallback) {
            return kRedoFp64;  // (register bins hold register columns only)
        }
    } else if (regs) {
        int st2;
        if constexpr (G == kWave && NMAX == 32)
            st2 = ls_solve_rows64<true>(L, s, m, n);
        else
            st2 = ls_solve_regs<G, (NMAX > 0 ? NMAX : 4), true>(L, s, m, n);
        if (st2 & (COL_SINGULAR | COL_UNDERDET)) return kDone;
        R = s.rq;
        ldr = L.capJ;
    } else {
        if (ls_solve<G>(L, s, m, n) & (COL_SINGULAR | COL_UNDERDET)) return kDone;
    }
    const bool dense = kMixed || regs;
    /* r = b - C m (into s.r) */
    csc_apply<G>(a, L, s, m, n, s.y, s.b0, s.r, dense);
    int rst = 1;  // 1: not refined in registers
    if constexpr (kMixed && NMAX > 0) {
        /* 16- and 64-lane groups, and 32-lane groups with |J| <= 8 */
        if (regs && (G != 32 || n <= 8)) rst = refine_adjoint_regs<G, (G == 64 ? 32 : 8)>(L, s, m, n);
        if (rst == kFallback) return kRedoFp64;
    }
    if (kMixed && rst != 0) {
        for (int i = lane; i < n; i += G) s.z[i] = s.g[i];
        __syncthreads();
        normal_solve<G>(s.aq32, L.ldq, s.g, n);
        bool ok = false;
        for (int it = 0; it < kRefineSteps; ++it) {
            csc_apply<G>(a, L, s, m, n, s.g, nullptr, s.b0, true);  // u = C lambda
            csc_apply_t<G>(a, L, s, m, n, s.b0, s.b, true);          // C^T u
            for (int i = lane; i < n; i += G) s.b[i] = s.z[i] - s.b[i];
            __syncthreads();
            normal_solve<G>(s.aq32, L.ldq, s.b, n);
            double dmx = 0.0, lmx = 0.0;
            for (int i = lane; i < n; i += G) {
                double l = s.g[i] + s.b[i];
                s.g[i] = l;
                dmx = fmax(dmx, fabs(s.b[i]));
                lmx = fmax(lmx, fabs(l));
            }
            dmx = Gr::max(dmx);
            lmx = Gr::max(lmx);
            __syncthreads();
            if (dmx <= 1e-10 * lmx) {
                ok = true;
                break;
            }
        }
        if (!ok) return kRedoFp64;
    } else if (!kMixed) {
        normal_solve<G>(R, ldr, s.g, n);
    }
    /* s = C lambda (into s.b0) */
    csc_apply<G>(a, L, s, m, n, s.g, nullptr, s.b0, dense);
    /* dL/dC_ij += r_i lambda_j - s_i m_j over the stored entries of J (the
     * column segments build_I left) */
    const int total = s.segoff[n];
    for (int e = lane; e < total; e += G) {
        const int t = s.sid[e];
        const int src = s.segs[t] + e - s.segoff[t];
        const int pos = lower_bound_i(s.I, m, a.A.rowind[src]);
        atomicAdd(&a.gC[src], __fma_rn(s.r[pos], s.g[t], -s.b0[pos] * s.y[t]));
    }
    if (a.gB && a.B.colptr) {
        int e0 = a.B.colptr[col], cnt = a.B.colptr[col + 1] - e0;
        for (int e = lane; e < cnt; e += G) {
            int row = a.B.rowind[e0 + e];
            int pos = lower_bound_i(s.I, m, row);
            a.gB[e0 + e] = (pos < m && s.I[pos] == row) ? s.b0[pos] : 0.0;
        }
    }
    __syncthreads();
    return kDone;
}

/* One column (Spai_Unrestrained<T>::SPAI_Column).  kMoveOn: it does not fit
 * this launch's layout; kRedoFp64: the mixed path declined a decision. */
/*
 * Spai_QRUpdate<T>::SPAI_Column with the dense update (-qr 2,
 * src/Spai_QRUpdate.imp:33-215, QRZerlegungLapack and qrupdate_voll in
 * src/qrupdates.cpp:37-327).  The factor of A(I,J) is kept across the
 * augmentation steps in the group's working set: rows are stored in the order
 * they arrived (sorted I, then each step's new rows at the bottom), columns in
 * the order they were added (new columns at the right).  A step applies the
 * existing reflectors to the new columns over the rows each reflector spans,
 * factors the block below the old R, re-applies every reflector to the stored
 * right-hand side and back-substitutes row by row (sequential, unfused, as
 * qrupdate_voll does); the solution is put back in ascending column order.
 */
template <int G, bool kTies>
__device__ __forceinline__ int solve_column_qrup(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                                 int col)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int ld = L.capI;
    double* wk = a.work ? s.wsum : nullptr;  // summed in LDS, stored with the column
    int p0 = a.p_ptr[col], n = a.p_ptr[col + 1] - p0;
    if (n > L.capJ) return kMoveOn;
    if (wk && lane < kNumWork) wk[lane] = 0.0;
    for (int i = lane; i < n; i += G) s.J[i] = a.p_ind[p0 + i];
    __syncthreads();
    int st = 0, solves = 0;
    /* first factorization: QRZerlegungLapack */
    int m = build_I<G, kTies>(a, L, s, n);
    if (m == kOverflow) return kMoveOn;
    st |= make_rhs<G>(a, s, m, col);
    gather_ahat<G, false>(a, L, s, m, n);
    for (int i = lane; i < m; i += G) {
        s.srow[i] = s.I[i];
        s.spos[i] = i;
        s.bperm[i] = s.b[i];
        s.yq[i] = s.b[i];
    }
    for (int c = lane; c < n; c += G) {
        s.colid[c] = s.J[c];
        s.reflrows[c] = m;
        s.tau[c] = 0.0;
    }
    __syncthreads();
    int k = m < n ? m : n;
    for (int i = 0; i < k; ++i) {
        double* ai = s.aqr + i * ld;
        double t = make_reflector<G>(ai, i, m);
        if (lane == 0) s.tau[i] = t;
        if (t != 0.0)
            for (int j = i + 1; j < n; ++j) apply_reflector<G>(ai, s.aqr + j * ld, i, m, t);
        __syncthreads();
    }
    for (int i = 0; i < k; ++i) {
        double t = s.tau[i];
        if (t != 0.0) apply_reflector<G>(s.aqr + i * ld, s.yq, i, m, t);
    }
    __syncthreads();
    if (m < n) {
        for (int i = lane; i < n; i += G) s.b[i] = 0.0;
        st |= COL_UNDERDET;
        __syncthreads();
    } else {
        bool zero = false;
        for (int i = lane; i < n; i += G) zero |= (s.aqr[i * ld + i] == 0.0);
        if (Gr::ballot(zero) != 0ull) {
            st |= COL_SINGULAR;
            for (int i = lane; i < n; i += G) s.b[i] = s.yq[i];
            __syncthreads();
        } else {
            for (int K = n - 1; K >= 0; --K) {
                double xk = s.yq[K];
                if (xk != 0.0) {
                    xk = xk / s.aqr[K * ld + K];
                    double nx = -xk;
                    __syncthreads();
                    if (lane == 0) s.yq[K] = xk;
                    for (int i = lane; i < K; i += G) s.yq[i] = __fma_rn(nx, s.aqr[K * ld + i], s.yq[i]);
                }
                __syncthreads();
            }
            for (int i = lane; i < n; i += G) s.b[i] = s.yq[i];
            __syncthreads();
        }
    }
    ++solves;
    count_solve<G>(a, s, m, n, wk);
    double rn = residual_norm<G>(a, L, s, m, n, false);
    int rows = m, cols = n;
    for (int step = 0; step < a.ns && rn > a.epsilon; ++step) {
        /* keep J, I and the storage positions of I */
        for (int i = lane; i < n; i += G) s.jold[i] = s.J[i];
        for (int i = lane; i < m; i += G) {
            s.iold[i] = s.I[i];
            s.sposold[i] = s.spos[i];
        }
        __syncthreads();
        const int nold = n, mold = m;
        int nn = augment<G, false, kTies, (G == 64 && !kTies) ? 2 : 1, (G == 64 && !kTies) ? 16 : 8,
                         (G == 64 && !kTies) ? 8 : 4>(a, L, s, col, m, n, rn, &st, wk);
        if (nn == kOverflow) return kMoveOn;
        if (nn < 0) break;
        n = nn;
        if (n > L.capJ) return kMoveOn;
        m = build_I<G, kTies>(a, L, s, n);
        if (m == kOverflow) return kMoveOn;
        st |= make_rhs<G>(a, s, m, col);
        /* storage positions: old rows keep theirs, new rows go to the bottom in
         * ascending order; b_perm grows by e_col on the new rows */
        int newr = 0;
        for (int base = 0; base < m; base += G) {
            int i = base + lane;
            bool isnew = false;
            int sp = 0;
            if (i < m) {
                int p = lower_bound_i(s.iold, mold, s.I[i]);
                isnew = !(p < mold && s.iold[p] == s.I[i]);
                if (!isnew) sp = s.sposold[p];
            }
            unsigned long long mask = Gr::ballot(isnew);
            if (isnew) {
                sp = rows + newr + __popcll(mask & Gr::lanes_below());
                s.srow[sp] = s.I[i];
                s.bperm[sp] = (s.I[i] == col) ? 1.0 : 0.0;
            }
            if (i < m) s.spos[i] = sp;
            newr += __popcll(mask);
        }
        const int rows_new = rows + newr;
        if (rows_new > L.capI) return kMoveOn;
        /* old columns: zero on the new rows */
        for (int c = 0; c < cols; ++c)
            for (int r = rows + lane; r < rows_new; r += G) s.aqr[c * ld + r] = 0.0;
        __syncthreads();
        /* new columns J \ J_old in ascending order, at the right */
        int newc = 0;
        for (int t = 0; t < n; ++t) {
            int j = s.J[t];
            int p = lower_bound_i(s.jold, nold, j);
            if (p < nold && s.jold[p] == j) continue;
            int c = cols + newc;
            ++newc;
            double* cc = s.aqr + c * ld;
            for (int r = lane; r < rows_new; r += G) cc[r] = 0.0;
            __syncthreads();
            int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
            for (int e = lane; e < cnt; e += G) {
                int q = lower_bound_i(s.I, m, a.A.rowind[e0 + e]);
                cc[s.spos[q]] = a.A.val[e0 + e];
            }
            if (lane == 0) {
                s.colid[c] = j;
                s.reflrows[c] = rows_new;
                s.tau[c] = 0.0;
            }
            __syncthreads();
            for (int q = 0; q < cols; ++q) {
                double t = s.tau[q];
                if (t != 0.0) apply_reflector<G>(s.aqr + q * ld, cc, q, s.reflrows[q], t);
            }
        }
        const int cols_new = cols + newc;
        /* factor the block below the old R (dgeqrf on B2) */
        for (int c = cols; c < cols_new; ++c) {
            if (c >= rows_new) break;
            double* ac = s.aqr + c * ld;
            double t = make_reflector<G>(ac, c, rows_new);
            if (lane == 0) s.tau[c] = t;
            if (t != 0.0)
                for (int u = c + 1; u < cols_new; ++u) apply_reflector<G>(ac, s.aqr + u * ld, c, rows_new, t);
            __syncthreads();
        }
        rows = rows_new;
        cols = cols_new;
        /* Q^T b from the stored right-hand side */
        for (int r = lane; r < rows; r += G) s.yq[r] = s.bperm[r];
        __syncthreads();
        for (int c = 0; c < cols; ++c) {
            double t = s.tau[c];
            if (t != 0.0) apply_reflector<G>(s.aqr + c * ld, s.yq, c, s.reflrows[c], t);
        }
        __syncthreads();
        if (rows < cols) {
            st |= COL_UNDERDET;
            for (int i = lane; i < cols; i += G) s.yq[i] = 0.0;
            __syncthreads();
        } else if (lane == 0) {
            /* row-oriented back substitution (src/qrupdates.cpp:306-314) */
            for (int i = cols - 1; i >= 0; --i) {
                double z = 0.0;
                for (int j = i + 1; j < cols; ++j) z = __dadd_rn(z, __dmul_rn(s.aqr[j * ld + i], s.yq[j]));
                double d = s.aqr[i * ld + i];
                if (d == 0.0) st |= COL_SINGULAR;
                s.yq[i] = __ddiv_rn(__dsub_rn(s.yq[i], z), d);
            }
        }
        st = Gr::bcast0(st);
        __syncthreads();
        /* solution in ascending column order */
        for (int t = lane; t < n; t += G) {
            int j = s.J[t];
            int c = 0;
            while (s.colid[c] != j) ++c;
            s.b[t] = s.yq[c];
        }
        __syncthreads();
        ++solves;
        count_solve<G>(a, s, m, n, wk);
        rn = residual_norm<G>(a, L, s, m, n, false);
        if (step == a.ns - 1) break;
    }
    size_t o = (size_t)col * a.cap;
    for (int i = lane; i < n; i += G) {
        a.m_idx[o + i] = s.J[i];
        a.m_val[o + i] = s.b[i];
    }
    if (lane == 0) {
        a.m_len[col] = n;
        a.res[col] = rn;
        a.nsolve[col] = solves;
        a.status[col] = st | a.status_or;
    }
    __syncthreads();
    if (a.work && lane < kNumWork) a.work[(size_t)col * kNumWork + lane] = s.wsum[lane];
    if (a.lwork) launch_totals(a, s.wsum, nullptr, lane, n);
    return kDone;
}

/*
 * Device pattern->R cache (Spai_Caching, src/Spai_Caching.imp:45-332).
 * Key: the reference's value hash of A(I,J) (Cache<double>::Compute_Key,
 * src/Cache.cpp:42-57: ret = 31 ret + mix + lo32 ^ hi32 over the values in
 * column-major order), evaluated as the equivalent polynomial in parallel.
 * A slot holds (m, n, key, state) and, in a.cache_data, A(I,J), its
 * Householder factor, tau, the right-hand side, the solution and the solve's
 * status.  A hit needs an equal key and an exactly equal A(I,J) (the
 * Compare_A_Hat confirmation); with an exactly equal right-hand side the
 * stored solution is reused, otherwise the stored factor is applied to the new
 * right-hand side.  Either way the result is bitwise the uncached one.  Slots
 * are claimed with a compare-and-swap (0 empty -> 1 writing) and published
 * after a release fence (-> 2 ready); they are never overwritten within a
 * build, so readers need no lock.
 */
constexpr int kCacheProbes = 8;

__device__ __forceinline__ uint32_t hash_double(double x)
{
    unsigned long long b = __double_as_longlong(x);
    return (uint32_t)b ^ (uint32_t)(b >> 32);
}

__device__ __forceinline__ uint32_t pow31(uint32_t e)
{
    uint32_t r = 1u, b = 31u;
    while (e) {
        if (e & 1u) r *= b;
        b *= b;
        e >>= 1;
    }
    return r;
}

template <int G>
__device__ __forceinline__ uint32_t cache_key(const double* A, int ld, int m, int n)
{
    using Gr = Group<G>;
    const uint32_t mix = 0x3FC0A3D7u ^ 0x0A3D70A4u;  // lo32 ^ hi32 of 0.13
    const uint32_t N = (uint32_t)(m * n);
    uint32_t acc = 0u;
    for (int t = Gr::lane(); t < m * n; t += G) {
        int j = t / m, i = t - j * m;
        acc += (mix + hash_double(A[j * ld + i])) * pow31(N - 1u - (uint32_t)t);
    }
    return (uint32_t)Gr::allreduce((int)acc, [](int x, int y) { return (int)((uint32_t)x + (uint32_t)y); });
}

/* doubles per slot: A, factor (m x n each), tau, x (n), b (m) */
__device__ __forceinline__ size_t cache_slot_doubles(const BuildArgs& a)
{
    return 2 * (size_t)a.cache_capI * a.cache_capJ + 2 * (size_t)a.cache_capJ + (size_t)a.cache_capI;
}

/* Least squares with the cache.  On entry s.aqr holds A(I,J) and s.b the
 * right-hand side; on exit as ls_solve. */
template <int G>
__device__ __forceinline__ int ls_solve_cached(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                               int m, int n)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int ld = L.capI;
    if (m > a.cache_capI || n > a.cache_capJ || m == 0 || n == 0) return ls_solve<G>(L, s, m, n);
    const uint32_t key = cache_key<G>(s.aqr, ld, m, n);
    const size_t sd = cache_slot_doubles(a);
    const int mask = a.cache_slots - 1;
    int slot = (int)(key & (uint32_t)mask);
    int claimed = -1;
    for (int p = 0; p < kCacheProbes; ++p, slot = (slot + 1) & mask) {
        int state = __hip_atomic_load(&a.cache_state[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (state == 0) {
            int won = 0;
            if (lane == 0) won = atomicCAS(&a.cache_state[slot], 0, 1) == 0;
            won = Gr::bcast0(won);
            if (won) {
                claimed = slot;
                break;
            }
            continue;  // another group took it
        }
        if (state != 2 || a.cache_key[slot] != key || a.cache_m[slot] != m || a.cache_n[slot] != n) continue;
        const double* d = a.cache_data + (size_t)slot * sd;
        bool diff = false;
        for (int t = lane; t < m * n; t += G) {
            int j = t / m, i = t - j * m;
            diff |= __double_as_longlong(d[t]) != __double_as_longlong(s.aqr[j * ld + i]);
        }
        if (Gr::ballot(diff) != 0ull) continue;
        /* hit */
        const double* fac = d + (size_t)a.cache_capI * a.cache_capJ;
        const double* tau = fac + (size_t)a.cache_capI * a.cache_capJ;
        const double* x = tau + a.cache_capJ;
        const double* b = x + a.cache_capJ;
        bool bdiff = false;
        for (int i = lane; i < m; i += G) bdiff |= __double_as_longlong(b[i]) != __double_as_longlong(s.b[i]);
        const int st_solve = a.cache_status[slot];
        if (lane == 0) atomicAdd(a.cache_hits, 1);
        if (Gr::ballot(bdiff) == 0ull) {
            for (int i = lane; i < n; i += G) s.b[i] = x[i];
            __syncthreads();
            return st_solve;
        }
        for (int t = lane; t < m * n; t += G) {
            int j = t / m, i = t - j * m;
            s.aqr[j * ld + i] = fac[t];
        }
        __syncthreads();
        int k = m < n ? m : n;
        for (int i = 0; i < k; ++i)
            if (tau[i] != 0.0) apply_reflector<G>(s.aqr + i * ld, s.b, i, m, tau[i]);
        __syncthreads();
        if (m < n) {
            for (int i = lane; i < n; i += G) s.b[i] = 0.0;
            __syncthreads();
            return COL_UNDERDET;
        }
        if (st_solve & COL_SINGULAR) return COL_SINGULAR;
        for (int K = n - 1; K >= 0; --K) {
            double xk = s.b[K];
            if (xk != 0.0) {
                xk = xk / s.aqr[K * ld + K];
                double nx = -xk;
                __syncthreads();
                if (lane == 0) s.b[K] = xk;
                for (int i = lane; i < K; i += G) s.b[i] = __fma_rn(nx, s.aqr[K * ld + i], s.b[i]);
            }
            __syncthreads();
        }
        return 0;
    }
    if (claimed < 0) return ls_solve<G>(L, s, m, n);
    /* miss with a claimed slot: store A and b, solve, store the factor */
    double* d = a.cache_data + (size_t)claimed * sd;
    double* fac = d + (size_t)a.cache_capI * a.cache_capJ;
    double* tau = fac + (size_t)a.cache_capI * a.cache_capJ;
    double* x = tau + a.cache_capJ;
    double* b = x + a.cache_capJ;
    for (int t = lane; t < m * n; t += G) {
        int j = t / m, i = t - j * m;
        d[t] = s.aqr[j * ld + i];
    }
    for (int i = lane; i < m; i += G) b[i] = s.b[i];
    __syncthreads();
    int st_solve = ls_solve<G>(L, s, m, n);
    for (int t = lane; t < m * n; t += G) {
        int j = t / m, i = t - j * m;
        fac[t] = s.aqr[j * ld + i];
    }
    for (int i = lane; i < n; i += G) {
        tau[i] = s.tau[i];
        x[i] = s.b[i];
    }
    if (lane == 0) {
        a.cache_key[claimed] = key;
        a.cache_m[claimed] = m;
        a.cache_n[claimed] = n;
        a.cache_status[claimed] = st_solve;
    }
    __threadfence();
    __syncthreads();
    if (lane == 0) __hip_atomic_store(&a.cache_state[claimed], 2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return st_solve;
}

/* A column that outgrows the launch's layout during `step`, or whose step a
 * guard of the mixed path declines, leaves the state of that step's start
 * for the next launch: J in its output slot, |J| in m_len, the status bits,
 * nsolve = -(step + 2) as the mark, and in builds with counters the work
 * counters in their row and the near-margin counts packed into res (each
 * below 2^20; read back exactly).  Every value the step computes depends
 * only on these, so the resumed column is bitwise the column built in one
 * launch.  A mixed launch leaves the FP64 state: every earlier step cleared
 * the guards, so its decisions were FP64's (and none was within the
 * near-margin band, which lies inside the guard band); index sets, work
 * counters and the status bits a mixed step sets (the target column outside
 * I) do not depend on the precision.  Without save_state the mark is 0: the
 * next launch starts afresh. */
template <int G>
__device__ __forceinline__ void save_column_state(const BuildArgs& a, const WaveScratch& s, int col, int step, int n,
                                                  int st, const int (&near)[3], double wk, bool counters)
{
    const int lane = Group<G>::lane();
    const size_t o = (size_t)col * a.cap;
    if (!a.save_state) {
        if (lane == 0) a.nsolve[col] = 0;
        __syncthreads();
        return;
    }
    for (int i = lane; i < n; i += G) a.m_idx[o + i] = s.J[i];
    if (counters && lane < kNumWork) a.work[(size_t)col * kNumWork + lane] = wk;
    if (lane == 0) {
        a.m_len[col] = n;
        a.status[col] = st;
        a.nsolve[col] = -(step + 2);
        if (counters) a.res[col] = near[0] + near[1] * 0x1p20 + near[2] * 0x1p40;
    }
    __syncthreads();
}

template <int G, bool kMixed, int NMAX, bool kTies, bool kTail = false>
__device__ __forceinline__ int solve_column(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                            int col)
{
    /* NMAX 20 or 24 in a 64-lane bin: the row-group register QR with eight
     * rows per lane (|I| <= 128, FP64 builds) */
    constexpr bool kRows128 = G == kWave && !kMixed && (NMAX == 20 || NMAX == 24);
    using Gr = Group<G>;
    const int lane = Gr::lane();
    double* wk = a.work ? s.wsum : nullptr;  // summed in LDS, stored with the column
    int st = 0, step0 = -1, n;
    int near[3] = {0, 0, 0};  // near-margin decisions (counters builds, FP64)
    const size_t o = (size_t)col * a.cap;
    const int mark = a.nsolve[col];
    const int* jsrc;
    if (mark < 0) {
        /* resume the state an earlier launch of this precision left at the
         * start of step -mark - 2 (see save_column_state) */
        step0 = -mark - 2;
        n = a.m_len[col];
        st = a.status[col];
        jsrc = a.m_idx + o;
    } else {
        const int p0 = a.p_ptr[col];
        n = a.p_ptr[col + 1] - p0;
        jsrc = a.p_ind + p0;
    }
    if (n > L.capJ) {
        /* a saved state stays for the next launch when that one resumes it */
        if (mark < 0 && !a.save_state && lane == 0) a.nsolve[col] = 0;
        return kMoveOn;
    }
    if (wk) {
        if (lane < kNumWork) wk[lane] = mark < 0 ? a.work[(size_t)col * kNumWork + lane] : 0.0;
        if (mark < 0) {
            const double pk = a.res[col];
            near[2] = (int)(pk * 0x1p-40);
            near[1] = (int)((pk - near[2] * 0x1p40) * 0x1p-20);
            near[0] = (int)(pk - near[2] * 0x1p40 - near[1] * 0x1p20);
        }
    }
    for (int i = lane; i < n; i += G) s.J[i] = jsrc[i];
    __syncthreads();
    int solves = step0 + 1;
    double rn = 1.0e50;
    /* the state at the start of the current step, left for the next launch
     * if the column outgrows this one's layout during the step */
    int moved = -2, st0 = st, near0[3] = {near[0], near[1], near[2]};
    bool redo = false;  // mixed: a guard declined the step, which FP64 redoes
    double wk0 = 0.0;
    for (int step = step0; step < a.ns; ++step) {
        st0 = st;
        near0[0] = near[0];
        near0[1] = near[1];
        near0[2] = near[2];
        if (wk && lane < kNumWork) wk0 = wk[lane];
        /* the FP64 bins pack each entry's position in I for the gather (in
         * the mixed bulk kernel it measured slower: 72 B of spill per lane
         * instead of 48) */
        bool packed = false;
        int m = build_I<G, kTies>(a, L, s, n, (kTies || kMixed) ? nullptr : &packed);
        if (m == kOverflow) {
            moved = step;
            break;
        }
        /* the generic kernel (I in HBM) finds rows through the LDS hash */
        const RowHash hI = kTies ? row_hash_build<G>(s, L, m, a.A.nrows) : RowHash{nullptr, 0, 0};
        gather_ahat<G, kMixed && NMAX == 0, kTies>(a, L, s, m, n, hI, packed);
        st |= make_rhs<G>(a, s, m, col);
        bool dense = kMixed;  // the mixed QR leaves the FP64 A(I,J) intact
        if (kMixed) {
            if (ls_solve_mixed<G, NMAX>(a, L, s, m, n) == kFallback) {
                redo = true;
                moved = step;
                break;
            }
        } else if (NMAX > 0 && m <= (kRows128 ? 2 * G : G) && n <= NMAX) {
            /* the register QR costs less than a cache lookup (which reads the
             * stored block and factor back from HBM): -cs serves the other
             * solves; both give the same bits */
            if constexpr (kRows128)
                st |= ls_solve_rows64<false, NMAX / 4, 8>(L, s, m, n);
            else if constexpr (G == kWave && (NMAX == 32 || NMAX == 16))
                st |= ls_solve_rows64<false, NMAX / 4>(L, s, m, n);
            else
                st |= ls_solve_regs<G, (NMAX > 0 ? NMAX : 4)>(L, s, m, n);
            dense = true;
        } else if (a.cache_slots > 0) {
            st |= ls_solve_cached<G>(a, L, s, m, n);
        } else {
            /* eight columns per pass in the 64-lane LDS bins (registers to
             * spare at their few waves per CU), sixteen in the widest bins
             * (one wave per SIMD) */
            st |= ls_solve<G, (G == kWave && !kTies) ? (kTail ? 16 : 8) : 4>(L, s, m, n);
        }
        ++solves;
        count_solve<G>(a, s, m, n, wk);
        /* the generic kernel's copy of A(I,J) gives the dense residual after
         * any QR (the same bits as the sparse one); the size bins' sparse
         * residual finds its rows through a hash in the raw-list region
         * (dead after the LDS QR) */
        const double* Ad = (kTies && !dense) ? s.a0 : nullptr;
        const RowHash hR = (!kTies && !dense) ? row_hash_build<G>(s, L, m, a.A.nrows) : hI;
        rn = residual_norm<G>(a, L, s, m, n, dense || Ad, hR, Ad);
        if (kMixed && !(fabs(rn - a.epsilon) > kGuard * a.epsilon)) {
            redo = true;
            moved = step;
            break;
        }
        if (!kMixed && wk) near[2] += fabs(rn - a.epsilon) <= kNearMargin * a.epsilon;
        if (rn <= a.epsilon || step == a.ns - 1) break;
        /* candidates in flight per lane: more where registers are plentiful
         * (64-lane LDS kernels without the register QR) */
        constexpr bool kWide = G == 64 && (NMAX == 0 || kRows128) && !kTies;
        int nn = augment<G, kMixed, kTies, kWide ? 2 : 1, (G == 64 && !kTies) ? 16 : 8, kWide ? 8 : 4>(
            a, L, s, col, m, n, rn, &st, wk, wk ? near : nullptr);
        if (nn == kOverflow) {
            moved = step;  // J is unchanged: augment merges only once every check has passed
            break;
        }
        if (nn == kFallback) {
            redo = true;
            moved = step;
            break;
        }
        if (nn < 0) break;
        n = nn;
    }
    if (moved > -2) {
        if (kMixed || a.save_state) save_column_state<G>(a, s, col, moved, n, st0, near0, wk0, wk != nullptr);
        return redo ? kRedoFp64 : kMoveOn;
    }
    for (int i = lane; i < n; i += G) {
        a.m_idx[o + i] = s.J[i];
        a.m_val[o + i] = s.b[i];
    }
    if (lane == 0) {
        a.m_len[col] = n;
        a.res[col] = rn;
        a.nsolve[col] = solves;
        a.status[col] = st | a.status_or;
    }
    __syncthreads();
    if (a.work && lane < kNumWork) a.work[(size_t)col * kNumWork + lane] = s.wsum[lane];
    if (a.lwork) launch_totals(a, s.wsum, near, lane, n);
    return kDone;
}

/* Append the columns whose `flag` is set (one flag per group, held by the
 * group's lane 0) to out[*count..], one atomic per wave. */
__device__ __forceinline__ void wave_append(bool flag, int col, int* out, int* count)
{
    unsigned long long mask = __ballot(flag);
    if (mask == 0ull) return;
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(count, __popcll(mask));
    base = __builtin_amdgcn_readlane(base, 0);
    if (flag) out[base + __popcll(mask & ((1ull << lane_id()) - 1ull))] = col;
}

/* kMode: 0 FP64 build, 1 mixed build, 2 FP64 adjoint, 3 mixed adjoint,
 * 4 FP64 build with the dense QR update (-qr 2).  The second launch bound is
 * the resident waves per SIMD the register allocation must allow where LDS
 * would admit more than the registers otherwise do: 5 for the FP64 8-wide
 * 32-lane kernel, 4 for its other modes and the generic (HBM working set)
 * kernels, 3 for the FP64 64-lane register kernel (each measured faster
 * despite the spill; the mixed 8-wide kernel is slower at 5), and 3 for the
 * 16-lane and the 16-wide 32-lane builds, whose bins LDS holds at 10-12
 * waves per CU.  kWaves > 0 sets it for the instantiations binned_kernel
 * picks by layout: 3 for the 16-wide 64-lane register kernels (bulk bins),
 * 2 for the eight-row register kernels (|I| <= 128) and the FP64 32-wide one
 * of bins capped at 8 waves per CU, 1 for bins whose LDS holds one wave per
 * SIMD. */
template <int G, bool kLds, int kMode, int NMAX, int kWaves = 0>
__global__ __launch_bounds__(64, kWaves > 0 ? kWaves
                                 : (G == 32 && NMAX == 8) ? (kMode == 0 ? 5 : 4)
                                 : ((G == 64 && NMAX == 32) ? (kMode == 0 ? 3 : 2)
                                    : ((kLds && kMode < 2 && (G == 16 || (G == 32 && NMAX == 16))) ? 3
                                       : (!kLds ? 4 : 1)))) void spai_column_kernel(BuildArgs a, ScratchLayout L, const int* queue,
                                                         const int* nqueue, int* head)
{
    extern __shared__ __align__(16) char smem[];
    /* an empty queue ends the wave before anything else: the working-set
     * pointers below are live across the whole kernel, and what the register
     * budget spills of them would otherwise be stored by every wave.  The
     * length does not change during the launch (waves append to other
     * queues only). */
    const int nq = *nqueue;
    if (nq <= 0) return;
    constexpr int NG = kWave / G;
    const int g = Group<G>::index();
    char* base = kLds ? smem + (size_t)g * L.per_wave
                      : a.scratch + ((size_t)blockIdx.x * NG + g) * L.per_wave;
    char* tie = L.tie_in_main ? base : a.tie_scratch + ((size_t)blockIdx.x * NG + g) * L.tie_per_group;
    WaveScratch s = wave_scratch(base, tie, L);
    if (kLds) {
        /* the hash of the scores lives in the raw-list buffer, dead from the
         * candidate union to the next I-set: capSort slots (a power of two,
         * at least 2 capI in every bin, so that I always fits at half load) */
        s.htab = (unsigned*)s.sbuf;
        s.hslots = L.capSort;
    } else {
        s.htab = (unsigned*)smem;
        s.hslots = kGenericHashSlots;
    }
    /* columns per chunk: one per group, `rounds` rounds.  Every wave starts on
     * the chunk of its block index; when the queue fits in one chunk per wave
     * (at most L.max_rounds columns per group) that static split is all, else
     * further chunks follow through the queue atomic.  All waves take the
     * same address, so cheap columns want few, large chunks; expensive ones
     * (large bins) want small chunks against the tail. */
    const long long groups = (long long)gridDim.x * NG;
    const long long cap = L.max_rounds > 0 ? L.max_rounds : 1;
    const int rounds = (int)max(1LL, min(cap, ((long long)nq + groups - 1) / groups));
    const int chunk = NG * rounds;
    const long long first = (long long)gridDim.x * chunk;  // chunks handed out by block index
    for (long long pass = 0;; ++pass) {
        long long q0 = 0;
        if (pass == 0) {
            /* workgroups go round-robin to the 8 XCDs: give XCD x the x-th
             * contiguous eighth of the first chunks, so that neighbouring
             * columns (which share rows of A) meet in the same L2 */
            const int per_xcd = (int)(gridDim.x / kXcds);
            const int b = blockIdx.x;
            const long long slot = b < per_xcd * kXcds ? (long long)(b % kXcds) * per_xcd + b / kXcds : b;
            q0 = slot * chunk;
        } else {
            /* a plain load first: once the queue is drained the waves leave
             * without queueing on the atomic */
            int h = 0;
            if (lane_id() == 0) {
                h = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (first + h < nq) h = atomicAdd(head, chunk);
            }
            q0 = first + __builtin_amdgcn_readlane(h, 0);
        }
        if (q0 >= nq) break;
        for (int round = 0; round < rounds; ++round) {
            const long long q = q0 + round * NG + g;
            int col = -1;
            int outcome = kDone;
            if (q < nq) {
                col = queue[q];
                if (kMode < 2)
                    outcome = solve_column<G, kMode == 1, NMAX, !kLds, kWaves == 1>(a, L, s, col);
                else if (kMode < 4)
                    outcome = backward_column<G, kMode == 3, NMAX, !kLds>(a, L, s, col);
                else
                    outcome = solve_column_qrup<G, !kLds>(a, L, s, col);
            }
            __syncthreads();
            const bool leader = Group<G>::lane() == 0;
            wave_append(leader && outcome == kMoveOn, col, a.overflow, a.n_overflow);
            if (kMode == 1 || kMode == 3) wave_append(leader && outcome == kRedoFp64, col, a.fallback, a.n_fallback);
        }
    }
}

}  // namespace

/* Number of distinct rows of the columns J (nj <= 8), counting stops at
 * limit + 1: a merge of the sorted row lists on one thread. */
__device__ __forceinline__ int distinct_rows_upto(const BuildArgs& a, int p0, int nj, int limit)
{
    int head[8], end[8], cur[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        head[t] = end[t] = 0;
        cur[t] = 0x7fffffff;
        if (t < nj) {
            const int j = a.p_ind[p0 + t];
            head[t] = a.A.colptr[j];
            end[t] = a.A.colptr[j + 1];
            if (head[t] < end[t]) cur[t] = a.A.rowind[head[t]];
        }
    }
    int count = 0;
    while (count <= limit) {
        int mn = 0x7fffffff;
#pragma unroll
        for (int t = 0; t < 8; ++t) mn = min(mn, cur[t]);
        if (mn == 0x7fffffff) break;
        ++count;
#pragma unroll
        for (int t = 0; t < 8; ++t)
            if (cur[t] == mn) {
                ++head[t];
                cur[t] = head[t] < end[t] ? a.A.rowind[head[t]] : 0x7fffffff;
            }
    }
    return count;
}

/* Route every column to the first bin whose |J|, raw row-list and |I|
 * capacities hold its start pattern (|I| is counted exactly when the raw list
 * exceeds the bin's |I| and |J0| <= 8; a column that still does not fit,
 * e.g. after augmentation, moves on by the overflow path). */
constexpr int kClassifyBlock = 1024;

__global__ __launch_bounds__(kClassifyBlock) void classify_kernel(BuildArgs a, int nbins, BinEdges edges, int* queues,
                                                       int* counts, int* gqueue, int* gcount, int n, int* done,
                                                       int* routed)
{
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    int bin = nbins;  // generic
    if (col < n) {
        if (a.nsolve) a.nsolve[col] = 0;  // no saved state: the column starts from its start pattern
        int p0 = a.p_ptr[col], nj = a.p_ptr[col + 1] - p0;
        long raw = 0;
        for (int e = p0; e < p0 + nj; ++e) {
            int j = a.p_ind[e];
            raw += a.A.colptr[j + 1] - a.A.colptr[j];
        }
        /* distinct rows, counted on demand up to limit + 1 (a count above
         * `limit` is only a lower bound) */
        int rows = -1, limit = -1;
        for (int b = 0; b < nbins; ++b) {
            if (nj > edges.capJ[b] || raw > edges.capSort[b]) continue;
            if (raw > edges.capI[b] && nj <= 8) {
                if (rows < 0 || (rows > limit && edges.capI[b] > limit)) {
                    limit = edges.capI[b];
                    rows = distinct_rows_upto(a, p0, nj, limit);
                }
                if (rows > edges.capI[b]) continue;
            }
            bin = b;
            break;
        }
    }
    /* one append per queue and block: lane b of each wave reserves the
     * wave's share of queue b in LDS, then thread b adds the block's count
     * of queue b to the queue (a per-wave atomic on the few queue counters
     * serializes thousands of times) */
    const int lane = lane_id();
    const bool live = col < n;
    unsigned long long mine = 0ull, lanes_b = 0ull;
    for (int b = 0; b <= nbins; ++b) {
        const unsigned long long mb = __ballot(live && bin == b);
        if (bin == b) mine = mb;
        if (lane == b) lanes_b = mb;
    }
    constexpr int kQueues = sizeof(BinEdges::capJ) / sizeof(int) + 1;  // the bins and the generic queue
    __shared__ int bcount[kQueues], bbase[kQueues];
    if ((int)threadIdx.x <= nbins) bcount[threadIdx.x] = 0;
    __syncthreads();
    int woff = 0;
    if (lane <= nbins && lanes_b != 0ull) woff = atomicAdd(&bcount[lane], __popcll(lanes_b));
    __syncthreads();
    if ((int)threadIdx.x <= nbins && bcount[threadIdx.x] > 0)
        bbase[threadIdx.x] = atomicAdd((int)threadIdx.x < nbins ? counts + threadIdx.x : gcount, bcount[threadIdx.x]);
    __syncthreads();
    const int off = __shfl(woff, bin);
    if (live) {
        int* out = bin < nbins ? queues + (size_t)bin * n : gqueue;
        out[bbase[bin] + off + __popcll(mine & ((1ull << lane) - 1ull))] = col;
    }
    /* the last block to finish records the routed queue lengths */
    __shared__ int last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(done, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (last && (int)threadIdx.x <= nbins) {
        __threadfence();
        int* c = (int)threadIdx.x < nbins ? counts + threadIdx.x : gcount;
        routed[threadIdx.x] = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void copy_ints_kernel(const int* src, int* dst, int count)
{
    int i = threadIdx.x;
    if (i < count) dst[i] = src[i];
}

using ColumnKernel = void (*)(BuildArgs, ScratchLayout, const int*, const int*, int*);

/* The LDS column kernel for a bin: register-resident QR (FP64 or mixed
 * build) when a group holds every row of the bin (|I| <= G) and the bin's
 * |J| fits the unrolled width; 32-lane bins whose |J| is at most 8 take the
 * 8-wide unrolling (fewer registers, more resident waves). */
static ColumnKernel binned_kernel(const ScratchLayout& L, int group, int mode, int wcap)
{
    const bool regs = mode != 4 && register_bin(group, L.capI, L.capJ);
#define MSPAI_PICK(MODE)                                                                            \
    if (group == 16) return &spai_column_kernel<16, true, MODE, 0>;                                 \
    if (group == 32) return &spai_column_kernel<32, true, MODE, 0>;                                 \
    return &spai_column_kernel<64, true, MODE, 0>
    switch (mode) {
    case 0:
        if (group == 16) return regs ? &spai_column_kernel<16, true, 0, 8> : &spai_column_kernel<16, true, 0, 0>;
        if (group == 32)
            return regs ? (L.capJ <= 8 ? &spai_column_kernel<32, true, 0, 8> : &spai_column_kernel<32, true, 0, 16>)
                        : &spai_column_kernel<32, true, 0, 0>;
        /* a bin capped at 8 resident waves per CU takes the FP64 register
         * kernel compiled for 2 waves per SIMD (256 VGPRs, no spills)
         * instead of 3 (168 VGPRs, some spilled) */
        if (regs && wcap > 0 && wcap <= 8) return &spai_column_kernel<64, true, 0, 32, 2>;
        /* |J| <= 16: the row-group QR with 4 column slots (fewer registers,
         * less spilled at 3 waves per SIMD) */
        if (regs && L.capJ <= 16) return &spai_column_kernel<64, true, 0, 16, 3>;
        /* 64 < |I| <= 128 and |J| <= 24: the register QR with eight rows
         * per lane (A(I,J) stays in LDS for the dense residual) */
        if (group == 64 && L.capI > 64 && L.capI <= 128 && L.capJ <= 24)
            return L.capJ <= 20 ? &spai_column_kernel<64, true, 0, 20, 2> : &spai_column_kernel<64, true, 0, 24, 2>;
        /* a bin whose LDS holds one wave per SIMD at most: the kernel
         * compiled for that (sixteen columns per QR pass) */
        if (!regs && L.per_wave > (160u << 10) / 5) return &spai_column_kernel<64, true, 0, 0, 1>;
        return regs ? &spai_column_kernel<64, true, 0, 32> : &spai_column_kernel<64, true, 0, 0>;
    case 1:
        if (group == 16) return regs ? &spai_column_kernel<16, true, 1, 8> : &spai_column_kernel<16, true, 1, 0>;
        if (group == 32)
            return regs ? (L.capJ <= 8 ? &spai_column_kernel<32, true, 1, 8> : &spai_column_kernel<32, true, 1, 16>)
                        : &spai_column_kernel<32, true, 1, 0>;
        /* |J| <= 16: the register QR and refinement unrolled to 16 columns,
         * compiled for 3 waves per SIMD (the 32-wide kernel needs 253 VGPRs,
         * 2 waves per SIMD, where LDS would hold 11 per CU) */
        if (regs && L.capJ <= 16) return &spai_column_kernel<64, true, 1, 16, 3>;
        return regs ? &spai_column_kernel<64, true, 1, 32> : &spai_column_kernel<64, true, 1, 0>;
    case 2:
        if (group == 16) return regs ? &spai_column_kernel<16, true, 2, 8> : &spai_column_kernel<16, true, 2, 0>;
        if (group == 32)
            return regs ? (L.capJ <= 8 ? &spai_column_kernel<32, true, 2, 8> : &spai_column_kernel<32, true, 2, 16>)
                        : &spai_column_kernel<32, true, 2, 0>;
        return regs ? &spai_column_kernel<64, true, 2, 32> : &spai_column_kernel<64, true, 2, 0>;
    case 3:
        if (group == 16) return regs ? &spai_column_kernel<16, true, 3, 8> : &spai_column_kernel<16, true, 3, 0>;
        if (group == 32)
            return regs ? (L.capJ <= 8 ? &spai_column_kernel<32, true, 3, 8> : &spai_column_kernel<32, true, 3, 16>)
                        : &spai_column_kernel<32, true, 3, 0>;
        return regs ? &spai_column_kernel<64, true, 3, 32> : &spai_column_kernel<64, true, 3, 0>;
    default: MSPAI_PICK(4);
    }
#undef MSPAI_PICK
}

int binned_blocks_per_cu(const ScratchLayout& L, int group, int mode, int wcap)
{
    static std::mutex mu;
    static std::map<std::tuple<const void*, size_t>, int> memo;
    const ColumnKernel k = binned_kernel(L, group, mode, wcap);
    const size_t lds = L.per_wave * (kWave / group);
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple((const void*)k, lds);
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k, kWave, lds) != hipSuccess) nb = 0;
    memo[key] = nb;
    return nb;
}

void launch_binned(const BuildArgs& a, const ScratchLayout& L, int group, int mode, const int* queue,
                   const int* nqueue, int* head, int nwaves, hipStream_t stream, int wcap)
{
    const size_t lds = L.per_wave * (kWave / group);
    hipLaunchKernelGGL(binned_kernel(L, group, mode, wcap), dim3(nwaves), dim3(kWave), lds, stream, a, L, queue, nqueue,
                       head);
}

void launch_classify(const BuildArgs& a, int nbins, const BinEdges& edges, int* queues, int* counts,
                     int* gqueue, int* gcount, int n, int* done, int* routed, hipStream_t stream)
{
    hipLaunchKernelGGL(classify_kernel, dim3((n + kClassifyBlock - 1) / kClassifyBlock), dim3(kClassifyBlock), 0,
                       stream, a, nbins, edges, queues, counts, gqueue, gcount, n, done, routed);
}

void launch_copy_ints(const int* src, int* dst, int count, hipStream_t stream)
{
    hipLaunchKernelGGL(copy_ints_kernel, dim3(1), dim3(64), 0, stream, src, dst, count);
}

void launch_fp64_generic(const BuildArgs& a, const ScratchLayout& L, int mode, const int* queue,
                         const int* nqueue, int* head, int nwaves, hipStream_t stream)
{
    const size_t lds = kGenericHashSlots * sizeof(unsigned);  // the score hash; the working set is in HBM
    if (mode == 2)
        hipLaunchKernelGGL((spai_column_kernel<64, false, 2, 0>), dim3(nwaves), dim3(kWave), lds, stream, a, L, queue,
                           nqueue, head);
    else if (mode == 4)
        hipLaunchKernelGGL((spai_column_kernel<64, false, 4, 0>), dim3(nwaves), dim3(kWave), lds, stream, a, L, queue,
                           nqueue, head);
    else
        hipLaunchKernelGGL((spai_column_kernel<64, false, 0, 0>), dim3(nwaves), dim3(kWave), lds, stream, a, L, queue,
                           nqueue, head);
}

}  // namespace mspai
