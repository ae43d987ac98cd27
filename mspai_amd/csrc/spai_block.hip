/*
 * Block SPAI (-bs != 1) on gfx950: one wave per block column.
 *
 * Spai_Unrestrained_Block<T>::SPAI_Column (src/Spai_Unrestrained_Block.imp:
 * 33-205) on the block matrix of Convert_To_Block_Matrix
 * (src/Matrix.imp:1516-1745): J and I are sets of block indices; A(I,J) is
 * the dense scalar expansion of the stored blocks (Create_Submatrix_AHat_Block,
 * src/Spai_Sub.imp:285-401); the right-hand side is the identity block of
 * block row k, bk = bsize[k] columns (Create_ek_Hat_Block, 691-733); one QR
 * serves all bk right-hand sides; the residual's Frobenius norm is compared
 * with eps*sqrt(bk); candidates are the block columns of the block rows of
 * I u {k}, scored by rho_j^2 = ||R||^2 - trace(N^T (A_j^T A_j)^-1 N) with
 * N = A(I,j)^T R (Compute_Rhos_Block, src/Spai_Sub.imp:1351-1466;
 * Compute_Numerator_Block, src/Spai_Sub.cpp:294-370).
 *
 * Layout: the block matrix in HBM as block CSC (block e column-major at
 * val[voff[e]]), its block row pattern as CSR, (A_j^T A_j)^-1 per block
 * column, and the start pattern over block columns.  The working set of a
 * column (index sets, A(I,J) twice, the bk right-hand sides, residual) is a
 * per-wave slot in HBM sized for the largest column; the kernel is
 * persistent, the waves take block columns from an atomic counter.
 *
 * Order of operations: the 64-lane order for the QR (row r on lane r % 64,
 * fused partial sums, xor butterfly), dgemv's column order for the residual,
 * sequential fused sums for the small dgemm products of the numerator and
 * of (A_j^T A_j)^-1 N, and the reference's own unfused sequential loops for
 * the trace term, the Frobenius sum, the mean and A_j^T A_j.
 */
#include "device_common.h"
#include "spai_host.h"

namespace mspai {

namespace {

constexpr int kBlkOverflow = -2;
constexpr int kBlkRegJ = 32;  // register solve: scalar columns of A(I,J)
constexpr int kBlkRegB = 4;   //   and right-hand sides (|I| <= 64)

struct BlkScratch {
    int* J;
    int* I;
    int* sbuf;
    int* jt;
    int* fr;
    int* fc;
    int* sel;
    int* stack;
    RhoPair* pairs;
    double* rho;
    double* ahat;
    double* aqr;
    double* b;
    double* b0;
    double* r;
    double* rb;
    double* x;
    double* nbuf;
};

__device__ __forceinline__ BlkScratch blk_scratch(char* ws, const BlockLayout& L)
{
    BlkScratch s;
    s.J = (int*)(ws + L.J);
    s.I = (int*)(ws + L.I);
    s.sbuf = (int*)(ws + L.sbuf);
    s.jt = (int*)(ws + L.jt);
    s.fr = (int*)(ws + L.fr);
    s.fc = (int*)(ws + L.fc);
    s.sel = (int*)(ws + L.sel);
    s.stack = (int*)(ws + L.stack);
    s.pairs = (RhoPair*)(ws + L.pairs);
    s.rho = (double*)(ws + L.rho);
    s.ahat = (double*)(ws + L.ahat);
    s.aqr = (double*)(ws + L.aqr);
    s.b = (double*)(ws + L.b);
    s.b0 = (double*)(ws + L.b0);
    s.r = (double*)(ws + L.r);
    s.rb = (double*)(ws + L.rb);
    s.x = (double*)(ws + L.x);
    s.nbuf = (double*)(ws + L.nbuf);
    return s;
}

using W = Group<kWave>;

/* out[0..count) = concatenation of ind[ptr[set[t]] .. ptr[set[t]+1]) over
 * t < nset (lane t reads set t's list), then the sorted distinct values into
 * dst.  Returns their number, or kBlkOverflow when the raw lists exceed cap. */
__device__ __forceinline__ int sorted_union(const int* ptr, const int* ind, const int* set, int nset, int extra,
                                            int* buf, int cap, int* dst)
{
    const int lane = W::lane();
    const int nl = nset + (extra >= 0 ? 1 : 0);
    int total = 0;
    for (int t = lane; t < nl; t += kWave) {
        int l = t < nset ? set[t] : extra;
        total += ptr[l + 1] - ptr[l];
    }
    total = W::sum_int(total);
    if (total > cap) return kBlkOverflow;
    int off = 0;
    for (int base = 0; base < nl; base += kWave) {
        int t = base + lane;
        int len = 0, e0 = 0;
        if (t < nl) {
            int l = t < nset ? set[t] : extra;
            e0 = ptr[l];
            len = ptr[l + 1] - e0;
        }
        int incl = W::scan_incl(len);
        int excl = incl - len;
        for (int e = 0; e < len; ++e) buf[off + excl + e] = ind[e0 + e];
        off += W::bcast_last(incl);
    }
    __syncthreads();
    return group_sort_unique<kWave>(buf, off, dst);
}

/* first[t] = scalar offset of block set[t]; returns the scalar length. */
__device__ __forceinline__ int prefix_sizes(const int* bsize, const int* set, int nset, int* first)
{
    const int lane = W::lane();
    int off = 0;
    for (int base = 0; base < nset; base += kWave) {
        int t = base + lane;
        int v = t < nset ? bsize[set[t]] : 0;
        int incl = W::scan_incl(v);
        if (t < nset) first[t] = off + incl - v;
        off += W::bcast_last(incl);
    }
    if (lane == 0) first[nset] = off;
    __syncthreads();
    return off;
}

/* Augment_Sparsity_Block.  Returns the new |J|, -1 when J~ is empty, or
 * kBlkOverflow. */
__device__ int augment_block(const BlockArgs& a, const BlockLayout& L, const BlkScratch& s, int col, int bk,
                             int mi, int m, int nj, double rn, int* st)
{
    const int lane = W::lane();
    const int ld = L.capIs;
    /* U = union of the block row patterns over I u {col}; J~ = U \ J */
    int p = lower_bound_i(s.I, mi, col);
    const bool col_in = p < mi && s.I[p] == col;
    int nu = sorted_union(a.rp_ptr, a.rp_ind, s.I, mi, col_in ? -1 : col, s.sbuf, L.capSort, s.sbuf);
    if (nu == kBlkOverflow || nu > L.capU) return kBlkOverflow;
    int c = 0;
    for (int base = 0; base < nu; base += kWave) {
        int t = base + lane;
        int u = t < nu ? s.sbuf[t] : 0;
        bool keep = false;
        if (t < nu) {
            int q = lower_bound_i(s.J, nj, u);
            keep = !(q < nj && s.J[q] == u);
        }
        unsigned long long mask = W::ballot(keep);
        int q2 = c + __popcll(mask & W::lanes_below());
        if (keep) s.jt[q2] = u;
        c += __popcll(mask);
    }
    __syncthreads();
    if (c == 0) return -1;
    /* Convert_To_Block: R in block-row order, each block bsize x bk */
    for (int k = lane; k < mi; k += kWave) {
        const int h = a.bsize[s.I[k]], r0 = s.fr[k];
        double* dst = s.rb + (size_t)r0 * bk;
        for (int q = 0; q < bk; ++q)
            for (int i = 0; i < h; ++i) dst[q * h + i] = s.r[(size_t)q * ld + r0 + i];
    }
    __syncthreads();
    /* rho per candidate, one candidate per lane */
    const int mb = L.maxbs;
    double* N = s.nbuf + (size_t)lane * 2 * mb * mb;
    double* Bm = N + (size_t)mb * mb;
    for (int t = lane; t < c; t += kWave) {
        const int j = s.jt[t], bj = a.bsize[j];
        for (int e = 0; e < bj * bk; ++e) N[e] = 0.0;
        for (int e = a.colptr[j]; e < a.colptr[j + 1]; ++e) {
            const int ib = a.rowind[e];
            const int q = lower_bound_i(s.I, mi, ib);
            if (q >= mi || s.I[q] != ib) continue;
            const int h = a.bsize[ib];
            const double* av = a.val + a.voff[e];
            const double* rblk = s.rb + (size_t)s.fr[q] * bk;
            for (int qq = 0; qq < bk; ++qq)
                for (int ii = 0; ii < bj; ++ii) {
                    double d = 0.0;
                    for (int tt = 0; tt < h; ++tt) d = __fma_rn(av[ii * h + tt], rblk[qq * h + tt], d);
                    N[qq * bj + ii] = __dadd_rn(N[qq * bj + ii], d);
                }
        }
        const double* inv = a.sqinv + a.sqoff[j];
        for (int qq = 0; qq < bk; ++qq)
            for (int ii = 0; ii < bj; ++ii) {
                double d = 0.0;
                for (int kk = 0; kk < bj; ++kk) d = __fma_rn(inv[kk * bj + ii], N[qq * bj + kk], d);
                Bm[qq * bj + ii] = d;
            }
        double term = 0.0;
        for (int qq = 0; qq < bk; ++qq)
            for (int ii = 0; ii < bj; ++ii) term = __dadd_rn(term, __dmul_rn(Bm[qq * bj + ii], N[qq * bj + ii]));
        s.rho[t] = sqrt(__dsub_rn(__dmul_rn(rn, rn), term));
    }
    __syncthreads();
    double mean = 0.0;
    if (lane == 0) {
        for (int t = 0; t < c; ++t) mean = __dadd_rn(mean, s.rho[t]);
        if (a.use_mean) mean = __ddiv_rn(mean, (double)c);
    }
    mean = W::bcast0(mean);
    if (isnan(mean)) {
        *st |= COL_NAN_RHO;
        return nj;
    }
    /* the first min(maxnew, #{rho <= mean}) after std::sort are taken */
    int qn = 0;
    for (int t = lane; t < c; t += kWave) qn += (s.rho[t] <= mean);
    qn = W::sum_int(qn);
    const int sneed = a.maxnew < qn ? a.maxnew : qn;
    if (sneed <= 0) return nj;
    if (nj + sneed > L.capJ) return kBlkOverflow;
    /* the sneed-th smallest score and the number of scores below it */
    double prev = -1.0, vstar = 0.0;
    int below = 0, ties = 0;
    for (int it = 0; it < c; ++it) {
        double v = INFINITY;
        for (int t = lane; t < c; t += kWave)
            if (s.rho[t] > prev) v = fmin(v, s.rho[t]);
        v = W::min(v);
        int cnt = 0;
        for (int t = lane; t < c; t += kWave) cnt += (s.rho[t] == v);
        cnt = W::sum_int(cnt);
        if (below + cnt >= sneed || cnt == 0) {
            vstar = v;
            ties = cnt;
            break;
        }
        below += cnt;
        prev = v;
    }
    if (sneed - below == ties) {
        int ns = 0;
        for (int base = 0; base < c; base += kWave) {
            int t = base + lane;
            bool keep = t < c && s.rho[t] <= vstar;
            unsigned long long mask = W::ballot(keep);
            int q2 = ns + __popcll(mask & W::lanes_below());
            if (keep) s.sel[q2] = s.jt[t];
            ns += __popcll(mask);
        }
    } else {
        /* a tie straddles the cut: std::sort's order decides */
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
    if (lane == 0) {
        int i = nj - 1, t = sneed - 1, w = nj + sneed - 1;
        while (t >= 0) {
            if (i >= 0 && s.J[i] > s.sel[t])
                s.J[w--] = s.J[i--];
            else
                s.J[w--] = s.sel[t--];
        }
    }
    __syncthreads();
    (void)m;
    return nj + sneed;
}

/* Approximative_Solution (block) with row r of A(I,J) and of the bk
 * right-hand sides in the registers of lane r (|I| <= 64, |J| <= NMAX,
 * bk <= NB): the operations and their order are the 64-lane order's (one row
 * per lane is its only slot), with no memory round trip or barrier between
 * reflectors.  Slots NMAX.. hold the right-hand sides; a reflector is applied
 * to four slots at a time so that their reductions interleave.  On exit
 * x[q*n + j] holds the solution (Q^T E when R has a zero pivot, 0 when
 * |I| < |J|).  Returns status bits.  Not inlined: its own register
 * allocation holds the 36-slot rows without spilling the kernel's state. */
template <int NMAX, int NB>
__device__ __attribute__((noinline)) int block_solve_regs(const double* ahat, const double* bsrc, int ld, int m, int n, int bk, double* x)
{
    constexpr int kB = 4;
    constexpr int NS = NMAX + NB;
    static_assert(NMAX % kB == 0 && NB % kB == 0, "register width");
    const int lane = W::lane();
    const bool row = lane < m;
    double a[NS];
#pragma unroll
    for (int j = 0; j < NMAX; ++j) a[j] = (row && j < n) ? ahat[(size_t)j * ld + lane] : 0.0;
#pragma unroll
    for (int q = 0; q < NB; ++q) a[NMAX + q] = (row && q < bk) ? bsrc[(size_t)q * ld + lane] : 0.0;
    const int k = m < n ? m : n;
    bool zero_diag = false;  // R_ii == 0 on this lane's row
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
            const double alpha = W::lane_value(ai, i);
            const double xn2 = W::sum_fixed(lane > i ? __fma_rn(ai, ai, 0.0) : 0.0);
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
                for (int jb = ib; jb < NS; jb += kB) {
                    const bool rhs = jb >= NMAX;
                    if (rhs ? jb - NMAX < bk : (jb + kB - 1 > i && jb < n)) {
                        double p[kB];
#pragma unroll
                        for (int t = 0; t < kB; ++t) p[t] = act ? __fma_rn(vi, a[jb + t], 0.0) : 0.0;
                        W::sum_bcast_n(p);
#pragma unroll
                        for (int t = 0; t < kB; ++t) {
                            const int j = jb + t;
                            const bool live = rhs ? j - NMAX < bk : (j > i && j < n);
                            if (live && act) a[j] = __fma_rn(vi, mt * p[t], a[j]);
                        }
                    }
                }
            }
            zero_diag |= lane == i && rii == 0.0;
        }
    }
    int st = 0;
    if (m < n) {
        st = COL_UNDERDET;
#pragma unroll
        for (int q = 0; q < NB; ++q) a[NMAX + q] = 0.0;
    } else if (W::ballot(zero_diag) != 0ull) {
        st = COL_SINGULAR;  // dtrtrs leaves Q^T E
    } else {
        /* dtrsm order per right-hand side; K over compile-time slots so R_iK
         * is a fixed register of lane i */
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            if (q < bk) {
                double b = a[NMAX + q];
#pragma unroll
                for (int K = NMAX - 1; K >= 0; --K) {
                    if (K < n) {
                        double xk = W::lane_value(b, K);
                        const double rkk = W::lane_value(a[K], K);
                        if (xk != 0.0) {
                            xk = xk / rkk;
                            if (lane == K) b = xk;
                            if (lane < K) b = __fma_rn(-xk, a[K], b);
                        }
                    }
                }
                a[NMAX + q] = b;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NB; ++q)
        if (q < bk && lane < n) x[q * n + lane] = a[NMAX + q];
    __syncthreads();
    return st;
}

/* One block column.  Returns false when it outgrew the working set. */
__device__ bool block_column(const BlockArgs& a, const BlockLayout& L, const BlkScratch& s, int col)
{
    const int lane = W::lane();
    const int ld = L.capIs;
    const int bk = a.bsize[col];
    const double eps = a.epsilon * sqrt((double)bk);  // adjust_epsilon
    const int p0 = a.p_ptr[col];
    int nj = a.p_ptr[col + 1] - p0;
    if (nj > L.capJ) return false;
    for (int i = lane; i < nj; i += kWave) s.J[i] = a.p_ind[p0 + i];
    __syncthreads();
    int st = 0, solves = 0, n = 0;
    double rn = 1.0e50;
    for (int step = -1; step < a.ns; ++step) {
        const int mi = sorted_union(a.colptr, a.rowind, s.J, nj, -1, s.sbuf, L.capSort, s.I);
        if (mi == kBlkOverflow || mi > L.capI) return false;
        const int m = prefix_sizes(a.bsize, s.I, mi, s.fr);
        n = prefix_sizes(a.bsize, s.J, nj, s.fc);
        if (m > L.capIs || n > L.capJs) return false;
        /* A(I,J): zero fill, then every stored block at its offsets */
        for (int c = 0; c < n; ++c)
            for (int r = lane; r < m; r += kWave) s.ahat[(size_t)c * ld + r] = 0.0;
        __syncthreads();
        for (int t = 0; t < nj; ++t) {
            const int j = s.J[t], bj = a.bsize[j], c0 = s.fc[t];
            for (int e = a.colptr[j] + lane; e < a.colptr[j + 1]; e += kWave) {
                const int ib = a.rowind[e];
                const int pos = lower_bound_i(s.I, mi, ib);
                const int h = a.bsize[ib], r0 = s.fr[pos];
                const double* blk = a.val + a.voff[e];
                for (int c = 0; c < bj; ++c)
                    for (int rr = 0; rr < h; ++rr) s.ahat[(size_t)(c0 + c) * ld + r0 + rr] = blk[c * h + rr];
            }
        }
        __syncthreads();
        /* E_k: the identity block at block row col's offset */
        for (int q = 0; q < bk; ++q)
            for (int r = lane; r < m; r += kWave) s.b[(size_t)q * ld + r] = 0.0;
        __syncthreads();
        const int pk = lower_bound_i(s.I, mi, col);
        if (pk < mi && s.I[pk] == col) {
            if (lane < bk) s.b[(size_t)lane * ld + s.fr[pk] + lane] = 1.0;
        } else {
            st |= COL_K_NOT_IN_I;
        }
        __syncthreads();
        for (int q = 0; q < bk; ++q)
            for (int r = lane; r < m; r += kWave) s.b0[(size_t)q * ld + r] = s.b[(size_t)q * ld + r];
        __syncthreads();
        if (m <= kWave && n <= kBlkRegJ && bk <= kBlkRegB) {
            st |= block_solve_regs<kBlkRegJ, kBlkRegB>(s.ahat, s.b, ld, m, n, bk, s.x);
        } else {
        for (int c = 0; c < n; ++c)
            for (int r = lane; r < m; r += kWave) s.aqr[(size_t)c * ld + r] = s.ahat[(size_t)c * ld + r];
        __syncthreads();
        /* QR of A(I,J); Q^T on every right-hand side */
        const int kq = m < n ? m : n;
        for (int i = 0; i < kq; ++i) {
            double* ai = s.aqr + (size_t)i * ld;
            const double tau = make_reflector<kWave>(ai, i, m);
            if (tau != 0.0) {
                for (int j = i + 1; j < n; ++j) apply_reflector<kWave>(ai, s.aqr + (size_t)j * ld, i, m, tau);
                for (int q = 0; q < bk; ++q) apply_reflector<kWave>(ai, s.b + (size_t)q * ld, i, m, tau);
            }
        }
        __syncthreads();
        if (m < n) {
            st |= COL_UNDERDET;
            for (int e = lane; e < n * bk; e += kWave) s.x[e] = 0.0;
        } else {
            bool zero = false;
            for (int i = lane; i < n; i += kWave) zero |= (s.aqr[(size_t)i * ld + i] == 0.0);
            if (W::ballot(zero) != 0ull) {
                st |= COL_SINGULAR;  // dtrtrs leaves Q^T E
            } else {
                for (int q = 0; q < bk; ++q) {
                    double* bq = s.b + (size_t)q * ld;
                    for (int K = n - 1; K >= 0; --K) {
                        double xk = bq[K];
                        if (xk != 0.0) {
                            xk = xk / s.aqr[(size_t)K * ld + K];
                            const double nx = -xk;
                            __syncthreads();
                            if (lane == 0) bq[K] = xk;
                            for (int i = lane; i < K; i += kWave)
                                bq[i] = __fma_rn(nx, s.aqr[(size_t)K * ld + i], bq[i]);
                        }
                        __syncthreads();
                    }
                }
            }
            for (int q = 0; q < bk; ++q)
                for (int j = lane; j < n; j += kWave) s.x[q * n + j] = s.b[(size_t)q * ld + j];
        }
        }
        __syncthreads();
        ++solves;
        /* R = A(I,J) X - E per right-hand side, then the Frobenius norm */
        for (int i = lane; i < m; i += kWave)
            for (int q = 0; q < bk; ++q) {
                double r = -s.b0[(size_t)q * ld + i];
                for (int jj = 0; jj < n; ++jj) {
                    const double xj = s.x[q * n + jj];
                    if (xj != 0.0) r = __fma_rn(xj, s.ahat[(size_t)jj * ld + i], r);
                }
                s.r[(size_t)q * ld + i] = r;
            }
        __syncthreads();
        double sum = 0.0;
        if (lane == 0)
            for (int q = 0; q < bk; ++q)
                for (int i = 0; i < m; ++i) {
                    const double v = s.r[(size_t)q * ld + i];
                    sum = __dadd_rn(sum, __dmul_rn(v, v));
                }
        rn = sqrt(W::bcast0(sum));
        if (rn <= eps || step == a.ns - 1) break;
        const int nn = augment_block(a, L, s, col, bk, mi, m, nj, rn, &st);
        if (nn == kBlkOverflow) return false;
        if (nn < 0) break;
        nj = nn;
    }
    /* Get_Solution_Vals_Block: block t of J as a bsize[J[t]] x bk block */
    const size_t ov = (size_t)col * a.capX;
    for (int t = 0; t < nj; ++t) {
        const int bj = a.bsize[s.J[t]], f = s.fc[t];
        for (int e = lane; e < bj * bk; e += kWave) {
            const int q = e / bj, jj = e - q * bj;
            a.m_val[ov + (size_t)f * bk + e] = s.x[q * n + f + jj];
        }
    }
    for (int t = lane; t < nj; t += kWave) a.m_idx[(size_t)col * a.capJ + t] = s.J[t];
    if (lane == 0) {
        a.m_len[col] = nj;
        a.res[col] = rn;
        a.nsolve[col] = solves;
        a.status[col] = st;
    }
    __syncthreads();
    return true;
}

__global__ __launch_bounds__(64, 4) void block_column_kernel(BlockArgs a, BlockLayout L)
{
    BlkScratch s = blk_scratch(a.scratch + (size_t)blockIdx.x * L.per_wave, L);
    for (;;) {
        int c = 0;
        if (lane_id() == 0) c = atomicAdd(a.head, 1);
        const int col = __builtin_amdgcn_readlane(c, 0);
        if (col >= a.nb) break;
        if (!block_column(a, L, s, col) && lane_id() == 0) {
            a.m_len[col] = 0;  // the scalar conversion reads only what was written
            atomicOr(a.error, 1);
        }
    }
}

/* Compute_Square_Inverse for every block column: S = A_j^T A_j by
 * Mult_Blocks_TN over the stored blocks (unfused, storage order), then S^-1
 * by Gauss-Jordan elimination with partial pivoting (first largest |pivot|),
 * S overwritten in tmp. */
__global__ void block_sq_inverse_kernel(BlockArgs a, double* tmp)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nb) return;
    const int bj = a.bsize[j];
    double* S = tmp + a.sqoff[j];
    double* inv = const_cast<double*>(a.sqinv) + a.sqoff[j];
    for (int e = 0; e < bj * bj; ++e) {
        S[e] = 0.0;
        inv[e] = 0.0;
    }
    for (int i = 0; i < bj; ++i) inv[i * bj + i] = 1.0;
    for (int e = a.colptr[j]; e < a.colptr[j + 1]; ++e) {
        const int h = a.bsize[a.rowind[e]];
        const double* v = a.val + a.voff[e];
        for (int i1 = 0; i1 < bj; ++i1)
            for (int j1 = 0; j1 < bj; ++j1)
                for (int k1 = 0; k1 < h; ++k1)
                    S[i1 + j1 * bj] = __dadd_rn(S[i1 + j1 * bj], __dmul_rn(v[k1 + i1 * h], v[k1 + j1 * h]));
    }
    for (int c = 0; c < bj; ++c) {
        int p = c;
        for (int r = c + 1; r < bj; ++r)
            if (fabs(S[r + c * bj]) > fabs(S[p + c * bj])) p = r;
        if (p != c)
            for (int k = 0; k < bj; ++k) {
                double t = S[c + k * bj];
                S[c + k * bj] = S[p + k * bj];
                S[p + k * bj] = t;
                t = inv[c + k * bj];
                inv[c + k * bj] = inv[p + k * bj];
                inv[p + k * bj] = t;
            }
        const double piv = S[c + c * bj];
        for (int k = 0; k < bj; ++k) {
            S[c + k * bj] = __ddiv_rn(S[c + k * bj], piv);
            inv[c + k * bj] = __ddiv_rn(inv[c + k * bj], piv);
        }
        for (int r = 0; r < bj; ++r) {
            if (r == c) continue;
            const double f = S[r + c * bj];
            if (f == 0.0) continue;
            for (int k = 0; k < bj; ++k) {
                S[r + k * bj] = __fma_rn(-f, S[c + k * bj], S[r + k * bj]);
                inv[r + k * bj] = __fma_rn(-f, inv[c + k * bj], inv[r + k * bj]);
            }
        }
    }
}

/* Scalar_Matrix, pass 1: nonzeros of every scalar column of M. */
__global__ void block_scalar_count_kernel(BlockArgs a, int* counts)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= a.nb) return;
    const int bk = a.bsize[col], nj = a.m_len[col];
    const size_t ov = (size_t)col * a.capX;
    for (int q = 0; q < bk; ++q) {
        int cnt = 0;
        size_t o = ov;
        for (int t = 0; t < nj; ++t) {
            const int bj = a.bsize[a.m_idx[(size_t)col * a.capJ + t]];
            for (int jj = 0; jj < bj; ++jj) cnt += a.m_val[o + (size_t)q * bj + jj] != 0.0;
            o += (size_t)bj * bk;
        }
        counts[a.bstart[col] + q] = cnt;
    }
}

/* Scalar_Matrix, pass 2: rows ascending (blocks in J order, rows within a
 * block in order), exact zeros dropped. */
__global__ void block_scalar_fill_kernel(BlockArgs a, const int* colptr, int* rowind, double* val)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= a.nb) return;
    const int bk = a.bsize[col], nj = a.m_len[col];
    const size_t ov = (size_t)col * a.capX;
    for (int q = 0; q < bk; ++q) {
        int w = colptr[a.bstart[col] + q];
        size_t o = ov;
        for (int t = 0; t < nj; ++t) {
            const int j = a.m_idx[(size_t)col * a.capJ + t];
            const int bj = a.bsize[j];
            for (int jj = 0; jj < bj; ++jj) {
                const double v = a.m_val[o + (size_t)q * bj + jj];
                if (v != 0.0) {
                    rowind[w] = a.bstart[j] + jj;
                    val[w] = v;
                    ++w;
                }
            }
            o += (size_t)bj * bk;
        }
    }
}

}  // namespace

void launch_block_sq_inverse(const BlockArgs& a, double* tmp, hipStream_t stream)
{
    if (a.nb <= 0) return;
    hipLaunchKernelGGL(block_sq_inverse_kernel, dim3((a.nb + 63) / 64), dim3(64), 0, stream, a, tmp);
}

void launch_block_columns(const BlockArgs& a, const BlockLayout& L, int nwaves, hipStream_t stream)
{
    if (a.nb <= 0 || nwaves <= 0) return;
    hipLaunchKernelGGL(block_column_kernel, dim3(nwaves), dim3(kWave), 0, stream, a, L);
}

void launch_block_scalar_count(const BlockArgs& a, int* counts, hipStream_t stream)
{
    if (a.nb <= 0) return;
    hipLaunchKernelGGL(block_scalar_count_kernel, dim3((a.nb + 127) / 128), dim3(128), 0, stream, a, counts);
}

void launch_block_scalar_fill(const BlockArgs& a, const int* colptr, int* rowind, double* val,
                              hipStream_t stream)
{
    if (a.nb <= 0) return;
    hipLaunchKernelGGL(block_scalar_fill_kernel, dim3((a.nb + 127) / 128), dim3(128), 0, stream, a, colptr,
                       rowind, val);
}

}  // namespace mspai
