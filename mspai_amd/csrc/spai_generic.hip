/*
 * Generic FP64 column kernel: one wave64 per column of M, any |I| x |J|,
 * working sets in a per-wave region of HBM.
 *
 * This is the reference-exact path for columns that do not fit the
 * register-resident size bins (spai_binned.hip).  One wave runs
 * Spai_Unrestrained<T>::SPAI_Column (src/Spai_Unrestrained.imp:33-133) for
 * one column: I-set, gather of A(I,J), Householder QR with Q^T b carried as an
 * extra column, back substitution, FP64 residual, and the augmentation step
 * (candidate union, rho scores, mean threshold, top-maxnew selection).
 * Floating-point operations follow the fixed 64-lane order (rows r on lane
 * r % 64, partials combined by the xor butterfly), which makes the result
 * bitwise reproducible; arithmetic the reference compiles itself (Sqrt_Sum,
 * Compute_Numerator, the rho formula and the mean) is sequential and unfused,
 * as there.
 *
 * Columns come from a persistent work queue (one atomic per column), the
 * device form of Load_Balance::Grab_M_Col (src/Load_Balance.imp:34-69).
 */
#include "device_common.h"

namespace mspai {

namespace {

struct WaveScratch {
    int* J;
    int* I;
    int* sbuf;
    int* ubuf;
    double* ahat;
    double* aqr;
    double* b;
    double* b0;
    double* r;
    double* tau;
    int* jt;
    double* rho;
    RhoPair* pairs;
    int* sel;
};

__device__ WaveScratch wave_scratch(const BuildArgs& a, int slot)
{
    char* ws = a.scratch + (size_t)slot * a.L.per_wave;
    WaveScratch s;
    s.J = (int*)(ws + a.L.J);
    s.I = (int*)(ws + a.L.I);
    s.sbuf = (int*)(ws + a.L.sbuf);
    s.ubuf = (int*)(ws + a.L.ubuf);
    s.ahat = (double*)(ws + a.L.ahat);
    s.aqr = (double*)(ws + a.L.aqr);
    s.b = (double*)(ws + a.L.b);
    s.b0 = (double*)(ws + a.L.b0);
    s.r = (double*)(ws + a.L.r);
    s.tau = (double*)(ws + a.L.tau);
    s.jt = (int*)(ws + a.L.jt);
    s.rho = (double*)(ws + a.L.rho);
    s.pairs = (RhoPair*)(ws + a.L.pairs);
    s.sel = (int*)(ws + a.L.sel);
    return s;
}

/* Get_I_Set (src/Spai_Sub.imp:65-122): sorted distinct rows of the columns in J. */
__device__ int build_I(const BuildArgs& a, const WaveScratch& s, int n)
{
    const int lane = lane_id();
    int off = 0;
    for (int jj = 0; jj < n; ++jj) {
        int j = s.J[jj];
        int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
        for (int e = lane; e < cnt; e += kWave) s.sbuf[off + e] = a.A.rowind[e0 + e];
        off += cnt;
    }
    __syncthreads();
    return wave_sort_unique(s.sbuf, off, s.I);
}

/* Create_Submatrix_AHat (src/Spai_Sub.imp:221-282): A(I,J) column-major,
 * leading dimension capI; also the working copy for the QR. */
__device__ void gather_ahat(const BuildArgs& a, const WaveScratch& s, int m, int n)
{
    const int lane = lane_id();
    const int ld = a.L.capI;
    for (int jj = 0; jj < n; ++jj)
        for (int i = lane; i < m; i += kWave) s.ahat[(size_t)jj * ld + i] = 0.0;
    __syncthreads();
    for (int jj = 0; jj < n; ++jj) {
        int j = s.J[jj];
        int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
        for (int e = lane; e < cnt; e += kWave) {
            int pos = lower_bound_i(s.I, m, a.A.rowind[e0 + e]);
            s.ahat[(size_t)jj * ld + pos] = a.A.val[e0 + e];
        }
    }
    __syncthreads();
    for (int jj = 0; jj < n; ++jj)
        for (int i = lane; i < m; i += kWave) s.aqr[(size_t)jj * ld + i] = s.ahat[(size_t)jj * ld + i];
    __syncthreads();
}

/* Create_ek_Hat / Create_bk_Hat (src/Spai_Sub.imp:629-658, 735-775).
 * Returns COL_K_NOT_IN_I when no entry of the target column lies in I. */
__device__ int make_rhs(const BuildArgs& a, const WaveScratch& s, int m, int col)
{
    const int lane = lane_id();
    for (int i = lane; i < m; i += kWave) s.b[i] = 0.0;
    __syncthreads();
    int hit = 0;
    if (!a.B.colptr) {
        int pos = lower_bound_i(s.I, m, col);
        hit = pos < m && s.I[pos] == col;
        if (hit && lane == 0) s.b[pos] = 1.0;
    } else {
        int e0 = a.B.colptr[col], cnt = a.B.colptr[col + 1] - e0;
        for (int e = lane; e < cnt; e += kWave) {
            int row = a.B.rowind[e0 + e];
            int pos = lower_bound_i(s.I, m, row);
            if (pos < m && s.I[pos] == row) {
                s.b[pos] = a.B.val[e0 + e];
                hit = 1;
            }
        }
        hit = __ballot(hit) != 0ull;
    }
    __syncthreads();
    for (int i = lane; i < m; i += kWave) s.b0[i] = s.b[i];
    __syncthreads();
    return hit ? 0 : COL_K_NOT_IN_I;
}

/* c <- (I - tau v v^T) c over rows i..m-1, v_i = 1 (dlarf('L')). */
__device__ void apply_reflector(const double* v, double* c, int i, int m, double tau)
{
    const int lane = lane_id();
    double p = 0.0;
    for (int r = lane; r < m; r += kWave) {
        if (r < i) continue;
        double vr = (r == i) ? 1.0 : v[r];
        p = __fma_rn(vr, c[r], p);
    }
    double w = wave_sum_fixed(p);
    double t = (-tau) * w;
    for (int r = lane; r < m; r += kWave) {
        if (r < i) continue;
        double vr = (r == i) ? 1.0 : v[r];
        c[r] = __fma_rn(vr, t, c[r]);
    }
    __syncthreads();
}

/* Approximative_Solution (src/Spai_Unrestrained.imp:135-195): QR, Q^T b,
 * dtrtrs.  On exit b[0..n) holds the solution.  Returns status bits. */
__device__ int ls_solve(const BuildArgs& a, const WaveScratch& s, int m, int n)
{
    const int lane = lane_id();
    const int ld = a.L.capI;
    int k = m < n ? m : n;
    for (int i = 0; i < k; ++i) {
        double* ai = s.aqr + (size_t)i * ld;
        double tau = 0.0;
        if (i + 1 < m) {
            double alpha = ai[i];
            double p = 0.0;
            for (int r = lane; r < m; r += kWave)
                if (r > i) p = __fma_rn(ai[r], ai[r], p);
            double xn2 = wave_sum_fixed(p);
            if (xn2 != 0.0) {
                double nrm = sqrt(__fma_rn(alpha, alpha, xn2));
                double beta = -copysign(nrm, alpha);
                tau = (beta - alpha) / beta;
                double scal = 1.0 / (alpha - beta);
                __syncthreads();
                for (int r = lane; r < m; r += kWave)
                    if (r > i) ai[r] = ai[r] * scal;
                if (lane == 0) ai[i] = beta;
                __syncthreads();
            }
        }
        if (lane == 0) s.tau[i] = tau;
        if (tau != 0.0) {
            for (int j = i + 1; j < n; ++j) apply_reflector(ai, s.aqr + (size_t)j * ld, i, m, tau);
            apply_reflector(ai, s.b, i, m, tau);
        }
    }
    __syncthreads();
    if (m < n) {
        for (int i = lane; i < n; i += kWave) s.b[i] = 0.0;
        __syncthreads();
        return COL_UNDERDET;
    }
    bool zero = false;
    for (int i = lane; i < n; i += kWave) zero |= (s.aqr[(size_t)i * ld + i] == 0.0);
    if (__ballot(zero) != 0ull) return COL_SINGULAR;
    for (int K = n - 1; K >= 0; --K) {
        double xk = s.b[K];
        if (xk != 0.0) {
            xk = xk / s.aqr[(size_t)K * ld + K];
            double nx = -xk;
            __syncthreads();
            if (lane == 0) s.b[K] = xk;
            for (int i = lane; i < K; i += kWave) s.b[i] = __fma_rn(nx, s.aqr[(size_t)K * ld + i], s.b[i]);
        }
        __syncthreads();
    }
    return 0;
}

/* Residual_Norm (src/Spai_Sub.imp:915-941): r = A_hat x - b by dgemv's
 * column order with beta = -1, then sqrt of the sequential sum of squares. */
__device__ double residual_norm(const BuildArgs& a, const WaveScratch& s, int m, int n)
{
    const int lane = lane_id();
    const int ld = a.L.capI;
    for (int i = lane; i < m; i += kWave) {
        double ri = -s.b0[i];
        for (int j = 0; j < n; ++j) {
            double xj = s.b[j];
            if (xj != 0.0) ri = __fma_rn(xj, s.ahat[(size_t)j * ld + i], ri);
        }
        s.r[i] = ri;
    }
    __syncthreads();
    double sum = 0.0;
    if (lane == 0) {
        for (int i = 0; i < m; ++i) {
            double v = s.r[i];
            sum = __dadd_rn(sum, __dmul_rn(v, v));
        }
    }
    sum = __shfl(sum, 0, kWave);
    return sqrt(sum);
}

/* Augment_Sparsity (src/Spai_Sub.imp:1010-1116) with Union_Nl_Sets and
 * Compute_Rhos.  Returns the new |J|, or -1 when J~ is empty. */
__device__ int augment(const BuildArgs& a, const WaveScratch& s, int col, int m, int n, double rn,
                       int* st, double* wk)
{
    const int lane = lane_id();
    /* L = I u {col}: append col when absent (order does not matter for the union) */
    int pos = lower_bound_i(s.I, m, col);
    bool col_in = pos < m && s.I[pos] == col;
    int nl = col_in ? m : m + 1;
    /* gather the row patterns of L */
    double acc3 = 0.0, acc4 = 0.0;  // work counters: exact integers, any order
    int off = 0;
    for (int base = 0; base < nl; base += kWave) {
        int li = base + lane;
        int l = li < m ? s.I[li] : col;
        int len = 0, e0 = 0;
        if (li < nl) {
            e0 = a.rp_ptr[l];
            len = a.rp_ptr[l + 1] - e0;
        }
        /* exclusive prefix of len over the wave */
        int incl = len;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            int v = __shfl_up(incl, d, kWave);
            if (lane >= d) incl += v;
        }
        int excl = incl - len;
        for (int e = 0; e < len; ++e) s.sbuf[off + excl + e] = a.rp_ind[e0 + e];
        if (li < nl) acc3 += 8.0 + 4.0 * len;
        off += __shfl(incl, kWave - 1, kWave);
    }
    __syncthreads();
    int nu = wave_sort_unique(s.sbuf, off, s.ubuf);
    const int* U = s.ubuf;
    /* J~ = U \ J */
    int c = 0;
    for (int base = 0; base < nu; base += kWave) {
        int t = base + lane;
        int u = t < nu ? U[t] : 0;
        bool keep = false;
        if (t < nu) {
            int p = lower_bound_i(s.J, n, u);
            keep = !(p < n && s.J[p] == u);
        }
        unsigned long long mask = __ballot(keep);
        int q = c + __popcll(mask & ((1ull << lane) - 1ull));
        if (keep) s.jt[q] = u;
        c += __popcll(mask);
    }
    __syncthreads();
    if (c == 0) {
        acc3 = wave_sum_fixed(acc3);
        if (wk && lane == 0) wk[3] += acc3;
        return -1;
    }
    /* rho_j = sqrt(||r||^2 - (r^T a_j)^2 / ||a_j||^2), numerator over I only */
    for (int t = lane; t < c; t += kWave) {
        int j = s.jt[t];
        double num = 0.0;
        int e0 = a.A.colptr[j], e1 = a.A.colptr[j + 1];
        for (int e = e0; e < e1; ++e) {
            int row = a.A.rowind[e];
            int p = lower_bound_i(s.I, m, row);
            if (p < m && s.I[p] == row) num = __dadd_rn(num, __dmul_rn(s.r[p], a.A.val[e]));
        }
        num = __dmul_rn(num, num);
        double ajn = a.ajn[j];
        double rho = __dsub_rn(__dmul_rn(rn, rn), __ddiv_rn(num, __dmul_rn(ajn, ajn)));
        s.rho[t] = sqrt(rho);
        acc4 += 16.0 + 12.0 * (e1 - e0);
    }
    __syncthreads();
    acc3 = wave_sum_fixed(acc3);
    acc4 = wave_sum_fixed(acc4);
    double mean = 0.0;
    if (lane == 0) {
        for (int t = 0; t < c; ++t) mean = __dadd_rn(mean, s.rho[t]);
        if (a.use_mean) mean = __ddiv_rn(mean, (double)c);
        if (wk) {
            wk[3] += acc3;
            wk[4] += acc4;
            wk[5] += (double)c;
        }
    }
    mean = __shfl(mean, 0, kWave);
    if (isnan(mean)) {
        *st |= COL_NAN_RHO;
        return n;
    }
    /* q = #{rho <= mean}; the first min(maxnew, q) after std::sort are taken */
    int q = 0;
    for (int t = lane; t < c; t += kWave) q += (s.rho[t] <= mean);
    q = wave_sum_int(q);
    int sneed = a.maxnew < q ? a.maxnew : q;
    if (sneed <= 0) return n;
    /* the sneed-th smallest score v* and the number of scores below it */
    double prev = -1.0, vstar = 0.0;
    int below = 0, ties = 0;
    for (int it = 0; it < c; ++it) {
        double v = INFINITY;
        for (int t = lane; t < c; t += kWave)
            if (s.rho[t] > prev) v = fmin(v, s.rho[t]);
        v = wave_min(v);
        int cnt = 0;
        for (int t = lane; t < c; t += kWave) cnt += (s.rho[t] == v);
        cnt = wave_sum_int(cnt);
        if (below + cnt >= sneed || cnt == 0) {
            vstar = v;
            ties = cnt;
            break;
        }
        below += cnt;
        prev = v;
    }
    int take_ties = sneed - below;
    if (take_ties == ties) {
        int ns = 0;
        for (int base = 0; base < c; base += kWave) {
            int t = base + lane;
            bool keep = t < c && s.rho[t] <= vstar;
            unsigned long long mask = __ballot(keep);
            int q2 = ns + __popcll(mask & ((1ull << lane) - 1ull));
            if (keep) s.sel[q2] = s.jt[t];
            ns += __popcll(mask);
        }
    } else {
        /* a tie straddles the cut: reproduce std::sort's order */
        if (lane == 0) {
            for (int t = 0; t < c; ++t) {
                s.pairs[t].idx = s.jt[t];
                s.pairs[t].rho = s.rho[t];
            }
            std_sort_rho(s.pairs, s.pairs + c);
            for (int t = 0; t < sneed; ++t) s.sel[t] = s.pairs[t].idx;
            /* sort the selected indices */
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

__device__ void count_solve(const BuildArgs& a, const WaveScratch& s, int m, int n, double* wk)
{
    if (!wk || lane_id() != 0) return;
    double dm = (double)m, dn = (double)n;
    double kk = dm < dn ? dm : dn;
    wk[0] = __dadd_rn(wk[0], __dsub_rn(2.0 * dm * dn * dn, (2.0 / 3.0) * kk * kk * kk));
    wk[1] = __dadd_rn(wk[1], (4.0 * dm * dn - 2.0 * dn * dn) + dn * dn + 2.0 * dm * dn);
    double g = 0.0;
    for (int jj = 0; jj < n; ++jj) {
        int j = s.J[jj];
        g += 8.0 + 12.0 * (a.A.colptr[j + 1] - a.A.colptr[j]);
    }
    wk[2] += g;
    wk[6] += dm;
    wk[7] += dn;
}

__device__ void solve_column(const BuildArgs& a, const WaveScratch& s, int col)
{
    const int lane = lane_id();
    double* wk = a.work ? a.work + (size_t)col * kNumWork : nullptr;
    if (wk && lane < kNumWork) wk[lane] = 0.0;
    int p0 = a.p_ptr[col], n = a.p_ptr[col + 1] - p0;
    for (int i = lane; i < n; i += kWave) s.J[i] = a.p_ind[p0 + i];
    __syncthreads();
    int st = 0, solves = 0;
    double rn = 1.0e50;
    for (int step = -1; step < a.ns; ++step) {
        int m = build_I(a, s, n);
        gather_ahat(a, s, m, n);
        st |= make_rhs(a, s, m, col);
        st |= ls_solve(a, s, m, n);
        ++solves;
        count_solve(a, s, m, n, wk);
        rn = residual_norm(a, s, m, n);
        if (rn <= a.epsilon || step == a.ns - 1) break;
        int nn = augment(a, s, col, m, n, rn, &st, wk);
        if (nn < 0) break;
        n = nn;
    }
    size_t o = (size_t)col * a.cap;
    for (int i = lane; i < n; i += kWave) {
        a.m_idx[o + i] = s.J[i];
        a.m_val[o + i] = s.b[i];
    }
    if (lane == 0) {
        a.m_len[col] = n;
        a.res[col] = rn;
        a.nsolve[col] = solves;
        a.status[col] = st;
    }
    __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(64) void spai_generic_fp64_kernel(BuildArgs a)
{
    WaveScratch s = wave_scratch(a, blockIdx.x);
    while (true) {
        int q = 0;
        if (lane_id() == 0) q = atomicAdd(a.counter, 1);
        q = __shfl(q, 0, kWave);
        if (q >= a.nqueue) break;
        solve_column(a, s, a.queue[q]);
    }
}

void launch_generic_fp64(const BuildArgs& a, int nwaves, hipStream_t stream)
{
    hipLaunchKernelGGL(spai_generic_fp64_kernel, dim3(nwaves), dim3(kWave), 0, stream, a);
}

}  // namespace mspai
