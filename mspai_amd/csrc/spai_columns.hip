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
    float* aq32;  // mixed path: FP32 copy of A(I,J) (aliases aqr)
    float* bq32;  // mixed path: FP32 right-hand side
    double* g;    // mixed path: A(I,J)^T r
    double* y;    // mixed path: triangular-solve work vector
};

__device__ __forceinline__ WaveScratch wave_scratch(char* ws, const ScratchLayout& L)
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
    s.pairs = (RhoPair*)(ws + L.pairs);
    s.sel = (int*)(ws + L.sel);
    s.stack = (int*)(ws + L.stack);
    s.aq32 = (float*)(ws + L.aqr);
    s.bq32 = (float*)(ws + L.bq32);
    s.g = (double*)(ws + L.g);
    s.y = (double*)(ws + L.y);
    return s;
}

constexpr int kOverflow = -2;
constexpr int kFallback = -3;

/* Mixed path guards.  A selection decision whose margin is below kGuard (on
 * the scale of ||r||^2 for the scores, of epsilon for the stop test) is not
 * trusted to the refined FP32 solution: the column is rebuilt by the FP64
 * kernels, which reproduce the oracle bit for bit.  kCondGuard bounds
 * min |R_ii| / max |R_ii| of the FP32 factor (below it, one or two
 * refinement steps are not certain to reach FP64 accuracy). */
constexpr double kGuard = 1e-10;
constexpr float kCondGuard = 1e-4f;
constexpr int kRefineSteps = 3;

/* Get_I_Set (src/Spai_Sub.imp:65-122): sorted distinct rows of the columns in
 * J.  Returns |I|, or kOverflow when the raw list or I exceeds the layout. */
template <int G>
__device__ __forceinline__ int build_I(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s, int n)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    int off = 0;
    for (int jj = 0; jj < n; ++jj) {
        int j = s.J[jj];
        off += a.A.colptr[j + 1] - a.A.colptr[j];
    }
    if (off > L.capSort) return kOverflow;
    off = 0;
    for (int jj = 0; jj < n; ++jj) {
        int j = s.J[jj];
        int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
        for (int e = lane; e < cnt; e += G) s.sbuf[off + e] = a.A.rowind[e0 + e];
        off += cnt;
    }
    __syncthreads();
    int m = group_sort_unique<G>(s.sbuf, off, s.sbuf);
    if (m > L.capI) return kOverflow;
    for (int i = lane; i < m; i += G) s.I[i] = s.sbuf[i];
    __syncthreads();
    return m;
}

/* Create_Submatrix_AHat (src/Spai_Sub.imp:221-282): A(I,J) column-major with
 * leading dimension capI, straight into the QR working array. */
template <int G, bool kMixed>
__device__ __forceinline__ void gather_ahat(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                            int m, int n)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    const int ld = L.capI;
    for (int jj = 0; jj < n; ++jj)
        for (int i = lane; i < m; i += G) {
            if (kMixed)
                s.aq32[jj * ld + i] = 0.f;
            else
                s.aqr[jj * ld + i] = 0.0;
        }
    __syncthreads();
    for (int jj = 0; jj < n; ++jj) {
        int j = s.J[jj];
        int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
        for (int e = lane; e < cnt; e += G) {
            int pos = lower_bound_i(s.I, m, a.A.rowind[e0 + e]);
            if (kMixed)
                s.aq32[jj * ld + pos] = (float)a.A.val[e0 + e];
            else
                s.aqr[jj * ld + pos] = a.A.val[e0 + e];
        }
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

/* Approximative_Solution (src/Spai_Unrestrained.imp:135-195): QR, Q^T b,
 * dtrtrs.  On exit b[0..n) holds the solution.  Returns status bits. */
template <int G>
__device__ __forceinline__ int ls_solve(const ScratchLayout& L, const WaveScratch& s, int m, int n)
{
    using Gr = Group<G>;
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
            for (int j = i + 1; j < n; ++j) apply_reflector<G>(ai, s.aqr + j * ld, i, m, tau);
            apply_reflector<G>(ai, s.b, i, m, tau);
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
template <int G>
__device__ __forceinline__ int ls_solve_mixed(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                              int m, int n)
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
    /* conditioning of the FP32 factor (also rejects zero and NaN pivots) */
    float dmax = 0.f, dmin = INFINITY;
    for (int i = lane; i < n; i += G) {
        float d = fabsf(q[i * ld + i]);
        dmax = fmaxf(dmax, d);
        dmin = fminf(dmin, d);
    }
    for (int o = G / 2; o >= 1; o >>= 1) {
        dmax = fmaxf(dmax, __shfl_xor(dmax, o, G));
        dmin = fminf(dmin, __shfl_xor(dmin, o, G));
    }
    if (!(dmin > kCondGuard * dmax)) return kFallback;
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
    for (int it = 0; it < kRefineSteps; ++it) {
        /* r = b0 - A x over I (FP64) */
        for (int i = lane; i < m; i += G) s.r[i] = s.b0[i];
        __syncthreads();
        for (int jj = 0; jj < n; ++jj) {
            double xj = s.b[jj];
            int j = s.J[jj];
            int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
            for (int e = lane; e < cnt; e += G) {
                int pos = lower_bound_i(s.I, m, a.A.rowind[e0 + e]);
                s.r[pos] = __fma_rn(-xj, a.A.val[e0 + e], s.r[pos]);
            }
            __syncthreads();
        }
        /* g = A^T r */
        for (int jj = 0; jj < n; ++jj) {
            int j = s.J[jj];
            int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
            double p = 0.0;
            for (int e = lane; e < cnt; e += G) {
                int pos = lower_bound_i(s.I, m, a.A.rowind[e0 + e]);
                p = __fma_rn(a.A.val[e0 + e], s.r[pos], p);
            }
            p = Gr::sum_fixed(p);
            if (lane == 0) s.g[jj] = p;
        }
        __syncthreads();
        /* R^T y = g (forward), R d = y (backward); R from the FP32 factor */
        for (int K = 0; K < n; ++K) {
            double yk = s.g[K] / (double)q[K * ld + K];
            __syncthreads();
            if (lane == 0) s.g[K] = yk;
            for (int i = K + 1 + lane; i < n; i += G) s.g[i] = __fma_rn(-yk, (double)q[i * ld + K], s.g[i]);
            __syncthreads();
        }
        for (int K = n - 1; K >= 0; --K) {
            double dk = s.g[K] / (double)q[K * ld + K];
            __syncthreads();
            if (lane == 0) s.g[K] = dk;
            for (int i = lane; i < K; i += G) s.g[i] = __fma_rn(-dk, (double)q[K * ld + i], s.g[i]);
            __syncthreads();
        }
        double dmx = 0.0, xmx = 0.0;
        for (int i = lane; i < n; i += G) {
            double x = s.b[i] + s.g[i];
            s.b[i] = x;
            dmx = fmax(dmx, fabs(s.g[i]));
            xmx = fmax(xmx, fabs(x));
        }
        for (int o = G / 2; o >= 1; o >>= 1) {
            dmx = fmax(dmx, __shfl_xor(dmx, o, G));
            xmx = fmax(xmx, __shfl_xor(xmx, o, G));
        }
        __syncthreads();
        /* each step shrinks the error by about kappa^2 u_32; a correction
         * below 1e-10 |x| leaves the next one below FP64 rounding */
        if (dmx <= 1e-10 * xmx) return 0;
    }
    return kFallback;  // not converged in kRefineSteps: let FP64 decide
}

/* Residual_Norm (src/Spai_Sub.imp:915-941): r = A(I,J) x - b in dgemv's column
 * order with beta = -1, formed from the sparse columns (a stored entry of
 * column j updates its row in order of j, as the dense product does; the
 * dense product's zero terms leave every nonzero residual unchanged), then
 * sqrt of the sequential sum of squares. */
template <int G>
__device__ __forceinline__ double residual_norm(const BuildArgs& a, const WaveScratch& s, int m, int n)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    for (int i = lane; i < m; i += G) s.r[i] = -s.b0[i];
    __syncthreads();
    for (int jj = 0; jj < n; ++jj) {
        double xj = s.b[jj];
        if (xj == 0.0) continue;
        int j = s.J[jj];
        int e0 = a.A.colptr[j], cnt = a.A.colptr[j + 1] - e0;
        for (int e = lane; e < cnt; e += G) {
            int pos = lower_bound_i(s.I, m, a.A.rowind[e0 + e]);
            s.r[pos] = __fma_rn(xj, a.A.val[e0 + e], s.r[pos]);
        }
        __syncthreads();
    }
    double sum = 0.0;
    if (lane == 0) {
        for (int i = 0; i < m; ++i) {
            double v = s.r[i];
            sum = __dadd_rn(sum, __dmul_rn(v, v));
        }
    }
    sum = Gr::bcast0(sum);
    return sqrt(sum);
}

/* Augment_Sparsity (src/Spai_Sub.imp:1010-1116) with Union_Nl_Sets and
 * Compute_Rhos.  Returns the new |J|, -1 when J~ is empty, or kOverflow. */
template <int G, bool kMixed>
__device__ __forceinline__ int augment(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s, int col,
                                       int m, int n, double rn, int* st, double* wk)
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
    /* gather the row patterns of L */
    int off = 0;
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
        for (int e = 0; e < len; ++e) s.sbuf[off + excl + e] = a.rp_ind[e0 + e];
        off += __shfl(incl, G - 1, G);
    }
    __syncthreads();
    int nu = group_sort_unique<G>(s.sbuf, off, s.sbuf);
    if (nu > L.capU) return kOverflow;
    /* J~ = U \ J */
    int c = 0;
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
    if (c == 0) {
        if (wk && lane == 0) wk[3] += acc3;
        return -1;
    }
    /* rho_j = sqrt(||r||^2 - (r^T a_j)^2 / ||a_j||^2), numerator over I only */
    for (int t = lane; t < c; t += G) {
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
    acc4 = Gr::sum_fixed(acc4);
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
    mean = Gr::bcast0(mean);
    if (isnan(mean)) {
        if (kMixed) return kFallback;
        *st |= COL_NAN_RHO;
        return n;
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
        if (c > L.capPairs) return kOverflow;
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

template <int G>
__device__ __forceinline__ void count_solve(const BuildArgs& a, const WaveScratch& s, int m, int n, double* wk)
{
    if (!wk || Group<G>::lane() != 0) return;
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

/* Outcome of one column in one launch. */
enum ColumnOutcome { kDone = 0, kMoveOn = 1, kRedoFp64 = 2 };

/* One column (Spai_Unrestrained<T>::SPAI_Column).  kMoveOn: it does not fit
 * this launch's layout; kRedoFp64: the mixed path declined a decision. */
template <int G, bool kMixed>
__device__ __forceinline__ int solve_column(const BuildArgs& a, const ScratchLayout& L, const WaveScratch& s,
                                            int col)
{
    using Gr = Group<G>;
    const int lane = Gr::lane();
    double* wk = a.work ? a.work + (size_t)col * kNumWork : nullptr;
    int p0 = a.p_ptr[col], n = a.p_ptr[col + 1] - p0;
    if (n > L.capJ) return kMoveOn;
    if (wk && lane < kNumWork) wk[lane] = 0.0;
    for (int i = lane; i < n; i += G) s.J[i] = a.p_ind[p0 + i];
    __syncthreads();
    int st = 0, solves = 0;
    double rn = 1.0e50;
    for (int step = -1; step < a.ns; ++step) {
        int m = build_I<G>(a, L, s, n);
        if (m == kOverflow) return kMoveOn;
        gather_ahat<G, kMixed>(a, L, s, m, n);
        st |= make_rhs<G>(a, s, m, col);
        if (kMixed) {
            if (ls_solve_mixed<G>(a, L, s, m, n) == kFallback) return kRedoFp64;
        } else {
            st |= ls_solve<G>(L, s, m, n);
        }
        ++solves;
        count_solve<G>(a, s, m, n, wk);
        rn = residual_norm<G>(a, s, m, n);
        if (kMixed && !(fabs(rn - a.epsilon) > kGuard * a.epsilon)) return kRedoFp64;
        if (rn <= a.epsilon || step == a.ns - 1) break;
        int nn = augment<G, kMixed>(a, L, s, col, m, n, rn, &st, wk);
        if (nn == kOverflow) return kMoveOn;
        if (nn == kFallback) return kRedoFp64;
        if (nn < 0) break;
        n = nn;
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
    base = __shfl(base, 0, kWave);
    if (flag) out[base + __popcll(mask & ((1ull << lane_id()) - 1ull))] = col;
}

template <int G, bool kLds, bool kMixed>
__global__ __launch_bounds__(64) void spai_column_kernel(BuildArgs a, ScratchLayout L, const int* queue,
                                                         const int* nqueue, int* head)
{
    extern __shared__ __align__(16) char smem[];
    constexpr int NG = kWave / G;
    /* columns taken per queue atomic: one per group, several rounds */
    constexpr int kRounds = G == 16 ? 8 : (G == 32 ? 4 : 2);
    const int g = Group<G>::index();
    char* base = kLds ? smem + (size_t)g * L.per_wave
                      : a.scratch + ((size_t)blockIdx.x * NG + g) * L.per_wave;
    WaveScratch s = wave_scratch(base, L);
    const int nq = *nqueue;
    while (true) {
        int q0 = 0;
        if (lane_id() == 0) q0 = atomicAdd(head, NG * kRounds);
        q0 = __shfl(q0, 0, kWave);
        if (q0 >= nq) break;
        for (int round = 0; round < kRounds; ++round) {
            int q = q0 + round * NG + g;
            int col = -1;
            int outcome = kDone;
            if (q < nq) {
                col = queue[q];
                outcome = solve_column<G, kMixed>(a, L, s, col);
            }
            __syncthreads();
            const bool leader = Group<G>::lane() == 0;
            wave_append(leader && outcome == kMoveOn, col, a.overflow, a.n_overflow);
            if (kMixed) wave_append(leader && outcome == kRedoFp64, col, a.fallback, a.n_fallback);
        }
    }
}

}  // namespace

/* Route every column to the first bin whose |J| and raw-row capacity hold
 * its start pattern (the exact size of I is only known once the column runs;
 * a column that still does not fit moves on by the overflow path). */
__global__ __launch_bounds__(256) void classify_kernel(BuildArgs a, int nbins, BinEdges edges, int* queues,
                                                       int* counts, int* gqueue, int* gcount, int n)
{
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    int bin = nbins;  // generic
    if (col < n) {
        int p0 = a.p_ptr[col], nj = a.p_ptr[col + 1] - p0;
        long raw = 0;
        for (int e = p0; e < p0 + nj; ++e) {
            int j = a.p_ind[e];
            raw += a.A.colptr[j + 1] - a.A.colptr[j];
        }
        for (int b = 0; b < nbins; ++b)
            if (nj <= edges.capJ[b] && raw <= edges.capSort[b]) {
                bin = b;
                break;
            }
    }
    for (int b = 0; b < nbins; ++b)
        wave_append(col < n && bin == b, col, queues + (size_t)b * n, counts + b);
    wave_append(col < n && bin == nbins, col, gqueue, gcount);
}

__global__ void copy_ints_kernel(const int* src, int* dst, int count)
{
    int i = threadIdx.x;
    if (i < count) dst[i] = src[i];
}

void launch_binned(const BuildArgs& a, const ScratchLayout& L, int group, bool mixed, const int* queue,
                   const int* nqueue, int* head, int nwaves, hipStream_t stream)
{
    size_t lds = L.per_wave * (kWave / group);
#define MSPAI_LAUNCH(G, MX)                                                                                    \
    hipLaunchKernelGGL((spai_column_kernel<G, true, MX>), dim3(nwaves), dim3(kWave), lds, stream, a, L, queue, \
                       nqueue, head)
    if (mixed) {
        if (group == 16)
            MSPAI_LAUNCH(16, true);
        else if (group == 32)
            MSPAI_LAUNCH(32, true);
        else
            MSPAI_LAUNCH(64, true);
    } else {
        if (group == 16)
            MSPAI_LAUNCH(16, false);
        else if (group == 32)
            MSPAI_LAUNCH(32, false);
        else
            MSPAI_LAUNCH(64, false);
    }
#undef MSPAI_LAUNCH
}

void launch_classify(const BuildArgs& a, int nbins, const BinEdges& edges, int* queues, int* counts,
                     int* gqueue, int* gcount, int n, hipStream_t stream)
{
    hipLaunchKernelGGL(classify_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, a, nbins, edges, queues,
                       counts, gqueue, gcount, n);
}

void launch_copy_ints(const int* src, int* dst, int count, hipStream_t stream)
{
    hipLaunchKernelGGL(copy_ints_kernel, dim3(1), dim3(64), 0, stream, src, dst, count);
}

void launch_fp64_generic(const BuildArgs& a, const ScratchLayout& L, const int* queue, const int* nqueue,
                         int* head, int nwaves, hipStream_t stream)
{
    hipLaunchKernelGGL((spai_column_kernel<64, false, false>), dim3(nwaves), dim3(kWave), 0, stream, a, L,
                       queue, nqueue, head);
}

}  // namespace mspai
