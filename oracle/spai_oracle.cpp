/*
 * CPU FP64 oracle for the MSPAI per-column least-squares path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  See spai_oracle.h for the list
 * of reference functions restated here; each function below cites the lines
 * it follows.
 *
 * Build: oracle/Makefile (g++ -O3 -march=x86-64-v3 -ffp-contract=off).
 * -ffp-contract=off matters: every fused multiply-add below is an explicit
 * __builtin_fma, every other product and sum is rounded separately.
 *
 * The 64-lane order
 * -----------------
 * The reference calls MKL for the QR (dgeqrf), Q^T b (dormqr), the triangular
 * solve (dtrtrs) and the residual (dgemv).  MKL publishes no operation order,
 * so none can be matched; the oracle fixes one and documents it here:
 *   * A dot product over rows r0..r1-1 of one column is formed in 64 partial
 *     sums.  Row r goes to partial r % 64; each partial accumulates its rows in
 *     ascending order with fma(x_r, y_r, p), starting from +0.  The partials are
 *     then combined by the butterfly p_l <- p_l + p_(l xor s) for
 *     s = 32, 16, 8, 4, 2, 1, and the result is p_0.
 *   * Householder QR follows LAPACK dgeqr2/dlarfg/dlarf: for column i,
 *     xn2 = sum_{r>i} a_ri^2 (64-lane order), tau = 0 when xn2 == 0 or i is the
 *     last row; otherwise nrm = sqrt(fma(alpha, alpha, xn2)),
 *     beta = -copysign(nrm, alpha), tau = (beta - alpha) / beta,
 *     v_r = a_ri * (1 / (alpha - beta)).  Each later column c (and the right
 *     hand side, which is how dormqr's Q^T b is formed) gets
 *     w = v^T c (64-lane order, v_i = 1), t = (-tau) * w, c_r = fma(v_r, t, c_r).
 *     LAPACK's safe-minimum rescaling loop in dlarfg is omitted (it only acts
 *     when |beta| < 2e-292).
 *   * Back substitution follows netlib dtrsm (column oriented):
 *     for K = n-1..0: if x_K != 0 { x_K /= R_KK; x_i = fma(-x_K, R_iK, x_i), i<K }.
 *     An exact zero on the diagonal returns without solving (dtrtrs info > 0).
 *   * The residual follows netlib dgemv with beta = -1: r_i = -b_i, then for
 *     each column j with x_j != 0, r_i = fma(x_j, A_ij, r_i).
 */
#include "spai_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <list>
#include <thread>
#include <vector>

namespace {

constexpr int kLanes = 64;

inline double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

/* 64-lane dot product (see header comment). */
double lane_dot(const double* x, const double* y, int r0, int r1)
{
    double p[kLanes];
    for (int l = 0; l < kLanes; ++l) p[l] = 0.0;
    for (int r = r0; r < r1; ++r) p[r & (kLanes - 1)] = fmad(x[r], y[r], p[r & (kLanes - 1)]);
    for (int s = kLanes / 2; s >= 1; s >>= 1)
        for (int l = 0; l < s; ++l) p[l] = p[l] + p[l + s];
    return p[0];
}

/* Apply H = I - tau v v^T (v_i = 1 implicit, v_r stored in v[r] for r > i)
 * to c over rows i..m-1: dlarf('L') restated in the 64-lane order. */
void apply_reflector(double* v, double* c, int i, int m, double tau)
{
    double save = v[i];
    v[i] = 1.0;
    double w = lane_dot(v, c, i, m);
    double t = (-tau) * w;
    for (int r = i; r < m; ++r) c[r] = fmad(v[r], t, c[r]);
    v[i] = save;
}

/* Householder vector of column i (dlarfg restated). Returns tau. */
double make_reflector(double* a, int i, int m)
{
    if (i + 1 >= m) return 0.0;
    double alpha = a[i];
    double xn2 = lane_dot(a, a, i + 1, m);
    if (xn2 == 0.0) return 0.0;
    double nrm = std::sqrt(fmad(alpha, alpha, xn2));
    double beta = -std::copysign(nrm, alpha);
    double tau = (beta - alpha) / beta;
    double scal = 1.0 / (alpha - beta);
    for (int r = i + 1; r < m; ++r) a[r] = a[r] * scal;
    a[i] = beta;
    return tau;
}

/* Column-oriented back substitution of dtrtrs/dtrsm (src/Spai_Sub.cpp:89-96). */
int back_substitute(int n, const double* R, int ldr, double* x)
{
    for (int i = 0; i < n; ++i)
        if (R[i * ldr + i] == 0.0) return ORC_ST_SINGULAR;
    for (int K = n - 1; K >= 0; --K) {
        if (x[K] != 0.0) {
            x[K] = x[K] / R[K * ldr + K];
            double nx = -x[K];
            for (int i = 0; i < K; ++i) x[i] = fmad(nx, R[K * ldr + i], x[i]);
        }
    }
    return 0;
}

/* Approximative_Solution (src/Spai_Unrestrained.imp:135-195): dgeqrf, dormqr
 * ('L','T'), compaction of R, dtrtrs('U','N','N'), restated in one pass with
 * b carried as an extra column.  A is column-major m x n. */
int ls_solve(int m, int n, double* A, double* tau, double* b)
{
    int st = 0;
    int k = std::min(m, n);
    for (int i = 0; i < k; ++i) {
        double* ai = A + (size_t)i * m;
        double t = make_reflector(ai, i, m);
        tau[i] = t;
        if (t != 0.0) {
            for (int j = i + 1; j < n; ++j) apply_reflector(ai, A + (size_t)j * m, i, m, t);
            apply_reflector(ai, b, i, m, t);
        }
    }
    if (m < n) {
        /* The reference's dtrtrs then reads an n x n R from an m x n array
         * (out of bounds); the oracle returns a zero solution and flags it. */
        for (int i = 0; i < n; ++i) b[i] = 0.0;
        return st | ORC_ST_UNDERDET;
    }
    st |= back_substitute(n, A, m, b);
    return st;
}

/* dgemv('N', alpha=1, beta=-1) into r, then Euclidean_Norm
 * (src/Spai_Sub.imp:915-991, src/Spai_Sub.cpp:138-187). */
double residual(int m, int n, const double* Ahat, const double* x, const double* b0, double* r)
{
    for (int i = 0; i < m; ++i) r[i] = -b0[i];
    for (int j = 0; j < n; ++j) {
        double xj = x[j];
        if (xj == 0.0) continue;
        const double* aj = Ahat + (size_t)j * m;
        for (int i = 0; i < m; ++i) r[i] = fmad(xj, aj[i], r[i]);
    }
    double sum = 0.0;
    for (int i = 0; i < m; ++i) {
        double v = r[i];
        sum += v * v;
    }
    return std::sqrt(sum);
}

struct RhoIdx {  // RHO_IDX, src/Matrix.h:139-142
    int idx;
    double rho;
};
struct RhoCmp {  // RHO_Comparator, src/Matrix.h:147-150
    bool operator()(const RhoIdx& a, const RhoIdx& b) const { return a.rho < b.rho; }
};

/* 32-bit value hash of the LRU cache, Cache<double>::Compute_Key and
 * Hash_Double (src/Cache.cpp:42-57, src/Cache.imp:163-177). */
uint32_t hash_double(double x)
{
    uint32_t p[2];
    std::memcpy(p, &x, sizeof p);
    return p[0] ^ p[1];
}
uint32_t compute_key(const double* v, size_t n)
{
    uint32_t ret = 0, mix = hash_double(0.13);
    while (n--) ret = ret * 31u + mix + hash_double(*v++);
    return ret;
}

struct CacheEntry {
    uint32_t key;
    int m, n;
    std::vector<double> ahat, aqr, tau, x;
    uint32_t key_b;
};

/* Per-thread working storage: the role of Matrix<T>'s scratch members
 * (src/Matrix.imp:177-283) and the bitvec/reset_vec of src/Spai.imp:111-117. */
struct Work {
    std::vector<uint32_t> bitvec;
    std::vector<uint32_t> reset;
    std::vector<int> I, J, L, U, Jt, Iold, Jold;
    std::vector<double> ahat, aqr, tau, b, b0, r;
    std::vector<RhoIdx> rhos;
    std::list<CacheEntry> cache;
    bool count_near = false;  // fresh-QR path: count near-margin decisions
    int near[3] = {0, 0, 0};
};

/* Margin at or below which a selection decision counts as near. */
constexpr double kNearMargin = 1e-12;

struct Problem {
    const orc_csc* A;
    const int* rp_ptr;
    const int* rp_ind;
    const int* p_ptr;
    const int* p_ind;
    const orc_csc* B;
    const orc_opts* o;
    int cap;
    const double* ajn;  // ||a_j|| as Euclidean_Norm computes it
};

/* Get_I_Set (src/Spai_Sub.imp:65-122): bitvec union of the row indices of
 * the columns in J, then std::sort. */
void get_I(const Problem& P, Work& w)
{
    w.I.clear();
    size_t nreset = 0;
    for (int j : w.J) {
        for (int e = P.A->colptr[j]; e < P.A->colptr[j + 1]; ++e) {
            int r = P.A->rowind[e];
            int pos = r >> 5, bit = r & 31;
            if (!(w.bitvec[pos] & (1u << bit))) {
                w.bitvec[pos] |= (1u << bit);
                w.I.push_back(r);
                if (nreset >= w.reset.size()) w.reset.resize(2 * nreset + 16);
                w.reset[nreset++] = pos;
            }
        }
    }
    std::sort(w.I.begin(), w.I.end());
    for (size_t i = 0; i < nreset; ++i) w.bitvec[w.reset[i]] = 0;
}

/* Create_Submatrix_AHat (src/Spai_Sub.imp:221-282): dense column-major
 * A(I,J), zero filled, by a merge of each column's sorted rows against I. */
void gather(const Problem& P, Work& w)
{
    int m = (int)w.I.size(), n = (int)w.J.size();
    w.ahat.assign((size_t)m * n, 0.0);
    for (int jj = 0; jj < n; ++jj) {
        int j = w.J[jj];
        int e = P.A->colptr[j], e1 = P.A->colptr[j + 1];
        double* col = &w.ahat[(size_t)jj * m];
        for (int i = 0; i < m && e < e1; ++i) {
            if (w.I[i] == P.A->rowind[e]) {
                col[i] = P.A->val[e];
                ++e;
            }
        }
    }
}

/* Create_ek_Hat (src/Spai_Sub.imp:629-658) for B = identity, or
 * Create_bk_Hat(A,M,B,...) (src/Spai_Sub.imp:735-775) for a target B. */
int make_rhs(const Problem& P, Work& w, int col)
{
    int m = (int)w.I.size();
    w.b.assign(m, 0.0);
    if (!P.B) {
        for (int i = 0; i < m; ++i) {
            if (w.I[i] > col) break;
            if (w.I[i] == col) {
                w.b[i] = 1.0;
                return 0;
            }
        }
        return ORC_ST_K_NOT_IN_I;
    }
    int e = P.B->colptr[col], e1 = P.B->colptr[col + 1], pos = 0, ii = 0;
    bool hit = false;
    while (e < e1 && ii < m) {
        int ib = P.B->rowind[e], iv = w.I[ii];
        if (ib < iv) {
            ++e;
        } else if (ib > iv) {
            ++pos;
            ++ii;
        } else {
            w.b[pos++] = P.B->val[e];
            hit = true;
            ++e;
            ++ii;
        }
    }
    return hit ? 0 : ORC_ST_K_NOT_IN_I;
}

/* Union_Nl_Sets (src/Spai_Sub.imp:124-200) over L = I u {col}
 * (src/Spai_Sub.imp:1044-1059), then J~ = U \ J (std::set_difference,
 * src/Spai_Sub.imp:1066-1069), then with an upper pattern J~ n UP(:,col)
 * (Index_Set::Set_Intersection, src/Spai_Sub.imp:1072-1094). */
void candidates(const Problem& P, Work& w, int col)
{
    w.L.assign(w.I.begin(), w.I.end());
    bool found = false;
    for (int r : w.I)
        if (r == col) found = true;
    if (!found) w.L.push_back(col);
    w.U.clear();
    size_t nreset = 0;
    for (int l : w.L) {
        for (int e = P.rp_ptr[l]; e < P.rp_ptr[l + 1]; ++e) {
            int c = P.rp_ind[e];
            int pos = c >> 5, bit = c & 31;
            if (!(w.bitvec[pos] & (1u << bit))) {
                w.bitvec[pos] |= (1u << bit);
                w.U.push_back(c);
                if (nreset >= w.reset.size()) w.reset.resize(2 * nreset + 16);
                w.reset[nreset++] = pos;
            }
        }
    }
    std::sort(w.U.begin(), w.U.end());
    for (size_t i = 0; i < nreset; ++i) w.bitvec[w.reset[i]] = 0;
    w.Jt.resize(w.U.size());
    auto last = std::set_difference(w.U.begin(), w.U.end(), w.J.begin(), w.J.end(), w.Jt.begin());
    w.Jt.resize(last - w.Jt.begin());
    if (P.o->up_ptr) {
        const int* u0 = P.o->up_ind + P.o->up_ptr[col];
        const int* u1 = P.o->up_ind + P.o->up_ptr[col + 1];
        auto kept = std::set_intersection(w.Jt.begin(), w.Jt.end(), u0, u1, w.Jt.begin());
        w.Jt.resize(kept - w.Jt.begin());
    }
}

/* Compute_Rhos (src/Spai_Sub.imp:1231-1348) with Compute_Numerator
 * (src/Spai_Sub.cpp:252-291), then the selection loop and std::sort(J) of
 * Augment_Sparsity (src/Spai_Sub.imp:1100-1113).  Returns false when J~ is
 * empty; sets *nan_mean when the mean is NaN. */
bool augment(const Problem& P, Work& w, int col, double rn, bool* nan_mean, double* work)
{
    candidates(P, w, col);
    if (work) {
        for (int l : w.L) work[3] += 8.0 + 4.0 * (P.rp_ptr[l + 1] - P.rp_ptr[l]);
    }
    if (w.Jt.empty()) return false;
    const orc_csc* A = P.A;
    int m = (int)w.I.size();
    double mean = 0.0;
    w.rhos.resize(w.Jt.size());
    for (size_t t = 0; t < w.Jt.size(); ++t) {
        int j = w.Jt[t];
        double ajn = P.ajn[j];
        double num = 0.0;
        for (int e = A->colptr[j]; e < A->colptr[j + 1]; ++e) {
            int r_idx = A->rowind[e];
            for (int ip = 0; ip < m; ++ip) {
                int ii = w.I[ip];
                if (r_idx < ii) break;
                if (ii == r_idx) num += w.r[ip] * A->val[e];
            }
        }
        num *= num;
        double rho = (rn * rn) - (num / (ajn * ajn));
        rho = std::sqrt(rho);
        mean += rho;
        w.rhos[t].idx = j;
        w.rhos[t].rho = rho;
        if (work) work[4] += 16.0 + 12.0 * (A->colptr[j + 1] - A->colptr[j]);
    }
    if (work) work[5] += (double)w.Jt.size();
    if (P.o->use_mean) mean /= (double)w.Jt.size();
    if (w.count_near && !std::isnan(mean))
        for (const RhoIdx& c : w.rhos) w.near[0] += std::fabs(c.rho - mean) <= kNearMargin * std::fabs(mean);
    std::sort(w.rhos.begin(), w.rhos.end(), RhoCmp());
    *nan_mean = std::isnan(mean);
    int el = 0;
    for (; el < P.o->maxnew && el < (int)w.Jt.size() && w.rhos[el].rho <= mean; ++el)
        w.J.push_back(w.rhos[el].idx);
    /* the cut: another candidate within the threshold was left out */
    if (w.count_near && el > 0 && el == P.o->maxnew && el < (int)w.Jt.size() && w.rhos[el].rho <= mean)
        w.near[1] += w.rhos[el].rho - w.rhos[el - 1].rho <= kNearMargin * w.rhos[el - 1].rho;
    std::sort(w.J.begin(), w.J.end());
    return true;
}

void count_solve(const Problem& P, const Work& w, double* work)
{
    if (!work) return;
    double m = (double)w.I.size(), n = (double)w.J.size();
    double kk = std::min(m, n);
    work[0] += 2.0 * m * n * n - (2.0 / 3.0) * kk * kk * kk;
    work[1] += (4.0 * m * n - 2.0 * n * n) + n * n + 2.0 * m * n;
    for (int j : w.J) work[2] += 8.0 + 12.0 * (P.A->colptr[j + 1] - P.A->colptr[j]);
    work[6] += m;
    work[7] += n;
}

/* One solve with the optional LRU cache of Spai_Caching::Approximative_Solution
 * (src/Spai_Caching.imp:169-332).  A hit needs equal keys and an exactly equal
 * A-hat (Compare_A_Hat); an equal right-hand-side key reuses the stored
 * solution, otherwise the stored factors are reapplied.  Both reproduce the
 * uncached result exactly. */
int solve_cached(const Problem& P, Work& w)
{
    int m = (int)w.I.size(), n = (int)w.J.size();
    if ((int)w.b.size() < n) w.b.resize(n, 0.0);  // |I| < |J|: room for the zero solution
    w.aqr = w.ahat;
    w.tau.assign(std::max(1, std::min(m, n)), 0.0);
    if (P.o->cache <= 0) return ls_solve(m, n, w.aqr.data(), w.tau.data(), w.b.data());
    uint32_t key = compute_key(w.ahat.data(), w.ahat.size());
    uint32_t key_b = compute_key(w.b.data(), w.b.size());
    for (auto it = w.cache.begin(); it != w.cache.end(); ++it) {
        if (it->key != key) continue;
        w.cache.splice(w.cache.begin(), w.cache, it);  // In_Cache moves the hit to the front
        CacheEntry& c = w.cache.front();
        if (c.m != m || c.n != n || c.ahat != w.ahat) break;  // Compare_A_Hat failed
        if (c.key_b == key_b) {
            for (int i = 0; i < n; ++i) w.b[i] = c.x[i];
            return 0;
        }
        w.aqr = c.aqr;
        int k = std::min(m, n);
        for (int i = 0; i < k; ++i)
            if (c.tau[i] != 0.0) apply_reflector(&w.aqr[(size_t)i * m], w.b.data(), i, m, c.tau[i]);
        if (m < n) {
            for (int i = 0; i < n; ++i) w.b[i] = 0.0;
            return ORC_ST_UNDERDET;
        }
        return back_substitute(n, w.aqr.data(), m, w.b.data());
    }
    int st = ls_solve(m, n, w.aqr.data(), w.tau.data(), w.b.data());
    CacheEntry c;
    c.key = key;
    c.m = m;
    c.n = n;
    c.ahat = w.ahat;
    c.aqr = w.aqr;
    c.tau = w.tau;
    c.x.assign(w.b.begin(), w.b.begin() + n);
    c.key_b = key_b;
    if ((int)w.cache.size() >= P.o->cache) w.cache.pop_back();
    w.cache.push_front(std::move(c));
    return st;
}

/* Spai_Unrestrained<T>::SPAI_Column (src/Spai_Unrestrained.imp:33-133). */
void spai_column(const Problem& P, Work& w, int col, int* len, int* idx, double* val,
                 double* rnorm, int* nsolve, int* status, double* work)
{
    const orc_opts* o = P.o;
    w.J.assign(P.p_ind + P.p_ptr[col], P.p_ind + P.p_ptr[col + 1]);  // Get_P_Col
    int st = 0, solves = 0;
    double rn = 1.0e50;
    for (int step = -1; step < o->max_impr_steps; ++step) {
        get_I(P, w);
        gather(P, w);
        st |= make_rhs(P, w, col);
        w.b0 = w.b;
        st |= solve_cached(P, w);
        ++solves;
        count_solve(P, w, work);
        int m = (int)w.I.size(), n = (int)w.J.size();
        w.r.resize(m);
        rn = residual(m, n, w.ahat.data(), w.b.data(), w.b0.data(), w.r.data());
        if (w.count_near) w.near[2] += std::fabs(rn - o->epsilon) <= kNearMargin * o->epsilon;
        if (rn <= o->epsilon || step == o->max_impr_steps - 1) break;
        bool nan_mean = false;
        if (!augment(P, w, col, rn, &nan_mean, work)) break;
        if (nan_mean) st |= ORC_ST_NAN_RHO;
    }
    int n = (int)w.J.size();
    *len = n;
    for (int i = 0; i < n; ++i) {  // Insert_Row_Solution (src/Com_Server.cpp:415-457)
        idx[i] = w.J[i];
        val[i] = w.b[i];
    }
    *rnorm = rn;
    *nsolve = solves;
    *status = st;
}

/* ---------------- dense QR-update path (-qr 2) ---------------- */

/* State of QR_Struktur (src/qrupdates.h) for the dense update.  Storage row
 * order is the order rows were appended (sorted I first, then each step's new
 * rows), column order the order columns were appended. */
struct QRState {
    int rows = 0, cols = 0;
    std::vector<std::vector<double>> col;  // col[c] has `rows_at_append` .. grows to rows
    std::vector<double> tau;
    std::vector<int> col_id;                // column index of A per stored column
    std::vector<int> row_id;                // row index of A per storage row
    std::vector<int> refl_rows;             // rows that existed when reflector c was made
    std::vector<double> bperm;              // right-hand side in storage row order
};

/* Back substitution of qrupdate_voll (src/qrupdates.cpp:306-314): row
 * oriented, sequential, no fused multiply-add. */
int qr_update_back_substitute(QRState& S, std::vector<double>& x)
{
    int n = S.cols;
    int st = 0;
    for (int i = n - 1; i >= 0; --i) {
        double z = 0.0;
        for (int j = i + 1; j < n; ++j) z += S.col[j][i] * x[j];
        x[i] = (x[i] - z) / S.col[i][i];
        if (S.col[i][i] == 0.0) st |= ORC_ST_SINGULAR;
    }
    return st;
}

/* Apply every stored reflector in order to a copy of bperm (the Q^T b loop of
 * qrupdate_voll, src/qrupdates.cpp:270-303). */
void qr_update_apply_all(QRState& S, std::vector<double>& y)
{
    for (int c = 0; c < S.cols; ++c) {
        if (S.tau[c] == 0.0) continue;
        apply_reflector(S.col[c].data(), y.data(), c, S.refl_rows[c], S.tau[c]);
    }
}

/* Spai_QRUpdate<T>::SPAI_Column (src/Spai_QRUpdate.imp:33-215) with
 * QRZerlegungLapack and qrupdate_voll (src/qrupdates.cpp:37-327). */
void spai_column_qrupdate(const Problem& P, Work& w, int col, int* len, int* idx, double* val,
                          double* rnorm, int* nsolve, int* status, double* work)
{
    const orc_opts* o = P.o;
    w.J.assign(P.p_ind + P.p_ptr[col], P.p_ind + P.p_ptr[col + 1]);
    int st = 0, solves = 0;
    get_I(P, w);
    st |= make_rhs(P, w, col);
    gather(P, w);
    int m = (int)w.I.size(), n = (int)w.J.size();
    QRState S;
    S.rows = m;
    S.cols = n;
    S.row_id = w.I;
    S.col_id = w.J;
    S.bperm = w.b;
    S.col.resize(n);
    S.tau.assign(n, 0.0);
    S.refl_rows.assign(n, m);
    for (int j = 0; j < n; ++j) S.col[j].assign(w.ahat.begin() + (size_t)j * m, w.ahat.begin() + (size_t)(j + 1) * m);
    /* QRZerlegungLapack: dgeqrf on A-hat, dlarf per reflector on b, dtrtrs. */
    std::vector<double> x;
    {
        std::vector<double> Aqr = w.ahat, tau(std::max(1, std::min(m, n)), 0.0), b = w.b;
        int k = std::min(m, n);
        for (int i = 0; i < k; ++i) {
            double t = make_reflector(&Aqr[(size_t)i * m], i, m);
            tau[i] = t;
            if (t != 0.0)
                for (int j = i + 1; j < n; ++j) apply_reflector(&Aqr[(size_t)i * m], &Aqr[(size_t)j * m], i, m, t);
        }
        for (int i = 0; i < n && i < k; ++i)
            if (tau[i] != 0.0) apply_reflector(&Aqr[(size_t)i * m], b.data(), i, m, tau[i]);
        if (m < n) {
            st |= ORC_ST_UNDERDET;
            b.assign(n, 0.0);
        } else {
            st |= back_substitute(n, Aqr.data(), m, b.data());
        }
        for (int j = 0; j < n; ++j) {
            S.col[j].assign(Aqr.begin() + (size_t)j * m, Aqr.begin() + (size_t)(j + 1) * m);
            S.tau[j] = j < k ? tau[j] : 0.0;
        }
        x.assign(b.begin(), b.begin() + n);
    }
    ++solves;
    count_solve(P, w, work);
    w.r.resize(m);
    /* w.b still holds e_k over sorted I (the solve above worked on a copy) */
    double rn = residual(m, n, w.ahat.data(), x.data(), w.b.data(), w.r.data());
    for (int step = 0; step < o->max_impr_steps && rn > o->epsilon; ++step) {
        w.Jold = w.J;
        w.Iold = w.I;
        bool nan_mean = false;
        if (!augment(P, w, col, rn, &nan_mean, work)) break;
        if (nan_mean) st |= ORC_ST_NAN_RHO;
        get_I(P, w);
        gather(P, w);
        m = (int)w.I.size();
        n = (int)w.J.size();
        st |= make_rhs(P, w, col);
        std::vector<double> b_sorted = w.b;
        std::vector<int> Iaug, Jaug;
        std::set_difference(w.I.begin(), w.I.end(), w.Iold.begin(), w.Iold.end(), std::back_inserter(Iaug));
        std::set_difference(w.J.begin(), w.J.end(), w.Jold.begin(), w.Jold.end(), std::back_inserter(Jaug));
        int newr = (int)Iaug.size(), newc = (int)Jaug.size();
        int oldrows = S.rows, oldcols = S.cols;
        /* extend storage rows and the permuted right-hand side */
        for (int r : Iaug) {  // b_perm grows by e_col on the new rows
            S.row_id.push_back(r);
            S.bperm.push_back(r == col ? 1.0 : 0.0);
        }
        S.rows = oldrows + newr;
        for (auto& c : S.col) c.resize(S.rows, 0.0);
        /* new columns: dense values on all storage rows */
        for (int t = 0; t < newc; ++t) {
            int j = Jaug[t];
            std::vector<double> c(S.rows, 0.0);
            for (int e = P.A->colptr[j]; e < P.A->colptr[j + 1]; ++e) {
                int r = P.A->rowind[e];
                for (int s = 0; s < S.rows; ++s)
                    if (S.row_id[s] == r) {
                        c[s] = P.A->val[e];
                        break;
                    }
            }
            /* apply the old reflectors over the rows they span */
            for (int q = 0; q < oldcols; ++q)
                if (S.tau[q] != 0.0) apply_reflector(S.col[q].data(), c.data(), q, S.refl_rows[q], S.tau[q]);
            S.col.push_back(std::move(c));
            S.col_id.push_back(j);
            S.tau.push_back(0.0);
            S.refl_rows.push_back(S.rows);
        }
        S.cols = oldcols + newc;
        /* dgeqrf on the lower block B2 = rows oldcols.. of the new columns */
        for (int t = 0; t < newc; ++t) {
            int c = oldcols + t;
            if (c >= S.rows) break;
            double tt = make_reflector(S.col[c].data(), c, S.rows);
            S.tau[c] = tt;
            if (tt != 0.0)
                for (int u = c + 1; u < S.cols; ++u) apply_reflector(S.col[c].data(), S.col[u].data(), c, S.rows, tt);
        }
        std::vector<double> y = S.bperm;
        qr_update_apply_all(S, y);
        std::vector<double> xs(y.begin(), y.begin() + S.cols);
        if (S.rows < S.cols) {
            st |= ORC_ST_UNDERDET;
            std::fill(xs.begin(), xs.end(), 0.0);
        } else {
            st |= qr_update_back_substitute(S, xs);
        }
        /* unpermute by column id (qsort with vergleicher, ids are unique) */
        std::vector<int> ord(S.cols);
        for (int c = 0; c < S.cols; ++c) ord[c] = c;
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return S.col_id[a] < S.col_id[b]; });
        x.resize(S.cols);
        for (int c = 0; c < S.cols; ++c) x[c] = xs[ord[c]];
        ++solves;
        count_solve(P, w, work);
        w.r.resize(m);
        rn = residual(m, n, w.ahat.data(), x.data(), b_sorted.data(), w.r.data());
        if (step == o->max_impr_steps - 1) break;
    }
    n = (int)w.J.size();
    *len = n;
    for (int i = 0; i < n; ++i) {
        idx[i] = w.J[i];
        val[i] = x[i];
    }
    *rnorm = rn;
    *nsolve = solves;
    *status = st;
}

}  // namespace

extern "C" int orc_spai(const orc_csc* A, const int* rp_ptr, const int* rp_ind, const int* p_ptr,
                        const int* p_ind, const orc_csc* B, const orc_opts* o, const int* cols,
                        int ncols_list, int cap, int* m_len, int* m_idx, double* m_val,
                        double* res_norm, int* nsolve, int* status, double* work, int* near)
{
    if (!A || !p_ptr || !p_ind || !o || cap <= 0) return -1;
    if (o->max_impr_steps > 0 && (!rp_ptr || !rp_ind)) return -2;
    if (o->qr != 0 && o->qr != 2) return -3;
    /* probing/target matrices are static-only, and QR updates need the
     * identity target (option checks, src/imspai.cpp:431-438) */
    if (B && (o->max_impr_steps > 0 || o->qr > 0)) return -5;
    int nlist = cols ? ncols_list : A->ncols;
    /* ||a_j|| exactly as Compute_Rhos computes it: Euclidean_Norm of the
     * column values, sequential, no fused multiply-add. */
    std::vector<double> ajn(A->ncols);
    for (int j = 0; j < A->ncols; ++j) {
        double s = 0.0;
        for (int e = A->colptr[j]; e < A->colptr[j + 1]; ++e) {
            double v = A->val[e];
            s += v * v;
        }
        ajn[j] = std::sqrt(s);
    }
    Problem P{A, rp_ptr, rp_ind, p_ptr, p_ind, B, o, cap, ajn.data()};
    int nt = std::max(1, o->threads);
    if (nt > nlist) nt = std::max(1, nlist);
    int err = 0;
    auto run = [&](int lo, int hi) {
        Work w;
        int bits = std::max(A->nrows, A->ncols);
        w.bitvec.assign((bits >> 5) + 1, 0u);
        for (int p = lo; p < hi; ++p) {
            int col = cols ? cols[p] : p;
            double* wk = work ? work + (size_t)p * ORC_NWORK : nullptr;
            if (wk)
                for (int q = 0; q < ORC_NWORK; ++q) wk[q] = 0.0;
            int plen = p_ptr[col + 1] - p_ptr[col];
            if (plen + o->maxnew * std::max(0, o->max_impr_steps) > cap) {
                err = -4;
                continue;
            }
            w.near[0] = w.near[1] = w.near[2] = 0;
            w.count_near = false;
            if (o->qr == 2 && o->max_impr_steps > 0 && o->maxnew > 0) {
                spai_column_qrupdate(P, w, col, &m_len[p], m_idx + (size_t)p * cap, m_val + (size_t)p * cap,
                                     &res_norm[p], &nsolve[p], &status[p], wk);
            } else {
                w.count_near = true;
                spai_column(P, w, col, &m_len[p], m_idx + (size_t)p * cap, m_val + (size_t)p * cap,
                            &res_norm[p], &nsolve[p], &status[p], wk);
            }
            if (near)
                for (int q = 0; q < 3; ++q) near[(size_t)p * 3 + q] = w.near[q];
        }
    };
    if (nt == 1) {
        run(0, nlist);
    } else {
        /* contiguous split: the first nlist % nt threads get one more column
         * (Distribution::Basic_Distribution, src/Distribution.cpp:39-72) */
        std::vector<std::thread> th;
        int base = nlist / nt, rem = nlist % nt, lo = 0;
        for (int t = 0; t < nt; ++t) {
            int cnt = base + (t < rem ? 1 : 0);
            th.emplace_back(run, lo, lo + cnt);
            lo += cnt;
        }
        for (auto& t : th) t.join();
    }
    return err;
}

extern "C" int orc_ls_solve(int m, int n, double* A, double* tau, double* b)
{
    return ls_solve(m, n, A, tau, b);
}

extern "C" double orc_lane_dot(const double* x, const double* y, int r0, int r1)
{
    return lane_dot(x, y, r0, r1);
}

extern "C" void orc_sort_rho(int n, int* idx, double* rho)
{
    std::vector<RhoIdx> v(n);
    for (int i = 0; i < n; ++i) v[i] = {idx[i], rho[i]};
    std::sort(v.begin(), v.end(), RhoCmp());
    for (int i = 0; i < n; ++i) {
        idx[i] = v[i].idx;
        rho[i] = v[i].rho;
    }
}

/* CPU FP64 reference gradient of the static solve: for each column k,
 * recompute the QR of C(I,J), solve R^T R lambda = g, and accumulate
 * dL/dC = r lambda^T - (C lambda) m^T (r = b - C m) and dL/db = C lambda. */
extern "C" int orc_spai_backward(const orc_csc* A, const orc_csc* B, const int* m_ptr, const int* m_ind,
                                 const double* m_val, const double* gM, double* gA, double* gB, int threads)
{
    int n = A->ncols;
    int nt = std::max(1, threads);
    std::vector<std::vector<double>> gA_t(nt, std::vector<double>(A->colptr[n], 0.0));
    std::vector<std::vector<double>> gB_t(nt);
    if (B && gB)
        for (auto& v : gB_t) v.assign(B->colptr[B->ncols], 0.0);
    auto run = [&](int t, int lo, int hi) {
        std::vector<uint32_t> bitvec((std::max(A->nrows, n) >> 5) + 1, 0u);
        std::vector<int> I, J;
        std::vector<double> ahat, aqr, tau, b, x, r, lam, s;
        for (int k = lo; k < hi; ++k) {
            J.assign(m_ind + m_ptr[k], m_ind + m_ptr[k + 1]);
            I.clear();
            for (int j : J)
                for (int e = A->colptr[j]; e < A->colptr[j + 1]; ++e) {
                    int rr = A->rowind[e];
                    if (!(bitvec[rr >> 5] & (1u << (rr & 31)))) {
                        bitvec[rr >> 5] |= 1u << (rr & 31);
                        I.push_back(rr);
                    }
                }
            std::sort(I.begin(), I.end());
            for (int rr : I) bitvec[rr >> 5] = 0;
            int mm = (int)I.size(), nn = (int)J.size();
            if (mm < nn || nn == 0) continue;
            ahat.assign((size_t)mm * nn, 0.0);
            for (int jj = 0; jj < nn; ++jj) {
                int j = J[jj];
                for (int e = A->colptr[j]; e < A->colptr[j + 1]; ++e) {
                    int pos = (int)(std::lower_bound(I.begin(), I.end(), A->rowind[e]) - I.begin());
                    ahat[(size_t)jj * mm + pos] = A->val[e];
                }
            }
            b.assign(mm, 0.0);
            if (B) {
                for (int e = B->colptr[k]; e < B->colptr[k + 1]; ++e) {
                    auto it = std::lower_bound(I.begin(), I.end(), B->rowind[e]);
                    if (it != I.end() && *it == B->rowind[e]) b[it - I.begin()] = B->val[e];
                }
            } else {
                auto it = std::lower_bound(I.begin(), I.end(), k);
                if (it != I.end() && *it == k) b[it - I.begin()] = 1.0;
            }
            aqr = ahat;
            tau.assign(nn, 0.0);
            std::vector<double> btmp = b;
            ls_solve(mm, nn, aqr.data(), tau.data(), btmp.data());
            x.assign(m_val + m_ptr[k], m_val + m_ptr[k + 1]);
            /* r = b - C m */
            r = b;
            for (int jj = 0; jj < nn; ++jj)
                for (int i = 0; i < mm; ++i) r[i] -= ahat[(size_t)jj * mm + i] * x[jj];
            /* R^T y = g, R lambda = y */
            lam.assign(gM + m_ptr[k], gM + m_ptr[k + 1]);
            for (int i = 0; i < nn; ++i) {
                double v = lam[i];
                for (int q = 0; q < i; ++q) v -= aqr[(size_t)i * mm + q] * lam[q];
                lam[i] = v / aqr[(size_t)i * mm + i];
            }
            for (int i = nn - 1; i >= 0; --i) {
                double v = lam[i];
                for (int q = i + 1; q < nn; ++q) v -= aqr[(size_t)q * mm + i] * lam[q];
                lam[i] = v / aqr[(size_t)i * mm + i];
            }
            s.assign(mm, 0.0);
            for (int jj = 0; jj < nn; ++jj)
                for (int i = 0; i < mm; ++i) s[i] += ahat[(size_t)jj * mm + i] * lam[jj];
            for (int jj = 0; jj < nn; ++jj) {
                int j = J[jj];
                for (int e = A->colptr[j]; e < A->colptr[j + 1]; ++e) {
                    int pos = (int)(std::lower_bound(I.begin(), I.end(), A->rowind[e]) - I.begin());
                    gA_t[t][e] += r[pos] * lam[jj] - s[pos] * x[jj];
                }
            }
            if (B && gB) {
                for (int e = B->colptr[k]; e < B->colptr[k + 1]; ++e) {
                    auto it = std::lower_bound(I.begin(), I.end(), B->rowind[e]);
                    if (it != I.end() && *it == B->rowind[e]) gB_t[t][e] += s[it - I.begin()];
                }
            }
        }
    };
    std::vector<std::thread> th;
    int base = n / nt, rem = n % nt, lo = 0;
    for (int t = 0; t < nt; ++t) {
        int cnt = base + (t < rem ? 1 : 0);
        th.emplace_back(run, t, lo, lo + cnt);
        lo += cnt;
    }
    for (auto& t : th) t.join();
    for (int e = 0; e < A->colptr[n]; ++e) {
        double v = 0.0;
        for (int t = 0; t < nt; ++t) v += gA_t[t][e];
        gA[e] = v;
    }
    if (B && gB)
        for (int e = 0; e < B->colptr[B->ncols]; ++e) {
            double v = 0.0;
            for (int t = 0; t < nt; ++t) v += gB_t[t][e];
            gB[e] = v;
        }
    return 0;
}
