/*
 * CPU FP64 oracle for block SPAI (-bs != 1).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library's block path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product never links or calls it.
 *
 * It restates, on the block matrix that Convert_To_Block_Matrix builds
 * (src/Matrix.imp:1516-1640):
 *   Spai_Unrestrained_Block<T>::SPAI_Column   src/Spai_Unrestrained_Block.imp:33-140
 *   Approximative_Solution (block)            src/Spai_Unrestrained_Block.imp:142-205
 *   Create_Submatrix_AHat_Block               src/Spai_Sub.imp:285-401
 *   Create_ek_Hat_Block                       src/Spai_Sub.imp:691-733
 *   Get_Solution_Vals_Scalar / _Block         src/Spai_Sub.imp:820-868
 *   Convert_To_Block                          src/Spai_Sub.imp:871-889
 *   Residual_Norm_Block, Sqrt_Sum_Matrix      src/Spai_Sub.imp:944-1008, src/Spai_Sub.cpp:190-200
 *   Augment_Sparsity_Block                    src/Spai_Sub.imp:1119-1229
 *   Compute_Rhos_Block                        src/Spai_Sub.imp:1351-1466
 *   Compute_Numerator_Block                   src/Spai_Sub.cpp:294-370
 *   Precomputation_Column_Square_Inverses,
 *   Compute_Square_Inverse, Mult_Blocks_TN    src/Matrix.imp:2945-3029, src/Matrix.cpp:556-579
 *
 * Arithmetic the reference compiles itself (Mult_Blocks_TN, the trace term of
 * the rho formula, Sqrt_Sum_Matrix, the mean) is reproduced operation for
 * operation without fused multiply-add.  Arithmetic it leaves to LAPACK/BLAS
 * (dgeqrf, dormqr, dtrtrs with several right-hand sides, the dgemm products,
 * dgetrf + dgetri of A_j^T A_j) has no published operation order; the oracle
 * fixes one -- the 64-lane order of spai_oracle.cpp for the QR, per-column
 * dgemv order for the residual, sequential fused sums for the small dgemm
 * products, Gauss-Jordan elimination with partial pivoting for the inverse --
 * and the GPU block kernel reproduces it bit for bit.
 */
#include "spai_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

constexpr int kLanes = 64;

inline double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

double lane_dot(const double* x, const double* y, int r0, int r1)
{
    double p[kLanes];
    for (int l = 0; l < kLanes; ++l) p[l] = 0.0;
    for (int r = r0; r < r1; ++r) p[r & (kLanes - 1)] = fmad(x[r], y[r], p[r & (kLanes - 1)]);
    for (int s = kLanes / 2; s >= 1; s >>= 1)
        for (int l = 0; l < s; ++l) p[l] = p[l] + p[l + s];
    return p[0];
}

void apply_reflector(double* v, double* c, int i, int m, double tau)
{
    double save = v[i];
    v[i] = 1.0;
    double w = lane_dot(v, c, i, m);
    double t = (-tau) * w;
    for (int r = i; r < m; ++r) c[r] = fmad(v[r], t, c[r]);
    v[i] = save;
}

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

struct RhoIdx {  // RHO_IDX, src/Matrix.h:139-142
    int idx;
    double rho;
};
struct RhoCmp {  // RHO_Comparator, src/Matrix.h:147-150
    bool operator()(const RhoIdx& a, const RhoIdx& b) const { return a.rho < b.rho; }
};

struct BProblem {
    const orc_bcsc* A;
    const int* rp_ptr;
    const int* rp_ind;
    const int* p_ptr;
    const int* p_ind;
    const orc_opts* o;
    const double* sqinv;          // (A_j^T A_j)^-1 per block column
    const long long* sqoff;       // offset of block column j's inverse
};

struct BWork {
    std::vector<uint32_t> bitvec;
    std::vector<int> touched;
    std::vector<int> I, J, L, U, Jt, fr, fc;
    std::vector<double> ahat, aqr, tau, b, b0, x, r, rb, N, Bm;
    std::vector<RhoIdx> rhos;
};

/* Union of the index lists of `sets` (bitvec union, then std::sort), the
 * pattern of Get_I_Set and Union_Nl_Sets (src/Spai_Sub.imp:65-200). */
void sorted_union(BWork& w, const std::vector<int>& sets, const int* ptr, const int* ind, std::vector<int>& out)
{
    out.clear();
    w.touched.clear();
    for (int s : sets) {
        for (int e = ptr[s]; e < ptr[s + 1]; ++e) {
            int v = ind[e];
            uint32_t& word = w.bitvec[v >> 5];
            uint32_t bit = 1u << (v & 31);
            if (!(word & bit)) {
                word |= bit;
                out.push_back(v);
                w.touched.push_back(v >> 5);
            }
        }
    }
    std::sort(out.begin(), out.end());
    for (int t : w.touched) w.bitvec[t] = 0;
}

/* first_index_set / first_index_row: scalar offset of each block of a list. */
int prefix_sizes(const orc_bcsc* A, const std::vector<int>& blocks, std::vector<int>& first)
{
    first.resize(blocks.size() + 1);
    first[0] = 0;
    for (size_t t = 0; t < blocks.size(); ++t) first[t + 1] = first[t] + A->bsize[blocks[t]];
    return first[blocks.size()];
}

/* Create_Submatrix_AHat_Block: dense column-major A(I,J) of scalar size
 * m x n; each stored block of A(:,J) lands at its block row's offset. */
void gather_block(const BProblem& P, BWork& w, int m, int n)
{
    const orc_bcsc* A = P.A;
    w.ahat.assign((size_t)m * n, 0.0);
    for (size_t jj = 0; jj < w.J.size(); ++jj) {
        int j = w.J[jj], bj = A->bsize[j];
        size_t i = 0;
        for (int e = A->colptr[j]; e < A->colptr[j + 1]; ++e) {
            int ib = A->rowind[e];
            while (i < w.I.size() && w.I[i] < ib) ++i;
            int h = A->bsize[ib];
            const double* blk = A->val + A->voff[e];
            for (int c = 0; c < bj; ++c)
                for (int rr = 0; rr < h; ++rr)
                    w.ahat[(size_t)(w.fc[jj] + c) * m + w.fr[i] + rr] = blk[c * h + rr];
        }
    }
}

/* Create_ek_Hat_Block: the identity block of block row `col` at its scalar
 * offset in each of the bk right-hand sides (the scan stops past col). */
int make_rhs_block(const BProblem& P, BWork& w, int col, int m, int bk)
{
    w.b.assign((size_t)m * bk, 0.0);
    for (size_t i = 0; i < w.I.size(); ++i) {
        if (w.I[i] > col) break;
        if (w.I[i] == col) {
            for (int q = 0; q < bk; ++q) w.b[(size_t)q * m + w.fr[i] + q] = 1.0;
            return 0;
        }
    }
    (void)P;
    return ORC_ST_K_NOT_IN_I;
}

/* QR of A(I,J), Q^T applied to every right-hand side, R^-1 per right-hand
 * side (dtrtrs with nrhs = bk).  On exit x (n x bk) holds the solution,
 * packed as Get_Solution_Vals_Scalar does. */
int solve_block(BWork& w, int m, int n, int bk)
{
    int st = 0;
    w.aqr = w.ahat;
    int k = std::min(m, n);
    w.tau.assign(std::max(1, k), 0.0);
    for (int i = 0; i < k; ++i) {
        double* ai = &w.aqr[(size_t)i * m];
        double t = make_reflector(ai, i, m);
        w.tau[i] = t;
        if (t != 0.0) {
            for (int j = i + 1; j < n; ++j) apply_reflector(ai, &w.aqr[(size_t)j * m], i, m, t);
            for (int q = 0; q < bk; ++q) apply_reflector(ai, &w.b[(size_t)q * m], i, m, t);
        }
    }
    w.x.assign((size_t)n * bk, 0.0);
    if (m < n) return st | ORC_ST_UNDERDET;
    bool singular = false;
    for (int i = 0; i < n; ++i)
        if (w.aqr[(size_t)i * m + i] == 0.0) singular = true;
    if (!singular) {
        for (int q = 0; q < bk; ++q) {
            double* xq = &w.b[(size_t)q * m];
            for (int K = n - 1; K >= 0; --K) {
                if (xq[K] != 0.0) {
                    xq[K] = xq[K] / w.aqr[(size_t)K * m + K];
                    double nx = -xq[K];
                    for (int i = 0; i < K; ++i) xq[i] = fmad(nx, w.aqr[(size_t)K * m + i], xq[i]);
                }
            }
        }
    } else {
        st |= ORC_ST_SINGULAR;  // dtrtrs returns info > 0 and leaves Q^T E
    }
    for (int q = 0; q < bk; ++q)
        for (int j = 0; j < n; ++j) w.x[(size_t)q * n + j] = w.b[(size_t)q * m + j];
    return st;
}

/* Residual_Norm_Block: R = A(I,J) X - E (per right-hand side in dgemv's
 * column order), then the Frobenius norm by Sqrt_Sum_Matrix. */
double residual_block(BWork& w, int m, int n, int bk)
{
    w.r.assign((size_t)m * bk, 0.0);
    for (int q = 0; q < bk; ++q) {
        double* rq = &w.r[(size_t)q * m];
        for (int i = 0; i < m; ++i) rq[i] = -w.b0[(size_t)q * m + i];
        for (int j = 0; j < n; ++j) {
            double xj = w.x[(size_t)q * n + j];
            if (xj == 0.0) continue;
            const double* aj = &w.ahat[(size_t)j * m];
            for (int i = 0; i < m; ++i) rq[i] = fmad(xj, aj[i], rq[i]);
        }
    }
    double sum = 0.0;
    for (size_t e = 0; e < (size_t)m * bk; ++e) {
        double v = w.r[e];
        sum += v * v;
    }
    return std::sqrt(sum);
}

/* Candidates J~ = (union of the block rows' patterns over I u {col}) \ J. */
void candidates_block(const BProblem& P, BWork& w, int col)
{
    w.L.assign(w.I.begin(), w.I.end());
    if (std::find(w.I.begin(), w.I.end(), col) == w.I.end()) w.L.push_back(col);
    sorted_union(w, w.L, P.rp_ptr, P.rp_ind, w.U);
    w.Jt.resize(w.U.size());
    auto last = std::set_difference(w.U.begin(), w.U.end(), w.J.begin(), w.J.end(), w.Jt.begin());
    w.Jt.resize(last - w.Jt.begin());
}

/* rho_j of Compute_Rhos_Block for candidate block column j:
 *   N = A(I,j)^T R (Compute_Numerator_Block, one dgemm per shared block row),
 *   B = (A_j^T A_j)^-1 N, term = sum_q sum_i B(i,q) N(i,q),
 *   rho = sqrt(||R||^2 - term).  rb is R in block order (Convert_To_Block). */
double rho_block(const BProblem& P, BWork& w, int j, int bk, double rn)
{
    const orc_bcsc* A = P.A;
    const int bj = A->bsize[j];
    w.N.assign((size_t)bj * bk, 0.0);
    size_t ir = 0;
    for (int e = A->colptr[j]; e < A->colptr[j + 1] && ir < w.I.size(); ++e) {
        int ib = A->rowind[e];
        while (ir < w.I.size() && w.I[ir] < ib) ++ir;
        if (ir == w.I.size() || w.I[ir] != ib) continue;
        int h = A->bsize[ib];
        const double* a = A->val + A->voff[e];
        const double* rblk = &w.rb[(size_t)w.fr[ir] * bk];
        for (int q = 0; q < bk; ++q)
            for (int ii = 0; ii < bj; ++ii) {
                double d = 0.0;
                for (int t = 0; t < h; ++t) d = fmad(a[ii * h + t], rblk[q * h + t], d);
                w.N[(size_t)q * bj + ii] = w.N[(size_t)q * bj + ii] + d;
            }
    }
    const double* inv = P.sqinv + P.sqoff[j];
    w.Bm.assign((size_t)bj * bk, 0.0);
    for (int q = 0; q < bk; ++q)
        for (int ii = 0; ii < bj; ++ii) {
            double d = 0.0;
            for (int kk = 0; kk < bj; ++kk) d = fmad(inv[kk * bj + ii], w.N[(size_t)q * bj + kk], d);
            w.Bm[(size_t)q * bj + ii] = d;
        }
    double term = 0.0;
    for (int q = 0; q < bk; ++q)
        for (int ii = 0; ii < bj; ++ii) term = term + w.Bm[(size_t)q * bj + ii] * w.N[(size_t)q * bj + ii];
    double rho = (rn * rn) - term;
    return std::sqrt(rho);
}

/* Convert_To_Block: R (m x bk, column-major) into block-row order, each block
 * column-major (bsize[I[k]] x bk). */
void residual_to_blocks(const BProblem& P, BWork& w, int m, int bk)
{
    w.rb.resize((size_t)m * bk);
    size_t o = 0;
    for (size_t k = 0; k < w.I.size(); ++k) {
        int h = P.A->bsize[w.I[k]];
        for (int q = 0; q < bk; ++q)
            for (int i = 0; i < h; ++i) w.rb[o++] = w.r[(size_t)q * m + w.fr[k] + i];
    }
}

/* Augment_Sparsity_Block.  Returns false when J~ is empty. */
bool augment_block(const BProblem& P, BWork& w, int col, int bk, int m, double rn, bool* nan_mean)
{
    candidates_block(P, w, col);
    if (w.Jt.empty()) return false;
    residual_to_blocks(P, w, m, bk);
    double mean = 0.0;
    w.rhos.resize(w.Jt.size());
    for (size_t t = 0; t < w.Jt.size(); ++t) {
        double rho = rho_block(P, w, w.Jt[t], bk, rn);
        mean += rho;
        w.rhos[t].idx = w.Jt[t];
        w.rhos[t].rho = rho;
    }
    if (P.o->use_mean) mean /= (double)w.Jt.size();
    std::sort(w.rhos.begin(), w.rhos.end(), RhoCmp());
    *nan_mean = std::isnan(mean);
    for (int el = 0; el < P.o->maxnew && el < (int)w.Jt.size() && w.rhos[el].rho <= mean; ++el)
        w.J.push_back(w.rhos[el].idx);
    std::sort(w.J.begin(), w.J.end());
    return true;
}

/* Spai_Unrestrained_Block<T>::SPAI_Column for block column `col`. */
void spai_block_column(const BProblem& P, BWork& w, int col, int* len, int* idx, double* val, double* rnorm,
                       int* nsolve, int* status)
{
    const orc_opts* o = P.o;
    const orc_bcsc* A = P.A;
    const int bk = A->bsize[col];
    const double eps = o->epsilon * std::sqrt((double)bk);  // adjust_epsilon
    w.J.assign(P.p_ind + P.p_ptr[col], P.p_ind + P.p_ptr[col + 1]);
    int st = 0, solves = 0;
    double rn = 1.0e50;
    for (int step = -1; step < o->max_impr_steps; ++step) {
        sorted_union(w, w.J, A->colptr, A->rowind, w.I);
        int m = prefix_sizes(A, w.I, w.fr);
        int n = prefix_sizes(A, w.J, w.fc);
        gather_block(P, w, m, n);
        st |= make_rhs_block(P, w, col, m, bk);
        w.b0 = w.b;
        st |= solve_block(w, m, n, bk);
        ++solves;
        rn = residual_block(w, m, n, bk);
        if (rn <= eps || step == o->max_impr_steps - 1) break;
        bool nan_mean = false;
        if (!augment_block(P, w, col, bk, m, rn, &nan_mean)) break;
        if (nan_mean) st |= ORC_ST_NAN_RHO;
    }
    /* Get_Solution_Vals_Block + Insert_Row_Solution_Block: block t of J as a
     * column-major bsize[J[t]] x bk block */
    int n = prefix_sizes(A, w.J, w.fc);
    *len = (int)w.J.size();
    size_t o2 = 0;
    for (size_t t = 0; t < w.J.size(); ++t) {
        idx[t] = w.J[t];
        int bj = A->bsize[w.J[t]];
        for (int q = 0; q < bk; ++q)
            for (int jj = 0; jj < bj; ++jj) val[o2++] = w.x[(size_t)q * n + w.fc[t] + jj];
    }
    *rnorm = rn;
    *nsolve = solves;
    *status = st;
}

/* Compute_Square_Inverse: S = A_j^T A_j by Mult_Blocks_TN over the stored
 * blocks (unfused, in storage order), then S^-1 by Gauss-Jordan elimination
 * with partial pivoting (first largest |pivot|). */
void square_inverse(const orc_bcsc* A, int j, double* inv)
{
    const int bj = A->bsize[j];
    std::vector<double> S((size_t)bj * bj, 0.0);
    for (int e = A->colptr[j]; e < A->colptr[j + 1]; ++e) {
        int h = A->bsize[A->rowind[e]];
        const double* a = A->val + A->voff[e];
        for (int i1 = 0; i1 < bj; ++i1)
            for (int j1 = 0; j1 < bj; ++j1)
                for (int k1 = 0; k1 < h; ++k1) S[i1 + (size_t)j1 * bj] += a[k1 + i1 * h] * a[k1 + j1 * h];
    }
    for (int i = 0; i < bj * bj; ++i) inv[i] = 0.0;
    for (int i = 0; i < bj; ++i) inv[i * bj + i] = 1.0;
    for (int c = 0; c < bj; ++c) {
        int p = c;
        for (int r = c + 1; r < bj; ++r)
            if (std::fabs(S[r + (size_t)c * bj]) > std::fabs(S[p + (size_t)c * bj])) p = r;
        if (p != c)
            for (int k = 0; k < bj; ++k) {
                std::swap(S[c + (size_t)k * bj], S[p + (size_t)k * bj]);
                std::swap(inv[c + k * bj], inv[p + k * bj]);
            }
        double piv = S[c + (size_t)c * bj];
        for (int k = 0; k < bj; ++k) {
            S[c + (size_t)k * bj] = S[c + (size_t)k * bj] / piv;
            inv[c + k * bj] = inv[c + k * bj] / piv;
        }
        for (int r = 0; r < bj; ++r) {
            if (r == c) continue;
            double f = S[r + (size_t)c * bj];
            if (f == 0.0) continue;
            for (int k = 0; k < bj; ++k) {
                S[r + (size_t)k * bj] = fmad(-f, S[c + (size_t)k * bj], S[r + (size_t)k * bj]);
                inv[r + k * bj] = fmad(-f, inv[c + k * bj], inv[r + k * bj]);
            }
        }
    }
}

std::vector<long long> square_offsets(const orc_bcsc* A)
{
    std::vector<long long> off(A->nb + 1, 0);
    for (int j = 0; j < A->nb; ++j) off[j + 1] = off[j] + (long long)A->bsize[j] * A->bsize[j];
    return off;
}

}  // namespace

extern "C" void orc_block_sq_inverse(const orc_bcsc* A, double* out)
{
    std::vector<long long> off = square_offsets(A);
    for (int j = 0; j < A->nb; ++j) square_inverse(A, j, out + off[j]);
}

extern "C" int orc_spai_block(const orc_bcsc* A, const int* rp_ptr, const int* rp_ind, const int* p_ptr,
                              const int* p_ind, const orc_opts* o, int capJ, long long capX, int* m_len,
                              int* m_idx, double* m_val, double* res_norm, int* nsolve, int* status)
{
    if (!A || !p_ptr || !p_ind || !o || capJ <= 0 || capX <= 0) return -1;
    if (o->max_impr_steps > 0 && (!rp_ptr || !rp_ind)) return -2;
    if (o->qr != 0 || o->cache != 0) return -3;
    std::vector<long long> off = square_offsets(A);
    std::vector<double> sqinv(off[A->nb]);
    for (int j = 0; j < A->nb; ++j) square_inverse(A, j, sqinv.data() + off[j]);
    BProblem P{A, rp_ptr, rp_ind, p_ptr, p_ind, o, sqinv.data(), off.data()};
    int maxbs = 1;
    for (int j = 0; j < A->nb; ++j) maxbs = std::max(maxbs, A->bsize[j]);
    int err = 0;
    auto run = [&](int lo, int hi) {
        BWork w;
        w.bitvec.assign((A->nb >> 5) + 1, 0u);
        for (int col = lo; col < hi; ++col) {
            int plen = p_ptr[col + 1] - p_ptr[col];
            /* J holds distinct block columns: never more than nb */
            if (std::min(A->nb, plen + o->maxnew * std::max(0, o->max_impr_steps)) > capJ ||
                (long long)capJ * maxbs * A->bsize[col] > capX) {
                err = -4;
                continue;
            }
            spai_block_column(P, w, col, &m_len[col], m_idx + (size_t)col * capJ, m_val + (size_t)col * capX,
                              &res_norm[col], &nsolve[col], &status[col]);
        }
    };
    int nt = std::max(1, std::min(o->threads, A->nb));
    if (nt == 1) {
        run(0, A->nb);
    } else {
        std::vector<std::thread> th;
        int base = A->nb / nt, rem = A->nb % nt, lo = 0;
        for (int t = 0; t < nt; ++t) {
            int cnt = base + (t < rem ? 1 : 0);
            th.emplace_back(run, lo, lo + cnt);
            lo += cnt;
        }
        for (auto& t : th) t.join();
    }
    return err;
}

extern "C" int orc_block_rhos(const orc_bcsc* A, const int* rp_ptr, const int* rp_ind, int col, const int* J,
                              int nj, double* rn_out, int* ncand, int* cand, double* rho, int cap)
{
    if (!A || !rp_ptr || !rp_ind || !J || col < 0 || col >= A->nb) return -1;
    std::vector<long long> off = square_offsets(A);
    std::vector<double> sqinv(off[A->nb]);
    for (int j = 0; j < A->nb; ++j) square_inverse(A, j, sqinv.data() + off[j]);
    orc_opts o{};
    BProblem P{A, rp_ptr, rp_ind, nullptr, nullptr, &o, sqinv.data(), off.data()};
    BWork w;
    w.bitvec.assign((A->nb >> 5) + 1, 0u);
    w.J.assign(J, J + nj);
    const int bk = A->bsize[col];
    sorted_union(w, w.J, A->colptr, A->rowind, w.I);
    int m = prefix_sizes(A, w.I, w.fr);
    int n = prefix_sizes(A, w.J, w.fc);
    gather_block(P, w, m, n);
    make_rhs_block(P, w, col, m, bk);
    w.b0 = w.b;
    solve_block(w, m, n, bk);
    double rn = residual_block(w, m, n, bk);
    *rn_out = rn;
    candidates_block(P, w, col);
    if ((int)w.Jt.size() > cap) return -2;
    residual_to_blocks(P, w, m, bk);
    for (size_t t = 0; t < w.Jt.size(); ++t) {
        cand[t] = w.Jt[t];
        rho[t] = rho_block(P, w, w.Jt[t], bk, rn);
    }
    *ncand = (int)w.Jt.size();
    return 0;
}
