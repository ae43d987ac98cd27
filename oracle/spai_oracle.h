/*
 * CPU FP64 oracle for the MSPAI per-column least-squares path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.  The product (mspai_amd/) never
 * links, imports or calls anything under oracle/.
 *
 * It restates the reference's algorithm (dmancy/mspai, MSPAI 1.2):
 *   Spai<T>::SPAI_Algorithm            src/Spai.imp:33-191
 *   Spai_Unrestrained<T>::SPAI_Column  src/Spai_Unrestrained.imp:33-133
 *   Approximative_Solution             src/Spai_Unrestrained.imp:135-195
 *   Spai_Sub<T> (I-set, gather, e_k, residual, augmentation, rho)
 *                                      src/Spai_Sub.imp:65-282,629-658,735-775,
 *                                      915-991,1010-1116,1231-1348;
 *                                      src/Spai_Sub.cpp:46-187,252-291
 *   QR-update path (-qr 2)             src/Spai_QRUpdate.imp:33-215,321-377;
 *                                      src/qrupdates.cpp:37-327,1032-1038
 *   LRU cache (-cs N)                  src/Spai_Caching.imp:45-332,
 *                                      src/Cache.imp:65-177, src/Cache.cpp:42-57
 *
 * Arithmetic that the reference compiles itself (Sqrt_Sum, Compute_Numerator,
 * the rho formula, the mean, the qrupdate back substitution) is reproduced
 * operation for operation, without fused multiply-add (the reference is built
 * -O0 for x86-64, which has no contraction).  Arithmetic the reference leaves
 * to MKL (dgeqrf, dormqr, dtrtrs, dgemv) has no published operation order;
 * the oracle fixes one, the "64-lane order" documented in spai_oracle.cpp, and
 * the GPU FP64 kernel reproduces it bit for bit.
 */
#ifndef MSPAI_ORACLE_H
#define MSPAI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* per-column status bits (shared meaning with the GPU library) */
enum {
    ORC_ST_SINGULAR = 1,  /* exact zero pivot in R: dtrtrs leaves Q^T b unsolved */
    ORC_ST_UNDERDET = 2,  /* |I| < |J| in some solve */
    ORC_ST_NAN_RHO  = 4,  /* mean of rho is NaN: no index selected */
    ORC_ST_K_NOT_IN_I = 16 /* target column has no entry in I (unit entry dropped) */
};

typedef struct {
    double epsilon;       /* -ep  */
    int maxnew;           /* -mn  */
    int max_impr_steps;   /* -ns  */
    int use_mean;         /* -um  */
    int qr;               /* -qr: 0 = plain, 2 = dense QR update */
    int cache;            /* -cs: LRU cache entries, 0 = off */
    int threads;          /* host threads, contiguous split of the column list */
    const int* up_ptr;    /* -up: upper pattern (CSC pattern, sorted), NULL = none */
    const int* up_ind;
} orc_opts;

typedef struct {
    int nrows, ncols;
    const int* colptr;    /* ncols+1 */
    const int* rowind;    /* sorted ascending within each column */
    const double* val;
} orc_csc;

/* Number of work counters written per column by orc_spai. */
#define ORC_NWORK 8

/*
 * Build the columns listed in `cols` (NULL = all ncols columns) of M with
 * A*M ~ B (B = NULL means the identity).
 *   rp_ptr/rp_ind : row pattern of A (CSR pattern, sorted column indices)
 *   p_ptr/p_ind   : start pattern J0 (CSC pattern, sorted)
 *   cap           : capacity per column of the output slots
 * Outputs (indexed by position in `cols`):
 *   m_len, m_idx[pos*cap ..], m_val[pos*cap ..], res_norm, nsolve, status,
 *   work[pos*ORC_NWORK ..] (may be NULL),
 *   near[pos*3 ..] (may be NULL): near-margin decisions of the fresh-QR
 *   path, margin at most 1e-12 relative: scores against the mean, the cut
 *   after the maxnew-th score (ties across it included), ||r|| against eps
 *   (zero on the QR-update path).
 * Returns 0, or a negative error code.
 */
int orc_spai(const orc_csc* A, const int* rp_ptr, const int* rp_ind,
             const int* p_ptr, const int* p_ind, const orc_csc* B,
             const orc_opts* o, const int* cols, int ncols_list, int cap,
             int* m_len, int* m_idx, double* m_val, double* res_norm,
             int* nsolve, int* status, double* work, int* near);

/* Dense least squares min ||A x - b|| in the oracle's order.
 * A: column-major m x n (overwritten by R and the Householder vectors),
 * tau: n, b: m (on exit b[0..n) = x).  Returns status bits. */
int orc_ls_solve(int m, int n, double* A, double* tau, double* b);

/* The 64-lane dot product sum_{r0<=r<r1} x[r]*y[r] in the oracle's order. */
double orc_lane_dot(const double* x, const double* y, int r0, int r1);

/* std::sort (libstdc++) of (idx, rho) pairs with the reference comparator
 * a.rho < b.rho (src/Matrix.h:135-150); for checking the GPU emulation. */
void orc_sort_rho(int n, int* idx, double* rho);

/* CPU reference gradient of a static solve (probing-capable): for every
 * column k of M, given g = dL/dm_k, accumulate dL/dC(i,j) into gA (indexed
 * like A's values) and dL/dB(i,k) into gB (indexed like B's values, may be
 * NULL).  M must come from orc_spai with ns = 0 on the same A, B, pattern. */
int orc_spai_backward(const orc_csc* A, const orc_csc* B, const int* m_ptr,
                      const int* m_ind, const double* m_val, const double* gM,
                      double* gA, double* gB, int threads);


/*
 * Block matrix, the layout Convert_To_Block_Matrix builds
 * (src/Matrix.imp:1516-1640): nb block columns (= block rows) of sizes
 * bsize[]; block column j holds its blocks at block rows
 * rowind[colptr[j]..colptr[j+1]) (ascending); block e is stored column-major
 * (bsize[rowind[e]] x bsize[j]) at val[voff[e] .. voff[e+1]).
 */
typedef struct {
    int nb;
    const int* bsize;
    const int* colptr;
    const int* rowind;
    const long long* voff;  /* nnzb + 1 */
    const double* val;
} orc_bcsc;

/*
 * Block SPAI (Spai_Unrestrained_Block, src/Spai_Unrestrained_Block.imp:33-205)
 * for every block column.  rp_ptr/rp_ind: block row pattern (CSR);
 * p_ptr/p_ind: block start pattern (CSC).  Outputs per block column `col`:
 * m_len[col] blocks, their block rows m_idx[col*capJ ..], and the values
 * m_val[col*capX ..] with block t a column-major bsize[J[t]] x bsize[col]
 * block (Get_Solution_Vals_Block).  Returns 0 or a negative error code.
 */
int orc_spai_block(const orc_bcsc* A, const int* rp_ptr, const int* rp_ind, const int* p_ptr,
                   const int* p_ind, const orc_opts* o, int capJ, long long capX, int* m_len,
                   int* m_idx, double* m_val, double* res_norm, int* nsolve, int* status);

/* (A_j^T A_j)^-1 of every block column, concatenated (sum of bsize^2). */
void orc_block_sq_inverse(const orc_bcsc* A, double* out);

/* One static solve of block column col on the block set J, then the rho of
 * every augmentation candidate (in candidate order).  For checking the rho
 * formula against brute-force least squares. */
int orc_block_rhos(const orc_bcsc* A, const int* rp_ptr, const int* rp_ind, int col, const int* J,
                   int nj, double* rn, int* ncand, int* cand, double* rho, int cap);

#ifdef __cplusplus
}
#endif
#endif
