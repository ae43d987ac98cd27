/*
 * mspai_gpu: MI355X (gfx950) builder for SPAI/MSPAI preconditioners.
 *
 * C ABI of the drop-in for the reference's per-column least-squares path.
 * The reference (dmancy/mspai) runs that path as
 *   Spai<T>::SPAI_Algorithm(A, M, B, P, UP, ep, mn, ns, hs, um, pk, pm, count, vb)
 *     src/Spai.h:89-153, src/Spai.imp:33-191, called at src/imspai.cpp:831-833
 * which loops Spai_Unrestrained<T>::SPAI_Column (src/Spai_Unrestrained.imp:33-133)
 * over every column.  Here the whole loop is one device build over a
 * GPU-resident CSC copy of A.  All functions return an mspai_status (0 = OK);
 * nothing throws across this boundary.  A context is not thread safe.
 */
#ifndef MSPAI_GPU_H
#define MSPAI_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mspai_ctx mspai_ctx;

/* Error codes.  The reference throws std::runtime_error for bad options
 * (src/imspai.cpp:239-480) and ignores LAPACK info (src/Spai_Sub.cpp:89-121);
 * both become codes here. */
typedef enum {
    MSPAI_OK = 0,
    MSPAI_ERR_ARG = -1,        /* bad argument or option (message: mspai_last_error) */
    MSPAI_ERR_STATE = -2,      /* call order (e.g. build before upload) */
    MSPAI_ERR_HIP = -3,        /* HIP runtime failure */
    MSPAI_ERR_NO_DEVICE = -4,  /* no gfx950 device visible */
    MSPAI_ERR_CAPACITY = -5,   /* a column exceeds the library's size limits */
    MSPAI_ERR_UNSUPPORTED = -6 /* option combination the reference rejects */
} mspai_status;

/* Per-column status bits (mspai_download_column_info). */
enum {
    MSPAI_COL_SINGULAR = 1,    /* exact zero pivot: Q^T b left unsolved, as dtrtrs does */
    MSPAI_COL_UNDERDET = 2,    /* |I| < |J| in some solve (solution set to 0) */
    MSPAI_COL_NAN_RHO = 4,     /* mean of the candidate scores is NaN: nothing added */
    MSPAI_COL_FALLBACK = 8,    /* mixed path: recomputed in FP64 (selection margin guard) */
    MSPAI_COL_K_NOT_IN_I = 16  /* target column has no entry in I (src/Spai_Sub.imp:629-658) */
};

/* Options: the mspai keys of Command_Line (defaults src/Command_Line.cpp:192-226,
 * PETSc keys src/imspai.cpp:87-218) that act on this path, plus GPU keys. */
typedef struct {
    double epsilon;        /* -ep  residual bound per column (default 0.4) */
    int maxnew;            /* -mn  max new indices per augmentation step (5) */
    int max_impr_steps;    /* -ns  augmentation steps (5); 0 = static SPAI */
    int use_mean;          /* -um  1: accept rho <= mean; 0: threshold = sum (1) */
    int qr;                /* -qr  0 = fresh QR per step, 2 = dense QR update (0) */
    int cache;             /* -cs  device pattern->R cache entries, 0 = off (0) */
    int precision;         /* GPU: 0 = FP64, 1 = FP32 QR + FP64 refinement (0) */
    int binning;           /* GPU: 1 = default size bins, 2 = autotuned bins (mspai_autotune),
                              0 = one generic bin (1) */
    int verbose;           /* -vb */
    int counters;          /* GPU: 1 = store the 8 work counters of every column (profiling,
                              mspai_work_totals); 0 = off, no counter traffic (0) */
} mspai_opts;

/* Fill `o` with the reference's defaults. */
void mspai_default_opts(mspai_opts* o);

/* Timing and counters of the last build. */
typedef struct {
    double build_ms;        /* device time of the whole build (HIP events) */
    double solve_ms;        /* device time of the column kernels only */
    long long nnz;          /* nnz of M */
    int ncols;
    int n_fallback;         /* mixed path: columns recomputed in FP64 */
    int n_overflow;         /* columns that outgrew their size bin and moved on */
    int n_cache_hits;       /* least-squares solves served by the pattern->R cache */
    double first_bin_ms;    /* device time of the first column-kernel launch (the first size bin,
                               where routing sends every column that fits it) */
} mspai_stats;

/* Context on HIP device `device` (replaces the MPI world of Com_Server). */
int mspai_ctx_create(int device, mspai_ctx** ctx);
int mspai_ctx_destroy(mspai_ctx* ctx);
const char* mspai_last_error(const mspai_ctx* ctx);

/*
 * Upload C (nrows x ncols CSC, row indices sorted per column) and its row
 * pattern (CSR: rowptr[nrows+1], colind sorted per row; NULL = derived on
 * the host).  Replaces Compressed_Lines (src/Compressed_Lines.h:52-101):
 * col_idcs/A/len_cols become rowind/val/colptr, row_idcs/len_rows become
 * colind/rowptr.  nrows > ncols holds probing rows (src/Read_mm_Matrix.cpp:532-667).
 * When the row pattern holds exactly C's entries (always so when derived)
 * and no row has more than 64, the candidate scores run over the rows of I
 * with C's values in row order (nnz doubles of device memory); otherwise
 * they run column-wise, with the same results.
 */
int mspai_upload_csc(mspai_ctx* ctx, int nrows, int ncols, const int* colptr, const int* rowind,
                     const double* val, const int* rowptr, const int* colind);

/* Replace only the values of the uploaded matrix (same pattern). */
int mspai_update_values(mspai_ctx* ctx, const double* val);

/* Start pattern J0 per column (Pattern::j_sets, src/Pattern.h:69-120,
 * Com_Server::Get_P_Col src/Com_Server.imp:445-470).  NULL pointers select
 * the diagonal pattern (-pp 1, src/Pattern.cpp:378-423). */
int mspai_set_pattern(mspai_ctx* ctx, const int* colptr, const int* rowind);

/* Upper pattern UP (-up, ncols columns, CSC pattern with sorted distinct
 * indices below ncols): each augmentation step of column k keeps only the
 * candidates in UP(:,k) (J~ = J~ n UP(:,k), Index_Set::Set_Intersection in
 * Augment_Sparsity, src/Spai_Sub.imp:1072-1094).  NULL pointers remove it.
 * Scalar SPAI only; a re-upload with another shape drops it. */
int mspai_set_upper_pattern(mspai_ctx* ctx, const int* colptr, const int* rowind);

/* Start pattern = the pattern of the last M built in this context: the
 * P_Memory reuse of repeated setups (P_Memory = M_REAL->To_Pattern,
 * src/imspai.cpp:850-854; between pattern recomputations the reference
 * re-solves on it with maxnew = max_impr_steps = 0, src/imspai.cpp:663-677).
 * Combine with mspai_update_values and a build with max_impr_steps = 0. */
int mspai_reuse_pattern(mspai_ctx* ctx);

/* Target B (nrows x ncols CSC) for MSPAI probing (Create_bk_Hat,
 * src/Spai_Sub.imp:735-775); colptr == NULL restores the identity. */
int mspai_set_target(mspai_ctx* ctx, const int* colptr, const int* rowind, const double* val);

/* Build M (Spai<T>::SPAI_Algorithm, src/Spai.imp:33-191).  M stays on the
 * device until downloaded.  `stats` may be NULL. */
int mspai_build(mspai_ctx* ctx, const mspai_opts* opts, mspai_stats* stats);

/*
 * mspai_build without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The inputs are those already in the context (matrix, start pattern --
 * set beforehand, since setting one allocates --, target); M is written to
 * the caller's device arrays d_colptr[ncols + 1], d_rowind[max_nnz] and
 * d_val[max_nnz], which must hold the worst case max_nnz >= ncols * (max |J0|
 * + maxnew * max_impr_steps), the capacity Init_Preconditioner reserves.
 * Everything else the build needs (queues, counters, output slots, per-column
 * results, working sets) lives in a static device array of
 * MSPAI_NOALLOC_SCRATCH_MIB MiB that each device holds until the process
 * ends; a problem that needs more fails with MSPAI_ERR_CAPACITY before any
 * launch (about 130 MiB for the random band n = 262,144 with 8 steps of 4).
 * FP64 or mixed, any binning; not with the cache (-cs), block SPAI or work
 * counters (MSPAI_ERR_UNSUPPORTED).  M is not kept in the context
 * (mspai_download_* and mspai_reuse_pattern do not see it).  Calls that
 * overlap on one GPU share the array and race.  Returns when M is written.
 */
#define MSPAI_NOALLOC_SCRATCH_MIB 256
int mspai_build_noalloc(mspai_ctx* ctx, const mspai_opts* opts, int* d_colptr, int* d_rowind, double* d_val,
                        long long max_nnz, mspai_stats* stats);

/* nnz of the last built M. */
int mspai_result_nnz(const mspai_ctx* ctx, long long* nnz);

/* Copy M out as CSC: colptr[ncols+1], rowind[nnz] (sorted J per column),
 * val[nnz] (Insert_Row_Solution, src/Com_Server.cpp:415-457). */
int mspai_download_csc(mspai_ctx* ctx, int* colptr, int* rowind, double* val);

/* Per-column results: final residual norm ||A(I,J) m - b(I)||, number of
 * least-squares solves, status bits, and 8 work counters per column (only
 * after a build with opts->counters set; else MSPAI_ERR_STATE when `work`
 * is not NULL).  Any pointer may be NULL. */
int mspai_download_column_info(mspai_ctx* ctx, double* res_norm, int* nsolve, int* status,
                               double* work);

/*
 * Device-pointer entry points for callers that keep their arrays on the GPU
 * (the PyTorch-ROCm test harness passes tensor storage).  No counterpart in
 * the reference, which has no device boundary.
 */
/* Replace C's values (nnz doubles, same pattern) from device memory. */
int mspai_update_values_device(mspai_ctx* ctx, const double* d_val);
/* Replace the target's values (same pattern) from device memory. */
int mspai_update_target_device(mspai_ctx* ctx, const double* d_val);
/* Copy the values of the last built M (nnz doubles, CSC order) to device memory. */
int mspai_result_values_device(mspai_ctx* ctx, double* d_out);

/*
 * Adjoint of the last build with its pattern held fixed: given
 * d_gM = dL/dM (device, nnz(M) doubles in M's CSC order) write
 * d_gC = dL/dC (device, nnz(C) doubles in C's CSC order) and, when a target
 * is set and d_gB is not NULL, d_gB = dL/dB (nnz(B) doubles).  Per column:
 * r = b - C(I,J) m, R^T R lambda = g, dL/dC(i,j) = r_i lambda_j -
 * (C lambda)_i m_j, dL/db = C lambda, with the QR recomputed (FP64, or FP32
 * with refinement when opts->precision is 1).  No counterpart in the
 * reference (it has no derivatives); MSPAI probing weights rho enter through
 * C's and B's probe rows, so dL/drho follows from gC and gB by the chain rule.
 * Accumulation into gC uses atomics, so its last bits depend on timing.
 */
int mspai_backward(mspai_ctx* ctx, const mspai_opts* opts, const double* d_gM, double* d_gC, double* d_gB,
                   mspai_stats* stats);

/*
 * Autotuner of the size bins (no counterpart in the reference): builds the
 * uploaded problem with each candidate bin table (lane-group widths, |J| and
 * |I| edges, resident-wave caps), `repeats` timed runs each after one warm-up,
 * and keeps the fastest for later builds with opts->binning == 2.  Results do
 * not depend on the table.  A new mspai_upload_csc clears the choice.
 */
#define MSPAI_MAX_TUNE 32
typedef struct {
    int ncandidates;
    int chosen;
    char chosen_name[32];
    double candidate_ms[MSPAI_MAX_TUNE];  /* best column-kernel time per table (-1: not tried) */
} mspai_tune_report;
int mspai_autotune(mspai_ctx* ctx, const mspai_opts* opts, int repeats, mspai_tune_report* report);

/* Sets the bin table that builds of `precision` (0 FP64, 1 mixed) with
 * opts->binning == 2 use, as the autotuner would (no counterpart in the
 * reference: for tests and for callers that tune once and store the choice).
 * `name` (32 bytes, may be NULL) receives the table's name, `ntables` (may be
 * NULL) the number of tables; table < 0 only reports that number.  Mixed
 * builds need a table with size bins (MSPAI_ERR_UNSUPPORTED otherwise). */
int mspai_set_bin_table(mspai_ctx* ctx, int precision, int table, char* name, int* ntables);

/* Sums over all columns of the 8 work counters of the last build (which
 * must have run with opts->counters set; else MSPAI_ERR_STATE):
 * QR flop, other solve flop (Q^T b, triangular solve, residual), gather
 * bytes, candidate-union bytes, rho bytes, candidates scored, sum |I|,
 * sum |J| (over all least-squares solves). */
int mspai_work_totals(mspai_ctx* ctx, double* totals);

/*
 * Per-launch profile of the last build or adjoint (no counterpart in the
 * reference): for every column-kernel launch in order, its name (precision,
 * size bin, lane-group width and |J|/|I| capacities, or the generic kernel),
 * its device time from HIP events on the library's stream, and -- after a
 * build with opts->counters set, else -1 -- the number of columns it
 * finished, the sums of their 8 work counters (as mspai_work_totals) and of
 * their near-margin decisions: FP64 selection decisions whose margin is at
 * most 1e-12 relative, so that another operation order could decide them
 * the other way (near[0]: a score against the mean threshold; near[1]: the
 * cut after the maxnew-th score, ties across it included; near[2]: ||r||
 * against epsilon), and the nonzeros of M they wrote.  A column that moves
 * on to a later launch is counted there only.  Writes min(max, launches)
 * entries and the count to *nlaunch.
 */
typedef struct {
    char name[48];
    double ms;
    double columns;
    double work[8];
    double near[3];
    double nnz;
} mspai_launch_stat;
int mspai_launch_stats(mspai_ctx* ctx, int max, int* nlaunch, mspai_launch_stat* out);

/* Size-bin routing of the last build: nbins, and for b = 0..nbins the
 * columns routed to bin b up front and the columns it processed (routed plus
 * overflow from bin b-1); bin nbins is the generic kernel.  Arrays hold
 * nbins + 1 entries (at most 8). */
int mspai_bin_counts(const mspai_ctx* ctx, int* nbins, int* routed, int* queued);

/* Sparse product on device vectors: y = op(X) x with X = A (which = 0, the
 * uploaded nrows x ncols matrix) or X = M (which = 1, the last built
 * preconditioner), op = transpose when `transpose` is nonzero.  This is the
 * preconditioner application of the Krylov solve (PC_MSPAI::Apply_MSPAI =
 * MatMult(PM, x, y), src/imspai.cpp:47-53) and the operator product beside
 * it.  Every y_i sums its terms in increasing index order with separate
 * multiply and add.  `stream` (a hipStream_t, or NULL) orders the product
 * after the caller's work on that stream; with NULL the call runs on the
 * context's stream and returns when y is written.  The first plain product
 * after A or M changed builds a CSR copy on the context's stream, ordered
 * after the caller's stream by an event.  Calls that change A or M
 * (upload, update_values, a build) must not overlap a product still in
 * flight on another stream: synchronize that stream first. */
int mspai_apply(mspai_ctx* ctx, int which, int transpose, const double* d_x, double* d_y, void* stream);


/*
 * Block SPAI (-bs, src/imspai.cpp:208-210, 746-760; Spai_Unrestrained_Block,
 * src/Spai_Unrestrained_Block.imp:33-205).
 *
 * mspai_set_block_size converts the uploaded (square) matrix into a block
 * matrix (Matrix<T>::Convert_Block_Matrix, src/Matrix.imp:1280-1516): bs > 1
 * constant blocks of bs columns, the last holding the remainder
 * (Find_Constant_Blocks); bs = 0 variable blocks from the runs of consecutive
 * rows at each diagonal (Find_Diagonal_Blocks, upper limit 1000000); bs = 1
 * returns to scalar SPAI.  In block mode the start pattern is over block
 * columns (mspai_set_pattern with nblocks columns; NULL = block diagonal, the
 * pattern the reference's -pp 1 gives the block matrix), mspai_build runs
 * block SPAI in FP64 against the identity, and M comes out as the scalar
 * matrix of Scalar_Matrix (src/Matrix.imp, exact zeros dropped): download,
 * apply and reuse_pattern (the block pattern) work as in scalar mode;
 * mspai_download_column_info fills the first nblocks entries (per block
 * column).  Blocks hold at most 32 columns.  The adjoint, the autotuner, QR
 * updates, the cache, mixed precision and targets are scalar-only.
 */
int mspai_set_block_size(mspai_ctx* ctx, int bs);

/* Block structure of the current block matrix: the number of blocks, stored
 * blocks, block sizes (nblocks), and the block pattern as CSC
 * (colptr[nblocks+1], rowind[nnzb]).  Any output may be NULL.  Returns
 * MSPAI_ERR_STATE outside block mode. */
int mspai_block_info(const mspai_ctx* ctx, int* nblocks, long long* nnzb, int* block_sizes, int* colptr,
                     int* rowind);

/* Block partition of an n x n CSC matrix without a context (host only):
 * Find_Constant_Blocks for bs > 0, Find_Diagonal_Blocks for bs = 0
 * (src/Matrix.imp:1345-1514).  sizes needs room for n entries. */
int mspai_block_sizes(int n, const int* colptr, const int* rowind, int bs, int* sizes, int* nblocks);

#ifdef __cplusplus
}
#endif
#endif
