/*
 * Host-side launchers of the column kernels (one per kernel family).
 */
#ifndef MSPAI_SPAI_HOST_H
#define MSPAI_SPAI_HOST_H

#include <hip/hip_runtime.h>

#include "device_common.h"

namespace mspai {

/* Column kernel with the working set in LDS: one column per group of
 * `group` lanes (16, 32 or 64), L.per_wave bytes of LDS per group; columns
 * that do not fit go to a.overflow.  mode 0: FP64 build; 1: FP32 QR + FP64
 * refinement, columns whose selection margins are too small go to
 * a.fallback; 2/3: the adjoint in FP64 / mixed; 4: FP64 build with the
 * dense QR update (-qr 2). */
void launch_binned(const BuildArgs& a, const ScratchLayout& L, int group, int mode, const int* queue,
                   const int* nqueue, int* head, int nwaves, hipStream_t stream, int wcap = 0);

/* Workgroups (one wave each) of that kernel that fit on one CU at once, by
 * its registers and LDS (0 if the runtime cannot tell). */
int binned_blocks_per_cu(const ScratchLayout& L, int group, int mode, int wcap = 0);

/* Route columns to bin queues: queue b < nbins at queues + b*n with its
 * length in counts[b]; the rest to gqueue (length gcount).  routed[0..nbins]
 * receives the lengths once routing is complete (counts[0..nbins), then
 * *gcount); *done must be 0 at launch. */
void launch_classify(const BuildArgs& a, int nbins, const BinEdges& edges, int* queues, int* counts,
                     int* gqueue, int* gcount, int n, int* done, int* routed, hipStream_t stream);

/* dst[0..count) = src[0..count) on the stream (count <= 64). */
void launch_copy_ints(const int* src, int* dst, int count, hipStream_t stream);

/* FP64 kernel (mode 0 build, 2 adjoint, 4 QR-update build) with the working
 * set in HBM (a.scratch, L.per_wave per wave). */
void launch_fp64_generic(const BuildArgs& a, const ScratchLayout& L, int mode, const int* queue,
                         const int* nqueue, int* head, int nwaves, hipStream_t stream);

/* Sparse products (apply.hip). */
void launch_fill_int(int* p, long long count, int value, hipStream_t stream);
/* col_of[e] = column of CSC entry e, order[e] = e. */
void launch_csr_entry_cols(const int* colptr, int ncols, int* col_of, int* order, hipStream_t stream);
/* rowptr[0..nrows] from the row indices of the entries sorted by row. */
void launch_csr_row_starts(long long nnz, int nrows, const unsigned* rows_sorted, int* rowptr, hipStream_t stream);
/* colind[k] = col_of[perm[k]], cval[k] = val[perm[k]]. */
void launch_csr_gather(long long nnz, const int* perm, const int* col_of, const double* val, int* colind,
                       double* cval, hipStream_t stream);
/* y[i] = sum over line i of val * x[ind] in storage order (mul, then add). */
void launch_line_dot(int nlines, const int* ptr, const int* ind, const double* val, const double* x, double* y,
                     hipStream_t stream);

/* Block SPAI (spai_block.hip). */
/* (A_j^T A_j)^-1 of every block column into a.sqinv (tmp: sum bsize^2 doubles). */
void launch_block_sq_inverse(const BlockArgs& a, double* tmp, hipStream_t stream);
/* Persistent block-column kernel on nwaves waves, L.per_wave bytes each at a.scratch. */
void launch_block_columns(const BlockArgs& a, const BlockLayout& L, int nwaves, hipStream_t stream);
/* Scalar_Matrix: per-scalar-column nonzero counts, then the CSC fill. */
void launch_block_scalar_count(const BlockArgs& a, int* counts, hipStream_t stream);
void launch_block_scalar_fill(const BlockArgs& a, const int* colptr, int* rowind, double* val,
                              hipStream_t stream);

}  // namespace mspai

#endif
