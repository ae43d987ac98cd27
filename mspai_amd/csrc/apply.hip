/*
 * Sparse products with the uploaded matrix A and the built preconditioner M
 * on the device: the preconditioner application of the Krylov solve
 * (PC_MSPAI::Apply_MSPAI = MatMult(PM, x, y), src/imspai.cpp:47-53) and the
 * operator product around it.
 *
 * A and M are held in CSC.  y = op^T x reads the columns directly (one
 * sequential dot product per column).  y = op x uses a CSR copy built on the
 * device on first use: a stable radix sort of the entries by row -- taken
 * in CSC order, so column order within a row -- gives each row's entries in
 * increasing column order, the order of a column-by-column CSC product on
 * the host, and the sorted rows give the row pointers.
 * Terms are multiplied and added separately (no fused multiply-add).
 */
#include <hip/hip_runtime.h>

#include "spai_host.h"

namespace mspai {
namespace {

/* Entry e of column j: col_of[e] = j, order[e] = e (the payload of the
 * stable sort of the entries by row). */
__global__ void csr_entry_cols_kernel(const int* colptr, int ncols, int* col_of, int* order)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncols) return;
    for (int e = colptr[j]; e < colptr[j + 1]; ++e) {
        col_of[e] = j;
        order[e] = e;
    }
}

/* Row pointers from the rows of the entries sorted by row: entry k (and the
 * end, k = nnz) starts every row after the previous entry's row up to its
 * own, so each of rowptr[0..nrows] is written once; empty rows get the
 * start of the next stored row. */
__global__ void csr_row_starts_kernel(long long nnz, int nrows, const unsigned* rows_sorted, int* rowptr)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nnz) return;
    const int prev = k == 0 ? -1 : (int)rows_sorted[k - 1];
    const int cur = k == nnz ? nrows : (int)rows_sorted[k];
    for (int r = prev + 1; r <= cur; ++r) rowptr[r] = (int)k;
}

/* CSR entry k is CSC entry perm[k]: its column and its value. */
__global__ void csr_gather_kernel(long long nnz, const int* perm, const int* col_of, const double* val,
                                  int* colind, double* cval)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int e = perm[k];
    colind[k] = col_of[e];
    cval[k] = val[e];
}

/* y_i = sum over the stored entries of line i (row of a CSR or column of a
 * CSC array) of val * x[index], in storage order.
 *
 * A block owns kLineBlock consecutive lines, whose entries are one contiguous
 * range of ind/val.  The range is read in chunks of kStage entries with
 * consecutive threads on consecutive entries (coalesced), each product
 * val * x[index] going to LDS; each thread then adds the products of its own
 * line that fall in the chunk, in storage order.  Chunks are taken in order,
 * so every line's sum is the same sequence of roundings as a plain loop over
 * its entries. */
constexpr int kLineBlock = 256;
constexpr int kStage = 2048;

__global__ __launch_bounds__(kLineBlock) void line_dot_kernel(int nlines, const int* ptr, const int* ind,
                                                              const double* val, const double* x, double* y)
{
    __shared__ double prod[kStage];
    const int i0 = blockIdx.x * kLineBlock;
    const int i = i0 + (int)threadIdx.x;
    const int iend = min(i0 + kLineBlock, nlines);
    const long long first = ptr[i0], last = ptr[iend];
    const long long b = i < nlines ? ptr[i] : last;
    const long long e = i < nlines ? ptr[i + 1] : last;
    double sum = 0.0;
    for (long long c0 = first; c0 < last; c0 += kStage) {
        const long long c1 = min(c0 + kStage, last);
        for (long long k = c0 + threadIdx.x; k < c1; k += kLineBlock) prod[k - c0] = __dmul_rn(val[k], x[ind[k]]);
        __syncthreads();
        const long long hi = min(e, c1);
        for (long long k = max(b, c0); k < hi; ++k) sum = __dadd_rn(sum, prod[k - c0]);
        __syncthreads();
    }
    if (i < nlines) y[i] = sum;
}

__global__ void fill_int_kernel(int* p, long long count, int value)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

}  // namespace

void launch_fill_int(int* p, long long count, int value, hipStream_t stream)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, p, count, value);
}

void launch_csr_entry_cols(const int* colptr, int ncols, int* col_of, int* order, hipStream_t stream)
{
    if (ncols <= 0) return;
    hipLaunchKernelGGL(csr_entry_cols_kernel, dim3((ncols + 255) / 256), dim3(256), 0, stream, colptr, ncols, col_of,
                       order);
}

void launch_csr_row_starts(long long nnz, int nrows, const unsigned* rows_sorted, int* rowptr, hipStream_t stream)
{
    hipLaunchKernelGGL(csr_row_starts_kernel, dim3((unsigned)((nnz + 1 + 255) / 256)), dim3(256), 0, stream, nnz,
                       nrows, rows_sorted, rowptr);
}

void launch_csr_gather(long long nnz, const int* perm, const int* col_of, const double* val, int* colind,
                       double* cval, hipStream_t stream)
{
    if (nnz <= 0) return;
    hipLaunchKernelGGL(csr_gather_kernel, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, stream, nnz, perm,
                       col_of, val, colind, cval);
}

void launch_line_dot(int nlines, const int* ptr, const int* ind, const double* val, const double* x, double* y,
                     hipStream_t stream)
{
    if (nlines <= 0) return;
    hipLaunchKernelGGL(line_dot_kernel, dim3((nlines + kLineBlock - 1) / kLineBlock), dim3(kLineBlock), 0, stream,
                       nlines, ptr, ind, val, x, y);
}

}  // namespace mspai
