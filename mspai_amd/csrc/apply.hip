/*
 * Sparse products with the uploaded matrix A and the built preconditioner M
 * on the device: the preconditioner application of the Krylov solve
 * (PC_MSPAI::Apply_MSPAI = MatMult(PM, x, y), src/imspai.cpp:47-53) and the
 * operator product around it.
 *
 * A and M are held in CSC.  y = op^T x reads the columns directly (one
 * sequential dot product per column).  y = op x uses a CSR copy built on the
 * device on first use: row counts, a scan, a fill, then every row sorted by
 * column index, so that each dot product sums its terms in increasing
 * column order -- the order of a column-by-column CSC product on the host.
 * Terms are multiplied and added separately (no fused multiply-add).
 */
#include <hip/hip_runtime.h>

#include "spai_host.h"

namespace mspai {
namespace {

/* counts[r] += 1 for every stored entry of row r (counts zeroed). */
__global__ void row_counts_kernel(const int* rowind, long long nnz, int* counts)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) atomicAdd(counts + rowind[e], 1);
}

/* Scatter the entries of column j into their rows (any order within a row;
 * cursor zeroed). */
__global__ void csr_fill_kernel(const int* colptr, const int* rowind, const double* val, int ncols,
                                const int* rowptr, int* cursor, int* colind, double* cval)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncols) return;
    for (int e = colptr[j]; e < colptr[j + 1]; ++e) {
        const int r = rowind[e];
        const int at = rowptr[r] + atomicAdd(cursor + r, 1);
        colind[at] = j;
        cval[at] = val[e];
    }
}

/* Insertion sort of every row by column index (rows are short). */
__global__ void csr_sort_rows_kernel(int nrows, const int* rowptr, int* colind, double* cval)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int b = rowptr[r], e = rowptr[r + 1];
    for (int i = b + 1; i < e; ++i) {
        const int c = colind[i];
        const double v = cval[i];
        int k = i - 1;
        while (k >= b && colind[k] > c) {
            colind[k + 1] = colind[k];
            cval[k + 1] = cval[k];
            --k;
        }
        colind[k + 1] = c;
        cval[k + 1] = v;
    }
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

void launch_row_counts(const int* rowind, long long nnz, int* counts, hipStream_t stream)
{
    if (nnz <= 0) return;
    hipLaunchKernelGGL(row_counts_kernel, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, stream, rowind, nnz,
                       counts);
}

void launch_csr_fill(const int* colptr, const int* rowind, const double* val, int ncols, const int* rowptr,
                     int* cursor, int* colind, double* cval, hipStream_t stream)
{
    if (ncols <= 0) return;
    hipLaunchKernelGGL(csr_fill_kernel, dim3((ncols + 255) / 256), dim3(256), 0, stream, colptr, rowind, val, ncols,
                       rowptr, cursor, colind, cval);
}

void launch_csr_sort_rows(int nrows, const int* rowptr, int* colind, double* cval, hipStream_t stream)
{
    if (nrows <= 0) return;
    hipLaunchKernelGGL(csr_sort_rows_kernel, dim3((nrows + 255) / 256), dim3(256), 0, stream, nrows, rowptr, colind,
                       cval);
}

void launch_line_dot(int nlines, const int* ptr, const int* ind, const double* val, const double* x, double* y,
                     hipStream_t stream)
{
    if (nlines <= 0) return;
    hipLaunchKernelGGL(line_dot_kernel, dim3((nlines + kLineBlock - 1) / kLineBlock), dim3(kLineBlock), 0, stream,
                       nlines, ptr, ind, val, x, y);
}

}  // namespace mspai
