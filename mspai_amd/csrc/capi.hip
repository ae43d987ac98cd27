/*
 * Host side of the mspai_gpu C ABI (include/mspai_gpu.h).
 *
 * One context owns one GPU-resident copy of C (CSC + row pattern), the start
 * pattern, the optional target B, the output slots of M and the per-wave
 * scratch.  mspai_build is the device form of Spai<T>::SPAI_Algorithm
 * (src/Spai.imp:33-191): the MPI column distribution (Distribution,
 * Load_Balance, Com_Server) becomes a persistent work queue over all columns
 * in one launch per size bin, and Insert_Row_Solution becomes a compaction of
 * the per-column slots into CSC.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mspai_gpu.h"
#include "device_common.h"
#include "spai_host.h"

using namespace mspai;

constexpr int kMaxBins = 4;
constexpr int kMaxLaunch = 2 * kMaxBins + 2;  // column-kernel launches of one build: mixed bins, FP64 bins, generic

/* Device scratch of mspai_build_noalloc (queues, counters, output slots,
 * per-column results, tie buffers, the scan workspace and the generic
 * kernel's working sets of one build), a static array so that the call
 * allocates nothing.  Each device the library runs on holds its own copy
 * until the process ends.  In host code `scratch` only names the symbol:
 * its device address comes from hipGetSymbolAddress, once per device (at
 * mspai_ctx_create), and that address is what the kernels receive.  Calls
 * on one device share the array: calls that overlap on one GPU race. */
constexpr size_t kNoAllocBytes = (size_t)MSPAI_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) char scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static char* g_noalloc_ptr[kMaxDevices] = {};

namespace {

template <class T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

__global__ void column_norms_kernel(DevCsc A, double* ajn)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.ncols) return;
    double s = 0.0;
    for (int e = A.colptr[j]; e < A.colptr[j + 1]; ++e) {
        double v = A.val[e];
        s = __dadd_rn(s, __dmul_rn(v, v));
    }
    ajn[j] = sqrt(s);
}

/* C's values in the order of its row pattern (which holds exactly C's
 * entries, sorted per row): each stored entry finds its column in its row. */
__global__ void row_values_kernel(DevCsc A, const int* rp_ptr, const int* rp_ind, double* rp_val)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.ncols) return;
    for (int e = A.colptr[j]; e < A.colptr[j + 1]; ++e) {
        const int i = A.rowind[e];
        int lo = rp_ptr[i], hi = rp_ptr[i + 1];
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (rp_ind[mid] < j) lo = mid + 1;
            else hi = mid;
        }
        if (lo < rp_ptr[i + 1] && rp_ind[lo] == j) rp_val[lo] = A.val[e];
    }
}

/* Sum the per-column work counters (one block, fixed order). */
__global__ void work_sum_kernel(const double* work, int n, double* out)
{
    __shared__ double part[256];
    int q = blockIdx.x;
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) s += work[(size_t)j * kNumWork + q];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = part[0];
}


/* M's CSC from the per-column outputs: colptr[1..ncols] holds the inclusive
 * scan of the lengths; this kernel also sets colptr[0] = 0 and *nnz, and
 * counts the columns the FP64 kernels built in a mixed build (*nfp64). */
/* Output slots -> CSC: kScatterLanes lanes per column (most columns hold
 * fewer entries), colptr from the scan of the lengths. */
constexpr int kScatterLanes = 16;

__global__ void scatter_columns_kernel(int ncols, int cap, const int* m_len, const int* m_idx,
                                       const double* m_val, int* colptr, int* rowind, double* val, int* nnz)
{
    const int col = blockIdx.x * (blockDim.x / kScatterLanes) + (int)(threadIdx.x / kScatterLanes);
    const int sub = threadIdx.x % kScatterLanes;
    if (col >= ncols) return;
    const int len = m_len[col], o = col == 0 ? 0 : colptr[col];
    if (sub == 0 && col == 0) colptr[0] = 0;
    if (sub == 0 && col == ncols - 1) *nnz = o + len;
    const size_t src = (size_t)col * cap;
    for (int i = sub; i < len; i += kScatterLanes) {
        rowind[o + i] = m_idx[src + i];
        val[o + i] = m_val[src + i];
    }
}

/* *out += the columns whose status has `bit`: a grid-stride count, one
 * atomic per block (a per-column atomic on one address serializes). */
__global__ __launch_bounds__(256) void count_status_kernel(int n, const int* status, int bit, int* out)
{
    __shared__ int part[256 / 64];
    int c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        c += (status[i] & bit) != 0;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < 256 / 64; ++w) t += part[w];
        if (t) atomicAdd(out, t);
    }
}

}  // namespace

struct mspai_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cu = 0;
    size_t lds_block = 64 * 1024;  // LDS one workgroup may hold (sharedMemPerBlock)
    // matrix C
    int nrows = 0, ncols = 0;
    long long nnz = 0;
    int maxcolnnz = 0, maxrownnz = 0;
    int *d_colptr = nullptr, *d_rowind = nullptr;
    double* d_val = nullptr;
    int *d_rp_ptr = nullptr, *d_rp_ind = nullptr;
    double* d_ajn = nullptr;
    double* d_rp_val = nullptr;  // C's values in row-pattern order (filled with the column norms)
    bool rp_match = false;       // the row pattern is C's own (then d_rp_val is allocated)
    // start pattern
    int *d_p_ptr = nullptr, *d_p_ind = nullptr;
    int plen_max = 0;
    int p_cols = 0;             // columns the start pattern was set for
    int *d_up_ptr = nullptr, *d_up_ind = nullptr;  // upper pattern (-up), CSC; nullptr: none
    int up_cols = 0;
    bool pattern_set = false;   // set by the caller (the implicit diagonal of a build is not)
    // target
    int *d_b_ptr = nullptr, *d_b_ind = nullptr;
    double* d_b_val = nullptr;
    int b_rows = 0, b_cols = 0;  // shape the target was set for
    // outputs
    int cap = 0;
    size_t out_cols = 0;
    int *d_m_len = nullptr, *d_m_idx = nullptr;
    double* d_m_val = nullptr;
    double* d_res = nullptr;
    int *d_nsolve = nullptr, *d_status = nullptr;
    double* d_work = nullptr;
    int *d_Mptr = nullptr, *d_Mind = nullptr;
    double* d_Mval = nullptr;
    long long nnzM = -1;        // nnz of the current M (-1: none, or its values are stale)
    long long pattern_nnz = -1;  // nnz of the last built M's pattern (kept across value updates)
    size_t Mcap = 0;
    // queues and scratch
    int* d_queues = nullptr;   // (kMaxBins + 1) queues of ncols entries
    int* d_counter = nullptr;  // FP64 queues: [0..8) lengths, [8..16) heads; [16..24) routed lengths;
                               // mixed queues: [24..32) lengths, [32..40) heads
    int* h_counter = nullptr;  // pinned host copy of d_counter, [63] = nnz(M) of the last build
    double* d_wsum = nullptr;  // work-counter totals
    char* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    char* d_tie = nullptr;
    size_t tie_bytes = 0;
    void* d_cub = nullptr;
    size_t cub_bytes = 0;
    // CSR copies for mspai_apply (built on first use; valid while *_csr_ok)
    int *d_acsr_ptr = nullptr, *d_acsr_ind = nullptr, *d_mcsr_ptr = nullptr, *d_mcsr_ind = nullptr;
    double *d_acsr_val = nullptr, *d_mcsr_val = nullptr;
    int acsr_rows_cap = 0, mcsr_rows_cap = 0;
    long long acsr_nnz_cap = 0, mcsr_nnz_cap = 0;
    int* d_csr_tmp = nullptr;  // 4 x csr_tmp_cap ints: columns, order, sorted rows, permutation
    long long csr_tmp_cap = 0;
    bool a_csr_ok = false, m_csr_ok = false;
    bool ajn_ok = false;  // d_ajn holds the column norms of the current values
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // [4..5]: first size bin
    hipEvent_t csr_ev = nullptr;  // orders a CSR copy on the context's stream with the caller's stream
    /* per-launch profile of the last build: events around every column-kernel
     * launch, and (builds with counters) the totals of the columns each finished */
    hipEvent_t lev[2 * kMaxLaunch] = {};
    int nlaunch = 0;
    char lname[kMaxLaunch][48] = {};
    double* d_lwork = nullptr;  // kMaxLaunch x kLaunchStat
    bool lwork_valid = false;
    /* mspai_build_noalloc: the build's buffers point into the static scratch
     * array for the duration of the call, and every path that would
     * allocate fails with MSPAI_ERR_CAPACITY instead */
    bool no_alloc = false;
    char* d_noalloc = nullptr;  // this device's address of the static scratch array
    /* pattern->R cache */
    int cache_slots = 0, cache_capI = 0, cache_capJ = 0;
    int* d_cache_state = nullptr;
    uint32_t* d_cache_key = nullptr;
    int* d_cache_m = nullptr;
    int* d_cache_n = nullptr;
    int* d_cache_status = nullptr;
    double* d_cache_data = nullptr;
    int tuned_table[2] = {-1, -1};  // autotuner's choice for this matrix (binning == 2), FP64 and mixed
    int last_table = 0;
    int last_nbins = 0;
    bool last_mixed = false;
    int last_binning = 1;
    bool last_counters = false;  // the last build stored per-column work counters
    long long nnzB = 0;
    int last_routed[8] = {0};
    int mixed_fp64_last = -1;  // FP64 queue length of the last mixed build (-1: none since the inputs changed)
    int last_queued[8] = {0};
    /* block SPAI (-bs): 1 = scalar, > 1 constant blocks, 0 = variable blocks */
    int block_bs = 1;
    int nb = 0, maxbs = 1, bcmax = 0, brmax = 0;
    long long nnzb = 0;
    std::vector<int> h_bsize, h_bcolptr, h_browind;
    int *d_bsize = nullptr, *d_bstart = nullptr, *d_bcolptr = nullptr, *d_browind = nullptr;
    int *d_brp_ptr = nullptr, *d_brp_ind = nullptr;
    long long *d_bvoff = nullptr, *d_sqoff = nullptr;
    double *d_bval = nullptr, *d_sqinv = nullptr, *d_sqtmp = nullptr;
    int *d_bm_len = nullptr, *d_bm_idx = nullptr;
    double* d_bm_val = nullptr;
    int bcapJ = 0;
    long long bcapX = 0;
    bool block_built = false;
    char* d_bscratch = nullptr;
    size_t bscratch_bytes = 0;
};

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);                        \
            return MSPAI_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

static int fail(mspai_ctx* ctx, int code, const std::string& msg)
{
    ctx->err = msg;
    return code;
}

template <class T>
static int upload(mspai_ctx* ctx, T*& d, const T* h, size_t n)
{
    if (ctx->no_alloc) return fail(ctx, MSPAI_ERR_CAPACITY, "this call allocates no device memory");
    dfree(d);
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void**)&d, n * sizeof(T)));
    if (h) HIPCHK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return MSPAI_OK;
}

/* Drop the block matrix and the block results (back to scalar SPAI). */
static void block_free(mspai_ctx* ctx)
{
    dfree(ctx->d_bsize);
    dfree(ctx->d_bstart);
    dfree(ctx->d_bcolptr);
    dfree(ctx->d_browind);
    dfree(ctx->d_brp_ptr);
    dfree(ctx->d_brp_ind);
    dfree(ctx->d_bvoff);
    dfree(ctx->d_sqoff);
    dfree(ctx->d_bval);
    dfree(ctx->d_sqinv);
    dfree(ctx->d_sqtmp);
    dfree(ctx->d_bm_len);
    dfree(ctx->d_bm_idx);
    dfree(ctx->d_bm_val);
    dfree(ctx->d_bscratch);
    ctx->bscratch_bytes = 0;
    ctx->bcapJ = 0;
    ctx->bcapX = 0;
    ctx->block_built = false;
    ctx->block_bs = 1;
    ctx->nb = 0;
    ctx->h_bsize.clear();
    ctx->h_bcolptr.clear();
    ctx->h_browind.clear();
}

static int block_convert(mspai_ctx* ctx);

extern "C" void mspai_default_opts(mspai_opts* o)
{
    /* Command_Line::Parameters defaults, src/Command_Line.cpp:192-226 */
    o->epsilon = 0.4;
    o->maxnew = 5;
    o->max_impr_steps = 5;
    o->use_mean = 1;
    o->qr = 0;
    o->cache = 0;
    o->precision = 0;
    o->binning = 1;
    o->verbose = 0;
    o->counters = 0;
}

extern "C" int mspai_ctx_create(int device, mspai_ctx** out)
{
    if (!out) return MSPAI_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return MSPAI_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MSPAI_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return MSPAI_ERR_NO_DEVICE;
    mspai_ctx* ctx = new mspai_ctx();
    ctx->device = device;
    ctx->num_cu = prop.multiProcessorCount;
    ctx->lds_block = std::min<size_t>(std::max<size_t>(prop.sharedMemPerBlock, 64 * 1024), 160 * 1024);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return MSPAI_ERR_HIP;
    }
    for (auto& e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return MSPAI_ERR_HIP;
        }
    for (auto& e : ctx->lev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return MSPAI_ERR_HIP;
        }
    if (hipEventCreateWithFlags(&ctx->csr_ev, hipEventDisableTiming) != hipSuccess) {
        delete ctx;
        return MSPAI_ERR_HIP;
    }
    if (device < kMaxDevices) {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (!g_noalloc_ptr[device]) {
            void* p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
                delete ctx;
                return MSPAI_ERR_HIP;
            }
            g_noalloc_ptr[device] = (char*)p;
        }
        ctx->d_noalloc = g_noalloc_ptr[device];
    }
    if (hipMalloc((void**)&ctx->d_counter, 64 * sizeof(int)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_counter, 64 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&ctx->d_wsum, kNumWork * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&ctx->d_lwork, kMaxLaunch * kLaunchStat * sizeof(double)) != hipSuccess) {
        delete ctx;
        return MSPAI_ERR_HIP;
    }
    *out = ctx;
    return MSPAI_OK;
}

extern "C" int mspai_ctx_destroy(mspai_ctx* ctx)
{
    if (!ctx) return MSPAI_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    dfree(ctx->d_colptr);
    dfree(ctx->d_rowind);
    dfree(ctx->d_val);
    dfree(ctx->d_rp_ptr);
    dfree(ctx->d_rp_ind);
    dfree(ctx->d_ajn);
    dfree(ctx->d_rp_val);
    dfree(ctx->d_p_ptr);
    dfree(ctx->d_p_ind);
    dfree(ctx->d_up_ptr);
    dfree(ctx->d_up_ind);
    dfree(ctx->d_b_ptr);
    dfree(ctx->d_b_ind);
    dfree(ctx->d_b_val);
    dfree(ctx->d_m_len);
    dfree(ctx->d_m_idx);
    dfree(ctx->d_m_val);
    dfree(ctx->d_res);
    dfree(ctx->d_nsolve);
    dfree(ctx->d_status);
    dfree(ctx->d_work);
    dfree(ctx->d_Mptr);
    dfree(ctx->d_Mind);
    dfree(ctx->d_Mval);
    dfree(ctx->d_queues);
    dfree(ctx->d_counter);
    if (ctx->h_counter) (void)hipHostFree(ctx->h_counter);
    dfree(ctx->d_wsum);
    dfree(ctx->d_lwork);
    dfree(ctx->d_scratch);
    dfree(ctx->d_tie);
    dfree(ctx->d_cache_state);
    dfree(ctx->d_cache_key);
    dfree(ctx->d_cache_m);
    dfree(ctx->d_cache_n);
    dfree(ctx->d_cache_status);
    dfree(ctx->d_cache_data);
    if (ctx->d_cub) (void)hipFree(ctx->d_cub);
    dfree(ctx->d_acsr_ptr);
    dfree(ctx->d_acsr_ind);
    dfree(ctx->d_acsr_val);
    dfree(ctx->d_mcsr_ptr);
    dfree(ctx->d_mcsr_ind);
    dfree(ctx->d_mcsr_val);
    dfree(ctx->d_csr_tmp);
    block_free(ctx);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->lev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->csr_ev) (void)hipEventDestroy(ctx->csr_ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MSPAI_OK;
}

extern "C" const char* mspai_last_error(const mspai_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int mspai_upload_csc(mspai_ctx* ctx, int nrows, int ncols, const int* colptr, const int* rowind,
                                const double* val, const int* rowptr, const int* colind)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (nrows <= 0 || ncols <= 0 || nrows < ncols || !colptr || !rowind || !val)
        return fail(ctx, MSPAI_ERR_ARG, "upload: need nrows >= ncols > 0 and colptr/rowind/val");
    HIPCHK(hipSetDevice(ctx->device));
    long long nnz = colptr[ncols];
    if (colptr[0] != 0 || nnz < 0) return fail(ctx, MSPAI_ERR_ARG, "upload: colptr must start at 0");
    int maxc = 0;
    for (int j = 0; j < ncols; ++j) {
        int a = colptr[j], b = colptr[j + 1];
        if (b < a) return fail(ctx, MSPAI_ERR_ARG, "upload: colptr not monotone");
        maxc = std::max(maxc, b - a);
        for (int e = a; e < b; ++e) {
            if (rowind[e] < 0 || rowind[e] >= nrows) return fail(ctx, MSPAI_ERR_ARG, "upload: row index out of range");
            if (e > a && rowind[e] <= rowind[e - 1])
                return fail(ctx, MSPAI_ERR_ARG, "upload: row indices must be sorted and distinct per column");
        }
    }
    std::vector<int> rp, ri;
    bool rp_match = true;  // a derived pattern is C's own
    if (rowptr && colind) {
        /* a given pattern is C's own when it holds exactly C's entries,
         * sorted per row (then the scores may run over its rows) */
        std::vector<int> cnt(nrows + 1, 0);
        for (long long e = 0; e < nnz; ++e) cnt[rowind[e] + 1]++;
        for (int i = 0; i < nrows && rp_match; ++i) rp_match = rowptr[i + 1] - rowptr[i] == cnt[i + 1];
        if (rp_match && rowptr[0] == 0) {
            std::vector<int> fill(nrows);
            for (int i = 0; i < nrows; ++i) fill[i] = rowptr[i];
            for (int j = 0; j < ncols && rp_match; ++j)
                for (int e = colptr[j]; e < colptr[j + 1]; ++e)
                    if (colind[fill[rowind[e]]++] != j) {
                        rp_match = false;
                        break;
                    }
        } else {
            rp_match = false;
        }
    }
    if (!rowptr || !colind) {
        /* CSR pattern of C: the row_idcs of Compressed_Lines */
        rp.assign(nrows + 1, 0);
        for (long long e = 0; e < nnz; ++e) rp[rowind[e] + 1]++;
        for (int i = 0; i < nrows; ++i) rp[i + 1] += rp[i];
        ri.resize(nnz);
        std::vector<int> fill(rp.begin(), rp.end() - 1);
        for (int j = 0; j < ncols; ++j)
            for (int e = colptr[j]; e < colptr[j + 1]; ++e) ri[fill[rowind[e]]++] = j;
        rowptr = rp.data();
        colind = ri.data();
    }
    int maxr = 0;
    for (int i = 0; i < nrows; ++i) maxr = std::max(maxr, rowptr[i + 1] - rowptr[i]);
    if (rowptr[nrows] != nnz) return fail(ctx, MSPAI_ERR_ARG, "upload: row pattern nnz differs from colptr");
    const bool reshaped = nrows != ctx->nrows || ncols != ctx->ncols;
    int rc;
    if ((rc = upload(ctx, ctx->d_colptr, colptr, ncols + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_rowind, rowind, nnz))) return rc;
    if ((rc = upload(ctx, ctx->d_val, val, nnz))) return rc;
    if ((rc = upload(ctx, ctx->d_rp_ptr, rowptr, nrows + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_rp_ind, colind, nnz))) return rc;
    if ((rc = upload<double>(ctx, ctx->d_ajn, nullptr, ncols))) return rc;
    if (rp_match && maxr <= kWave) {
        if ((rc = upload<double>(ctx, ctx->d_rp_val, nullptr, nnz))) return rc;
    } else {
        dfree(ctx->d_rp_val);
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->rp_match = rp_match && maxr <= kWave;
    ctx->nrows = nrows;
    ctx->ncols = ncols;
    ctx->nnz = nnz;
    ctx->maxcolnnz = maxc;
    ctx->maxrownnz = maxr;
    ctx->nnzM = -1;
    ctx->mixed_fp64_last = -1;
    ctx->a_csr_ok = ctx->m_csr_ok = ctx->ajn_ok = false;
    ctx->pattern_nnz = -1;
    ctx->tuned_table[0] = ctx->tuned_table[1] = -1;
    if (ctx->block_bs != 1) {
        /* a new matrix leaves block mode; its block pattern no longer applies */
        block_free(ctx);
        ctx->pattern_set = false;
    }
    if (!ctx->pattern_set || reshaped) {
        /* a start pattern or target of another shape would be read past its end */
        dfree(ctx->d_p_ptr);
        dfree(ctx->d_p_ind);
        ctx->pattern_set = false;
        ctx->p_cols = 0;
    }
    if (reshaped) {
        dfree(ctx->d_up_ptr);
        dfree(ctx->d_up_ind);
        ctx->up_cols = 0;
        dfree(ctx->d_b_ptr);
        dfree(ctx->d_b_ind);
        dfree(ctx->d_b_val);
        ctx->nnzB = 0;
        ctx->b_rows = ctx->b_cols = 0;
    }
    return MSPAI_OK;
}

/* The start pattern and the target must have been set for the current
 * matrix (columns of C, or block columns in block mode; B with C's shape). */
static int check_shapes(mspai_ctx* ctx)
{
    const int pc = ctx->block_bs != 1 ? ctx->nb : ctx->ncols;
    if (ctx->d_p_ptr && ctx->p_cols != pc)
        return fail(ctx, MSPAI_ERR_STATE, "start pattern was set for " + std::to_string(ctx->p_cols) +
                                              " columns, the matrix has " + std::to_string(pc));
    if (ctx->d_b_ptr && (ctx->b_rows != ctx->nrows || ctx->b_cols != ctx->ncols))
        return fail(ctx, MSPAI_ERR_STATE, "target was set for another matrix shape");
    if (ctx->d_up_ptr && ctx->up_cols != ctx->ncols)
        return fail(ctx, MSPAI_ERR_STATE, "upper pattern was set for another matrix shape");
    if (ctx->d_up_ptr && ctx->block_bs != 1)
        return fail(ctx, MSPAI_ERR_ARG, "the upper pattern applies to scalar SPAI only");
    return MSPAI_OK;
}

extern "C" int mspai_update_values(mspai_ctx* ctx, const double* val)
{
    if (!ctx || !val || !ctx->d_val) return MSPAI_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(ctx->d_val, val, ctx->nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->nnzM = -1;
    ctx->mixed_fp64_last = -1;
    ctx->a_csr_ok = ctx->m_csr_ok = ctx->ajn_ok = false;
    if (ctx->block_bs != 1) return block_convert(ctx);  // Convert_Block_Matrix_Update
    return MSPAI_OK;
}

extern "C" int mspai_set_pattern(mspai_ctx* ctx, const int* colptr, const int* rowind)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (ctx->ncols <= 0) return fail(ctx, MSPAI_ERR_STATE, "set_pattern: upload the matrix first");
    HIPCHK(hipSetDevice(ctx->device));
    const int n = ctx->block_bs != 1 ? ctx->nb : ctx->ncols;  // block mode: a pattern over block columns
    std::vector<int> dp, di;
    if (!colptr || !rowind) {
        dp.resize(n + 1);
        di.resize(n);
        for (int j = 0; j <= n; ++j) dp[j] = j;
        for (int j = 0; j < n; ++j) di[j] = j;
        colptr = dp.data();
        rowind = di.data();
    }
    int pmax = 0;
    for (int j = 0; j < n; ++j) {
        int a = colptr[j], b = colptr[j + 1];
        if (b < a) return fail(ctx, MSPAI_ERR_ARG, "set_pattern: colptr not monotone");
        pmax = std::max(pmax, b - a);
        for (int e = a; e < b; ++e) {
            if (rowind[e] < 0 || rowind[e] >= n) return fail(ctx, MSPAI_ERR_ARG, "set_pattern: index out of range");
            if (e > a && rowind[e] <= rowind[e - 1])
                return fail(ctx, MSPAI_ERR_ARG, "set_pattern: indices must be sorted and distinct");
        }
    }
    int rc;
    if ((rc = upload(ctx, ctx->d_p_ptr, colptr, n + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_p_ind, rowind, (size_t)colptr[n]))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->plen_max = pmax;
    ctx->p_cols = n;
    ctx->pattern_set = true;
    return MSPAI_OK;
}

/* -up: the candidates of column k are restricted to UP(:,k)
 * (src/Spai_Sub.imp:1072-1094). */
extern "C" int mspai_set_upper_pattern(mspai_ctx* ctx, const int* colptr, const int* rowind)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (ctx->ncols <= 0) return fail(ctx, MSPAI_ERR_STATE, "set_upper_pattern: upload the matrix first");
    HIPCHK(hipSetDevice(ctx->device));
    if (!colptr || !rowind) {
        dfree(ctx->d_up_ptr);
        dfree(ctx->d_up_ind);
        ctx->up_cols = 0;
        return MSPAI_OK;
    }
    if (ctx->block_bs != 1) return fail(ctx, MSPAI_ERR_ARG, "set_upper_pattern: scalar SPAI only");
    const int n = ctx->ncols;
    if (colptr[0] != 0) return fail(ctx, MSPAI_ERR_ARG, "set_upper_pattern: colptr must start at 0");
    for (int j = 0; j < n; ++j) {
        const int a = colptr[j], b = colptr[j + 1];
        if (b < a) return fail(ctx, MSPAI_ERR_ARG, "set_upper_pattern: colptr not monotone");
        for (int e = a; e < b; ++e) {
            if (rowind[e] < 0 || rowind[e] >= n)
                return fail(ctx, MSPAI_ERR_ARG, "set_upper_pattern: index out of range");
            if (e > a && rowind[e] <= rowind[e - 1])
                return fail(ctx, MSPAI_ERR_ARG, "set_upper_pattern: indices must be sorted and distinct");
        }
    }
    int rc;
    if ((rc = upload(ctx, ctx->d_up_ptr, colptr, n + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_up_ind, rowind, (size_t)colptr[n]))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->up_cols = n;
    return MSPAI_OK;
}

extern "C" int mspai_reuse_pattern(mspai_ctx* ctx)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (ctx->pattern_nnz < 0) return fail(ctx, MSPAI_ERR_STATE, "reuse_pattern: build M first");
    HIPCHK(hipSetDevice(ctx->device));
    if (ctx->block_bs != 1) {
        /* P_Memory of a block build is the block pattern of M */
        if (!ctx->block_built) return fail(ctx, MSPAI_ERR_STATE, "reuse_pattern: build M first");
        const int nb = ctx->nb, cj = ctx->bcapJ;
        std::vector<int> len(nb), idx((size_t)nb * cj), bp(nb + 1, 0), bi;
        HIPCHK(hipMemcpy(len.data(), ctx->d_bm_len, nb * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(idx.data(), ctx->d_bm_idx, (size_t)nb * cj * sizeof(int), hipMemcpyDeviceToHost));
        for (int c = 0; c < nb; ++c) {
            bp[c + 1] = bp[c] + len[c];
            bi.insert(bi.end(), idx.begin() + (size_t)c * cj, idx.begin() + (size_t)c * cj + len[c]);
        }
        if (bi.empty()) bi.push_back(0);
        return mspai_set_pattern(ctx, bp.data(), bi.data());
    }
    std::vector<int> ptr(ctx->ncols + 1), ind(std::max<long long>(ctx->pattern_nnz, 1));
    HIPCHK(hipMemcpy(ptr.data(), ctx->d_Mptr, ptr.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (ctx->pattern_nnz > 0)
        HIPCHK(hipMemcpy(ind.data(), ctx->d_Mind, ctx->pattern_nnz * sizeof(int), hipMemcpyDeviceToHost));
    return mspai_set_pattern(ctx, ptr.data(), ind.data());
}

extern "C" int mspai_set_target(mspai_ctx* ctx, const int* colptr, const int* rowind, const double* val)
{
    if (!ctx) return MSPAI_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    /* M and the FP64 queue length of the last mixed build belong to the old target */
    ctx->nnzM = -1;
    ctx->mixed_fp64_last = -1;
    if (!colptr) {
        dfree(ctx->d_b_ptr);
        dfree(ctx->d_b_ind);
        dfree(ctx->d_b_val);
        ctx->nnzB = 0;
        ctx->b_rows = ctx->b_cols = 0;
        return MSPAI_OK;
    }
    if (ctx->ncols <= 0) return fail(ctx, MSPAI_ERR_STATE, "set_target: upload the matrix first");
    if (!rowind || !val) return MSPAI_ERR_ARG;
    int n = ctx->ncols;
    for (int j = 0; j < n; ++j)
        for (int e = colptr[j]; e < colptr[j + 1]; ++e) {
            if (rowind[e] < 0 || rowind[e] >= ctx->nrows) return fail(ctx, MSPAI_ERR_ARG, "set_target: row out of range");
            if (e > colptr[j] && rowind[e] <= rowind[e - 1])
                return fail(ctx, MSPAI_ERR_ARG, "set_target: rows must be sorted per column");
        }
    int rc;
    if ((rc = upload(ctx, ctx->d_b_ptr, colptr, n + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_b_ind, rowind, (size_t)colptr[n]))) return rc;
    if ((rc = upload(ctx, ctx->d_b_val, val, (size_t)colptr[n]))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->nnzB = colptr[n];
    ctx->b_rows = ctx->nrows;
    ctx->b_cols = n;
    return MSPAI_OK;
}

/* Option checks of PC_MSPAI::PCSetFromOptions_SPAI (src/imspai.cpp:239-480)
 * that concern this path. */
static int check_opts(mspai_ctx* ctx, const mspai_opts* o)
{
    if (o->max_impr_steps < 0) return fail(ctx, MSPAI_ERR_ARG, "Illegal number of improvement steps (ns < 0)");
    if (o->epsilon < 0) return fail(ctx, MSPAI_ERR_ARG, "Illegal epsilon (ep < 0)");
    if (o->maxnew < 0) return fail(ctx, MSPAI_ERR_ARG, "Illegal maxnew (mn < 0)");
    if (o->cache < 0 || o->cache > 10000000) return fail(ctx, MSPAI_ERR_ARG, "Illegal cache size parameter");
    if (o->qr < 0 || o->qr > 5) return fail(ctx, MSPAI_ERR_ARG, "Illegal parameter for qr");
    if (o->qr == 1 || o->qr > 2)
        return fail(ctx, MSPAI_ERR_UNSUPPORTED, "qr 1/3/4/5 select the sparse QR variants, which this library does not build");
    if (o->qr == 2 && o->precision == 1)
        return fail(ctx, MSPAI_ERR_UNSUPPORTED, "the QR update (-qr 2) runs in FP64 only");
    if (o->cache > 0 && o->qr > 0)
        return fail(ctx, MSPAI_ERR_ARG, "QR-Updates and Caching optimization cannot be used simultaneously");
    if (o->use_mean < 0 || o->use_mean > 1) return fail(ctx, MSPAI_ERR_ARG, "Illegal value for using mean value");
    if (o->precision < 0 || o->precision > 1) return fail(ctx, MSPAI_ERR_ARG, "precision must be 0 (FP64) or 1 (mixed)");
    if (ctx->d_b_ptr && (o->max_impr_steps > 0 || o->qr > 0))
        return fail(ctx, MSPAI_ERR_UNSUPPORTED,
                    "Probing constraints and/or target matrix approximations are only supported in static MSPAI");
    return MSPAI_OK;
}

static int ensure_outputs(mspai_ctx* ctx, int cap)
{
    size_t n = ctx->ncols;
    if (ctx->cap == cap && ctx->out_cols == n && ctx->d_m_len) return MSPAI_OK;
    if (ctx->no_alloc) return fail(ctx, MSPAI_ERR_CAPACITY, "outputs: this call allocates no device memory");
    dfree(ctx->d_m_len);
    dfree(ctx->d_m_idx);
    dfree(ctx->d_m_val);
    dfree(ctx->d_res);
    dfree(ctx->d_nsolve);
    dfree(ctx->d_status);
    dfree(ctx->d_work);
    dfree(ctx->d_Mptr);
    dfree(ctx->d_Mind);
    dfree(ctx->d_Mval);
    dfree(ctx->d_queues);
    HIPCHK(hipMalloc((void**)&ctx->d_m_len, n * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_m_idx, n * cap * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_m_val, n * cap * sizeof(double)));
    HIPCHK(hipMalloc((void**)&ctx->d_res, n * sizeof(double)));
    HIPCHK(hipMalloc((void**)&ctx->d_nsolve, n * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_status, n * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_work, n * kNumWork * sizeof(double)));
    HIPCHK(hipMalloc((void**)&ctx->d_Mptr, (n + 1) * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_Mind, n * cap * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_Mval, n * cap * sizeof(double)));
    HIPCHK(hipMalloc((void**)&ctx->d_queues, (2 * kMaxBins + 1) * n * sizeof(int)));
    ctx->cap = cap;
    ctx->out_cols = n;
    ctx->Mcap = n * cap;
    /* scan workspace for the compaction */
    size_t need = 0;
    hipcub::DeviceScan::InclusiveSum(nullptr, need, ctx->d_m_len, ctx->d_Mptr + 1, (int)n, ctx->stream);
    need = std::max<size_t>(need, kNumWork * sizeof(double));
    if (need > ctx->cub_bytes) {
        if (ctx->d_cub) (void)hipFree(ctx->d_cub);
        ctx->d_cub = nullptr;
        HIPCHK(hipMalloc(&ctx->d_cub, need));
        ctx->cub_bytes = need;
    }
    return MSPAI_OK;
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

static size_t next_pow2(size_t x)
{
    size_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

/* Buffers a launch needs besides the FP64 build's (bit flags). */
enum LayoutNeeds { kNeedMixed = 1, kNeedBackward = 2, kNeedQrUpdate = 4, kNeedA0 = 8 };

/* Byte offsets of one group's working set holding the given sizes.  The
 * std::sort tie-path buffers (pairs, stack) sit in the main region when
 * tie_in_main, else in a separate per-group HBM region of tie_per_group
 * bytes (LDS bins: the path is rare). */
static ScratchLayout make_layout(int capJ, size_t capI, size_t capU, size_t capSort, int maxnew, size_t capPairs,
                                 int needs, bool tie_in_main, int group = 0, size_t capSid = 0)
{
    ScratchLayout L{};
    capI = std::max<size_t>(capI, 1);
    capU = std::max<size_t>(capU, 1);
    capSort = next_pow2(std::max<size_t>(capSort, 1));
    L.capJ = capJ;
    L.capI = (int)capI;
    L.capU = (int)capU;
    L.capSort = (int)capSort;
    L.capPairs = (int)std::max<size_t>(1, capPairs);
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o = align_up(o + bytes, 16);
        return at;
    };
    size_t capB = std::max<size_t>(capI, (size_t)capJ);
    L.J = take((size_t)capJ * 4 + 16);
    L.I = take(capI * 4);
    if (needs & kNeedQrUpdate) {
        /* the QR update keeps its factor in aqr across the steps */
        L.sbuf = take(capSort * 4);
        L.aqr = take(capI * capJ * 8);
        L.jt = take(capU * 4);
        L.rho = take(capU * 8);
    } else {
        /* A(I,J) lives from the gather to the residual, the raw row lists
         * (sbuf) in Get_I_Set before the gather and in the augmentation
         * after the residual, with the candidates and their scores (jt,
         * rho): one region holds either */
        const size_t sb = align_up(capSort * 4, 16), jb = align_up(capU * 4, 16);
        const size_t at = take(std::max(capI * capJ * 8, sb + jb + capU * 8));
        L.aqr = at;
        L.sbuf = at;
        L.jt = at + sb;
        L.rho = at + sb + jb;
    }
    L.b = take(capB * 8);
    L.b0 = take(capB * 8);
    L.r = take(capB * 8);
    L.tau = take((size_t)capJ * 8);
    L.sel = take((size_t)std::max(1, maxnew) * 4);
    L.segs = take(((size_t)capJ + 1) * 4);
    L.segoff = take(((size_t)capJ + 1) * 4);
    /* the flat entries of J's columns: at most |J| * max column nnz, and
     * never more than the raw list they are sorted in */
    L.capSid = (int)std::min<size_t>(capSort, capSid > 0 ? capSid : capSort);
    L.sid = take((size_t)L.capSid * 4);
    L.wsum = take(8 * 8);
    /* register bins convert A(I,J) to FP32 as they load it: aq32 holds R only */
    const bool regbin = register_bin(group, (int)capI, capJ);
    /* the mixed builds' 64-lane LDS bins factor the Gram matrix instead of
     * an FP32 copy of A(I,J) (not the adjoint, whose solve is the QR's) */
    const bool gram = (needs & kNeedMixed) && !(needs & kNeedBackward) && group == 64 && !regbin;
    L.gram = gram;
    L.ldq = regbin ? capJ : (gram ? kGramLd : (int)capI);
    if (needs & kNeedMixed) L.aq32 = take((size_t)L.ldq * (gram ? kGramLd : capJ) * 4);
    if (needs & kNeedBackward) L.rq = take((size_t)capJ * capJ * 8);
    if (needs & kNeedA0) {
        L.a0 = take(capI * capJ * 8);
        L.keep_a0 = 1;
    }
    /* the FP32 right-hand side of the LDS QR and the vectors of its
     * refinement, and the adjoint's vectors (the register QR and its
     * refinement keep theirs in registers) */
    if ((needs & kNeedBackward) || ((needs & kNeedMixed) && !regbin)) {
        L.bq32 = take(capB * 4);
        L.g = take((size_t)capJ * 8);
        L.y = take((size_t)capJ * 8);
        L.z = take((size_t)capJ * 8);
    }
    if (needs & kNeedQrUpdate) {
        L.srow = take(capB * 4);
        L.spos = take(capB * 4);
        L.iold = take(capB * 4);
        L.sposold = take(capB * 4);
        L.jold = take((size_t)capJ * 4);
        L.colid = take((size_t)capJ * 4);
        L.reflrows = take((size_t)capJ * 4);
        L.bperm = take(capB * 8);
        L.yq = take(capB * 8);
    }
    size_t tie_pairs = (size_t)L.capPairs * sizeof(RhoPair);
    size_t tie_stack = 3 * introsort_detail::kStackFrames * sizeof(int);
    if (tie_in_main) {
        L.pairs = take(tie_pairs);
        L.stack = take(tie_stack);
        L.tie_in_main = 1;
        L.tie_per_group = 0;
    } else {
        L.pairs = 0;
        L.stack = align_up(tie_pairs, 16);
        L.tie_in_main = 0;
        L.tie_per_group = align_up(L.stack + tie_stack, 256);
    }
    L.per_wave = align_up(o, 256);
    /* chunk size cap: cheap columns (small bins) in large chunks, expensive
     * ones in small chunks */
    L.max_rounds = capI <= 32 ? 16 : (capI <= 64 ? 8 : 2);
    return L;
}

/* Worst-case layout of the generic kernel: |J| <= cap,
 * |I| <= min(nrows, cap * max nnz per column), raw candidate list
 * <= (|I| + 1) * max nnz per row. */
static ScratchLayout generic_layout(const mspai_ctx* ctx, int cap, int maxnew, bool adaptive, int needs)
{
    size_t rawI = (size_t)cap * std::max(1, ctx->maxcolnnz);
    size_t capI = std::max<size_t>(std::min<size_t>(ctx->nrows, rawI), 1);
    size_t rawU = adaptive ? (capI + 1) * std::max(1, ctx->maxrownnz) : 1;
    size_t capU = adaptive ? std::min<size_t>(ctx->ncols, rawU) : 1;
    /* the FP64 build's generic kernel keeps a copy of A(I,J) for the dense
     * residual (its QR runs in HBM and overwrites the working copy) */
    const int a0 = (needs & (kNeedQrUpdate | kNeedBackward)) ? 0 : kNeedA0;
    return make_layout(cap, capI, capU, std::max(rawI, rawU), maxnew, capU, needs | a0, true, 0, rawI);
}

/* One size bin: a lane-group width, a resident-wave cap and the LDS layout
 * of one group. */
struct Bin {
    int group;
    int wcap;
    ScratchLayout L;
    int fp64;  // a bin of a mixed build's chain that runs the FP64 kernel (FP64 layout)
};

/* A size-bin table: per bin the lane-group width, the |J|, |I|, distinct
 * candidate and raw-list capacities, and a cap on resident waves per CU
 * (0 = as many as LDS allows, at most 32); `mixed` > 0 limits a mixed
 * build's FP32 chain to the first `mixed` bins (the rest run in FP64). */
struct BinEdge {
    int group, J, I, U, sort, wcap;
};
struct BinTable {
    const char* name;
    int count;
    BinEdge e[4];
    int mixed;
    int fp64mask;  // bins (bit t for entry t) that a mixed build runs in FP64 inside its chain
};

/* Candidate tables of the autotuner; table 0 is the default. */
static const BinTable kTables[] = {
    {"g16-g32-g64-g64", 4,
     {{16, 8, 16, 64, 128, 0}, {32, 16, 32, 128, 256, 0}, {64, 32, 64, 256, 512, 0},
      {64, 40, 128, 512, 1024, 0}}},
    {"g32-g64-g64", 3,
     {{32, 16, 32, 128, 256, 0}, {64, 32, 64, 256, 512, 0}, {64, 40, 128, 512, 1024, 0}}},
    {"g16wide-g64-g64", 3,
     {{16, 16, 16, 128, 256, 0}, {64, 32, 64, 256, 512, 0}, {64, 40, 128, 512, 1024, 0}}},
    {"g64-g64", 2, {{64, 32, 64, 256, 512, 0}, {64, 40, 128, 512, 1024, 0}}},
    {"g16-g64", 2, {{16, 8, 16, 64, 128, 0}, {64, 40, 128, 512, 1024, 0}}},
    {"g16-g32-g64-g64/occ12", 4,
     {{16, 8, 16, 64, 128, 12}, {32, 16, 32, 128, 256, 12}, {64, 32, 64, 256, 512, 8},
      {64, 40, 128, 512, 1024, 4}}},
    {"g16-g64/occ8", 2, {{16, 8, 16, 64, 128, 8}, {64, 40, 128, 512, 1024, 0}}},
    {"g32-g64-g64/occ8", 3, {{32, 16, 32, 128, 256, 8}, {64, 32, 64, 256, 512, 0}, {64, 40, 128, 512, 1024, 0}}},
    {"g16-g64/occ10", 2, {{16, 8, 16, 64, 128, 10}, {64, 40, 128, 512, 1024, 0}}},
    /* long candidate lists: a last bin of |I| <= 384 and raw lists to 8192
     * (about 117 KB of LDS per wave at |J| <= 33, one wave per CU) keeps the
     * widest columns of the random band out of the HBM working set */
    {"g16-g64-g64-g64x384", 4,
     {{16, 8, 16, 64, 128, 0}, {64, 32, 64, 512, 1024, 0}, {64, 40, 128, 1024, 2048, 0},
      {64, 40, 384, 4096, 8192, 0}}},
    {"g16-g64-g64x192-g64x384", 4,
     {{16, 8, 16, 64, 128, 0}, {64, 32, 64, 512, 1024, 0}, {64, 40, 192, 2048, 4096, 0},
      {64, 40, 384, 4096, 8192, 0}}},
    {"g32-g64-g64-g64x384", 4,
     {{32, 16, 32, 256, 512, 0}, {64, 32, 64, 512, 1024, 0}, {64, 40, 128, 1024, 2048, 0},
      {64, 40, 384, 4096, 8192, 0}}},
    {"g64-g64-g64x384", 3,
     {{64, 32, 64, 512, 1024, 0}, {64, 40, 128, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}},
    /* no size bins: every column in the generic kernel */
    {"generic", 0, {}},
    /* |J| <= 16 in the bulk bin: half the A(I,J) region, about 14 KB of LDS
     * per wave, 11 resident waves per CU instead of 7 */
    {"g64x16-g64-g64x384", 3,
     {{64, 16, 64, 512, 1024, 0}, {64, 40, 128, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}},
    {"g16-g64x16-g64-g64x384", 4,
     {{16, 8, 16, 64, 128, 0}, {64, 16, 64, 512, 1024, 0}, {64, 40, 128, 1024, 2048, 0},
      {64, 40, 384, 4096, 8192, 0}}},
    /* the same with the bulk bin at 8 resident waves per CU: its FP64
     * kernel then keeps every register (no spills to scratch) */
    {"g64x16/occ8-g64-g64x384", 3,
     {{64, 16, 64, 512, 1024, 8}, {64, 40, 128, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}},
    {"g16-g64x16/occ8-g64-g64x384", 4,
     {{16, 8, 16, 64, 128, 0}, {64, 16, 64, 512, 1024, 8}, {64, 40, 128, 1024, 2048, 0},
      {64, 40, 384, 4096, 8192, 0}}},
    /* a narrower second bin (|J| <= 24 or 20, |I| <= 96 or 112: about 26 KB,
     * 6 resident waves per CU instead of 3) ahead of the wide ones */
    {"g64x16-g64x24i96-g64i160-g64x384", 4,
     {{64, 16, 64, 512, 1024, 0}, {64, 24, 96, 1024, 2048, 0}, {64, 40, 160, 2048, 4096, 0},
      {64, 40, 384, 4096, 8192, 0}}},
    {"g64x16-g64x24i96-g64x384", 3,
     {{64, 16, 64, 512, 1024, 0}, {64, 24, 96, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}},
    {"g64x16-g64x20i112-g64x384", 3,
     {{64, 16, 64, 512, 1024, 0}, {64, 20, 112, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}},
    /* the same two with the FP32 QR in the bulk bin only: the wider bins'
     * FP64 kernels run faster than their mixed ones (identical tables in
     * FP64 builds, which the autotuner skips) */
    {"g64x16-g64x24i96-g64x384/m1", 3,
     {{64, 16, 64, 512, 1024, 0}, {64, 24, 96, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}, 1},
    {"g64x16-g64x20i112-g64x384/m1", 3,
     {{64, 16, 64, 512, 1024, 0}, {64, 20, 112, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}, 1},
    /* (u) the bulk bin with room for 384 candidates instead of 512: about
     * 1.5 KB less LDS per wave, one more resident wave per CU */
    {"g64x16u-g64x20i112-g64x384", 3,
     {{64, 16, 64, 384, 1024, 0}, {64, 20, 112, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}},
    {"g64x16u-g64x20i112-g64x384/m1", 3,
     {{64, 16, 64, 384, 1024, 0}, {64, 20, 112, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}, 1},
    /* a second bin of |J| <= 24, |I| <= 128 (the eight-row register QR),
     * alone or after the |J| <= 20, |I| <= 112 one */
    {"g64x16u-g64x24i128-g64x384", 3,
     {{64, 16, 64, 384, 1024, 0}, {64, 24, 128, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}},
    {"g64x16u-g64x24i128-g64x384/m1", 3,
     {{64, 16, 64, 384, 1024, 0}, {64, 24, 128, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}, 1},
    {"g64x16u-x20i112-x24i128-x384", 4,
     {{64, 16, 64, 384, 1024, 0}, {64, 20, 112, 1024, 2048, 0}, {64, 24, 128, 1024, 2048, 0},
      {64, 40, 384, 4096, 8192, 0}}},
    {"g64x16u-x20i112-x24i128-x384/m1", 4,
     {{64, 16, 64, 384, 1024, 0}, {64, 20, 112, 1024, 2048, 0}, {64, 24, 128, 1024, 2048, 0},
      {64, 40, 384, 4096, 8192, 0}}, 1},
    /* mixed builds: FP32 in the bulk bin and (the Gram matrix on the matrix
     * cores) in the widest one, the second bin in FP64 on its eight-row
     * register QR (f) */
    {"g64x16u-g64x24i128f-g64x384", 3,
     {{64, 16, 64, 384, 1024, 0}, {64, 24, 128, 1024, 2048, 0}, {64, 40, 384, 4096, 8192, 0}}, 0, 0x2},
    {"g64x16u-x20i112f-x24i128f-x384", 4,
     {{64, 16, 64, 384, 1024, 0}, {64, 20, 112, 1024, 2048, 0}, {64, 24, 128, 1024, 2048, 0},
      {64, 40, 384, 4096, 8192, 0}}, 0, 0x6},
};
constexpr int kNumTables = sizeof(kTables) / sizeof(kTables[0]);

/* LDS layouts of a table's bins, smallest first: columns that outgrow a bin
 * move to the next one; what outgrows the last runs generic.  A bin that
 * already holds every column (by the worst-case bounds) ends the list. */
static std::vector<Bin> bin_layouts(const mspai_ctx* ctx, int cap, int maxnew, bool adaptive,
                                    const ScratchLayout& G, const BinTable& table, int needs)
{
    std::vector<Bin> out;
    for (int t = 0; t < table.count; ++t) {
        if ((needs & kNeedMixed) && table.mixed > 0 && (int)out.size() >= table.mixed) break;
        const BinEdge& e = table.e[t];
        int capJ = std::min(cap, e.J);
        size_t capI = std::min<size_t>(G.capI, e.I);
        size_t capU = adaptive ? std::min<size_t>(G.capU, e.U) : 1;
        size_t capSort = std::min<size_t>(G.capSort, e.sort);
        /* a mixed chain's FP64 bin: FP64 layout (the mixed build, not its adjoint) */
        const int fp64 = (needs & kNeedMixed) && !(needs & kNeedBackward) && ((table.fp64mask >> t) & 1);
        const int bneeds = fp64 ? (needs & ~kNeedMixed) : needs;
        /* ties of scores across the cut are ordered in the generic kernel only */
        ScratchLayout L = make_layout(capJ, capI, capU, capSort, maxnew, 1, bneeds, false, e.group,
                                      (size_t)capJ * std::max(1, ctx->maxcolnnz));
        if (L.per_wave * (64 / e.group) > ctx->lds_block) continue;
        out.push_back({e.group, e.wcap, L, fp64});
        if (capJ == cap && L.capI == G.capI && L.capSort >= G.capSort && (!adaptive || L.capU >= G.capU)) break;
    }
    return out;
}

/* The bin table of a build: the autotuner's choice for the build's precision
 * (binning == 2), else table 0. */
static int table_for(const mspai_ctx* ctx, const mspai_opts* o)
{
    const int t = ctx->tuned_table[o->precision == 1 ? 1 : 0];
    return o->binning == 2 && t >= 0 ? t : 0;
}

static int ensure_scratch(mspai_ctx* ctx, size_t need)
{
    if (need <= ctx->scratch_bytes) return MSPAI_OK;
    if (ctx->no_alloc) return fail(ctx, MSPAI_ERR_CAPACITY, "working sets: this call allocates no device memory");
    dfree(ctx->d_scratch);
    HIPCHK(hipMalloc((void**)&ctx->d_scratch, need));
    ctx->scratch_bytes = need;
    return MSPAI_OK;
}

/* Resident waves per CU for a bin's kernel in `mode`: what its registers
 * and 160 KB of LDS allow (runtime occupancy query), at most 32, and at most
 * the bin's cap.  A persistent launch of more waves than fit only adds
 * waves that find the queue drained. */
static int waves_per_cu(const Bin& b, int mode)
{
    size_t lds = b.L.per_wave * (64 / b.group);
    int w = (int)std::min<size_t>(32, std::max<size_t>(1, (160 * 1024) / lds));
    const int occ = binned_blocks_per_cu(b.L, b.group, mode, b.wcap);
    if (occ > 0) w = std::min(w, occ);
    if (b.wcap > 0) w = std::min(w, b.wcap);
    return w;
}

/* Per-launch profile: an event before and after each column-kernel launch,
 * the launch's name, and its slot of the counter totals (builds with
 * counters only; a.work is set exactly then). */
static int begin_launch(mspai_ctx* ctx, BuildArgs& a, const char* what, int mode, int bin, const Bin* b)
{
    const int k = ctx->nlaunch;
    if (k >= kMaxLaunch) return fail(ctx, MSPAI_ERR_STATE, "more column-kernel launches than the profile holds");
    static const char* kModes[] = {"FP64", "mixed", "adjoint FP64", "adjoint mixed", "FP64 QR update"};
    if (b)
        std::snprintf(ctx->lname[k], sizeof ctx->lname[k], "%s bin %d G=%d |J|<=%d |I|<=%d", kModes[mode], bin,
                      b->group, b->L.capJ, b->L.capI);
    else
        std::snprintf(ctx->lname[k], sizeof ctx->lname[k], "%s %s", kModes[mode], what);
    a.lwork = a.work ? ctx->d_lwork + (size_t)k * kLaunchStat : nullptr;
    HIPCHK(hipEventRecord(ctx->lev[2 * k], ctx->stream));
    return MSPAI_OK;
}

static int end_launch(mspai_ctx* ctx)
{
    HIPCHK(hipEventRecord(ctx->lev[2 * ctx->nlaunch + 1], ctx->stream));
    ctx->nlaunch++;
    return MSPAI_OK;
}

/* HBM tie-path buffers of the LDS bins: one region per resident group (the
 * mixed chain's bins in their mode, the FP64 chain's in FP64). */
static size_t tie_bytes_needed(const mspai_ctx* ctx, const std::vector<Bin>& bins, const std::vector<Bin>& fbins,
                               bool mixed, int fp64_mode)
{
    size_t need = 0;
    for (const Bin& b : fbins)
        need = std::max(need, (size_t)ctx->num_cu * waves_per_cu(b, fp64_mode) * (64 / b.group) * b.L.tie_per_group);
    if (mixed)
        for (const Bin& b : bins)
            need = std::max(need, (size_t)ctx->num_cu * waves_per_cu(b, b.fp64 ? fp64_mode : (fp64_mode == 2 ? 3 : 1)) *
                                      (64 / b.group) * b.L.tie_per_group);
    return need;
}

/* Most FP64 columns of a mixed build that the generic kernel takes alone. */
constexpr int kDirectFp64 = 4096;

/* Launch routing and the bin chain over the columns whose patterns are in
 * a.p_ptr/p_ind.  `bins` are the mixed chain's bins in a mixed build (else
 * the FP64 chain's), `fbins` the FP64 chain's.  Queues: FP64 queue b at
 * d_queues + b*n (b = 0..nf, nf = generic), mixed queue b at d_queues +
 * (kMaxBins + 1 + b)*n.  Routing fills the first queues; a bin's overflow
 * appends to the next queue of its chain; the mixed chain ends in FP64
 * queue 0, which also takes the columns the mixed guards decline.  Columns
 * go on from the step they reached (save_state) except in the adjoint and
 * the QR update, which restart them.  Counters: FP64 [0..8) lengths,
 * [8..16) heads, [16..24) routed lengths, mixed [24..32) lengths, [32..40)
 * heads, [40] blocks of the routing kernel done. */
static int run_chain(mspai_ctx* ctx, BuildArgs& a, const std::vector<Bin>& bins, const std::vector<Bin>& fbins,
                     const ScratchLayout& G, int gwaves, bool mixed, int fp64_mode, int n)
{
    const int nb = (int)bins.size();
    const int nf = (int)fbins.size();
    const std::vector<Bin>& routed = mixed ? bins : fbins;  // the chain the routing kernel fills
    BinEdges edges{};
    for (size_t b = 0; b < routed.size(); ++b) {
        edges.capJ[b] = routed[b].L.capJ;
        edges.capSort[b] = routed[b].L.capSid;  // the I-set raw list of the start pattern
        edges.capI[b] = routed[b].L.capI;
    }
    int* counts = ctx->d_counter;
    int* heads = ctx->d_counter + 8;
    int* mcounts = ctx->d_counter + 24;
    int* mheads = ctx->d_counter + 32;
    int* fq = ctx->d_queues;
    int* mq = ctx->d_queues + (size_t)(kMaxBins + 1) * n;
    int rc;
    launch_fill_int(ctx->d_counter, 64, 0, ctx->stream);  // a kernel, not a memset node
    HIPCHK(hipGetLastError());
    ctx->nlaunch = 0;
    ctx->lwork_valid = a.work != nullptr;
    if (a.work) HIPCHK(hipMemsetAsync(ctx->d_lwork, 0, kMaxLaunch * kLaunchStat * sizeof(double), ctx->stream));
    a.status_or = 0;
    const size_t tie_need = tie_bytes_needed(ctx, bins, fbins, mixed, fp64_mode);
    if (tie_need > ctx->tie_bytes) {
        if (ctx->no_alloc) return fail(ctx, MSPAI_ERR_CAPACITY, "tie buffers: this call allocates no device memory");
        dfree(ctx->d_tie);
        HIPCHK(hipMalloc((void**)&ctx->d_tie, tie_need));
        ctx->tie_bytes = tie_need;
    }
    a.tie_scratch = ctx->d_tie;
    if (mixed) {
        /* FP64 bins that hold nothing the last mixed bin cannot: the mixed
         * chain's overflow passes them by (declined columns start at bin 0) */
        const ScratchLayout& Lm = bins[nb - 1].L;
        int skip = 0;
        while (skip < nf && fbins[skip].L.capJ <= Lm.capJ && fbins[skip].L.capI <= Lm.capI &&
               fbins[skip].L.capSort <= Lm.capSort && fbins[skip].L.capU <= Lm.capU &&
               fbins[skip].L.capSid <= Lm.capSid)
            ++skip;
        /* the columns for FP64 go to the generic kernel in one launch when
         * the last mixed build had few of them and no FP64 bin is wider than
         * the mixed ones: the FP64 bins would mostly launch on empty queues
         * (not after a Gram bin, whose declined columns are wide: the LDS
         * bins take them faster). Both routes give the same bits. */
        const bool direct = fp64_mode == 0 && skip == nf && ctx->mixed_fp64_last >= 0 &&
                            ctx->mixed_fp64_last <= kDirectFp64 && !Lm.gram;
        const int to = direct ? 0 : skip;  // FP64 queue of the mixed chain's overflow
        launch_classify(a, nb, edges, mq, mcounts, fq, counts, n, ctx->d_counter + 40, ctx->d_counter + 16,
                        ctx->stream);
        HIPCHK(hipGetLastError());
        a.fallback = fq;
        a.n_fallback = counts;
        for (int b = 0; b < nb; ++b) {
            const int grp = bins[b].group;
            a.overflow = b + 1 < nb ? mq + (size_t)(b + 1) * n : fq + (size_t)to * n;
            /* overflow and declined steps go on from the step they reached,
             * in the next mixed bin or in the FP64 chain */
            a.save_state = fp64_mode != 2;
            a.n_overflow = b + 1 < nb ? mcounts + b + 1 : counts + to;
            /* a bin the table runs in FP64 inside the mixed chain passes its
             * overflow on like the others and declines nothing */
            const int mode = bins[b].fp64 ? fp64_mode : (fp64_mode == 2 ? 3 : 1);
            int per_cu = waves_per_cu(bins[b], mode);
            const ScratchLayout& L = bins[b].L;
            if (b == 0) HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
            if ((rc = begin_launch(ctx, a, nullptr, mode, b, &bins[b]))) return rc;
            launch_binned(a, L, grp, mode, mq + (size_t)b * n, mcounts + b, mheads + b, ctx->num_cu * per_cu,
                          ctx->stream, bins[b].wcap);
            HIPCHK(hipGetLastError());
            if ((rc = end_launch(ctx))) return rc;
            if (b == 0) HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
        }
        a.status_or = COL_FALLBACK;
        a.save_state = 0;
        if (direct) {
            a.fallback = nullptr;
            a.n_fallback = nullptr;
            a.overflow = fq;  // never written: the generic layout holds every column
            a.n_overflow = counts + 7;
            if ((rc = begin_launch(ctx, a, "generic (HBM working set)", fp64_mode, -1, nullptr))) return rc;
            launch_fp64_generic(a, G, fp64_mode, fq, counts, heads + nf, gwaves, ctx->stream);
            HIPCHK(hipGetLastError());
            return end_launch(ctx);
        }
    } else {
        launch_classify(a, nf, edges, fq, counts, fq + (size_t)nf * n, counts + nf, n, ctx->d_counter + 40,
                        ctx->d_counter + 16, ctx->stream);
        HIPCHK(hipGetLastError());
    }
    a.fallback = nullptr;
    a.n_fallback = nullptr;
    for (int b = 0; b < nf; ++b) {
        const int grp = fbins[b].group;
        a.overflow = fq + (size_t)(b + 1) * n;
        a.n_overflow = counts + b + 1;
        int per_cu = waves_per_cu(fbins[b], fp64_mode);
        const ScratchLayout& L = fbins[b].L;
        const bool first = b == 0 && !mixed;
        a.save_state = fp64_mode == 0;  // FP64 builds resume in the next bin (not the QR update or the adjoint)
        if (first) HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
        if ((rc = begin_launch(ctx, a, nullptr, fp64_mode, b, &fbins[b]))) return rc;
        launch_binned(a, L, grp, fp64_mode, fq + (size_t)b * n, counts + b, heads + b,
                      ctx->num_cu * per_cu, ctx->stream, fbins[b].wcap);
        HIPCHK(hipGetLastError());
        if ((rc = end_launch(ctx))) return rc;
        if (first) HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
    }
    a.overflow = fq;  // never written: the generic layout holds every column
    a.n_overflow = counts + 7;
    a.save_state = 0;
    if (nf == 0) HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));  // no bins: the generic kernel is the first
    if ((rc = begin_launch(ctx, a, "generic (HBM working set)", fp64_mode, -1, nullptr))) return rc;
    launch_fp64_generic(a, G, fp64_mode, fq + (size_t)nf * n, counts + nf, heads + nf, gwaves, ctx->stream);
    HIPCHK(hipGetLastError());
    if ((rc = end_launch(ctx))) return rc;
    if (nf == 0) HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
    return MSPAI_OK;
}

/* Queue the copy of the counters to the pinned host array (after the last
 * kernel of a build on the stream; read once the stream is synchronized). */
static int fetch_counters(mspai_ctx* ctx)
{
    HIPCHK(hipMemcpyAsync(ctx->h_counter, ctx->d_counter, 64 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    return MSPAI_OK;
}

/* Read the queue counters of the last chain (fetched with fetch_counters)
 * into stats and the context. */
static int read_chain_counts(mspai_ctx* ctx, int nbins, bool mixed, mspai_stats* stats)
{
    const int* cnt = ctx->h_counter;
    ctx->last_nbins = nbins;
    stats->n_overflow = 0;
    stats->n_fallback = 0;
    if (mixed) {
        int moved = 0;
        for (int b = 1; b < nbins; ++b) moved += cnt[24 + b] - cnt[16 + b];
        stats->n_overflow = moved;
        stats->n_fallback = cnt[62];  // columns with COL_FALLBACK, counted by the compaction
        for (int b = 0; b < 8; ++b) {
            ctx->last_routed[b] = cnt[16 + b];
            ctx->last_queued[b] = b < nbins ? cnt[24 + b] : (b == nbins ? cnt[0] : 0);
        }
    } else {
        for (int b = 0; b <= nbins; ++b) stats->n_overflow += cnt[b] - cnt[16 + b];
        for (int b = 0; b < 8; ++b) {
            ctx->last_routed[b] = cnt[16 + b];
            ctx->last_queued[b] = cnt[b];
        }
    }
    return MSPAI_OK;
}

/* Slots of the pattern->R cache: -cs entries rounded up to a power of two,
 * at most 2^18; each slot holds up to 32 x 16 blocks. */
static int ensure_cache(mspai_ctx* ctx, int entries, int capI, int capJ)
{
    if (ctx->no_alloc) return fail(ctx, MSPAI_ERR_UNSUPPORTED, "the cache (-cs) allocates device memory");
    int slots = 1;
    while (slots < entries && slots < (1 << 18)) slots <<= 1;
    if (slots == ctx->cache_slots && capI == ctx->cache_capI && capJ == ctx->cache_capJ) return MSPAI_OK;
    dfree(ctx->d_cache_state);
    dfree(ctx->d_cache_key);
    dfree(ctx->d_cache_m);
    dfree(ctx->d_cache_n);
    dfree(ctx->d_cache_status);
    dfree(ctx->d_cache_data);
    size_t sd = 2 * (size_t)capI * capJ + 2 * (size_t)capJ + (size_t)capI;
    HIPCHK(hipMalloc((void**)&ctx->d_cache_state, slots * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_cache_key, slots * sizeof(uint32_t)));
    HIPCHK(hipMalloc((void**)&ctx->d_cache_m, slots * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_cache_n, slots * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_cache_status, slots * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_cache_data, slots * sd * sizeof(double)));
    ctx->cache_slots = slots;
    ctx->cache_capI = capI;
    ctx->cache_capJ = capJ;
    return MSPAI_OK;
}

/* Generic-kernel sizing: keep its HBM working sets within 16 GiB. */
static int prepare_generic(mspai_ctx* ctx, const ScratchLayout& G, int* gwaves)
{
    int w = std::max(1, ctx->num_cu) * 16;  // 4 waves per SIMD, the generic kernels' register budget
    /* without allocation: as many waves as the given working-set region holds */
    const size_t budget = ctx->no_alloc ? ctx->scratch_bytes : (size_t)16 << 30;
    while (w > 1 && (size_t)w * G.per_wave > budget) w /= 2;
    if ((size_t)w * G.per_wave > budget) return fail(ctx, MSPAI_ERR_CAPACITY, "per-column working set too large");
    *gwaves = w;
    return ensure_scratch(ctx, (size_t)w * G.per_wave);
}

static BuildArgs base_args(mspai_ctx* ctx)
{
    BuildArgs a{};
    a.A = DevCsc{ctx->nrows, ctx->ncols, ctx->d_colptr, ctx->d_rowind, ctx->d_val};
    a.rp_ptr = ctx->d_rp_ptr;
    a.rp_ind = ctx->d_rp_ind;
    a.rp_val = ctx->rp_match ? ctx->d_rp_val : nullptr;  // filled with the column norms, before any use
    a.rp_maxlen = ctx->maxrownnz;
    a.B = DevCsc{ctx->nrows, ctx->ncols, ctx->d_b_ptr, ctx->d_b_ind, ctx->d_b_val};
    a.ajn = ctx->d_ajn;
    a.scratch = ctx->d_scratch;
    return a;
}

static int build_block(mspai_ctx* ctx, const mspai_opts* opts, mspai_stats* stats);

extern "C" int mspai_build(mspai_ctx* ctx, const mspai_opts* opts, mspai_stats* stats)
{
    if (!ctx || !opts) return MSPAI_ERR_ARG;
    if (ctx->ncols <= 0) return fail(ctx, MSPAI_ERR_STATE, "build: upload the matrix first");
    int rc;
    if ((rc = check_opts(ctx, opts))) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    if ((rc = check_shapes(ctx))) return rc;
    if (ctx->block_bs != 1) return build_block(ctx, opts, stats);
    if (!ctx->d_p_ptr) {
        if ((rc = mspai_set_pattern(ctx, nullptr, nullptr))) return rc;
        ctx->pattern_set = false;  // the implicit diagonal: rebuilt for the next matrix
    }
    const int n = ctx->ncols;
    const int ns = opts->max_impr_steps;
    const int cap = ctx->plen_max + opts->maxnew * ns;
    if (cap <= 0) return fail(ctx, MSPAI_ERR_ARG, "build: empty start pattern and no augmentation");
    if ((rc = ensure_outputs(ctx, cap))) return rc;
    const bool adaptive = ns > 0;
    const bool qrup = opts->qr == 2 && ns > 0 && opts->maxnew > 0;  // opt_level 3, src/imspai.cpp:483-491
    const int needs = (qrup ? kNeedQrUpdate : 0) | (opts->precision == 1 && !qrup ? kNeedMixed : 0);
    const ScratchLayout G = generic_layout(ctx, cap, opts->maxnew, adaptive, needs);
    std::vector<Bin> bins, fbins;
    if (opts->binning && !ctx->d_up_ptr) {  // an upper pattern runs unbinned
        int t = table_for(ctx, opts);
        bins = bin_layouts(ctx, cap, opts->maxnew, adaptive, G, kTables[t], needs);
        /* the FP64 chain of a mixed build: the same table in FP64 layouts
         * (no FP32 copies), which may keep bins the mixed layouts exceed */
        fbins = (needs & kNeedMixed) ? bin_layouts(ctx, cap, opts->maxnew, adaptive, G, kTables[t],
                                                   needs & ~kNeedMixed)
                                     : bins;
        ctx->last_table = t;
    }
    int gwaves = 0;
    if ((rc = prepare_generic(ctx, G, &gwaves))) return rc;

    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    BuildArgs a = base_args(ctx);
    if (!ctx->ajn_ok) {  // once per set of values
        hipLaunchKernelGGL(column_norms_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, a.A, ctx->d_ajn);
        if (a.rp_val)
            hipLaunchKernelGGL(row_values_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, a.A, a.rp_ptr,
                               a.rp_ind, ctx->d_rp_val);
        HIPCHK(hipGetLastError());
        ctx->ajn_ok = true;
    }
    a.p_ptr = ctx->d_p_ptr;
    a.p_ind = ctx->d_p_ind;
    a.up_ptr = ctx->d_up_ptr;
    a.up_ind = ctx->d_up_ind;
    a.epsilon = opts->epsilon;
    a.maxnew = opts->maxnew;
    a.ns = ns;
    a.use_mean = opts->use_mean;
    a.cap = cap;
    a.m_len = ctx->d_m_len;
    a.m_idx = ctx->d_m_idx;
    a.m_val = ctx->d_m_val;
    a.res = ctx->d_res;
    a.nsolve = ctx->d_nsolve;
    a.status = ctx->d_status;
    a.work = opts->counters ? ctx->d_work : nullptr;  // opt-in: 64 B per column of writes
    const bool mixed = opts->precision == 1 && !bins.empty() && !qrup;
    if (opts->cache > 0 && !qrup && opts->precision == 0) {
        const int cI = std::min(G.capI, 32), cJ = std::min(cap, 16);
        if ((rc = ensure_cache(ctx, opts->cache, cI, cJ))) return rc;
        HIPCHK(hipMemsetAsync(ctx->d_cache_state, 0, ctx->cache_slots * sizeof(int), ctx->stream));
        a.cache_slots = ctx->cache_slots;
        a.cache_capI = ctx->cache_capI;
        a.cache_capJ = ctx->cache_capJ;
        a.cache_state = ctx->d_cache_state;
        a.cache_key = ctx->d_cache_key;
        a.cache_m = ctx->d_cache_m;
        a.cache_n = ctx->d_cache_n;
        a.cache_status = ctx->d_cache_status;
        a.cache_data = ctx->d_cache_data;
        a.cache_hits = ctx->d_counter + 48;
    }
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    if ((rc = run_chain(ctx, a, bins, fbins, G, gwaves, mixed, qrup ? 4 : 0, n))) return rc;
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));

    /* compaction: colptr[0] = 0, colptr[1..n] = inclusive scan of the lengths */
    size_t cub_bytes = ctx->cub_bytes;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(ctx->d_cub, cub_bytes, ctx->d_m_len, ctx->d_Mptr + 1, n,
                                             ctx->stream));
    constexpr int kScatterCols = 256 / kScatterLanes;
    hipLaunchKernelGGL(scatter_columns_kernel, dim3((n + kScatterCols - 1) / kScatterCols), dim3(256), 0,
                       ctx->stream, n, cap, ctx->d_m_len, ctx->d_m_idx, ctx->d_m_val, ctx->d_Mptr, ctx->d_Mind,
                       ctx->d_Mval, ctx->d_counter + 63);
    if (mixed)
        hipLaunchKernelGGL(count_status_kernel, dim3(std::min(256, (n + 255) / 256)), dim3(256), 0, ctx->stream, n,
                           ctx->d_status, COL_FALLBACK, ctx->d_counter + 62);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    if ((rc = fetch_counters(ctx))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));

    const int last = ctx->h_counter[63];
    ctx->nnzM = last;
    ctx->pattern_nnz = last;
    ctx->m_csr_ok = false;
    ctx->last_mixed = mixed;
    ctx->mixed_fp64_last = mixed ? ctx->h_counter[62] : -1;
    ctx->last_binning = opts->binning;
    ctx->last_counters = opts->counters != 0;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float t_all = 0, t_solve = 0;
        HIPCHK(hipEventElapsedTime(&t_all, ctx->ev[0], ctx->ev[3]));
        HIPCHK(hipEventElapsedTime(&t_solve, ctx->ev[1], ctx->ev[2]));
        stats->build_ms = t_all;
        stats->solve_ms = t_solve;
        float t_first = 0;
        HIPCHK(hipEventElapsedTime(&t_first, ctx->ev[4], ctx->ev[5]));
        stats->first_bin_ms = t_first;
        stats->nnz = last;
        stats->ncols = n;
        if ((rc = read_chain_counts(ctx, (int)(mixed ? bins.size() : fbins.size()), mixed, stats))) return rc;
        stats->n_cache_hits = a.cache_slots > 0 ? ctx->h_counter[48] : 0;
    }
    return MSPAI_OK;
}

extern "C" int mspai_build_noalloc(mspai_ctx* ctx, const mspai_opts* opts, int* d_colptr, int* d_rowind,
                                   double* d_val, long long max_nnz, mspai_stats* stats)
{
    if (!ctx || !opts || !d_colptr || !d_rowind || !d_val) return MSPAI_ERR_ARG;
    if (ctx->ncols <= 0) return fail(ctx, MSPAI_ERR_STATE, "build_noalloc: upload the matrix first");
    if (!ctx->d_noalloc) return fail(ctx, MSPAI_ERR_STATE, "build_noalloc: no scratch address for this device");
    if (ctx->block_bs != 1 || opts->cache > 0 || opts->counters)
        return fail(ctx, MSPAI_ERR_UNSUPPORTED, "build_noalloc: not with block SPAI, the cache (-cs) or work counters");
    if (!ctx->d_p_ptr)
        return fail(ctx, MSPAI_ERR_STATE, "build_noalloc: set the start pattern first (mspai_set_pattern allocates)");
    int rc;
    if ((rc = check_opts(ctx, opts))) return rc;
    if ((rc = check_shapes(ctx))) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = ctx->ncols;
    const int cap = ctx->plen_max + opts->maxnew * opts->max_impr_steps;
    if (cap <= 0) return fail(ctx, MSPAI_ERR_ARG, "build: empty start pattern and no augmentation");
    if (max_nnz < (long long)(n * cap))
        return fail(ctx, MSPAI_ERR_CAPACITY, "build_noalloc: max_nnz must be at least ncols * (max |J0| + mn * ns) = " +
                                                 std::to_string(n * cap));
    /* carve the static array: the build's fixed buffers first, then the
     * tie buffers of the size bins, the scan workspace, and what is left
     * for the generic kernel's working sets */
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t at = off;
        off = align_up(off + bytes, 256);
        return ctx->d_noalloc + at;
    };
    const size_t fixed = align_up(64 * sizeof(int), 256) + align_up((2 * kMaxBins + 1) * n * sizeof(int), 256) +
                         align_up(n * sizeof(int), 256) + align_up(n * cap * sizeof(int), 256) +
                         align_up(n * cap * sizeof(double), 256) + align_up(n * sizeof(double), 256) +
                         2 * align_up(n * sizeof(int), 256);
    const bool adaptive = opts->max_impr_steps > 0;
    const bool qrup = opts->qr == 2 && adaptive && opts->maxnew > 0;
    const int needs = (qrup ? kNeedQrUpdate : 0) | (opts->precision == 1 && !qrup ? kNeedMixed : 0);
    const ScratchLayout G = generic_layout(ctx, cap, opts->maxnew, adaptive, needs);
    std::vector<Bin> bins, fbins;
    if (opts->binning && !ctx->d_up_ptr) {  // an upper pattern runs unbinned
        int t = table_for(ctx, opts);
        bins = bin_layouts(ctx, cap, opts->maxnew, adaptive, G, kTables[t], needs);
        fbins = (needs & kNeedMixed) ? bin_layouts(ctx, cap, opts->maxnew, adaptive, G, kTables[t],
                                                   needs & ~kNeedMixed)
                                     : bins;
    }
    const bool mixed = opts->precision == 1 && !bins.empty() && !qrup;
    const size_t tie = align_up(tie_bytes_needed(ctx, bins, fbins, mixed, qrup ? 4 : 0), 256);
    size_t cub = 0;
    hipcub::DeviceScan::InclusiveSum(nullptr, cub, (const int*)nullptr, (int*)nullptr, (int)n, ctx->stream);
    cub = align_up(std::max<size_t>(cub, 256), 256);
    if (fixed + tie + cub + G.per_wave > kNoAllocBytes)
        return fail(ctx, MSPAI_ERR_CAPACITY, "build_noalloc: the problem needs more than the " +
                                                 std::to_string(MSPAI_NOALLOC_SCRATCH_MIB) + " MiB scratch array");
    /* the context's own buffers step aside for the call */
    struct Saved {
        int *counter, *queues, *m_len, *m_idx, *nsolve, *status, *Mptr, *Mind;
        double *m_val, *res, *work, *Mval;
        char *tie, *scr;
        void* cub;
        size_t tie_bytes, cub_bytes, scr_bytes, Mcap, out_cols;
        int cap;
    } sv{ctx->d_counter, ctx->d_queues, ctx->d_m_len, ctx->d_m_idx, ctx->d_nsolve, ctx->d_status, ctx->d_Mptr,
         ctx->d_Mind, ctx->d_m_val, ctx->d_res, ctx->d_work, ctx->d_Mval, ctx->d_tie, ctx->d_scratch, ctx->d_cub,
         ctx->tie_bytes, ctx->cub_bytes, ctx->scratch_bytes, ctx->Mcap, ctx->out_cols, ctx->cap};
    ctx->d_counter = (int*)take(64 * sizeof(int));
    ctx->d_queues = (int*)take((2 * kMaxBins + 1) * n * sizeof(int));
    ctx->d_m_len = (int*)take(n * sizeof(int));
    ctx->d_m_idx = (int*)take(n * cap * sizeof(int));
    ctx->d_m_val = (double*)take(n * cap * sizeof(double));
    ctx->d_res = (double*)take(n * sizeof(double));
    ctx->d_nsolve = (int*)take(n * sizeof(int));
    ctx->d_status = (int*)take(n * sizeof(int));
    ctx->d_tie = take(tie);
    ctx->tie_bytes = tie;
    ctx->d_cub = take(cub);
    ctx->cub_bytes = cub;
    ctx->d_scratch = ctx->d_noalloc + off;
    ctx->scratch_bytes = kNoAllocBytes - off;
    ctx->d_work = nullptr;
    ctx->d_Mptr = d_colptr;
    ctx->d_Mind = d_rowind;
    ctx->d_Mval = d_val;
    ctx->Mcap = (size_t)max_nnz;
    ctx->cap = cap;
    ctx->out_cols = n;
    ctx->no_alloc = true;
    rc = mspai_build(ctx, opts, stats);
    /* back to the context's buffers; M lives in the caller's arrays */
    ctx->d_counter = sv.counter;
    ctx->d_queues = sv.queues;
    ctx->d_m_len = sv.m_len;
    ctx->d_m_idx = sv.m_idx;
    ctx->d_m_val = sv.m_val;
    ctx->d_res = sv.res;
    ctx->d_nsolve = sv.nsolve;
    ctx->d_status = sv.status;
    ctx->d_tie = sv.tie;
    ctx->tie_bytes = sv.tie_bytes;
    ctx->d_cub = sv.cub;
    ctx->cub_bytes = sv.cub_bytes;
    ctx->d_scratch = sv.scr;
    ctx->scratch_bytes = sv.scr_bytes;
    ctx->d_work = sv.work;
    ctx->d_Mptr = sv.Mptr;
    ctx->d_Mind = sv.Mind;
    ctx->d_Mval = sv.Mval;
    ctx->Mcap = sv.Mcap;
    ctx->cap = sv.cap;
    ctx->out_cols = sv.out_cols;
    ctx->no_alloc = false;
    ctx->nnzM = -1;
    ctx->pattern_nnz = -1;
    ctx->m_csr_ok = false;
    ctx->lwork_valid = false;
    return rc;
}

extern "C" int mspai_update_values_device(mspai_ctx* ctx, const double* d_val)
{
    if (!ctx || !d_val || !ctx->d_val) return MSPAI_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(ctx->d_val, d_val, ctx->nnz * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->nnzM = -1;
    ctx->mixed_fp64_last = -1;
    ctx->a_csr_ok = ctx->m_csr_ok = ctx->ajn_ok = false;
    if (ctx->block_bs != 1) return block_convert(ctx);
    return MSPAI_OK;
}

extern "C" int mspai_update_target_device(mspai_ctx* ctx, const double* d_val)
{
    if (!ctx || !d_val) return MSPAI_ERR_ARG;
    if (!ctx->d_b_ptr) return fail(ctx, MSPAI_ERR_STATE, "update_target: set a target first");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(ctx->d_b_val, d_val, ctx->nnzB * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->nnzM = -1;
    ctx->mixed_fp64_last = -1;
    return MSPAI_OK;
}

extern "C" int mspai_result_values_device(mspai_ctx* ctx, double* d_out)
{
    if (!ctx || !d_out) return MSPAI_ERR_ARG;
    if (ctx->nnzM < 0) return fail(ctx, MSPAI_ERR_STATE, "result: build first");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(d_out, ctx->d_Mval, ctx->nnzM * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSPAI_OK;
}

extern "C" int mspai_backward(mspai_ctx* ctx, const mspai_opts* opts, const double* d_gM, double* d_gC,
                              double* d_gB, mspai_stats* stats)
{
    if (!ctx || !opts || !d_gM || !d_gC) return MSPAI_ERR_ARG;
    if (ctx->block_bs != 1) return fail(ctx, MSPAI_ERR_UNSUPPORTED, "backward: not available for block SPAI");
    if (ctx->nnzM < 0) return fail(ctx, MSPAI_ERR_STATE, "backward: build M first (values unchanged since)");
    if (d_gB && !ctx->d_b_ptr) return fail(ctx, MSPAI_ERR_ARG, "backward: gB given but no target set");
    HIPCHK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = check_shapes(ctx))) return rc;
    const int n = ctx->ncols;
    const int cap = ctx->cap;
    const int needs = kNeedBackward | (opts->precision == 1 ? kNeedMixed : 0);
    const ScratchLayout G = generic_layout(ctx, cap, 1, false, needs);
    std::vector<Bin> bins;
    if (opts->binning) {
        int t = table_for(ctx, opts);
        bins = bin_layouts(ctx, cap, 1, false, G, kTables[t], needs);
    }
    int gwaves = 0;
    if ((rc = prepare_generic(ctx, G, &gwaves))) return rc;
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    HIPCHK(hipMemsetAsync(d_gC, 0, ctx->nnz * sizeof(double), ctx->stream));
    if (d_gB) HIPCHK(hipMemsetAsync(d_gB, 0, ctx->nnzB * sizeof(double), ctx->stream));
    BuildArgs a = base_args(ctx);
    a.p_ptr = ctx->d_Mptr;
    a.p_ind = ctx->d_Mind;
    a.m_in = ctx->d_Mval;
    a.gM = d_gM;
    a.gC = d_gC;
    a.gB = d_gB;
    a.cap = cap;
    const bool mixed = opts->precision == 1 && !bins.empty();
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    if ((rc = run_chain(ctx, a, bins, bins, G, gwaves, mixed, 2, n))) return rc;
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    if ((rc = fetch_counters(ctx))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float t_all = 0, t_solve = 0;
        HIPCHK(hipEventElapsedTime(&t_all, ctx->ev[0], ctx->ev[3]));
        float t_first = 0;
        HIPCHK(hipEventElapsedTime(&t_first, ctx->ev[4], ctx->ev[5]));
        stats->first_bin_ms = t_first;
        HIPCHK(hipEventElapsedTime(&t_solve, ctx->ev[1], ctx->ev[2]));
        stats->build_ms = t_all;
        stats->solve_ms = t_solve;
        stats->nnz = ctx->nnz;
        stats->ncols = n;
        if ((rc = read_chain_counts(ctx, (int)bins.size(), mixed, stats))) return rc;
    }
    return MSPAI_OK;
}

extern "C" int mspai_autotune(mspai_ctx* ctx, const mspai_opts* opts, int repeats, mspai_tune_report* report)
{
    if (!ctx || !opts) return MSPAI_ERR_ARG;
    if (ctx->block_bs != 1) return fail(ctx, MSPAI_ERR_UNSUPPORTED, "autotune: not available for block SPAI");
    if (repeats < 1) repeats = 1;
    mspai_opts o = *opts;
    o.binning = 2;
    int best = 0;
    double best_ms = 1e300;
    if (report) std::memset(report, 0, sizeof(*report));
    const int prec = o.precision == 1 ? 1 : 0;
    for (int t = 0; t < kNumTables && t < MSPAI_MAX_TUNE; ++t) {
        if ((prec == 1 && kTables[t].count == 0) || (prec == 0 && (kTables[t].mixed > 0 || kTables[t].fp64mask))) {
            /* the mixed kernels run in size bins: a bin-less table is all
             * FP64; a limit on the mixed chain changes nothing in FP64 */
            if (report) report->candidate_ms[t] = -1.0;
            continue;
        }
        ctx->tuned_table[prec] = t;
        double tmin = 1e300;
        for (int r = 0; r < repeats + 1; ++r) {  // first run warms up
            mspai_stats st;
            int rc = mspai_build(ctx, &o, &st);
            if (rc) {
                ctx->tuned_table[prec] = -1;
                return rc;
            }
            if (r > 0) tmin = std::min(tmin, st.solve_ms);
        }
        if (report) report->candidate_ms[t] = tmin;
        if (tmin < best_ms) {
            best_ms = tmin;
            best = t;
        }
    }
    ctx->tuned_table[prec] = best;
    if (report) {
        report->ncandidates = std::min(kNumTables, MSPAI_MAX_TUNE);
        report->chosen = best;
        std::snprintf(report->chosen_name, sizeof(report->chosen_name), "%s", kTables[best].name);
    }
    return MSPAI_OK;
}

extern "C" int mspai_set_bin_table(mspai_ctx* ctx, int precision, int table, char* name, int* ntables)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (ntables) *ntables = kNumTables;
    if (table < 0) return MSPAI_OK;  // the count only
    if (table >= kNumTables || precision < 0 || precision > 1)
        return fail(ctx, MSPAI_ERR_ARG, "set_bin_table: table or precision out of range");
    if (precision == 1 && kTables[table].count == 0)
        return fail(ctx, MSPAI_ERR_UNSUPPORTED, "set_bin_table: mixed builds run in size bins");
    ctx->tuned_table[precision] = table;
    if (name) std::snprintf(name, 32, "%s", kTables[table].name);
    return MSPAI_OK;
}

extern "C" int mspai_work_totals(mspai_ctx* ctx, double* totals)
{
    if (!ctx || !totals) return MSPAI_ERR_ARG;
    if (ctx->nnzM < 0) return fail(ctx, MSPAI_ERR_STATE, "work_totals: build first");
    if (!ctx->last_counters)
        return fail(ctx, MSPAI_ERR_STATE, "work_totals: the last build ran with counters off");
    HIPCHK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(work_sum_kernel, dim3(kNumWork), dim3(256), 0, ctx->stream, ctx->d_work, ctx->ncols,
                       ctx->d_wsum);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(totals, ctx->d_wsum, kNumWork * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSPAI_OK;
}

extern "C" int mspai_launch_stats(mspai_ctx* ctx, int max, int* nlaunch, mspai_launch_stat* out)
{
    if (!ctx || !nlaunch || (max > 0 && !out)) return MSPAI_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *nlaunch = ctx->nlaunch;
    std::vector<double> tot((size_t)kMaxLaunch * kLaunchStat, -1.0);
    if (ctx->lwork_valid)
        HIPCHK(hipMemcpy(tot.data(), ctx->d_lwork, tot.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < ctx->nlaunch && k < max; ++k) {
        mspai_launch_stat& o = out[k];
        std::memset(&o, 0, sizeof o);
        std::snprintf(o.name, sizeof o.name, "%s", ctx->lname[k]);
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ctx->lev[2 * k], ctx->lev[2 * k + 1]));
        o.ms = ms;
        const double* t = tot.data() + (size_t)k * kLaunchStat;
        for (int q = 0; q < kNumWork; ++q) o.work[q] = t[q];
        o.columns = t[kNumWork];
        for (int q = 0; q < 3; ++q) o.near[q] = t[kNumWork + 1 + q];
        o.nnz = t[kNumWork + 4];
    }
    return MSPAI_OK;
}

extern "C" int mspai_bin_counts(const mspai_ctx* ctx, int* nbins, int* routed, int* queued)
{
    if (!ctx || !nbins) return MSPAI_ERR_ARG;
    *nbins = ctx->last_nbins;
    for (int b = 0; b <= ctx->last_nbins && b < 8; ++b) {
        if (routed) routed[b] = ctx->last_routed[b];
        if (queued) queued[b] = ctx->last_queued[b];
    }
    return MSPAI_OK;
}

extern "C" int mspai_result_nnz(const mspai_ctx* ctx, long long* nnz)
{
    if (!ctx || !nnz) return MSPAI_ERR_ARG;
    if (ctx->nnzM < 0) return MSPAI_ERR_STATE;
    *nnz = ctx->nnzM;
    return MSPAI_OK;
}

extern "C" int mspai_download_csc(mspai_ctx* ctx, int* colptr, int* rowind, double* val)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (ctx->nnzM < 0) return fail(ctx, MSPAI_ERR_STATE, "download: build first");
    HIPCHK(hipSetDevice(ctx->device));
    if (colptr) HIPCHK(hipMemcpy(colptr, ctx->d_Mptr, (ctx->ncols + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (rowind) HIPCHK(hipMemcpy(rowind, ctx->d_Mind, ctx->nnzM * sizeof(int), hipMemcpyDeviceToHost));
    if (val) HIPCHK(hipMemcpy(val, ctx->d_Mval, ctx->nnzM * sizeof(double), hipMemcpyDeviceToHost));
    return MSPAI_OK;
}

extern "C" int mspai_download_column_info(mspai_ctx* ctx, double* res_norm, int* nsolve, int* status,
                                          double* work)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (ctx->nnzM < 0) return fail(ctx, MSPAI_ERR_STATE, "download: build first");
    HIPCHK(hipSetDevice(ctx->device));
    size_t n = ctx->ncols;
    if (res_norm) HIPCHK(hipMemcpy(res_norm, ctx->d_res, n * sizeof(double), hipMemcpyDeviceToHost));
    if (nsolve) HIPCHK(hipMemcpy(nsolve, ctx->d_nsolve, n * sizeof(int), hipMemcpyDeviceToHost));
    if (status) HIPCHK(hipMemcpy(status, ctx->d_status, n * sizeof(int), hipMemcpyDeviceToHost));
    if (work && !ctx->last_counters)
        return fail(ctx, MSPAI_ERR_STATE, "download: the last build ran with counters off");
    if (work) HIPCHK(hipMemcpy(work, ctx->d_work, n * kNumWork * sizeof(double), hipMemcpyDeviceToHost));
    return MSPAI_OK;
}

/* Device buffer of at least `need` elements, grown (never shrunk) in steps
 * of a quarter so that copies of similar size reuse it. */
template <typename T>
static int grow_buffer(mspai_ctx* ctx, T*& p, long long need, long long& cap)
{
    if (need <= cap && p) return MSPAI_OK;
    dfree(p);
    cap = 0;
    const long long want = std::max<long long>(need + need / 4, 1);
    HIPCHK(hipMalloc((void**)&p, want * sizeof(T)));
    cap = want;
    return MSPAI_OK;
}

/* CSR copy of a CSC matrix (nrows x ncols) on the stream: a stable radix
 * sort of the entries by row, taken in CSC order (columns ascending), leaves
 * each row's entries in increasing column order; the sorted rows give the
 * row pointers and a gather moves columns and values. */
static int build_csr(mspai_ctx* ctx, int nrows, int ncols, long long nnz, const int* colptr, const int* rowind,
                     const double* val, int*& rp, int*& ri, double*& rv, int& rows_cap, long long& nnz_cap,
                     hipStream_t st)
{
    int rc;
    long long rcap = rows_cap;
    if ((rc = grow_buffer(ctx, rp, (long long)nrows + 1, rcap))) return rc;
    rows_cap = (int)std::min<long long>(rcap, 0x7fffffff);
    long long ncap = nnz_cap;
    if (!ri || !rv || nnz > nnz_cap) {
        dfree(ri);
        dfree(rv);
        long long c1 = 0, c2 = 0;
        if ((rc = grow_buffer(ctx, ri, nnz, c1)) || (rc = grow_buffer(ctx, rv, nnz, c2))) return rc;
        ncap = std::min(c1, c2);
    }
    nnz_cap = ncap;
    if ((rc = grow_buffer(ctx, ctx->d_csr_tmp, 4 * nnz, ctx->csr_tmp_cap))) return rc;
    const long long tcap = ctx->csr_tmp_cap / 4;
    int* const col_of = ctx->d_csr_tmp;
    int* const order = col_of + tcap;
    unsigned* const rows_sorted = reinterpret_cast<unsigned*>(order + tcap);
    int* const perm = reinterpret_cast<int*>(rows_sorted + tcap);
    const int bits = 32 - __builtin_clz((unsigned)std::max(nrows - 1, 1));
    if (nnz > 0) {
        size_t need = 0;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, reinterpret_cast<const unsigned*>(rowind),
                                                  rows_sorted, order, perm, (int)nnz, 0, bits, st));
        if (need > ctx->cub_bytes) {
            if (ctx->d_cub) (void)hipFree(ctx->d_cub);
            ctx->d_cub = nullptr;
            ctx->cub_bytes = 0;
            HIPCHK(hipMalloc(&ctx->d_cub, need));
            ctx->cub_bytes = need;
        }
        launch_csr_entry_cols(colptr, ncols, col_of, order, st);
        size_t bytes = ctx->cub_bytes;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(ctx->d_cub, bytes, reinterpret_cast<const unsigned*>(rowind),
                                                  rows_sorted, order, perm, (int)nnz, 0, bits, st));
        launch_csr_gather(nnz, perm, col_of, val, ri, rv, st);
    }
    launch_csr_row_starts(nnz, nrows, rows_sorted, rp, st);
    HIPCHK(hipGetLastError());
    return MSPAI_OK;
}

extern "C" int mspai_apply(mspai_ctx* ctx, int which, int transpose, const double* d_x, double* d_y, void* stream)
{
    if (!ctx || !d_x || !d_y || (which != 0 && which != 1)) return MSPAI_ERR_ARG;
    if (ctx->ncols <= 0) return fail(ctx, MSPAI_ERR_STATE, "apply: upload the matrix first");
    if (which == 1 && ctx->nnzM < 0) return fail(ctx, MSPAI_ERR_STATE, "apply: build M first");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    const int n = ctx->ncols;
    const int nrows = which == 0 ? ctx->nrows : n;
    const int* cp = which == 0 ? ctx->d_colptr : ctx->d_Mptr;
    const int* ri = which == 0 ? ctx->d_rowind : ctx->d_Mind;
    const double* cv = which == 0 ? ctx->d_val : ctx->d_Mval;
    int rc;
    if (transpose) {
        launch_line_dot(n, cp, ri, cv, d_x, d_y, st);
    } else {
        bool& ok = which == 0 ? ctx->a_csr_ok : ctx->m_csr_ok;
        int*& rp = which == 0 ? ctx->d_acsr_ptr : ctx->d_mcsr_ptr;
        int*& rind = which == 0 ? ctx->d_acsr_ind : ctx->d_mcsr_ind;
        double*& rv = which == 0 ? ctx->d_acsr_val : ctx->d_mcsr_val;
        if (!ok) {
            const long long nnz = which == 0 ? ctx->nnz : ctx->nnzM;
            int& rcap = which == 0 ? ctx->acsr_rows_cap : ctx->mcsr_rows_cap;
            long long& ncap = which == 0 ? ctx->acsr_nnz_cap : ctx->mcsr_nnz_cap;
            /* the copy runs on the context's stream, after the caller's
             * work on `st` and before the product: its scan and sort share
             * the context's library workspace with the builds */
            if (st != ctx->stream) {
                HIPCHK(hipEventRecord(ctx->csr_ev, st));
                HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->csr_ev, 0));
            }
            if ((rc = build_csr(ctx, nrows, n, nnz, cp, ri, cv, rp, rind, rv, rcap, ncap, ctx->stream))) return rc;
            if (st != ctx->stream) {
                HIPCHK(hipEventRecord(ctx->csr_ev, ctx->stream));
                HIPCHK(hipStreamWaitEvent(st, ctx->csr_ev, 0));
            }
            ok = true;
        }
        launch_line_dot(nrows, rp, rind, rv, d_x, d_y, st);
    }
    HIPCHK(hipGetLastError());
    if (!stream) HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSPAI_OK;
}


/* ------------------------------------------------------------------ block SPAI */

namespace {

constexpr int kMaxBlock = 32;                        // largest block size of the device path
constexpr size_t kBlockScratchBudget = 8ull << 30;   // HBM for the block kernel's working sets

/* Initial_Run_Length (src/Matrix.imp:1430-1462): rows i, i+1, ... stored
 * consecutively in column i from its diagonal entry on (1 without one). */
int initial_run_length(const int* cp, const int* ri, int i)
{
    int k = cp[i];
    while (k < cp[i + 1] && ri[k] != i) ++k;
    if (k == cp[i + 1]) return 1;
    int bs = 1, ii = i + 1;
    for (++k; k < cp[i + 1]; ++k) {
        if (ri[k] != ii) return bs;
        ++bs;
        ++ii;
    }
    return bs;
}

/* Check_Next_Run (src/Matrix.imp:1464-1514): the run of column i_next from
 * row i must reach i_next's diagonal; then min(bs, run), else 0. */
int check_next_run(const int* cp, const int* ri, int i, int i_next, int bs)
{
    int k = cp[i_next];
    while (k < cp[i_next + 1] && ri[k] != i) ++k;
    if (k == cp[i_next + 1]) return 0;
    int nbs = 1, ii = i + 1;
    for (++k; k < cp[i_next + 1]; ++k) {
        if (ri[k] != ii) break;
        ++nbs;
        ++ii;
    }
    if (nbs > i_next - i) return nbs > bs ? bs : nbs;
    return 0;
}

/* Find_Diagonal_Block (src/Matrix.imp:1404-1428): last column of the block
 * that starts at column i. */
int find_diagonal_block(int n, const int* cp, const int* ri, int i, int limit)
{
    int max_bs = initial_run_length(cp, ri, i);
    int bs = max_bs;
    if (max_bs == 1) return i;
    if (limit && max_bs > limit) bs = max_bs = limit;
    for (int k = 0; k < max_bs; ++k) {
        if (k > 0 && i + k == n) return i + k - 1;
        bs = check_next_run(cp, ri, i, i + k, bs);
        if (!bs) return i + k - 1;
    }
    return i + max_bs - 1;
}

}  // namespace

extern "C" int mspai_block_sizes(int n, const int* colptr, const int* rowind, int bs, int* sizes, int* nblocks)
{
    if (n <= 0 || !sizes || !nblocks || bs < 0) return MSPAI_ERR_ARG;
    int nbk = 0;
    if (bs > 0) {
        /* Find_Constant_Blocks (src/Matrix.imp:1345-1363) */
        for (int i = 0; i < n / bs; ++i) sizes[nbk++] = bs;
        if (n % bs) sizes[nbk++] = n % bs;
    } else {
        /* Find_Diagonal_Blocks (src/Matrix.imp:1381-1402), upper limit of src/imspai.cpp:754-755 */
        if (!colptr || !rowind) return MSPAI_ERR_ARG;
        for (int i = 0; i < n;) {
            int iend = find_diagonal_block(n, colptr, rowind, i, 1000000);
            sizes[nbk++] = iend - i + 1;
            i = iend + 1;
        }
    }
    *nblocks = nbk;
    return MSPAI_OK;
}

/* Convert_To_Block_Matrix (src/Matrix.imp:1516-1745) of the uploaded matrix
 * with block sizes ctx->h_bsize: block CSC with zero-filled column-major
 * blocks, the block row pattern, and the offsets of (A_j^T A_j)^-1. */
static int block_convert(mspai_ctx* ctx)
{
    const int n = ctx->ncols, nb = (int)ctx->h_bsize.size();
    const long long nnz = ctx->nnz;
    std::vector<int> cp(n + 1), ri(std::max<long long>(nnz, 1));
    std::vector<double> v(std::max<long long>(nnz, 1));
    HIPCHK(hipMemcpy(cp.data(), ctx->d_colptr, (n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ri.data(), ctx->d_rowind, nnz * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(v.data(), ctx->d_val, nnz * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int> start(nb + 1, 0), bmap(n);
    int maxbs = 1;
    for (int b = 0; b < nb; ++b) {
        start[b + 1] = start[b] + ctx->h_bsize[b];
        maxbs = std::max(maxbs, ctx->h_bsize[b]);
        for (int i = start[b]; i < start[b + 1]; ++i) bmap[i] = b;
    }
    if (maxbs > kMaxBlock)
        return fail(ctx, MSPAI_ERR_UNSUPPORTED,
                    "set_block_size: block of " + std::to_string(maxbs) + " columns; the device path takes at most " +
                        std::to_string(kMaxBlock));
    std::vector<int> bcp(nb + 1, 0), bri, pos(nb, -1);
    std::vector<long long> voff(1, 0);
    std::vector<double> bv;
    int bcmax = 0;
    for (int jb = 0; jb < nb; ++jb) {
        const int w = ctx->h_bsize[jb];
        std::vector<int> rows;
        for (int j = start[jb]; j < start[jb + 1]; ++j)
            for (int e = cp[j]; e < cp[j + 1]; ++e) {
                int ib = bmap[ri[e]];
                if (pos[ib] < 0) {
                    pos[ib] = 0;
                    rows.push_back(ib);
                }
            }
        std::sort(rows.begin(), rows.end());
        const size_t base = bv.size();
        for (size_t t = 0; t < rows.size(); ++t) {
            const int ib = rows[t], h = ctx->h_bsize[ib];
            pos[ib] = (int)(voff.back() - (long long)base);
            bv.resize(bv.size() + (size_t)h * w, 0.0);
            voff.push_back(voff.back() + (long long)h * w);
            bri.push_back(ib);
        }
        for (int j = start[jb]; j < start[jb + 1]; ++j)
            for (int e = cp[j]; e < cp[j + 1]; ++e) {
                const int ib = bmap[ri[e]], h = ctx->h_bsize[ib];
                bv[base + pos[ib] + (size_t)(j - start[jb]) * h + (ri[e] - start[ib])] = v[e];
            }
        for (int ib : rows) pos[ib] = -1;
        bcp[jb + 1] = (int)bri.size();
        bcmax = std::max(bcmax, (int)rows.size());
    }
    /* block row structure: the transpose of the block pattern */
    const long long nnzb = bcp[nb];
    std::vector<int> rp(nb + 1, 0), rind(std::max<long long>(nnzb, 1));
    for (long long e = 0; e < nnzb; ++e) rp[bri[e] + 1]++;
    int brmax = 0;
    for (int b = 0; b < nb; ++b) {
        brmax = std::max(brmax, rp[b + 1]);
        rp[b + 1] += rp[b];
    }
    {
        std::vector<int> fill(rp.begin(), rp.end() - 1);
        for (int jb = 0; jb < nb; ++jb)
            for (int e = bcp[jb]; e < bcp[jb + 1]; ++e) rind[fill[bri[e]]++] = jb;
    }
    std::vector<long long> sqoff(nb + 1, 0);
    for (int b = 0; b < nb; ++b) sqoff[b + 1] = sqoff[b] + (long long)ctx->h_bsize[b] * ctx->h_bsize[b];
    ctx->h_bcolptr = bcp;
    ctx->h_browind = bri;
    if (bri.empty()) bri.push_back(0);  // uploads take at least one element
    if (bv.empty()) bv.push_back(0.0);
    int rc;
    if ((rc = upload(ctx, ctx->d_bsize, ctx->h_bsize.data(), nb))) return rc;
    if ((rc = upload(ctx, ctx->d_bstart, start.data(), nb + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_bcolptr, bcp.data(), nb + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_browind, bri.data(), bri.size()))) return rc;
    if ((rc = upload(ctx, ctx->d_bvoff, voff.data(), voff.size()))) return rc;
    if ((rc = upload(ctx, ctx->d_bval, bv.data(), bv.size()))) return rc;
    if ((rc = upload(ctx, ctx->d_brp_ptr, rp.data(), nb + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_brp_ind, rind.data(), nnzb))) return rc;
    if ((rc = upload(ctx, ctx->d_sqoff, sqoff.data(), nb + 1))) return rc;
    if ((rc = upload<double>(ctx, ctx->d_sqinv, nullptr, sqoff[nb]))) return rc;
    if ((rc = upload<double>(ctx, ctx->d_sqtmp, nullptr, sqoff[nb]))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->nb = nb;
    ctx->maxbs = maxbs;
    ctx->bcmax = bcmax;
    ctx->brmax = brmax;
    ctx->nnzb = nnzb;
    return MSPAI_OK;
}

extern "C" int mspai_set_block_size(mspai_ctx* ctx, int bs)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (ctx->ncols <= 0) return fail(ctx, MSPAI_ERR_STATE, "set_block_size: upload the matrix first");
    if (bs < 0) return fail(ctx, MSPAI_ERR_ARG, "Illegal block size (bs < 0)");
    HIPCHK(hipSetDevice(ctx->device));
    const bool was_block = ctx->block_bs != 1;
    if (bs == 1) {
        if (was_block) {
            block_free(ctx);
            dfree(ctx->d_p_ptr);
            dfree(ctx->d_p_ind);
            ctx->pattern_set = false;
        }
        ctx->nnzM = -1;
    ctx->mixed_fp64_last = -1;
        return MSPAI_OK;
    }
    if (ctx->nrows != ctx->ncols)
        return fail(ctx, MSPAI_ERR_UNSUPPORTED, "block SPAI needs a square matrix (no probing rows)");
    const int n = ctx->ncols;
    std::vector<int> sizes(n);
    int nbk = 0;
    std::vector<int> cp, ri;
    if (bs == 0) {
        cp.resize(n + 1);
        ri.resize(std::max<long long>(ctx->nnz, 1));
        HIPCHK(hipMemcpy(cp.data(), ctx->d_colptr, (n + 1) * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ri.data(), ctx->d_rowind, ctx->nnz * sizeof(int), hipMemcpyDeviceToHost));
    }
    int rc = mspai_block_sizes(n, cp.empty() ? nullptr : cp.data(), ri.empty() ? nullptr : ri.data(), bs,
                               sizes.data(), &nbk);
    if (rc) return fail(ctx, rc, "set_block_size: block detection failed");
    block_free(ctx);
    sizes.resize(nbk);
    ctx->h_bsize = sizes;
    if ((rc = block_convert(ctx))) {
        block_free(ctx);
        return rc;
    }
    ctx->block_bs = bs;
    /* a start pattern set before now is over scalar columns: back to the
     * block diagonal unless the caller sets a block pattern */
    dfree(ctx->d_p_ptr);
    dfree(ctx->d_p_ind);
    ctx->pattern_set = false;
    ctx->nnzM = -1;
    ctx->mixed_fp64_last = -1;
    ctx->pattern_nnz = -1;
    ctx->m_csr_ok = false;
    return MSPAI_OK;
}

extern "C" int mspai_block_info(const mspai_ctx* ctx, int* nblocks, long long* nnzb, int* block_sizes, int* colptr,
                                int* rowind)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (ctx->block_bs == 1) return MSPAI_ERR_STATE;
    if (nblocks) *nblocks = ctx->nb;
    if (nnzb) *nnzb = ctx->nnzb;
    if (block_sizes) std::copy(ctx->h_bsize.begin(), ctx->h_bsize.end(), block_sizes);
    if (colptr) std::copy(ctx->h_bcolptr.begin(), ctx->h_bcolptr.end(), colptr);
    if (rowind) std::copy(ctx->h_browind.begin(), ctx->h_browind.end(), rowind);
    return MSPAI_OK;
}

/* Spai<T>::SPAI_Algorithm with Spai_Unrestrained_Block (Switch_Algorithm,
 * src/Switch_Algorithm.imp:65-69) and Scalar_Matrix at the end
 * (src/imspai.cpp:858-870). */
static int build_block(mspai_ctx* ctx, const mspai_opts* opts, mspai_stats* stats)
{
    if (opts->precision != 0) return fail(ctx, MSPAI_ERR_UNSUPPORTED, "block SPAI runs in FP64 only");
    if (opts->qr != 0 || opts->cache != 0)
        return fail(ctx, MSPAI_ERR_UNSUPPORTED, "block SPAI takes neither QR updates (-qr) nor the cache (-cs)");
    if (ctx->d_b_ptr)
        return fail(ctx, MSPAI_ERR_UNSUPPORTED, "block SPAI builds against the identity target only");
    int rc;
    if (!ctx->d_p_ptr) {
        if ((rc = mspai_set_pattern(ctx, nullptr, nullptr))) return rc;
    }
    const int n = ctx->ncols, nb = ctx->nb, mb = ctx->maxbs;
    const int ns = opts->max_impr_steps;
    const int capJ = std::min<long long>(nb, (long long)ctx->plen_max + (long long)opts->maxnew * ns);
    if (capJ <= 0) return fail(ctx, MSPAI_ERR_ARG, "build: empty start pattern and no augmentation");
    const long long capX = (long long)capJ * mb * mb;
    BlockLayout L{};
    L.maxbs = mb;
    L.capJ = capJ;
    L.capJs = (int)std::min<long long>(n, (long long)capJ * mb);
    L.capI = (int)std::min<long long>(nb, (long long)capJ * std::max(1, ctx->bcmax));
    L.capIs = (int)std::min<long long>(n, (long long)L.capI * mb);
    const long long rawI = (long long)capJ * std::max(1, ctx->bcmax);
    const long long rawU = ns > 0 ? (long long)(L.capI + 1) * std::max(1, ctx->brmax) : 1;
    const long long rawmax = std::max<long long>(std::max(rawI, rawU), 64);
    if (rawmax > (1 << 26)) return fail(ctx, MSPAI_ERR_CAPACITY, "block build: index lists too long");
    L.capSort = (int)next_pow2((size_t)rawmax);
    L.capU = L.capSort;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    L.J = take(capJ * sizeof(int));
    L.I = take(L.capSort * sizeof(int));
    L.sbuf = take(L.capSort * sizeof(int));
    L.jt = take(L.capSort * sizeof(int));
    L.fr = take((L.capSort + 1) * sizeof(int));
    L.fc = take((capJ + 1) * sizeof(int));
    L.sel = take(capJ * sizeof(int));
    L.stack = take(3 * introsort_detail::kStackFrames * sizeof(int));
    L.pairs = take(L.capSort * sizeof(RhoPair));
    L.rho = take(L.capSort * sizeof(double));
    L.ahat = take((size_t)L.capIs * L.capJs * sizeof(double));
    L.aqr = take((size_t)L.capIs * L.capJs * sizeof(double));
    L.b = take((size_t)L.capIs * mb * sizeof(double));
    L.b0 = take((size_t)L.capIs * mb * sizeof(double));
    L.r = take((size_t)L.capIs * mb * sizeof(double));
    L.rb = take((size_t)L.capIs * mb * sizeof(double));
    L.x = take((size_t)L.capJs * mb * sizeof(double));
    L.nbuf = take((size_t)kWave * 2 * mb * mb * sizeof(double));
    L.per_wave = off;
    long long nwaves = std::min<long long>(nb, (long long)ctx->num_cu * 16);  // 4 per SIMD (the kernel's register budget)
    nwaves = std::min<long long>(nwaves, (long long)(kBlockScratchBudget / L.per_wave));
    if (nwaves < 1) return fail(ctx, MSPAI_ERR_CAPACITY, "block build: working set of one column exceeds the budget");
    const size_t sbytes = (size_t)nwaves * L.per_wave;
    if (sbytes > ctx->bscratch_bytes) {
        dfree(ctx->d_bscratch);
        ctx->bscratch_bytes = 0;
        HIPCHK(hipMalloc((void**)&ctx->d_bscratch, sbytes));
        ctx->bscratch_bytes = sbytes;
    }
    if ((rc = ensure_outputs(ctx, (int)std::min<long long>(n, (long long)capJ * mb)))) return rc;
    if (ctx->bcapJ != capJ || ctx->bcapX != capX || !ctx->d_bm_len) {
        dfree(ctx->d_bm_len);
        dfree(ctx->d_bm_idx);
        dfree(ctx->d_bm_val);
        HIPCHK(hipMalloc((void**)&ctx->d_bm_len, (size_t)nb * sizeof(int)));
        HIPCHK(hipMalloc((void**)&ctx->d_bm_idx, (size_t)nb * capJ * sizeof(int)));
        HIPCHK(hipMalloc((void**)&ctx->d_bm_val, (size_t)nb * capX * sizeof(double)));
        ctx->bcapJ = capJ;
        ctx->bcapX = capX;
    }
    BlockArgs a{};
    a.nb = nb;
    a.bsize = ctx->d_bsize;
    a.bstart = ctx->d_bstart;
    a.colptr = ctx->d_bcolptr;
    a.rowind = ctx->d_browind;
    a.voff = ctx->d_bvoff;
    a.val = ctx->d_bval;
    a.rp_ptr = ctx->d_brp_ptr;
    a.rp_ind = ctx->d_brp_ind;
    a.p_ptr = ctx->d_p_ptr;
    a.p_ind = ctx->d_p_ind;
    a.sqinv = ctx->d_sqinv;
    a.sqoff = ctx->d_sqoff;
    a.epsilon = opts->epsilon;
    a.maxnew = opts->maxnew;
    a.ns = ns;
    a.use_mean = opts->use_mean;
    a.capJ = capJ;
    a.capX = capX;
    a.m_len = ctx->d_bm_len;
    a.m_idx = ctx->d_bm_idx;
    a.m_val = ctx->d_bm_val;
    a.res = ctx->d_res;
    a.nsolve = ctx->d_nsolve;
    a.status = ctx->d_status;
    a.head = ctx->d_counter + 56;
    a.error = ctx->d_counter + 57;
    a.scratch = ctx->d_bscratch;

    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_counter + 56, 0, 2 * sizeof(int), ctx->stream));
    ctx->last_counters = false;  // the block kernel keeps no work counters
    ctx->nlaunch = 0;            // nor a per-launch profile
    ctx->lwork_valid = false;
    if (ns > 0) launch_block_sq_inverse(a, ctx->d_sqtmp, ctx->stream);
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    launch_block_columns(a, L, (int)nwaves, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    /* Scalar_Matrix: counts per scalar column, scan, fill */
    launch_block_scalar_count(a, ctx->d_m_len, ctx->stream);
    size_t cub_bytes = ctx->cub_bytes;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(ctx->d_cub, cub_bytes, ctx->d_m_len, ctx->d_Mptr + 1, n, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_Mptr, 0, sizeof(int), ctx->stream));
    launch_block_scalar_fill(a, ctx->d_Mptr, ctx->d_Mind, ctx->d_Mval, ctx->stream);
    HIPCHK(hipGetLastError());
    launch_copy_ints(ctx->d_Mptr + n, ctx->d_counter + 63, 1, ctx->stream);
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    if ((rc = fetch_counters(ctx))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->h_counter[57])
        return fail(ctx, MSPAI_ERR_CAPACITY, "block build: a block column outgrew its working set");
    const int last = ctx->h_counter[63];
    ctx->nnzM = last;
    ctx->pattern_nnz = last;
    ctx->m_csr_ok = false;
    ctx->block_built = true;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float t_all = 0, t_solve = 0;
        HIPCHK(hipEventElapsedTime(&t_all, ctx->ev[0], ctx->ev[3]));
        HIPCHK(hipEventElapsedTime(&t_solve, ctx->ev[1], ctx->ev[2]));
        stats->build_ms = t_all;
        stats->solve_ms = t_solve;
        stats->first_bin_ms = t_solve;  // the block column kernel
        stats->nnz = last;
        stats->ncols = nb;
    }
    return MSPAI_OK;
}
