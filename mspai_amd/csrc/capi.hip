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
#include <string>
#include <vector>

#include "../../include/mspai_gpu.h"
#include "device_common.h"
#include "spai_host.h"

using namespace mspai;

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

__global__ void iota_kernel(int* q, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = i;
}

__global__ void scatter_columns_kernel(int ncols, int cap, const int* m_len, const int* m_idx,
                                       const double* m_val, const int* colptr, int* rowind, double* val)
{
    int col = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (col >= ncols) return;
    int len = m_len[col], o = colptr[col];
    size_t src = (size_t)col * cap;
    for (int i = lane; i < len; i += 64) {
        rowind[o + i] = m_idx[src + i];
        val[o + i] = m_val[src + i];
    }
}

}  // namespace

struct mspai_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cu = 0;
    // matrix C
    int nrows = 0, ncols = 0;
    long long nnz = 0;
    int maxcolnnz = 0, maxrownnz = 0;
    int *d_colptr = nullptr, *d_rowind = nullptr;
    double* d_val = nullptr;
    int *d_rp_ptr = nullptr, *d_rp_ind = nullptr;
    double* d_ajn = nullptr;
    // start pattern
    int *d_p_ptr = nullptr, *d_p_ind = nullptr;
    int plen_max = 0;
    bool pattern_set = false;
    // target
    int *d_b_ptr = nullptr, *d_b_ind = nullptr;
    double* d_b_val = nullptr;
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
    long long nnzM = -1;
    size_t Mcap = 0;
    // queues and scratch
    int* d_queue = nullptr;
    int* d_counter = nullptr;  // [0] queue head, [1] overflow count
    int* d_overflow = nullptr;
    char* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    void* d_cub = nullptr;
    size_t cub_bytes = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
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
    dfree(d);
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void**)&d, n * sizeof(T)));
    if (h) HIPCHK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return MSPAI_OK;
}

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
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return MSPAI_ERR_HIP;
    }
    for (auto& e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return MSPAI_ERR_HIP;
        }
    if (hipMalloc((void**)&ctx->d_counter, 4 * sizeof(int)) != hipSuccess) {
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
    dfree(ctx->d_p_ptr);
    dfree(ctx->d_p_ind);
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
    dfree(ctx->d_queue);
    dfree(ctx->d_counter);
    dfree(ctx->d_overflow);
    dfree(ctx->d_scratch);
    if (ctx->d_cub) (void)hipFree(ctx->d_cub);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
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
    int rc;
    if ((rc = upload(ctx, ctx->d_colptr, colptr, ncols + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_rowind, rowind, nnz))) return rc;
    if ((rc = upload(ctx, ctx->d_val, val, nnz))) return rc;
    if ((rc = upload(ctx, ctx->d_rp_ptr, rowptr, nrows + 1))) return rc;
    if ((rc = upload(ctx, ctx->d_rp_ind, colind, nnz))) return rc;
    if ((rc = upload<double>(ctx, ctx->d_ajn, nullptr, ncols))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->nrows = nrows;
    ctx->ncols = ncols;
    ctx->nnz = nnz;
    ctx->maxcolnnz = maxc;
    ctx->maxrownnz = maxr;
    ctx->nnzM = -1;
    if (!ctx->pattern_set) {
        dfree(ctx->d_p_ptr);
        dfree(ctx->d_p_ind);
    }
    return MSPAI_OK;
}

extern "C" int mspai_update_values(mspai_ctx* ctx, const double* val)
{
    if (!ctx || !val || !ctx->d_val) return MSPAI_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(ctx->d_val, val, ctx->nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->nnzM = -1;
    return MSPAI_OK;
}

extern "C" int mspai_set_pattern(mspai_ctx* ctx, const int* colptr, const int* rowind)
{
    if (!ctx) return MSPAI_ERR_ARG;
    if (ctx->ncols <= 0) return fail(ctx, MSPAI_ERR_STATE, "set_pattern: upload the matrix first");
    HIPCHK(hipSetDevice(ctx->device));
    int n = ctx->ncols;
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
    ctx->pattern_set = true;
    return MSPAI_OK;
}

extern "C" int mspai_set_target(mspai_ctx* ctx, const int* colptr, const int* rowind, const double* val)
{
    if (!ctx) return MSPAI_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    if (!colptr) {
        dfree(ctx->d_b_ptr);
        dfree(ctx->d_b_ind);
        dfree(ctx->d_b_val);
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
    dfree(ctx->d_queue);
    dfree(ctx->d_overflow);
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
    HIPCHK(hipMalloc((void**)&ctx->d_queue, n * sizeof(int)));
    HIPCHK(hipMalloc((void**)&ctx->d_overflow, n * sizeof(int)));
    ctx->cap = cap;
    ctx->out_cols = n;
    ctx->Mcap = n * cap;
    /* scan workspace for the compaction */
    size_t need = 0;
    hipcub::DeviceScan::InclusiveSum(nullptr, need, ctx->d_m_len, ctx->d_Mptr + 1, (int)n, ctx->stream);
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

/* Size the generic kernel's per-wave scratch from worst-case bounds:
 * |J| <= cap, |I| <= min(nrows, cap * max nnz per column),
 * raw candidate list <= (|I| + 1) * max nnz per row. */
static ScratchLayout generic_layout(const mspai_ctx* ctx, int cap, int maxnew, bool adaptive)
{
    ScratchLayout L{};
    size_t rawI = (size_t)cap * std::max(1, ctx->maxcolnnz);
    size_t capI = std::min<size_t>(ctx->nrows, rawI);
    capI = std::max<size_t>(capI, cap);
    size_t rawU = adaptive ? (capI + 1) * std::max(1, ctx->maxrownnz) : 1;
    size_t capU = adaptive ? std::min<size_t>(ctx->ncols, rawU) : 1;
    size_t capSort = next_pow2(std::max(rawI, rawU));
    L.capJ = cap;
    L.capI = (int)capI;
    L.capU = (int)capU;
    L.capSort = (int)capSort;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    L.J = take((size_t)cap * 4 + 64);
    L.I = take(capI * 4);
    L.sbuf = take(capSort * 4);
    L.ubuf = take(capU * 4);
    L.ahat = take(capI * cap * 8);
    L.aqr = take(capI * cap * 8);
    L.b = take(capI * 8);
    L.b0 = take(capI * 8);
    L.r = take(capI * 8);
    L.tau = take((size_t)cap * 8);
    L.jt = take(capU * 4);
    L.rho = take(capU * 8);
    L.pairs = take(capU * sizeof(RhoPair));
    L.sel = take((size_t)std::max(1, maxnew) * 4);
    L.per_wave = o;
    return L;
}

extern "C" int mspai_build(mspai_ctx* ctx, const mspai_opts* opts, mspai_stats* stats)
{
    if (!ctx || !opts) return MSPAI_ERR_ARG;
    if (ctx->ncols <= 0) return fail(ctx, MSPAI_ERR_STATE, "build: upload the matrix first");
    int rc;
    if ((rc = check_opts(ctx, opts))) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->d_p_ptr) {
        if ((rc = mspai_set_pattern(ctx, nullptr, nullptr))) return rc;
    }
    const int n = ctx->ncols;
    const int ns = opts->max_impr_steps;
    const int cap = ctx->plen_max + opts->maxnew * ns;
    if (cap <= 0) return fail(ctx, MSPAI_ERR_ARG, "build: empty start pattern and no augmentation");
    if ((rc = ensure_outputs(ctx, cap))) return rc;
    const bool adaptive = ns > 0;
    ScratchLayout L = generic_layout(ctx, cap, opts->maxnew, adaptive);
    int nwaves = std::max(1, ctx->num_cu) * 8;
    /* keep the scratch within 16 GiB; fewer resident waves if needed */
    const size_t budget = (size_t)16 << 30;
    while (nwaves > ctx->num_cu && (size_t)nwaves * L.per_wave > budget) nwaves /= 2;
    size_t need = (size_t)nwaves * L.per_wave;
    if (need > budget) return fail(ctx, MSPAI_ERR_CAPACITY, "build: per-column working set too large");
    if (need > ctx->scratch_bytes) {
        dfree(ctx->d_scratch);
        HIPCHK(hipMalloc((void**)&ctx->d_scratch, need));
        ctx->scratch_bytes = need;
    }

    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    DevCsc A{ctx->nrows, ctx->ncols, ctx->d_colptr, ctx->d_rowind, ctx->d_val};
    hipLaunchKernelGGL(column_norms_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, A, ctx->d_ajn);
    hipLaunchKernelGGL(iota_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_queue, n);
    HIPCHK(hipMemsetAsync(ctx->d_counter, 0, 4 * sizeof(int), ctx->stream));

    BuildArgs a{};
    a.A = A;
    a.rp_ptr = ctx->d_rp_ptr;
    a.rp_ind = ctx->d_rp_ind;
    a.p_ptr = ctx->d_p_ptr;
    a.p_ind = ctx->d_p_ind;
    a.B = DevCsc{ctx->nrows, ctx->ncols, ctx->d_b_ptr, ctx->d_b_ind, ctx->d_b_val};
    a.ajn = ctx->d_ajn;
    a.epsilon = opts->epsilon;
    a.maxnew = opts->maxnew;
    a.ns = ns;
    a.use_mean = opts->use_mean;
    a.cap = cap;
    a.queue = ctx->d_queue;
    a.nqueue = n;
    a.counter = ctx->d_counter;
    a.overflow = ctx->d_overflow;
    a.n_overflow = ctx->d_counter + 1;
    a.m_len = ctx->d_m_len;
    a.m_idx = ctx->d_m_idx;
    a.m_val = ctx->d_m_val;
    a.res = ctx->d_res;
    a.nsolve = ctx->d_nsolve;
    a.status = ctx->d_status;
    a.work = ctx->d_work;
    a.scratch = ctx->d_scratch;
    a.L = L;

    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    launch_generic_fp64(a, nwaves, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));

    /* compaction: colptr[0] = 0, colptr[1..n] = inclusive scan of the lengths */
    size_t cub_bytes = ctx->cub_bytes;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(ctx->d_cub, cub_bytes, ctx->d_m_len, ctx->d_Mptr + 1, n,
                                             ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_Mptr, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(scatter_columns_kernel, dim3((n + 3) / 4), dim3(256), 0, ctx->stream, n, cap,
                       ctx->d_m_len, ctx->d_m_idx, ctx->d_m_val, ctx->d_Mptr, ctx->d_Mind, ctx->d_Mval);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));

    int last = 0;
    HIPCHK(hipMemcpy(&last, ctx->d_Mptr + n, sizeof(int), hipMemcpyDeviceToHost));
    ctx->nnzM = last;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float t_all = 0, t_solve = 0;
        HIPCHK(hipEventElapsedTime(&t_all, ctx->ev[0], ctx->ev[3]));
        HIPCHK(hipEventElapsedTime(&t_solve, ctx->ev[1], ctx->ev[2]));
        stats->build_ms = t_all;
        stats->solve_ms = t_solve;
        stats->nnz = last;
        stats->ncols = n;
        std::vector<double> w((size_t)n * kNumWork);
        HIPCHK(hipMemcpy(w.data(), ctx->d_work, w.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int q = 0; q < kNumWork; ++q) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += w[(size_t)j * kNumWork + q];
            stats->work[q] = s;
        }
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
    if (work) HIPCHK(hipMemcpy(work, ctx->d_work, n * kNumWork * sizeof(double), hipMemcpyDeviceToHost));
    return MSPAI_OK;
}
