/*
 * mspai_gpu: stand-alone builder in the shape of the reference's mspai
 * command line (options of doc/mspai-1.2.pdf section 2.3, defaults of
 * Command_Line::Parameters, src/Command_Line.cpp:192-226):
 *
 *   mspai_gpu <matrix.mtx> [-ep eps] [-mn maxnew] [-ns steps] [-um 0|1]
 *             [-qr 0|2] [-cs entries] [-ch 0|1] [-bs size] [-pp 1|2]
 *             [-out file] [-vb 0|1] [-prec fp64|mixed] [-bins 0|1|2]
 *
 * -pp 1: diagonal start pattern (default); -pp 2: pattern of A (with -bs:
 * of the block matrix).  -bs: block SPAI, bs > 1 constant blocks, 0 variable
 * blocks (src/imspai.cpp:208-210, 746-760).  -ch 1: the hash-table variant
 * (Spai_Update, src/Spai_Update.imp:45-146), whose M is the static solve on
 * the start pattern -- the stored factors it keeps between setups never reach
 * M, since the second Insert_Row_Solution overwrites the first.  The
 * preconditioner M with A M ~ I is written in the reference's MatrixMarket
 * layout (Write_Matrix_To_File), default file precond.mtx.  Errors are
 * reported on stderr with exit status 1, as the reference's
 * std::runtime_error messages are.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mspai_gpu.h"
#include "mmio.h"

namespace {

void check(int rc, mspai_ctx* ctx, const char* what)
{
    if (rc != MSPAI_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                                 "): " + (ctx ? mspai_last_error(ctx) : ""));
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr,
                     "usage: %s <matrix.mtx> [-ep eps] [-mn maxnew] [-ns steps] [-um 0|1] [-qr 0|2] "
                     "[-cs entries] [-ch 0|1] [-bs size] [-pp 1|2] [-out file] [-vb 0|1] [-prec fp64|mixed] "
                     "[-bins 0|1|2]\n",
                     argv[0]);
        return 1;
    }
    mspai_ctx* ctx = nullptr;
    try {
        mspai_opts o;
        mspai_default_opts(&o);
        std::string out = "precond.mtx";
        int pp = 1, ch = 0, bs = 1;
        for (int k = 2; k < argc; ++k) {
            std::string key = argv[k];
            if (k + 1 >= argc) throw std::runtime_error("option " + key + " needs a value");
            const char* v = argv[++k];
            if (key == "-ep") o.epsilon = std::atof(v);
            else if (key == "-mn") o.maxnew = std::atoi(v);
            else if (key == "-ns") o.max_impr_steps = std::atoi(v);
            else if (key == "-um") o.use_mean = std::atoi(v);
            else if (key == "-qr") o.qr = std::atoi(v);
            else if (key == "-cs") o.cache = std::atoi(v);
            else if (key == "-ch") ch = std::atoi(v);
            else if (key == "-bs") bs = std::atoi(v);
            else if (key == "-pp") pp = std::atoi(v);
            else if (key == "-out") out = v;
            else if (key == "-vb") o.verbose = std::atoi(v);
            else if (key == "-prec") o.precision = std::strcmp(v, "mixed") == 0 ? 1 : 0;
            else if (key == "-bins") o.binning = std::atoi(v);
            else throw std::runtime_error("unknown option " + key);
        }
        if (pp != 1 && pp != 2) throw std::runtime_error("-pp: only 1 (diagonal) and 2 (pattern of A) are built");
        /* option checks of src/imspai.cpp:280-313 */
        if (ch < 0 || ch > 1)
            throw std::runtime_error("Illegal hash size parameter: use 1 for using a hashtable and 0 for not using one");
        if (o.cache > 0 && ch > 0)
            throw std::runtime_error("Caching and Hashing cannot be used simultaneously");
        if (ch > 0 && o.qr > 0) throw std::runtime_error("QR-Updates and Hashing optimization cannot be used simultaneously");
        if (bs < 0) throw std::runtime_error("Illegal block size (bs < 0)");
        if (ch > 0) o.max_impr_steps = 0;  // Spai_Update: static solve on the start pattern
        mspai::CscMatrix A = mspai::read_matrix_market(argv[1]);
        check(mspai_ctx_create(0, &ctx), ctx, "mspai_ctx_create");
        check(mspai_upload_csc(ctx, A.nrows, A.ncols, A.colptr.data(), A.rowind.data(), A.val.data(), nullptr,
                               nullptr),
              ctx, "mspai_upload_csc");
        if (bs != 1) check(mspai_set_block_size(ctx, bs), ctx, "mspai_set_block_size");
        if (pp == 2 && bs != 1) {
            int nb = 0;
            long long nnzb = 0;
            check(mspai_block_info(ctx, &nb, &nnzb, nullptr, nullptr, nullptr), ctx, "mspai_block_info");
            std::vector<int> bp(nb + 1), bi(nnzb > 0 ? nnzb : 1);
            check(mspai_block_info(ctx, nullptr, nullptr, nullptr, bp.data(), bi.data()), ctx, "mspai_block_info");
            check(mspai_set_pattern(ctx, bp.data(), bi.data()), ctx, "mspai_set_pattern");
        } else if (pp == 2) {
            check(mspai_set_pattern(ctx, A.colptr.data(), A.rowind.data()), ctx, "mspai_set_pattern");
        } else {
            check(mspai_set_pattern(ctx, nullptr, nullptr), ctx, "mspai_set_pattern");
        }
        if (o.binning == 2 && bs == 1) check(mspai_autotune(ctx, &o, 1, nullptr), ctx, "mspai_autotune");
        mspai_stats st;
        check(mspai_build(ctx, &o, &st), ctx, "mspai_build");
        long long nnz = 0;
        check(mspai_result_nnz(ctx, &nnz), ctx, "mspai_result_nnz");
        mspai::CscMatrix M;
        M.nrows = A.ncols;
        M.ncols = A.ncols;
        M.colptr.resize(A.ncols + 1);
        M.rowind.resize(nnz);
        M.val.resize(nnz);
        check(mspai_download_csc(ctx, M.colptr.data(), M.rowind.data(), M.val.data()), ctx, "mspai_download_csc");
        mspai::write_preconditioner(out, M);
        if (o.verbose)
            std::printf("n=%d nnz(A)=%d nnz(M)=%lld build %.3f ms (columns %.3f ms), overflow %d, fp64 %d\n",
                        A.ncols, A.colptr[A.ncols], nnz, st.build_ms, st.solve_ms, st.n_overflow, st.n_fallback);
        mspai_ctx_destroy(ctx);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "mspai_gpu: %s\n", e.what());
        if (ctx) mspai_ctx_destroy(ctx);
        return 1;
    }
}
