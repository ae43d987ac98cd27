/*
 * mspai_gpu: stand-alone builder in the shape of the reference's mspai
 * command line (doc/mspai-1.2.pdf sections 2.2-2.4; defaults of
 * Command_Line::Parameters, src/Command_Line.cpp:192-226):
 *
 *   mspai_gpu <matrix.mtx> [<pattern.mtx|-1|-2> [<target.mtx|-1>]]
 *             [-ep eps] [-mn maxnew] [-ns steps] [-um 0|1] [-qr 0|2]
 *             [-cs entries] [-ch 0|1] [-bs size] [-pp 1|2] [-Ce file]
 *             [-Be file] [-rho r] [-schur 0] [-out file] [-vb 0|1]
 *             [-prec fp64|mixed] [-bins 0|1|2]
 *
 * Start pattern: a MatrixMarket pattern file (column j lists J0 of column
 * j), -1 the diagonal (default), -2 the pattern of A without probing rows
 * (Matrix::To_Pattern, src/Matrix.imp:285-356); -pp 1|2 is the same choice
 * as an option.  Target: a square MatrixMarket file, or -1 for the identity
 * (default).  -Ce/-Be: probing rows appended below C and B, scaled by -rho
 * (Read_Matrix_File / Read_Target_File, src/Read_mm_Matrix.cpp:115-420).
 * -bs: block SPAI, bs > 1 constant blocks, 0 variable blocks
 * (src/imspai.cpp:208-210, 746-760).  -ch 1: the hash-table variant
 * (Spai_Update, src/Spai_Update.imp:45-146), whose M is the static solve on
 * the start pattern -- the stored factors it keeps between setups never reach
 * M, since the second Insert_Row_Solution overwrites the first.  The
 * preconditioner M with C M ~ B is written in the reference's MatrixMarket
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

/* The option checks of PC_MSPAI::PCSetFromOptions_SPAI
 * (src/imspai.cpp:280-313, 345-352, 367-395, 442-465) that concern this
 * path; `target` is true when a target file is given. */
void check_options(const mspai_opts& o, int ch, int bs, double rho, int schur, bool ce, bool be, bool target,
                   bool fixed_pattern)
{
    if (ch < 0 || ch > 1)
        throw std::runtime_error("Illegal hash size parameter: use 1 for using a hashtable and 0 for not using one");
    if (o.cache > 0 && ch > 0) throw std::runtime_error("Caching and Hashing cannot be used simultaneously");
    if (ch > 0 && o.qr > 0) throw std::runtime_error("QR-Updates and Hashing optimization cannot be used simultaneously");
    if (bs < 0) throw std::runtime_error("Illegal block size (bs < 0)");
    if (rho < 0.0)
        throw std::runtime_error("Illegal parameter for rho: " + std::to_string(rho) +
                                 " Allowed are: 0.0 <= rho (default is rho = 1.0)");
    if (schur < 0 || schur > 1)
        throw std::runtime_error("Illegal value for using schur value: " + std::to_string(schur) +
                                 " Allowed are: 0: don't use schur probing (default), 1: use schur probing");
    if (schur && target)
        throw std::runtime_error("Arbitrary target matrices are not allowed when using schur probing!");
    if (schur && !fixed_pattern)
        throw std::runtime_error("No identity pattern or pattern filled where A has nnz is allowed when using "
                                 "schur probing! Use a valid start pattern file instead.");
    if ((ce || be || target) && (o.max_impr_steps > 0 || o.qr > 0))
        throw std::runtime_error("Probing constraints and/or target matrix approximations are only supported in "
                                 "static MSPAI yet! Please set the number of improvement steps to 0.");
    if (ce != be)
        throw std::runtime_error("Only one probing condition defined! Please set a probing condition for C (-Ce) "
                                 "and a probing condition for B (-Be)!");
    if (!ce && !be && schur)
        throw std::runtime_error("No probing condition defined but schur probing requested! Please set a probing "
                                 "condition for C (-Ce) and a probing condition for B (-Be)!");
    if (schur)
        throw std::runtime_error("schur probing (a preconditioner with fewer columns than C) is not built by "
                                 "mspai_gpu");
    if (bs != 1 && (ce || target))
        throw std::runtime_error("block SPAI (-bs) runs without probing rows or a target matrix");
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr,
                     "usage: %s <matrix.mtx> [<pattern.mtx|-1|-2> [<target.mtx|-1>]] [-ep eps] [-mn maxnew] "
                     "[-ns steps] [-um 0|1] [-qr 0|2] [-cs entries] [-ch 0|1] [-bs size] [-pp 1|2] [-Ce file] "
                     "[-Be file] [-rho r] [-schur 0] [-up file|1|2] [-out file] [-vb 0|1] [-prec fp64|mixed] "
                     "[-bins 0|1|2]\n",
                     argv[0]);
        return 1;
    }
    mspai_ctx* ctx = nullptr;
    try {
        mspai_opts o;
        mspai_default_opts(&o);
        std::string out = "precond.mtx", pattern_file, target_file, ce_file, be_file, up_arg;
        int pp = 1, ch = 0, bs = 1, schur = 0;
        double rho = 1.0;
        int k = 2;
        /* positional start pattern and target (doc/mspai-1.2.pdf section 2.2) */
        for (int pos = 0; pos < 2 && k < argc; ++pos, ++k) {
            std::string v = argv[k];
            if (v == "-1" || v == "-2") {
                if (pos == 0) pp = v == "-1" ? 1 : 2;
                else if (v == "-2") throw std::runtime_error("target: a file, or -1 for the identity");
            } else if (!v.empty() && v[0] != '-') {
                (pos == 0 ? pattern_file : target_file) = v;
            } else {
                break;
            }
        }
        for (; k < argc; ++k) {
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
            else if (key == "-Ce") ce_file = v;
            else if (key == "-Be") be_file = v;
            else if (key == "-rho") rho = std::atof(v);
            else if (key == "-schur") schur = std::atoi(v);
            else if (key == "-up") up_arg = v;
            else if (key == "-out") out = v;
            else if (key == "-vb") o.verbose = std::atoi(v);
            else if (key == "-prec") o.precision = std::strcmp(v, "mixed") == 0 ? 1 : 0;
            else if (key == "-bins") o.binning = std::atoi(v);
            else throw std::runtime_error("unknown option " + key);
        }
        if (pp != 1 && pp != 2) throw std::runtime_error("-pp: only 1 (diagonal) and 2 (pattern of A) are built");
        check_options(o, ch, bs, rho, schur, !ce_file.empty(), !be_file.empty(), !target_file.empty(),
                      !pattern_file.empty());
        if (ch > 0) o.max_impr_steps = 0;  // Spai_Update: static solve on the start pattern
        mspai::CscMatrix A = mspai::read_system(argv[1], ce_file, rho);
        const int n = A.ncols;
        const bool targeted = !be_file.empty() || !target_file.empty();
        mspai::CscMatrix B;
        if (targeted) B = mspai::read_target(A, target_file, be_file, rho);
        mspai::CscMatrix P;
        if (!pattern_file.empty()) {
            P = mspai::read_matrix_market(pattern_file);
            if (bs == 1 && P.ncols != n)
                throw std::runtime_error("start pattern must have as many columns as the system matrix");
        }
        check(mspai_ctx_create(0, &ctx), ctx, "mspai_ctx_create");
        check(mspai_upload_csc(ctx, A.nrows, A.ncols, A.colptr.data(), A.rowind.data(), A.val.data(), nullptr,
                               nullptr),
              ctx, "mspai_upload_csc");
        if (bs != 1) check(mspai_set_block_size(ctx, bs), ctx, "mspai_set_block_size");
        if (!pattern_file.empty()) {
            check(mspai_set_pattern(ctx, P.colptr.data(), P.rowind.data()), ctx, "mspai_set_pattern");
        } else if (pp == 2 && bs != 1) {
            int nb = 0;
            long long nnzb = 0;
            check(mspai_block_info(ctx, &nb, &nnzb, nullptr, nullptr, nullptr), ctx, "mspai_block_info");
            std::vector<int> bp(nb + 1), bi(nnzb > 0 ? nnzb : 1);
            check(mspai_block_info(ctx, nullptr, nullptr, nullptr, bp.data(), bi.data()), ctx, "mspai_block_info");
            check(mspai_set_pattern(ctx, bp.data(), bi.data()), ctx, "mspai_set_pattern");
        } else if (pp == 2) {
            /* To_Pattern: the entries of C in the probing rows are left out */
            std::vector<int> pc(n + 1, 0), pi;
            pi.reserve(A.rowind.size());
            for (int j = 0; j < n; ++j) {
                for (int e = A.colptr[j]; e < A.colptr[j + 1] && A.rowind[e] < n; ++e) pi.push_back(A.rowind[e]);
                pc[j + 1] = (int)pi.size();
            }
            if (pi.empty()) pi.push_back(0);
            check(mspai_set_pattern(ctx, pc.data(), pi.data()), ctx, "mspai_set_pattern");
        } else {
            check(mspai_set_pattern(ctx, nullptr, nullptr), ctx, "mspai_set_pattern");
        }
        if (targeted)
            check(mspai_set_target(ctx, B.colptr.data(), B.rowind.data(), B.val.data()), ctx, "mspai_set_target");
        if (!up_arg.empty()) {
            /* -up: a MatrixMarket pattern file, 1 (diagonal) or 2 (the pattern
             * of A), Generate_Pattern in src/imspai.cpp:797-810 */
            if (bs != 1) throw std::runtime_error("-up: scalar SPAI only");
            std::vector<int> uc(n + 1, 0), ui;
            if (up_arg == "1") {
                for (int j = 0; j < n; ++j) {
                    ui.push_back(j);
                    uc[j + 1] = j + 1;
                }
            } else if (up_arg == "2") {
                for (int j = 0; j < n; ++j) {
                    for (int e = A.colptr[j]; e < A.colptr[j + 1] && A.rowind[e] < n; ++e) ui.push_back(A.rowind[e]);
                    uc[j + 1] = (int)ui.size();
                }
            } else {
                mspai::CscMatrix U = mspai::read_matrix_market(up_arg);
                if (U.ncols != n) throw std::runtime_error("-up: the pattern needs as many columns as the matrix");
                uc = U.colptr;
                ui = U.rowind;
            }
            if (ui.empty()) ui.push_back(0);
            check(mspai_set_upper_pattern(ctx, uc.data(), ui.data()), ctx, "mspai_set_upper_pattern");
        }
        if (o.binning == 2 && bs == 1) check(mspai_autotune(ctx, &o, 1, nullptr), ctx, "mspai_autotune");
        mspai_stats st;
        check(mspai_build(ctx, &o, &st), ctx, "mspai_build");
        long long nnz = 0;
        check(mspai_result_nnz(ctx, &nnz), ctx, "mspai_result_nnz");
        mspai::CscMatrix M;
        M.nrows = n;
        M.ncols = n;
        M.colptr.resize(n + 1);
        M.rowind.resize(nnz);
        M.val.resize(nnz);
        check(mspai_download_csc(ctx, M.colptr.data(), M.rowind.data(), M.val.data()), ctx, "mspai_download_csc");
        mspai::write_preconditioner(out, M);
        if (o.verbose)
            std::printf("n=%d nnz(C)=%d nnz(M)=%lld build %.3f ms (columns %.3f ms), overflow %d, fp64 %d\n", n,
                        A.colptr[n], nnz, st.build_ms, st.solve_ms, st.n_overflow, st.n_fallback);
        mspai_ctx_destroy(ctx);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "mspai_gpu: %s\n", e.what());
        if (ctx) mspai_ctx_destroy(ctx);
        return 1;
    }
}
