// Test helper for tests/test_mmio.py (mspai_amd/host/mmio.cpp):
//   mmio_roundtrip <in.mtx> <out.mtx>
//       reads a MatrixMarket file with the library's reader and writes it
//       back with the reference-format writer;
//   mmio_roundtrip --probe <A.mtx> <Ce.mtx|-> <Be.mtx|-> <rho> <T.mtx|-> <C.txt> <B.txt>
//       reads the system with probing rows and the target, and writes C and B
//       as "nrows ncols nnz" then one "row col value" line per entry
//       (0-based, column order, %.17g so that values round-trip exactly).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

#include "mmio.h"

static void dump(const char* path, const mspai::CscMatrix& M)
{
    FILE* f = std::fopen(path, "w");
    std::fprintf(f, "%d %d %d\n", M.nrows, M.ncols, M.colptr[M.ncols]);
    for (int j = 0; j < M.ncols; ++j)
        for (int e = M.colptr[j]; e < M.colptr[j + 1]; ++e) std::fprintf(f, "%d %d %.17g\n", M.rowind[e], j, M.val[e]);
    std::fclose(f);
}

int main(int argc, char** argv)
{
    try {
        if (argc == 9 && std::strcmp(argv[1], "--probe") == 0) {
            auto opt = [](const char* s) { return std::string(std::strcmp(s, "-") == 0 ? "" : s); };
            mspai::CscMatrix C = mspai::read_system(argv[2], opt(argv[3]), std::atof(argv[5]));
            mspai::CscMatrix B = mspai::read_target(C, opt(argv[6]), opt(argv[4]), std::atof(argv[5]));
            dump(argv[7], C);
            dump(argv[8], B);
            return 0;
        }
        if (argc != 3) return 2;
        mspai::CscMatrix A = mspai::read_matrix_market(argv[1]);
        mspai::write_preconditioner(argv[2], A);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
