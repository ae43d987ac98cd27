// Reads a MatrixMarket file with the library's reader and writes it back with
// the reference-format writer (mspai_amd/host/mmio.cpp); used by
// tests/test_mmio.py.
#include <cstdio>
#include <exception>

#include "mmio.h"

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    try {
        mspai::CscMatrix A = mspai::read_matrix_market(argv[1]);
        mspai::write_preconditioner(argv[2], A);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
