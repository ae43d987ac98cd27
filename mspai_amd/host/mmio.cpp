#include "mmio.h"

#include <algorithm>
#include <cstdio>
#include <cctype>
#include <cstring>
#include <stdexcept>

namespace mspai {

namespace {

struct Entry {
    int i, j;
    double v;
};

}  // namespace

CscMatrix read_matrix_market(const std::string& path)
{
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw std::runtime_error("cannot open " + path);
    char line[1024];
    if (!std::fgets(line, sizeof line, f)) {
        std::fclose(f);
        throw std::runtime_error(path + ": empty file");
    }
    std::string banner(line);
    for (char& c : banner) c = (char)std::tolower((unsigned char)c);
    if (banner.rfind("%%matrixmarket", 0) != 0 || banner.find("matrix") == std::string::npos ||
        banner.find("coordinate") == std::string::npos) {
        std::fclose(f);
        throw std::runtime_error(path + ": not a MatrixMarket coordinate matrix");
    }
    if (banner.find("complex") != std::string::npos || banner.find("hermitian") != std::string::npos) {
        std::fclose(f);
        throw std::runtime_error(path + ": complex matrices are not supported");
    }
    const bool pattern = banner.find("pattern") != std::string::npos;
    const bool symmetric = banner.find("symmetric") != std::string::npos;
    /* skip comments (Skip_Header) */
    do {
        if (!std::fgets(line, sizeof line, f)) {
            std::fclose(f);
            throw std::runtime_error(path + ": missing size line");
        }
    } while (line[0] == '%');
    int m = 0, n = 0;
    long nnz = 0;
    if (std::sscanf(line, "%d %d %ld", &m, &n, &nnz) != 3 || m <= 0 || n <= 0 || nnz < 0) {
        std::fclose(f);
        throw std::runtime_error(path + ": bad size line");
    }
    std::vector<Entry> ent;
    ent.reserve(symmetric ? 2 * nnz : nnz);
    for (long k = 0; k < nnz; ++k) {
        if (!std::fgets(line, sizeof line, f)) {
            std::fclose(f);
            throw std::runtime_error(path + ": fewer entries than declared");
        }
        for (char* c = line; *c; ++c)
            if (*c == ',') *c = ' ';
        int i = 0, j = 0;
        double v = 1.0;
        int got = pattern ? std::sscanf(line, "%d %d", &i, &j) : std::sscanf(line, "%d %d %le", &i, &j, &v);
        if (got != (pattern ? 2 : 3) || i < 1 || j < 1 || i > m || j > n) {
            std::fclose(f);
            throw std::runtime_error(path + ": bad entry line " + std::to_string(k + 1));
        }
        ent.push_back({i - 1, j - 1, v});
        if (symmetric && i != j) ent.push_back({j - 1, i - 1, v});
    }
    std::fclose(f);
    std::sort(ent.begin(), ent.end(), [](const Entry& a, const Entry& b) {
        return a.j < b.j || (a.j == b.j && a.i < b.i);
    });
    CscMatrix A;
    A.nrows = m;
    A.ncols = n;
    A.symmetric = symmetric;
    A.colptr.assign(n + 1, 0);
    int pi = -1, pj = -1;
    for (const Entry& e : ent) {
        if (e.i == pi && e.j == pj) {
            A.val.back() += e.v;  // repeated entry: summed
            continue;
        }
        A.rowind.push_back(e.i);
        A.val.push_back(e.v);
        A.colptr[e.j + 1]++;
        pi = e.i;
        pj = e.j;
    }
    for (int j = 0; j < n; ++j) A.colptr[j + 1] += A.colptr[j];
    return A;
}

void write_preconditioner(const std::string& path, const CscMatrix& M)
{
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("Failed writing preconditioner to file " + path);
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%d %d %d\n", M.nrows, M.ncols, M.colptr[M.ncols]);
    for (int j = 0; j < M.ncols; ++j)
        for (int e = M.colptr[j]; e < M.colptr[j + 1]; ++e) {
            double v = M.val[e];
            std::fprintf(f, "%d %d%s%.13e\n", M.rowind[e] + 1, j + 1, v < 0.0 ? " " : "  ", v);
        }
    std::fclose(f);
}

}  // namespace mspai
