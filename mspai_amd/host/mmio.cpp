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

/* The entries of one MatrixMarket coordinate file, 0-based, in file order;
 * a symmetric file gives each off-diagonal entry twice (the mirrored one
 * second), a skew-symmetric file its mirror with the opposite sign. */
struct MmFile {
    int m = 0, n = 0;
    bool symmetric = false, skew = false;
    std::vector<Entry> ent;
};

MmFile read_entries(const std::string& path)
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
    MmFile r;
    const bool pattern = banner.find("pattern") != std::string::npos;
    /* "skew-symmetric" contains "symmetric": test it first */
    r.skew = banner.find("skew-symmetric") != std::string::npos;
    r.symmetric = !r.skew && banner.find("symmetric") != std::string::npos;
    /* skip comments (Skip_Header) */
    do {
        if (!std::fgets(line, sizeof line, f)) {
            std::fclose(f);
            throw std::runtime_error(path + ": missing size line");
        }
    } while (line[0] == '%');
    long nnz = 0;
    if (std::sscanf(line, "%d %d %ld", &r.m, &r.n, &nnz) != 3 || r.m <= 0 || r.n <= 0 || nnz < 0) {
        std::fclose(f);
        throw std::runtime_error(path + ": bad size line");
    }
    r.ent.reserve(r.symmetric || r.skew ? 2 * nnz : nnz);
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
        if (got != (pattern ? 2 : 3) || i < 1 || j < 1 || i > r.m || j > r.n) {
            std::fclose(f);
            throw std::runtime_error(path + ": bad entry line " + std::to_string(k + 1));
        }
        if (r.skew && i == j) {
            std::fclose(f);
            throw std::runtime_error(path + ": skew-symmetric file with a diagonal entry");
        }
        r.ent.push_back({i - 1, j - 1, v});
        if ((r.symmetric || r.skew) && i != j) r.ent.push_back({j - 1, i - 1, r.skew ? -v : v});
    }
    std::fclose(f);
    return r;
}

/* Entries -> CSC, ordered by column then row (Col_Comparator,
 * src/Read_mm_Matrix.h:73-82); a repeated entry is summed. */
CscMatrix to_csc(int m, int n, std::vector<Entry>& ent)
{
    std::stable_sort(ent.begin(), ent.end(), [](const Entry& a, const Entry& b) {
        return a.j < b.j || (a.j == b.j && a.i < b.i);
    });
    CscMatrix A;
    A.nrows = m;
    A.ncols = n;
    A.colptr.assign(n + 1, 0);
    int pi = -1, pj = -1;
    for (const Entry& e : ent) {
        if (e.i == pi && e.j == pj) {
            A.val.back() += e.v;
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

/* Data_To_Memory (src/Read_mm_Matrix.cpp:532-599) for a probing file: its
 * entry (i, j, v) becomes row `dim` + i of column j with value v * rho
 * (symmetric files mirrored with the same shift and scale). */
void append_probing(const MmFile& p, int dim, double rho, std::vector<Entry>& ent)
{
    for (const Entry& e : p.ent) ent.push_back({e.i + dim, e.j, e.v * rho});
}

}  // namespace

CscMatrix read_matrix_market(const std::string& path)
{
    MmFile r = read_entries(path);
    CscMatrix A = to_csc(r.m, r.n, r.ent);
    A.symmetric = r.symmetric;
    return A;
}

CscMatrix read_system(const std::string& matrix, const std::string& ce_path, double rho)
{
    MmFile a = read_entries(matrix);
    if (a.m != a.n)
        throw std::runtime_error("Could not read matrix size and nnz's properly! Matrix must be square");
    int rows = a.m;
    if (!ce_path.empty()) {
        MmFile p = read_entries(ce_path);
        if (p.n != a.n)
            throw std::runtime_error("Mismatch in matrix and probing dimensions! Probing matrix and system/target "
                                     "matrix must have the same number of columns in case of inverse or explicit "
                                     "probing.");
        append_probing(p, a.n, rho, a.ent);
        rows += p.m;
    }
    CscMatrix C = to_csc(rows, a.n, a.ent);
    C.symmetric = a.symmetric && ce_path.empty();
    return C;
}

CscMatrix read_target(const CscMatrix& C, const std::string& target, const std::string& be_path, double rho)
{
    const int n = C.ncols;
    std::vector<Entry> ent;
    int rows = n;
    if (!target.empty()) {
        MmFile t = read_entries(target);
        if (t.m != t.n)
            throw std::runtime_error("Could not read matrix size and nnz's properly! Matrix must be square");
        if (t.n != n)
            throw std::runtime_error("Mismatch in system matrix and target matrix dimensions! Target matrix and "
                                     "system matrix must have the same number of columns.");
        ent = std::move(t.ent);
        rows = t.m;
    } else {
        /* Generate_Identity_Data (src/Read_mm_Matrix.cpp:601-667) */
        ent.reserve(n);
        for (int i = 0; i < n; ++i) ent.push_back({i, i, 1.0});
    }
    if (!be_path.empty()) {
        MmFile p = read_entries(be_path);
        if (p.n != n)
            throw std::runtime_error("Mismatch in matrix and probing dimensions! Probing matrix and system/target "
                                     "matrix must have the same number of columns.");
        append_probing(p, n, rho, ent);
        rows += p.m;
        if (rows != C.nrows)
            throw std::runtime_error("Mismatch in probing dimensions! Probing matrix Ce and probing matrix Be "
                                     "must have the same number of rows!");
    }
    return to_csc(rows, n, ent);
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
