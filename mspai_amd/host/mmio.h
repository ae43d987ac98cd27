/*
 * MatrixMarket input and preconditioner output in the reference's formats.
 *
 *   read_matrix_market  follows Read_mm_Matrix (src/Read_mm_Matrix.cpp:115-231,
 *                       532-599, 669-845): "coordinate real general|symmetric"
 *                       files, 1-based indices, commas accepted as separators,
 *                       symmetric files expanded to both triangles, entries
 *                       ordered by column then row (Col_Comparator,
 *                       src/Read_mm_Matrix.h:73-82).  Skew-symmetric files are
 *                       expanded with the mirrored entry negated (the format's
 *                       meaning; the reference's reader takes them as general
 *                       and keeps only the stored triangle).
 *   read_system         Read_Matrix_File with a probing file (-Ce,
 *                       src/Read_mm_Matrix.cpp:115-231, 432-599): the square
 *                       system matrix A with the probing rows rho * Ce appended
 *                       below it, C = [A; rho Ce] ((n + k) x n).  The file holds
 *                       the probing vectors transposed, one per row.
 *   read_target         Read_Target_File (src/Read_mm_Matrix.cpp:233-420):
 *                       B = [T; rho Be] from a square target file T, or
 *                       [I; rho Be] (Generate_Identity_Data, :601-667) when no
 *                       target file is given; without -Be, B = T or I.
 *   write_preconditioner follows Matrix<double>::Write_Matrix_To_File
 *                       (src/Matrix.cpp:250-342): header, "m n nnz", then per
 *                       entry "row col" + (" " for a negative value, else two
 *                       spaces) + "%.13e", columns in order.
 */
#ifndef MSPAI_MMIO_H
#define MSPAI_MMIO_H

#include <string>
#include <vector>

namespace mspai {

struct CscMatrix {
    int nrows = 0, ncols = 0;
    std::vector<int> colptr, rowind;
    std::vector<double> val;
    bool symmetric = false;  // the file declared "symmetric"
};

/* All readers throw std::runtime_error with the reason on malformed input
 * or mismatched dimensions (the reference's messages). */
CscMatrix read_matrix_market(const std::string& path);

/* ce_path empty: C = A. */
CscMatrix read_system(const std::string& matrix, const std::string& ce_path, double rho);

/* target empty: the identity; be_path empty: no probing rows.  C is the
 * system read_system returned (its shape is checked against B's). */
CscMatrix read_target(const CscMatrix& C, const std::string& target, const std::string& be_path, double rho);

/* Throws std::runtime_error when the file cannot be written. */
void write_preconditioner(const std::string& path, const CscMatrix& M);

}  // namespace mspai

#endif
