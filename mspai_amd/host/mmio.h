/*
 * MatrixMarket input and preconditioner output in the reference's formats.
 *
 *   read_matrix_market  follows Read_mm_Matrix (src/Read_mm_Matrix.cpp:115-231,
 *                       532-599, 669-845): "coordinate real general|symmetric"
 *                       files, 1-based indices, commas accepted as separators,
 *                       symmetric files expanded to both triangles, entries
 *                       ordered by column then row (Col_Comparator,
 *                       src/Read_mm_Matrix.h:73-82).
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

/* Throws std::runtime_error with the reason on malformed input. */
CscMatrix read_matrix_market(const std::string& path);

/* Throws std::runtime_error when the file cannot be written. */
void write_preconditioner(const std::string& path, const CscMatrix& M);

}  // namespace mspai

#endif
