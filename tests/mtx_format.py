"""The reference's preconditioner file layout (Matrix<double>::
Write_Matrix_To_File, src/Matrix.cpp:250-342), restated for comparisons."""
import numpy as np
import scipy.sparse as sp


def format_mtx(M) -> str:
    M = sp.csc_matrix(M)
    M.sort_indices()
    out = ["%%MatrixMarket matrix coordinate real general\n", f"{M.shape[0]} {M.shape[1]} {M.nnz}\n"]
    for j in range(M.shape[1]):
        for e in range(M.indptr[j], M.indptr[j + 1]):
            v = float(M.data[e])
            out.append("%d %d%s%.13e\n" % (M.indices[e] + 1, j + 1, " " if v < 0.0 else "  ", v))
    return "".join(out)


def write_input_mtx(path, A, symmetric=False):
    """A MatrixMarket input file (1-based, full precision)."""
    A = sp.coo_matrix(A)
    if symmetric:
        keep = A.row >= A.col
        rows, cols, vals = A.row[keep], A.col[keep], A.data[keep]
    else:
        rows, cols, vals = A.row, A.col, A.data
    with open(path, "w") as f:
        f.write("%%%%MatrixMarket matrix coordinate real %s\n" % ("symmetric" if symmetric else "general"))
        f.write("%% written by the test suite\n")
        f.write(f"{A.shape[0]} {A.shape[1]} {rows.size}\n")
        for i, j, v in zip(rows, cols, vals):
            f.write("%d %d %.17g\n" % (i + 1, j + 1, v))
