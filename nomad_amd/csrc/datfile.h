// Rating-matrix container and the train.dat/test.dat binary format.
//
// The on-disk layout is the PETSc-style sparse-matrix file that the reference
// reads in Code/nomad/nb_data.cpp:19-176 and that Scripts/convert.py:101-104
// writes:
//   int32 class_id               1211216 (MAT) or 1015 (LONG)
//   MAT : int32 rows, int32 cols, int32 nnz
//   LONG: int32 rows, int32 cols, int64 nnz
//   int32  nnz_per_row[rows]
//   int32  col_index[nnz]        row-major order (file order within a row kept)
//   double value[nnz]
//
// In memory the whole matrix is kept as one column-major (CSC) structure with
// the rows of each column in ascending order (the order nb_data.cpp:136-145
// produces when it walks the rows of its part).  A contiguous row block's
// entries of column j are then a contiguous sub-range of column j, which is
// what every kernel of this library relies on.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace nomad {

constexpr int32_t kMatFileClassId = 1211216;   // nb_data.cpp:14
constexpr int32_t kLongFileClassId = 1015;     // nb_data.cpp:15

struct Csc {
    int32_t m = 0, n = 0;
    std::vector<int64_t> colptr;  // n + 1
    std::vector<int32_t> rowidx;  // nnz, ascending within a column
    std::vector<double> val;      // nnz
    int64_t nnz() const { return colptr.empty() ? 0 : colptr.back(); }
};

// Row-major triplet view used by the writers (rows ascending; columns in the
// order they should appear in the file).
struct Csr {
    int32_t m = 0, n = 0;
    std::vector<int64_t> rowptr;  // m + 1
    std::vector<int32_t> colidx;
    std::vector<double> val;
    int64_t nnz() const { return rowptr.empty() ? 0 : rowptr.back(); }
};

// Reads a .dat file into CSC.  Returns false with a message on any format or
// I/O error (the reference exits on these; the library reports them).
bool read_dat(const std::string& path, Csc& out, std::string& err);

// Reads only the header: rows, cols, nnz.
bool read_dat_header(const std::string& path, int32_t& m, int32_t& n, int64_t& nnz,
                     std::string& err);

// Writes CSR to a .dat file.  long_format selects class id 1015 (int64 nnz);
// otherwise 1211216 is used, which requires nnz < 2^31.
bool write_dat(const std::string& path, const Csr& a, bool long_format, std::string& err);

// CSR -> CSC with rows ascending inside each column (stable in row order, so a
// duplicated (i, j) keeps its file order, as nb_data.cpp:136-145 does).
void csr_to_csc(const Csr& a, Csc& out);

// CSC -> CSR (columns ascending inside each row).
void csc_to_csr(const Csc& a, Csr& out);

}  // namespace nomad
