// Test driver for include/nomad_csc.h: reads the P row parts of a .dat file
// the way the reference does (the oracle's restatement of read_data,
// Code/nomad/nb_data.cpp:19-176) into a struct with the members of the
// reference's Data (Code/nomad/nomad_body.h:34-46), merges them with
// nomad_csc_from_parts and writes the global CSC as raw arrays:
// int64 colptr[n + 1], int32 rowidx[nnz], double val[nnz].
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nomad_csc.h"
#include "nomad_oracle.h"

struct Data {
    std::vector<int> col_offset;
    std::vector<int> row_idx;
    std::vector<double> row_val;
    int num_rows = 0, num_cols = 0;
    long long num_nonzero = 0;
    int min_row_index = 0, max_row_index = 0, local_num_rows = 0;
    long long local_num_nonzero = 0;
};

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    const int P = std::atoi(argv[2]);
    std::vector<Data> parts(P);
    int ncols = 0;
    for (int p = 0; p < P; ++p) {
        Data& d = parts[p];
        int64_t tot = 0, ln = 0;
        if (orc_dat_part_info(argv[1], p, P, &d.num_rows, &d.num_cols, &tot, &d.min_row_index, &d.local_num_rows,
                              &ln))
            return 3;
        ncols = d.num_cols;
        d.col_offset.resize(size_t(ncols) + 1);
        d.row_idx.resize(size_t(ln));
        d.row_val.resize(size_t(ln));
        if (orc_dat_part_read(argv[1], p, P, d.col_offset.data(), d.row_idx.data(), d.row_val.data())) return 4;
    }
    const nomad_global_csc g = nomad_csc_from_parts(parts.data(), P, ncols);
    FILE* f = std::fopen(argv[3], "wb");
    if (!f) return 5;
    std::fwrite(g.colptr.data(), 8, g.colptr.size(), f);
    std::fwrite(g.rowidx.data(), 4, g.rowidx.size(), f);
    std::fwrite(g.val.data(), 8, g.val.size(), f);
    std::fclose(f);
    return 0;
}
