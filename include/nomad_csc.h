/*
 * nomad_csc.h -- header-only C++ helper: the reference's per-thread rating
 * parts merged into the one global CSC that nomad_upload_train /
 * nomad_upload_test (nomad_hip.h) take.
 *
 * The reference reads part p of P = ranks * threads as a CSC with local row
 * ids (struct Data, Code/nomad/nomad_body.h:34-46, filled by read_data,
 * Code/nomad/nb_data.cpp:19-176): col_offset[num_cols + 1], row_idx (row -
 * min_row_index), row_val, over the contiguous rows [min_row_index,
 * min_row_index + local_num_rows).  Parts of consecutive indexes hold
 * consecutive row ranges, so concatenating each column's entries part after
 * part gives rows ascending inside every column, which is what the library
 * requires.
 *
 * Part is any type with those members (the reference's Data with its TBB
 * allocators, or an equivalent); parts[0 .. nparts) must be in row order and
 * cover the rows they hold without overlap.  No library call is made; the
 * result owns plain std::vectors.
 */
#ifndef NOMAD_CSC_H_
#define NOMAD_CSC_H_

#ifdef __cplusplus

#include <stdint.h>

#include <stdexcept>
#include <vector>

struct nomad_global_csc {
    std::vector<int64_t> colptr; /* ncols + 1 */
    std::vector<int32_t> rowidx; /* global row ids, ascending in each column */
    std::vector<double> val;
    int64_t nnz = 0;
};

template <class Part>
nomad_global_csc nomad_csc_from_parts(const Part* parts, int nparts, int ncols) {
    nomad_global_csc g;
    g.colptr.assign(size_t(ncols) + 1, 0);
    for (int p = 0; p < nparts; ++p) {
        if (int64_t(parts[p].col_offset.size()) != int64_t(ncols) + 1)
            throw std::invalid_argument("nomad_csc_from_parts: a part's col_offset does not have ncols + 1 entries");
        if (p > 0 && parts[p].min_row_index < parts[p - 1].min_row_index + parts[p - 1].local_num_rows)
            throw std::invalid_argument("nomad_csc_from_parts: parts are not in row order");
        for (int j = 0; j < ncols; ++j) g.colptr[size_t(j) + 1] += parts[p].col_offset[j + 1] - parts[p].col_offset[j];
    }
    for (int j = 0; j < ncols; ++j) g.colptr[size_t(j) + 1] += g.colptr[size_t(j)];
    g.nnz = g.colptr[size_t(ncols)];
    g.rowidx.resize(size_t(g.nnz));
    g.val.resize(size_t(g.nnz));
    std::vector<int64_t> fill(g.colptr.begin(), g.colptr.end() - 1);
    for (int p = 0; p < nparts; ++p) {
        const Part& d = parts[p];
        for (int j = 0; j < ncols; ++j)
            for (int o = d.col_offset[j]; o < d.col_offset[j + 1]; ++o) {
                const int64_t e = fill[size_t(j)]++;
                g.rowidx[size_t(e)] = int32_t(d.row_idx[o] + d.min_row_index);
                g.val[size_t(e)] = double(d.row_val[o]);
            }
    }
    return g;
}

#endif /* __cplusplus */

#endif /* NOMAD_CSC_H_ */
