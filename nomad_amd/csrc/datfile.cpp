// train.dat / test.dat reader and writer.  Format: Code/nomad/nb_data.cpp:19-176
// (reader) and Scripts/convert.py:101-104 (writer).
#include "datfile.h"

#include <cstdio>
#include <cstring>
#include <memory>

namespace nomad {

namespace {

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

bool read_exact(FILE* f, void* dst, size_t bytes) {
    char* p = static_cast<char*>(dst);
    while (bytes > 0) {
        size_t chunk = bytes > (size_t(1) << 30) ? (size_t(1) << 30) : bytes;
        size_t got = std::fread(p, 1, chunk, f);
        if (got != chunk) return false;
        p += got;
        bytes -= got;
    }
    return true;
}

bool write_exact(FILE* f, const void* src, size_t bytes) {
    const char* p = static_cast<const char*>(src);
    while (bytes > 0) {
        size_t chunk = bytes > (size_t(1) << 30) ? (size_t(1) << 30) : bytes;
        size_t put = std::fwrite(p, 1, chunk, f);
        if (put != chunk) return false;
        p += put;
        bytes -= put;
    }
    return true;
}

bool read_header(FILE* f, const std::string& path, int32_t& m, int32_t& n, int64_t& nnz,
                 std::string& err) {
    int32_t id = 0;
    if (!read_exact(f, &id, sizeof(id))) {
        err = "cannot read class id from " + path;
        return false;
    }
    if (id == kMatFileClassId) {  // nb_data.cpp:46-57
        int32_t h[3];
        if (!read_exact(f, h, sizeof(h))) {
            err = "cannot read header from " + path;
            return false;
        }
        m = h[0];
        n = h[1];
        nnz = h[2];
    } else if (id == kLongFileClassId) {  // nb_data.cpp:59-73
        int32_t h[2];
        if (!read_exact(f, h, sizeof(h)) || !read_exact(f, &nnz, sizeof(nnz))) {
            err = "cannot read header from " + path;
            return false;
        }
        m = h[0];
        n = h[1];
    } else {
        err = std::to_string(id) + " does not identify a valid binary matrix file: " + path;
        return false;
    }
    if (m < 0 || n < 0 || nnz < 0) {
        err = "negative dimension in header of " + path;
        return false;
    }
    return true;
}

}  // namespace

bool read_dat_header(const std::string& path, int32_t& m, int32_t& n, int64_t& nnz,
                     std::string& err) {
    FilePtr f(std::fopen(path.c_str(), "rb"));
    if (!f) {
        err = "cannot open file: " + path;
        return false;
    }
    return read_header(f.get(), path, m, n, nnz, err);
}

bool read_dat(const std::string& path, Csc& out, std::string& err) {
    FilePtr f(std::fopen(path.c_str(), "rb"));
    if (!f) {
        err = "cannot open file: " + path;
        return false;
    }
    int32_t m, n;
    int64_t nnz;
    if (!read_header(f.get(), path, m, n, nnz, err)) return false;

    std::vector<int32_t> row_nnz(m);
    if (!read_exact(f.get(), row_nnz.data(), sizeof(int32_t) * size_t(m))) {
        err = "error in reading nnz values from " + path;
        return false;
    }
    int64_t sum = 0;
    for (int32_t i = 0; i < m; ++i) {
        if (row_nnz[i] < 0) {
            err = "negative row length in " + path;
            return false;
        }
        sum += row_nnz[i];
    }
    if (sum != nnz) {
        err = "row lengths do not add up to nnz in " + path;
        return false;
    }
    std::vector<int32_t> col(nnz);
    if (!read_exact(f.get(), col.data(), sizeof(int32_t) * size_t(nnz))) {
        err = "error in reading column indices from " + path;
        return false;
    }
    std::vector<double> v(nnz);
    if (!read_exact(f.get(), v.data(), sizeof(double) * size_t(nnz))) {
        err = "error in reading matrix values from " + path;
        return false;
    }
    for (int64_t e = 0; e < nnz; ++e) {
        if (col[e] < 0 || col[e] >= n) {
            err = "column index out of range in " + path;
            return false;
        }
    }
    // Counting sort by column, stable in row order.
    out.m = m;
    out.n = n;
    out.colptr.assign(size_t(n) + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) out.colptr[col[e] + 1]++;
    for (int32_t j = 0; j < n; ++j) out.colptr[j + 1] += out.colptr[j];
    out.rowidx.resize(nnz);
    out.val.resize(nnz);
    std::vector<int64_t> fill(out.colptr.begin(), out.colptr.end() - 1);
    int64_t e = 0;
    for (int32_t i = 0; i < m; ++i) {
        for (int32_t t = 0; t < row_nnz[i]; ++t, ++e) {
            int64_t d = fill[col[e]]++;
            out.rowidx[d] = i;
            out.val[d] = v[e];
        }
    }
    return true;
}

bool write_dat(const std::string& path, const Csr& a, bool long_format, std::string& err) {
    const int64_t nnz = a.nnz();
    if (!long_format && nnz > INT32_MAX) {
        err = "nnz does not fit the 1211216 format; use the long (1015) format";
        return false;
    }
    FilePtr f(std::fopen(path.c_str(), "wb"));
    if (!f) {
        err = "cannot create file: " + path;
        return false;
    }
    bool ok = true;
    if (long_format) {
        int32_t h[3] = {kLongFileClassId, a.m, a.n};
        ok = write_exact(f.get(), h, sizeof(h)) && write_exact(f.get(), &nnz, sizeof(nnz));
    } else {
        int32_t h[4] = {kMatFileClassId, a.m, a.n, int32_t(nnz)};
        ok = write_exact(f.get(), h, sizeof(h));
    }
    std::vector<int32_t> row_nnz(a.m);
    for (int32_t i = 0; i < a.m; ++i) {
        int64_t len = a.rowptr[i + 1] - a.rowptr[i];
        if (len > INT32_MAX) {
            err = "row too long for the format";
            return false;
        }
        row_nnz[i] = int32_t(len);
    }
    ok = ok && write_exact(f.get(), row_nnz.data(), sizeof(int32_t) * size_t(a.m));
    ok = ok && write_exact(f.get(), a.colidx.data(), sizeof(int32_t) * size_t(nnz));
    ok = ok && write_exact(f.get(), a.val.data(), sizeof(double) * size_t(nnz));
    if (!ok) {
        err = "write error on " + path;
        return false;
    }
    return true;
}

void csr_to_csc(const Csr& a, Csc& out) {
    const int64_t nnz = a.nnz();
    out.m = a.m;
    out.n = a.n;
    out.colptr.assign(size_t(a.n) + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) out.colptr[a.colidx[e] + 1]++;
    for (int32_t j = 0; j < a.n; ++j) out.colptr[j + 1] += out.colptr[j];
    out.rowidx.resize(nnz);
    out.val.resize(nnz);
    std::vector<int64_t> fill(out.colptr.begin(), out.colptr.end() - 1);
    for (int32_t i = 0; i < a.m; ++i) {
        for (int64_t e = a.rowptr[i]; e < a.rowptr[i + 1]; ++e) {
            int64_t d = fill[a.colidx[e]]++;
            out.rowidx[d] = i;
            out.val[d] = a.val[e];
        }
    }
}

void csc_to_csr(const Csc& a, Csr& out) {
    const int64_t nnz = a.nnz();
    out.m = a.m;
    out.n = a.n;
    out.rowptr.assign(size_t(a.m) + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) out.rowptr[a.rowidx[e] + 1]++;
    for (int32_t i = 0; i < a.m; ++i) out.rowptr[i + 1] += out.rowptr[i];
    out.colidx.resize(nnz);
    out.val.resize(nnz);
    std::vector<int64_t> fill(out.rowptr.begin(), out.rowptr.end() - 1);
    for (int32_t j = 0; j < a.n; ++j) {
        for (int64_t e = a.colptr[j]; e < a.colptr[j + 1]; ++e) {
            int64_t d = fill[a.rowidx[e]]++;
            out.colidx[d] = j;
            out.val[d] = a.val[e];
        }
    }
}

}  // namespace nomad
