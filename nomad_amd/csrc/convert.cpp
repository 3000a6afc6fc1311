// Rating-file tooling around the train.dat / test.dat format.
//
// 1. TSV -> .dat conversion, the procedure of the reference's
//    Scripts/convert.py: two passes over "user<TAB>item<TAB>rating" lines of
//    the training and the test file; every distinct user and item id gets an
//    index after a shuffle with Python's random generator seeded 12345;
//    each file is written as a row-major matrix (class id 1211216) whose rows
//    are the user indexes, duplicate (user, item) lines summed and columns
//    ascending inside a row, as scipy's csr_matrix builds it; user_ids.txt
//    and item_ids.txt list the ids in index order.  The ids enter the shuffle
//    in order of first appearance (train, then test): convert.py starts from
//    the iteration order of a Python 2 set, an artefact of that interpreter's
//    string hashing, which this conversion does not reproduce.
//
// 2. The reference's synthetic generators (Code/nomad/synthgen-conf.cpp,
//    Code/nomad/synthgen.cpp, pl-dis.hpp, gen.sh), restated on the same
//    libstdc++ distributions drawn in the same order (std::mt19937,
//    std::normal_distribution, std::discrete_distribution,
//    std::uniform_real_distribution), so one seed gives the same files as the
//    reference built with the same standard library.
#include "convert.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <set>
#include <sstream>
#include <unordered_map>

#include "datfile.h"

namespace nomad {

// ---- Python's random module: MT19937 seeded by init_by_array ------------------
PyRandom::PyRandom(uint64_t seed) {
    // random.seed(int): the absolute value split into 32-bit words, low first
    std::vector<uint32_t> key;
    do {
        key.push_back(uint32_t(seed & 0xFFFFFFFFu));
        seed >>= 32;
    } while (seed);
    // init_genrand(19650218)
    mt_[0] = 19650218u;
    for (int i = 1; i < 624; ++i) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + uint32_t(i);
    // init_by_array(key)
    int i = 1, j = 0;
    const int klen = int(key.size());
    for (int k = std::max(624, klen); k; --k) {
        mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1664525u)) + key[size_t(j)] + uint32_t(j);
        ++i;
        ++j;
        if (i >= 624) {
            mt_[0] = mt_[623];
            i = 1;
        }
        if (j >= klen) j = 0;
    }
    for (int k = 623; k; --k) {
        mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1566083941u)) - uint32_t(i);
        ++i;
        if (i >= 624) {
            mt_[0] = mt_[623];
            i = 1;
        }
    }
    mt_[0] = 0x80000000u;
    idx_ = 624;
}

uint32_t PyRandom::next32() {
    if (idx_ >= 624) {
        for (int k = 0; k < 624; ++k) {
            const uint32_t y = (mt_[k] & 0x80000000u) | (mt_[(k + 1) % 624] & 0x7FFFFFFFu);
            mt_[k] = mt_[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908B0DFu : 0u);
        }
        idx_ = 0;
    }
    uint32_t y = mt_[idx_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9D2C5680u;
    y ^= (y << 15) & 0xEFC60000u;
    y ^= y >> 18;
    return y;
}

double PyRandom::random() {
    // genrand_res53: 27 + 26 bits
    const uint32_t a = next32() >> 5, b = next32() >> 6;
    return (double(a) * 67108864.0 + double(b)) * (1.0 / 9007199254740992.0);
}

// ---- TSV -> .dat -------------------------------------------------------------------
namespace {

struct Triplets {
    std::vector<int64_t> user, item;
    std::vector<double> val;
};

// One pass over a TSV file: fields 0 and 1 are the ids, field 2 the rating.
template <class F>
bool scan_tsv(const std::string& path, F&& per_line, std::string& err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open " + path;
        return false;
    }
    std::string line;
    int64_t no = 0;
    while (std::getline(in, line)) {
        ++no;
        const size_t t1 = line.find('\t');
        const size_t t2 = t1 == std::string::npos ? t1 : line.find('\t', t1 + 1);
        if (t2 == std::string::npos) {
            err = path + ":" + std::to_string(no) + ": fewer than three tab-separated fields";
            return false;
        }
        const size_t t3 = line.find('\t', t2 + 1);
        const std::string rating = line.substr(t2 + 1, t3 == std::string::npos ? std::string::npos : t3 - t2 - 1);
        char* end = nullptr;
        const double v = std::strtod(rating.c_str(), &end);
        if (end == rating.c_str()) {
            err = path + ":" + std::to_string(no) + ": rating is not a number";
            return false;
        }
        if (!per_line(line.substr(0, t1), line.substr(t1 + 1, t2 - t1 - 1), v)) return false;
    }
    return true;
}

// scipy's csr_matrix((vals, (rows, cols))): duplicates summed, columns
// ascending inside each row.
Csr to_csr(int32_t m, int32_t n, const Triplets& t) {
    std::vector<int64_t> order(t.val.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = int64_t(i);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return t.user[size_t(a)] != t.user[size_t(b)] ? t.user[size_t(a)] < t.user[size_t(b)]
                                                      : t.item[size_t(a)] < t.item[size_t(b)];
    });
    Csr c;
    c.m = m;
    c.n = n;
    c.rowptr.assign(size_t(m) + 1, 0);
    for (size_t q = 0; q < order.size(); ++q) {
        const size_t e = size_t(order[q]);
        if (q > 0) {
            const size_t p = size_t(order[q - 1]);
            if (t.user[p] == t.user[e] && t.item[p] == t.item[e]) {
                c.val.back() += t.val[e];
                continue;
            }
        }
        c.colidx.push_back(int32_t(t.item[e]));
        c.val.push_back(t.val[e]);
        c.rowptr[size_t(t.user[e]) + 1]++;
    }
    for (int32_t i = 0; i < m; ++i) c.rowptr[size_t(i) + 1] += c.rowptr[size_t(i)];
    return c;
}

}  // namespace

bool convert_tsv(const std::string& train_path, const std::string& test_path, const std::string& out_dir,
                 uint64_t seed, int64_t& nusers, int64_t& nitems, std::string& err) {
    // first pass: distinct ids, in order of first appearance
    std::vector<std::string> users, items;
    std::unordered_map<std::string, int64_t> uidx, iidx;
    auto collect = [&](const std::string& u, const std::string& i, double) {
        if (uidx.emplace(u, int64_t(users.size())).second) users.push_back(u);
        if (iidx.emplace(i, int64_t(items.size())).second) items.push_back(i);
        return true;
    };
    if (!scan_tsv(train_path, collect, err) || !scan_tsv(test_path, collect, err)) return false;
    if (users.size() >= (size_t(1) << 31) || items.size() >= (size_t(1) << 31)) {
        err = "more than 2^31 - 1 distinct users or items";
        return false;
    }
    // random.seed(seed); random.shuffle(users); random.shuffle(items) (Python 2:
    // j = int(random() * (i + 1)) for i = n-1 .. 1)
    PyRandom rng(seed);
    auto shuffle = [&](std::vector<std::string>& v) {
        for (size_t i = v.size(); i-- > 1;) {
            const size_t j = size_t(rng.random() * double(i + 1));
            std::swap(v[i], v[j]);
        }
    };
    shuffle(users);
    shuffle(items);
    for (size_t q = 0; q < users.size(); ++q) uidx[users[q]] = int64_t(q);
    for (size_t q = 0; q < items.size(); ++q) iidx[items[q]] = int64_t(q);
    nusers = int64_t(users.size());
    nitems = int64_t(items.size());
    // second pass per file: triplets with the new indexes, written row-major
    for (const auto& [src, name] : {std::make_pair(train_path, std::string("train.dat")),
                                    std::make_pair(test_path, std::string("test.dat"))}) {
        Triplets t;
        auto add = [&](const std::string& u, const std::string& i, double v) {
            t.user.push_back(uidx.at(u));
            t.item.push_back(iidx.at(i));
            t.val.push_back(v);
            return true;
        };
        if (!scan_tsv(src, add, err)) return false;
        const Csr c = to_csr(int32_t(nusers), int32_t(nitems), t);
        if (!write_dat(out_dir + "/" + name, c, false, err)) return false;
    }
    for (const auto& [name, ids] : {std::make_pair(std::string("user_ids.txt"), &users),
                                    std::make_pair(std::string("item_ids.txt"), &items)}) {
        std::ofstream o(out_dir + "/" + name);
        if (!o) {
            err = "cannot write " + out_dir + "/" + name;
            return false;
        }
        for (const std::string& s : *ids) o << s << "\n";
    }
    return true;
}

// ---- the reference's synthetic generators ------------------------------------------
bool synthgen_conf(int32_t nrow, int32_t ncol, int32_t seed, const std::string& row_path,
                   const std::string& col_path, const std::string& sep, std::string& err) {
    if (nrow < 1 || ncol < 1) {
        err = "synthgen_conf: nrow and ncol must be positive";
        return false;
    }
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    // frequency of each row degree 0..ncol: int(N(20, 5)) redrawn until in [0, ncol)
    std::normal_distribution<> dis_row(20, 5);
    auto gen_row = [&]() {
        int v = static_cast<int>(dis_row(rng));
        while (v < 0 || v >= ncol) v = static_cast<int>(dis_row(rng));
        return v;
    };
    {
        std::ofstream f(row_path);
        if (!f) {
            err = "cannot write " + row_path;
            return false;
        }
        for (int i = 0; i <= ncol; ++i) f << i << sep << gen_row() << "\n";
    }
    // frequency of each column degree 0..nrow: int of a power law with xmin 1
    // and exponent 2.2, xmin * (1 - U)^(1 / (1 - alpha)) (pl-dis.hpp),
    // redrawn while above nrow
    std::uniform_real_distribution<double> unif;
    const double ex = 1.0 / (1 - 2.2);
    auto gen_col = [&]() {
        int v = static_cast<int>(1.0 * std::pow(1.0 - unif(rng), ex));
        while (v > nrow) v = static_cast<int>(1.0 * std::pow(1.0 - unif(rng), ex));
        return v;
    };
    {
        std::ofstream f(col_path);
        if (!f) {
            err = "cannot write " + col_path;
            return false;
        }
        for (int i = 0; i <= nrow; ++i) f << i << sep << gen_col() << "\n";
    }
    return true;
}

namespace {

// "degree<sep>frequency" lines until the first empty one
bool read_freqs(const std::string& path, const std::string& sep, std::vector<int>& deg, std::vector<int>& freq,
                std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::string line;
    while (std::getline(f, line)) {
        const size_t p = line.find(sep);
        if (line.empty()) break;
        if (p == std::string::npos) {
            err = path + ": a line without the separator";
            return false;
        }
        deg.push_back(std::stoi(line.substr(0, p)));
        freq.push_back(std::stoi(line.substr(p + sep.size())));
    }
    return true;
}

}  // namespace

bool synthgen(const std::string& row_path, const std::string& col_path, const std::string& out_dir, int32_t nrow,
              int32_t ncol, int32_t dim, int32_t seed, int64_t& train_nnz, int64_t& test_nnz, std::string& err) {
    std::vector<int> row_deg, row_freq, col_deg, col_freq;
    if (!read_freqs(row_path, ",", row_deg, row_freq, err) || !read_freqs(col_path, ",", col_deg, col_freq, err))
        return false;
    if (row_deg.empty() || col_deg.empty() || dim < 1 || nrow < 1 || ncol < 1) {
        err = "synthgen: empty degree file or non-positive size";
        return false;
    }
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    std::discrete_distribution<> col_degree_dist(col_freq.begin(), col_freq.end());
    std::discrete_distribution<> row_degree_dist(row_freq.begin(), row_freq.end());
    std::normal_distribution<> normal_dist(0, 1);
    // each column's weight is a column degree drawn from the column file
    std::vector<int> column_weights;
    column_weights.reserve(size_t(ncol));
    for (int i = 0; i < ncol; ++i) column_weights.push_back(col_deg[size_t(col_degree_dist(rng))]);
    const int64_t positive = std::count_if(column_weights.begin(), column_weights.end(), [](int w) { return w > 0; });
    if (positive == 0) {
        err = "synthgen: every column weight is zero";
        return false;
    }
    std::discrete_distribution<> col_index_dist(column_weights.begin(), column_weights.end());
    std::vector<double> H(size_t(dim) * size_t(ncol));
    for (double& x : H) x = normal_dist(rng);

    // rows of each split: degrees (unsigned int), column indexes (int), values (double)
    std::vector<uint32_t> size_tr, size_te;
    std::vector<int32_t> ind_tr, ind_te;
    std::vector<double> val_tr, val_te;
    std::vector<double> u(static_cast<size_t>(dim));
    train_nnz = test_nnz = 0;
    for (int r = 0; r < nrow; ++r) {
        for (double& x : u) x = normal_dist(rng);
        const int row_degree = std::min(row_deg[size_t(row_degree_dist(rng))] * 5 / 4, ncol);
        const unsigned train_degree = static_cast<int>(row_degree * 0.8);
        const unsigned test_degree = unsigned(row_degree) - train_degree;
        if (int64_t(std::max(train_degree, test_degree)) > positive) {
            err = "synthgen: a row needs more distinct columns than have positive weight";
            return false;
        }
        std::set<int> tr_cols, te_cols;
        for (unsigned j = 0; j < unsigned(row_degree); ++j) {
            for (;;) {
                const int c = col_index_dist(rng);
                std::set<int>& s = j < train_degree ? tr_cols : te_cols;
                if (s.insert(c).second) break;
            }
        }
        for (const auto& [cols, ind, val] :
             {std::make_tuple(&tr_cols, &ind_tr, &val_tr), std::make_tuple(&te_cols, &ind_te, &val_te)}) {
            for (int c : *cols) {
                double t = 0;
                for (int q = 0; q < dim; ++q) t += u[size_t(q)] * H[size_t(dim) * size_t(c) + size_t(q)];
                t += normal_dist(rng) * 0.1;
                ind->push_back(c);
                val->push_back(t);
            }
        }
        size_tr.push_back(train_degree);
        size_te.push_back(test_degree);
        train_nnz += train_degree;
        test_nnz += test_degree;
    }
    // gen.sh: header (class id 1015, rows, cols, int64 nnz), sizes, indexes,
    // values concatenated into train.dat / test.dat
    for (const auto& [name, sz, ind, val, nnz] :
         {std::make_tuple(std::string("train.dat"), &size_tr, &ind_tr, &val_tr, train_nnz),
          std::make_tuple(std::string("test.dat"), &size_te, &ind_te, &val_te, test_nnz)}) {
        std::FILE* f = std::fopen((out_dir + "/" + name).c_str(), "wb");
        if (!f) {
            err = "cannot write " + out_dir + "/" + name;
            return false;
        }
        const int32_t head[3] = {kLongFileClassId, nrow, ncol};
        const int64_t n64 = nnz;
        bool ok = std::fwrite(head, 4, 3, f) == 3 && std::fwrite(&n64, 8, 1, f) == 1 &&
                  std::fwrite(sz->data(), 4, sz->size(), f) == sz->size() &&
                  std::fwrite(ind->data(), 4, ind->size(), f) == ind->size() &&
                  std::fwrite(val->data(), 8, val->size(), f) == val->size();
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) {
            err = "write error on " + out_dir + "/" + name;
            return false;
        }
    }
    return true;
}

}  // namespace nomad
