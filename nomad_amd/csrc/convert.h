// Rating-file tooling: TSV -> .dat conversion and the reference's synthetic
// generators (convert.cpp), checkpoint files (checkpoint.cpp).
#pragma once

#include <cstdint>
#include <string>

namespace nomad {

// Python's random module (MT19937 seeded by init_by_array with the seed's
// 32-bit words, random() = 53-bit genrand_res53), as random.seed(seed) and
// random.random() produce it.
class PyRandom {
public:
    explicit PyRandom(uint64_t seed);
    uint32_t next32();
    double random();

private:
    uint32_t mt_[624];
    int idx_;
};

// Scripts/convert.py: train/test TSV files -> out_dir/{train,test}.dat,
// user_ids.txt, item_ids.txt.
bool convert_tsv(const std::string& train_path, const std::string& test_path, const std::string& out_dir,
                 uint64_t seed, int64_t& nusers, int64_t& nitems, std::string& err);

// Code/nomad/synthgen-conf.cpp: the row and column degree-frequency files.
bool synthgen_conf(int32_t nrow, int32_t ncol, int32_t seed, const std::string& row_path,
                   const std::string& col_path, const std::string& sep, std::string& err);

// Code/nomad/synthgen.cpp + gen.sh: out_dir/train.dat and test.dat (class id
// 1015) from the degree files.
bool synthgen(const std::string& row_path, const std::string& col_path, const std::string& out_dir, int32_t nrow,
              int32_t ncol, int32_t dim, int32_t seed, int64_t& train_nnz, int64_t& test_nnz, std::string& err);

// Checkpoint files of the reference's layout (checkpoint.cpp): W in nparts
// .state files, the columns as .msg records, an empty .smsg.
bool write_checkpoint(const std::string& dir, int job, int epoch, int nparts, int32_t m, int32_t n, int32_t k,
                      const double* W, const double* H, std::string& err);
bool read_checkpoint(const std::string& dir, int job, int epoch, int nparts, int32_t m, int32_t n, int32_t k,
                     double* W, double* H, std::string& err);

}  // namespace nomad
