// Synthetic rating matrices with a planted low-rank truth.
//
// Cells are distinct.  Degree models:
//   uniform : cells drawn uniformly over the m x n grid;
//   zipf    : row i gets degree proportional to rank_r(i)^-row_exp (ranks from a
//             random permutation), clipped to [1, n] and rescaled to hit the
//             exact total; each row then draws its columns without replacement
//             with probability proportional to rank_c(j)^-col_exp.
// Ratings:
//   rounded : clip(round(center + scale * u_i.v_j + eps), lo, hi), eps ~ N(0, noise_std^2)
//   affine  : lo + (hi - lo) * (x - min x) / (max x - min x), x = u_i.v_j + eps
// Truth entries are i.i.d. N(0, factor_var).  A uniformly random subset of
// exactly nnz_test cells is held out as the test matrix.
//
// Every random stream is keyed by (seed, row) so the output does not depend on
// the number of host threads used to make it.
#pragma once

#include <cstdint>

#include "datfile.h"

namespace nomad {

enum SynthDegree : int32_t { kDegreeUniform = 0, kDegreeZipf = 1 };
enum SynthRating : int32_t { kRatingRounded = 0, kRatingAffine = 1 };

struct SynthSpec {
    int32_t m = 0, n = 0;
    int64_t nnz_train = 0, nnz_test = 0;
    int32_t rank = 1;
    double factor_var = 1.0;
    double noise_std = 0.0;
    int32_t degree = kDegreeUniform;
    double row_exp = 1.0, col_exp = 1.0;
    int32_t rating = kRatingRounded;
    double center = 0.0, scale = 1.0, lo = 1.0, hi = 5.0;
    uint64_t seed = 0;
    int32_t threads = 0;  // 0 = hardware concurrency
};

// Fills train and test (row-major, columns ascending within a row).  Returns
// false with a message if the spec cannot be met (e.g. more cells than m*n).
bool synth_generate(const SynthSpec& spec, Csr& train, Csr& test, std::string& err);

}  // namespace nomad
