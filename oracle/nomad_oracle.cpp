// CPU oracle for the NOMAD SGD path -- TEST INFRASTRUCTURE ONLY.
//
// Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
// the checker; nothing under nomad_amd/ links or calls it.
//
// What pins each function (besides agreeing with the GPU path):
//   orc_dat_part_*   reference reader nb_data.cpp:19-176 restated; pinned by
//                    hand-written byte fixtures in tests (both class ids) and
//                    libmf's smalldata fixture converted to the format.
//   orc_init_*       nb_main.cpp:70-98, 444-456, pool.cpp:32-38; pinned by an
//                    independent MT19937-64 (checked against the standard's
//                    10000th output), the generate_canonical formula and a
//                    Python restatement of libstdc++'s std::shuffle (two
//                    swaps per draw, Lemire's bounded draw) in
//                    tests/test_oracle.py.
//   orc_step         closed form checked in tests.
//   orc_sgd_strat    ARITH_REF: nb_main.cpp:514-544 restated; pinned by a
//                    hand-computed one-rating update (k = 2) and by an
//                    independent numpy FP64 replay of the serial order on small
//                    inputs (also of a B = 4 tile order, with the default and
//                    with uneven given row bounds).  ARITH_GPU: the same algorithm with the GPU's
//                    fused update and 64-lane tree; pinned against ARITH_REF
//                    (agreement to rounding) and closed forms.
//   orc_sgd_async    nb_main.cpp:415-578; pinned by invariants (update count =
//   orc_async_*      nnz * epochs, T = 1 equals the serial replay bit-for-bit;
//                    bf16 at T = 1 equals a float32 serial replay rounded by
//                    torch's bf16 conversion, bit-for-bit).  The segments
//                    (nb_main.cpp:109-328): at T = 1, segments cut by update
//                    budgets at arbitrary points equal one uninterrupted serial
//                    replay bit-for-bit, with visit counts = visits (a count
//                    reset by set_perm, or columns restarted out of order,
//                    fails it); at T > 1 the per-thread counts sum to the
//                    updates' column visits and no thread's count of a
//                    column exceeds the column's completed cycles plus the
//                    segments run.
//   orc_sse          nb_main.cpp:601-667; pinned by numpy FP64.
//   orc_loss_grad    sgd-mpi-omp.cpp:122-164 objective; pinned by the
//                    reference DSGD binary's printed objective (oracle/_ref)
//                    and by central finite differences in tests.
//   orc_ccd          ccdr1-mpi-omp.cpp:334-558; pinned by the reference CCD++
//                    binary's printed objective trace (tests/golden).
#include "nomad_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <mutex>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// bf16 helpers (round to nearest even / stochastic by 16 random bits)
uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
uint16_t bf16_sr(float f, uint32_t r16) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += (r16 & 0xFFFFu);
    return uint16_t(u >> 16);
}
float bf16_to_f(uint16_t b) {
    uint32_t u = uint32_t(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
// stochastic-rounding bits: hash of (seed, a, b) XOR a per-lane key
uint32_t sr_bits(uint32_t seed, uint32_t a, uint32_t b, uint32_t lane) {
    const uint32_t lane_key = mix32(seed ^ ((lane + 1u) * 0x85EBCA6Bu));
    return mix32(seed ^ mix32(a * 0x9E3779B1u + b)) ^ lane_key;
}

// ---------------------------------------------------------------------------
// The GPU kernels' dot product: lane l holds elements (2l, 2l+1); per lane
// p = x0*y0 then p = fma(x1, y1, p); the 64 partials are combined as
//   Q_q = (p[4q]+p[4q+1]) + (p[4q+2]+p[4q+3]);  R_r = (Q[4r+3]+Q[4r+2]) + (Q[4r+1]+Q[4r]);
//   total = (R3+R2) + (R1+R0).
// Rows of more than 128 elements (ldf > 128): lane l also holds elements
// (128+2l, 128+2l+1) (zero past the row) and adds p = fma(x2, y2, p),
// p = fma(x3, y3, p) before the tree.
template <class A>
A gpu_dot(const A* h, const A* w, int ldf) {
    A p[64];
    const int npairs = ldf / 2;
    for (int l = 0; l < 64; ++l) {
        if (l < npairs) {
            A t = h[2 * l] * w[2 * l];
            p[l] = std::fma(h[2 * l + 1], w[2 * l + 1], t);
        } else {
            p[l] = A(0);
        }
        if (npairs > 64) {
            const int q = l + 64;
            const bool in = q < npairs;
            p[l] = std::fma(in ? h[2 * q] : A(0), in ? w[2 * q] : A(0), p[l]);
            p[l] = std::fma(in ? h[2 * q + 1] : A(0), in ? w[2 * q + 1] : A(0), p[l]);
        }
    }
    A Q[16], R[4];
    for (int q = 0; q < 16; ++q) Q[q] = (p[4 * q] + p[4 * q + 1]) + (p[4 * q + 2] + p[4 * q + 3]);
    for (int r = 0; r < 4; ++r) R[r] = (Q[4 * r + 3] + Q[4 * r + 2]) + (Q[4 * r + 1] + Q[4 * r]);
    return (R[3] + R[2]) + (R[1] + R[0]);
}

// reference dot: std::inner_product(h, h + k, w, -r) -- a left fold from -r
template <class A>
A ref_err(const A* h, const A* w, int k, A r) {
    A acc = -r;
    for (int d = 0; d < k; ++d) acc = acc + h[d] * w[d];
    return acc;
}

// One rating, reference order (nb_main.cpp:528-538).
template <class A>
void ref_update(A* w, A* h, int k, A r, A step, A reg) {
    const A err = ref_err(h, w, k, r);
    for (int d = 0; d < k; ++d) {
        const A t = w[d];
        w[d] -= step * (err * h[d] + reg * t);
        h[d] -= step * (err * t + reg * h[d]);
    }
}

// One rating, the GPU's fused form (device_common.hpp sgd_step).
template <class A>
void gpu_update(A* w, A* h, int ldf, A r, A step, A reg) {
    const A err = gpu_dot(h, w, ldf) - r;
    for (int d = 0; d < ldf; ++d) {
        const A a = std::fma(err, h[d], reg * w[d]);
        const A b = std::fma(err, w[d], reg * h[d]);
        w[d] = std::fma(-step, a, w[d]);
        h[d] = std::fma(-step, b, h[d]);
    }
}

int ldf_of(int k) { return (k + 1) & ~1; }

}  // namespace

// The exported functions below have C linkage through nomad_oracle.h.

// ---------------------------------------------------------------------------
// nb_data.cpp:19-176
static int read_part(const char* path, int part, int nparts, int* nrows_o, int* ncols_o, int64_t* total_o,
                     int* min_row_o, int* local_rows_o, int64_t* local_nnz_o, int* col_offset, int* row_idx,
                     double* row_val) {
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f) return -1;
    int id = 0;
    if (!f.read(reinterpret_cast<char*>(&id), 4)) return -2;
    int nrows, ncols;
    long long total;
    if (id == 1211216) {
        int h[3];
        if (!f.read(reinterpret_cast<char*>(h), 12)) return -2;
        nrows = h[0];
        ncols = h[1];
        total = h[2];
    } else if (id == 1015) {
        int h[2];
        if (!f.read(reinterpret_cast<char*>(h), 8)) return -2;
        nrows = h[0];
        ncols = h[1];
        if (!f.read(reinterpret_cast<char*>(&total), 8)) return -2;
    } else {
        return -3;
    }
    const int per = nrows / nparts + ((nrows % nparts > 0) ? 1 : 0);
    const int min_row = per * part;
    const int max_row = std::min(per * (part + 1), nrows);
    const int local_rows = std::max(0, max_row - min_row);
    std::vector<int> row_nnz(nrows);
    if (!f.read(reinterpret_cast<char*>(row_nnz.data()), std::streamsize(nrows) * 4)) return -2;
    long long begin_skip = 0, local = 0;
    for (int i = 0; i < std::min(min_row, nrows); ++i) begin_skip += row_nnz[i];
    for (int i = min_row; i < max_row; ++i) local += row_nnz[i];
    const long long end_skip = total - local - begin_skip;
    if (nrows_o) *nrows_o = nrows;
    if (ncols_o) *ncols_o = ncols;
    if (total_o) *total_o = total;
    if (min_row_o) *min_row_o = min_row;
    if (local_rows_o) *local_rows_o = local_rows;
    if (local_nnz_o) *local_nnz_o = local;
    if (!col_offset) return 0;
    f.seekg(begin_skip * 4, std::ios_base::cur);
    std::vector<int> col(local);
    if (!f.read(reinterpret_cast<char*>(col.data()), std::streamsize(local) * 4)) return -2;
    f.seekg(end_skip * 4 + begin_skip * 8, std::ios_base::cur);
    std::vector<double> val(local);
    if (!f.read(reinterpret_cast<char*>(val.data()), std::streamsize(local) * 8)) return -2;
    std::vector<std::vector<int>> ri(ncols);
    std::vector<std::vector<double>> rv(ncols);
    long long e = 0;
    for (int i = min_row; i < max_row; ++i) {
        for (int t = 0; t < row_nnz[i]; ++t, ++e) {
            ri[col[e]].push_back(i - min_row);
            rv[col[e]].push_back(val[e]);
        }
    }
    int off = 0;
    col_offset[0] = 0;
    for (int j = 0; j < ncols; ++j) {
        std::copy(ri[j].begin(), ri[j].end(), row_idx + off);
        std::copy(rv[j].begin(), rv[j].end(), row_val + off);
        off += int(ri[j].size());
        col_offset[j + 1] = off;
    }
    return 0;
}

int orc_dat_part_info(const char* path, int part, int nparts, int* nrows, int* ncols, int64_t* total_nnz,
                      int* min_row, int* local_rows, int64_t* local_nnz) {
    return read_part(path, part, nparts, nrows, ncols, total_nnz, min_row, local_rows, local_nnz, nullptr,
                     nullptr, nullptr);
}

int orc_dat_part_read(const char* path, int part, int nparts, int* col_offset, int* row_idx, double* row_val) {
    return read_part(path, part, nparts, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, col_offset,
                     row_idx, row_val);
}

// ---------------------------------------------------------------------------
// nb_main.cpp:444-456 (W, per thread part) and 70-93 (H, one rank)
int orc_init_factors(int m, int n, int k, int nthreads, uint64_t seed, double* W, double* H) {
    if (k <= 0 || nthreads <= 0) return -1;
    const int num_parts = nthreads;  // one rank
    const int per = m / num_parts + ((m % num_parts > 0) ? 1 : 0);
    for (int t = 0; t < nthreads && W; ++t) {
        const int min_row = per * t;
        const int max_row = std::min(per * (t + 1), m);
        const int local_rows = std::max(0, max_row - min_row);
        std::mt19937_64 rng(seed + 0 * 131 + t + 1);
        std::uniform_real_distribution<double> init_dist(0, 1.0 / std::sqrt(double(k)));
        double* latent_rows = W + int64_t(min_row) * k;
        for (int64_t i = 0; i < int64_t(local_rows) * k; ++i) latent_rows[i] = init_dist(rng);
    }
    if (H) {
        std::mt19937_64 rng(seed + 0 * 131 + 139);
        std::uniform_real_distribution<double> init_dist(0, 1.0 / std::sqrt(double(k)));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < k; ++j) H[int64_t(i) * k + j] = init_dist(rng);
    }
    return 0;
}

// pool.cpp:32-38 with rng_temp(seed + rank + 137), nb_main.cpp:79-88
int orc_init_perms(int n, int nthreads, uint64_t seed, int* perms) {
    std::mt19937_64 rng_temp(seed + 0 + 137);
    for (int c = 0; c < n; ++c) {
        int* perm = perms + int64_t(c) * nthreads;
        for (int i = 0; i < nthreads; ++i) perm[i] = i;
        std::shuffle(perm, perm + nthreads, rng_temp);
    }
    return 0;
}

// nb_main.cpp:514-515 (double build); the float build (nomad_body.hpp:313-320)
// holds lrate/drate as float and rounds the step to float.
double orc_step(int storage, double lrate, double drate, int count) {
    if (storage == ORC_FP64)
        return lrate * 1.5 / (1.0 + drate * std::pow(count + 1, 1.5));
    const float lr = float(lrate), dr = float(drate);
    const float s = float(lr * 1.5 / (1.0 + dr * std::pow(count + 1, 1.5)));
    return double(s);
}

// ---------------------------------------------------------------------------
template <class A>
static void strat_replay(int arith, int m, int n, int k, const int64_t* colptr, const int32_t* rowidx,
                         const double* val, std::vector<A>& W, std::vector<A>& H, int B, const int32_t* rbound,
                         int epoch0, int nepochs, double lrate, double drate, double reg, int storage) {
    const int ldf = ldf_of(k);
    const int cpb = n / B + ((n % B) > 0 ? 1 : 0);
    const A regA = storage == ORC_FP64 ? A(reg) : A(float(reg));
    for (int e = 0; e < nepochs; ++e) {
        const A step = A(orc_step(storage, lrate, drate, epoch0 + e));
        for (int s = 0; s < B; ++s) {
            for (int b = 0; b < B; ++b) {
                const int cb = (b + s) % B;
                const int rlo = rbound[b], rhi = rbound[b + 1];
                const int clo = cb * cpb, chi = std::min(n, clo + cpb);
                for (int j = clo; j < chi; ++j) {
                    A* h = &H[int64_t(j) * ldf];
                    for (int64_t o = colptr[j]; o < colptr[j + 1]; ++o) {
                        const int i = rowidx[o];
                        if (i < rlo || i >= rhi) continue;
                        A* w = &W[int64_t(i) * ldf];
                        const A r = A(val[o]);
                        if (arith == ORC_ARITH_REF)
                            ref_update(w, h, k, r, step, regA);
                        else
                            gpu_update(w, h, ldf, r, step, regA);
                    }
                }
            }
        }
    }
}

// bf16 storage: rows held as bf16, arithmetic in float with the GPU order; h
// stays in float for the whole (column, row block) visit and is rounded when
// the visit ends; w is rounded after each update.
static void strat_replay_bf16(int rounding, uint32_t sr_seed, int m, int n, int k, const int64_t* colptr,
                              const int32_t* rowidx, const double* val, std::vector<uint16_t>& W,
                              std::vector<uint16_t>& H, int B, const int32_t* rbound, int epoch0, int nepochs,
                              double lrate, double drate, double reg) {
    const int ldf = ldf_of(k);
    const int npairs = ldf / 2;
    const int cpb = n / B + ((n % B) > 0 ? 1 : 0);
    const float regf = float(reg);
    std::vector<float> h(ldf), w(ldf);
    auto store = [&](uint16_t* dst, const float* src, uint32_t a, uint32_t bkey) {
        for (int l = 0; l < npairs; ++l) {
            // (pairs 64.. of a wide row: the bits of pair l - 64, rehashed)
            uint32_t rnd = l < 64 ? sr_bits(sr_seed, a, bkey, uint32_t(l))
                                  : mix32(sr_bits(sr_seed, a, bkey, uint32_t(l - 64)) ^ 0x5BD1E995u);
            if (rounding == ORC_ROUND_STOCHASTIC) {
                dst[2 * l] = bf16_sr(src[2 * l], rnd);
                dst[2 * l + 1] = bf16_sr(src[2 * l + 1], rnd >> 16);
            } else {
                dst[2 * l] = bf16_rne(src[2 * l]);
                dst[2 * l + 1] = bf16_rne(src[2 * l + 1]);
            }
        }
    };
    for (int e = 0; e < nepochs; ++e) {
        const int count = epoch0 + e;
        const float step = float(orc_step(ORC_FP32, lrate, drate, count));
        for (int s = 0; s < B; ++s) {
            for (int b = 0; b < B; ++b) {
                const int cb = (b + s) % B;
                const int rlo = rbound[b], rhi = rbound[b + 1];
                const int clo = cb * cpb, chi = std::min(n, clo + cpb);
                for (int j = clo; j < chi; ++j) {
                    bool any = false;
                    const uint32_t bkey = uint32_t(j) * 0x01000193u + uint32_t(count);
                    for (int64_t o = colptr[j]; o < colptr[j + 1]; ++o) {
                        const int i = rowidx[o];
                        if (i < rlo || i >= rhi) continue;
                        if (!any) {
                            for (int d = 0; d < ldf; ++d) h[d] = bf16_to_f(H[int64_t(j) * ldf + d]);
                            any = true;
                        }
                        for (int d = 0; d < ldf; ++d) w[d] = bf16_to_f(W[int64_t(i) * ldf + d]);
                        gpu_update(w.data(), h.data(), ldf, float(val[o]), step, regf);
                        store(&W[int64_t(i) * ldf], w.data(), uint32_t(i), bkey);
                    }
                    if (any) store(&H[int64_t(j) * ldf], h.data(), 0x80000000u | uint32_t(b), bkey);
                }
            }
        }
    }
}

// One element pair as the kernels store it in BF16: element 2l from the low
// 16 bits of sr_bits(seed, a, b, l), element 2l+1 from the high 16 bits
// (stochastic), or round to nearest even.
void orc_bf16_store_pair(float x0, float x1, int rounding, uint32_t seed, uint32_t a, uint32_t b, uint32_t lane,
                         float* out) {
    const uint32_t rnd = sr_bits(seed, a, b, lane);
    const uint16_t u0 = rounding == ORC_ROUND_STOCHASTIC ? bf16_sr(x0, rnd) : bf16_rne(x0);
    const uint16_t u1 = rounding == ORC_ROUND_STOCHASTIC ? bf16_sr(x1, rnd >> 16) : bf16_rne(x1);
    out[0] = bf16_to_f(u0);
    out[1] = bf16_to_f(u1);
}

int64_t orc_sgd_strat(int storage, int arith, int rounding, uint32_t sr_seed, int m, int n, int k,
                      const int64_t* colptr, const int32_t* rowidx, const double* val, double* W, double* H,
                      int B, const int32_t* row_bounds, int epoch0, int nepochs, double lrate, double drate,
                      double reg) {
    if (m <= 0 || n <= 0 || k <= 0 || B <= 0 || nepochs < 0) return -1;
    const int ldf = ldf_of(k);
    // row block b = rows [rb[b], rb[b + 1]): the caller's bounds, or ceil(m / B) rows each
    std::vector<int32_t> rb(size_t(B) + 1);
    const int rpb = m / B + ((m % B) > 0 ? 1 : 0);
    for (int b = 0; b <= B; ++b) rb[size_t(b)] = row_bounds ? row_bounds[b] : std::min(m, b * rpb);
    if (rb[0] != 0 || rb[size_t(B)] != m) return -1;
    for (int b = 0; b < B; ++b)
        if (rb[size_t(b) + 1] < rb[size_t(b)]) return -1;
    if (storage == ORC_FP32 || storage == ORC_FP64) {
        auto run = [&](auto zero) {
            using A = decltype(zero);
            std::vector<A> w(size_t(m) * ldf, A(0)), h(size_t(n) * ldf, A(0));
            for (int64_t i = 0; i < m; ++i)
                for (int d = 0; d < k; ++d) w[i * ldf + d] = A(W[i * k + d]);
            for (int64_t j = 0; j < n; ++j)
                for (int d = 0; d < k; ++d) h[j * ldf + d] = A(H[j * k + d]);
            strat_replay<A>(arith, m, n, k, colptr, rowidx, val, w, h, B, rb.data(), epoch0, nepochs, lrate, drate,
                            reg, storage);
            for (int64_t i = 0; i < m; ++i)
                for (int d = 0; d < k; ++d) W[i * k + d] = double(w[i * ldf + d]);
            for (int64_t j = 0; j < n; ++j)
                for (int d = 0; d < k; ++d) H[j * k + d] = double(h[j * ldf + d]);
        };
        if (storage == ORC_FP32)
            run(0.0f);
        else
            run(0.0);
    } else if (storage == ORC_BF16) {
        std::vector<uint16_t> w(size_t(m) * ldf, 0), h(size_t(n) * ldf, 0);
        for (int64_t i = 0; i < m; ++i)
            for (int d = 0; d < k; ++d) w[i * ldf + d] = bf16_rne(float(W[i * k + d]));
        for (int64_t j = 0; j < n; ++j)
            for (int d = 0; d < k; ++d) h[j * ldf + d] = bf16_rne(float(H[j * k + d]));
        strat_replay_bf16(rounding, sr_seed, m, n, k, colptr, rowidx, val, w, h, B, rb.data(), epoch0, nepochs,
                          lrate, drate, reg);
        for (int64_t i = 0; i < m; ++i)
            for (int d = 0; d < k; ++d) W[i * k + d] = bf16_to_f(w[i * ldf + d]);
        for (int64_t j = 0; j < n; ++j)
            for (int d = 0; d < k; ++d) H[j * k + d] = bf16_to_f(h[j * ldf + d]);
    } else {
        return -1;
    }
    return colptr[n] * int64_t(nepochs);
}

// ---------------------------------------------------------------------------
// Asynchronous NOMAD with T CPU threads, one rank (nb_main.cpp:415-578), as a
// state kept across timeout segments (nb_main.cpp:109-328).
namespace {

struct Token {
    int col;
    int pos;
    int64_t cycles;        // completed route cycles (pos wrapped to 0)
    std::vector<int> perm;  // ColumnData::perm_ (pool.h:11-37)
};

struct TokQueue {
    std::mutex mu;
    std::deque<Token*> q;
    void push(Token* t) {
        std::lock_guard<std::mutex> g(mu);
        q.push_back(t);
    }
    bool try_pop(Token*& t) {
        std::lock_guard<std::mutex> g(mu);
        if (q.empty()) return false;
        t = q.front();
        q.pop_front();
        return true;
    }
};

struct AsyncBase {
    virtual ~AsyncBase() {}
    virtual int segment(double seconds, int64_t max_updates, int64_t epochs, orc_async_stats* st) = 0;
    virtual void get(double* W, double* H) const = 0;
    virtual void counts(int32_t* out) const = 0;
};

template <class A>
struct AsyncState : AsyncBase {
    int T, m, n, k, storage;
    double lrate, drate;
    A regA;
    // per-thread CSC with local rows (nb_data.cpp:131-169), W block and the
    // thread's visit count per column (col_update_counts, nb_main.cpp:458-459)
    struct Part {
        int min_row = 0, rows = 0;
        std::vector<int64_t> off;
        std::vector<int> ri;
        std::vector<A> rv;
        std::vector<A> w;
        std::vector<int32_t> counts;
    };
    std::vector<Part> parts;
    std::vector<Token> toks;
    std::vector<std::vector<A>> hv;
    std::vector<TokQueue> queues;
    std::vector<Token*> parked;
    std::mt19937_64 rng;  // NomadBody::rng, mt19937_64(seed + rank * 131 + 139) (nb_init.cpp:219)

    // bf16 storage (A = float): every stored factor value is rounded to the
    // nearest bf16 -- the initial factors, a row after each of its updates and
    // a column at the end of each visit; arithmetic stays fp32
    void round_store(A* v, int len) const {
        if (storage == ORC_BF16)
            for (int d = 0; d < len; ++d) v[d] = A(bf16_to_f(bf16_rne(float(v[d]))));
    }

    AsyncState(int storage_, int T_, int m_, int n_, int k_, const int64_t* colptr, const int32_t* rowidx,
               const double* val, const double* W, const double* H, uint64_t seed, double lrate_, double drate_,
               double reg)
        : T(T_), m(m_), n(n_), k(k_), storage(storage_), lrate(lrate_), drate(drate_),
          regA(storage_ == ORC_FP64 ? A(reg) : A(float(reg))), parts(T_), toks(n_), hv(n_), queues(T_),
          rng(seed + 139) {
        const int per = m / T + ((m % T) > 0 ? 1 : 0);  // nb_data.cpp:83
        for (int t = 0; t < T; ++t) {
            Part& P = parts[t];
            P.min_row = std::min(per * t, m);
            const int max_row = std::min(per * (t + 1), m);
            P.rows = std::max(0, max_row - P.min_row);
            P.off.assign(size_t(n) + 1, 0);
            P.w.resize(size_t(P.rows) * k);
            for (int64_t e = 0; e < int64_t(P.rows) * k; ++e) P.w[e] = A(W[int64_t(P.min_row) * k + e]);
            for (int r = 0; r < P.rows; ++r) round_store(&P.w[int64_t(r) * k], k);
            P.counts.assign(n, 0);
        }
        // one pass: count, then place (rows stay ascending inside a column)
        for (int j = 0; j < n; ++j)
            for (int64_t o = colptr[j]; o < colptr[j + 1]; ++o) parts[rowidx[o] / per].off[j + 1]++;
        for (int t = 0; t < T; ++t) {
            Part& P = parts[t];
            for (int j = 0; j < n; ++j) P.off[j + 1] += P.off[j];
            P.ri.resize(P.off[n]);
            P.rv.resize(P.off[n]);
        }
        {
            std::vector<std::vector<int64_t>> fill(T);
            for (int t = 0; t < T; ++t) fill[t].assign(parts[t].off.begin(), parts[t].off.end() - 1);
            for (int j = 0; j < n; ++j)
                for (int64_t o = colptr[j]; o < colptr[j + 1]; ++o) {
                    const int t = rowidx[o] / per;
                    const int64_t d = fill[t][j]++;
                    parts[t].ri[d] = rowidx[o] - parts[t].min_row;
                    parts[t].rv[d] = A(val[o]);
                }
        }
        // initial placement: set_perm with mt19937_64(seed + rank + 137), columns
        // in order, each pushed to the queue of its first thread (nb_main.cpp:70-97)
        std::mt19937_64 rng_temp(seed + 137);
        for (int j = 0; j < n; ++j) {
            Token& tk = toks[j];
            tk.col = j;
            tk.pos = 0;
            tk.cycles = 0;
            tk.perm.resize(T);
            for (int i = 0; i < T; ++i) tk.perm[i] = i;
            std::shuffle(tk.perm.begin(), tk.perm.end(), rng_temp);
            hv[j].resize(k);
            for (int d = 0; d < k; ++d) hv[j][d] = A(H[int64_t(j) * k + d]);
            round_store(hv[j].data(), k);
            queues[tk.perm[tk.pos]].push(&tk);
        }
    }

    int segment(double seconds, int64_t max_updates, int64_t epochs, orc_async_stats* st) override {
        int64_t target = INT64_MAX;
        if (epochs > 0) {
            int64_t mn = INT64_MAX;
            for (const Token& t : toks) mn = std::min(mn, t.cycles);
            target = (n ? mn : 0) + epochs;
        }
        std::atomic<bool> stop{false};
        std::atomic<int> nparked{0};
        std::atomic<int64_t> total_updates{0}, total_fail{0};
        std::mutex park_mu;
        const auto t0 = std::chrono::steady_clock::now();
        auto worker = [&](int t) {
            Part& P = parts[t];
            int64_t nupd = 0, nfail = 0, npoll = 0;
            // the stop flag is read before each pop, so a visit always completes
            // (flag_train_stop, nb_main.cpp:472, 480-570)
            while (!stop.load(std::memory_order_relaxed)) {
                Token* tk;
                if (queues[t].try_pop(tk)) {
                    const int j = tk->col;
                    const A step = A(orc_step(storage, lrate, drate, P.counts[j]));  // nb_main.cpp:513-515
                    A* col = hv[j].data();
                    int64_t vis = 0;
                    for (int64_t o = P.off[j]; o < P.off[j + 1]; ++o) {
                        A* row = &P.w[int64_t(P.ri[o]) * k];
                        ref_update(row, col, k, P.rv[o], step, regA);
                        round_store(row, k);
                        ++vis;
                    }
                    round_store(col, k);
                    nupd += vis;
                    P.counts[j]++;  // nb_main.cpp:544
                    tk->pos++;      // nb_main.cpp:552-565 at one rank
                    bool park = false;
                    if (tk->pos >= T) {
                        tk->pos = 0;
                        tk->cycles++;
                        park = tk->cycles >= target;
                    }
                    if (park) {
                        {
                            std::lock_guard<std::mutex> g(park_mu);
                            parked.push_back(tk);
                        }
                        if (nparked.fetch_add(1) + 1 == n) stop = true;
                    } else {
                        queues[tk->perm[tk->pos % T]].push(tk);
                    }
                    if (max_updates > 0 && total_updates.fetch_add(vis) + vis >= max_updates) stop = true;
                } else {
                    ++nfail;
                    std::this_thread::yield();
                }
                if (seconds > 0 && (++npoll & 63) == 0) {
                    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                    if (el >= seconds) stop = true;
                }
            }
            if (max_updates <= 0) total_updates += nupd;
            total_fail += nfail;
        };
        if (T == 1) {
            worker(0);
        } else {
            std::vector<std::thread> th;
            for (int t = 0; t < T; ++t) th.emplace_back(worker, t);
            if (seconds > 0) {
                while (!stop) {
                    std::this_thread::sleep_for(std::chrono::milliseconds(1));
                    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                    if (el >= seconds) stop = true;
                }
            }
            for (auto& x : th) x.join();
        }
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        // gather every column, queues in thread order (nb_main.cpp:140-178), then
        // parked ones; each starts a fresh route: set_perm (pool.cpp:32-38) with
        // the member generator, pushed to its first thread (nb_main.cpp:313-321)
        std::vector<Token*> all;
        for (int t = 0; t < T; ++t) {
            Token* tk;
            while (queues[t].try_pop(tk)) all.push_back(tk);
        }
        for (Token* tk : parked) all.push_back(tk);
        parked.clear();
        for (Token* tk : all) {
            tk->pos = 0;
            for (int i = 0; i < T; ++i) tk->perm[i] = i;
            std::shuffle(tk->perm.begin(), tk->perm.end(), rng);
            queues[tk->perm[0]].push(tk);
        }
        int64_t mn = INT64_MAX;
        for (const Token& t : toks) mn = std::min(mn, t.cycles);
        if (st) {
            st->updates = total_updates;
            st->failures = total_fail;
            st->seconds = el;
            st->min_cycles = n ? mn : 0;
        }
        return int(all.size()) == n ? 0 : -2;
    }

    void get(double* W, double* H) const override {
        if (W)
            for (int t = 0; t < T; ++t)
                for (int64_t e = 0; e < int64_t(parts[t].rows) * k; ++e)
                    W[int64_t(parts[t].min_row) * k + e] = double(parts[t].w[e]);
        if (H)
            for (int j = 0; j < n; ++j)
                for (int d = 0; d < k; ++d) H[int64_t(j) * k + d] = double(hv[j][d]);
    }

    void counts(int32_t* out) const override {
        for (int t = 0; t < T; ++t) std::copy(parts[t].counts.begin(), parts[t].counts.end(), out + int64_t(t) * n);
    }
};

}  // namespace

struct orc_async {
    AsyncBase* impl;
};

orc_async* orc_async_create(int storage, int T, int m, int n, int k, const int64_t* colptr, const int32_t* rowidx,
                            const double* val, const double* W, const double* H, uint64_t seed, double lrate,
                            double drate, double reg) {
    if (T <= 0 || m <= 0 || n <= 0 || k <= 0) return nullptr;
    AsyncBase* b = nullptr;
    if (storage == ORC_FP32 || storage == ORC_BF16)
        b = new AsyncState<float>(storage, T, m, n, k, colptr, rowidx, val, W, H, seed, lrate, drate, reg);
    else if (storage == ORC_FP64)
        b = new AsyncState<double>(storage, T, m, n, k, colptr, rowidx, val, W, H, seed, lrate, drate, reg);
    else
        return nullptr;
    return new orc_async{b};
}

int orc_async_segment(orc_async* s, double seconds, int64_t max_updates, int64_t epochs, orc_async_stats* stats) {
    if (!s || (!(seconds > 0) && max_updates <= 0 && epochs <= 0)) return -1;
    return s->impl->segment(seconds, max_updates, epochs, stats);
}

int orc_async_get(const orc_async* s, double* W, double* H) {
    if (!s) return -1;
    s->impl->get(W, H);
    return 0;
}

int orc_async_counts(const orc_async* s, int32_t* counts) {
    if (!s || !counts) return -1;
    s->impl->counts(counts);
    return 0;
}

void orc_async_destroy(orc_async* s) {
    if (!s) return;
    delete s->impl;
    delete s;
}

int orc_sgd_async(int storage, int T, int m, int n, int k, const int64_t* colptr, const int32_t* rowidx,
                  const double* val, double* W, double* H, uint64_t seed, double lrate, double drate,
                  double reg, double seconds, int64_t epochs, orc_async_stats* stats) {
    if (!(seconds > 0) && epochs <= 0) return -1;
    orc_async* s = orc_async_create(storage, T, m, n, k, colptr, rowidx, val, W, H, seed, lrate, drate, reg);
    if (!s) return -1;
    const int rc = orc_async_segment(s, seconds, 0, epochs, stats);
    orc_async_get(s, W, H);
    orc_async_destroy(s);
    return rc;
}

// ---------------------------------------------------------------------------
double orc_sse(int storage, int arith, int m, int n, int k, const int64_t* colptr, const int32_t* rowidx,
               const double* val, const double* W, const double* H) {
    (void)m;
    const int ldf = ldf_of(k);
    double sse = 0.0;
    auto run = [&](auto zero) {
        using A = decltype(zero);
        std::vector<A> h(ldf), w(ldf);
        for (int j = 0; j < n; ++j) {
            for (int d = 0; d < ldf; ++d) h[d] = d < k ? A(H[int64_t(j) * k + d]) : A(0);
            for (int64_t o = colptr[j]; o < colptr[j + 1]; ++o) {
                const int i = rowidx[o];
                for (int d = 0; d < ldf; ++d) w[d] = d < k ? A(W[int64_t(i) * k + d]) : A(0);
                A err;
                if (arith == ORC_ARITH_REF) {
                    // nb_main.cpp:645-648: cur_error = -r; cur_error += col[i]*row[i]
                    err = -A(val[o]);
                    for (int d = 0; d < k; ++d) err += h[d] * w[d];
                } else {
                    err = gpu_dot(h.data(), w.data(), ldf) - A(val[o]);
                }
                sse += double(err) * double(err);
            }
        }
    };
    if (storage == ORC_FP64)
        run(0.0);
    else
        run(0.0f);  // fp32 and bf16 (values already rounded by the caller)
    return sse;
}

// ---------------------------------------------------------------------------
double orc_loss_grad(int m, int n, int k, const int64_t* colptr, const int32_t* rowidx, const double* val,
                     const double* W, const double* H, double reg, double* gW, double* gH) {
    std::vector<int64_t> degr(m, 0), degc(n, 0);
    if (gW) std::fill(gW, gW + int64_t(m) * k, 0.0);
    if (gH) std::fill(gH, gH + int64_t(n) * k, 0.0);
    double loss = 0.0;
    for (int j = 0; j < n; ++j) {
        const double* h = H + int64_t(j) * k;
        for (int64_t o = colptr[j]; o < colptr[j + 1]; ++o) {
            const int i = rowidx[o];
            const double* w = W + int64_t(i) * k;
            double dot = 0.0;
            for (int d = 0; d < k; ++d) dot += w[d] * h[d];
            const double e = dot - val[o];
            loss += e * e;
            degr[i]++;
            degc[j]++;
            if (gW)
                for (int d = 0; d < k; ++d) gW[int64_t(i) * k + d] += 2.0 * e * h[d];
            if (gH)
                for (int d = 0; d < k; ++d) gH[int64_t(j) * k + d] += 2.0 * e * w[d];
        }
    }
    double regsum = 0.0;
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int d = 0; d < k; ++d) s += W[int64_t(i) * k + d] * W[int64_t(i) * k + d];
        regsum += double(degr[i]) * s;
        if (gW)
            for (int d = 0; d < k; ++d) gW[int64_t(i) * k + d] += 2.0 * reg * double(degr[i]) * W[int64_t(i) * k + d];
    }
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int d = 0; d < k; ++d) s += H[int64_t(j) * k + d] * H[int64_t(j) * k + d];
        regsum += double(degc[j]) * s;
        if (gH)
            for (int d = 0; d < k; ++d) gH[int64_t(j) * k + d] += 2.0 * reg * double(degc[j]) * H[int64_t(j) * k + d];
    }
    return loss + reg * regsum;
}

// ---------------------------------------------------------------------------
// CCD++ (ccdr1-mpi-omp.cpp:334-558), one process.  Rhat = R - W H^T is kept
// for every rating in CSC order and in CSR order.
static double rank_one_update(const int64_t* ptr, const int32_t* idx, const double* rhat, int c,
                              const std::vector<double>& u, double lambda, double vj, double* redvar) {
    // ccdr1-mpi-omp.cpp:334-358 (do_nmf = 0)
    double g = 0, h = lambda;
    if (ptr[c + 1] == ptr[c]) return 0;
    for (int64_t e = ptr[c]; e < ptr[c + 1]; ++e) {
        const int i = idx[e];
        g += u[i] * rhat[e];
        h += u[i] * u[i];
    }
    const double newvj = g / h;
    const double delta = vj - newvj;
    *redvar += h * delta * delta;
    return newvj;
}

int orc_ccd(int m, int n, int k, const int64_t* colptr, const int32_t* rowidx, const double* val, double* W,
            double* H, double lambda, int iters, int inner, double eps, double* obj, const int64_t* te_colptr,
            const int32_t* te_rowidx, const double* te_val, double* rmse) {
    if (m <= 0 || n <= 0 || k <= 0 || iters < 0 || inner <= 0) return -1;
    // test RMSE of the current factors (the reference keeps the test residual
    // up to date with UpdateRating, ccdr1-mpi-omp.cpp:444-447, 517-519)
    auto test_rmse = [&]() {
        double s = 0;
        for (int j = 0; j < n; ++j)
            for (int64_t e = te_colptr[j]; e < te_colptr[j + 1]; ++e) {
                double dot = 0;
                for (int d = 0; d < k; ++d) dot += W[int64_t(te_rowidx[e]) * k + d] * H[int64_t(j) * k + d];
                const double r = te_val[e] - dot;
                s += r * r;
            }
        return std::sqrt(s / double(te_colptr[n]));
    };
    if (rmse)
        for (int64_t q = 0; q < int64_t(iters) * k; ++q) rmse[q] = std::nan("");
    const int64_t nnz = colptr[n];
    // CSR copy with the CSC->CSR position map
    std::vector<int64_t> rowptr(size_t(m) + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) rowptr[rowidx[e] + 1]++;
    for (int i = 0; i < m; ++i) rowptr[i + 1] += rowptr[i];
    std::vector<int32_t> colidx(nnz);
    std::vector<int64_t> pos(nnz);
    {
        std::vector<int64_t> fill(rowptr.begin(), rowptr.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) {
                const int64_t d = fill[rowidx[e]]++;
                colidx[d] = j;
                pos[e] = d;
            }
    }
    std::vector<double> rc(nnz), rr(nnz);
    for (int j = 0; j < n; ++j)
        for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) {
            const int i = rowidx[e];
            double dot = 0;
            for (int d = 0; d < k; ++d) dot += W[int64_t(i) * k + d] * H[int64_t(j) * k + d];
            rc[e] = val[e] - dot;
            rr[pos[e]] = rc[e];
        }
    auto nnz_col = [&](int j) { return double(colptr[j + 1] - colptr[j]); };
    auto nnz_row = [&](int i) { return double(rowptr[i + 1] - rowptr[i]); };
    double reg = 0;
    for (int t = 0; t < k; ++t) {
        for (int i = 0; i < m; ++i) reg += W[int64_t(i) * k + t] * W[int64_t(i) * k + t] * nnz_row(i);
        for (int j = 0; j < n; ++j) reg += H[int64_t(j) * k + t] * H[int64_t(j) * k + t] * nnz_col(j);
    }
    std::vector<double> u(m), v(n);
    if (obj)
        for (int64_t q = 0; q < int64_t(iters) * k; ++q) obj[q] = std::nan("");
    for (int oiter = 1; oiter <= iters; ++oiter) {
        // both restart with every outer iteration (ccdr1-mpi-omp.cpp:424-428)
        double fundec_max = 0;
        int early_stop = 0;
        for (int t = 0; t < k; ++t) {
            if (early_stop >= 5) break;
            for (int i = 0; i < m; ++i) u[i] = W[int64_t(i) * k + t];
            for (int j = 0; j < n; ++j) v[j] = H[int64_t(j) * k + t];
            // Rhat += u v^T (ccdr1-mpi-omp.cpp:444-447, UpdateRating add)
            for (int j = 0; j < n; ++j)
                for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) {
                    rc[e] += u[rowidx[e]] * v[j];
                    rr[pos[e]] += v[j] * u[rowidx[e]];
                }
            for (int i = 0; i < m; ++i) reg -= u[i] * u[i] * nnz_row(i);
            for (int j = 0; j < n; ++j) reg -= v[j] * v[j] * nnz_col(j);
            for (int iter = 1; iter <= inner; ++iter) {
                double fundec_cur = 0;
                for (int j = 0; j < n; ++j)
                    v[j] = rank_one_update(colptr, rowidx, rc.data(), j, u, lambda * nnz_col(j), v[j], &fundec_cur);
                for (int i = 0; i < m; ++i)
                    u[i] = rank_one_update(rowptr.data(), colidx.data(), rr.data(), i, v, lambda * nnz_row(i), u[i],
                                           &fundec_cur);
                if (fundec_cur < fundec_max * eps) {
                    if (iter == 1) early_stop += 1;
                    break;
                }
                if (!(oiter == 1 && t == 0 && iter == 1)) fundec_max = std::max(fundec_max, fundec_cur);
            }
            for (int i = 0; i < m; ++i) W[int64_t(i) * k + t] = u[i];
            for (int j = 0; j < n; ++j) H[int64_t(j) * k + t] = v[j];
            // Rhat -= u v^T; the printed loss is the row-order sum (UpdateRating on Rt)
            for (int j = 0; j < n; ++j)
                for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) rc[e] -= u[rowidx[e]] * v[j];
            double loss = 0;
            for (int i = 0; i < m; ++i)
                for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
                    rr[e] -= v[colidx[e]] * u[i];
                    loss += rr[e] * rr[e];
                }
            for (int i = 0; i < m; ++i) reg += u[i] * u[i] * nnz_row(i);
            for (int j = 0; j < n; ++j) reg += v[j] * v[j] * nnz_col(j);
            if (obj) obj[int64_t(oiter - 1) * k + t] = loss + reg * lambda;
            if (rmse && te_colptr && te_colptr[n] > 0) rmse[int64_t(oiter - 1) * k + t] = test_rmse();
        }
    }
    return 0;
}


