// Synthetic rating matrices with a planted low-rank truth (see synth.h).
#include "synth.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <thread>
#include <vector>

namespace nomad {

namespace {

inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** (Blackman & Vigna), seeded through splitmix64.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double u01() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double u01_open() { return (double(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t bound) {  // Lemire's multiply-shift (slightly biased for huge bounds)
        return uint64_t((__uint128_t(next()) * bound) >> 64);
    }
    double normal() {  // Box-Muller, one value per call
        double u1 = u01_open(), u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline uint64_t stream_key(uint64_t seed, uint64_t tag, uint64_t idx) {
    uint64_t x = seed ^ (tag * 0xD1B54A32D192ED03ull);
    uint64_t a = splitmix64(x);
    x = a ^ (idx * 0x9E3779B97F4A7C15ull);
    return splitmix64(x);
}

template <class F>
void parallel_for(int64_t count, int threads, F&& fn) {
    if (threads <= 1 || count < 1024) {
        for (int64_t i = 0; i < count; ++i) fn(i, 0);
        return;
    }
    std::atomic<int64_t> next{0};
    const int64_t chunk = std::max<int64_t>(64, count / (int64_t(threads) * 64));
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&, t] {
            for (;;) {
                int64_t b = next.fetch_add(chunk);
                if (b >= count) break;
                int64_t e = std::min(count, b + chunk);
                for (int64_t i = b; i < e; ++i) fn(i, t);
            }
        });
    }
    for (auto& th : pool) th.join();
}

// Vose alias table.
struct Alias {
    std::vector<double> prob;
    std::vector<int32_t> alias;
    void build(const std::vector<double>& w) {
        const size_t n = w.size();
        prob.assign(n, 0.0);
        alias.assign(n, 0);
        double sum = std::accumulate(w.begin(), w.end(), 0.0);
        std::vector<double> p(n);
        std::vector<int32_t> small, large;
        for (size_t i = 0; i < n; ++i) {
            p[i] = w[i] * double(n) / sum;
            (p[i] < 1.0 ? small : large).push_back(int32_t(i));
        }
        while (!small.empty() && !large.empty()) {
            int32_t s = small.back(), l = large.back();
            small.pop_back();
            prob[s] = p[s];
            alias[s] = l;
            p[l] = (p[l] + p[s]) - 1.0;
            if (p[l] < 1.0) {
                large.pop_back();
                small.push_back(l);
            }
        }
        for (int32_t l : large) prob[l] = 1.0;
        for (int32_t s : small) prob[s] = 1.0;
    }
    int32_t sample(Rng& r) const {
        uint64_t i = r.below(prob.size());
        return r.u01() < prob[i] ? int32_t(i) : alias[i];
    }
};

void random_permutation(std::vector<int32_t>& perm, Rng& r) {
    std::iota(perm.begin(), perm.end(), 0);
    for (size_t i = perm.size(); i > 1; --i) std::swap(perm[i - 1], perm[r.below(i)]);
}

// Row degrees for the Zipf model: clip(round(s * w_i), 1, n) with s found by
// bisection so the sum hits total, then a +-1 fix-up on random unclipped rows.
bool zipf_row_degrees(const SynthSpec& sp, int64_t total, std::vector<int64_t>& deg,
                      std::string& err) {
    const int32_t m = sp.m, n = sp.n;
    if (total > int64_t(m) * n) {
        err = "more cells requested than the grid holds";
        return false;
    }
    if (total < m) {
        err = "zipf model needs at least one cell per row";
        return false;
    }
    Rng r(stream_key(sp.seed, 1, 0));
    std::vector<int32_t> perm(m);
    random_permutation(perm, r);
    std::vector<double> w(m);
    for (int32_t i = 0; i < m; ++i) w[i] = std::pow(double(perm[i]) + 1.0, -sp.row_exp);
    auto degree_sum = [&](double s) {
        int64_t acc = 0;
        for (int32_t i = 0; i < m; ++i) {
            double d = std::nearbyint(s * w[i]);
            acc += int64_t(std::min<double>(std::max<double>(d, 1.0), double(n)));
        }
        return acc;
    };
    double lo = 0.0, hi = 1.0;
    while (degree_sum(hi) < total) hi *= 2.0;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        (degree_sum(mid) < total ? lo : hi) = mid;
    }
    deg.resize(m);
    int64_t acc = 0;
    for (int32_t i = 0; i < m; ++i) {
        double d = std::nearbyint(hi * w[i]);
        deg[i] = int64_t(std::min<double>(std::max<double>(d, 1.0), double(n)));
        acc += deg[i];
    }
    int64_t diff = total - acc;  // small after bisection
    int64_t guard = 0;
    while (diff != 0) {
        int32_t i = int32_t(r.below(m));
        if (diff > 0 && deg[i] < n) {
            deg[i]++;
            diff--;
        } else if (diff < 0 && deg[i] > 1) {
            deg[i]--;
            diff++;
        }
        if (++guard > int64_t(1) << 40) {
            err = "cannot fit row degrees";
            return false;
        }
    }
    return true;
}

}  // namespace

bool synth_generate(const SynthSpec& sp, Csr& train, Csr& test, std::string& err) {
    const int32_t m = sp.m, n = sp.n;
    const int64_t total = sp.nnz_train + sp.nnz_test;
    if (m <= 0 || n <= 0 || sp.rank <= 0 || sp.nnz_train < 0 || sp.nnz_test < 0) {
        err = "invalid synthetic spec";
        return false;
    }
    if (total > int64_t(m) * int64_t(n)) {
        err = "more cells requested than the grid holds";
        return false;
    }
    int threads = sp.threads > 0 ? sp.threads : int(std::thread::hardware_concurrency());
    if (threads <= 0) threads = 1;

    // ---- cell positions, row-major, columns ascending within a row ----
    std::vector<int64_t> rowptr(size_t(m) + 1, 0);
    std::vector<int32_t> cols;
    if (sp.degree == kDegreeUniform) {
        // Distinct uniform cells: draw ids, sort, unique, top up, then drop a
        // random surplus.
        Rng r(stream_key(sp.seed, 2, 0));
        const uint64_t grid = uint64_t(m) * uint64_t(n);
        std::vector<uint64_t> ids;
        ids.reserve(size_t(total + total / 16 + 16));
        while (int64_t(ids.size()) < total) {
            int64_t need = total - int64_t(ids.size());
            for (int64_t t = 0; t < need + need / 16 + 16; ++t) ids.push_back(r.below(grid));
            std::sort(ids.begin(), ids.end());
            ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        }
        // drop surplus uniformly
        for (int64_t surplus = int64_t(ids.size()) - total; surplus > 0; --surplus) {
            size_t k = size_t(r.below(ids.size()));
            ids[k] = UINT64_MAX;
            std::swap(ids[k], ids.back());
            ids.pop_back();
        }
        std::sort(ids.begin(), ids.end());
        cols.resize(total);
        for (int64_t e = 0; e < total; ++e) {
            int32_t i = int32_t(ids[e] / uint64_t(n));
            cols[e] = int32_t(ids[e] % uint64_t(n));
            rowptr[i + 1]++;
        }
        for (int32_t i = 0; i < m; ++i) rowptr[i + 1] += rowptr[i];
    } else if (sp.degree == kDegreeZipf) {
        std::vector<int64_t> deg;
        if (!zipf_row_degrees(sp, total, deg, err)) return false;
        for (int32_t i = 0; i < m; ++i) rowptr[i + 1] = rowptr[i] + deg[i];
        cols.resize(total);
        Rng rc(stream_key(sp.seed, 3, 0));
        std::vector<int32_t> cperm(n);
        random_permutation(cperm, rc);
        std::vector<double> cw(n);
        for (int32_t j = 0; j < n; ++j) cw[j] = std::pow(double(cperm[j]) + 1.0, -sp.col_exp);
        Alias alias;
        alias.build(cw);
        std::vector<std::vector<uint8_t>> mark(threads, std::vector<uint8_t>(n, 0));
        std::vector<std::vector<std::pair<double, int32_t>>> keys(threads);
        parallel_for(m, threads, [&](int64_t i, int tid) {
            Rng r(stream_key(sp.seed, 4, uint64_t(i)));
            const int64_t d = deg[i];
            int32_t* out = cols.data() + rowptr[i];
            if (d == n) {
                for (int32_t j = 0; j < n; ++j) out[j] = j;
                return;
            }
            if (d * 4 > n) {
                // Weighted sampling without replacement by exponential races
                // (Efraimidis-Spirakis): keep the d smallest -log(u)/w_j.
                auto& kv = keys[tid];
                kv.resize(n);
                for (int32_t j = 0; j < n; ++j) kv[j] = {-std::log(r.u01_open()) / cw[j], j};
                std::nth_element(kv.begin(), kv.begin() + d, kv.end());
                for (int64_t t = 0; t < d; ++t) out[t] = kv[t].second;
            } else {
                auto& mk = mark[tid];
                int64_t got = 0;
                while (got < d) {
                    int32_t j = alias.sample(r);
                    if (!mk[j]) {
                        mk[j] = 1;
                        out[got++] = j;
                    }
                }
                for (int64_t t = 0; t < d; ++t) mk[out[t]] = 0;
            }
            std::sort(out, out + d);
        });
    } else {
        err = "unknown degree model";
        return false;
    }

    // ---- planted truth and ratings ----
    const int32_t rk = sp.rank;
    const double fstd = std::sqrt(sp.factor_var);
    std::vector<float> U(size_t(m) * rk), V(size_t(n) * rk);
    parallel_for(m, threads, [&](int64_t i, int) {
        Rng r(stream_key(sp.seed, 5, uint64_t(i)));
        for (int32_t t = 0; t < rk; ++t) U[size_t(i) * rk + t] = float(fstd * r.normal());
    });
    parallel_for(n, threads, [&](int64_t j, int) {
        Rng r(stream_key(sp.seed, 6, uint64_t(j)));
        for (int32_t t = 0; t < rk; ++t) V[size_t(j) * rk + t] = float(fstd * r.normal());
    });
    std::vector<double> x(total);
    parallel_for(m, threads, [&](int64_t i, int) {
        Rng r(stream_key(sp.seed, 7, uint64_t(i)));
        const float* u = &U[size_t(i) * rk];
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            const float* v = &V[size_t(cols[e]) * rk];
            double acc = 0.0;
            for (int32_t t = 0; t < rk; ++t) acc += double(u[t]) * double(v[t]);
            x[e] = acc + sp.noise_std * r.normal();
        }
    });
    if (sp.rating == kRatingRounded) {
        for (int64_t e = 0; e < total; ++e) {
            double v = std::nearbyint(sp.center + sp.scale * x[e]);
            x[e] = std::min(std::max(v, sp.lo), sp.hi);
        }
    } else if (sp.rating == kRatingAffine) {
        double mn = total ? *std::min_element(x.begin(), x.end()) : 0.0;
        double mx = total ? *std::max_element(x.begin(), x.end()) : 1.0;
        double span = mx > mn ? mx - mn : 1.0;
        for (int64_t e = 0; e < total; ++e) x[e] = sp.lo + (sp.hi - sp.lo) * (x[e] - mn) / span;
    } else {
        err = "unknown rating model";
        return false;
    }

    // ---- exact-size uniform test split ----
    std::vector<uint8_t> is_test(total, 0);
    {
        Rng r(stream_key(sp.seed, 8, 0));
        if (sp.nnz_test * 2 <= total) {
            for (int64_t got = 0; got < sp.nnz_test;) {
                uint64_t e = r.below(uint64_t(total));
                if (!is_test[e]) {
                    is_test[e] = 1;
                    ++got;
                }
            }
        } else {
            std::fill(is_test.begin(), is_test.end(), 1);
            for (int64_t got = 0; got < sp.nnz_train;) {
                uint64_t e = r.below(uint64_t(total));
                if (is_test[e]) {
                    is_test[e] = 0;
                    ++got;
                }
            }
        }
    }
    auto split = [&](Csr& out, uint8_t want) {
        out.m = m;
        out.n = n;
        out.rowptr.assign(size_t(m) + 1, 0);
        for (int32_t i = 0; i < m; ++i) {
            int64_t c = 0;
            for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) c += (is_test[e] == want);
            out.rowptr[i + 1] = out.rowptr[i] + c;
        }
        out.colidx.resize(out.rowptr[m]);
        out.val.resize(out.rowptr[m]);
        int64_t d = 0;
        for (int64_t e = 0; e < total; ++e) {
            if (is_test[e] == want) {
                out.colidx[d] = cols[e];
                out.val[d] = x[e];
                ++d;
            }
        }
    };
    split(train, 0);
    split(test, 1);
    return true;
}

}  // namespace nomad
