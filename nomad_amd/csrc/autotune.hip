// Autotuner of the launch configuration (nomad_autotune in nomad_hip.h).
//
// The reference fixes its worker count by the --nthreads option
// (Code/nomad/nomad_option.cpp:16-77) and has no tuner.  Here the worker
// count (async mode: row blocks, hence workgroup shape and ring capacity) or
// the strata count (strat mode) is picked by timing short trials of each
// candidate on the caller's own problem.  Only the public C ABI is used, so a
// trial is exactly what a caller would run with that configuration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "nomad_hip.h"

namespace nomad_host {
int fail(int code, const std::string& msg);
}

namespace {

struct Candidate {
    int32_t nworkers;
    int32_t flags;
};

// Times one candidate: returns NOMAD_OK and the rate, or the error code.
int run_trial(const nomad_params& base, const Candidate& cand, const int64_t* colptr, const int32_t* rowidx,
              const double* val, int64_t nnz, const std::vector<double>& W0, const std::vector<double>& H0,
              double trial_seconds, nomad_tune_trial* out) {
    nomad_params p = base;
    p.nworkers = cand.nworkers;
    p.flags = (base.flags & ~NOMAD_FLAG_BALANCE) | cand.flags;
    out->nworkers = cand.nworkers;
    out->flags = cand.flags;
    out->updates_per_s = 0.0;
    nomad_ctx* c = nullptr;
    int rc = nomad_create(&p, &c);
    if (rc == NOMAD_OK) rc = nomad_upload_train(c, colptr, rowidx, val, nnz);
    if (rc == NOMAD_OK) rc = nomad_set_factors(c, W0.data(), H0.data());
    if (rc == NOMAD_OK) rc = nomad_get_nworkers(c, &out->nworkers);
    nomad_train_stats st{};
    if (rc == NOMAD_OK) {
        if (p.mode == NOMAD_MODE_ASYNC) {
            rc = nomad_train(c, trial_seconds, 0, &st);  // warm-up: builds the routes
            if (rc == NOMAD_OK) rc = nomad_train(c, trial_seconds, 0, &st);
        } else {
            rc = nomad_train_batch(&c, 1, 1, 0, &st);
            if (rc == NOMAD_OK) rc = nomad_train_batch(&c, 1, 1, 0, &st);
        }
    }
    if (rc == NOMAD_OK && st.train_seconds > 0) out->updates_per_s = double(st.updates) / st.train_seconds;
    out->status = rc;
    nomad_destroy(c);
    return rc;
}

}  // namespace

extern "C" int nomad_autotune(const nomad_params* params, const int64_t* colptr, const int32_t* rowidx,
                              const double* val, int64_t nnz, double trial_seconds, nomad_tune_trial* best,
                              nomad_tune_trial* trials, int32_t max_trials, int32_t* ntrials) {
    using nomad_host::fail;
    if (!params || !colptr || !best || (max_trials > 0 && !trials) || max_trials < 0)
        return fail(NOMAD_EINVAL, "nomad_autotune: null or negative argument");
    if (params->mode == NOMAD_MODE_ASYNC && !(trial_seconds > 0))
        return fail(NOMAD_EINVAL, "nomad_autotune: async trials need trial_seconds > 0");
    if (params->m <= 0 || params->n <= 0 || params->k <= 0)
        return fail(NOMAD_EINVAL, "nomad_autotune: m, n, k must be positive");
    if (ntrials) *ntrials = 0;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, params->device) != hipSuccess || cus <= 0)
        return fail(NOMAD_EHIP, "nomad_autotune: device query failed");

    std::vector<Candidate> cands;
    if (params->mode == NOMAD_MODE_ASYNC) {
        for (int per_cu : {2, 4, 6, 8, 12, 16})
            for (int32_t fl : {NOMAD_FLAG_BALANCE, 0}) cands.push_back({int32_t(per_cu * cus), fl});
    } else {
        int32_t last = -1;
        for (int num : {1, 2, 4, 8, 16}) {  // B = num / 4 x compute units
            const int32_t B = std::max(1, std::min(params->m, num * cus / 4));
            if (B != last) {
                cands.push_back({B, NOMAD_FLAG_BALANCE});
                cands.push_back({B, 0});
            }
            last = B;
        }
    }

    // one start point for every trial: U(0, 1/sqrt(k)) from the seed
    std::mt19937_64 rng(params->seed);
    std::uniform_real_distribution<double> u(0.0, 1.0 / std::sqrt(double(params->k)));
    std::vector<double> W0(size_t(params->m) * params->k), H0(size_t(params->n) * params->k);
    for (double& x : W0) x = u(rng);
    for (double& x : H0) x = u(rng);

    int32_t run = 0;
    bool any = false;
    int first_err = NOMAD_OK;
    std::string first_msg;
    for (const Candidate& cd : cands) {
        nomad_tune_trial t{};
        const int rc = run_trial(*params, cd, colptr, rowidx, val, nnz, W0, H0, trial_seconds, &t);
        if (run < max_trials) trials[run] = t;
        ++run;
        if (rc == NOMAD_OK) {
            if (!any || t.updates_per_s > best->updates_per_s) *best = t;
            any = true;
        } else if (first_err == NOMAD_OK) {
            first_err = rc;
            first_msg = nomad_last_error();
        }
    }
    if (ntrials) *ntrials = run;
    if (!any) return fail(first_err, "nomad_autotune: every trial failed; first error: " + first_msg);
    return NOMAD_OK;
}

extern "C" int nomad_autotune_batch(const nomad_batch_problem* probs, int32_t count, int32_t* strata_out,
                                    int32_t* flags_out, double* updates_per_s) {
    using nomad_host::fail;
    if (!probs || count <= 0 || count > 256 || !strata_out)
        return fail(NOMAD_EINVAL, "nomad_autotune_batch: need 1..256 problems and an output array");
    for (int32_t i = 0; i < count; ++i) {
        const nomad_params& p = probs[i].params;
        if (p.mode != NOMAD_MODE_STRAT || p.m <= 0 || p.n <= 0 || p.k <= 0 || !probs[i].colptr)
            return fail(NOMAD_EINVAL, "nomad_autotune_batch: problems must be valid strat-mode problems");
    }
    // start points, drawn once per problem
    std::vector<std::vector<double>> W0(count), H0(count);
    for (int32_t i = 0; i < count; ++i) {
        const nomad_params& p = probs[i].params;
        std::mt19937_64 rng(p.seed);
        std::uniform_real_distribution<double> u(0.0, 1.0 / std::sqrt(double(p.k)));
        W0[i].resize(size_t(p.m) * p.k);
        H0[i].resize(size_t(p.n) * p.k);
        for (double& x : W0[i]) x = u(rng);
        for (double& x : H0[i]) x = u(rng);
    }
    double best_rate = -1.0;
    int32_t best_B = 0;
    std::string last_msg;
    int last_rc = NOMAD_OK;
    for (int32_t B : {64, 128, 256, 512, 1024}) {
        std::vector<nomad_ctx*> cs(static_cast<size_t>(count), nullptr);
        int rc = NOMAD_OK;
        for (int32_t i = 0; i < count && rc == NOMAD_OK; ++i) {
            nomad_params p = probs[i].params;
            p.nworkers = std::min(B, p.m);
            rc = nomad_create(&p, &cs[i]);
            if (rc == NOMAD_OK) rc = nomad_upload_train(cs[i], probs[i].colptr, probs[i].rowidx, probs[i].val,
                                                        probs[i].nnz);
            if (rc == NOMAD_OK) rc = nomad_set_factors(cs[i], W0[i].data(), H0[i].data());
        }
        std::vector<nomad_train_stats> st(static_cast<size_t>(count));
        if (rc == NOMAD_OK) rc = nomad_train_batch(cs.data(), count, 1, 0, st.data());
        if (rc == NOMAD_OK) rc = nomad_train_batch(cs.data(), count, 1, 0, st.data());
        if (rc == NOMAD_OK) {
            double upd = 0.0;
            for (const auto& x : st) upd += double(x.updates);
            const double rate = st[0].train_seconds > 0 ? upd / st[0].train_seconds : 0.0;
            if (rate > best_rate) {
                best_rate = rate;
                best_B = B;
            }
        } else {
            last_rc = rc;
            last_msg = nomad_last_error();
        }
        for (nomad_ctx* c : cs) nomad_destroy(c);
    }
    if (best_B == 0) return fail(last_rc, "nomad_autotune_batch: every trial failed; last error: " + last_msg);
    std::vector<int32_t> Bsel(static_cast<size_t>(count));
    for (int32_t i = 0; i < count; ++i) Bsel[size_t(i)] = std::min(best_B, probs[i].params.m);

    std::vector<int32_t> Fsel(static_cast<size_t>(count));
    for (int32_t i = 0; i < count; ++i) Fsel[size_t(i)] = probs[i].params.flags;

    // Second stage (greedy): a batched launch lasts at least as long as its
    // slowest problem's chain of sub-epochs.  A sparse problem's chain is
    // shorter with more strata, and a skewed one's with row blocks balanced
    // by ratings (or, where balancing leaves blocks of many light rows that
    // no longer fit the LDS staging, without); but every extra tile costs a
    // ticket and a publication, so problems are changed one at a time.
    // Each problem's own epoch (a batched launch of one) is timed at the
    // common count; the problems within 40 % of the slowest, slowest first,
    // then try 1x, 2x and 4x the common count with the balance flag as given
    // and toggled (when flags_out is given; else as given only), and keep
    // the move that makes the batched epoch of all problems shortest, if it
    // is at least 1 % shorter than the best so far (each timing: the shorter
    // of two epochs after a warm-up epoch).
    auto make = [&](int32_t i, int32_t B, int32_t flags, nomad_ctx** out) -> int {
        nomad_params p = probs[i].params;
        p.nworkers = B;
        p.flags = flags;
        *out = nullptr;
        int rc = nomad_create(&p, out);
        if (rc == NOMAD_OK) rc = nomad_upload_train(*out, probs[i].colptr, probs[i].rowidx, probs[i].val, probs[i].nnz);
        if (rc == NOMAD_OK) rc = nomad_set_factors(*out, W0[size_t(i)].data(), H0[size_t(i)].data());
        return rc;
    };
    auto timed = [](nomad_ctx** cs, int32_t n, double* sec) -> int {
        std::vector<nomad_train_stats> st(static_cast<size_t>(n));
        int rc = nomad_train_batch(cs, n, 1, 0, st.data());
        double best = 0.0;
        for (int rep = 0; rep < 2 && rc == NOMAD_OK; ++rep) {
            rc = nomad_train_batch(cs, n, 1, 0, st.data());
            if (rc == NOMAD_OK && (rep == 0 || st[0].train_seconds < best)) best = st[0].train_seconds;
        }
        *sec = best;
        return rc;
    };
    if (count > 1) {
        std::vector<nomad_ctx*> cur(static_cast<size_t>(count), nullptr);
        std::vector<double> own(static_cast<size_t>(count), 0.0);
        int rc = NOMAD_OK;
        for (int32_t i = 0; i < count && rc == NOMAD_OK; ++i) {
            rc = make(i, Bsel[size_t(i)], Fsel[size_t(i)], &cur[size_t(i)]);
            if (rc == NOMAD_OK) rc = timed(&cur[size_t(i)], 1, &own[size_t(i)]);
        }
        double T = 0.0;
        if (rc == NOMAD_OK) rc = timed(cur.data(), count, &T);
        if (rc == NOMAD_OK && T > 0.0) {
            const double own_max = *std::max_element(own.begin(), own.end());
            std::vector<int32_t> cand;
            for (int32_t i = 0; i < count; ++i)
                if (own[size_t(i)] >= 0.6 * own_max) cand.push_back(i);
            std::stable_sort(cand.begin(), cand.end(),
                             [&](int32_t x, int32_t y) { return own[size_t(x)] > own[size_t(y)]; });
            for (int32_t i : cand) {
                const int32_t f0 = Fsel[size_t(i)];
                nomad_ctx* best_ctx = nullptr;
                double best_t = 0.99 * T;
                int32_t best_b = 0, best_f = 0;
                for (int32_t mult : {1, 2, 4}) {
                    const int32_t B = std::min(best_B * mult, probs[i].params.m);
                    if (mult > 1 && B <= std::min(best_B * mult / 2, probs[i].params.m)) break;  // capped at m
                    for (int32_t f : {f0, f0 ^ NOMAD_FLAG_BALANCE}) {
                        if (f != f0 && !flags_out) continue;
                        if (B == Bsel[size_t(i)] && f == f0) continue;  // the current configuration
                        nomad_ctx* nc = nullptr;
                        if (make(i, B, f, &nc) != NOMAD_OK) {
                            nomad_destroy(nc);
                            continue;
                        }
                        nomad_ctx* old = cur[size_t(i)];
                        cur[size_t(i)] = nc;
                        double t = 0.0;
                        const bool ok = timed(cur.data(), count, &t) == NOMAD_OK && t > 0.0;
                        cur[size_t(i)] = old;
                        if (ok && t < best_t) {
                            nomad_destroy(best_ctx);
                            best_ctx = nc;
                            best_t = t;
                            best_b = B;
                            best_f = f;
                        } else {
                            nomad_destroy(nc);
                        }
                    }
                }
                if (best_ctx) {
                    nomad_destroy(cur[size_t(i)]);
                    cur[size_t(i)] = best_ctx;
                    T = best_t;
                    Bsel[size_t(i)] = best_b;
                    Fsel[size_t(i)] = best_f;
                }
            }
            double upd = 0.0;
            for (int32_t i = 0; i < count; ++i) upd += double(probs[i].nnz);
            best_rate = upd / T;
        }
        for (nomad_ctx* c : cur) nomad_destroy(c);
    }
    for (int32_t i = 0; i < count; ++i) {
        strata_out[i] = Bsel[size_t(i)];
        if (flags_out) flags_out[i] = Fsel[size_t(i)];
    }
    if (updates_per_s) *updates_per_s = best_rate;
    return NOMAD_OK;
}
