// Device helpers shared by the SGD kernels: step size, the per-rating chain
// over a chunk of ratings, and ring helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "device_common.hpp"
#include "kernel_args.h"

namespace nomad_dev {

using nomad_k::SgdArgs;
using nomad_k::VisitDesc;

constexpr int kPrefetch = 8;  // rows of W in flight per worker inside a visit

template <class A>
__device__ __forceinline__ A step_for(const SgdArgs& a, int32_t count) {
    if (count < a.nsteps) return static_cast<const A*>(a.steps)[count];
    // past the table: same formula as the host table (nb_main.cpp:514-515)
    const double s = a.lrate * 1.5 / (1.0 + a.drate * pow(double(count) + 1.0, 1.5));
    return A(s);
}

// Row access of W / H for the stratified kernels: plain (DC = false) when
// every reader of a row runs after a kernel boundary, device-coherent (the
// sc1 cache policy of relaxed agent-scope atomics, DC = true) when a row
// moves between waves on different XCDs inside one launch.  The values are
// the same.
template <int S, bool DC>
__device__ __forceinline__ typename Store<S>::pair row_load(const typename Store<S>::elem* row, int lane,
                                                            int npairs) {
    return Store<S>::template load<DC ? kCpDevice : 0>(row, lane, npairs);
}
template <int S, bool DC>
__device__ __forceinline__ void row_store(typename Store<S>::elem* row, int lane, int npairs,
                                          typename Store<S>::pair p, uint32_t rnd, int round_mode) {
    Store<S>::template store<DC ? kCpDevice : 0>(row, lane, npairs, p, rnd, round_mode);
}

// SGD steps for the first `cnt` (<= 64) ratings held one per lane in
// (myrow, myval), in lane order, for column j whose factor pair h this lane
// holds.  Up to kPrefetch rows of W are kept in flight.
template <int S, bool DC = false>
__device__ __forceinline__ void run_chunk(const SgdArgs& a, typename Store<S>::pair& h, int32_t myrow,
                                          typename Store<S>::rval myval, int32_t cnt,
                                          typename Store<S>::acc step, typename Store<S>::acc reg,
                                          uint32_t sr_b) {
    using T = Store<S>;
    using E = typename T::elem;
    using Aa = typename T::acc;
    const int lane = lane_id();
    E* __restrict__ W = static_cast<E*>(a.W);
    const int npairs = a.ldf >> 1;
    const int64_t ldf = a.ldf;
    const uint32_t lkey = sr_lane_key(a.sr_seed, uint32_t(lane));
    typename T::pair w[kPrefetch];
#pragma unroll
    for (int d = 0; d < kPrefetch; ++d) {
        w[d] = row_load<S, DC>(W + int64_t(readlane_i32(myrow, min(d, cnt - 1))) * ldf, lane, npairs);
    }
    for (int32_t t0 = 0; t0 < cnt; t0 += kPrefetch) {
#pragma unroll
        for (int d = 0; d < kPrefetch; ++d) {
            const int32_t t = t0 + d;
            if (t < cnt) {
                const int32_t i = readlane_i32(myrow, t);
                const Aa r = Aa(readlane_val(myval, t));
                sgd_step<typename T::pair, Aa>(w[d], h, r, step, reg);
                const uint32_t rnd = sr_bits_k(a.sr_seed, uint32_t(i), sr_b, lkey);
                row_store<S, DC>(W + int64_t(i) * ldf, lane, npairs, w[d], rnd, a.round_mode);
            }
            // refill slot d with the row kPrefetch ratings ahead (clamped: past
            // the end it re-reads a row of this chunk and is unused)
            const int32_t tn = min(t + kPrefetch, cnt - 1);
            w[d] = row_load<S, DC>(W + int64_t(readlane_i32(myrow, tn)) * ldf, lane, npairs);
        }
    }
}

// Walks ratings [start, start + len) of column j for one worker.  h is the
// column's factor pair held by this lane; it is updated in place.
template <int S>
__device__ __forceinline__ void run_visit(const SgdArgs& a, typename Store<S>::pair& h, int32_t j,
                                          int64_t start, int32_t len, typename Store<S>::acc step,
                                          typename Store<S>::acc reg, uint32_t count) {
    using R = typename Store<S>::rval;
    const int lane = lane_id();
    const int32_t* __restrict__ ridx = a.rowidx + start;
    const R* __restrict__ rv = static_cast<const R*>(a.val) + start;
    const uint32_t sr_b = uint32_t(j) * 0x01000193u + count;
    for (int32_t base = 0; base < len; base += kWave) {
        const int32_t cnt = min(kWave, len - base);
        const int32_t myrow = ridx[base + min(lane, cnt - 1)];
        const R myval = rv[base + min(lane, cnt - 1)];
        run_chunk<S>(a, h, myrow, myval, cnt, step, reg, sr_b);
    }
}

// run_visit with the first chunk's rows and values already loaded (one per
// lane, lanes past the visit's length holding the last rating's), so that
// the caller can fetch them ahead, during the previous visit.
template <int S, bool DC = false>
__device__ __forceinline__ void run_visit_pre(const SgdArgs& a, typename Store<S>::pair& h, int32_t j,
                                              int64_t start, int32_t len, int32_t row0,
                                              typename Store<S>::rval val0, typename Store<S>::acc step,
                                              typename Store<S>::acc reg, uint32_t count) {
    using R = typename Store<S>::rval;
    const int lane = lane_id();
    const int32_t* __restrict__ ridx = a.rowidx + start;
    const R* __restrict__ rv = static_cast<const R*>(a.val) + start;
    const uint32_t sr_b = uint32_t(j) * 0x01000193u + count;
    run_chunk<S, DC>(a, h, row0, val0, min(kWave, len), step, reg, sr_b);
    for (int32_t base = kWave; base < len; base += kWave) {
        const int32_t cnt = min(kWave, len - base);
        const int32_t myrow = ridx[base + min(lane, cnt - 1)];
        const R myval = rv[base + min(lane, cnt - 1)];
        run_chunk<S, DC>(a, h, myrow, myval, cnt, step, reg, sr_b);
    }
}

// lower_bound of `key` in the ascending rows rowidx[lo, hi)
__device__ __forceinline__ int64_t row_lower_bound(const int32_t* rowidx, int64_t lo, int64_t hi,
                                                   int32_t key) {
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (rowidx[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(int(uint32_t(v)));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(int(uint32_t(v >> 32)));
    return (uint64_t(hi) << 32) | lo;
}

}  // namespace nomad_dev
