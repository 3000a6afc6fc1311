// NOMAD SGD kernels for gfx950.
//
// A worker is one wavefront that owns a contiguous row block of W (the
// reference's per-thread part, Code/nomad/nb_data.cpp:83-89).  A visit of
// column j by worker p is the reference's per-token loop
// (Code/nomad/nb_main.cpp:513-544): fix the step size from the visit count,
// then walk the worker's ratings of column j in ascending row order, each one
// an SGD step on (w_i, h_j) with h_j carried in registers from rating to
// rating.
//
//   sgd_strat_kernel : one launch per sub-epoch of a B x B stratification;
//                      wave b owns tile (row block b, column block (b+s) % B).
//   sgd_async_kernel : persistent; column tokens circulate through per-worker
//                      multi-producer / single-consumer rings in HBM
//                      (nb_main.cpp:491-565 with TBB queues replaced).
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernel_args.h"

using namespace nomad_dev;
using nomad_k::SgdArgs;
using nomad_k::VisitDesc;

namespace {

constexpr int kPrefetch = 8;  // rows of W in flight per worker inside a visit

template <class A>
__device__ __forceinline__ A step_for(const SgdArgs& a, int32_t count) {
    if (count < a.nsteps) return static_cast<const A*>(a.steps)[count];
    // past the table: same formula as the host table (nb_main.cpp:514-515)
    const double s = a.lrate * 1.5 / (1.0 + a.drate * pow(double(count) + 1.0, 1.5));
    return A(s);
}

// Walks ratings [start, start + len) of column j for one worker.  h is the
// column's factor pair held by this lane; it is updated in place.
template <int S>
__device__ __forceinline__ void run_visit(const SgdArgs& a, typename Store<S>::pair& h, int32_t j,
                                          int64_t start, int32_t len, typename Store<S>::acc step,
                                          typename Store<S>::acc reg, uint32_t count) {
    using T = Store<S>;
    using E = typename T::elem;
    using Aa = typename T::acc;
    using R = typename T::rval;
    const int lane = lane_id();
    E* __restrict__ W = static_cast<E*>(a.W);
    const int32_t* __restrict__ ridx = a.rowidx + start;
    const R* __restrict__ rv = static_cast<const R*>(a.val) + start;
    const int npairs = a.ldf >> 1;
    const int64_t ldf = a.ldf;
    const uint32_t sr_b = uint32_t(j) * 0x01000193u + count;

    for (int32_t base = 0; base < len; base += kWave) {
        const int32_t cnt = min(kWave, len - base);
        const int32_t myrow = ridx[base + min(lane, cnt - 1)];
        const R myval = rv[base + min(lane, cnt - 1)];
        typename T::pair w[kPrefetch];
#pragma unroll
        for (int d = 0; d < kPrefetch; ++d) {
            w[d] = T::load(W + int64_t(readlane_i32(myrow, min(d, cnt - 1))) * ldf, lane, npairs);
        }
        for (int32_t t0 = 0; t0 < cnt; t0 += kPrefetch) {
#pragma unroll
            for (int d = 0; d < kPrefetch; ++d) {
                const int32_t t = t0 + d;
                if (t < cnt) {
                    const int32_t i = readlane_i32(myrow, t);
                    const Aa r = Aa(readlane_val(myval, t));
                    sgd_step<typename T::pair, Aa>(w[d], h, r, step, reg);
                    const uint32_t rnd = sr_bits(a.sr_seed, uint32_t(i), sr_b, uint32_t(lane));
                    T::store(W + int64_t(i) * ldf, lane, npairs, w[d], rnd, a.round_mode);
                }
                // refill slot d with the row kPrefetch ratings ahead (clamped:
                // past the end it re-reads a row of this chunk and is unused)
                const int32_t tn = min(t + kPrefetch, cnt - 1);
                w[d] = T::load(W + int64_t(readlane_i32(myrow, tn)) * ldf, lane, npairs);
            }
        }
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

__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(int(uint32_t(v)));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(int(uint32_t(v >> 32)));
    return (uint64_t(hi) << 32) | lo;
}

// Multi-producer push of column j onto worker q's ring.  The agent-scope
// release fence makes this wave's earlier writes of h_j and of the column's
// routing state visible before the slot is published.
__device__ __forceinline__ void ring_push(const SgdArgs& a, int32_t q, int32_t j) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane_id() == 0) {
        const uint32_t t =
            __hip_atomic_fetch_add(a.qtail + q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t v = (uint64_t((t >> a.qlog2) + 1u) << 32) | uint32_t(j);
        uint64_t* slot = a.qslot + (int64_t(q) << a.qlog2) + (t & ((1u << a.qlog2) - 1u));
        __hip_atomic_store(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Stratified schedule: tile (b, (b + substep) % B).  Columns of the tile in
// ascending order, rows of each column in ascending order; the step size uses
// the epoch as visit count (every (column, row block) pair is visited once per
// epoch, as with the reference's per-thread col_update_counts at one rank).
template <int S>
__global__ __launch_bounds__(64) void sgd_strat_kernel(SgdArgs a) {
    using T = Store<S>;
    using E = typename T::elem;
    using Aa = typename T::acc;
    const int b = blockIdx.x;
    if (b >= a.nblocks) return;
    const int lane = lane_id();
    const int cb = (b + a.substep) % a.nblocks;
    const int32_t rlo = b * a.rpb;
    const int32_t rhi = min(a.m, rlo + a.rpb);
    const int32_t clo = cb * a.cpb;
    const int32_t chi = min(a.n, clo + a.cpb);
    if (rlo >= rhi || clo >= chi) return;
    const Aa step = step_for<Aa>(a, a.epoch);
    const Aa reg = Aa(a.reg);
    E* __restrict__ H = static_cast<E*>(a.H);
    for (int32_t j = clo; j < chi; ++j) {
        const int64_t c0 = a.colptr[j], c1 = a.colptr[j + 1];
        const int64_t s = row_lower_bound(a.rowidx, c0, c1, rlo);
        const int64_t e = row_lower_bound(a.rowidx, s, c1, rhi);
        if (e <= s) continue;
        E* hrow = H + int64_t(j) * a.ldf;
        typename T::pair h = T::load(hrow, lane, a.ldf >> 1);
        run_visit<S>(a, h, j, s, int32_t(e - s), step, reg, uint32_t(a.epoch));
        const uint32_t rnd = sr_bits(a.sr_seed, 0x80000000u | uint32_t(b),
                                     uint32_t(j) * 0x01000193u + uint32_t(a.epoch), uint32_t(lane));
        T::store(hrow, lane, a.ldf >> 1, h, rnd, a.round_mode);
    }
}

// ---------------------------------------------------------------------------
// Asynchronous NOMAD.  Worker p = block p (one wave).  It pops column ids from
// its ring in FIFO order; for each it runs the visit described by the column's
// current route entry, bumps that entry's visit count, advances the column's
// route position and pushes the column to the next worker of its route.  A
// column that completes route cycle target_cycle is parked instead.  Every
// wave leaves the loop when all active columns are parked or when the
// wall-clock deadline passes.
template <int S>
__global__ __launch_bounds__(64) void sgd_async_kernel(SgdArgs a) {
    using T = Store<S>;
    using E = typename T::elem;
    using Aa = typename T::acc;
    const int p = blockIdx.x;
    if (p >= a.nblocks) return;
    const int lane = lane_id();
    const uint32_t mask = (1u << a.qlog2) - 1u;
    const uint64_t* ring = a.qslot + (int64_t(p) << a.qlog2);
    E* __restrict__ H = static_cast<E*>(a.H);
    const Aa reg = Aa(a.reg);
    uint32_t head = a.qhead[p];
    const uint64_t t_end = wall_clock64() + a.deadline;
    int64_t n_upd = 0, n_vis = 0, n_fail = 0;
    bool hit_deadline = false;

    for (;;) {
        const uint64_t sv = uniform64(ld_relaxed(ring + (head & mask)));
        if (uint32_t(sv >> 32) == (head >> a.qlog2) + 1u) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            ++head;
            const int32_t j = int32_t(uint32_t(sv));
            const int32_t pos = a.tok_pos[j];
            const int64_t vb = a.vptr[j];
            const int32_t nv = int32_t(a.vptr[j + 1] - vb);
            const int64_t vi = vb + pos;
            const VisitDesc vd = a.visits[vi];
            const int32_t count = a.vcnt[vi];
            const Aa step = step_for<Aa>(a, count);
            E* hrow = H + int64_t(j) * a.ldf;
            typename T::pair h = T::load(hrow, lane, a.ldf >> 1);
            run_visit<S>(a, h, j, vd.start, vd.len, step, reg, uint32_t(count));
            const uint32_t rnd = sr_bits(a.sr_seed, 0x80000000u | uint32_t(p),
                                         uint32_t(j) * 0x01000193u + uint32_t(count), uint32_t(lane));
            T::store(hrow, lane, a.ldf >> 1, h, rnd, a.round_mode);
            n_upd += vd.len;
            ++n_vis;
            if (lane == 0) a.vcnt[vi] = count + 1;
            int32_t npos = pos + 1;
            bool park = false;
            if (npos == nv) {
                npos = 0;
                const int32_t cyc = a.tok_cycle[j] + 1;
                if (lane == 0) a.tok_cycle[j] = cyc;
                park = cyc >= a.target_cycle;
            }
            if (lane == 0) a.tok_pos[j] = npos;
            if (park) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                if (lane == 0)
                    __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                ring_push(a, a.visits[vb + npos].worker, j);
            }
            if (wall_clock64() > t_end) {
                hit_deadline = true;
                break;
            }
            continue;
        }
        ++n_fail;
        if (int32_t(ld_relaxed(a.done)) >= a.ncols_active) break;
        if (wall_clock64() > t_end) {
            hit_deadline = true;
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    if (lane == 0) {
        a.qhead[p] = head;
        a.stats[3 * p + 0] += n_upd;
        a.stats[3 * p + 1] += n_vis;
        a.stats[3 * p + 2] += n_fail;
        if (hit_deadline) __hip_atomic_store(a.timed_out, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Appends per-worker token lists to the rings before a launch: worker p gets
// list[off[p] .. off[p+1]) in order.  One thread per worker; plain stores are
// published by the kernel boundary.
__global__ void seed_rings_kernel(SgdArgs a, const int32_t* list, const int64_t* off) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nblocks) return;
    const uint32_t mask = (1u << a.qlog2) - 1u;
    uint32_t t = a.qtail[p];
    for (int64_t e = off[p]; e < off[p + 1]; ++e, ++t) {
        a.qslot[(int64_t(p) << a.qlog2) + (t & mask)] =
            (uint64_t((t >> a.qlog2) + 1u) << 32) | uint32_t(list[e]);
    }
    a.qtail[p] = t;
}

// Explicit instantiations launched from capi.hip.
template __global__ void sgd_strat_kernel<NOMAD_STORE_FP32>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_BF16>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_FP64>(SgdArgs);
template __global__ void sgd_async_kernel<NOMAD_STORE_FP32>(SgdArgs);
template __global__ void sgd_async_kernel<NOMAD_STORE_BF16>(SgdArgs);
template __global__ void sgd_async_kernel<NOMAD_STORE_FP64>(SgdArgs);
