// Deterministic stratified SGD kernel for gfx950.
//
// A worker is one wavefront that owns a contiguous row block of W (the
// reference's per-thread part, Code/nomad/nb_data.cpp:83-89).  A visit of
// column j by a worker is the reference's per-token loop
// (Code/nomad/nb_main.cpp:513-544): fix the step size from the visit count,
// then walk the worker's ratings of column j in ascending row order, each one
// an SGD step on (w_i, h_j) with h_j carried in registers from rating to
// rating.
//
//   sgd_strat_kernel : one launch per sub-epoch of a B x B stratification;
//                      wave b owns tile (row block b, column block (b+s) % B).
//   sgd_batch_kernel : many stratified problems in one persistent launch;
//                      waves take tiles in ticket order and wait for the two
//                      tiles of the previous sub-epoch that the tile reads.
// The asynchronous token kernel is in async_kernel.hip.
#include <hip/hip_runtime.h>

#include <climits>

#include "sgd_common.hpp"

using namespace nomad_dev;
using nomad_k::BatchArgs;
using nomad_k::kBatchMax;
using nomad_k::SgdArgs;
using nomad_k::VisitDesc;


// ---------------------------------------------------------------------------
// Stratified schedule: tile (b, (b + substep) % B).  Columns of the tile in
// ascending order, rows of each column in ascending order; the step size uses
// the epoch as visit count (every (column, row block) pair is visited once per
// epoch, as with the reference's per-thread col_update_counts at one rank).
// The tile's non-empty column visits are listed ahead (tile_ptr / visits), so
// empty (column, row block) pairs cost nothing.  While a visit runs, the next
// two visits' h rows and first 64 rows / values are already being fetched:
// their columns differ from the running one's (a column is visited once per
// tile), so their h is not being written; their W rows may be, and are
// loaded only when the visit starts.
// The single-problem kernel and the batched kernel both run a tile through
// this function, so a problem's arithmetic is the same in either launch.
//
// When the tile's row block of W fits in the wave's LDS buffer (lds, with
// room for lds_pairs stored pairs), the block is copied there at the start,
// the tile's ratings read and write their rows there, and the block is copied
// back at the end: the rows' values, and the order in which the steps see
// them, are those of the HBM path, so the result is the same bit for bit,
// but a step waits for LDS instead of HBM.
template <int S, bool DC, int RM>
__device__ __forceinline__ void strat_tile(const SgdArgs& a, int b, int substep, int epoch,
                                           typename Store<S>::raw* lds, int lds_pairs) {
    using T = Store<S>;
    using E = typename T::elem;
    using Aa = typename T::acc;
    using R = typename T::rval;
    const int lane = lane_id();
    const int B = a.nblocks;
    const int cb = (b + substep) % B;
    const int64_t* tp = a.tile_ptr + int64_t(b) * B + cb;
    const int64_t v0 = tp[0], v1 = tp[1];
    if (v0 >= v1) return;
    const Aa step = step_for<Aa>(a, epoch);
    const Aa reg = Aa(a.reg);
    E* __restrict__ H = static_cast<E*>(a.H);
    const R* __restrict__ rv = static_cast<const R*>(a.val);
    const int npairs = a.ldf >> 1;
    const int32_t rlo = b * a.rpb, rhi = min(a.m, (b + 1) * a.rpb);
    const RowBlock rb = row_block<S>(a, rlo, rhi, lane);
    const uint32_t lkey = sr_lane_key(a.sr_seed, uint32_t(lane));
    const int32_t block_pairs = (rhi - rlo) * npairs;
    const bool in_lds = lds != nullptr && block_pairs <= lds_pairs;
    constexpr int CP = DC ? kCpDevice : 0;
    if (in_lds) {
        // the block's rows are contiguous in W: block_pairs stored pairs,
        // eight per lane in flight
        constexpr int kU = 8;
        for (int32_t q0 = 0; q0 < block_pairs; q0 += kU * kWave) {
            typename T::raw t[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) t[u] = T::template load_raw_b<CP>(rb.rsrc, (q0 + u * kWave + lane) * T::kPairBytes);
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (q0 + u * kWave + lane < block_pairs) lds[q0 + u * kWave + lane] = t[u];
        }
    }
    // Visit descriptors 64 at a time, one per lane (the current batch and
    // the next), so that no descriptor costs a memory round trip inside the
    // loop; a visit's column factor and its first 64 rows and values are
    // requested two visits ahead.  Indexes past the tile are clamped to v1
    // (the next tile's first visit or the list's sentinel, in bounds) and
    // what they load is dropped.
    struct Descs {
        int64_t start;
        int32_t len, col;
    };
    auto load_descs = [&](int64_t vb) {
        const VisitDesc dd = a.visits[min(vb + lane, v1)];
        return Descs{dd.start, dd.len, dd.worker};
    };
    int64_t vb = v0;
    Descs dA = load_descs(vb), dB = load_descs(vb + kWave);
    auto desc = [&](int64_t v, int64_t& st, int32_t& ln, int32_t& col) {
        const int u = int(min(v, v1) - vb);
        const Descs& q = u < kWave ? dA : dB;
        const int l = u & (kWave - 1);
        st = int64_t((uint64_t(uint32_t(readlane_i32(int32_t(uint64_t(q.start) >> 32), l))) << 32) |
                     uint32_t(readlane_i32(int32_t(q.start), l)));
        ln = readlane_i32(q.len, l);
        col = readlane_i32(q.col, l);
    };
    struct Pre {
        typename T::pair h;
        int32_t row;
        R val;
    };
    auto prefetch = [&](int64_t v) {
        int64_t st;
        int32_t ln, col;
        desc(v, st, ln, col);
        Pre p;
        p.h = row_load<S, DC>(H + int64_t(col) * a.ldf, lane, npairs);
        const int64_t fi = st + min(lane, max(ln, 1) - 1);
        p.row = a.rowidx[fi];
        p.val = rv[fi];
        return p;
    };
    Pre c0 = prefetch(v0), c1 = prefetch(v0 + 1);
    for (int64_t v = v0; v < v1; ++v) {
        if (v - vb == kWave) {
            vb += kWave;
            dA = dB;
            dB = load_descs(vb + kWave);
        }
        const Pre c2 = prefetch(v + 2);
        int64_t st;
        int32_t ln, j;
        desc(v, st, ln, j);
        typename T::pair h = c0.h;
        if (in_lds)
            run_visit_lds<S, RM>(a, lds, rlo, npairs, h, j, st, ln, c0.row, c0.val, step, reg, uint32_t(epoch), lkey);
        else
            run_visit_pre<S, DC, RM>(a, rb, h, j, st, ln, c0.row, c0.val, step, reg, uint32_t(epoch), lkey);
        const uint32_t rnd = sr_bits_k(a.sr_seed, 0x80000000u | uint32_t(b),
                                       uint32_t(j) * 0x01000193u + uint32_t(epoch), lkey);
        row_store<S, DC>(H + int64_t(j) * a.ldf, lane, npairs, h, rnd, RM);
        c0 = c1;
        c1 = c2;
    }
    if (in_lds) {
        // the block back to W (the range check drops the lanes past it)
        for (int32_t q = lane; q < block_pairs + kWave - 1; q += kWave) {
            const int32_t qq = min(q, block_pairs - 1);
            T::template store_raw_b<CP>(rb.rsrc, q < block_pairs ? qq * T::kPairBytes : rb.drop(), lds[qq]);
        }
    }
}

// (RM: the rounding of bf16 stores, a template argument so that no store
// branches on it; fp32 and fp64 ignore it)
template <int S, int RM>
__global__ __launch_bounds__(64) void sgd_strat_kernel(SgdArgs a) {
    using raw = typename Store<S>::raw;
    constexpr int kPairs = kTileLdsBytes / int(sizeof(raw));
    __shared__ raw wl[kPairs];
    const int b = blockIdx.x;
    if (b >= a.nblocks) return;
    strat_tile<S, false, RM>(a, b, a.substep, a.epoch, wl, kPairs);
}

// ---------------------------------------------------------------------------
// Batched stratified sweep.  Problem p (its own ratings, factors, rank, step
// table and strata count B_p, args.epoch = its first epoch) runs R_p = E * B_p
// rounds; round r of p is sub-epoch r % B_p of epoch args.epoch + r / B_p,
// i.e. the B_p tiles (r, b).  Tile (r, b) reads W block b and H block
// (b + r) % B_p, which the tiles (r - 1, b) and (r - 1, (b + 1) % B_p) wrote
// last; every tile that touched either block earlier is an ancestor of those
// two.  So tile (r, b) may start once both have finished, and it then sees
// exactly the values the one-launch-per-sub-epoch schedule gives it.
//
// Tickets: round by round (all problems' tiles of round r, then round r + 1),
// problems sorted by R_p descending so that the problems still running in a
// round are a prefix of the sorted list.  A wave takes the next ticket, waits
// for the two parent tiles (their tickets are smaller, so they were taken by
// waves that are running), runs the tile and publishes it.  The smallest
// unfinished ticket can always run, so the launch needs no co-residency and
// finishes for any grid size.
//
// Publication: the next tile on a block may run on another XCD, whose L2 is
// not this one's.  W and H rows are therefore read and written with
// device-coherent (sc1) accesses, and a tile's progress counter is raised
// only after all of the wave's stores have completed (s_waitcnt 0), so a
// wave that sees the counter reads the new rows; no L2 writeback or
// invalidate is needed.  Counters live in uncached memory and are polled
// with device-coherent loads.
template <int S>
__global__ __launch_bounds__(256) void sgd_batch_kernel(BatchArgs ba) {
    __shared__ int64_t seg_t0[kBatchMax];
    __shared__ int32_t seg_r0[kBatchMax];
    __shared__ int32_t seg_w[kBatchMax];
    __shared__ int32_t boff[kBatchMax + 1];
    __shared__ int32_t order[kBatchMax];
    using raw = typename Store<S>::raw;
    constexpr int kPairs = kTileLdsBytes / int(sizeof(raw));
    __shared__ raw wl[4][kPairs];  // per wave (256-thread workgroups)
    const int Q = ba.nsorted, G = ba.nseg;
    for (int i = threadIdx.x; i <= Q; i += blockDim.x) {
        if (i < G) {
            seg_t0[i] = ba.seg_t0[i];
            seg_r0[i] = ba.seg_r0[i];
            seg_w[i] = ba.seg_w[i];
        }
        boff[i] = ba.boff[i];
        if (i < Q) order[i] = ba.order[i];
    }
    __syncthreads();
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const uint64_t t_end = wall_clock64() + ba.deadline;
    for (;;) {
        uint64_t t = 0;
        if (lane == 0) t = __hip_atomic_fetch_add(ba.ticket, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t = uniform64(t);
        if (t >= uint64_t(ba.nitems)) break;
        // segment: rounds of constant width (the same problems still running)
        int sg = 0;
        while (sg + 1 < G && uint64_t(seg_t0[sg + 1]) <= t) ++sg;
        const uint64_t rel = t - uint64_t(seg_t0[sg]);
        const int32_t r = seg_r0[sg] + int32_t(rel / uint32_t(seg_w[sg]));
        const int32_t local = int32_t(rel % uint32_t(seg_w[sg]));
        int q = 0;
        while (q + 1 < Q && boff[q + 1] <= local) ++q;
        const int p = __builtin_amdgcn_readfirstlane(order[q]);
        const int b = __builtin_amdgcn_readfirstlane(local - boff[q]);
        const SgdArgs a = ba.probs[p];  // by value: its fields stay in registers across the row stores
        const int B = a.nblocks;
        uint32_t* done = ba.done + ba.done_off[p];
        if (r > 0) {
            const int b2 = b + 1 == B ? 0 : b + 1;
            int ok = 1;
            if (lane == 0) {
                while (ld_dc(done + b) < uint32_t(r) || ld_dc(done + b2) < uint32_t(r)) {
                    if (wall_clock64() > t_end) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(4);
                }
            }
            ok = __builtin_amdgcn_readfirstlane(ok);
            if (!ok) {
                if (lane == 0) st_dc(ba.timed_out, 1u);
                break;
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);  // no row access moves above the wait
        }
        if constexpr (S == NOMAD_STORE_BF16) {  // problems may round differently
            if (a.round_mode == NOMAD_ROUND_STOCHASTIC)
                strat_tile<S, true, NOMAD_ROUND_STOCHASTIC>(a, b, r % B, a.epoch + r / B, wl[wave], kPairs);
            else
                strat_tile<S, true, NOMAD_ROUND_NEAREST>(a, b, r % B, a.epoch + r / B, wl[wave], kPairs);
        } else {
            strat_tile<S, true, NOMAD_ROUND_NEAREST>(a, b, r % B, a.epoch + r / B, wl[wave], kPairs);
        }
        wait_all_mem();
        if (lane == 0) st_dc(done + b, uint32_t(r + 1));
    }
}

// Explicit instantiations launched from capi.hip.
template __global__ void sgd_strat_kernel<NOMAD_STORE_FP32, NOMAD_ROUND_NEAREST>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_BF16, NOMAD_ROUND_NEAREST>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_BF16, NOMAD_ROUND_STOCHASTIC>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_FP64, NOMAD_ROUND_NEAREST>(SgdArgs);
template __global__ void sgd_batch_kernel<NOMAD_STORE_FP32>(BatchArgs);
template __global__ void sgd_batch_kernel<NOMAD_STORE_BF16>(BatchArgs);
template __global__ void sgd_batch_kernel<NOMAD_STORE_FP64>(BatchArgs);

