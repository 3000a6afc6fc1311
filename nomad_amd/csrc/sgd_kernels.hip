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
#include <type_traits>

#include "sgd_common.hpp"

using namespace nomad_dev;
using nomad_k::BatchArgs;
using nomad_k::kBatchEntries;
using nomad_k::SgdArgs;
using nomad_k::VisitDesc;


// ---------------------------------------------------------------------------
// Stratified schedule: tile (b, (b + substep) % B).  Columns of the tile in
// ascending order, rows of each column in ascending order; the step size uses
// the epoch as visit count (every (column, row block) pair is visited once per
// epoch, as with the reference's per-thread col_update_counts at one rank).
// The tile's non-empty column visits are listed ahead (tile_ptr / visits), so
// empty (column, row block) pairs cost nothing.  A strat context keeps its
// ratings in tile order (a tile's visits one after the other, each visit's
// rows ascending), so that the ratings of consecutive visits are contiguous.
// The single-problem kernel and the batched kernel both run a tile through
// this function, so a problem's arithmetic is the same in either launch.
//
// Visits run in groups of at most kK, each group as long as the first 64
// ratings of each of its visits lie among the group's first 64 * kQ (the
// first visit of a group always does).  A group's h rows and first 64 * kQ
// ratings are staged in LDS one group ahead: at the start of group g the
// registers holding group g + 1 (requested at the start of group g - 1) are
// written to its LDS buffer and group g + 2 is requested into them, with its
// descriptors requested one group earlier still.  Every wait for memory is
// then for requests a whole group old, and no register is moved while its
// load is in flight (a register copy of an in-flight load waits for it).  The
// h rows may be read early: a column is visited once per tile, and the tiles
// that wrote it earlier are complete; W rows are not read early.  The loop
// is written out once for W in LDS and once for W in HBM, so that neither
// holds the other's loads (a register that a load of the other path may
// still be writing would cost a wait for all loads).
//
// When the rows of W that the tile touches (trows, at most kStageRows) fit
// in the wave's LDS buffer, they are copied there at the start, packed in
// ascending row order; each rating finds its row's place through tslot,
// reads and writes it there, and the rows are copied back at the end: the
// rows' values, and the order in which the steps see them, are those of the
// HBM path, so the result is the same bit for bit, but a step waits for LDS
// instead of HBM.
template <int S, int WB = kTileLdsBytes>
struct TileLds {
    using T = Store<S>;
    using raw = typename T::raw;
    using R = typename T::rval;
    static constexpr int kB = base_storage(S);
    // staged W, WB bytes (the two-pair layout does not stage W: its rows take
    // the HBM path); a multiple of 64 entries
    static constexpr int kW = T::kNP == 1 ? WB / int(sizeof(raw)) : kWave;
    static_assert(kW % kWave == 0, "the staged W copy is loaded 64 entries at a time");
    static constexpr int kK = kB == NOMAD_STORE_FP64 ? 4 : 8;  // visits per group
    static constexpr int kQ = kB == NOMAD_STORE_BF16 ? (WB > kTileLdsBytes ? 3 : 4) : (kB == NOMAD_STORE_FP32 ? 2 : 1);  // x64 ratings
    struct Group {
        raw h[kK * kWave];      // visit u's h pair of lane l at [u * 64 + l]
        int32_t key[kQ * kWave];  // rating's LDS slot (W staged) or row
        R val[kQ * kWave];
    };
    raw w[kW];
    Group g[2];
};

template <int S, bool DC, int RM, class TL>
__device__ __forceinline__ void strat_tile(const SgdArgs& a, int b, int substep, int epoch, TL& L) {
    using T = Store<S>;
    using E = typename T::elem;
    using Aa = typename T::acc;
    using R = typename T::rval;
    using raw = typename T::raw;
    constexpr int kK = TL::kK, kQ = TL::kQ;
    const int lane = lane_id();
    const int B = a.nblocks;
    const int cb = (b + substep) % B;
    const int64_t tile = int64_t(b) * B + cb;
    // (every load below goes through gbl(): the batched kernel reads its
    // arguments from memory, so their pointers are generic)
    const int64_t v0 = gbl(a.tile_ptr)[tile], v1 = gbl(a.tile_ptr)[tile + 1];
    const int64_t tr0 = gbl(a.trow_ptr)[tile];
    const int32_t nrows = int32_t(gbl(a.trow_ptr)[tile + 1] - tr0);
    if (v0 >= v1) return;
    const Aa step = step_for<Aa>(a, epoch);
    const Aa reg = Aa(a.reg);
    E* __restrict__ H = static_cast<E*>(a.H);
    const gptr<R> rv = gbl(static_cast<const R*>(a.val));
    const int npairs = a.ldf >> 1;
    const int32_t rlo = gbl(a.wbound)[b], rhi = gbl(a.wbound)[b + 1];
    const RowBlock<S> rb = row_block<S>(a, rlo, rhi, lane);
    const uint32_t lkey = sr_lane_key(a.sr_seed, uint32_t(lane));
    const int32_t staged = nrows * npairs;
    // (the last entry of the LDS copy stays zero: the lanes past a row use it)
    const bool in_lds = T::kNP == 1 && nrows <= kStageRows && staged < TL::kW;
    constexpr int CP = DC ? kCpDevice : 0;

    // Group descriptors: lane u <= kK holds visit f + u (clamped to v1, the
    // next tile's first visit or the list's sentinel: in bounds, and its
    // start is this tile's end of ratings).
    struct GDesc {
        int64_t start;
        int32_t len, col;
    };
    auto load_gdesc = [&](int64_t f) {
        const gptr<VisitDesc> dd = gbl(a.visits) + min(f + min(lane, kK), v1);
        return GDesc{dd->start, dd->len, dd->worker};
    };
    auto start_of = [](const GDesc& d, int u) {
        return int64_t((uint64_t(uint32_t(readlane_i32(int32_t(uint64_t(d.start) >> 32), u))) << 32) |
                       uint32_t(readlane_i32(int32_t(d.start), u)));
    };
    // visits in the group that starts at f (d: its descriptors)
    auto group_len = [&](const GDesc& d, int64_t f) {
        const int64_t gs = start_of(d, 0);
        const bool ok = lane < kK && d.start - gs + min(d.len, kWave) <= kQ * kWave;
        const int nv = int(__builtin_ctzll(~__ballot(ok)));
        return int(min(int64_t(nv), v1 - f));
    };

    // Stored pair q of the LDS copy of W is pair q % npairs of the tile's row
    // q / npairs; the row ids sit one per lane in trw[] (row s in trw[s / 64]
    // of lane s % 64) and reach the lane that needs them by a lane permute.
    // (q + 0.5) / npairs in fp32 rounds to within 2^-20 of the exact
    // quotient, whose distance to the next integer is at least 0.5 / npairs.
    int32_t trw[kStageRows / kWave];
    const float inv_np = 1.0f / float(npairs);
    auto row_of_slot = [&](int32_t sq) {
        int32_t r = __shfl(trw[0], sq & (kWave - 1));
#pragma unroll
        for (int g = 1; g < kStageRows / kWave; ++g)
            if (nrows > g * kWave) {
                const int32_t rg = __shfl(trw[g], sq & (kWave - 1));
                r = (sq >> 6) == g ? rg : r;
            }
        return r;
    };
    auto staged_row = [&](int32_t q, int32_t& pr) {
        const int32_t sq = min(int32_t((float(q) + 0.5f) * inv_np), nrows - 1);
        pr = q - sq * npairs;
        return row_of_slot(sq);
    };

    // the first three groups' descriptors from one batch of 64 (a group has
    // at most kK visits, so the third ends before visit 3 * kK < 64)
    const GDesc d64 = [&] {
        const gptr<VisitDesc> dd = gbl(a.visits) + min(v0 + lane, v1);
        return GDesc{dd->start, dd->len, dd->worker};
    }();
    if (in_lds) {
#pragma unroll
        for (int g = 0; g < kStageRows / kWave; ++g) trw[g] = gbl(a.trows)[tr0 + min(g * kWave + lane, nrows - 1)];
    }
    // d64 seen from visit f: lane u gets d64's lane f - v0 + u (< 64, as
    // f - v0 <= 2 kK), which holds visit min(f + u, v1)
    auto shift = [&](int64_t f) {
        const int src = int(f - v0) + min(lane, kK);
        GDesc d;
        d.start = int64_t((uint64_t(uint32_t(__shfl(int32_t(uint64_t(d64.start) >> 32), src))) << 32) |
                          uint32_t(__shfl(int32_t(d64.start), src)));
        d.len = __shfl(d64.len, src);
        d.col = __shfl(d64.col, src);
        return d;
    };

    auto run = [&](auto w_in_lds) {
        constexpr bool LW = decltype(w_in_lds)::value;
        // staging registers of one group: kK h pairs, kQ ratings per lane
        struct Stage {
            raw h[kK];
            int32_t key[kQ];
            R val[kQ];
        };
        auto request = [&](const GDesc& d, int nv, Stage& sg) {
#pragma unroll
            for (int u = 0; u < kK; ++u) {
                const E* row = H + int64_t(readlane_i32(d.col, min(u, nv - 1))) * a.ldf;
                sg.h[u] = T::template load_raw_b<CP>(row_rsrc(row, npairs * T::kPairBytes),
                                                     T::lane_voff(lane, kWave * T::kNP));
            }
            const int64_t gs = start_of(d, 0);
            const int32_t n = int32_t(min(start_of(d, nv) - gs, int64_t(kQ * kWave)));
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const int64_t e = gs + min(q * kWave + lane, max(n - 1, 0));
                sg.key[q] = LW ? int32_t(gbl(a.tslot)[e]) : gbl(a.rowidx)[e];
                sg.val[q] = rv[e];
            }
        };
        auto publish = [&](const Stage& sg, typename TL::Group& G) {
#pragma unroll
            for (int u = 0; u < kK; ++u) G.h[u * kWave + lane] = sg.h[u];
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                G.key[q * kWave + lane] = sg.key[q];
                G.val[q * kWave + lane] = sg.val[q];
            }
        };
        if constexpr (LW) {
            // every staged pair in one round of loads (kW <= kU * 64)
            constexpr int kU = TL::kW / kWave;
            raw t[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (u * kWave >= staged) break;
                const int32_t q = u * kWave + lane;
                int32_t pr;
                const int32_t r = staged_row(q, pr);
                t[u] = T::template load_raw_b<CP>(rb.rsrc, q < staged ? T::voff_bytes(rb.soff(r) + pr * T::kPairBytes) : rb.drop());
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (u * kWave >= staged) break;
                if (u * kWave + lane < staged) L.w[u * kWave + lane] = t[u];
            }
            if (lane == 0) L.w[TL::kW - 1] = raw{};
        }
        // groups g (dc, from fc, nc visits), g + 1 (d1, f1, n1), g + 2 (d2, f2)
        int64_t fc = v0;
        GDesc dc = shift(fc);
        int nc = group_len(dc, fc);
        int64_t f1 = fc + nc;
        GDesc d1 = shift(f1);
        int n1 = f1 < v1 ? group_len(d1, f1) : 1;
        int64_t f2 = f1 + n1;
        GDesc d2 = shift(f2);
        Stage sg;
        request(dc, nc, sg);
        publish(sg, L.g[0]);
        request(d1, n1, sg);
        for (int p = 0; fc < v1; p ^= 1) {
            // group g + 1 to LDS, group g + 2 requested, descriptors of g + 3
            publish(sg, L.g[p ^ 1]);
            const int n2 = f2 < v1 ? group_len(d2, f2) : 1;
            request(d2, n2, sg);
            const int64_t f3 = f2 + n2;
            const GDesc d3 = load_gdesc(f3);
            const typename TL::Group& G = L.g[p];
            const int64_t gs = start_of(dc, 0);
            for (int u = 0; u < nc; ++u) {
                const int64_t st = start_of(dc, u);
                const int32_t ln = readlane_i32(dc.len, u), j = readlane_i32(dc.col, u);
                typename T::pair h = T::unpack(G.h[u * kWave + lane]);
                // the visit's first 64 ratings are staged (the group's rule)
                const int32_t o = int32_t(st - gs) + min(lane, ln - 1);
                const int32_t key = G.key[o];
                const R val = G.val[o];
                if constexpr (LW) {
                    const int32_t row = RM == NOMAD_ROUND_STOCHASTIC ? row_of_slot(key) : 0;
                    run_visit_lds<S, RM>(a, L.w, npairs, TL::kW - 1, h, j, st, ln, row, key, val, step, reg,
                                         uint32_t(epoch),
                                         lkey);
                } else {
                    run_visit_pre<S, DC, RM>(a, rb, h, j, st, ln, key, val, step, reg, uint32_t(epoch), lkey);
                }
                const uint32_t rnd = sr_bits_k(a.sr_seed, 0x80000000u | uint32_t(b),
                                               uint32_t(j) * 0x01000193u + uint32_t(epoch), lkey);
                row_store<S, DC>(H + int64_t(j) * a.ldf, lane, npairs, h, rnd, RM);
            }
            fc = f1;
            nc = n1;
            dc = d1;
            f1 = f2;
            n1 = n2;
            d1 = d2;
            f2 = f3;
            d2 = d3;
        }
    };
    if constexpr (T::kNP == 1) {
        if (in_lds)
            run(std::true_type{});
        else
            run(std::false_type{});
    } else {
        run(std::false_type{});
    }
    if constexpr (T::kNP == 1) {
        if (!in_lds) return;
        // the rows back to W (the range check drops the lanes past them)
        for (int32_t q0 = 0; q0 < staged; q0 += kWave) {
            const int32_t q = q0 + lane;
            int32_t pr;
            const int32_t r = staged_row(q, pr);
            T::template store_raw_b<CP>(rb.rsrc, q < staged ? T::voff_bytes(rb.soff(r) + pr * T::kPairBytes) : rb.drop(),
                                        L.w[min(q, staged - 1)]);
        }
    }
}

// (RM: the rounding of bf16 stores, a template argument so that no store
// branches on it; fp32 and fp64 ignore it)
template <int S, int RM>
__global__ __launch_bounds__(64) void sgd_strat_kernel(SgdArgs a) {
    __shared__ TileLds<S> tl;
    const int b = blockIdx.x;
    if (b >= a.nblocks) return;
    strat_tile<S, false, RM>(a, b, a.substep, a.epoch, tl);
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
// Tickets: problems advance in proportion to their strata.  Problem p runs
// s_p rounds per ticket round ("stride"; s_p = ceil(B_p / Bref), so that a
// problem of many thin strata does not run on alone after the others have
// finished): it appears as s_p entries (p, j), j < s_p, and entry (p, j) of
// ticket round u is round u * s_p + j of p.  Entries are sorted by their
// ticket rounds descending, so that those still running in a ticket round are
// a prefix of the sorted list, and within equal counts by the midpoint of
// their share of the ticket round, (j + 1/2) / s_p, so that the rounds of a
// strided problem are spread between the other problems' tiles.  Round
// u * s_p + j - 1 of p then always has smaller tickets than round u * s_p + j.
// A wave takes the next ticket, waits for the two parent tiles (their tickets
// are smaller, so they were taken by waves that are running), runs the tile
// and publishes it.  The smallest unfinished ticket can always run, so the
// launch needs no co-residency and finishes for any grid size.
//
// Publication: the next tile on a block may run on another XCD, whose L2 is
// not this one's.  W and H rows are therefore read and written with
// device-coherent (sc1) accesses, and a tile's progress counter is raised
// only after all of the wave's stores have completed (s_waitcnt 0), so a
// wave that sees the counter reads the new rows; no L2 writeback or
// invalidate is needed.  Counters live in uncached memory and are polled
// with device-coherent loads.
// Staged W per wave in the batched kernel: bf16 takes the LDS that two
// 4-wave workgroups per CU leave over, and 1 KiB more from staging 3 x 64
// ratings per visit group instead of 4 x 64 (11,520 B instead of 8 KiB),
// so that more tiles keep their rows in LDS.
template <int S>
constexpr int batch_stage_bytes() {
    return S == NOMAD_STORE_BF16 ? 11520 : kTileLdsBytes;
}

template <int S>
__global__ __launch_bounds__(256) void sgd_batch_kernel(BatchArgs ba) {
    __shared__ int64_t seg_t0[kBatchEntries];
    __shared__ int32_t seg_r0[kBatchEntries];
    __shared__ int32_t seg_w[kBatchEntries];
    __shared__ int32_t boff[kBatchEntries + 1];
    __shared__ int32_t order[kBatchEntries];
    __shared__ TileLds<S, batch_stage_bytes<S>()> tl[4];  // per wave (256-thread workgroups)
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
        // segment: ticket rounds of constant width (the same entries still
        // running); the last segment with seg_t0 <= t
        int sg = 0, sh = G;
        while (sh - sg > 1) {
            const int mid = (sg + sh) >> 1;
            if (uint64_t(seg_t0[mid]) <= t)
                sg = mid;
            else
                sh = mid;
        }
        const uint64_t rel = t - uint64_t(seg_t0[sg]);
        const int32_t u = seg_r0[sg] + int32_t(rel / uint32_t(seg_w[sg]));
        const int32_t local = int32_t(rel % uint32_t(seg_w[sg]));
        int q = 0, qh = Q;  // the last entry with boff <= local
        while (qh - q > 1) {
            const int mid = (q + qh) >> 1;
            if (boff[mid] <= local)
                q = mid;
            else
                qh = mid;
        }
        const uint32_t ent = uint32_t(__builtin_amdgcn_readfirstlane(order[q]));
        const int p = int(ent & 0xFFu);
        const int32_t r = u * int32_t(ent >> 18) + int32_t((ent >> 8) & 0x3FFu);  // the problem's round
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
        if constexpr (base_storage(S) == NOMAD_STORE_BF16) {  // problems may round differently
            if (a.round_mode == NOMAD_ROUND_STOCHASTIC)
                strat_tile<S, true, NOMAD_ROUND_STOCHASTIC>(a, b, r % B, a.epoch + r / B, tl[wave]);
            else
                strat_tile<S, true, NOMAD_ROUND_NEAREST>(a, b, r % B, a.epoch + r / B, tl[wave]);
        } else {
            strat_tile<S, true, NOMAD_ROUND_NEAREST>(a, b, r % B, a.epoch + r / B, tl[wave]);
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
// rows of more than 128 elements (two element pairs per lane)
template __global__ void sgd_strat_kernel<NOMAD_STORE_FP32 | kWideBit, NOMAD_ROUND_NEAREST>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_BF16 | kWideBit, NOMAD_ROUND_NEAREST>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_BF16 | kWideBit, NOMAD_ROUND_STOCHASTIC>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_FP64 | kWideBit, NOMAD_ROUND_NEAREST>(SgdArgs);
template __global__ void sgd_batch_kernel<NOMAD_STORE_FP32 | kWideBit>(BatchArgs);
template __global__ void sgd_batch_kernel<NOMAD_STORE_BF16 | kWideBit>(BatchArgs);
template __global__ void sgd_batch_kernel<NOMAD_STORE_FP64 | kWideBit>(BatchArgs);

