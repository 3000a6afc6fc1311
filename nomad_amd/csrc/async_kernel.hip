// Asynchronous NOMAD token kernel for gfx950 (two-level).
//
// The reference runs T threads per machine; a column token visits every
// thread of its machine in the order of its permutation and then moves to
// another machine (Code/nomad/nb_main.cpp:491-565, nb_sendrecv.cpp:189-204).
// Here a workgroup of NW waves is the machine and each wave a thread:
//
//   * the workgroup owns NW consecutive row blocks of W, wave w the rows
//     [wbound[g*NW+w], wbound[g*NW+w+1]) (the reference splits rows evenly,
//     nb_data.cpp:83-89; the host may instead balance the blocks by ratings);
//   * columns arrive on the workgroup's ring in HBM (multi-producer, one
//     consumer position in LDS).  An idle wave claims a batch of them, copies
//     each h_j into an LDS slot (FP32, or FP64 for double storage) together
//     with a window of the column's first 64 ratings in this row range and
//     the offsets at which each wave's rows start inside that window, and
//     hands the slot to the wave owning the column's first row here;
//   * a wave takes every slot waiting in its inbox whose piece (its rows of
//     the column) lies inside the slot's window, up to 64 ratings in all, and
//     runs their SGD steps as one stream (nb_main.cpp:520-542; rows ascending
//     within a column, columns in inbox order).  Rows of W are prefetched 8-32
//     ratings ahead across column boundaries; a row that occurs twice within
//     that distance takes the value of its earlier update from a per-wave LDS
//     ring of the rows the stream has stored.
//     Slots whose piece is not inside the window are run one at a time,
//     reading their ratings from HBM and refilling the window;
//   * h goes back to the slot and the slot to the wave owning the column's
//     next row; after the last piece h_j returns to HBM and the column is
//     pushed to the next workgroup of its route; after the last workgroup of
//     the route the column's cycle count advances, and a column whose count
//     reaches target_cycle is parked.
//
// Step size: the reference keeps one visit count per (thread, column)
// (nb_main.cpp:458, 514-515, 544).  Here a workgroup's waves see a column in
// one hop, so the count is kept per (column, workgroup) pair in HBM: read when
// the hop is claimed (the step follows from it), written back one higher when
// the column leaves the workgroup.  At the deadline of a timed segment every
// workgroup stops claiming and finishes the hops it holds, so every visit is
// whole and counted, as a reference thread finishes its column before it sees
// the stop flag (nb_main.cpp:472-570); the host then starts every column at
// the head of a fresh route, as set_perm does (pool.cpp:32-38).
//
// Cross-workgroup data (h_j, the column's TokInfo) moves with device-coherent
// (sc1) loads and stores; a producer waits for all its memory operations to
// complete before it publishes ring slots, so no L2 writeback or invalidate
// is needed.  W rows are private to their wave.
#include <hip/hip_runtime.h>

#include "sgd_common.hpp"

using namespace nomad_dev;
using nomad_k::RouteDesc;
using nomad_k::TokInfo;
using nomad_k::async_slots;

namespace {

constexpr int kFetchBatch = 16;  // columns claimed from the ring per fetch
constexpr int kBatch = 16;       // inbox slots examined per stream
// Rows of W in flight inside a stream, and updated rows kept for forwarding
// (per wave, in LDS: a ring of as many rows as this prefetch distance).  At
// 8 waves 4 rows for fp32 and bf16: the room that lets a workgroup hold 128
// token slots, and no slower than 8 or 16 there (bf16, Netflix shape: 4.18e9
// updates/s against 4.12e9 with 8); otherwise the rings of a workgroup take
// 64 KiB.
template <int S, int NW>
constexpr int stream_depth() {
    // (32 KiB of rings in the two-pair layout, whose slots are larger)
    return (NW == 8 && base_storage(S) != NOMAD_STORE_FP64)
               ? 4
               : (64 * 1024 / Store<S>::kNP) / (NW * kWave * 2 * int(sizeof(typename Store<S>::acc)) * Store<S>::kNP);
}

// W rows in flight inside one long piece: deep enough to cover HBM latency
// with one wave per SIMD (4 waves per workgroup), less where more waves share
// the SIMD (at 8 waves 8 rows were as fast as 16)
template <int S, int NW>
constexpr int long_prefetch_depth() {
    return (NW <= 4 ? 32 : 8) / Store<S>::kNP;
}

template <int S, int NW, int SLOTS>
struct AsyncShared {
    using A = typename Store<S>::acc;
    using R = typename Store<S>::rval;
    static constexpr int kPbw = (NW + 4) & ~3;
    static constexpr int kVd = stream_depth<S, NW>();
    typename Store<S>::raw vic[NW][kVd][kWave];  // per wave: the last kVd updated rows of its stream,
                                                 // as stored
    A h[SLOTS][2 * kWave * Store<S>::kNP];  // column factors of resident tokens
    int32_t wrow[SLOTS][kWave];  // window of the column's ratings here: rows
    R wval[SLOTS][kWave];        //                                    values
    int32_t col[SLOTS];
    uint32_t start[SLOTS];       // CSC offset of the column's ratings in this row range
    int32_t len[SLOTS];
    int32_t cur[SLOTS];          // ratings of the range already processed
    int32_t wb[SLOTS];           // the window holds range offsets [wb, we)
    int32_t we[SLOTS];
    int32_t cyc[SLOTS];          // completed route cycles of the column
    int32_t cs[SLOTS];           // visit counter of the (column, workgroup) pair
    int32_t vc[SLOTS];           //   its value at the claim
    A stp[SLOTS];                //   and the step size it gives
    int32_t npos[SLOTS];         // route position of the column's next hop
    int32_t nwg[SLOTS];          //   its workgroup
    uint32_t nstart[SLOTS];      //   its CSC range
    int32_t nlen[SLOTS];
    int32_t ncs[SLOTS];          //   its visit counter
    uint8_t pb[SLOTS][kPbw];     // pb[s][w]: window entries with a row below wave w's first row
    uint32_t inbox[NW][SLOTS];   // per wave: (lap << 16) | slot
    uint32_t tpush[SLOTS];       // wall clock (low 32 bits) when the slot was last handed over
    uint32_t itail[NW];
    uint32_t freebits[SLOTS / 32];
    uint32_t lock;               // serialises ring claims and slot allocation
    uint32_t ghead;              // consumer position of this workgroup's ring
    uint32_t stop;               // set at the deadline: no more claims
};

__device__ __forceinline__ uint32_t lds_ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Hand slot s to wave w of this workgroup.  Called by a whole wave; the
// workgroup-scope release orders this wave's earlier LDS writes of the slot.
template <int SLOTS, class L>
__device__ __forceinline__ void inbox_push(L& sh, int w, int s, bool timing) {
    if (lane_id() == 0) {
        if (timing) sh.tpush[s] = uint32_t(wall_clock64());
        const uint32_t t =
            __hip_atomic_fetch_add(&sh.itail[w], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t lap = ((t / uint32_t(SLOTS)) + 1u) & 0xFFFFu;
        __hip_atomic_store(&sh.inbox[w][t & uint32_t(SLOTS - 1)], (lap << 16) | uint32_t(s), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// Wave of this workgroup that owns (wave-uniform) row r.
__device__ __forceinline__ int wave_of(int32_t lane_bound, int32_t r) {
    return __popcll(__ballot(r >= lane_bound));
}

// Same for a per-lane row.
template <int NW>
__device__ __forceinline__ int wave_of_lane(int32_t lane_bound, int32_t r) {
    int w = 0;
#pragma unroll
    for (int b = 0; b < NW - 1; ++b) w += r >= readlane_i32(lane_bound, b) ? 1 : 0;
    return w;
}

// Offsets of the waves' rows inside a freshly written window of slot s:
// lane l holds the window's row l (valid for l < wlen).  pb[s][w] = number of
// window rows below the first row of wave w.
template <int NW, class L>
__device__ __forceinline__ void write_bounds(L& sh, int s, int32_t myrow, int32_t wlen, int32_t lane_bound) {
    const int lane = lane_id();
    int32_t mine = lane == NW ? wlen : 0;
#pragma unroll
    for (int w = 1; w < NW; ++w) {
        const int32_t c = __popcll(__ballot(lane < wlen && myrow < readlane_i32(lane_bound, w - 1)));
        if (lane == w) mine = c;
    }
    if (lane <= NW) sh.pb[s][lane] = uint8_t(mine);
}

__device__ __forceinline__ void tok_store(TokInfo* p, uint32_t start, int32_t len, int32_t pos, int32_t cycle,
                                          int32_t cslot) {
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
    st_dc(q, (uint64_t(uint32_t(len)) << 32) | start);
    st_dc(q + 1, (uint64_t(uint32_t(cycle)) << 32) | uint32_t(pos));
    st_dc(q + 2, uint64_t(uint32_t(cslot)));
}

}  // namespace

template <int S, int NW, int SLOTS, int RM>
__global__ __launch_bounds__(NW * 64) void sgd_async_kernel(SgdArgs a) {
    using T = Store<S>;
    using E = typename T::elem;
    using Aa = typename T::acc;
    using R = typename T::rval;
    using P = typename T::pair;
    static_assert(SLOTS % 32 == 0 && (SLOTS & (SLOTS - 1)) == 0, "SLOTS must be a power of two >= 32");
    static_assert(NW <= 16 && kBatch <= kWave, "layout limits");
    constexpr int kPf = stream_depth<S, NW>();
    static_assert(kWave % kPf == 0 && kWave % long_prefetch_depth<S, NW>() == 0, "depths must divide the wave");
    __shared__ AsyncShared<S, NW, SLOTS> sh;
    const int g = blockIdx.x;
    // wave-uniform values are made scalar so that branches on them stay
    // scalar (a branch the compiler must treat as divergent ends every
    // outstanding load with a full wait)
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int lane = lane_id();
    const int npairs = a.ldf >> 1;
    const int64_t ldf = a.ldf;
    const uint32_t lkey = sr_lane_key(a.sr_seed, uint32_t(lane));
    // Stochastic-rounding bits of the W stores: the lane's key plus a Weyl
    // sequence of the wave (a scalar that advances by the golden-ratio
    // constant per rating, started from a hash of seed, wave and launch), so
    // a rating costs one vector add for its bits; the 16-bit halves of the
    // sum are equidistributed per lane.  (Column stores keep the hashed bits.)
    uint32_t sr_acc = mix32(a.sr_seed ^ mix32(uint32_t(g * NW + wave) * 0x9E3779B1u + uint32_t(a.epoch)));
    const int32_t* wb_g = a.wbound + int64_t(g) * NW;
    const int32_t row_lo = __builtin_amdgcn_readfirstlane(wb_g[wave]);
    const int32_t row_hi = __builtin_amdgcn_readfirstlane(wb_g[wave + 1]);
    // lane t < NW - 1 holds the first row of wave t + 1 of this workgroup
    const int32_t lane_bound = lane < NW - 1 ? wb_g[lane + 1] : INT_MAX;
    const uint32_t gmask = (1u << a.qlog2) - 1u;
    const int slot_cap = a.slot_cap > 0 && a.slot_cap < SLOTS ? a.slot_cap : SLOTS;
    const uint64_t* ring = a.qslot + (int64_t(g) << a.qlog2);
    E* __restrict__ W = static_cast<E*>(a.W);
    E* __restrict__ H = static_cast<E*>(a.H);
    const int32_t* __restrict__ rowidx = a.rowidx;
    const R* __restrict__ val = static_cast<const R*>(a.val);
    const Aa reg = Aa(a.reg);
    // This wave's rows of W through one buffer resource.  A row's byte offset
    // within the block is the scalar offset of its load or store, always a
    // row of this wave; the hardware range check covers the per-lane offset
    // (not the scalar one), which is the lane's pair offset, 2^31 for lanes
    // past the row, or drop_voff with a scalar offset of 0 for a store to drop
    // (out-of-range loads read zero, stores are dropped).
    const int rowbytes = a.ldf * int(sizeof(E));
    const __amdgpu_buffer_rsrc_t wrs = __builtin_amdgcn_make_buffer_rsrc(
        W + int64_t(row_lo) * ldf, 0, (row_hi - row_lo) * rowbytes, 0x00020000);
    const typename T::voff_t wvoff = T::lane_voff(lane, npairs);
    const Aa sel63 = lane == kWave - 1 ? Aa(1) : Aa(0);  // where a rating enters the dot product's tree

    for (int i = threadIdx.x; i < NW * SLOTS; i += blockDim.x) (&sh.inbox[0][0])[i] = 0u;
    if (threadIdx.x < NW) sh.itail[threadIdx.x] = 0u;
    if (threadIdx.x < SLOTS / 32) sh.freebits[threadIdx.x] = 0xFFFFFFFFu;
    if (threadIdx.x == 0) {
        sh.lock = 0u;
        sh.stop = 0u;
        sh.ghead = a.qhead[int64_t(g) * nomad_k::kQStride];
    }
    __syncthreads();

    // Per-phase wall-clock counters only when asked for: each clock read is a
    // scalar memory round trip.  The deadline is checked every 64 streams or
    // visits and every 8 idle polls; the first wave past it stops the
    // workgroup's claims.
    const bool timing = a.timing != 0;
    auto clk = [timing]() -> uint64_t { return timing ? wall_clock64() : 0; };
    const uint64_t t_end = wall_clock64() + a.deadline;
    uint32_t ihead = 0;
    int64_t n_upd = 0, n_vis = 0, n_idle = 0, n_fetch = 0, n_fetched = 0, n_exit = 0, n_res = 0, n_blocked = 0;
    int64_t tk_visit = 0, tk_exit = 0, tk_fetch = 0, tk_chunk = 0, n_depth = 0, n_work = 0;
    int64_t tk_transit_l = 0, tk_inbox_l = 0;  // per lane
    bool hit_deadline = false;
    int idle = 0;
    auto past_deadline = [&]() {
        if (!hit_deadline && wall_clock64() > t_end) {
            hit_deadline = true;
            if (lane == 0) __hip_atomic_store(&sh.stop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    };

    for (;;) {
        // ---- slots waiting in this wave's inbox -----------------------------
        uint32_t pe = 0;
        bool rdy = false;
        if (lane < kBatch) {
            pe = lds_ld(&sh.inbox[wave][(ihead + uint32_t(lane)) & (SLOTS - 1)]);
            rdy = (pe >> 16) == ((((ihead + uint32_t(lane)) / SLOTS) + 1u) & 0xFFFFu);
        }
        const int nr = int(__builtin_ctzll(~__ballot(rdy)));  // lanes >= kBatch are never ready
        if (nr > 0) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            idle = 0;
            const uint64_t t_v0 = clk();
            n_depth += int64_t(lds_ld(&sh.itail[wave]) - ihead);
            // lane t < nr: the t-th waiting slot and its piece in the window
            const int ts = int(pe & uint32_t(SLOTS - 1));
            int32_t t_wb = 0, t_lo = 0, t_hi = 0, t_cnt = 0;
            bool t_ok = false;
            if (lane < nr) {
                const int32_t cur = sh.cur[ts];
                t_wb = sh.wb[ts];
                const int32_t we = sh.we[ts];
                const int32_t wlen = we - t_wb;
                t_lo = cur - t_wb;
                t_hi = sh.pb[ts][wave + 1];
                t_ok = cur >= t_wb && cur < we && (t_hi < wlen || we == sh.len[ts]) && t_hi > t_lo;
                t_cnt = t_ok ? t_hi - t_lo : 0;
                if (timing) tk_inbox_l += int64_t(uint32_t(t_v0) - sh.tpush[ts]);
            }
            // the longest run of leading slots whose pieces fit one 64-rating stream
            int nt = 0;
            int32_t total = 0, t_pref = 0;
            for (int t = 0; t < nr; ++t) {
                const int32_t c = readlane_i32(t_cnt, t);
                if (c == 0 || total + c > kWave) break;
                if (lane == t) t_pref = total;
                total += c;
                ++nt;
            }
            if (nt > 0) {
                // ============ stream of nt pieces, total <= 64 ratings ===========
                ihead += uint32_t(nt);
                n_vis += nt;
                n_upd += total;
                ++n_work;
                // lane l < total: rating l of the stream, from piece my_t
                int my_t = 0;
                for (int t = 1; t < nt; ++t) my_t += lane >= readlane_i32(t_pref, t) ? 1 : 0;
                const int ms = __shfl(ts, my_t);
                const int32_t mo = __shfl(t_lo, my_t) + lane - __shfl(t_pref, my_t);
                int32_t srow = INT_MAX;
                R sval = R(0);
                if (lane < total) {
                    srow = sh.wrow[ms][mo];
                    sval = sh.wval[ms][mo];
                }
                // byte offset of each rating's row inside this wave's block: it
                // goes to the scalar offset of the row's load and store (always
                // a row of this wave; the per-lane offset keeps the range check)
                const int32_t soffv = lane < total ? (srow - row_lo) * rowbytes : 0;
                // The stream's first W rows are requested as soon as the rows are
                // known; the rest of the set-up runs while they arrive.  Each
                // prologue load is preceded by a store that the range check
                // drops, so the prologue issues memory operations in the same
                // pattern as the loop and the compiler's wait counts stay as
                // loose as the prefetch distance allows.
                const uint64_t t_c0 = clk();
                typename T::raw wp[kPf];
                int32_t soff[kPf];  // row offset of the load into wp[d] (scalar)
#pragma unroll
                for (int d = 0; d < kPf; ++d) {
                    soff[d] = readlane_i32(soffv, min(d, total - 1));
                    // (distinct offsets: stores to one offset would be merged away)
                    T::store_b(wrs, T::drop_voff(wvoff, d), 0, T::zero(), 0u, RM);
                    wp[d] = T::load_raw_b(wrs, wvoff, soff[d]);
                }
                Aa t_step = Aa(0);
                uint32_t t_srb = 0;
                int32_t t_cyc = 0;
                if (lane < nt) {
                    t_cyc = sh.cyc[ts];
                    t_step = sh.stp[ts];
                    t_srb = uint32_t(sh.col[ts]) * 0x01000193u + uint32_t(t_cyc);
                }
                // distance back to the latest earlier rating of the same row
                // inside the prefetch window (0: none).  Its prefetched copy
                // predates that rating's update, so the row is taken from the
                // ring of updated rows instead.
                // (rows of one column are distinct: a single piece has none)
                // (all shuffles first, into registers of their own, so that they
                // are in flight together rather than one LDS round trip each)
                int32_t pdist = 0;
                if (nt > 1) {
                    int32_t prev[kPf];
#pragma unroll
                    for (int b = 1; b < kPf; ++b) prev[b] = __shfl_up(srow, b);
                    __builtin_amdgcn_sched_barrier(0);  // (keeps the compiler from interleaving them)
#pragma unroll
                    for (int b = kPf - 1; b >= 1; --b)
                        if (lane >= b && lane < total && prev[b] == srow) pdist = b;
                }
                // scalar mask over the stream's ratings: the first rating of each
                // piece after the first
                const uint64_t pmask = __ballot(lane > 0 && lane < total && __shfl_up(my_t, 1) != my_t);

                int cur_t = 0;
                int s_cur = readlane_i32(ts, 0);
                P h;
                h = T::lds_get(sh.h[s_cur], lane);
                // per piece: c1 = 1 - step * reg and -step (the step's two coefficients)
                const Aa t_c1 = Aa(1) - t_step * reg;
                Aa c1 = readlane_val(t_c1, 0);
                Aa nstep = -__shfl(t_step, 0);  // (kept in a vector register: the error is scalar)
                // lane u: the scalar offset of the row that rating u prefetches
                // for rating u + kPf (the last rating's row past the stream)
                const int32_t soffN = __shfl(soffv, min(lane + kPf, total - 1));
                // kPf ratings per iteration.  The loop is left before the first
                // rating past the stream (an exit edge: every path through the
                // body issues the same memory operations, so the compiler's
                // wait counts stay exact).
                for (int u0 = 0;; u0 += kPf) {
                    const int left = total - u0;
                    const uint32_t pm = uint32_t(pmask >> u0);  // piece starts of this iteration
#pragma unroll
                    for (int d = 0; d < kPf; ++d) {
                        if (d >= left) goto stream_done;
                        const int u = u0 + d;
                        if ((pm >> d) & 1u) {  // next piece: park h, take the next column's
                            T::lds_put(sh.h[s_cur], lane, h);
                            ++cur_t;
                            s_cur = readlane_i32(ts, cur_t);
                            h = T::lds_get(sh.h[s_cur], lane);
                            c1 = readlane_val(t_c1, cur_t);
                            nstep = -__shfl(t_step, cur_t);
                        }
                        // a row updated pd ratings ago is read back from the ring
                        // of updated rows
                        P w = T::unpack(wp[d]);
                        const int pd = readlane_i32(pdist, u);
                        P fw = T::zero();
                        if (pd != 0) fw = T::unpack(sh.vic[wave][(d - pd + kPf) % kPf][lane]);
                        w = T::select(pd != 0, fw, w);
                        sgd_step_b(w, h, Aa(readlane_val(sval, u)), sel63, c1, nstep);
                        sr_acc += 0x9E3779B9u;
                        const typename T::raw wr = T::pack(w, lkey + sr_acc, RM);
                        // (the offset is uniform; for bf16 the compiler carries some
                        // soff[d] across the loop in a vector register, and without
                        // the readfirstlane it wraps the store in a per-value loop)
                        T::store_raw_b(wrs, wvoff, wr, __builtin_amdgcn_readfirstlane(soff[d]));
                        sh.vic[wave][d][lane] = wr;
                        soff[d] = readlane_i32(soffN, u);
                        wp[d] = T::load_raw_b(wrs, wvoff, soff[d]);
                    }
                }
            stream_done:
                T::lds_put(sh.h[s_cur], lane, h);
                tk_chunk += int64_t(clk() - t_c0);

                // ---- hand each slot on: next wave here, or out of the workgroup
                const int32_t ncur = t_wb + t_hi;
                bool t_exit = false;
                int t_nw = 0;
                if (lane < nt) {
                    t_exit = ncur == sh.len[ts];
                    // an eligible piece that does not end the range ends inside the window
                    if (!t_exit) t_nw = wave_of_lane<NW>(lane_bound, sh.wrow[ts][t_hi]);
                    sh.cur[ts] = ncur;
                    if (timing && !t_exit) sh.tpush[ts] = uint32_t(wall_clock64());
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if (lane < nt && !t_exit) {
                    const uint32_t p =
                        __hip_atomic_fetch_add(&sh.itail[t_nw], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const uint32_t lap = ((p / uint32_t(SLOTS)) + 1u) & 0xFFFFu;
                    __hip_atomic_store(&sh.inbox[t_nw][p & uint32_t(SLOTS - 1)], (lap << 16) | uint32_t(ts),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                const uint64_t xmask = __ballot(t_exit);
                const uint64_t t_v1 = clk();
                tk_visit += int64_t(t_v1 - t_v0);
                if (xmask != 0) {
                    // columns leaving the workgroup: h_j and TokInfo to HBM, one
                    // wait for all of them, then their ring entries
                    int32_t t_j = 0, t_np = 0, t_ncyc = 0;
                    bool t_park = false;
                    if (t_exit) {
                        t_j = sh.col[ts];
                        t_np = sh.npos[ts];
                        t_ncyc = t_cyc + (t_np == 0 ? 1 : 0);
                        t_park = t_np == 0 && t_ncyc >= a.target_cycle;
                        // the pair's visit is complete: its counter goes up by one
                        st_dc(a.vcnt + sh.cs[ts], sh.vc[ts] + 1);
                    }
                    for (uint64_t mm = xmask; mm != 0; mm &= mm - 1) {
                        const int t = int(__builtin_ctzll(mm));
                        const int s = readlane_i32(ts, t);
                        const int32_t j = readlane_i32(t_j, t);
                        const P hx = T::lds_get(sh.h[s], lane);
                        const uint32_t rnd = sr_bits_k(a.sr_seed, 0x80000000u | uint32_t(g),
                                                     uint32_t(readlane_i32(int32_t(t_srb), t)), lkey);
                        T::store_dc(H + int64_t(j) * ldf, lane, npairs, hx, rnd, RM);
                        if (lane == 0) {
                            tok_store(a.tok + j, sh.nstart[s], sh.nlen[s], readlane_i32(t_np, t),
                                      readlane_i32(t_ncyc, t), sh.ncs[s]);
                            if (timing) st_dc(a.tok_stamp + j, uint64_t(wall_clock64()));
                        }
                    }
                    // the ring positions are reserved while the stores drain (a
                    // reserved entry is not consumed before it is published)
                    int32_t q = 0;
                    uint32_t p = 0;
                    if (t_exit && !t_park) {
                        q = sh.nwg[ts];
                        p = __hip_atomic_fetch_add(a.qtail + int64_t(q) * nomad_k::kQStride, 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                    }
                    wait_all_mem();
                    if (t_exit) {
                        if (t_park) {
                            __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        } else {
                            const uint64_t v = (uint64_t((p >> a.qlog2) + 1u) << 32) | uint32_t(t_j);
                            st_dc(a.qslot + (int64_t(q) << a.qlog2) + (p & gmask), v);
                        }
                        __hip_atomic_fetch_or(&sh.freebits[ts >> 5], 1u << (ts & 31), __ATOMIC_RELEASE,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    n_exit += __popcll(xmask);
                    tk_exit += int64_t(clk() - t_v1);
                }
                if ((n_work & 63) == 0) past_deadline();
                continue;
            }

            // ============ one slot whose piece is not inside its window ==========
            ++ihead;
            ++n_work;
            const int s = readlane_i32(ts, 0);
            const int32_t j = sh.col[s];
            const uint32_t vstart = sh.start[s];
            const int32_t vlen = sh.len[s];
            int32_t cur = sh.cur[s];
            int32_t wb = sh.wb[s], we = sh.we[s];
            const int32_t cycle = sh.cyc[s];
            const Aa step = sh.stp[s];
            const uint32_t sr_b = uint32_t(j) * 0x01000193u + uint32_t(cycle);
            const Aa c1 = Aa(1) - step * reg;
            P h = T::lds_get(sh.h[s], lane);
            // Stream through the column's ratings from cur while the rows are
            // this wave's: two 64-rating chunks of the CSC range in registers,
            // W rows kPfL ratings ahead across the chunk boundary.  Rows of
            // one column are distinct, so no update needs forwarding.
            constexpr int kPfL = long_prefetch_depth<S, NW>();
            const int64_t cbase = int64_t(vstart);
            int32_t c0 = cur;  // range offset of chunk A
            // the 64 ratings from range offset o (rows past the range: INT_MAX).
            // The loads are not under a branch: a load only some lanes issue
            // leaves the compiler unsure of what is in flight, and the loop
            // after a chunk change would then start with a full wait.
            auto load_chunk = [&](int32_t o, int32_t& r, R& v) {
                const int64_t e = cbase + min(o + lane, vlen - 1);
                const int32_t r0 = rowidx[e];
                const R v0 = val[e];
                const bool in = o + lane < vlen;
                r = in ? r0 : INT_MAX;
                v = in ? v0 : R(0);
            };
            int32_t rA, rB;
            R vA, vB;
            load_chunk(c0, rA, vA);
            load_chunk(c0 + kWave, rB, vB);
            const uint64_t t_c0 = clk();
            // Scalar byte offsets of the rows (see the stream): for lane u of
            // chunk A its own row, and for the prefetch the row kPfL ratings
            // ahead (in chunk B past A's end).  Rows past the piece (another
            // wave's, or past the range) are never updated; their prefetch
            // slots read the wave's first row.
            auto row_offset = [&](int32_t r) { return r < row_hi ? (r - row_lo) * rowbytes : 0; };
            auto ahead_offsets = [&]() {
                const int32_t ra = __shfl(rA, (lane + kPfL) & (kWave - 1));
                const int32_t rb = __shfl(rB, (lane + kPfL) & (kWave - 1));
                return row_offset(lane + kPfL < kWave ? ra : rb);
            };
            int32_t offN = ahead_offsets();
            typename T::raw wp[kPfL];
            int32_t soff[kPfL];
            {
                const int32_t offA = row_offset(rA);
#pragma unroll
                for (int d = 0; d < kPfL; ++d) {
                    soff[d] = readlane_i32(offA, d);
                    T::store_b(wrs, T::drop_voff(wvoff, d), 0, T::zero(), 0u, RM);  // dropped: see the stream prologue
                    wp[d] = T::load_raw_b(wrs, wvoff, soff[d]);
                }
            }
            const Aa nstep = -step;
            // piece ratings in chunk A: rows ascend, so they are a prefix
            int32_t cntA = __popcll(__ballot(rA < row_hi));
            int32_t cnt = 0;
            // full groups of kPfL ratings (u0 + d < kWave); the piece's last
            // cntA mod kPfL ratings follow the chunk loop
            auto rating = [&](int uu, int d) {
                P w = T::unpack(wp[d]);
                sgd_step_b(w, h, Aa(readlane_val(vA, uu)), sel63, c1, nstep);
                sr_acc += 0x9E3779B9u;
                T::store_raw_b(wrs, wvoff, T::pack(w, lkey + sr_acc, RM), soff[d]);
                soff[d] = readlane_i32(offN, uu);
                wp[d] = T::load_raw_b(wrs, wvoff, soff[d]);
            };
            for (;;) {
                const int full = cntA & ~(kPfL - 1);
                for (int u0 = 0; u0 < full; u0 += kPfL) {
#pragma unroll
                    for (int d = 0; d < kPfL; ++d) rating(u0 + d, d);
                }
                cnt += cntA;
                if (cntA < kWave) break;  // the piece ends in this chunk
                // the piece goes on in the next chunk
                rA = rB;
                vA = vB;
                c0 += kWave;
                load_chunk(c0 + kWave, rB, vB);
                offN = ahead_offsets();
                cntA = __popcll(__ballot(rA < row_hi));
                if (cntA == 0) break;
            }
            {
                // the piece's last cntA mod kPfL ratings, after its chunk loop
                // (here the exits leave the piece, so no loop sees them)
                const int full = cntA & ~(kPfL - 1);
#pragma unroll
                for (int d = 0; d < kPfL; ++d) {
                    if (full + d >= cntA) break;
                    rating(full + d, d);
                }
            }
            // first rating past the piece: position cntA of chunk A
            const int stop_u = cntA;
            const int32_t stop_row = readlane_i32(rA, stop_u & (kWave - 1));
            tk_chunk += int64_t(clk() - t_c0);
            n_upd += cnt;
            cur += cnt;
            const int32_t next_row = cur < vlen ? stop_row : -1;
            if (next_row >= 0) {
                // window of the following 64 ratings for the waves that follow
                const int src = stop_u + lane;  // < 128
                const int32_t wr_a = __shfl(rA, src & (kWave - 1)), wr_b = __shfl(rB, src & (kWave - 1));
                const R wv_a = __shfl(vA, src & (kWave - 1)), wv_b = __shfl(vB, src & (kWave - 1));
                const int32_t myrow = src < kWave ? wr_a : wr_b;
                const R myval = src < kWave ? wv_a : wv_b;
                sh.wrow[s][lane] = myrow;
                sh.wval[s][lane] = myval;
                wb = cur;
                we = cur + min(kWave, vlen - cur);
                write_bounds<NW>(sh, s, myrow, we - wb, lane_bound);
            }
            ++n_vis;
            const uint64_t t_v1 = clk();
            tk_visit += int64_t(t_v1 - t_v0);
            if (next_row >= 0) {
                // more of this column lives in a later row block of this workgroup
                T::lds_put(sh.h[s], lane, h);
                if (lane == 0) {
                    sh.cur[s] = cur;
                    sh.wb[s] = wb;
                    sh.we[s] = we;
                }
                inbox_push<SLOTS>(sh, wave_of(lane_bound, next_row), s, timing);
            } else {
                // the column leaves this workgroup
                const int32_t np = sh.npos[s];
                const int32_t ncyc = cycle + (np == 0 ? 1 : 0);
                const bool park = np == 0 && ncyc >= a.target_cycle;
                const uint32_t rnd = sr_bits_k(a.sr_seed, 0x80000000u | uint32_t(g), sr_b, lkey);
                T::store_dc(H + int64_t(j) * ldf, lane, npairs, h, rnd, RM);
                if (lane == 0) {
                    tok_store(a.tok + j, sh.nstart[s], sh.nlen[s], np, ncyc, sh.ncs[s]);
                    st_dc(a.vcnt + sh.cs[s], sh.vc[s] + 1);  // the pair's visit is complete
                }
                if (park) {
                    wait_all_mem();
                    if (lane == 0) __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    const int32_t q = sh.nwg[s];
                    if (timing && lane == 0) st_dc(a.tok_stamp + j, uint64_t(wall_clock64()));
                    // the ring position is reserved while the stores drain
                    uint32_t t = 0;
                    if (lane == 0)
                        t = __hip_atomic_fetch_add(a.qtail + int64_t(q) * nomad_k::kQStride, 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                    wait_all_mem();
                    if (lane == 0) {
                        const uint64_t v = (uint64_t((t >> a.qlog2) + 1u) << 32) | uint32_t(j);
                        st_dc(a.qslot + (int64_t(q) << a.qlog2) + (t & gmask), v);
                    }
                }
                if (lane == 0)
                    __hip_atomic_fetch_or(&sh.freebits[s >> 5], 1u << (s & 31), __ATOMIC_RELEASE,
                                          __HIP_MEMORY_SCOPE_WORKGROUP);
                ++n_exit;
                tk_exit += int64_t(clk() - t_v1);
            }
            if ((n_work & 63) == 0) past_deadline();
            continue;
        }

        // ---- past the deadline: no claims; leave once every slot is free -----
        // (taken under the claim lock, so no claim that started before the
        // stop can allocate a slot after this wave found them all free)
        if (__builtin_amdgcn_readfirstlane(lds_ld(&sh.stop)) != 0u) {
            uint32_t lk = 1u, used = 1u;
            if (lane == 0) {
                lk = __hip_atomic_exchange(&sh.lock, 1u, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (lk == 0u) {
                    used = 0u;
#pragma unroll
                    for (int w = 0; w < SLOTS / 32; ++w) used |= ~lds_ld(&sh.freebits[w]);
                    __hip_atomic_store(&sh.lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            if (__builtin_amdgcn_readfirstlane(used) == 0u) break;
            ++n_idle;
            __builtin_amdgcn_s_sleep(2);
            continue;
        }

        // ---- idle: claim columns from this workgroup's ring ------------------
        uint32_t lk = 1u;
        if (lane == 0) lk = __hip_atomic_exchange(&sh.lock, 1u, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        lk = __builtin_amdgcn_readfirstlane(lk);
        int nb = 0;
        int32_t myslot = 0;
        uint64_t sv = 0;
        // (the stop flag is read again under the lock: a wave that stops the
        // workgroup takes the lock before it checks for free slots)
        if (lk == 0u && __builtin_amdgcn_readfirstlane(lds_ld(&sh.stop)) != 0u) {
            if (lane == 0) __hip_atomic_store(&sh.lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            lk = 1u;
        }
        if (lk == 0u) {
            // Shared LDS words are read with atomic loads and every value that
            // selects a slot is made wave-uniform: lane 0 allocates, the wave
            // reads its result back.  (Plain reads of words another lane just
            // changed are data races the compiler may resolve per lane.)
            const uint32_t gh = __builtin_amdgcn_readfirstlane(lds_ld(&sh.ghead));
            // the batch's entries in one round trip (the lock is held meanwhile)
            {
                bool valid = false;
                if (lane < kFetchBatch) {
                    sv = ld_dc(ring + ((gh + uint32_t(lane)) & gmask));
                    valid = uint32_t(sv >> 32) == (((gh + uint32_t(lane)) >> a.qlog2) + 1u);
                }
                const uint64_t vmask = __ballot(valid);
                nb = int(__builtin_ctzll(~vmask));  // lanes >= kFetchBatch are never valid
                if ((vmask >> nb) != 0) ++n_blocked;  // published entries behind an unpublished one
            }
            int nfree = 0;
#pragma unroll
            for (int w = 0; w < SLOTS / 32; ++w) nfree += __popc(lds_ld(&sh.freebits[w]));
            nfree = __builtin_amdgcn_readfirstlane(nfree);
            // at most slot_cap columns resident (the --ptoken cap)
            nb = min(nb, min(nfree - (SLOTS - slot_cap), kFetchBatch));
            nb = max(nb, 0);
            ++n_fetch;
            n_res += SLOTS - nfree;
            for (int t = 0; t < nb; ++t) {
                int s = 0;
                if (lane == 0) {
                    for (int w = 0; w < SLOTS / 32; ++w) {
                        const uint32_t bits = lds_ld(&sh.freebits[w]);
                        if (bits) {
                            s = w * 32 + __builtin_ctz(bits);
                            break;
                        }
                    }
                    __hip_atomic_fetch_and(&sh.freebits[s >> 5], ~(1u << (s & 31)), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                s = __builtin_amdgcn_readfirstlane(s);
                if (lane == t) myslot = s;
            }
            if (lane == 0) {
                __hip_atomic_store(&sh.ghead, gh + uint32_t(nb), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_store(&sh.lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (nb > 0) {
            const uint64_t t_f0 = clk();
            idle = 0;
            n_fetched += nb;
            // lane t < nb: metadata of claimed column t
            const int32_t jj = int32_t(uint32_t(sv));
            uint32_t tstart = 0;
            int32_t tlen = 1, tpos = 0, tcyc = 0, nwg = 0, nlen = 1, tcs = 0, tvc = 0, ncs = 0;
            Aa tstp = Aa(0);
            uint32_t nstart = 0;
            int32_t np = 0;
            // Everything that depends only on the column index is requested
            // first -- its route range and the h rows of the first four
            // columns -- so that it shares a round trip with the token
            // records; what those records select (next hop, visit count,
            // rating window) shares the next one, and the step is computed
            // from the count after the window loads are issued.
            P hv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                hv[u] = T::load_dc(H + int64_t(readlane_i32(jj, min(u, nb - 1))) * ldf, lane, npairs);
            if (lane < nb) {
                const int64_t vb = a.vptr[jj];
                const int32_t nv = int32_t(a.vptr[jj + 1] - vb);
                const uint64_t* q = reinterpret_cast<const uint64_t*>(a.tok + jj);
                const uint64_t w0 = ld_dc(q);
                const uint64_t w1 = ld_dc(q + 1);
                const uint64_t w2 = ld_dc(q + 2);
                tstart = uint32_t(w0);
                tlen = int32_t(w0 >> 32);
                tpos = int32_t(uint32_t(w1));
                tcyc = int32_t(w1 >> 32);
                tcs = int32_t(uint32_t(w2));
                np = tpos + 1 == nv ? 0 : tpos + 1;
                const RouteDesc nd = a.route[vb + np];
                nwg = nd.unit;
                nstart = nd.start;
                nlen = nd.len;
                ncs = nd.cslot;
                tvc = ld_dc(a.vcnt + tcs);
                if (timing) {
                    const uint64_t stamp = ld_dc(a.tok_stamp + jj);
                    if (stamp != 0) tk_transit_l += int64_t(wall_clock64() - stamp);
                }
            }
            for (int t0 = 0; t0 < nb; t0 += 4) {
                int32_t wr[4];
                R wv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = min(t0 + u, nb - 1);
                    const uint32_t st = uint32_t(readlane_i32(int32_t(tstart), t));
                    const int32_t ln = readlane_i32(tlen, t);
                    if (t0 > 0) hv[u] = T::load_dc(H + int64_t(readlane_i32(jj, t)) * ldf, lane, npairs);
                    const int32_t o = min(lane, ln - 1);
                    wr[u] = rowidx[int64_t(st) + o];
                    wv[u] = val[int64_t(st) + o];
                }
                if (t0 == 0) tstp = lane < nb ? Aa(step_of_count(a.lrate, a.drate, tvc)) : Aa(0);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = t0 + u;
                    if (t < nb) {
                        const int s = readlane_i32(myslot, t);
                        const int32_t ln = readlane_i32(tlen, t);
                        T::lds_put(sh.h[s], lane, hv[u]);
                        sh.wrow[s][lane] = wr[u];
                        sh.wval[s][lane] = wv[u];
                        write_bounds<NW>(sh, s, wr[u], min(kWave, ln), lane_bound);
                        if (lane == 0) {
                            sh.col[s] = readlane_i32(jj, t);
                            sh.start[s] = uint32_t(readlane_i32(int32_t(tstart), t));
                            sh.len[s] = ln;
                            sh.cur[s] = 0;
                            sh.wb[s] = 0;
                            sh.we[s] = min(kWave, ln);
                            sh.cyc[s] = readlane_i32(tcyc, t);
                            sh.cs[s] = readlane_i32(tcs, t);
                            sh.vc[s] = readlane_i32(tvc, t);
                            sh.stp[s] = readlane_val(tstp, t);
                            sh.npos[s] = readlane_i32(np, t);
                            sh.nwg[s] = readlane_i32(nwg, t);
                            sh.nstart[s] = uint32_t(readlane_i32(int32_t(nstart), t));
                            sh.nlen[s] = readlane_i32(nlen, t);
                            sh.ncs[s] = readlane_i32(ncs, t);
                        }
                        inbox_push<SLOTS>(sh, wave_of(lane_bound, readlane_i32(wr[u], 0)), s, timing);
                    }
                }
            }
            tk_fetch += int64_t(clk() - t_f0);
            continue;
        }
        ++n_idle;
        // the shared completion counter is read only now and then: every idle
        // wave polling one word would load a single memory channel
        if ((idle & 15) == 15 && int32_t(ld_dc(a.done)) >= a.ncols_active) break;
        if ((idle & 7) == 7) past_deadline();
        if (idle < 8)
            __builtin_amdgcn_s_sleep(2);
        else
            __builtin_amdgcn_s_sleep(20);
        ++idle;
    }

    // ---- exit: every slot is free here (a deadline drains the claimed hops;
    // otherwise every column is parked); h of any column still resident would
    // go back to HBM ----------------------------------------------------------
    __syncthreads();
    for (int s = wave; s < SLOTS; s += NW) {
        if (lds_ld(&sh.freebits[s >> 5]) & (1u << (s & 31))) continue;
        const int32_t j = sh.col[s];
        const P h = T::lds_get(sh.h[s], lane);
        const uint32_t rnd = sr_bits_k(a.sr_seed, 0x80000000u | uint32_t(g),
                                     uint32_t(j) * 0x01000193u + uint32_t(sh.cyc[s]), lkey);
        T::store_dc(H + int64_t(j) * ldf, lane, npairs, h, rnd, RM);
    }
    // lanes hold their own sums; fold them into lane 0
    for (int off = 32; off > 0; off >>= 1) {
        tk_transit_l += __shfl_down(tk_transit_l, off, 64);
        tk_inbox_l += __shfl_down(tk_inbox_l, off, 64);
    }
    if (lane == 0) {
        int64_t* st = a.stats + int64_t(g * NW + wave) * nomad_k::kStatWidth;  // per worker
        const int64_t v[16] = {n_upd, n_vis, n_idle, n_fetch, n_fetched, n_exit, n_res, n_blocked,
                               tk_visit, tk_exit, tk_fetch, tk_chunk, tk_transit_l, tk_inbox_l, n_depth, n_work};
        for (int i = 0; i < 16; ++i)
            if (v[i]) __hip_atomic_fetch_add(st + i, v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (hit_deadline) __hip_atomic_store(a.timed_out, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0) a.qhead[int64_t(g) * nomad_k::kQStride] = lds_ld(&sh.ghead);
}

// Appends per-workgroup column lists to the rings before a launch: ring g
// gets list[off[g] .. off[g+1]) in order.  One thread per ring; plain stores
// are published by the kernel boundary.
__global__ void seed_rings_kernel(SgdArgs a, const int32_t* list, const int64_t* off) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nblocks) return;
    const uint32_t mask = (1u << a.qlog2) - 1u;
    uint32_t t = a.qtail[int64_t(p) * nomad_k::kQStride];
    for (int64_t e = off[p]; e < off[p + 1]; ++e, ++t) {
        a.qslot[(int64_t(p) << a.qlog2) + (t & mask)] = (uint64_t((t >> a.qlog2) + 1u) << 32) | uint32_t(list[e]);
    }
    a.qtail[int64_t(p) * nomad_k::kQStride] = t;
}

// The rounding of bf16 stores is a template argument, so the per-rating
// store carries no branch on it; fp32 and fp64 ignore it.
// (storage | kWideBit: rows of more than 128 elements, two pairs per lane)
#define NOMAD_ASYNC_INST(NW, WB)                                                                          \
    template __global__ void sgd_async_kernel<NOMAD_STORE_FP32 | WB, NW, async_slots<NOMAD_STORE_FP32 | WB, NW>(), \
                                              NOMAD_ROUND_NEAREST>(SgdArgs);                                   \
    template __global__ void sgd_async_kernel<NOMAD_STORE_BF16 | WB, NW, async_slots<NOMAD_STORE_BF16 | WB, NW>(), \
                                              NOMAD_ROUND_NEAREST>(SgdArgs);                                   \
    template __global__ void sgd_async_kernel<NOMAD_STORE_BF16 | WB, NW, async_slots<NOMAD_STORE_BF16 | WB, NW>(), \
                                              NOMAD_ROUND_STOCHASTIC>(SgdArgs);                                \
    template __global__ void sgd_async_kernel<NOMAD_STORE_FP64 | WB, NW, async_slots<NOMAD_STORE_FP64 | WB, NW>(), \
                                              NOMAD_ROUND_NEAREST>(SgdArgs);
NOMAD_ASYNC_INST(4, 0)
NOMAD_ASYNC_INST(8, 0)
NOMAD_ASYNC_INST(16, 0)
NOMAD_ASYNC_INST(4, kWideBit)
NOMAD_ASYNC_INST(8, kWideBit)
NOMAD_ASYNC_INST(16, kWideBit)
