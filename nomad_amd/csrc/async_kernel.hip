// Asynchronous NOMAD token kernel for gfx950 (two-level).
//
// The reference runs T threads per machine; a column token visits every
// thread of its machine in the order of its permutation and then moves to
// another machine (Code/nomad/nb_main.cpp:491-565, nb_sendrecv.cpp:189-204).
// Here a workgroup of NW waves is the machine and each wave a thread:
//
//   * the workgroup owns NW * rpb consecutive rows of W, wave w the rpb rows
//     of its sub-block (the reference's per-part split, nb_data.cpp:83-89);
//   * columns arrive on the workgroup's ring in HBM (multi-producer, one
//     consumer position in LDS).  An idle wave claims a batch of them, copies
//     each h_j into an LDS slot (FP32, or FP64 for double storage) together
//     with the column's first 64 ratings of this row range, and hands the
//     slot to the wave owning the column's first row here;
//   * a wave pops a slot from its LDS inbox, runs the SGD steps of its piece
//     of the column (nb_main.cpp:520-542, rows ascending), writes h back to
//     the slot and hands it to the wave owning the column's next row;
//   * after the last piece h_j returns to HBM and the column is pushed to the
//     next workgroup of its route; after the last workgroup of the route the
//     column's cycle count (the step-size counter) advances, and a column
//     whose count reaches target_cycle is parked.
//
// Cross-workgroup data (h_j, the column's TokInfo) moves with device-coherent
// (sc1) loads and stores; a producer waits for all its memory operations to
// complete before it publishes the ring slot, so no L2 writeback or
// invalidate is needed.  W rows are private to their wave.
#include <hip/hip_runtime.h>

#include "sgd_common.hpp"

using namespace nomad_dev;
using nomad_k::TokInfo;

namespace {

constexpr int kFetchBatch = 16;  // columns claimed from the ring per fetch

template <int S, int NW, int SLOTS>
struct AsyncShared {
    using A = typename Store<S>::acc;
    using R = typename Store<S>::rval;
    A h[SLOTS][2 * kWave];       // column factors of resident tokens
    int32_t wrow[SLOTS][kWave];  // window of the visit's ratings: rows
    R wval[SLOTS][kWave];        //                                values
    int32_t col[SLOTS];
    uint32_t start[SLOTS];       // CSC offset of the visit in this row range
    int32_t len[SLOTS];
    int32_t cur[SLOTS];          // ratings of the visit already processed
    int32_t wb[SLOTS];           // the window holds visit offsets [wb, we)
    int32_t we[SLOTS];
    int32_t cyc[SLOTS];          // completed route cycles of the column
    int32_t npos[SLOTS];         // route position of the column's next visit
    int32_t nwg[SLOTS];          //   its workgroup
    uint32_t nstart[SLOTS];      //   its CSC range
    int32_t nlen[SLOTS];
    uint32_t inbox[NW][SLOTS];   // per wave: (lap << 16) | slot
    uint32_t itail[NW];
    uint32_t freebits[SLOTS / 32];
    uint32_t lock;               // serialises ring claims and slot allocation
    uint32_t ghead;              // consumer position of this workgroup's ring
};

__device__ __forceinline__ uint32_t lds_ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Hand slot s to wave w of this workgroup.  Called by a whole wave; the
// workgroup-scope release orders this wave's earlier LDS writes of the slot.
template <int SLOTS, class L>
__device__ __forceinline__ void inbox_push(L& sh, int w, int s) {
    if (lane_id() == 0) {
        const uint32_t t =
            __hip_atomic_fetch_add(&sh.itail[w], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t lap = ((t / uint32_t(SLOTS)) + 1u) & 0xFFFFu;
        __hip_atomic_store(&sh.inbox[w][t & uint32_t(SLOTS - 1)], (lap << 16) | uint32_t(s), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

__device__ __forceinline__ void tok_store(TokInfo* p, uint32_t start, int32_t len, int32_t pos, int32_t cycle) {
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
    st_dc(q, (uint64_t(uint32_t(len)) << 32) | start);
    st_dc(q + 1, (uint64_t(uint32_t(cycle)) << 32) | uint32_t(pos));
}

}  // namespace

template <int S, int NW, int SLOTS>
__global__ __launch_bounds__(NW * 64) void sgd_async_kernel(SgdArgs a) {
    using T = Store<S>;
    using E = typename T::elem;
    using Aa = typename T::acc;
    using R = typename T::rval;
    static_assert(SLOTS % 32 == 0 && (SLOTS & (SLOTS - 1)) == 0, "SLOTS must be a power of two >= 32");
    __shared__ AsyncShared<S, NW, SLOTS> sh;
    const int g = blockIdx.x;
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    const int npairs = a.ldf >> 1;
    const int32_t wg_row0 = g * NW * a.rpb;
    const int32_t row_hi = min(a.m, wg_row0 + (wave + 1) * a.rpb);
    const uint32_t gmask = (1u << a.qlog2) - 1u;
    const uint64_t* ring = a.qslot + (int64_t(g) << a.qlog2);
    E* __restrict__ H = static_cast<E*>(a.H);
    const int32_t* __restrict__ rowidx = a.rowidx;
    const R* __restrict__ val = static_cast<const R*>(a.val);
    const Aa reg = Aa(a.reg);

    for (int i = threadIdx.x; i < NW * SLOTS; i += blockDim.x) (&sh.inbox[0][0])[i] = 0u;
    if (threadIdx.x < NW) sh.itail[threadIdx.x] = 0u;
    if (threadIdx.x < SLOTS / 32) sh.freebits[threadIdx.x] = 0xFFFFFFFFu;
    if (threadIdx.x == 0) {
        sh.lock = 0u;
        sh.ghead = a.qhead[g];
    }
    __syncthreads();

    const uint64_t t_end = wall_clock64() + a.deadline;
    uint32_t ihead = 0;
    int64_t n_upd = 0, n_vis = 0, n_idle = 0, n_fetch = 0, n_fetched = 0, n_exit = 0, n_res = 0, n_blocked = 0;
    int64_t tk_visit = 0, tk_exit = 0, tk_fetch = 0, tk_chunk = 0;
    int64_t tk_transit_l = 0;  // per lane
    bool hit_deadline = false;
    int idle = 0;

    for (;;) {
        // ---- a slot waiting in this wave's inbox -------------------------------
        const uint32_t e = __builtin_amdgcn_readfirstlane(lds_ld(&sh.inbox[wave][ihead & (SLOTS - 1)]));
        if ((e >> 16) == (((ihead / SLOTS) + 1u) & 0xFFFFu)) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const uint64_t t_v0 = wall_clock64();
            ++ihead;
            idle = 0;
            const int s = int(e & 0xFFFFu);
            const int32_t j = sh.col[s];
            const uint32_t vstart = sh.start[s];
            const int32_t vlen = sh.len[s];
            int32_t cur = sh.cur[s];
            int32_t wb = sh.wb[s], we = sh.we[s];
            const int32_t cycle = sh.cyc[s];
            const Aa step = step_for<Aa>(a, cycle);
            const uint32_t sr_b = uint32_t(j) * 0x01000193u + uint32_t(cycle);
            typename T::pair h;
            h.x = sh.h[s][2 * lane];
            h.y = sh.h[s][2 * lane + 1];
            int32_t next_row = -1;
            for (;;) {  // this wave's piece of the column, up to 64 ratings at a time
                const int32_t left = vlen - cur;
                const int32_t idx = cur + lane;
                const bool inwin = idx >= wb && idx < we;
                int32_t myrow = INT_MAX;
                R myval = R(0);
                if (lane < left) {
                    if (inwin) {
                        myrow = sh.wrow[s][idx - wb];
                        myval = sh.wval[s][idx - wb];
                    } else {
                        myrow = rowidx[int64_t(vstart) + idx];
                        myval = val[int64_t(vstart) + idx];
                    }
                }
                const bool from_hbm = __ballot(lane < left && !inwin) != 0;
                const int32_t cnt = __popcll(__ballot(myrow < row_hi));
                const uint64_t t_c0 = wall_clock64();
                if (cnt > 0) run_chunk<S>(a, h, myrow, myval, cnt, step, reg, sr_b);
                tk_chunk += int64_t(wall_clock64() - t_c0);
                if (from_hbm) {  // keep this chunk for the waves that follow
                    if (lane < left) {
                        sh.wrow[s][lane] = myrow;
                        sh.wval[s][lane] = myval;
                    }
                    wb = cur;
                    we = cur + min(kWave, left);
                    // later reads of the window by other lanes of this wave see these writes
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
                n_upd += cnt;
                cur += cnt;
                if (cnt < kWave || cur >= vlen) {
                    if (cur < vlen) next_row = readlane_i32(myrow, cnt);
                    break;
                }
            }
            ++n_vis;
            const uint64_t t_v1 = wall_clock64();
            tk_visit += int64_t(t_v1 - t_v0);
            if (next_row >= 0) {
                // more of this column lives in a later sub-block of this workgroup
                sh.h[s][2 * lane] = h.x;
                sh.h[s][2 * lane + 1] = h.y;
                if (lane == 0) {
                    sh.cur[s] = cur;
                    sh.wb[s] = wb;
                    sh.we[s] = we;
                }
                inbox_push<SLOTS>(sh, (next_row - wg_row0) / a.rpb, s);
            } else {
                // the column leaves this workgroup
                const int32_t np = sh.npos[s];
                const int32_t ncyc = cycle + (np == 0 ? 1 : 0);
                const bool park = np == 0 && ncyc >= a.target_cycle;
                const uint32_t rnd = sr_bits(a.sr_seed, 0x80000000u | uint32_t(g), sr_b, uint32_t(lane));
                T::store_dc(H + int64_t(j) * a.ldf, lane, npairs, h, rnd, a.round_mode);
                if (lane == 0) tok_store(a.tok + j, sh.nstart[s], sh.nlen[s], np, ncyc);
                if (park) {
                    wait_all_mem();
                    if (lane == 0) __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    const int32_t q = sh.nwg[s];
                    // h_j and TokInfo complete before a ring position is even
                    // reserved, so the consumer never waits long on a reserved
                    // but unpublished entry
                    if (lane == 0) st_dc(a.tok_stamp + j, uint64_t(wall_clock64()));
                    wait_all_mem();
                    if (lane == 0) {
                        const uint32_t t =
                            __hip_atomic_fetch_add(a.qtail + q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        const uint64_t v = (uint64_t((t >> a.qlog2) + 1u) << 32) | uint32_t(j);
                        st_dc(a.qslot + (int64_t(q) << a.qlog2) + (t & gmask), v);
                    }
                }
                if (lane == 0)
                    __hip_atomic_fetch_or(&sh.freebits[s >> 5], 1u << (s & 31), __ATOMIC_RELEASE,
                                          __HIP_MEMORY_SCOPE_WORKGROUP);
                ++n_exit;
                tk_exit += int64_t(wall_clock64() - t_v1);
            }
            if (wall_clock64() > t_end) {
                hit_deadline = true;
                break;
            }
            continue;
        }

        // ---- idle: claim columns from this workgroup's ring ------------------
        uint32_t lk = 1u;
        if (lane == 0) lk = __hip_atomic_exchange(&sh.lock, 1u, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        lk = __builtin_amdgcn_readfirstlane(lk);
        int nb = 0;
        int32_t myslot = 0;
        uint64_t sv = 0;
        if (lk == 0u) {
            // Shared LDS words are read with atomic loads and every value that
            // selects a slot is made wave-uniform: lane 0 allocates, the wave
            // reads its result back.  (Plain reads of words another lane just
            // changed are data races the compiler may resolve per lane.)
            const uint32_t gh = __builtin_amdgcn_readfirstlane(lds_ld(&sh.ghead));
            sv = ld_dc(ring + ((gh + uint32_t(lane)) & gmask));
            const bool valid = uint32_t(sv >> 32) == (((gh + uint32_t(lane)) >> a.qlog2) + 1u);
            const uint64_t vmask = __ballot(valid);
            nb = vmask == ~0ull ? kWave : int(__builtin_ctzll(~vmask));
            if (nb < kWave && (vmask >> nb) != 0) ++n_blocked;  // published entries behind an unpublished one
            int nfree = 0;
#pragma unroll
            for (int w = 0; w < SLOTS / 32; ++w) nfree += __popc(lds_ld(&sh.freebits[w]));
            nfree = __builtin_amdgcn_readfirstlane(nfree);
            nb = min(nb, min(nfree, kFetchBatch));
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
            const uint64_t t_f0 = wall_clock64();
            idle = 0;
            n_fetched += nb;
            // lane t < nb: metadata of claimed column t
            const int32_t jj = int32_t(uint32_t(sv));
            uint32_t tstart = 0;
            int32_t tlen = 1, tpos = 0, tcyc = 0, nwg = 0, nlen = 1;
            uint32_t nstart = 0;
            int32_t np = 0;
            if (lane < nb) {
                const uint64_t* q = reinterpret_cast<const uint64_t*>(a.tok + jj);
                const uint64_t w0 = ld_dc(q);
                const uint64_t w1 = ld_dc(q + 1);
                tstart = uint32_t(w0);
                tlen = int32_t(w0 >> 32);
                tpos = int32_t(uint32_t(w1));
                tcyc = int32_t(w1 >> 32);
                const int64_t vb = a.vptr[jj];
                const int32_t nv = int32_t(a.vptr[jj + 1] - vb);
                np = tpos + 1 == nv ? 0 : tpos + 1;
                const VisitDesc nd = a.visits[vb + np];
                nwg = nd.worker;
                nstart = uint32_t(nd.start);
                nlen = nd.len;
                const uint64_t stamp = ld_dc(a.tok_stamp + jj);
                if (stamp != 0) tk_transit_l += int64_t(wall_clock64() - stamp);
            }
            for (int t0 = 0; t0 < nb; t0 += 4) {
                typename T::pair hv[4];
                int32_t wr[4];
                R wv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = min(t0 + u, nb - 1);
                    const int32_t j = readlane_i32(jj, t);
                    const uint32_t st = uint32_t(readlane_i32(int32_t(tstart), t));
                    const int32_t ln = readlane_i32(tlen, t);
                    hv[u] = T::load_dc(H + int64_t(j) * a.ldf, lane, npairs);
                    const int32_t o = min(lane, ln - 1);
                    wr[u] = rowidx[int64_t(st) + o];
                    wv[u] = val[int64_t(st) + o];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = t0 + u;
                    if (t < nb) {
                        const int s = readlane_i32(myslot, t);
                        const int32_t ln = readlane_i32(tlen, t);
                        sh.h[s][2 * lane] = hv[u].x;
                        sh.h[s][2 * lane + 1] = hv[u].y;
                        sh.wrow[s][lane] = wr[u];
                        sh.wval[s][lane] = wv[u];
                        if (lane == 0) {
                            sh.col[s] = readlane_i32(jj, t);
                            sh.start[s] = uint32_t(readlane_i32(int32_t(tstart), t));
                            sh.len[s] = ln;
                            sh.cur[s] = 0;
                            sh.wb[s] = 0;
                            sh.we[s] = min(kWave, ln);
                            sh.cyc[s] = readlane_i32(tcyc, t);
                            sh.npos[s] = readlane_i32(np, t);
                            sh.nwg[s] = readlane_i32(nwg, t);
                            sh.nstart[s] = uint32_t(readlane_i32(int32_t(nstart), t));
                            sh.nlen[s] = readlane_i32(nlen, t);
                        }
                        inbox_push<SLOTS>(sh, (readlane_i32(wr[u], 0) - wg_row0) / a.rpb, s);
                    }
                }
            }
            tk_fetch += int64_t(wall_clock64() - t_f0);
            continue;
        }
        ++n_idle;
        if (int32_t(ld_dc(a.done)) >= a.ncols_active) break;
        if (wall_clock64() > t_end) {
            hit_deadline = true;
            break;
        }
        if (idle < 8)
            __builtin_amdgcn_s_sleep(2);
        else
            __builtin_amdgcn_s_sleep(20);
        ++idle;
    }

    // ---- exit: columns still resident in LDS go back to HBM ---------------------
    __syncthreads();
    for (int s = wave; s < SLOTS; s += NW) {
        if (lds_ld(&sh.freebits[s >> 5]) & (1u << (s & 31))) continue;
        const int32_t j = sh.col[s];
        typename T::pair h;
        h.x = sh.h[s][2 * lane];
        h.y = sh.h[s][2 * lane + 1];
        const uint32_t rnd = sr_bits(a.sr_seed, 0x80000000u | uint32_t(g),
                                     uint32_t(j) * 0x01000193u + uint32_t(sh.cyc[s]), uint32_t(lane));
        T::store_dc(H + int64_t(j) * a.ldf, lane, npairs, h, rnd, a.round_mode);
    }
    // lanes hold their own transit sums; fold them into lane 0
    for (int off = 32; off > 0; off >>= 1) tk_transit_l += __shfl_down(tk_transit_l, off, 64);
    if (lane == 0) {
        int64_t* st = a.stats + int64_t(g) * nomad_k::kStatWidth;
        const int64_t v[13] = {n_upd, n_vis, n_idle, n_fetch, n_fetched, n_exit, n_res, n_blocked,
                               tk_visit, tk_exit, tk_fetch, tk_chunk, tk_transit_l};
        for (int i = 0; i < 13; ++i)
            if (v[i]) __hip_atomic_fetch_add(st + i, v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (hit_deadline) __hip_atomic_store(a.timed_out, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0) a.qhead[g] = lds_ld(&sh.ghead);
}

// Appends per-workgroup column lists to the rings before a launch: ring g
// gets list[off[g] .. off[g+1]) in order.  One thread per ring; plain stores
// are published by the kernel boundary.
__global__ void seed_rings_kernel(SgdArgs a, const int32_t* list, const int64_t* off) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nblocks) return;
    const uint32_t mask = (1u << a.qlog2) - 1u;
    uint32_t t = a.qtail[p];
    for (int64_t e = off[p]; e < off[p + 1]; ++e, ++t) {
        a.qslot[(int64_t(p) << a.qlog2) + (t & mask)] = (uint64_t((t >> a.qlog2) + 1u) << 32) | uint32_t(list[e]);
    }
    a.qtail[p] = t;
}

template __global__ void sgd_async_kernel<NOMAD_STORE_FP32, 16, 128>(SgdArgs);
template __global__ void sgd_async_kernel<NOMAD_STORE_BF16, 16, 128>(SgdArgs);
template __global__ void sgd_async_kernel<NOMAD_STORE_FP64, 16, 64>(SgdArgs);
