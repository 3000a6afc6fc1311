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

template <class A>
__device__ __forceinline__ A step_for(const SgdArgs& a, int32_t count) {
    if (count < a.nsteps) return gbl(static_cast<const A*>(a.steps))[count];
    // past the table: same formula as the host table (nb_main.cpp:514-515)
    const double s = a.lrate * 1.5 / (1.0 + a.drate * pow(double(count) + 1.0, 1.5));
    return A(s);
}

// Step size of visit count c (nb_main.cpp:514-515) computed on the device, so
// that a count read from memory needs no further table lookup:
// lrate * 1.5 / (1 + drate * (c + 1)^1.5) in double.  (c + 1)^1.5 is x * sqrt(x)
// with the rounding errors of the square root and of the product added back
// (both are exact fma residuals), which gives the correctly rounded power.
// Rounded to float it equals the host's step table (std::pow) for every count
// below 2e7 at the learning and decay rates tried; in double it differs from
// glibc's pow, which is not always correctly rounded, in the last bit of
// about one count in a thousand.
__device__ __forceinline__ double step_of_count(double lrate, double drate, int32_t c) {
    const double x = double(c) + 1.0;
    const double s = __builtin_sqrt(x);
    const double r = __builtin_fma(-s, s, x);  // x - s^2, exact
    const double p = x * s;
    const double e = __builtin_fma(x, s, -p);  // x s - p, exact
    const double x15 = p + (e + x * (r / (2.0 * s)));
    return lrate * 1.5 / (1.0 + drate * x15);
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

// W rows of one row block [rlo, rlo + rows) through one buffer resource:
// lane l's pair of row r sits at byte offset wvoff + (r - rlo) * rowbytes,
// wvoff = l * pair bytes (or 2^31 for lanes past the row, so that they are
// out of range; two such offsets in the two-pair layout), and drop() puts
// every lane out of range (a store the range check drops).  The row part
// (r - rlo) * rowbytes goes to the scalar offset of a load or store (it is
// always a row of the block; the hardware range check covers the per-lane
// offset only).
template <int S>
struct RowBlock {
    using voff_t = typename Store<S>::voff_t;
    __amdgpu_buffer_rsrc_t rsrc;
    voff_t wvoff;
    int32_t rlo;
    int32_t rowbytes;
    __device__ __forceinline__ int soff(int32_t r) const { return (r - rlo) * rowbytes; }
    __device__ __forceinline__ voff_t drop(int d = 0) const { return Store<S>::drop_voff(wvoff, d); }
};

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(int(uint32_t(v)));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(int(uint32_t(v >> 32)));
    return (uint64_t(hi) << 32) | lo;
}

// (the descriptor's fields are read from the first lane: they are uniform,
// but where the arguments were loaded from memory into vector registers the
// descriptor would otherwise be built in vector registers too, and the
// compiler then wraps every buffer access in a loop over its distinct
// values -- twelve instructions and an exec-mask round trip per access)
template <int S>
__device__ __forceinline__ RowBlock<S> row_block(const SgdArgs& a, int32_t rlo, int32_t rhi, int lane) {
    using T = Store<S>;
    using E = typename T::elem;
    RowBlock<S> rb;
    rlo = __builtin_amdgcn_readfirstlane(rlo);
    rb.rowbytes = __builtin_amdgcn_readfirstlane(a.ldf * int(sizeof(E)));
    const int32_t bytes = __builtin_amdgcn_readfirstlane((rhi - rlo) * rb.rowbytes);
    const uint64_t base = uniform64(reinterpret_cast<uint64_t>(static_cast<E*>(a.W)) + uint64_t(rlo) * rb.rowbytes);
    rb.rsrc = __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(base), 0, bytes, 0x00020000);
    rb.wvoff = T::lane_voff(lane, a.ldf >> 1);
    rb.rlo = rlo;
    return rb;
}

// rows of W in flight inside a visit of the stratified kernels; the batched
// kernel's device-coherent loads miss the L2, so it keeps fewer in flight
// (past a visit's end the prefetch slots re-read its last row)
template <int S, bool DC>
constexpr int strat_depth() {
    return (base_storage(S) == NOMAD_STORE_FP64 || DC) ? 8 / Store<S>::kNP : 16 / Store<S>::kNP;
}

// SGD steps for the first `cnt` (<= 64) ratings held one per lane in
// (myrow, myval), in lane order, for the column whose factor pair h this lane
// holds; the rows lie in row block rb.  strat_depth rows of W are in flight.
// The loop body has no branch around a memory operation (slots past cnt are
// idle: their store is dropped by the range check and h keeps its value) and
// the prologue issues a dropped store before each load, so the compiler's
// waits are exactly the prefetch distance.  A column's rows are distinct, so
// no prefetched row is updated before its rating uses it.
template <int S, bool DC, int RM>
__device__ __forceinline__ void run_chunk(const SgdArgs& a, const RowBlock<S>& rb, typename Store<S>::pair& h,
                                          int32_t myrow, typename Store<S>::rval myval, int32_t cnt,
                                          typename Store<S>::acc step, typename Store<S>::acc reg, uint32_t sr_b,
                                          uint32_t lkey) {
    using T = Store<S>;
    using Aa = typename T::acc;
    using P = typename T::pair;
    constexpr int kD = strat_depth<S, DC>();
    constexpr int CP = DC ? kCpDevice : 0;
    typename T::raw wp[kD];
#pragma unroll
    for (int d = 0; d < kD; ++d) {
        T::template store_raw_b<CP>(rb.rsrc, rb.drop(d), T::pack(T::zero(), 0u, 0));
        wp[d] = T::template load_raw_b<CP>(rb.rsrc, rb.wvoff, rb.soff(readlane_i32(myrow, min(d, cnt - 1))));
    }
    // groups of kG ratings; a group that starts past the chunk leaves both
    // loops at once (an exit edge keeps the loop's wait counts exact)
    constexpr int kG = 4;
    for (int32_t t0 = 0;; t0 += kD) {
#pragma unroll
        for (int gq = 0; gq < kD / kG; ++gq) {
            if (t0 + gq * kG >= cnt) goto chunk_done;
#pragma unroll
            for (int q = 0; q < kG; ++q) {
                const int d = gq * kG + q;
                const int32_t t = t0 + d;
                const bool real = t < cnt;
                const int32_t uc = real ? t : cnt - 1;
                const int32_t i = readlane_i32(myrow, uc);
                P w = T::unpack(wp[d]);
                const P h0 = h;
                sgd_step(w, h, Aa(readlane_val(myval, uc)), step, reg);
                h = T::select(real, h, h0);
                const uint32_t rnd = sr_bits_k(a.sr_seed, uint32_t(i), sr_b, lkey);
                T::template store_raw_b<CP>(rb.rsrc, real ? rb.wvoff : rb.drop(), T::pack(w, rnd, RM),
                                            real ? rb.soff(i) : 0);
                wp[d] = T::template load_raw_b<CP>(rb.rsrc, rb.wvoff,
                                                   rb.soff(readlane_i32(myrow, min(t + kD, cnt - 1))));
            }
        }
    }
chunk_done:;
}

// The ratings [start, start + len) of column j in row block rb, with the
// first chunk's rows and values already loaded (one per lane, lanes past the
// visit's length holding the last rating's), so that the caller can fetch
// them ahead, during the previous visit.
template <int S, bool DC, int RM>
__device__ __forceinline__ void run_visit_pre(const SgdArgs& a, const RowBlock<S>& rb, typename Store<S>::pair& h,
                                              int32_t j, int64_t start, int32_t len, int32_t row0,
                                              typename Store<S>::rval val0, typename Store<S>::acc step,
                                              typename Store<S>::acc reg, uint32_t count, uint32_t lkey) {
    using R = typename Store<S>::rval;
    const int lane = lane_id();
    const gptr<int32_t> ridx = gbl(a.rowidx) + start;
    const gptr<R> rv = gbl(static_cast<const R*>(a.val)) + start;
    const uint32_t sr_b = uint32_t(j) * 0x01000193u + count;
    run_chunk<S, DC, RM>(a, rb, h, row0, val0, min(kWave, len), step, reg, sr_b, lkey);
    for (int32_t base = kWave; base < len; base += kWave) {
        const int32_t cnt = min(kWave, len - base);
        const int32_t myrow = ridx[base + min(lane, cnt - 1)];
        const R myval = rv[base + min(lane, cnt - 1)];
        run_chunk<S, DC, RM>(a, rb, h, myrow, myval, cnt, step, reg, sr_b, lkey);
    }
}

// ---- a row block of W held in LDS for a whole tile -----------------------------
// Row r of the block is at wl[(r - rlo) * npairs + lane] as the stored pair
// (the packed bf16 word, or the fp32 / fp64 pair): what the row's HBM words
// would hold, so a rating reads exactly the value the previous update of the
// row stored.  Lanes past the row read zero and store nothing, as the
// buffer accesses' range check does.
constexpr int kTileLdsBytes = 8192;  // per wave
constexpr int kStageRows = 256;      // most rows of W a tile stages in LDS

// run_chunk with the rows in LDS, rating t's row at slot myslot[t]: the same
// steps in the same order (a column's rows are distinct, so a row loaded
// ahead is never stale), the stored forms written back to LDS.  Lanes past
// the row (lane >= npairs) read and write the zero entry wl[zslot] instead
// of a branch around the access: they hold h = 0, so they read 0 and write
// back a zero (of either sign), and the loop body stays one basic block
// (the scalar rounding hash then fills the reduction's wait slots).
template <int S, int RM>
__device__ __forceinline__ void run_chunk_lds(const SgdArgs& a, typename Store<S>::raw* wl, int npairs,
                                              int32_t zslot, typename Store<S>::pair& h, int32_t myrow,
                                              int32_t myslot, typename Store<S>::rval myval, int32_t cnt,
                                              typename Store<S>::acc step, typename Store<S>::acc reg,
                                              uint32_t sr_b, uint32_t lkey) {
    using T = Store<S>;
    using Aa = typename T::acc;
    using P = typename T::pair;
    constexpr int kD = 4;
    const int lane = lane_id();
    const bool inr = lane < npairs;
    auto at = [&](int32_t slot) { return inr ? slot * npairs + lane : zslot; };
    typename T::raw wp[kD];
#pragma unroll
    for (int d = 0; d < kD; ++d) wp[d] = wl[at(readlane_i32(myslot, min(d, cnt - 1)))];
    for (int32_t t0 = 0;; t0 += kD) {
#pragma unroll
        for (int d = 0; d < kD; ++d) {
            const int32_t t = t0 + d;
            if (t >= cnt) goto chunk_done;
            const int32_t i = readlane_i32(myrow, t);
            P w = T::unpack(wp[d]);
            sgd_step(w, h, Aa(readlane_val(myval, t)), step, reg);
            const uint32_t rnd = sr_bits_k(a.sr_seed, uint32_t(i), sr_b, lkey);
            wl[at(readlane_i32(myslot, t))] = T::pack(w, rnd, RM);
            wp[d] = wl[at(readlane_i32(myslot, min(t + kD, cnt - 1)))];
        }
    }
chunk_done:;
}

template <int S, int RM>
__device__ __forceinline__ void run_visit_lds(const SgdArgs& a, typename Store<S>::raw* wl, int npairs,
                                              int32_t zslot, typename Store<S>::pair& h, int32_t j, int64_t start, int32_t len,
                                              int32_t row0, int32_t slot0, typename Store<S>::rval val0,
                                              typename Store<S>::acc step, typename Store<S>::acc reg,
                                              uint32_t count, uint32_t lkey) {
    using R = typename Store<S>::rval;
    const int lane = lane_id();
    const gptr<int32_t> ridx = gbl(a.rowidx) + start;
    const gptr<R> rv = gbl(static_cast<const R*>(a.val)) + start;
    const gptr<uint16_t> sl = gbl(a.tslot) + start;
    const uint32_t sr_b = uint32_t(j) * 0x01000193u + count;
    run_chunk_lds<S, RM>(a, wl, npairs, zslot, h, row0, slot0, val0, min(kWave, len), step, reg, sr_b, lkey);
    for (int32_t base = kWave; base < len; base += kWave) {
        const int32_t cnt = min(kWave, len - base);
        const int32_t myrow = ridx[base + min(lane, cnt - 1)];
        const int32_t myslot = sl[base + min(lane, cnt - 1)];
        const R myval = rv[base + min(lane, cnt - 1)];
        // (the chunk loop is shared with the first chunk, whose ratings come
        // from LDS: waiting here keeps its waits off the first chunk)
        wait_vm();
        run_chunk_lds<S, RM>(a, wl, npairs, zslot, h, myrow, myslot, myval, cnt, step, reg, sr_b, lkey);
    }
}

}  // namespace nomad_dev
