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
// The asynchronous token kernel is in async_kernel.hip.
#include <hip/hip_runtime.h>

#include <climits>

#include "sgd_common.hpp"

using namespace nomad_dev;
using nomad_k::SgdArgs;
using nomad_k::VisitDesc;


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

// Explicit instantiations launched from capi.hip.
template __global__ void sgd_strat_kernel<NOMAD_STORE_FP32>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_BF16>(SgdArgs);
template __global__ void sgd_strat_kernel<NOMAD_STORE_FP64>(SgdArgs);

