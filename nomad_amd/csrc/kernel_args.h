// Kernel argument blocks shared by the host side (capi.hip) and the kernels.
#pragma once

#include <stdint.h>

namespace nomad_k {

// One non-empty (column, worker) visit: the worker's ratings of the column are
// the contiguous CSC range [start, start + len) (rows ascending).  In the
// stratified schedule's tile lists `worker` holds the column instead.
struct VisitDesc {
    int64_t start;
    int32_t len;
    int32_t worker;
};

// One hop of an async column route: the column's ratings in workgroup
// `unit`'s rows are the CSC range [start, start + len); `cslot` indexes the
// (column, workgroup) pair's visit counter.  A column's hops are listed in
// route order; the counter index stays with the pair when the host
// re-permutes the route.
struct RouteDesc {
    uint32_t start;
    int32_t len;
    int32_t unit;
    int32_t cslot;
};

// Per-column token state kept in HBM between holders: the CSC range of the
// column's next hop, its position in the column's route, the number of
// completed route cycles (for parking after a number of epochs) and the
// hop's visit-counter index.  Written with device-coherent stores before the
// column is published on a ring.
struct alignas(8) TokInfo {
    uint32_t start;
    int32_t len;
    int32_t pos;
    int32_t cycle;
    int32_t cslot;
    int32_t pad;
};

// uint32 words between consecutive workgroups' ring counters (qtail, qhead):
// each counter sits in its own 128-byte line
constexpr int kQStride = 32;

// LDS token slots per workgroup of the async kernel (storage S = NOMAD_STORE_*
// 0 fp32, 1 bf16, 2 fp64; NW waves): 128 for fp32 and bf16 at 8 waves, where
// a 4-deep stream prefetch leaves room for them in the 160 KiB, 32 for fp64,
// 64 otherwise.  More resident columns keep a workgroup's waves busy: at 8
// waves bf16 ran 19 % and fp32 17 % faster with 128 slots than with 64.
template <int S, int NW>
constexpr int async_slots() {
    // (S & 16: rows of more than 128 elements, whose factors take twice the
    // room per slot)
    return (S & 16) ? ((S & 15) == 2 ? 32 : NW == 8 ? 64 : 32) : (S == 2 ? 32 : NW == 8 ? 128 : 64);
}

// per-workgroup counters of the async kernel
enum AsyncStat : int { kStatUpdates = 0, kStatVisits, kStatIdle, kStatFetches, kStatFetched, kStatExits,
                       kStatResident, kStatBlocked, kStatTicksVisit, kStatTicksExit, kStatTicksFetch,
                       kStatTicksChunk, kStatTicksTransit, kStatTicksInbox, kStatDepth, kStatWork, kStatWidth = 16 };

struct SgdArgs {
    // problem
    int32_t m, n, k, ldf;
    int32_t nblocks;      // workers (async) or strata B (strat)
    int32_t rpb;          // rows per row block
    int32_t cpb;          // columns per column block (strat)
    int32_t round_mode;   // NOMAD_ROUND_*
    uint32_t sr_seed;
    int32_t nsteps;       // entries in the step table
    void* W;              // m x ldf, storage type
    void* H;              // n x ldf, storage type
    const int64_t* colptr;
    const int32_t* rowidx;  // (strat: the ratings in tile order, indexed by visits[].start)
    const void* val;      // float (fp32/bf16 storage) or double (fp64)
    const void* steps;    // step table indexed by visit count: float or double
    double lrate, drate;  // for counts past the table
    double reg;

    // async schedule (two-level: workgroups of waves)
    const int32_t* wbound;     // nblocks * waves + 1 first rows of the worker row blocks
                               // (strat: nblocks + 1 first rows of the strata's row blocks)
    const int64_t* vptr;       // n + 1 offsets into route
    const RouteDesc* route;    // per column, its hops in route order
    int32_t* vcnt;             // per (column, workgroup) pair: visits so far (the step-size counter)
    TokInfo* tok;              // per column: its next hop and completed cycles
    uint64_t* tok_stamp;       // per column: wall clock when it was last published (diagnostics)
    uint64_t* qslot;           // nblocks x (1 << qlog2) ring slots: (lap+1) << 32 | column
    uint32_t* qtail;           // per workgroup (stride kQStride), producers' reservation counter
    uint32_t* qhead;           // per workgroup (stride kQStride), consumer position (kept across launches)
    uint32_t* done;            // columns parked after reaching target_cycle
    int64_t* stats;            // per workgroup: kStatWidth counters
    uint32_t* timed_out;       // set when a workgroup stopped on the deadline
    int32_t qlog2;
    int32_t target_cycle;      // park a column when its cycle count reaches this
    int32_t ncols_active;      // columns circulating in this launch
    int32_t timing;            // nonzero: collect the per-phase wall-clock counters
    int32_t slot_cap;          // > 0: at most this many columns resident per workgroup
    uint64_t deadline;         // wall_clock64 ticks after its start at which a worker stops

    // strat schedule
    const VisitDesc* visits;   // strat tile lists; .worker = column, .start in tile order
    const int64_t* tile_ptr;   // B * B + 1: first visit of tile (b, cb) at [b * B + cb] in visits
    const int64_t* trow_ptr;   // B * B + 1: the tile's distinct rows at trows[trow_ptr[t] .. trow_ptr[t + 1])
    const int32_t* trows;      //   ascending
    const uint16_t* tslot;     // per rating (tile order): its row's rank among its tile's distinct rows
    int32_t substep;
    int32_t epoch;             // strat: the epoch; async: the launch index
};

// Batched stratified launch: at most kBatchMax problems per launch, and at
// most kBatchEntries (problem, stride offset) entries: the sum over problems
// of their rounds per ticket round.  order[q] packs entry q as problem
// (bits 0-7) | offset j (bits 8-17) | stride s (bits 18-27), s <= kBatchEntries.
constexpr int kBatchMax = 256;
constexpr int kBatchEntries = 256;  // the LDS tables of 256 entries keep two workgroups per CU

struct BatchArgs {
    const SgdArgs* __restrict__ probs;  // per problem; .epoch = its first epoch, .nblocks = its strata B
    const int64_t* seg_t0;    // nseg: first ticket of each run of equally wide rounds
    const int32_t* seg_r0;    // nseg: first ticket round of the run
    const int32_t* seg_w;     // nseg: tiles per ticket round in the run
    const int32_t* boff;      // nsorted + 1: offset of each sorted entry's tiles in a ticket round
    const int32_t* order;     // nsorted: sorted position -> packed entry
    const int64_t* done_off;  // per problem: first of its B progress counters in done
    uint32_t* done;           // per (problem, row block): rounds completed (uncached)
    uint64_t* ticket;         // next ticket (uncached)
    uint32_t* timed_out;      // set when a wave gave up waiting (uncached)
    int64_t nitems;           // tickets in the launch
    int32_t nsorted, nseg;    // entries (<= kBatchEntries), segments (<= nsorted)
    uint64_t deadline;        // wall_clock64 ticks a wave may run before it gives up waiting
};

struct SseArgs {
    int32_t m, n, k, ldf;
    const void* W;
    const void* H;
    const int64_t* colptr;
    const int32_t* rowidx;
    const void* val;
    int64_t nnz;
    int64_t chunk;         // ratings per wave
    double* partial;       // one FP64 partial per wave
};

}  // namespace nomad_k
