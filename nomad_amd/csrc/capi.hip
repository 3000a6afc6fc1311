// Host side of the C ABI declared in include/nomad_hip.h.
//
// Owns the device copies of the ratings and factors, builds the column routes
// (the reference's per-token thread permutation, Code/nomad/pool.cpp:32-38,
// nb_main.cpp:79-97 and 313-321), seeds the device rings and launches the
// kernels of sgd_kernels.hip / eval_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <new>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "kernel_args.h"
#include "nomad_hip.h"

using nomad_k::RouteDesc;
using nomad_k::SgdArgs;
using nomad_k::SseArgs;
using nomad_k::TokInfo;
using nomad_k::VisitDesc;

template <int S, int RM>
__global__ void sgd_strat_kernel(SgdArgs);
template <int S>
__global__ void sgd_batch_kernel(nomad_k::BatchArgs);
template <int S, int NW, int SLOTS, int RM>
__global__ void sgd_async_kernel(SgdArgs);
template <int S>
__global__ void sse_kernel(SseArgs);
__global__ void seed_rings_kernel(SgdArgs, const int32_t*, const int64_t*);

// dst[i] = src[idx[i]] (a permutation of the ratings' values)
template <class R>
__global__ void gather_kernel(const R* __restrict__ src, const int64_t* __restrict__ idx, R* __restrict__ dst,
                              int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        dst[i] = src[idx[i]];
}

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define NOMAD_HIP_TRY(expr)                                                                   \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(NOMAD_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

size_t elem_bytes(int storage) {
    return storage == NOMAD_STORE_FP64 ? 8 : storage == NOMAD_STORE_BF16 ? 2 : 4;
}

uint16_t bf16_from_float_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

float float_from_bf16(uint16_t b) {
    uint32_t u = uint32_t(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

template <class T>
int dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess)
        return fail(NOMAD_ENOMEM, std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                                      " bytes): " + hipGetErrorString(e));
    return NOMAD_OK;
}

template <class T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Memory that workgroups on different XCDs read and write during one kernel
// (column rings, token records, h_j) must not sit in an XCD's L2 as a stale
// copy: each XCD has its own L2 and plain (coarse-grained) device memory is
// only made coherent at kernel boundaries.  Such buffers are therefore
// allocated uncached, so every access reaches memory.  Uncached blocks are
// taken from a process-wide pool and are never returned to the HIP allocator
// while the process runs (freeing uncached memory that kernels have read has
// corrupted later allocations on this platform); a destroyed context hands
// its blocks back to the pool for reuse.
struct CoherentBlock {
    int device;
    void* ptr;
    size_t bytes;
    bool used;
};
std::mutex g_coh_mu;
std::vector<CoherentBlock> g_coh;

int coh_alloc_bytes(int device, void** p, size_t bytes) {
    *p = nullptr;
    bytes = std::max<size_t>(4096, (bytes + 4095) & ~size_t(4095));
    std::lock_guard<std::mutex> g(g_coh_mu);
    CoherentBlock* best = nullptr;
    for (auto& b : g_coh)
        if (!b.used && b.device == device && b.bytes >= bytes && (!best || b.bytes < best->bytes)) best = &b;
    if (best) {
        best->used = true;
        *p = best->ptr;
        return NOMAD_OK;
    }
    hipError_t e = hipExtMallocWithFlags(p, bytes, hipDeviceMallocUncached);
    if (e != hipSuccess)
        return fail(NOMAD_ENOMEM, std::string("uncached hipExtMallocWithFlags(") + std::to_string(bytes) +
                                      " bytes): " + hipGetErrorString(e));
    g_coh.push_back(CoherentBlock{device, *p, bytes, true});
    return NOMAD_OK;
}

template <class T>
int coh_alloc(int device, T** p, size_t count) {
    return coh_alloc_bytes(device, reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

template <class T>
void coh_free(T*& p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(g_coh_mu);
    for (auto& b : g_coh)
        if (b.ptr == static_cast<void*>(p)) b.used = false;
    p = nullptr;
}

// Device memory of nomad_eval_noalloc: the per-wave partial sums of the SSE
// kernel, a fixed array in the library's code object instead of memory the
// context allocates.  The runtime places one copy on each GPU the library
// runs on, kept until the process ends.  In host code `scratch` only names
// the symbol: its device address is looked up once per device, with that
// device current (nomad_create), and only that address is handed to kernels
// and copies.  Every call on a GPU shares the array, so calls that overlap on
// one GPU race; each call writes every partial it reads before reading it.
constexpr int64_t kScratchPartials = int64_t(1) << 20;  // 8 MiB of doubles
}  // namespace
static __device__ __align__(256) double scratch[kScratchPartials];
namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // per device ordinal: scratch's address there

// Looks up scratch's address on `device` (which must be current) once.
int resolve_scratch(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (int(g_scratch_ptr.size()) <= device) g_scratch_ptr.resize(size_t(device) + 1, nullptr);
    if (g_scratch_ptr[size_t(device)]) return NOMAD_OK;
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(NOMAD_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[size_t(device)] = p;
    return NOMAD_OK;
}

double* scratch_on(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return device >= 0 && device < int(g_scratch_ptr.size()) ? g_scratch_ptr[size_t(device)] : nullptr;
}

struct DevCsc {
    int64_t nnz = -1;  // -1 = not uploaded
    int64_t* colptr = nullptr;
    int32_t* rowidx = nullptr;
    void* val = nullptr;
};

}  // namespace

// waves (row blocks) per workgroup of the async kernel: 4, 8 or 16
constexpr int kWaveChoices[3] = {4, 8, 16};
constexpr int kStatWidth = nomad_k::kStatWidth;

// error reporting shared with the library's other translation units
namespace nomad_host {
int fail(int code, const std::string& msg) { return ::fail(code, msg); }
}  // namespace nomad_host

// Contiguous row blocks minimising the largest block's rating count relative
// to its capacity cap[b] (all 1 when cap is null): the smallest bound L for
// which closing block b before the row that would take it past L * cap[b]
// needs at most P blocks (a row heavier than L * cap[b] gets a block of its
// own), found by bisection; if that split has fewer than P blocks, the block
// with the most rows is halved (by ratings) until there are P, so no block
// is empty.  O(nnz + m log nnz) host time.
extern "C" int nomad_row_blocks(const int32_t* rowidx, int64_t nnz, int32_t m, int32_t P, const double* cap,
                                int32_t* bounds) {
    if (!bounds || m <= 0 || P <= 0 || P > m || nnz < 0 || (nnz > 0 && !rowidx))
        return fail(NOMAD_EINVAL, "nomad_row_blocks: need 0 < P <= m, nnz >= 0 and the row indices");
    if (cap)
        for (int32_t b = 0; b < P; ++b)
            if (!(cap[b] > 0 && cap[b] <= 1)) return fail(NOMAD_EINVAL, "nomad_row_blocks: capacities must be in (0, 1]");
    auto limit = [&](int64_t L, int64_t b) -> double { return cap && b < P ? double(L) * cap[b] : double(L); };
    std::vector<int64_t> deg(size_t(m), 0);
    for (int64_t e = 0; e < nnz; ++e) {
        if (rowidx[e] < 0 || rowidx[e] >= m) return fail(NOMAD_EINVAL, "nomad_row_blocks: row index out of range");
        ++deg[size_t(rowidx[e])];
    }
    auto blocks_for = [&](int64_t L) {
        int64_t k = 1, cur = 0;
        for (int32_t r = 0; r < m; ++r) {
            if (cur > 0 && double(cur + deg[size_t(r)]) > limit(L, k - 1)) {
                ++k;
                cur = 0;
            }
            cur += deg[size_t(r)];
        }
        return k;
    };
    double cmin = 1.0;
    if (cap)
        for (int32_t b = 0; b < P; ++b) cmin = std::min(cmin, cap[b]);
    int64_t lo = (nnz + P - 1) / P, hi = std::max<int64_t>(lo, int64_t(double(std::max<int64_t>(nnz, 1)) / cmin) + 1);
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (blocks_for(mid) <= P)
            hi = mid;
        else
            lo = mid + 1;
    }
    std::vector<int32_t> start{0};
    int64_t cur = 0;
    for (int32_t r = 0; r < m; ++r) {
        if (cur > 0 && double(cur + deg[size_t(r)]) > limit(lo, int64_t(start.size()) - 1)) {
            start.push_back(r);
            cur = 0;
        }
        cur += deg[size_t(r)];
    }
    // fewer than P blocks: halve the block with the most rows, by ratings
    // (every split leaves at least one row on each side), largest first
    if (int32_t(start.size()) < P) {
        std::vector<std::pair<int32_t, int32_t>> heap;  // (rows, first row), max-heap on rows
        for (size_t b = 0; b < start.size(); ++b)
            heap.push_back({(b + 1 < start.size() ? start[b + 1] : m) - start[b], start[b]});
        std::make_heap(heap.begin(), heap.end());
        int32_t nblk = int32_t(heap.size());
        while (nblk < P) {
            std::pop_heap(heap.begin(), heap.end());
            const auto [rows, b0] = heap.back();
            heap.pop_back();
            const int32_t b1 = b0 + rows;  // rows >= 2: P <= m and every block has a row
            int64_t tot = 0;
            for (int32_t r = b0; r < b1; ++r) tot += deg[size_t(r)];
            int32_t cut = b0 + 1;
            int64_t acc = deg[size_t(b0)];
            while (cut < b1 - 1 && 2 * (acc + deg[size_t(cut)]) <= tot) acc += deg[size_t(cut++)];
            heap.push_back({cut - b0, b0});
            std::push_heap(heap.begin(), heap.end());
            heap.push_back({b1 - cut, cut});
            std::push_heap(heap.begin(), heap.end());
            ++nblk;
        }
        start.clear();
        for (const auto& blk : heap) start.push_back(blk.second);
        std::sort(start.begin(), start.end());
    }
    for (int32_t b = 0; b < P; ++b) bounds[b] = start[size_t(b)];
    bounds[P] = m;
    return NOMAD_OK;
}

struct nomad_ctx {
    nomad_params p{};
    int P = 0;  // row blocks: workers (async) or strata (strat)
    int rpb = 0, cpb = 0, ldf = 0;
    int units = 0;        // async: workgroups (route units)
    int nw = 16;          // async: waves (row blocks) per workgroup
    std::vector<int32_t> h_wbound;  // first row of each row block, P + 1 entries (async workers, strat strata)
    std::vector<int32_t> row_unit;  // async: workgroup owning each row; strat: row block of each row
    int32_t* d_wbound = nullptr;
    size_t esz = 4, rvsz = 4;
    uint32_t sr_seed = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double wall_hz = 100e6;

    void* dW = nullptr;
    void* dH = nullptr;
    bool h_coherent = false;  // dH lives in the uncached pool (async mode)
    DevCsc tr, te;
    std::vector<int64_t> h_colptr;  // training CSC kept on the host for routes
    std::vector<int32_t> h_rowidx;

    void* d_steps = nullptr;
    int nsteps = 0;
    double* d_partial = nullptr;
    int64_t partial_cap = 0;

    // async routes and rings
    bool routes_ready = false;
    bool routes_used = false;  // a launch has run on the current routes
    int64_t V = 0;
    std::vector<int64_t> h_vptr;
    std::vector<RouteDesc> h_route;  // per column its hops in route order
    int64_t* d_vptr = nullptr;
    RouteDesc* d_route = nullptr;
    int32_t* d_vcnt = nullptr;       // per (column, workgroup) pair, index RouteDesc::cslot: visits so far
    TokInfo* d_tok = nullptr;
    uint64_t* d_tok_stamp = nullptr;
    uint64_t* d_qslot = nullptr;
    uint32_t* d_qtail = nullptr;
    uint32_t* d_qhead = nullptr;
    int qlog2 = 0;
    uint32_t* d_done = nullptr;
    uint32_t* d_timed_out = nullptr;
    int64_t* d_stats = nullptr;
    std::vector<int64_t> stats_prev;
    std::vector<int64_t> last_unit_stats;  // per workgroup, last call
    int64_t last_stats[nomad_k::kStatWidth] = {0};
    int32_t ncols_active = 0;
    std::mt19937_64 rng_route;   // initial permutations (nb_main.cpp:79)
    std::mt19937_64 rng_reperm;  // re-permutation after a segment (nb_init.cpp:219, nb_main.cpp:316)
    int segments = 0;
    int64_t async_launches = 0;  // launch index of the async kernel (salts its rounding bits)
    int64_t cycle_base = 0;      // route cycles completed by every active column
    int64_t strat_epochs = 0;
    // strat: per tile (row block b, column block cb) its non-empty column
    // visits, columns ascending; d_tile_ptr has B * B + 1 entries
    bool tiles_ready = false;
    int64_t* d_tile_ptr = nullptr;
    VisitDesc* d_tile_visits = nullptr;
    int64_t* d_trow_ptr = nullptr;  // per tile its distinct rows ...
    int32_t* d_trows = nullptr;
    uint16_t* d_tslot = nullptr;    // ... and each rating's rank among them
    int32_t* d_tm_row = nullptr;    // the training ratings in tile order (strat)
    char* d_tm_val = nullptr;

    ~nomad_ctx() {
        dev_free(dW);
        if (h_coherent)
            coh_free(dH);
        else
            dev_free(dH);
        for (DevCsc* c : {&tr, &te}) {
            dev_free(c->colptr);
            dev_free(c->rowidx);
            dev_free(c->val);
        }
        dev_free(d_steps);
        dev_free(d_partial);
        dev_free(d_vptr);
        dev_free(d_route);
        dev_free(d_vcnt);
        dev_free(d_tile_ptr);
        dev_free(d_tile_visits);
        dev_free(d_trow_ptr);
        dev_free(d_trows);
        dev_free(d_tslot);
        dev_free(d_tm_row);
        dev_free(d_tm_val);
        dev_free(d_wbound);
        coh_free(d_tok);
        coh_free(d_tok_stamp);
        coh_free(d_qslot);
        coh_free(d_qtail);
        coh_free(d_qhead);
        coh_free(d_done);
        coh_free(d_timed_out);
        dev_free(d_stats);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Step table indexed by the visit count c (Code/nomad/nb_main.cpp:514-515):
//   step(c) = lrate * 1.5 / (1 + drate * (c + 1)^1.5)
// computed in double with the host libm; for float storage lrate and drate are
// first rounded to float and the step is rounded to float, as in the float
// build of the legacy updater (Code/nomad/nomad_body.hpp:313-320).
int ensure_steps(nomad_ctx* c, int64_t need) {
    if (need <= c->nsteps) return NOMAD_OK;
    int n = int(std::max<int64_t>(need + 1024, 4096));
    const bool is64 = c->p.storage == NOMAD_STORE_FP64;
    std::vector<char> buf(size_t(n) * (is64 ? 8 : 4));
    for (int i = 0; i < n; ++i) {
        if (is64) {
            double s = c->p.lrate * 1.5 / (1.0 + c->p.drate * std::pow(double(i) + 1.0, 1.5));
            std::memcpy(buf.data() + size_t(i) * 8, &s, 8);
        } else {
            const float lr = float(c->p.lrate), dr = float(c->p.drate);
            float s = float(double(lr) * 1.5 / (1.0 + double(dr) * std::pow(double(i) + 1.0, 1.5)));
            std::memcpy(buf.data() + size_t(i) * 4, &s, 4);
        }
    }
    dev_free(c->d_steps);
    int rc = dev_alloc(reinterpret_cast<char**>(&c->d_steps), buf.size());
    if (rc) return rc;
    NOMAD_HIP_TRY(hipMemcpy(c->d_steps, buf.data(), buf.size(), hipMemcpyHostToDevice));
    c->nsteps = n;
    return NOMAD_OK;
}

int upload_csc(nomad_ctx* c, DevCsc& d, const int64_t* colptr, const int32_t* rowidx,
               const double* val, int64_t nnz, bool keep_host) {
    const int32_t m = c->p.m, n = c->p.n;
    if (!colptr || (nnz > 0 && (!rowidx || !val)) || nnz < 0)
        return fail(NOMAD_EINVAL, "upload: null pointer or negative nnz");
    if (colptr[0] != 0 || colptr[n] != nnz) return fail(NOMAD_EINVAL, "upload: colptr does not span nnz");
    for (int32_t j = 0; j < n; ++j) {
        if (colptr[j + 1] < colptr[j]) return fail(NOMAD_EINVAL, "upload: colptr not monotone");
        for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) {
            if (rowidx[e] < 0 || rowidx[e] >= m) return fail(NOMAD_EINVAL, "upload: row index out of range");
            if (e > colptr[j] && rowidx[e] < rowidx[e - 1])
                return fail(NOMAD_EINVAL, "upload: rows not ascending within a column");
        }
    }
    dev_free(d.colptr);
    dev_free(d.rowidx);
    dev_free(d.val);
    int rc;
    if ((rc = dev_alloc(&d.colptr, size_t(n) + 1))) return rc;
    if ((rc = dev_alloc(&d.rowidx, size_t(nnz)))) return rc;
    if ((rc = dev_alloc(reinterpret_cast<char**>(&d.val), size_t(nnz) * c->rvsz))) return rc;
    NOMAD_HIP_TRY(hipMemcpy(d.colptr, colptr, (size_t(n) + 1) * 8, hipMemcpyHostToDevice));
    if (nnz > 0) {
        NOMAD_HIP_TRY(hipMemcpy(d.rowidx, rowidx, size_t(nnz) * 4, hipMemcpyHostToDevice));
        if (c->rvsz == 8) {
            NOMAD_HIP_TRY(hipMemcpy(d.val, val, size_t(nnz) * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(val, val + nnz);
            NOMAD_HIP_TRY(hipMemcpy(d.val, f.data(), size_t(nnz) * 4, hipMemcpyHostToDevice));
        }
    }
    d.nnz = nnz;
    if (keep_host) {
        c->h_colptr.assign(colptr, colptr + n + 1);
        c->h_rowidx.assign(rowidx, rowidx + nnz);
        c->routes_ready = false;
        c->tiles_ready = false;
    }
    return NOMAD_OK;
}

// Tile lists of the stratified schedule: row block b = rows [h_wbound[b],
// h_wbound[b + 1]), column block cb = columns [cb * cpb, (cb + 1) * cpb).  The
// visits of row block b are listed by ascending column, so tile (b, cb) is
// the contiguous run of them whose column lies in column block cb.  One pass
// over the ratings; a final sentinel visit (len 0) lets a tile loop read one
// descriptor ahead without a bounds test.
int ensure_tiles(nomad_ctx* c) {
    if (c->tiles_ready) return NOMAD_OK;
    const int32_t n = c->p.n, B = c->P, m = c->p.m;
    // Row blocks: ceil(m / B) rows each, or with NOMAD_FLAG_BALANCE the
    // contiguous split whose largest training-rating count is smallest
    // (nomad_row_blocks): a sub-epoch waits for its heaviest tile, so blocks
    // of equal ratings shorten it.
    c->h_wbound.assign(size_t(B) + 1, m);
    if (c->p.flags & NOMAD_FLAG_BALANCE) {
        const int brc = nomad_row_blocks(c->h_rowidx.data(), int64_t(c->h_rowidx.size()), m, B, nullptr,
                                         c->h_wbound.data());
        if (brc) return brc;  // (its message is set)
    } else {
        for (int b = 0; b <= B; ++b) c->h_wbound[size_t(b)] = int32_t(std::min<int64_t>(int64_t(b) * c->rpb, m));
    }
    int32_t maxrows = 0;
    c->row_unit.assign(size_t(m), 0);
    for (int32_t b = 0; b < B; ++b) {
        maxrows = std::max(maxrows, c->h_wbound[size_t(b) + 1] - c->h_wbound[size_t(b)]);
        for (int32_t r = c->h_wbound[size_t(b)]; r < c->h_wbound[size_t(b) + 1]; ++r) c->row_unit[size_t(r)] = b;
    }
    // a row block is addressed through one buffer resource (32-bit range)
    if (int64_t(maxrows) * c->ldf * int64_t(c->esz) >= (int64_t(1) << 31))
        return fail(NOMAD_EINVAL, "strat mode: a row block exceeds 2 GiB of W; use more strata");
    const std::vector<int32_t>& blk = c->row_unit;
    std::vector<std::vector<VisitDesc>> per(static_cast<size_t>(B));
    for (int32_t j = 0; j < n; ++j) {
        const int64_t c0 = c->h_colptr[j], c1 = c->h_colptr[j + 1];
        int64_t e = c0;
        while (e < c1) {
            const int32_t b = blk[size_t(c->h_rowidx[e])];
            int64_t f = e + 1;
            while (f < c1 && blk[size_t(c->h_rowidx[f])] == b) ++f;
            per[b].push_back(VisitDesc{e, int32_t(f - e), j});
            e = f;
        }
    }
    std::vector<int64_t> tptr(size_t(B) * B + 1, 0);
    std::vector<VisitDesc> vis;
    size_t total = 0;
    for (const auto& v : per) total += v.size();
    vis.reserve(total + 1);
    for (int32_t b = 0; b < B; ++b) {
        size_t x = 0;
        const auto& v = per[b];
        for (int32_t cb = 0; cb < B; ++cb) {
            tptr[size_t(b) * B + cb] = int64_t(vis.size());
            const int32_t chi = std::min<int64_t>(n, int64_t(cb + 1) * c->cpb);
            while (x < v.size() && v[x].worker < chi) vis.push_back(v[x++]);
        }
    }
    tptr[size_t(B) * B] = int64_t(vis.size());
    vis.push_back(VisitDesc{0, 0, 0});
    // Each tile's distinct rows (ascending) and every rating's rank among
    // them: a kernel may hold just those rows in LDS for the tile.  Ranks
    // past 65535 are stored as 65535 (such a tile never fits in LDS).
    const int64_t nnz = c->tr.nnz;
    std::vector<int64_t> trp(size_t(B) * B + 1, 0);
    std::vector<int32_t> trows;
    std::vector<uint16_t> tslot(static_cast<size_t>(std::max<int64_t>(nnz, 1)), 0);
    {
        std::vector<int32_t> mark(static_cast<size_t>(std::max(maxrows, 1)), -1), rows_t;
        for (int32_t t = 0; t < B * B; ++t) {
            const int32_t rlo = c->h_wbound[size_t(t / B)];
            trp[size_t(t)] = int64_t(trows.size());
            rows_t.clear();
            for (int64_t v = tptr[size_t(t)]; v < tptr[size_t(t) + 1]; ++v)
                for (int64_t e = vis[size_t(v)].start; e < vis[size_t(v)].start + vis[size_t(v)].len; ++e) {
                    int32_t& mk = mark[size_t(c->h_rowidx[size_t(e)] - rlo)];
                    if (mk < 0) {
                        mk = 0;
                        rows_t.push_back(c->h_rowidx[size_t(e)]);
                    }
                }
            std::sort(rows_t.begin(), rows_t.end());
            for (size_t q = 0; q < rows_t.size(); ++q) mark[size_t(rows_t[q] - rlo)] = int32_t(q);
            for (int64_t v = tptr[size_t(t)]; v < tptr[size_t(t) + 1]; ++v)
                for (int64_t e = vis[size_t(v)].start; e < vis[size_t(v)].start + vis[size_t(v)].len; ++e)
                    tslot[size_t(e)] = uint16_t(std::min(mark[size_t(c->h_rowidx[size_t(e)] - rlo)], 65535));
            for (int32_t r : rows_t) mark[size_t(r - rlo)] = -1;
            trows.insert(trows.end(), rows_t.begin(), rows_t.end());
        }
        trp[size_t(B) * B] = int64_t(trows.size());
    }
    // The ratings in tile order (tile after tile, visit after visit, rows
    // ascending), so that consecutive visits' ratings are contiguous; the
    // visits then index this order, and the sentinel's start is its end.
    // One padding entry lets a kernel read one past the last rating.
    const size_t ntm = static_cast<size_t>(nnz) + 1;
    std::vector<int64_t> perm(ntm, 0);  // tile-order position -> CSC position
    std::vector<int32_t> tm_row(ntm, 0);
    std::vector<uint16_t> tm_slot(ntm, 0);
    {
        int64_t o = 0;
        for (size_t v = 0; v + 1 < vis.size(); ++v) {
            const int64_t s0 = vis[v].start;
            vis[v].start = o;
            for (int64_t e = s0; e < s0 + vis[v].len; ++e, ++o) {
                perm[size_t(o)] = e;
                tm_row[size_t(o)] = c->h_rowidx[size_t(e)];
                tm_slot[size_t(o)] = tslot[size_t(e)];
            }
        }
        vis.back().start = o;
    }
    dev_free(c->d_tile_ptr);
    dev_free(c->d_tile_visits);
    dev_free(c->d_trow_ptr);
    dev_free(c->d_trows);
    dev_free(c->d_tslot);
    dev_free(c->d_tm_row);
    dev_free(c->d_tm_val);
    dev_free(c->d_wbound);
    int rc;
    if ((rc = dev_alloc(&c->d_wbound, c->h_wbound.size()))) return rc;
    NOMAD_HIP_TRY(hipMemcpy(c->d_wbound, c->h_wbound.data(), c->h_wbound.size() * 4, hipMemcpyHostToDevice));
    if ((rc = dev_alloc(&c->d_tile_ptr, tptr.size())) || (rc = dev_alloc(&c->d_tile_visits, vis.size())) ||
        (rc = dev_alloc(&c->d_trow_ptr, trp.size())) || (rc = dev_alloc(&c->d_trows, std::max<size_t>(trows.size(), 1))) ||
        (rc = dev_alloc(&c->d_tslot, ntm)) || (rc = dev_alloc(&c->d_tm_row, ntm)) ||
        (rc = dev_alloc(&c->d_tm_val, ntm * size_t(c->rvsz))))
        return rc;
    {
        int64_t* d_perm = nullptr;
        if ((rc = dev_alloc(&d_perm, ntm))) return rc;
        NOMAD_HIP_TRY(hipMemcpy(d_perm, perm.data(), ntm * 8, hipMemcpyHostToDevice));
        const int threads = 256;
        const int blocks = int(std::min<int64_t>((int64_t(ntm) + threads - 1) / threads, 8192));
        if (nnz > 0) {
            if (c->rvsz == 8)
                hipLaunchKernelGGL(gather_kernel<double>, dim3(blocks), dim3(threads), 0, c->stream,
                                   static_cast<const double*>(c->tr.val), d_perm,
                                   reinterpret_cast<double*>(c->d_tm_val), int64_t(ntm));
            else
                hipLaunchKernelGGL(gather_kernel<float>, dim3(blocks), dim3(threads), 0, c->stream,
                                   static_cast<const float*>(c->tr.val), d_perm,
                                   reinterpret_cast<float*>(c->d_tm_val), int64_t(ntm));
            NOMAD_HIP_TRY(hipGetLastError());
        } else {
            NOMAD_HIP_TRY(hipMemsetAsync(c->d_tm_val, 0, ntm * size_t(c->rvsz), c->stream));
        }
        NOMAD_HIP_TRY(hipStreamSynchronize(c->stream));
        dev_free(d_perm);
    }
    NOMAD_HIP_TRY(hipMemcpy(c->d_tile_ptr, tptr.data(), tptr.size() * 8, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(c->d_tile_visits, vis.data(), vis.size() * sizeof(VisitDesc), hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(c->d_trow_ptr, trp.data(), trp.size() * 8, hipMemcpyHostToDevice));
    if (!trows.empty())
        NOMAD_HIP_TRY(hipMemcpy(c->d_trows, trows.data(), trows.size() * 4, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(c->d_tslot, tm_slot.data(), ntm * 2, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(c->d_tm_row, tm_row.data(), ntm * 4, hipMemcpyHostToDevice));
    c->tiles_ready = true;
    return NOMAD_OK;
}

// Row blocks of the async workers.  Default: ceil(m / P) rows each, the
// reference's split (Code/nomad/nb_data.cpp:83-89).  With NOMAD_FLAG_BALANCE
// the blocks are contiguous row ranges whose largest training-rating count,
// relative to the block's capacity, is as small as a contiguous split into P
// blocks allows (nomad_row_blocks).
void build_partition(nomad_ctx* c) {
    const int32_t m = c->p.m;
    const int P = c->P;
    c->h_wbound.assign(size_t(P) + 1, m);
    if (c->p.flags & NOMAD_FLAG_BALANCE) {
        // The last wave of a workgroup hands every column on to the next
        // workgroup; those exits take about 0.37x the rest of its work at 8
        // waves per workgroup (device counters, Netflix shape), 0.046 x waves
        // in general, so its block's capacity is 1 / (1 + 0.046 x waves).
        std::vector<double> cap(size_t(P), 1.0);
        const int nw = c->nw > 0 ? c->nw : 1;
        if (nw > 1)
            for (int b = nw - 1; b < P; b += nw) cap[size_t(b)] = 1.0 / (1.0 + 0.046 * nw);
        nomad_row_blocks(c->h_rowidx.data(), int64_t(c->h_rowidx.size()), m, P, cap.data(), c->h_wbound.data());
    } else {
        for (int b = 0; b <= P; ++b) c->h_wbound[size_t(b)] = int32_t(std::min<int64_t>(int64_t(b) * c->rpb, m));
    }
    c->row_unit.assign(size_t(m), 0);
    for (int g = 0; g < c->units; ++g)
        for (int32_t r = c->h_wbound[size_t(g) * c->nw]; r < c->h_wbound[size_t(g + 1) * c->nw]; ++r)
            c->row_unit[size_t(r)] = g;
}

// Visit lists: for column j, the non-empty (column, worker) pairs ordered by
// the column's worker permutation.  perm is a std::shuffle of 0..P-1 drawn
// from rng, one per column in column order, like ColumnData::set_perm
// (Code/nomad/pool.cpp:32-38) with P workers in place of the reference's
// threads.  Workers with no rating of the column are skipped: their visit in
// the reference is an empty loop.
// Re-permutation after a segment (the reference re-shuffles every column's
// route after each evaluation, nb_main.cpp:313-321): each column's runs are
// put in a fresh uniformly random order by a Fisher-Yates shuffle.  The runs
// (the column's non-empty workgroup ranges) do not change, so this is
// O(visits); columns are shuffled in 64 fixed chunks on up to 16 host
// threads, each column from its own splitmix64 stream keyed by one draw of
// the re-permutation generator and the column, so the result does not
// depend on the thread count.
uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void reperm_routes(nomad_ctx* c) {
    const int32_t n = c->p.n;
    const uint64_t key = c->rng_reperm();
    constexpr int kChunks = 64;
    auto chunk = [&](int q) {
        const int32_t j0 = int32_t(int64_t(n) * q / kChunks), j1 = int32_t(int64_t(n) * (q + 1) / kChunks);
        for (int32_t j = j0; j < j1; ++j) {
            uint64_t st = key ^ (uint64_t(j) * 0xD1B54A32D192ED03ull);
            RouteDesc* v = c->h_route.data() + c->h_vptr[j];
            const int64_t len = c->h_vptr[j + 1] - c->h_vptr[j];
            for (int64_t i = len - 1; i > 0; --i) {
                const int64_t r = int64_t(splitmix64(st) % uint64_t(i + 1));
                std::swap(v[i], v[r]);
            }
        }
    };
    const int nth = int(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t)
        th.emplace_back([&, t] {
            for (int q = t; q < kChunks; q += nth) chunk(q);
        });
    for (int q = 0; q < kChunks; q += nth) chunk(q);
    for (auto& x : th) x.join();
}

// Initial routes: per column a shuffle of the workgroups with the route
// generator (the reference's set_perm, pool.cpp:32-38, with mt19937_64
// seeded seed + 137, nb_main.cpp:79), restricted to the workgroups that hold
// ratings of the column; each such (column, workgroup) pair is one run of
// the CSC column (rows ascending, so a workgroup's rows are contiguous).
// The pair's visit counter is its index in workgroup order: entry
// vptr[j] + i counts the visits of column j's i-th workgroup (ascending).
void build_routes(nomad_ctx* c, std::mt19937_64& rng) {
    const int32_t n = c->p.n, P = c->units;
    std::vector<int32_t> perm(P), rank(P);
    std::vector<RouteDesc> runs;
    std::vector<int64_t> vptr(size_t(n) + 1, 0);
    std::vector<RouteDesc> route;
    for (int32_t j = 0; j < n; ++j) {
        for (int32_t q = 0; q < P; ++q) perm[q] = q;
        std::shuffle(perm.begin(), perm.end(), rng);
        for (int32_t q = 0; q < P; ++q) rank[perm[q]] = q;
        runs.clear();
        const int64_t c0 = c->h_colptr[j], c1 = c->h_colptr[j + 1];
        for (int64_t e = c0; e < c1;) {
            const int32_t wk = c->row_unit[c->h_rowidx[e]];
            int64_t f = e;
            while (f < c1 && c->row_unit[c->h_rowidx[f]] == wk) ++f;
            runs.push_back(RouteDesc{uint32_t(e), int32_t(f - e), wk, int32_t(int64_t(route.size()) + int64_t(runs.size()))});
            e = f;
        }
        std::sort(runs.begin(), runs.end(),
                  [&](const RouteDesc& a, const RouteDesc& b) { return rank[a.unit] < rank[b.unit]; });
        route.insert(route.end(), runs.begin(), runs.end());
        vptr[j + 1] = int64_t(route.size());
    }
    c->h_vptr.swap(vptr);
    c->h_route.swap(route);
    c->V = int64_t(c->h_route.size());
}

// Row blocks, routes, rings, token records and visit counters of an async
// context, built once per upload of the training data (before the first
// launch, or when the host restores visit counters).
int build_async_state(nomad_ctx* c) {
    int rc;
    const int32_t n = c->p.n, P = c->units;
    std::vector<TokInfo> tok(static_cast<size_t>(n));
    {
        c->h_vptr.clear();
        c->h_route.clear();
        if (c->tr.nnz >= (int64_t(1) << 32)) return fail(NOMAD_EINVAL, "async mode: 2^32 or more ratings");
        build_partition(c);
        build_routes(c, c->rng_route);
        // ring capacity: more than the number of columns that can ever be
        // queued at one workgroup (its distinct columns)
        std::vector<int32_t> per_unit(P, 0);
        c->ncols_active = 0;
        for (int32_t j = 0; j < n; ++j) {
            if (c->h_vptr[j + 1] > c->h_vptr[j]) ++c->ncols_active;
            for (int64_t v = c->h_vptr[j]; v < c->h_vptr[j + 1]; ++v) per_unit[c->h_route[v].unit]++;
        }
        const int32_t maxq = P ? *std::max_element(per_unit.begin(), per_unit.end()) : 0;
        c->qlog2 = 1;
        while ((int64_t(1) << c->qlog2) < int64_t(maxq) + 1) ++c->qlog2;
        if (c->V >= (int64_t(1) << 31)) return fail(NOMAD_EINVAL, "async mode: 2^31 or more route hops");
        dev_free(c->d_vptr);
        dev_free(c->d_route);
        dev_free(c->d_vcnt);
        dev_free(c->d_wbound);
        coh_free(c->d_tok);
        coh_free(c->d_tok_stamp);
        coh_free(c->d_qslot);
        coh_free(c->d_qtail);
        coh_free(c->d_qhead);
        coh_free(c->d_done);
        coh_free(c->d_timed_out);
        dev_free(c->d_stats);
        if ((rc = dev_alloc(&c->d_vptr, size_t(n) + 1))) return rc;
        if ((rc = dev_alloc(&c->d_route, size_t(c->V)))) return rc;
        // visit counters start at zero, like the reference's per-thread
        // col_update_counts (nb_main.cpp:458-459), and persist across segments
        if ((rc = dev_alloc(&c->d_vcnt, size_t(c->V)))) return rc;
        NOMAD_HIP_TRY(hipMemset(c->d_vcnt, 0, size_t(c->V) * 4));
        if ((rc = dev_alloc(&c->d_wbound, c->h_wbound.size()))) return rc;
        NOMAD_HIP_TRY(hipMemcpy(c->d_wbound, c->h_wbound.data(), c->h_wbound.size() * 4, hipMemcpyHostToDevice));
        if ((rc = coh_alloc(c->p.device, &c->d_tok, size_t(n)))) return rc;
        if ((rc = coh_alloc(c->p.device, &c->d_tok_stamp, size_t(n)))) return rc;
        NOMAD_HIP_TRY(hipMemset(c->d_tok_stamp, 0, size_t(n) * 8));
        if ((rc = coh_alloc(c->p.device, &c->d_qslot, size_t(P) << c->qlog2))) return rc;
        if ((rc = coh_alloc(c->p.device, &c->d_qtail, size_t(P) * nomad_k::kQStride))) return rc;
        if ((rc = coh_alloc(c->p.device, &c->d_qhead, size_t(P) * nomad_k::kQStride))) return rc;
        if ((rc = coh_alloc(c->p.device, &c->d_done, 1))) return rc;
        if ((rc = coh_alloc(c->p.device, &c->d_timed_out, 1))) return rc;
        // device counters: one record per worker (wave)
        const size_t nstat = size_t(P) * size_t(c->nw) * kStatWidth;
        if ((rc = dev_alloc(&c->d_stats, nstat))) return rc;
        NOMAD_HIP_TRY(hipMemset(c->d_qslot, 0, (size_t(P) << c->qlog2) * 8));
        NOMAD_HIP_TRY(hipMemset(c->d_qtail, 0, size_t(P) * nomad_k::kQStride * 4));
        NOMAD_HIP_TRY(hipMemset(c->d_qhead, 0, size_t(P) * nomad_k::kQStride * 4));
        NOMAD_HIP_TRY(hipMemset(c->d_stats, 0, nstat * 8));
        c->stats_prev.assign(nstat, 0);
        for (auto& t : tok) t = TokInfo{0, 0, 0, 0, 0, 0};
        c->segments = 0;
        c->cycle_base = 0;
    }
    for (int32_t j = 0; j < n; ++j) {
        if (c->h_vptr[j + 1] > c->h_vptr[j]) {
            const RouteDesc& v = c->h_route[c->h_vptr[j]];
            tok[j].start = v.start;
            tok[j].len = v.len;
            tok[j].cslot = v.cslot;
        }
    }
    NOMAD_HIP_TRY(hipMemcpy(c->d_vptr, c->h_vptr.data(), c->h_vptr.size() * 8, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(c->d_route, c->h_route.data(), size_t(c->V) * sizeof(RouteDesc), hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(c->d_tok, tok.data(), size_t(n) * sizeof(TokInfo), hipMemcpyHostToDevice));
    c->routes_ready = true;
    c->routes_used = false;
    return NOMAD_OK;
}

// Routes, rings and token records for the next launch.  Every launch starts
// every active column at the first workgroup of its route; routes a launch
// has used are re-drawn first, as the reference re-permutes every column
// after each evaluation (nb_main.cpp:313-321, set_perm pool.cpp:32-38).
int setup_async(nomad_ctx* c) {
    if (!c->routes_ready) return build_async_state(c);
    if (!c->routes_used) return NOMAD_OK;
    const int32_t n = c->p.n, P = c->units;
    std::vector<TokInfo> tok(static_cast<size_t>(n));
    NOMAD_HIP_TRY(hipMemcpy(tok.data(), c->d_tok, size_t(n) * sizeof(TokInfo), hipMemcpyDeviceToHost));
    reperm_routes(c);
    // drop whatever is still queued: head = tail on every ring
    NOMAD_HIP_TRY(hipMemcpy(c->d_qhead, c->d_qtail, size_t(P) * nomad_k::kQStride * 4, hipMemcpyDeviceToDevice));
    for (int32_t j = 0; j < n; ++j) {
        if (c->h_vptr[j + 1] > c->h_vptr[j]) {
            const RouteDesc& v = c->h_route[c->h_vptr[j]];
            tok[j].start = v.start;
            tok[j].len = v.len;
            tok[j].cslot = v.cslot;
        }
        tok[j].pos = 0;
    }
    NOMAD_HIP_TRY(hipMemcpy(c->d_route, c->h_route.data(), size_t(c->V) * sizeof(RouteDesc), hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(c->d_tok, tok.data(), size_t(n) * sizeof(TokInfo), hipMemcpyHostToDevice));
    c->routes_used = false;
    return NOMAD_OK;
}

SgdArgs base_args(nomad_ctx* c) {
    SgdArgs a{};
    a.m = c->p.m;
    a.n = c->p.n;
    a.k = c->p.k;
    a.ldf = c->ldf;
    a.nblocks = c->p.mode == NOMAD_MODE_ASYNC ? c->units : c->P;
    a.rpb = c->rpb;
    a.cpb = c->cpb;
    a.round_mode = c->p.rounding;
    a.sr_seed = c->sr_seed;
    a.nsteps = c->nsteps;
    a.W = c->dW;
    a.H = c->dH;
    a.colptr = c->tr.colptr;
    a.rowidx = c->tr.rowidx;
    a.val = c->tr.val;
    a.steps = c->d_steps;
    if (c->p.mode == NOMAD_MODE_STRAT) {
        a.wbound = c->d_wbound;
        a.tile_ptr = c->d_tile_ptr;
        a.visits = c->d_tile_visits;
        a.rowidx = c->d_tm_row;
        a.val = c->d_tm_val;
        a.trow_ptr = c->d_trow_ptr;
        a.trows = c->d_trows;
        a.tslot = c->d_tslot;
    }
    a.lrate = c->p.storage == NOMAD_STORE_FP64 ? c->p.lrate : double(float(c->p.lrate));
    a.drate = c->p.storage == NOMAD_STORE_FP64 ? c->p.drate : double(float(c->p.drate));
    a.reg = c->p.storage == NOMAD_STORE_FP64 ? c->p.reg : double(float(c->p.reg));
    return a;
}

template <int S>
hipError_t launch_strat(const SgdArgs& a, hipStream_t st) {
    if constexpr ((S & 15) == NOMAD_STORE_BF16) {
        if (a.round_mode == NOMAD_ROUND_STOCHASTIC) {
            hipLaunchKernelGGL((sgd_strat_kernel<S, NOMAD_ROUND_STOCHASTIC>), dim3(a.nblocks), dim3(64), 0, st, a);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((sgd_strat_kernel<S, NOMAD_ROUND_NEAREST>), dim3(a.nblocks), dim3(64), 0, st, a);
    return hipGetLastError();
}

template <int S>
hipError_t launch_batch(const nomad_k::BatchArgs& a, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(sgd_batch_kernel<S>, dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int S>
hipError_t batch_occupancy(int* per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, sgd_batch_kernel<S>, 256, 0);
}

using nomad_k::async_slots;

// async kernel instances: rounding is a template argument for bf16 only
template <int S, int RM>
hipError_t launch_async_rm(const SgdArgs& a, int nw, hipStream_t st) {
    switch (nw) {
        case 4:
            hipLaunchKernelGGL((sgd_async_kernel<S, 4, async_slots<S, 4>(), RM>), dim3(a.nblocks), dim3(4 * 64), 0,
                               st, a);
            break;
        case 8:
            hipLaunchKernelGGL((sgd_async_kernel<S, 8, async_slots<S, 8>(), RM>), dim3(a.nblocks), dim3(8 * 64), 0,
                               st, a);
            break;
        default:
            hipLaunchKernelGGL((sgd_async_kernel<S, 16, async_slots<S, 16>(), RM>), dim3(a.nblocks),
                               dim3(16 * 64), 0, st, a);
    }
    return hipGetLastError();
}

template <int S>
hipError_t launch_async(const SgdArgs& a, int nw, int rounding, hipStream_t st) {
    if constexpr ((S & 15) == NOMAD_STORE_BF16) {
        if (rounding == NOMAD_ROUND_STOCHASTIC) return launch_async_rm<S, NOMAD_ROUND_STOCHASTIC>(a, nw, st);
    }
    return launch_async_rm<S, NOMAD_ROUND_NEAREST>(a, nw, st);
}

template <int S, int RM>
hipError_t async_occupancy_rm(int nw, int* per_cu) {
    switch (nw) {
        case 4:
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, sgd_async_kernel<S, 4, async_slots<S, 4>(), RM>,
                                                                4 * 64, 0);
        case 8:
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, sgd_async_kernel<S, 8, async_slots<S, 8>(), RM>,
                                                                8 * 64, 0);
        default:
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(
                per_cu, sgd_async_kernel<S, 16, async_slots<S, 16>(), RM>, 16 * 64, 0);
    }
}

template <int S>
hipError_t async_occupancy(int nw, int rounding, int* per_cu) {
    if constexpr ((S & 15) == NOMAD_STORE_BF16) {
        if (rounding == NOMAD_ROUND_STOCHASTIC) return async_occupancy_rm<S, NOMAD_ROUND_STOCHASTIC>(nw, per_cu);
    }
    return async_occupancy_rm<S, NOMAD_ROUND_NEAREST>(nw, per_cu);
}

template <int S>
hipError_t launch_sse(const SseArgs& a, int64_t waves, hipStream_t st) {
    const int64_t blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(sse_kernel<S>, dim3(unsigned(blocks)), dim3(256), 0, st, a);
    return hipGetLastError();
}

// Kernel storage code: the storage type, plus kWideLayout when a factor row
// has more than 128 elements (the kernels then hold two element pairs per
// lane: pairs l and l + 64).
constexpr int kWideLayout = 16;
int kernel_code(int storage, int ldf) { return storage | (ldf > 128 ? kWideLayout : 0); }

template <class F>
hipError_t dispatch_storage(int code, F&& f) {
    switch (code) {
        case NOMAD_STORE_FP32: return f(std::integral_constant<int, NOMAD_STORE_FP32>{});
        case NOMAD_STORE_BF16: return f(std::integral_constant<int, NOMAD_STORE_BF16>{});
        case NOMAD_STORE_FP64: return f(std::integral_constant<int, NOMAD_STORE_FP64>{});
        case NOMAD_STORE_FP32 | kWideLayout: return f(std::integral_constant<int, NOMAD_STORE_FP32 | kWideLayout>{});
        case NOMAD_STORE_BF16 | kWideLayout: return f(std::integral_constant<int, NOMAD_STORE_BF16 | kWideLayout>{});
        default: return f(std::integral_constant<int, NOMAD_STORE_FP64 | kWideLayout>{});
    }
}

double watchdog_seconds() {
    const char* s = std::getenv("NOMAD_WATCHDOG_SECONDS");
    double v = s ? std::atof(s) : 0.0;
    return v > 0 ? v : 120.0;
}

int train_async(nomad_ctx* c, double seconds, int64_t epochs, nomad_train_stats* st) {
    int rc = setup_async(c);
    if (rc) return rc;
    const int32_t n = c->p.n, P = c->units;
    // seed every active column at the first worker of its (new) route, in
    // column order
    std::vector<int64_t> off(size_t(P) + 1, 0);
    for (int32_t j = 0; j < n; ++j)
        if (c->h_vptr[j + 1] > c->h_vptr[j]) off[c->h_route[c->h_vptr[j]].unit + 1]++;
    for (int32_t q = 0; q < P; ++q) off[q + 1] += off[q];
    std::vector<int32_t> list(static_cast<size_t>(off[P]));
    {
        std::vector<int64_t> fill(off.begin(), off.end() - 1);
        for (int32_t j = 0; j < n; ++j)
            if (c->h_vptr[j + 1] > c->h_vptr[j]) list[fill[c->h_route[c->h_vptr[j]].unit]++] = j;
    }
    int32_t* d_list = nullptr;
    int64_t* d_off = nullptr;
    if ((rc = dev_alloc(&d_list, list.size()))) return rc;
    if ((rc = dev_alloc(&d_off, off.size()))) {
        dev_free(d_list);
        return rc;
    }
    auto cleanup = [&] {
        dev_free(d_list);
        dev_free(d_off);
    };
    if (hipMemcpy(d_list, list.data(), list.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_off, off.data(), off.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return fail(NOMAD_EHIP, "copy of ring seeds failed");
    }
    const int64_t target = epochs > 0 ? c->cycle_base + epochs : int64_t(INT_MAX);
    SgdArgs a = base_args(c);
    a.wbound = c->d_wbound;
    a.vptr = c->d_vptr;
    a.route = c->d_route;
    a.vcnt = c->d_vcnt;
    a.tok = c->d_tok;
    a.tok_stamp = c->d_tok_stamp;
    a.qslot = c->d_qslot;
    a.qtail = c->d_qtail;
    a.qhead = c->d_qhead;
    a.done = c->d_done;
    a.stats = c->d_stats;
    a.timed_out = c->d_timed_out;
    a.qlog2 = c->qlog2;
    a.target_cycle = int32_t(std::min<int64_t>(target, INT_MAX));
    a.ncols_active = c->ncols_active;
    a.timing = (c->p.flags & NOMAD_FLAG_TIMING) ? 1 : 0;
    a.slot_cap = c->p.ptoken;
    a.epoch = int32_t(c->async_launches++ & 0x7FFFFFFF);
    const double limit = seconds > 0 ? seconds : watchdog_seconds();
    a.deadline = uint64_t(limit * c->wall_hz);
    hipError_t e = hipMemsetAsync(c->d_done, 0, 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->d_timed_out, 0, 4, c->stream);
    if (e == hipSuccess) {
        const int threads = 256;
        hipLaunchKernelGGL(seed_rings_kernel, dim3((P + threads - 1) / threads), dim3(threads), 0,
                           c->stream, a, d_list, d_off);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    if (e == hipSuccess)
        e = dispatch_storage(kernel_code(c->p.storage, c->ldf),
                             [&](auto s) { return launch_async<decltype(s)::value>(a, c->nw, c->p.rounding, c->stream); });
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) return fail(NOMAD_EHIP, std::string("async training launch: ") + hipGetErrorString(e));
    float ms = 0.f;
    NOMAD_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    uint32_t done = 0, timed_out = 0;
    NOMAD_HIP_TRY(hipMemcpy(&done, c->d_done, 4, hipMemcpyDeviceToHost));
    NOMAD_HIP_TRY(hipMemcpy(&timed_out, c->d_timed_out, 4, hipMemcpyDeviceToHost));
    std::vector<int64_t> stats(size_t(P) * size_t(c->nw) * kStatWidth);  // per worker
    NOMAD_HIP_TRY(hipMemcpy(stats.data(), c->d_stats, stats.size() * 8, hipMemcpyDeviceToHost));
    int64_t sum[kStatWidth] = {0};
    for (int32_t q = 0; q < P * c->nw; ++q)
        for (int i = 0; i < kStatWidth; ++i) sum[i] += stats[size_t(q) * kStatWidth + i] - c->stats_prev[size_t(q) * kStatWidth + i];
    c->last_unit_stats.resize(stats.size());
    for (size_t i = 0; i < stats.size(); ++i) c->last_unit_stats[i] = stats[i] - c->stats_prev[i];
    c->stats_prev.swap(stats);
    for (int i = 0; i < kStatWidth; ++i) c->last_stats[i] = sum[i];
    const int64_t upd = sum[nomad_k::kStatUpdates], vis = sum[nomad_k::kStatVisits], fails = sum[nomad_k::kStatIdle];
    ++c->segments;
    c->routes_used = true;
    const bool all_parked = int32_t(done) >= c->ncols_active;
    if (all_parked && epochs > 0) {
        c->cycle_base = target;
    } else {
        std::vector<TokInfo> tok(static_cast<size_t>(n));
        NOMAD_HIP_TRY(hipMemcpy(tok.data(), c->d_tok, size_t(n) * sizeof(TokInfo), hipMemcpyDeviceToHost));
        int64_t mn = INT64_MAX;
        for (int32_t j = 0; j < n; ++j)
            if (c->h_vptr[j + 1] > c->h_vptr[j]) mn = std::min<int64_t>(mn, tok[j].cycle);
        c->cycle_base = mn == INT64_MAX ? c->cycle_base : mn;
    }
    if (st) {
        st->updates = upd;
        st->visits = vis;
        st->failed_pops = fails;
        st->epochs_done = c->cycle_base;
        st->train_seconds = ms * 1e-3;
        const double ks = double(c->p.k) * double(c->esz);
        st->alg_bytes = double(upd) * (4.0 + double(c->rvsz) + 2.0 * ks) + double(vis) * 2.0 * ks;
    }
    if (seconds <= 0 && !all_parked)
        return fail(NOMAD_ETIMEOUT, "async training stopped by the device watchdog after " +
                                        std::to_string(limit) + " s before finishing its epochs");
    (void)timed_out;
    return NOMAD_OK;
}

// Statistics of `epochs` stratified epochs: every rating once per epoch;
// visits = non-empty (column, row block) pairs.
void strat_stats(const nomad_ctx* c, int64_t epochs, double seconds, nomad_train_stats* st) {
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    int64_t vis = 0;
    for (int32_t j = 0; j < c->p.n; ++j) {
        int32_t last = -1;
        for (int64_t e = c->h_colptr[j]; e < c->h_colptr[j + 1]; ++e) {
            const int32_t b = c->row_unit[size_t(c->h_rowidx[e])];
            if (b != last) {
                ++vis;
                last = b;
            }
        }
    }
    st->updates = c->tr.nnz * epochs;
    st->visits = vis * epochs;
    st->failed_pops = 0;
    st->epochs_done = c->strat_epochs;
    st->train_seconds = seconds;
    const double ks = double(c->p.k) * double(c->esz);
    st->alg_bytes = double(st->updates) * (4.0 + double(c->rvsz) + 2.0 * ks) + double(st->visits) * 2.0 * ks;
}

int train_strat(nomad_ctx* c, int64_t epochs, nomad_train_stats* st) {
    if (epochs <= 0) return fail(NOMAD_EINVAL, "strat mode needs epochs > 0");
    int rc = ensure_steps(c, c->strat_epochs + epochs + 2);
    if (rc) return rc;
    if ((rc = ensure_tiles(c))) return rc;
    SgdArgs a = base_args(c);
    NOMAD_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int64_t ep = 0; ep < epochs; ++ep) {
        a.epoch = int32_t(c->strat_epochs + ep);
        for (int s = 0; s < c->P; ++s) {
            a.substep = s;
            hipError_t e = dispatch_storage(kernel_code(c->p.storage, c->ldf),
                                            [&](auto t) { return launch_strat<decltype(t)::value>(a, c->stream); });
            if (e != hipSuccess) return fail(NOMAD_EHIP, std::string("strat launch: ") + hipGetErrorString(e));
        }
    }
    NOMAD_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    NOMAD_HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    NOMAD_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->strat_epochs += epochs;
    strat_stats(c, epochs, ms * 1e-3, st);
    return NOMAD_OK;
}

// Batched stratified training (see sgd_batch_kernel): every context runs
// `epochs` epochs of its own B x B schedule in one persistent launch.
int train_batch(nomad_ctx* const* cs, int32_t count, int64_t epochs, int32_t waves, nomad_train_stats* stats) {
    using nomad_k::BatchArgs;
    if (!cs || count <= 0 || count > nomad_k::kBatchMax || epochs <= 0 || waves < 0)
        return fail(NOMAD_EINVAL, "nomad_train_batch: need 1..256 contexts, epochs > 0 and waves >= 0");
    for (int32_t i = 0; i < count; ++i) {
        const nomad_ctx* c = cs[i];
        if (!c) return fail(NOMAD_EINVAL, "nomad_train_batch: null context");
        if (c->p.mode != NOMAD_MODE_STRAT) return fail(NOMAD_EINVAL, "nomad_train_batch: contexts must be in strat mode");
        if (c->p.device != cs[0]->p.device || c->p.storage != cs[0]->p.storage ||
            (c->ldf > 128) != (cs[0]->ldf > 128))
            return fail(NOMAD_EINVAL,
                        "nomad_train_batch: contexts must share device, storage type and row layout (k <= 128 or > 128)");
        if (c->tr.nnz < 0) return fail(NOMAD_ESTATE, "nomad_train_batch: a context has no training data");
        for (int32_t j = 0; j < i; ++j)
            if (cs[j] == c) return fail(NOMAD_EINVAL, "nomad_train_batch: a context appears twice");
        if (epochs * int64_t(c->P) >= (int64_t(1) << 31))
            return fail(NOMAD_EINVAL, "nomad_train_batch: epochs * strata exceeds 2^31 rounds");
    }
    nomad_ctx* c0 = cs[0];
    NOMAD_HIP_TRY(hipSetDevice(c0->p.device));
    std::vector<SgdArgs> probs(static_cast<size_t>(count));
    std::vector<int64_t> rounds(static_cast<size_t>(count)), done_off(static_cast<size_t>(count));
    int64_t ndone = 0;
    for (int32_t i = 0; i < count; ++i) {
        nomad_ctx* c = cs[i];
        int rc = ensure_steps(c, c->strat_epochs + epochs + 2);
        if (rc) return rc;
        if ((rc = ensure_tiles(c))) return rc;
        NOMAD_HIP_TRY(hipStreamSynchronize(c->stream));
        probs[i] = base_args(c);
        probs[i].epoch = int32_t(c->strat_epochs);
        probs[i].substep = 0;
        rounds[i] = epochs * c->P;
        done_off[i] = ndone;
        ndone += c->P;
    }
    // strides: problem i runs s_i = ceil(B_i / Bref) of its rounds per ticket
    // round, Bref the smallest strata count whose strides fit kBatchEntries
    // (the largest count always does: every stride is then 1)
    std::vector<int32_t> Bs(static_cast<size_t>(count));
    for (int32_t i = 0; i < count; ++i) Bs[size_t(i)] = cs[i]->P;
    std::sort(Bs.begin(), Bs.end());
    int32_t Bref = Bs.back();
    for (int32_t cand : Bs) {
        int64_t sum = 0;
        for (int32_t i = 0; i < count; ++i) sum += (cs[i]->P + cand - 1) / cand;
        if (sum <= nomad_k::kBatchEntries) {
            Bref = cand;
            break;
        }
    }
    struct Entry {
        int32_t p, j, s;
        int64_t ur;  // ticket rounds the entry runs: ceil((rounds_p - j) / s)
    };
    std::vector<Entry> ent;
    for (int32_t i = 0; i < count; ++i) {
        const int32_t si = (cs[i]->P + Bref - 1) / Bref;
        for (int32_t j = 0; j < si; ++j) ent.push_back({i, j, si, (rounds[i] - j + si - 1) / si});
    }
    // sorted by ticket rounds descending (those still running in a ticket
    // round are a prefix), then by the midpoint (2j + 1) / 2s of the entry's
    // share of a ticket round, ascending, then by ratings per row block,
    // ascending: a round's light tiles go first and free their waves for
    // the next round's while the heavy ones run (measured on a sweep of 32
    // problems: 40.0 ms per epoch light first, 40.7 in problem order, 44.4
    // heavy first)
    std::vector<double> heavy(static_cast<size_t>(count));
    for (int32_t i = 0; i < count; ++i) heavy[size_t(i)] = double(cs[i]->tr.nnz) / double(cs[i]->P);
    std::stable_sort(ent.begin(), ent.end(), [&](const Entry& x, const Entry& y) {
        if (x.ur != y.ur) return x.ur > y.ur;
        const int64_t kx = int64_t(2 * x.j + 1) * y.s, ky = int64_t(2 * y.j + 1) * x.s;
        if (kx != ky) return kx < ky;
        return heavy[size_t(x.p)] < heavy[size_t(y.p)];
    });
    const int32_t E = int32_t(ent.size());
    std::vector<int32_t> order(static_cast<size_t>(E)), boff(static_cast<size_t>(E) + 1, 0);
    for (int32_t q = 0; q < E; ++q) {
        const Entry& x = ent[size_t(q)];
        order[size_t(q)] = x.p | (x.j << 8) | (x.s << 18);
        boff[size_t(q) + 1] = boff[size_t(q)] + cs[x.p]->P;
    }
    std::vector<int64_t> seg_t0;
    std::vector<int32_t> seg_r0, seg_w;
    int64_t t = 0, r_prev = 0;
    for (int32_t A = E; A >= 1; --A) {  // A entries running in ticket rounds [r_prev, that of sorted A-1)
        const int64_t r_end = ent[size_t(A) - 1].ur;
        if (r_end <= r_prev) continue;
        seg_t0.push_back(t);
        seg_r0.push_back(int32_t(r_prev));
        seg_w.push_back(boff[size_t(A)]);
        t += (r_end - r_prev) * boff[size_t(A)];
        r_prev = r_end;
    }
    const int64_t nitems = t;

    int cus = 0, per_cu = 0;
    NOMAD_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c0->p.device));
    NOMAD_HIP_TRY(dispatch_storage(kernel_code(c0->p.storage, c0->ldf), [&](auto tag) {
        return batch_occupancy<decltype(tag)::value>(&per_cu);
    }));
    // default grid: one wave per tile of the widest round, up to the waves
    // that fit on the device at once (more would only wait)
    int64_t want = waves > 0 ? waves : std::min<int64_t>(boff[size_t(E)], int64_t(cus) * std::max(per_cu, 1) * 4);
    const int blocks = int(std::max<int64_t>(1, std::min<int64_t>((want + 3) / 4, 65535)));

    struct Scratch {
        SgdArgs* probs = nullptr;
        int64_t* seg_t0 = nullptr;
        int32_t* seg_r0 = nullptr;
        int32_t* seg_w = nullptr;
        int32_t* boff = nullptr;
        int32_t* order = nullptr;
        int64_t* done_off = nullptr;
        uint32_t* done = nullptr;
        uint64_t* ticket = nullptr;
        uint32_t* timed_out = nullptr;
        ~Scratch() {
            dev_free(probs);
            dev_free(seg_t0);
            dev_free(seg_r0);
            dev_free(seg_w);
            dev_free(boff);
            dev_free(order);
            dev_free(done_off);
            coh_free(done);
            coh_free(ticket);
            coh_free(timed_out);
        }
    } w;
    int rc;
    const size_t nseg = seg_t0.size();
    if ((rc = dev_alloc(&w.probs, size_t(count))) || (rc = dev_alloc(&w.seg_t0, nseg)) ||
        (rc = dev_alloc(&w.seg_r0, nseg)) || (rc = dev_alloc(&w.seg_w, nseg)) ||
        (rc = dev_alloc(&w.boff, boff.size())) || (rc = dev_alloc(&w.order, order.size())) ||
        (rc = dev_alloc(&w.done_off, done_off.size())) || (rc = coh_alloc(c0->p.device, &w.done, size_t(ndone))) ||
        (rc = coh_alloc(c0->p.device, &w.ticket, 1)) || (rc = coh_alloc(c0->p.device, &w.timed_out, 1)))
        return rc;
    NOMAD_HIP_TRY(hipMemcpy(w.probs, probs.data(), probs.size() * sizeof(SgdArgs), hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(w.seg_t0, seg_t0.data(), nseg * 8, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(w.seg_r0, seg_r0.data(), nseg * 4, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(w.seg_w, seg_w.data(), nseg * 4, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(w.boff, boff.data(), boff.size() * 4, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(w.order, order.data(), order.size() * 4, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemcpy(w.done_off, done_off.data(), done_off.size() * 8, hipMemcpyHostToDevice));
    NOMAD_HIP_TRY(hipMemsetAsync(w.done, 0, size_t(ndone) * 4, c0->stream));
    NOMAD_HIP_TRY(hipMemsetAsync(w.ticket, 0, 8, c0->stream));
    NOMAD_HIP_TRY(hipMemsetAsync(w.timed_out, 0, 4, c0->stream));

    BatchArgs ba{};
    ba.probs = w.probs;
    ba.seg_t0 = w.seg_t0;
    ba.seg_r0 = w.seg_r0;
    ba.seg_w = w.seg_w;
    ba.boff = w.boff;
    ba.order = w.order;
    ba.done_off = w.done_off;
    ba.done = w.done;
    ba.ticket = w.ticket;
    ba.timed_out = w.timed_out;
    ba.nitems = nitems;
    ba.nsorted = E;
    ba.nseg = int32_t(nseg);
    ba.deadline = uint64_t(watchdog_seconds() * c0->wall_hz);
    NOMAD_HIP_TRY(hipEventRecord(c0->ev0, c0->stream));
    hipError_t e = dispatch_storage(kernel_code(c0->p.storage, c0->ldf), [&](auto tag) {
        return launch_batch<decltype(tag)::value>(ba, blocks, c0->stream);
    });
    if (e != hipSuccess) return fail(NOMAD_EHIP, std::string("batch launch: ") + hipGetErrorString(e));
    NOMAD_HIP_TRY(hipEventRecord(c0->ev1, c0->stream));
    NOMAD_HIP_TRY(hipStreamSynchronize(c0->stream));
    float ms = 0.f;
    NOMAD_HIP_TRY(hipEventElapsedTime(&ms, c0->ev0, c0->ev1));
    uint32_t timed_out = 0;
    NOMAD_HIP_TRY(hipMemcpy(&timed_out, w.timed_out, 4, hipMemcpyDeviceToHost));
    if (timed_out)
        return fail(NOMAD_ETIMEOUT, "batched training stopped by the device watchdog after " +
                                        std::to_string(watchdog_seconds()) + " s; factors are partial");
    for (int32_t i = 0; i < count; ++i) {
        cs[i]->strat_epochs += epochs;
        strat_stats(cs[i], epochs, ms * 1e-3, stats ? stats + i : nullptr);
    }
    return NOMAD_OK;
}

// SSE of one rating matrix: one FP64 partial per wave of 512 ratings, added
// on the host in wave order.  The partials go to the context's own buffer
// (allocated on first need) or, with `fixed`, to the library's scratch array.
int eval_one(nomad_ctx* c, const DevCsc& d, double* sse, bool fixed) {
    *sse = 0.0;
    if (d.nnz <= 0) return NOMAD_OK;
    const int64_t chunk = 512;
    const int64_t waves = (d.nnz + chunk - 1) / chunk;
    double* partial = nullptr;
    if (fixed) {
        // the launch covers whole blocks of 4 waves, each writing its partial
        if (waves + 4 > kScratchPartials)
            return fail(NOMAD_EINVAL, "nomad_eval_noalloc: " + std::to_string(d.nnz) +
                                          " ratings exceed the fixed scratch (" +
                                          std::to_string((kScratchPartials - 4) * chunk) + " at most)");
        partial = scratch_on(c->p.device);
        if (!partial) return fail(NOMAD_ESTATE, "nomad_eval_noalloc: scratch address not resolved");
    } else {
        if (waves > c->partial_cap) {
            dev_free(c->d_partial);
            int rc = dev_alloc(&c->d_partial, size_t(waves + 4));
            if (rc) return rc;
            c->partial_cap = waves + 4;
        }
        partial = c->d_partial;
    }
    SseArgs a{};
    a.m = c->p.m;
    a.n = c->p.n;
    a.k = c->p.k;
    a.ldf = c->ldf;
    a.W = c->dW;
    a.H = c->dH;
    a.colptr = d.colptr;
    a.rowidx = d.rowidx;
    a.val = d.val;
    a.nnz = d.nnz;
    a.chunk = chunk;
    a.partial = partial;
    hipError_t e = dispatch_storage(kernel_code(c->p.storage, c->ldf),
                                    [&](auto t) { return launch_sse<decltype(t)::value>(a, waves, c->stream); });
    if (e != hipSuccess) return fail(NOMAD_EHIP, std::string("sse launch: ") + hipGetErrorString(e));
    std::vector<double> part(static_cast<size_t>(waves));
    NOMAD_HIP_TRY(hipMemcpyAsync(part.data(), partial, size_t(waves) * 8, hipMemcpyDeviceToHost, c->stream));
    NOMAD_HIP_TRY(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double v : part) s += v;
    *sse = s;
    return NOMAD_OK;
}

}  // namespace

extern "C" {

const char* nomad_last_error(void) { return g_last_error.c_str(); }

int nomad_abi_version(void) { return NOMAD_ABI_VERSION; }

int nomad_create(const nomad_params* p, nomad_ctx** out) {
    if (!p || !out) return fail(NOMAD_EINVAL, "nomad_create: null argument");
    *out = nullptr;
    if (p->m <= 0 || p->n <= 0 || p->k <= 0) return fail(NOMAD_EINVAL, "nomad_create: m, n, k must be positive");
    if (p->k > 256)
        return fail(NOMAD_EINVAL, "nomad_create: k > 256 is not supported (at most two element pairs per lane)");
    if (p->storage < NOMAD_STORE_FP32 || p->storage > NOMAD_STORE_FP64)
        return fail(NOMAD_EINVAL, "nomad_create: unknown storage type");
    if (p->mode != NOMAD_MODE_ASYNC && p->mode != NOMAD_MODE_STRAT)
        return fail(NOMAD_EINVAL, "nomad_create: unknown mode");
    if (p->rounding != NOMAD_ROUND_NEAREST && p->rounding != NOMAD_ROUND_STOCHASTIC)
        return fail(NOMAD_EINVAL, "nomad_create: unknown rounding");
    if (p->nworkers < 0 || p->ptoken < 0 || (p->flags & ~(NOMAD_FLAG_BALANCE | NOMAD_FLAG_TIMING)) != 0)
        return fail(NOMAD_EINVAL, "nomad_create: negative nworkers/ptoken or unknown flags");
    if (!(p->lrate > 0) || !(p->drate >= 0) || !(p->reg >= 0))
        return fail(NOMAD_EINVAL, "nomad_create: lrate must be > 0, drate and reg >= 0");
    int ndev = 0;
    NOMAD_HIP_TRY(hipGetDeviceCount(&ndev));
    if (p->device < 0 || p->device >= ndev) return fail(NOMAD_EINVAL, "nomad_create: no such device");
    NOMAD_HIP_TRY(hipSetDevice(p->device));

    nomad_ctx* c = new (std::nothrow) nomad_ctx();
    if (!c) return fail(NOMAD_ENOMEM, "nomad_create: out of host memory");
    c->p = *p;
    c->esz = elem_bytes(p->storage);
    c->rvsz = p->storage == NOMAD_STORE_FP64 ? 8 : 4;
    c->ldf = (p->k + 1) & ~1;
    c->sr_seed = uint32_t(p->seed ^ (p->seed >> 32)) ^ 0xA5A5A5A5u;
    c->rng_route.seed(p->seed + 137);
    c->rng_reperm.seed(p->seed + 139);
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, p->device) == hipSuccess && rate_khz > 0)
        c->wall_hz = double(rate_khz) * 1e3;

    auto bail = [&](int rc) {
        delete c;
        return rc;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(NOMAD_EHIP, "nomad_create: stream/event creation failed"));
    // the device address of nomad_eval_noalloc's scratch array, once per device
    if (int rs = resolve_scratch(p->device)) return bail(rs);

    int P = p->nworkers;
    if (p->mode == NOMAD_MODE_ASYNC) {
        // P workers = G workgroups of nw waves, all co-resident: nw is the
        // smallest of 4, 8, 16 that spreads P over the CUs
        int cus = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess)
            return bail(fail(NOMAD_EHIP, "nomad_create: device query failed"));
        if (P <= 0) P = 8 * cus;  // default: two worker waves per SIMD
        int nw = kWaveChoices[2];
        for (int choice : kWaveChoices)
            if (int64_t(choice) * cus >= P) {
                nw = choice;
                break;
            }
        hipError_t oe = dispatch_storage(kernel_code(p->storage, (p->k + 1) & ~1), [&](auto t) {
            return async_occupancy<decltype(t)::value>(nw, p->rounding, &per_cu);
        });
        if (oe != hipSuccess || per_cu < 1) return bail(fail(NOMAD_EHIP, "nomad_create: occupancy query failed"));
        const int gmax = cus * per_cu;
        int G = (P + nw - 1) / nw;
        G = std::max(1, std::min(G, gmax));
        G = std::max(1, std::min(G, (p->m + nw - 1) / nw));
        P = G * nw;
        c->units = G;
        c->nw = nw;
    } else if (P == 0) {
        P = 64;
    }
    if (p->mode == NOMAD_MODE_STRAT) P = std::max(1, std::min(P, p->m));
    c->P = P;
    c->rpb = p->m / P + ((p->m % P) > 0 ? 1 : 0);  // nb_data.cpp:83
    if (p->mode == NOMAD_MODE_STRAT) c->cpb = p->n / P + ((p->n % P) > 0 ? 1 : 0);

    int rc;
    if ((rc = dev_alloc(reinterpret_cast<char**>(&c->dW), size_t(p->m) * c->ldf * c->esz))) return bail(rc);
    if (p->mode == NOMAD_MODE_ASYNC) {
        c->h_coherent = true;
        if ((rc = coh_alloc(p->device, reinterpret_cast<char**>(&c->dH), size_t(p->n) * c->ldf * c->esz)))
            return bail(rc);
    } else if ((rc = dev_alloc(reinterpret_cast<char**>(&c->dH), size_t(p->n) * c->ldf * c->esz))) {
        return bail(rc);
    }
    if (hipMemset(c->dW, 0, size_t(p->m) * c->ldf * c->esz) != hipSuccess ||
        hipMemset(c->dH, 0, size_t(p->n) * c->ldf * c->esz) != hipSuccess)
        return bail(fail(NOMAD_EHIP, "nomad_create: memset failed"));
    *out = c;
    return NOMAD_OK;
}

int nomad_destroy(nomad_ctx* c) {
    if (!c) return NOMAD_OK;
    (void)hipSetDevice(c->p.device);
    (void)hipStreamSynchronize(c->stream);
    delete c;
    return NOMAD_OK;
}

int nomad_get_nworkers(nomad_ctx* c, int32_t* nworkers) {
    if (!c || !nworkers) return fail(NOMAD_EINVAL, "nomad_get_nworkers: null argument");
    *nworkers = c->P;
    return NOMAD_OK;
}

int nomad_get_counters(nomad_ctx* c, int64_t* out, int32_t n) {
    if (!c || !out || n < 0) return fail(NOMAD_EINVAL, "nomad_get_counters: bad argument");
    for (int32_t i = 0; i < std::min<int32_t>(n, kStatWidth); ++i) out[i] = c->last_stats[i];
    for (int64_t i = kStatWidth; i < n && size_t(i - kStatWidth) < c->last_unit_stats.size(); ++i)
        out[i] = c->last_unit_stats[size_t(i - kStatWidth)];
    return NOMAD_OK;
}

int nomad_get_route_state(nomad_ctx* c, int64_t* vptr, int32_t* unit, int32_t* nratings, int32_t* count,
                          int32_t* route_pos, int32_t* tok_pos, int32_t* tok_cycle, int64_t* nentries) {
    if (!c) return fail(NOMAD_EINVAL, "nomad_get_route_state: null context");
    if (c->p.mode != NOMAD_MODE_ASYNC) return fail(NOMAD_ESTATE, "nomad_get_route_state: not an async context");
    if (c->tr.nnz < 0) return fail(NOMAD_ESTATE, "nomad_get_route_state: no training data uploaded");
    NOMAD_HIP_TRY(hipSetDevice(c->p.device));
    NOMAD_HIP_TRY(hipStreamSynchronize(c->stream));
    int rc;
    if (!c->routes_ready && (rc = build_async_state(c))) return rc;
    const int32_t n = c->p.n;
    if (nentries) *nentries = c->V;
    if (vptr) std::copy(c->h_vptr.begin(), c->h_vptr.end(), vptr);
    // entries in workgroup order: entry cslot of the route's hop with that index
    if (unit || nratings || route_pos) {
        for (int32_t j = 0; j < n; ++j)
            for (int64_t e = c->h_vptr[j]; e < c->h_vptr[j + 1]; ++e) {
                const RouteDesc& r = c->h_route[size_t(e)];
                if (unit) unit[r.cslot] = r.unit;
                if (nratings) nratings[r.cslot] = r.len;
                if (route_pos) route_pos[r.cslot] = int32_t(e - c->h_vptr[j]);
            }
    }
    if (count) NOMAD_HIP_TRY(hipMemcpy(count, c->d_vcnt, size_t(c->V) * 4, hipMemcpyDeviceToHost));
    if (tok_pos || tok_cycle) {
        std::vector<TokInfo> tok(static_cast<size_t>(n));
        NOMAD_HIP_TRY(hipMemcpy(tok.data(), c->d_tok, size_t(n) * sizeof(TokInfo), hipMemcpyDeviceToHost));
        for (int32_t j = 0; j < n; ++j) {
            if (tok_pos) tok_pos[j] = tok[j].pos;
            if (tok_cycle) tok_cycle[j] = tok[j].cycle;
        }
    }
    return NOMAD_OK;
}

int nomad_set_visit_counts(nomad_ctx* c, const int32_t* count, int64_t nentries) {
    if (!c || !count) return fail(NOMAD_EINVAL, "nomad_set_visit_counts: null argument");
    if (c->p.mode != NOMAD_MODE_ASYNC) return fail(NOMAD_ESTATE, "nomad_set_visit_counts: not an async context");
    if (c->tr.nnz < 0) return fail(NOMAD_ESTATE, "nomad_set_visit_counts: no training data uploaded");
    NOMAD_HIP_TRY(hipSetDevice(c->p.device));
    NOMAD_HIP_TRY(hipStreamSynchronize(c->stream));
    int rc;
    if (!c->routes_ready && (rc = build_async_state(c))) return rc;
    if (nentries != c->V) return fail(NOMAD_EINVAL, "nomad_set_visit_counts: nentries != the context's pairs");
    for (int64_t e = 0; e < nentries; ++e)
        if (count[e] < 0) return fail(NOMAD_EINVAL, "nomad_set_visit_counts: negative count");
    NOMAD_HIP_TRY(hipMemcpy(c->d_vcnt, count, size_t(nentries) * 4, hipMemcpyHostToDevice));
    return NOMAD_OK;
}

int nomad_get_progress(nomad_ctx* c, int64_t* epochs) {
    if (!c || !epochs) return fail(NOMAD_EINVAL, "nomad_get_progress: null argument");
    *epochs = c->p.mode == NOMAD_MODE_STRAT ? c->strat_epochs : c->cycle_base;
    return NOMAD_OK;
}

int nomad_set_progress(nomad_ctx* c, int64_t epochs) {
    if (!c) return fail(NOMAD_EINVAL, "nomad_set_progress: null context");
    if (epochs < 0 || epochs >= INT_MAX) return fail(NOMAD_EINVAL, "nomad_set_progress: epochs out of range");
    if (c->p.mode == NOMAD_MODE_STRAT) {
        c->strat_epochs = epochs;
        return NOMAD_OK;
    }
    if (c->tr.nnz < 0) return fail(NOMAD_ESTATE, "nomad_set_progress: no training data uploaded");
    NOMAD_HIP_TRY(hipSetDevice(c->p.device));
    NOMAD_HIP_TRY(hipStreamSynchronize(c->stream));
    int rc;
    if (!c->routes_ready && (rc = build_async_state(c))) return rc;
    const int32_t n = c->p.n;
    std::vector<TokInfo> tok(static_cast<size_t>(n));
    NOMAD_HIP_TRY(hipMemcpy(tok.data(), c->d_tok, size_t(n) * sizeof(TokInfo), hipMemcpyDeviceToHost));
    for (auto& t : tok) t.cycle = int32_t(epochs);
    NOMAD_HIP_TRY(hipMemcpy(c->d_tok, tok.data(), size_t(n) * sizeof(TokInfo), hipMemcpyHostToDevice));
    c->cycle_base = epochs;
    return NOMAD_OK;
}

int nomad_upload_train(nomad_ctx* c, const int64_t* colptr, const int32_t* rowidx, const double* val,
                       int64_t nnz) {
    if (!c) return fail(NOMAD_EINVAL, "nomad_upload_train: null context");
    NOMAD_HIP_TRY(hipSetDevice(c->p.device));
    return upload_csc(c, c->tr, colptr, rowidx, val, nnz, true);
}

int nomad_upload_test(nomad_ctx* c, const int64_t* colptr, const int32_t* rowidx, const double* val,
                      int64_t nnz) {
    if (!c) return fail(NOMAD_EINVAL, "nomad_upload_test: null context");
    NOMAD_HIP_TRY(hipSetDevice(c->p.device));
    return upload_csc(c, c->te, colptr, rowidx, val, nnz, false);
}

int nomad_set_factors(nomad_ctx* c, const double* W, const double* H) {
    if (!c) return fail(NOMAD_EINVAL, "nomad_set_factors: null context");
    NOMAD_HIP_TRY(hipSetDevice(c->p.device));
    const int k = c->p.k, ldf = c->ldf;
    for (int which = 0; which < 2; ++which) {
        const double* src = which == 0 ? W : H;
        if (!src) continue;
        const int64_t rows = which == 0 ? c->p.m : c->p.n;
        std::vector<char> buf(size_t(rows) * ldf * c->esz, 0);
        for (int64_t r = 0; r < rows; ++r) {
            for (int t = 0; t < k; ++t) {
                const double v = src[r * k + t];
                const size_t o = size_t(r) * ldf + t;
                if (c->p.storage == NOMAD_STORE_FP64) {
                    reinterpret_cast<double*>(buf.data())[o] = v;
                } else if (c->p.storage == NOMAD_STORE_FP32) {
                    reinterpret_cast<float*>(buf.data())[o] = float(v);
                } else {
                    reinterpret_cast<uint16_t*>(buf.data())[o] = bf16_from_float_rne(float(v));
                }
            }
        }
        NOMAD_HIP_TRY(hipMemcpy(which == 0 ? c->dW : c->dH, buf.data(), buf.size(), hipMemcpyHostToDevice));
    }
    return NOMAD_OK;
}

int nomad_get_factors(nomad_ctx* c, double* W, double* H) {
    if (!c) return fail(NOMAD_EINVAL, "nomad_get_factors: null context");
    NOMAD_HIP_TRY(hipSetDevice(c->p.device));
    NOMAD_HIP_TRY(hipStreamSynchronize(c->stream));
    const int k = c->p.k, ldf = c->ldf;
    for (int which = 0; which < 2; ++which) {
        double* dst = which == 0 ? W : H;
        if (!dst) continue;
        const int64_t rows = which == 0 ? c->p.m : c->p.n;
        std::vector<char> buf(size_t(rows) * ldf * c->esz);
        NOMAD_HIP_TRY(hipMemcpy(buf.data(), which == 0 ? c->dW : c->dH, buf.size(), hipMemcpyDeviceToHost));
        for (int64_t r = 0; r < rows; ++r) {
            for (int t = 0; t < k; ++t) {
                const size_t o = size_t(r) * ldf + t;
                double v;
                if (c->p.storage == NOMAD_STORE_FP64)
                    v = reinterpret_cast<const double*>(buf.data())[o];
                else if (c->p.storage == NOMAD_STORE_FP32)
                    v = reinterpret_cast<const float*>(buf.data())[o];
                else
                    v = float_from_bf16(reinterpret_cast<const uint16_t*>(buf.data())[o]);
                dst[r * k + t] = v;
            }
        }
    }
    return NOMAD_OK;
}

int nomad_train(nomad_ctx* c, double seconds, int64_t epochs, nomad_train_stats* st) {
    if (!c) return fail(NOMAD_EINVAL, "nomad_train: null context");
    if (c->tr.nnz < 0) return fail(NOMAD_ESTATE, "nomad_train: no training data uploaded");
    if (!(seconds > 0) && epochs <= 0) return fail(NOMAD_EINVAL, "nomad_train: need seconds > 0 or epochs > 0");
    NOMAD_HIP_TRY(hipSetDevice(c->p.device));
    if (st) std::memset(st, 0, sizeof(*st));
    if (c->p.mode == NOMAD_MODE_STRAT) return train_strat(c, epochs, st);
    return train_async(c, seconds, epochs, st);
}

int nomad_train_batch(nomad_ctx* const* ctxs, int32_t count, int64_t epochs, int32_t waves,
                      nomad_train_stats* stats) {
    return train_batch(ctxs, count, epochs, waves, stats);
}

namespace {
int eval_both(nomad_ctx* c, nomad_eval_stats* st, bool fixed) {
    NOMAD_HIP_TRY(hipSetDevice(c->p.device));
    std::memset(st, 0, sizeof(*st));
    auto t0 = std::chrono::steady_clock::now();
    int rc;
    if ((rc = eval_one(c, c->tr, &st->train_sse, fixed))) return rc;
    if ((rc = eval_one(c, c->te, &st->test_sse, fixed))) return rc;
    st->train_count = std::max<int64_t>(c->tr.nnz, 0);
    st->test_count = std::max<int64_t>(c->te.nnz, 0);
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return NOMAD_OK;
}
}  // namespace

int nomad_eval(nomad_ctx* c, nomad_eval_stats* st) {
    if (!c || !st) return fail(NOMAD_EINVAL, "nomad_eval: null argument");
    return eval_both(c, st, false);
}

int nomad_eval_noalloc(nomad_ctx* c, nomad_eval_stats* st) {
    if (!c || !st) return fail(NOMAD_EINVAL, "nomad_eval_noalloc: null argument");
    return eval_both(c, st, true);
}

}  // extern "C"
