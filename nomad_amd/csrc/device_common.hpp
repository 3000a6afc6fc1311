// Device-side building blocks shared by the SGD, eval and gradient kernels.
//
// Factor layout: a factor row of k <= 128 elements is spread over the 64 lanes
// of a wavefront as element pairs, lane l holding elements (2l, 2l+1).  One
// row is therefore one coalesced load per wave: 8 B per lane in FP32, 4 B per
// lane in BF16 (two packed bf16), 16 B per lane in FP64.  Lanes past the row
// hold zeros and never store.
//
// Wave dot product: each lane forms p = x0*y0 then p = fma(x1, y1, p); the 64
// partials are summed by a fixed tree that is reproduced bit-for-bit by the
// CPU oracle (oracle/nomad_oracle.cpp, gpu_dot):
//   Q_q   = (p[4q] + p[4q+1]) + (p[4q+2] + p[4q+3])          (quad_perm DPP)
//   R_r   = (Q_{4r+3} + Q_{4r+2}) + (Q_{4r+1} + Q_{4r})        (row_shr:4, :8)
//   total = (R_3 + R_2) + (R_1 + R_0)                          (row_bcast:15, :31)
// read back from lane 63.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nomad_hip.h"

namespace nomad_dev {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// ---- DPP helpers (gfx9 encodings) ----
// quad_perm [1,0,3,2] = 0xB1, quad_perm [2,3,0,1] = 0x4E, row_shr:4 = 0x114,
// row_shr:8 = 0x118, row_bcast:15 = 0x142, row_bcast:31 = 0x143.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f32(float x) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, ROW_MASK, 0xF, true));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const int lo = __builtin_amdgcn_update_dpp(0, int(uint32_t(u)), CTRL, ROW_MASK, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, int(uint32_t(u >> 32)), CTRL, ROW_MASK, 0xF, true);
    return __builtin_bit_cast(double, (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

// The tree of the two broadcast steps is written with every row enabled:
// rows 0 and 2 (row_bcast:15) and rows 0 and 1 (row_bcast:31) then also take
// an addition, but no lane that lane 63's total reads is changed before it is
// read, so lane 63 holds the same bits as with row masks 0xA and 0xC, and each
// step is one DPP add instead of a move, a DPP move and an add.
__device__ __forceinline__ float wave_tree_total(float x) {
    x = x + dpp_f32<0xB1, 0xF>(x);
    x = x + dpp_f32<0x4E, 0xF>(x);
    x = x + dpp_f32<0x114, 0xF>(x);
    x = x + dpp_f32<0x118, 0xF>(x);
    x = x + dpp_f32<0x142, 0xF>(x);
    x = x + dpp_f32<0x143, 0xF>(x);
    return x;  // the total in lane 63
}

__device__ __forceinline__ float wave_tree_sum(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, wave_tree_total(x)), 63));
}

__device__ __forceinline__ double wave_tree_sum(double x) {
    x = x + dpp_f64<0xB1, 0xF>(x);
    x = x + dpp_f64<0x4E, 0xF>(x);
    x = x + dpp_f64<0x114, 0xF>(x);
    x = x + dpp_f64<0x118, 0xF>(x);
    x = x + dpp_f64<0x142, 0xA>(x);
    x = x + dpp_f64<0x143, 0xC>(x);
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readlane(int(uint32_t(u)), 63);
    const uint32_t hi = __builtin_amdgcn_readlane(int(uint32_t(u >> 32)), 63);
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// ---- bf16 <-> f32 ----
__device__ __forceinline__ float bf16_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t u) { return __uint_as_float(u & 0xFFFF0000u); }

// Round an element pair to the nearest bf16 (ties to even) and pack it, x in
// the low half: one v_cvt_pk_bf16_f32 on gfx950.  For every finite input it
// equals adding 0x7FFF plus the kept lsb below the cut and truncating (the
// oracle's form; denormals are preserved in these kernels); a NaN comes out
// as a quiet NaN.
typedef __bf16 nomad_bf16x2 __attribute__((ext_vector_type(2)));
typedef float nomad_f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf16_pack_rne(float x, float y) {
    const nomad_f32x2 v = {x, y};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, nomad_bf16x2));
}

// Stochastic rounding: add 16 uniform random bits below the bf16 cut, truncate.
__device__ __forceinline__ uint32_t f32_to_bf16_sr(float f, uint32_t r16) {
    uint32_t u = __float_as_uint(f);
    u += (r16 & 0xFFFFu);
    return u >> 16;
}

// 32-bit avalanche mixer (lowbias32) used to derive stochastic-rounding bits
// from (seed, row/column, update counter, lane).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Stochastic-rounding bits of element pair `lane` of row (or column) a at
// update key b: a hash of (seed, a, b), which is the same for the whole wave
// and runs on the scalar unit, XOR a per-lane key fixed for the wave.  Each
// lane's bits are as uniform as the hash; the vector cost per store is one
// XOR.
__device__ __forceinline__ uint32_t sr_lane_key(uint32_t seed, uint32_t lane) {
    return mix32(seed ^ ((lane + 1u) * 0x85EBCA6Bu));
}
__device__ __forceinline__ uint32_t sr_bits_k(uint32_t seed, uint32_t a, uint32_t b, uint32_t lkey) {
    return mix32(seed ^ mix32(a * 0x9E3779B1u + b)) ^ lkey;
}
__device__ __forceinline__ uint32_t sr_bits(uint32_t seed, uint32_t a, uint32_t b, uint32_t lane) {
    return sr_bits_k(seed, a, b, sr_lane_key(seed, lane));
}

// ---- storage traits: load/store one factor row as per-lane pairs ----
// Rows have an even leading dimension ldf >= k; the padding element (k odd)
// is stored as zero and stays zero under the update.  npairs = ldf / 2 lanes
// carry data; the others read the all-zero block g_zero_pairs and never
// store.  A load is an address select plus one memory instruction, with no
// data-dependent select after it, so several rows can be in flight.
__device__ __align__(16) static double g_zero_pairs[2 * kWave];  // zero-initialised, never written

// Buffer resource over one factor row with `bytes` valid bytes (gfx9 raw
// buffer: stride 0, dword data format).  Accesses past `bytes` are dropped
// (stores) or read as zero (loads) by the hardware's range check, so lanes
// beyond the row need neither a branch nor an address select, and the
// compiler counts every load and store exactly when it places its waits.
// The row address must be wave-uniform (it lives in scalar registers): it
// and the size are read from the first lane, since a descriptor the
// compiler builds in vector registers (from values it cannot prove
// uniform) makes it wrap every access in a loop over distinct values.
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* row, int bytes) {
    const uint64_t v = reinterpret_cast<uint64_t>(row);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(int(uint32_t(v)));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(int(uint32_t(v >> 32)));
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>((uint64_t(hi) << 32) | lo), 0,
                                             __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}

// Cache-policy operand of a buffer access that makes it device-coherent:
// the sc1 bit, the encoding the compiler gives relaxed agent-scope atomic
// loads and stores on gfx950 (see ld_dc / st_dc below).
constexpr int kCpDevice = 16;

// Device-coherent accesses: relaxed agent-scope atomics, i.e. loads and
// stores with the sc1 bit, which are coherent across the XCDs' separate L2s
// without the L2 writeback / invalidate that an agent-scope fence costs.
template <class T>
__device__ __forceinline__ T ld_dc(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_dc(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until every memory operation this wave has issued has completed
// (s_waitcnt vmcnt(0) expcnt(0) lgkmcnt(0)); with device-coherent stores that
// means they are visible to every XCD.  The signal fences keep the compiler
// from moving memory operations across the wait.
__device__ __forceinline__ void wait_all_mem() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// s_waitcnt vmcnt(0) alone (expcnt and lgkmcnt at their maxima, 7 and 15).
// Placed right after loads whose registers a later shared loop reads, it
// retires them there, so that the loop's own waits need not cover them.
__device__ __forceinline__ void wait_vm() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <int STORE>
struct Store;

template <>
struct Store<NOMAD_STORE_FP32> {
    using elem = float;
    using acc = float;
    using rval = float;
    struct pair {
        float x, y;
    };
    template <int CP = 0>
    static __device__ __forceinline__ pair load(const elem* row, int lane, int npairs) {
        const u32x2_t v = __builtin_amdgcn_raw_buffer_load_b64(row_rsrc(row, npairs * 8), lane * 8, 0, CP);
        return pair{__uint_as_float(v.x), __uint_as_float(v.y)};
    }
    template <int CP = 0>
    static __device__ __forceinline__ void store(elem* row, int lane, int npairs, pair p, uint32_t,
                                                 int) {
        const u32x2_t v = {__float_as_uint(p.x), __float_as_uint(p.y)};
        __builtin_amdgcn_raw_buffer_store_b64(v, row_rsrc(row, npairs * 8), lane * 8, 0, CP);
    }
    // rows of a block through one buffer resource: byte offset voff (per
    // lane, out of range for lanes past the row) + soff (row within block)
    static constexpr int kPairBytes = 8;

    // ---- layout helpers shared with the wide layout (see Wide below) ----
    static constexpr int kNP = 1;  // element pairs per lane
    using voff_t = int;            // per-lane byte offset of a row's pair(s) in a buffer
    static __device__ __forceinline__ voff_t lane_voff(int lane, int npairs) {
        return int(lane < npairs ? uint32_t(lane * kPairBytes) : 0x80000000u);
    }
    // every lane out of range (a store the range check drops); d selects
    // one of 64 distinct such offsets
    static __device__ __forceinline__ voff_t drop_voff(voff_t v, int d) {
        return int(uint32_t(v) + 0x7FFFFFFFu) - d * 1024;
    }
    // the lane's pair(s) at a given byte offset
    static __device__ __forceinline__ voff_t voff_bytes(int off) { return off; }
    static __device__ __forceinline__ pair zero() { return pair{acc(0), acc(0)}; }
    static __device__ __forceinline__ pair select(bool c, pair a, pair b) {
        return pair{c ? a.x : b.x, c ? a.y : b.y};
    }
    // a column factor held in LDS as acc[2 * 64 * kNP]
    static __device__ __forceinline__ pair lds_get(const acc* base, int lane) {
        return pair{base[2 * lane], base[2 * lane + 1]};
    }
    static __device__ __forceinline__ void lds_put(acc* base, int lane, pair p) {
        base[2 * lane] = p.x;
        base[2 * lane + 1] = p.y;
    }
    static __device__ __forceinline__ pair load_b(__amdgpu_buffer_rsrc_t r, int voff, int soff = 0) {
        const u32x2_t v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        return pair{__uint_as_float(v.x), __uint_as_float(v.y)};
    }
    // a loaded pair as it sits in registers before use (see Store<BF16>)
    using raw = pair;
    template <int CP = 0>
    static __device__ __forceinline__ raw load_raw_b(__amdgpu_buffer_rsrc_t r, int voff, int soff = 0) {
        const u32x2_t v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, CP);
        return pair{__uint_as_float(v.x), __uint_as_float(v.y)};
    }
    static __device__ __forceinline__ pair unpack(raw v) { return v; }
    // the stored form of a pair (rounding for bf16) and its store
    static __device__ __forceinline__ raw pack(pair p, uint32_t, int) { return p; }
    template <int CP = 0>
    static __device__ __forceinline__ void store_raw_b(__amdgpu_buffer_rsrc_t r, int voff, raw p, int soff = 0) {
        const u32x2_t v = {__float_as_uint(p.x), __float_as_uint(p.y)};
        __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, CP);
    }
    static __device__ __forceinline__ void store_b(__amdgpu_buffer_rsrc_t r, int voff, int soff, pair p, uint32_t,
                                                   int) {
        const u32x2_t v = {__float_as_uint(p.x), __float_as_uint(p.y)};
        __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, 0);
    }
    // device-coherent (sc1) variants for rows that move between workers
    static __device__ __forceinline__ pair load_dc(const elem* row, int lane, int npairs) {
        const uint64_t* src = lane < npairs ? reinterpret_cast<const uint64_t*>(row)
                                            : reinterpret_cast<const uint64_t*>(g_zero_pairs);
        const uint64_t u = ld_dc(src + lane);
        return pair{__uint_as_float(uint32_t(u)), __uint_as_float(uint32_t(u >> 32))};
    }
    static __device__ __forceinline__ void store_dc(elem* row, int lane, int npairs, pair p, uint32_t,
                                                    int) {
        if (lane < npairs)
            st_dc(reinterpret_cast<uint64_t*>(row) + lane,
                  (uint64_t(__float_as_uint(p.y)) << 32) | __float_as_uint(p.x));
    }
};

template <>
struct Store<NOMAD_STORE_FP64> {
    using elem = double;
    using acc = double;
    using rval = double;
    struct pair {
        double x, y;
    };
    template <int CP = 0>
    static __device__ __forceinline__ pair load(const elem* row, int lane, int npairs) {
        const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(row_rsrc(row, npairs * 16), lane * 16, 0, CP);
        return pair{__builtin_bit_cast(double, (uint64_t(v.y) << 32) | v.x),
                    __builtin_bit_cast(double, (uint64_t(v.w) << 32) | v.z)};
    }
    template <int CP = 0>
    static __device__ __forceinline__ void store(elem* row, int lane, int npairs, pair p, uint32_t,
                                                 int) {
        const uint64_t a = __builtin_bit_cast(uint64_t, p.x), b = __builtin_bit_cast(uint64_t, p.y);
        const u32x4_t v = {uint32_t(a), uint32_t(a >> 32), uint32_t(b), uint32_t(b >> 32)};
        __builtin_amdgcn_raw_buffer_store_b128(v, row_rsrc(row, npairs * 16), lane * 16, 0, CP);
    }
    static constexpr int kPairBytes = 16;

    // ---- layout helpers shared with the wide layout (see Wide below) ----
    static constexpr int kNP = 1;  // element pairs per lane
    using voff_t = int;            // per-lane byte offset of a row's pair(s) in a buffer
    static __device__ __forceinline__ voff_t lane_voff(int lane, int npairs) {
        return int(lane < npairs ? uint32_t(lane * kPairBytes) : 0x80000000u);
    }
    // every lane out of range (a store the range check drops); d selects
    // one of 64 distinct such offsets
    static __device__ __forceinline__ voff_t drop_voff(voff_t v, int d) {
        return int(uint32_t(v) + 0x7FFFFFFFu) - d * 1024;
    }
    // the lane's pair(s) at a given byte offset
    static __device__ __forceinline__ voff_t voff_bytes(int off) { return off; }
    static __device__ __forceinline__ pair zero() { return pair{acc(0), acc(0)}; }
    static __device__ __forceinline__ pair select(bool c, pair a, pair b) {
        return pair{c ? a.x : b.x, c ? a.y : b.y};
    }
    // a column factor held in LDS as acc[2 * 64 * kNP]
    static __device__ __forceinline__ pair lds_get(const acc* base, int lane) {
        return pair{base[2 * lane], base[2 * lane + 1]};
    }
    static __device__ __forceinline__ void lds_put(acc* base, int lane, pair p) {
        base[2 * lane] = p.x;
        base[2 * lane + 1] = p.y;
    }
    static __device__ __forceinline__ pair load_b(__amdgpu_buffer_rsrc_t r, int voff, int soff = 0) {
        const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
        return pair{__builtin_bit_cast(double, (uint64_t(v.y) << 32) | v.x),
                    __builtin_bit_cast(double, (uint64_t(v.w) << 32) | v.z)};
    }
    using raw = pair;
    template <int CP = 0>
    static __device__ __forceinline__ raw load_raw_b(__amdgpu_buffer_rsrc_t r, int voff, int soff = 0) {
        const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, CP);
        return pair{__builtin_bit_cast(double, (uint64_t(v.y) << 32) | v.x),
                    __builtin_bit_cast(double, (uint64_t(v.w) << 32) | v.z)};
    }
    static __device__ __forceinline__ pair unpack(raw v) { return v; }
    // the stored form of a pair (rounding for bf16) and its store
    static __device__ __forceinline__ raw pack(pair p, uint32_t, int) { return p; }
    template <int CP = 0>
    static __device__ __forceinline__ void store_raw_b(__amdgpu_buffer_rsrc_t r, int voff, raw p, int soff = 0) {
        const uint64_t a = __builtin_bit_cast(uint64_t, p.x), b = __builtin_bit_cast(uint64_t, p.y);
        const u32x4_t v = {uint32_t(a), uint32_t(a >> 32), uint32_t(b), uint32_t(b >> 32)};
        __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, CP);
    }
    static __device__ __forceinline__ void store_b(__amdgpu_buffer_rsrc_t r, int voff, int soff, pair p, uint32_t,
                                                   int) {
        const uint64_t a = __builtin_bit_cast(uint64_t, p.x), b = __builtin_bit_cast(uint64_t, p.y);
        const u32x4_t v = {uint32_t(a), uint32_t(a >> 32), uint32_t(b), uint32_t(b >> 32)};
        __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, 0);
    }
    static __device__ __forceinline__ pair load_dc(const elem* row, int lane, int npairs) {
        const uint64_t* src = lane < npairs ? reinterpret_cast<const uint64_t*>(row)
                                            : reinterpret_cast<const uint64_t*>(g_zero_pairs);
        const uint64_t a = ld_dc(src + 2 * lane);
        const uint64_t b = ld_dc(src + 2 * lane + 1);
        return pair{__builtin_bit_cast(double, a), __builtin_bit_cast(double, b)};
    }
    static __device__ __forceinline__ void store_dc(elem* row, int lane, int npairs, pair p, uint32_t,
                                                    int) {
        if (lane < npairs) {
            st_dc(reinterpret_cast<uint64_t*>(row) + 2 * lane, __builtin_bit_cast(uint64_t, p.x));
            st_dc(reinterpret_cast<uint64_t*>(row) + 2 * lane + 1, __builtin_bit_cast(uint64_t, p.y));
        }
    }
};

template <>
struct Store<NOMAD_STORE_BF16> {
    using elem = uint16_t;
    using acc = float;
    using rval = float;
    struct pair {
        float x, y;
    };
    // element pair (2l, 2l+1) is one aligned 32-bit word
    template <int CP = 0>
    static __device__ __forceinline__ pair load(const elem* row, int lane, int npairs) {
        const uint32_t u = __builtin_amdgcn_raw_buffer_load_b32(row_rsrc(row, npairs * 4), lane * 4, 0, CP);
        return pair{bf16_lo(u), bf16_hi(u)};
    }
    template <int CP = 0>
    static __device__ __forceinline__ void store(elem* row, int lane, int npairs, pair p, uint32_t rnd,
                                                 int round_mode) {
        const uint32_t word = round_mode == NOMAD_ROUND_STOCHASTIC
                                  ? f32_to_bf16_sr(p.x, rnd) | (f32_to_bf16_sr(p.y, rnd >> 16) << 16)
                                  : bf16_pack_rne(p.x, p.y);
        __builtin_amdgcn_raw_buffer_store_b32(word, row_rsrc(row, npairs * 4), lane * 4, 0, CP);
    }
    static constexpr int kPairBytes = 4;

    // ---- layout helpers shared with the wide layout (see Wide below) ----
    static constexpr int kNP = 1;  // element pairs per lane
    using voff_t = int;            // per-lane byte offset of a row's pair(s) in a buffer
    static __device__ __forceinline__ voff_t lane_voff(int lane, int npairs) {
        return int(lane < npairs ? uint32_t(lane * kPairBytes) : 0x80000000u);
    }
    // every lane out of range (a store the range check drops); d selects
    // one of 64 distinct such offsets
    static __device__ __forceinline__ voff_t drop_voff(voff_t v, int d) {
        return int(uint32_t(v) + 0x7FFFFFFFu) - d * 1024;
    }
    // the lane's pair(s) at a given byte offset
    static __device__ __forceinline__ voff_t voff_bytes(int off) { return off; }
    static __device__ __forceinline__ pair zero() { return pair{acc(0), acc(0)}; }
    static __device__ __forceinline__ pair select(bool c, pair a, pair b) {
        return pair{c ? a.x : b.x, c ? a.y : b.y};
    }
    // a column factor held in LDS as acc[2 * 64 * kNP]
    static __device__ __forceinline__ pair lds_get(const acc* base, int lane) {
        return pair{base[2 * lane], base[2 * lane + 1]};
    }
    static __device__ __forceinline__ void lds_put(acc* base, int lane, pair p) {
        base[2 * lane] = p.x;
        base[2 * lane + 1] = p.y;
    }
    static __device__ __forceinline__ pair load_b(__amdgpu_buffer_rsrc_t r, int voff, int soff = 0) {
        const uint32_t u = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
        return pair{bf16_lo(u), bf16_hi(u)};
    }
    // Prefetched rows stay packed until used: widening right after the load
    // would make the compiler wait for the load there, which defeats the
    // prefetch.
    using raw = uint32_t;
    template <int CP = 0>
    static __device__ __forceinline__ raw load_raw_b(__amdgpu_buffer_rsrc_t r, int voff, int soff = 0) {
        return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, CP);
    }
    static __device__ __forceinline__ pair unpack(raw u) { return pair{bf16_lo(u), bf16_hi(u)}; }
    // (stochastic: f32_to_bf16_sr of both elements, the two upper halves
    // joined by one byte permute)
    static __device__ __forceinline__ raw pack(pair p, uint32_t rnd, int round_mode) {
        if (round_mode == NOMAD_ROUND_STOCHASTIC) {
            const uint32_t lo = __float_as_uint(p.x) + (rnd & 0xFFFFu);
            const uint32_t hi = __float_as_uint(p.y) + (rnd >> 16);
            return __builtin_amdgcn_perm(hi, lo, 0x07060302u);
        }
        return bf16_pack_rne(p.x, p.y);
    }
    template <int CP = 0>
    static __device__ __forceinline__ void store_raw_b(__amdgpu_buffer_rsrc_t r, int voff, raw v, int soff = 0) {
        __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, CP);
    }
    static __device__ __forceinline__ void store_b(__amdgpu_buffer_rsrc_t r, int voff, int soff, pair p,
                                                   uint32_t rnd, int round_mode) {
        const uint32_t word = round_mode == NOMAD_ROUND_STOCHASTIC
                                  ? f32_to_bf16_sr(p.x, rnd) | (f32_to_bf16_sr(p.y, rnd >> 16) << 16)
                                  : bf16_pack_rne(p.x, p.y);
        __builtin_amdgcn_raw_buffer_store_b32(word, r, voff, soff, 0);
    }
    static __device__ __forceinline__ pair load_dc(const elem* row, int lane, int npairs) {
        const uint32_t* src = lane < npairs ? reinterpret_cast<const uint32_t*>(row)
                                            : reinterpret_cast<const uint32_t*>(g_zero_pairs);
        const uint32_t u = ld_dc(src + lane);
        return pair{bf16_lo(u), bf16_hi(u)};
    }
    static __device__ __forceinline__ void store_dc(elem* row, int lane, int npairs, pair p, uint32_t rnd,
                                                    int round_mode) {
        const uint32_t word = round_mode == NOMAD_ROUND_STOCHASTIC
                                  ? f32_to_bf16_sr(p.x, rnd) | (f32_to_bf16_sr(p.y, rnd >> 16) << 16)
                                  : bf16_pack_rne(p.x, p.y);
        if (lane < npairs) st_dc(reinterpret_cast<uint32_t*>(row) + lane, word);
    }
};

// ---- rows of more than 128 elements: two element pairs per lane ----
// Lane l holds element pairs l and l + 64 of the row (elements 2l, 2l + 1
// and 128 + 2l, 128 + 2l + 1), each stored as in the one-pair layout, so a
// row is two coalesced accesses per wave; ldf <= 256.  The second pair is
// the first pair's layout applied to the row from element 128 on.  A lane's
// partial of a dot product adds the second pair's two products after the
// first pair's.  Stochastic rounding draws the second pair's bits as
// mix32(bits of the first pair ^ kWideSalt).
constexpr int kWideBit = 16;  // storage code | kWideBit selects this layout
constexpr uint32_t kWideSalt = 0x5BD1E995u;
__device__ __forceinline__ uint32_t wide_bits(uint32_t rnd) { return mix32(rnd ^ kWideSalt); }

template <class BP>
struct WPair {
    BP a, b;
};
struct WVoff {
    int a, b;
};

template <class B>
struct Wide {
    using elem = typename B::elem;
    using acc = typename B::acc;
    using rval = typename B::rval;
    using pair = WPair<typename B::pair>;
    using raw = WPair<typename B::raw>;
    using voff_t = WVoff;
    static constexpr int kNP = 2;
    static constexpr int kPairBytes = B::kPairBytes;        // per element pair
    static constexpr int kHalfBytes = kWave * B::kPairBytes;  // pair l to pair l + 64
    static constexpr int kHalfElems = 2 * kWave;

    template <int CP = 0>
    static __device__ __forceinline__ pair load(const elem* row, int lane, int npairs) {
        return pair{B::template load<CP>(row, lane, npairs),
                    B::template load<CP>(row + kHalfElems, lane, npairs - kWave)};
    }
    template <int CP = 0>
    static __device__ __forceinline__ void store(elem* row, int lane, int npairs, pair p, uint32_t rnd, int rm) {
        B::template store<CP>(row, lane, npairs, p.a, rnd, rm);
        B::template store<CP>(row + kHalfElems, lane, npairs - kWave, p.b, wide_bits(rnd), rm);
    }
    static __device__ __forceinline__ voff_t lane_voff(int lane, int npairs) {
        return voff_t{B::lane_voff(lane, npairs), int(lane + kWave < npairs ? uint32_t((lane + kWave) * kPairBytes)
                                                                             : 0x80000000u)};
    }
    static __device__ __forceinline__ voff_t drop_voff(voff_t v, int d) {
        return voff_t{B::drop_voff(v.a, d), B::drop_voff(v.b, d)};
    }
    static __device__ __forceinline__ voff_t voff_bytes(int off) { return voff_t{off, off + kHalfBytes}; }
    template <int CP = 0>
    static __device__ __forceinline__ raw load_raw_b(__amdgpu_buffer_rsrc_t r, voff_t v, int soff = 0) {
        return raw{B::template load_raw_b<CP>(r, v.a, soff), B::template load_raw_b<CP>(r, v.b, soff)};
    }
    static __device__ __forceinline__ pair unpack(raw u) { return pair{B::unpack(u.a), B::unpack(u.b)}; }
    static __device__ __forceinline__ raw pack(pair p, uint32_t rnd, int rm) {
        return raw{B::pack(p.a, rnd, rm), B::pack(p.b, wide_bits(rnd), rm)};
    }
    template <int CP = 0>
    static __device__ __forceinline__ void store_raw_b(__amdgpu_buffer_rsrc_t r, voff_t v, raw p, int soff = 0) {
        B::template store_raw_b<CP>(r, v.a, p.a, soff);
        B::template store_raw_b<CP>(r, v.b, p.b, soff);
    }
    static __device__ __forceinline__ void store_b(__amdgpu_buffer_rsrc_t r, voff_t v, int soff, pair p,
                                                   uint32_t rnd, int rm) {
        B::store_b(r, v.a, soff, p.a, rnd, rm);
        B::store_b(r, v.b, soff, p.b, wide_bits(rnd), rm);
    }
    static __device__ __forceinline__ pair load_dc(const elem* row, int lane, int npairs) {
        return pair{B::load_dc(row, lane, npairs), B::load_dc(row + kHalfElems, lane, npairs - kWave)};
    }
    static __device__ __forceinline__ void store_dc(elem* row, int lane, int npairs, pair p, uint32_t rnd,
                                                    int rm) {
        B::store_dc(row, lane, npairs, p.a, rnd, rm);
        B::store_dc(row + kHalfElems, lane, npairs - kWave, p.b, wide_bits(rnd), rm);
    }
    static __device__ __forceinline__ pair zero() { return pair{B::zero(), B::zero()}; }
    static __device__ __forceinline__ pair select(bool c, pair x, pair y) {
        return pair{B::select(c, x.a, y.a), B::select(c, x.b, y.b)};
    }
    static __device__ __forceinline__ pair lds_get(const acc* base, int lane) {
        return pair{B::lds_get(base, lane), B::lds_get(base + kHalfElems, lane)};
    }
    static __device__ __forceinline__ void lds_put(acc* base, int lane, pair p) {
        B::lds_put(base, lane, p.a);
        B::lds_put(base + kHalfElems, lane, p.b);
    }
};
template <>
struct Store<NOMAD_STORE_FP32 | kWideBit> : Wide<Store<NOMAD_STORE_FP32>> {};
template <>
struct Store<NOMAD_STORE_BF16 | kWideBit> : Wide<Store<NOMAD_STORE_BF16>> {};
template <>
struct Store<NOMAD_STORE_FP64 | kWideBit> : Wide<Store<NOMAD_STORE_FP64>> {};

// the storage type of a kernel's storage code (without the layout bit)
constexpr int base_storage(int s) { return s & ~kWideBit; }

// Fused multiply-add in the arithmetic type itself (a single rounding in
// FP32 for float, in FP64 for double).
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// The asynchronous kernel's step, written as w' = c1 w + c2 h,
// h' = c1 h + c2 w (old w) with c1 = 1 - step * reg fixed per visit and
// c2 = -step * err: four operations per element instead of six (one
// rounding less); its results are compared by RMSE and not bit for bit.
// The rating enters the dot product's tree: `sel63`
// is 1 in lane 63 and 0 elsewhere, so lane 63's partial starts from -r, its
// tree total is the error (dot - r) itself, and one readlane brings it to the
// scalar unit.
template <class P, class A>
__device__ __forceinline__ void sgd_step_b(P& w, P& h, A r, A sel63, A c1, A nstep) {
    A p = fma_t(h.x, w.x, -r * sel63);
    p = fma_t(h.y, w.y, p);
    const A c2 = nstep * wave_tree_sum(p);
    const P w0 = w;
    w.x = fma_t(c2, h.x, c1 * w.x);
    w.y = fma_t(c2, h.y, c1 * w.y);
    h.x = fma_t(c2, w0.x, c1 * h.x);
    h.y = fma_t(c2, w0.y, c1 * h.y);
}

// One SGD step on (w_i, h_j) given the rating r.  Per lane, elements x and y:
//   p   = h.x*w.x ; p = fma(h.y, w.y, p) ; dot = tree(p) ; err = dot - r
//   a   = fma(err, h, reg*w)     b = fma(err, w, reg*h)
//   w'  = fma(-step, a, w)       h' = fma(-step, b, h)
// This is the reference update of Code/nomad/nb_main.cpp:528-538 (w and h
// updated simultaneously from their old values) in fused form.
template <class P, class A>
__device__ __forceinline__ A sgd_step(P& w, P& h, A r, A step, A reg) {
    A p = h.x * w.x;
    p = fma_t(h.y, w.y, p);
    const A dot = wave_tree_sum(p);
    const A err = dot - r;
    const A ax = fma_t(err, h.x, reg * w.x);
    const A ay = fma_t(err, h.y, reg * w.y);
    const A bx = fma_t(err, w.x, reg * h.x);
    const A by = fma_t(err, w.y, reg * h.y);
    w.x = fma_t(-step, ax, w.x);
    w.y = fma_t(-step, ay, w.y);
    h.x = fma_t(-step, bx, h.x);
    h.y = fma_t(-step, by, h.y);
    return err;
}

template <class A>
__device__ __forceinline__ A wave_dot(A hx, A hy, A wx, A wy) {
    A p = hx * wx;
    p = fma_t(hy, wy, p);
    return wave_tree_sum(p);
}

// The same three functions for two pairs per lane: the partial takes the
// second pair's products after the first's; the updates are per element.
template <class BP, class A>
__device__ __forceinline__ void sgd_step_b(WPair<BP>& w, WPair<BP>& h, A r, A sel63, A c1, A nstep) {
    A p = fma_t(h.a.x, w.a.x, -r * sel63);
    p = fma_t(h.a.y, w.a.y, p);
    p = fma_t(h.b.x, w.b.x, p);
    p = fma_t(h.b.y, w.b.y, p);
    const A c2 = nstep * wave_tree_sum(p);
    const WPair<BP> w0 = w;
    w.a.x = fma_t(c2, h.a.x, c1 * w.a.x);
    w.a.y = fma_t(c2, h.a.y, c1 * w.a.y);
    w.b.x = fma_t(c2, h.b.x, c1 * w.b.x);
    w.b.y = fma_t(c2, h.b.y, c1 * w.b.y);
    h.a.x = fma_t(c2, w0.a.x, c1 * h.a.x);
    h.a.y = fma_t(c2, w0.a.y, c1 * h.a.y);
    h.b.x = fma_t(c2, w0.b.x, c1 * h.b.x);
    h.b.y = fma_t(c2, w0.b.y, c1 * h.b.y);
}

template <class BP, class A>
__device__ __forceinline__ A sgd_step(WPair<BP>& w, WPair<BP>& h, A r, A step, A reg) {
    A p = h.a.x * w.a.x;
    p = fma_t(h.a.y, w.a.y, p);
    p = fma_t(h.b.x, w.b.x, p);
    p = fma_t(h.b.y, w.b.y, p);
    const A err = wave_tree_sum(p) - r;
    auto upd = [&](A& wv, A& hv) {
        const A aa = fma_t(err, hv, reg * wv);
        const A bb = fma_t(err, wv, reg * hv);
        wv = fma_t(-step, aa, wv);
        hv = fma_t(-step, bb, hv);
    };
    upd(w.a.x, h.a.x);
    upd(w.a.y, h.a.y);
    upd(w.b.x, h.b.x);
    upd(w.b.y, h.b.y);
    return err;
}

// dot product of a lane's pair(s) over the wave
template <class P>
__device__ __forceinline__ auto wave_dot_p(const P& h, const P& w) {
    return wave_dot(h.x, h.y, w.x, w.y);
}
template <class BP>
__device__ __forceinline__ auto wave_dot_p(const WPair<BP>& h, const WPair<BP>& w) {
    auto p = h.a.x * w.a.x;
    p = fma_t(h.a.y, w.a.y, p);
    p = fma_t(h.b.x, w.b.x, p);
    p = fma_t(h.b.y, w.b.y, p);
    return wave_tree_sum(p);
}

__device__ __forceinline__ int readlane_i32(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

// A pointer into device memory as a global-address-space pointer.  Pointers
// read from memory (rather than passed as kernel arguments) are generic, and
// loads through them are flat loads; a flat load may complete out of order
// with LDS accesses, so every wait the compiler places for one is vmcnt(0)
// and lgkmcnt(0), which serialises loads issued ahead.  Loads through the
// converted pointer are global loads with exact wait counts.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> gbl(const T* p) {
    return (gptr<T>)p;
}
__device__ __forceinline__ float readlane_f32(float x, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}
__device__ __forceinline__ double readlane_f64(double x, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readlane(int(uint32_t(u)), l);
    const uint32_t hi = __builtin_amdgcn_readlane(int(uint32_t(u >> 32)), l);
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
__device__ __forceinline__ float readlane_val(float x, int l) { return readlane_f32(x, l); }
__device__ __forceinline__ double readlane_val(double x, int l) { return readlane_f64(x, l); }
}  // namespace nomad_dev
