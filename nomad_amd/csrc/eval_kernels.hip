// Sum of squared errors over a CSC rating matrix (the "test-RMSE SDDMM").
//
// Reference: the test pass of Code/nomad/nb_main.cpp:601-667, which for every
// column token and every stored rating (i, j, r) forms err = <h_j, w_i> - r and
// accumulates err^2 in double; the per-thread sums are added and the RMSE
// printed as sqrt(sse / count) (nb_main.cpp:228-275).
//
// Mapping: wave w owns the rating range [w * chunk, (w + 1) * chunk) of the
// CSC arrays, finds the column of its first rating by binary search over
// colptr and walks forward; h_j is reloaded when the column changes.  The dot
// runs in the storage's arithmetic type with the shared wave tree; err^2 is
// accumulated in FP64 and written as one partial per wave, which the host adds
// in wave order, so the result is deterministic.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernel_args.h"

using namespace nomad_dev;
using nomad_k::SseArgs;

namespace {
constexpr int kEvalPrefetch = 8;
}

template <int S>
__global__ __launch_bounds__(256) void sse_kernel(SseArgs a) {
    using T = Store<S>;
    using E = typename T::elem;
    using Aa = typename T::acc;
    using R = typename T::rval;
    const int lane = lane_id();
    const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t e0 = wave * a.chunk;
    const int64_t e1 = min(a.nnz, e0 + a.chunk);
    double sse = 0.0;
    if (e0 < e1) {
        const E* __restrict__ W = static_cast<const E*>(a.W);
        const E* __restrict__ H = static_cast<const E*>(a.H);
        const R* __restrict__ val = static_cast<const R*>(a.val);
        const int npairs = a.ldf >> 1;
        // column containing e0: last j with colptr[j] <= e0
        int32_t lo = 0, hi = a.n;  // invariant: colptr[lo] <= e0 < colptr[hi]
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (a.colptr[mid] <= e0)
                lo = mid;
            else
                hi = mid;
        }
        int32_t j = lo;
        int64_t cend = a.colptr[j + 1];
        typename T::pair h = T::load(H + int64_t(j) * a.ldf, lane, npairs);
        for (int64_t base = e0; base < e1; base += kWave) {
            const int32_t cnt = int32_t(min<int64_t>(kWave, e1 - base));
            const int32_t myrow = a.rowidx[base + min(lane, cnt - 1)];
            const R myval = val[base + min(lane, cnt - 1)];
            typename T::pair w[kEvalPrefetch];
#pragma unroll
            for (int d = 0; d < kEvalPrefetch; ++d)
                w[d] = T::load(W + int64_t(readlane_i32(myrow, min(d, cnt - 1))) * a.ldf, lane, npairs);
            for (int32_t t0 = 0; t0 < cnt; t0 += kEvalPrefetch) {
#pragma unroll
                for (int d = 0; d < kEvalPrefetch; ++d) {
                    const int32_t t = t0 + d;
                    if (t < cnt) {
                        const int64_t e = base + t;
                        if (e >= cend) {
                            do {
                                ++j;
                                cend = a.colptr[j + 1];
                            } while (e >= cend);
                            h = T::load(H + int64_t(j) * a.ldf, lane, npairs);
                        }
                        const Aa r = Aa(readlane_val(myval, t));
                        const Aa err = wave_dot_p(h, w[d]) - r;
                        sse += double(err) * double(err);
                    }
                    const int32_t tn = min(t + kEvalPrefetch, cnt - 1);
                    w[d] = T::load(W + int64_t(readlane_i32(myrow, tn)) * a.ldf, lane, npairs);
                }
            }
        }
    }
    if (lane == 0) a.partial[wave] = sse;
}

template __global__ void sse_kernel<NOMAD_STORE_FP32>(SseArgs);
template __global__ void sse_kernel<NOMAD_STORE_BF16>(SseArgs);
template __global__ void sse_kernel<NOMAD_STORE_FP64>(SseArgs);
template __global__ void sse_kernel<NOMAD_STORE_FP32 | kWideBit>(SseArgs);
template __global__ void sse_kernel<NOMAD_STORE_BF16 | kWideBit>(SseArgs);
template __global__ void sse_kernel<NOMAD_STORE_FP64 | kWideBit>(SseArgs);
