// Objective and gradient of the factorisation on the GPU:
//   L = sum_{(i,j)} (w_i.h_j - r_ij)^2 + reg (sum_i deg(i) |w_i|^2 + sum_j deg(j) |h_j|^2)
//   dL/dw_i = 2 sum_j e_ij h_j + 2 reg deg(i) w_i,   e_ij = w_i.h_j - r_ij
//   dL/dh_j = 2 sum_i e_ij w_i + 2 reg deg(j) h_j
// (the objective the reference's competitors print, calobj_mpi,
// Code/competitor/sgd-mpi-omp.cpp:122-164; it equals the NOMAD loss with the
// per-rating regulariser summed).  W, H and the gradients are device arrays
// (m x ldf, n x ldf row-major, ldf = k rounded up to even) owned by the caller
// -- typically torch tensors -- and the kernels run on the caller's stream.
//
// dL/dH and the squared errors come from a pass over the CSC, dL/dW from a
// pass over a CSR copy; every column (row) is split into tasks of at most
// kTask ratings, one wave per task, a wave holding the line's factor as lane
// pairs and walking its ratings with the other side's rows prefetched.  The
// partial sums of a line are added in task order, the squared errors in a
// fixed tree: the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "nomad_hip.h"

namespace nomad_host {
int fail(int code, const std::string& msg);
}

using namespace nomad_dev;

namespace {

constexpr int kTask = 1024;
constexpr int kPre = 8;  // rows in flight per wave
constexpr int kRedBlocks = 256;

struct ObjTask {
    int64_t e0;
    int32_t len;
    int32_t line;
};

// the storage traits of the arithmetic type; WIDE: rows of more than 128
// elements, element pairs l and l + 64 in lane l
template <class A, bool WIDE>
struct Acc {
    using T = Store<(sizeof(A) == 8 ? NOMAD_STORE_FP64 : NOMAD_STORE_FP32) | (WIDE ? kWideBit : 0)>;
    using pair = typename T::pair;
};

// acc += e * y per element, and the lane's elements written to a plain row
// (lanes past the row write nothing)
template <class P, class A>
__device__ __forceinline__ void axpy_p(A e, const P& y, P& acc) {
    acc.x = fma_t(e, y.x, acc.x);
    acc.y = fma_t(e, y.y, acc.y);
}
template <class BP, class A>
__device__ __forceinline__ void axpy_p(A e, const WPair<BP>& y, WPair<BP>& acc) {
    axpy_p(e, y.a, acc.a);
    axpy_p(e, y.b, acc.b);
}
template <class P, class A>
__device__ __forceinline__ void row_put(A* dst, int lane, int npairs, const P& v) {
    if (lane < npairs) {
        dst[2 * lane] = v.x;
        dst[2 * lane + 1] = v.y;
    }
}
template <class BP, class A>
__device__ __forceinline__ void row_put(A* dst, int lane, int npairs, const WPair<BP>& v) {
    row_put(dst, lane, npairs, v.a);
    row_put(dst + 2 * kWave, lane, npairs - kWave, v.b);
}

// One wave per task: own = the task line's factor (column j: h_j, row i: w_i),
// other = the opposite side.  part[t] (ldf values) = sum over the task's
// ratings of e * other_row; sq[t] = sum of e^2 (FP64).
template <class A, bool WIDE>
__global__ void obj_pass(const ObjTask* tasks, int ntasks, const int32_t* idx, const A* val, const A* own,
                         const A* other, int ldf, A* part, double* sq) {
    using T = typename Acc<A, WIDE>::T;
    using P = typename Acc<A, WIDE>::pair;
    const int lane = threadIdx.x & 63;
    // wave-uniform (scalar) task index: the task's line then addresses its
    // row through scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(int((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int npairs = ldf >> 1;
    for (int t = wave; t < ntasks; t += nwaves) {
        const ObjTask tk = tasks[t];
        const P x = T::load(own + int64_t(tk.line) * ldf, lane, npairs);
        P acc = T::zero();
        double s2 = 0;
        for (int base = 0; base < tk.len; base += kWave) {
            const int cnt = min(kWave, tk.len - base);
            const int o = min(lane, cnt - 1);
            const int32_t myrow = idx[tk.e0 + base + o];
            const A myval = val[tk.e0 + base + o];
            P y[kPre];
#pragma unroll
            for (int d = 0; d < kPre; ++d) y[d] = T::load(other + int64_t(readlane_i32(myrow, min(d, cnt - 1))) * ldf,
                                                          lane, npairs);
            for (int q0 = 0; q0 < cnt; q0 += kPre) {
#pragma unroll
                for (int d = 0; d < kPre; ++d) {
                    const int q = q0 + d;
                    if (q < cnt) {
                        const A e = wave_dot_p(x, y[d]) - A(readlane_val(myval, q));
                        axpy_p(e, y[d], acc);
                        s2 += double(e) * double(e);
                    }
                    const int qn = min(q + kPre, cnt - 1);
                    y[d] = T::load(other + int64_t(readlane_i32(myrow, qn)) * ldf, lane, npairs);
                }
            }
        }
        row_put(part + int64_t(t) * ldf, lane, npairs, acc);
        if (lane == 0) sq[t] = s2;
    }
}

// grad[line] = 2 sum of the line's task partials + 2 reg deg(line) own[line];
// nrm[line] = deg(line) |own[line]|^2 (FP64).  Grid-stride over (line, element).
template <class A>
__global__ void obj_finalize(int nlines, int ldf, const int32_t* task_ptr, const int64_t* ptr, const A* part,
                             const A* own, A reg, A* grad, double* nrm) {
    const int64_t total = int64_t(nlines) * ldf;
    for (int64_t gid = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; gid < total;
         gid += int64_t(gridDim.x) * blockDim.x) {
        const int c = int(gid / ldf), d = int(gid % ldf);
        A s = A(0);
        for (int t = task_ptr[c]; t < task_ptr[c + 1]; ++t) s += part[int64_t(t) * ldf + d];
        const A deg = A(ptr[c + 1] - ptr[c]);
        const A x = own[gid];
        grad[gid] = A(2) * s + A(2) * reg * deg * x;
        if (nrm) nrm[gid] = double(deg) * double(x) * double(x);
    }
}

__global__ void obj_sum_stage1(int64_t n, const double* a, double* part) {
    __shared__ double s[256];
    double acc = 0;
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        acc += a[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void obj_sum_stage2(int n, const double* part, double* out, int slot) {
    __shared__ double s[256];
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[slot] = s[0];
}

void make_tasks(int nlines, const int64_t* ptr, std::vector<ObjTask>& tasks, std::vector<int32_t>& task_ptr) {
    tasks.clear();
    task_ptr.assign(size_t(nlines) + 1, 0);
    for (int c = 0; c < nlines; ++c) {
        for (int64_t e = ptr[c]; e < ptr[c + 1]; e += kTask)
            tasks.push_back(ObjTask{e, int32_t(std::min<int64_t>(kTask, ptr[c + 1] - e)), c});
        task_ptr[size_t(c) + 1] = int32_t(tasks.size());
    }
}

}  // namespace

struct nomad_obj_ctx {
    nomad_obj_params p{};
    int ldf = 0;
    size_t esz = 8;
    std::vector<void*> bufs;
    int64_t* colptr = nullptr;
    int32_t* rowidx = nullptr;
    void* cval = nullptr;
    int64_t* rowptr = nullptr;
    int32_t* colidx = nullptr;
    void* rval = nullptr;
    ObjTask *ctasks = nullptr, *rtasks = nullptr;
    int32_t *ctask_ptr = nullptr, *rtask_ptr = nullptr;
    int nct = 0, nrt = 0;
    void* part = nullptr;     // max(nct, nrt) x ldf
    double* sq = nullptr;     // max(nct, nrt)
    double* nrm = nullptr;    // max(m, n) x ldf
    double* red = nullptr;    // kRedBlocks
    double* out = nullptr;    // 4 slots
    void *hW = nullptr, *hH = nullptr, *hgW = nullptr, *hgH = nullptr;  // staging of nomad_obj_eval_host
    ~nomad_obj_ctx() {
        for (void* b : bufs) (void)hipFree(b);
    }
    template <class T>
    int alloc(T** ptr, size_t count) {
        *ptr = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return nomad_host::fail(NOMAD_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        bufs.push_back(*ptr);
        return NOMAD_OK;
    }
    template <class T>
    int upload(T** ptr, const T* src, size_t count) {
        int rc = alloc(ptr, count);
        if (rc) return rc;
        if (count && hipMemcpy(*ptr, src, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return nomad_host::fail(NOMAD_EHIP, "hipMemcpy to device failed");
        return NOMAD_OK;
    }
};

extern "C" int nomad_obj_create(const nomad_obj_params* p, const int64_t* colptr, const int32_t* rowidx,
                                const double* val, int64_t nnz, nomad_obj_ctx** out) {
    using nomad_host::fail;
    if (!p || !colptr || !rowidx || !val || !out) return fail(NOMAD_EINVAL, "nomad_obj_create: null argument");
    *out = nullptr;
    const int m = p->m, n = p->n, k = p->k;
    if (m <= 0 || n <= 0 || k <= 0 || k > 256) return fail(NOMAD_EINVAL, "nomad_obj_create: need m, n > 0, 0 < k <= 256");
    if (p->storage != NOMAD_STORE_FP32 && p->storage != NOMAD_STORE_FP64)
        return fail(NOMAD_EINVAL, "nomad_obj_create: storage must be FP32 or FP64");
    if (nnz < 0 || colptr[0] != 0 || colptr[n] != nnz) return fail(NOMAD_EINVAL, "nomad_obj_create: bad colptr");
    for (int j = 0; j < n; ++j)
        if (colptr[j + 1] < colptr[j]) return fail(NOMAD_EINVAL, "nomad_obj_create: colptr not ascending");
    for (int64_t e = 0; e < nnz; ++e)
        if (rowidx[e] < 0 || rowidx[e] >= m) return fail(NOMAD_EINVAL, "nomad_obj_create: row index out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || p->device < 0 || p->device >= ndev)
        return fail(NOMAD_EHIP, "nomad_obj_create: no such HIP device");
    if (hipSetDevice(p->device) != hipSuccess) return fail(NOMAD_EHIP, "nomad_obj_create: hipSetDevice failed");
    auto* c = new (std::nothrow) nomad_obj_ctx();
    if (!c) return fail(NOMAD_ENOMEM, "nomad_obj_create: out of host memory");
    c->p = *p;
    c->ldf = (k + 1) & ~1;
    c->esz = p->storage == NOMAD_STORE_FP64 ? 8 : 4;
    // CSR copy with values in row order
    std::vector<int64_t> rowptr(size_t(m) + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) rowptr[size_t(rowidx[e]) + 1]++;
    for (int i = 0; i < m; ++i) rowptr[size_t(i) + 1] += rowptr[size_t(i)];
    std::vector<int32_t> colidx(static_cast<size_t>(nnz));
    std::vector<double> rv(static_cast<size_t>(nnz));
    {
        std::vector<int64_t> fill(rowptr.begin(), rowptr.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) {
                const int64_t d = fill[size_t(rowidx[e])]++;
                colidx[size_t(d)] = j;
                rv[size_t(d)] = val[e];
            }
    }
    std::vector<ObjTask> ct, rt;
    std::vector<int32_t> ctp, rtp;
    make_tasks(n, colptr, ct, ctp);
    make_tasks(m, rowptr.data(), rt, rtp);
    int rc;
    auto up_vals = [&](void** dst, const double* src) -> int {
        if (c->esz == 8) return c->upload(reinterpret_cast<double**>(dst), src, size_t(nnz));
        std::vector<float> f(static_cast<size_t>(nnz));
        for (int64_t e = 0; e < nnz; ++e) f[size_t(e)] = float(src[e]);
        return c->upload(reinterpret_cast<float**>(dst), f.data(), f.size());
    };
    const size_t maxt = std::max(ct.size(), rt.size());
    if ((rc = c->upload(&c->colptr, colptr, size_t(n) + 1)) || (rc = c->upload(&c->rowidx, rowidx, size_t(nnz))) ||
        (rc = up_vals(&c->cval, val)) || (rc = c->upload(&c->rowptr, rowptr.data(), rowptr.size())) ||
        (rc = c->upload(&c->colidx, colidx.data(), colidx.size())) || (rc = up_vals(&c->rval, rv.data())) ||
        (rc = c->upload(&c->ctasks, ct.data(), ct.size())) || (rc = c->upload(&c->rtasks, rt.data(), rt.size())) ||
        (rc = c->upload(&c->ctask_ptr, ctp.data(), ctp.size())) ||
        (rc = c->upload(&c->rtask_ptr, rtp.data(), rtp.size())) ||
        (rc = c->alloc(reinterpret_cast<char**>(&c->part), maxt * size_t(c->ldf) * c->esz)) ||
        (rc = c->alloc(&c->sq, maxt)) || (rc = c->alloc(&c->nrm, size_t(std::max(m, n)) * size_t(c->ldf))) ||
        (rc = c->alloc(&c->red, kRedBlocks)) || (rc = c->alloc(&c->out, 4))) {
        delete c;
        return rc;
    }
    c->nct = int(ct.size());
    c->nrt = int(rt.size());
    *out = c;
    return NOMAD_OK;
}

extern "C" int nomad_obj_destroy(nomad_obj_ctx* c) {
    if (!c) return NOMAD_OK;
    (void)hipSetDevice(c->p.device);
    (void)hipDeviceSynchronize();
    delete c;
    return NOMAD_OK;
}

namespace {

template <class A>
int obj_eval_t(nomad_obj_ctx* c, const A* W, const A* H, double reg, A* gW, A* gH, double* loss, hipStream_t s) {
    const int m = c->p.m, n = c->p.n, ldf = c->ldf;
    A* part = static_cast<A*>(c->part);
    auto grid = [](int64_t count, int threads) {
        return unsigned(std::max<int64_t>(1, std::min<int64_t>(16384, (count + threads - 1) / threads)));
    };
    // columns: dL/dH, squared errors, column norms
    // (rows of more than 128 elements: two element pairs per lane)
    auto pass = ldf > 128 ? obj_pass<A, true> : obj_pass<A, false>;
    hipLaunchKernelGGL(pass, dim3(grid(int64_t(c->nct) * 64, 256)), dim3(256), 0, s, c->ctasks, c->nct, c->rowidx,
                       static_cast<const A*>(c->cval), H, W, ldf, part, c->sq);
    hipLaunchKernelGGL(obj_sum_stage1, dim3(kRedBlocks), dim3(256), 0, s, int64_t(c->nct), c->sq, c->red);
    hipLaunchKernelGGL(obj_sum_stage2, dim3(1), dim3(256), 0, s, kRedBlocks, c->red, c->out, 0);
    hipLaunchKernelGGL((obj_finalize<A>), dim3(grid(int64_t(n) * ldf, 256)), dim3(256), 0, s, n, ldf, c->ctask_ptr,
                       c->colptr, part, H, A(reg), gH, c->nrm);
    hipLaunchKernelGGL(obj_sum_stage1, dim3(kRedBlocks), dim3(256), 0, s, int64_t(n) * ldf, c->nrm, c->red);
    hipLaunchKernelGGL(obj_sum_stage2, dim3(1), dim3(256), 0, s, kRedBlocks, c->red, c->out, 1);
    // rows: dL/dW, row norms
    hipLaunchKernelGGL(pass, dim3(grid(int64_t(c->nrt) * 64, 256)), dim3(256), 0, s, c->rtasks, c->nrt, c->colidx,
                       static_cast<const A*>(c->rval), W, H, ldf, part, c->sq);
    hipLaunchKernelGGL((obj_finalize<A>), dim3(grid(int64_t(m) * ldf, 256)), dim3(256), 0, s, m, ldf, c->rtask_ptr,
                       c->rowptr, part, W, A(reg), gW, c->nrm);
    hipLaunchKernelGGL(obj_sum_stage1, dim3(kRedBlocks), dim3(256), 0, s, int64_t(m) * ldf, c->nrm, c->red);
    hipLaunchKernelGGL(obj_sum_stage2, dim3(1), dim3(256), 0, s, kRedBlocks, c->red, c->out, 2);
    if (hipGetLastError() != hipSuccess) return nomad_host::fail(NOMAD_EHIP, "nomad_obj_eval: launch failed");
    double o[4];
    if (hipMemcpyAsync(o, c->out, sizeof(o), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return nomad_host::fail(NOMAD_EHIP, "nomad_obj_eval: kernel or copy failed");
    if (loss) *loss = o[0] + reg * (o[1] + o[2]);
    return NOMAD_OK;
}

}  // namespace

extern "C" int nomad_obj_eval(nomad_obj_ctx* c, const void* W, const void* H, double reg, void* gW, void* gH,
                              double* loss, void* stream) {
    if (!c || !W || !H || !gW || !gH) return nomad_host::fail(NOMAD_EINVAL, "nomad_obj_eval: null argument");
    if (hipSetDevice(c->p.device) != hipSuccess) return nomad_host::fail(NOMAD_EHIP, "nomad_obj_eval: hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (c->p.storage == NOMAD_STORE_FP64)
        return obj_eval_t<double>(c, static_cast<const double*>(W), static_cast<const double*>(H), reg,
                                  static_cast<double*>(gW), static_cast<double*>(gH), loss, s);
    return obj_eval_t<float>(c, static_cast<const float*>(W), static_cast<const float*>(H), reg,
                             static_cast<float*>(gW), static_cast<float*>(gH), loss, s);
}

// Host-array convenience: W (m x k), H (n x k) row-major double in, the
// gradients (same shapes) out; staged through device buffers of the storage
// type owned by the context.
extern "C" int nomad_obj_eval_host(nomad_obj_ctx* c, const double* W, const double* H, double reg, double* gW,
                                   double* gH, double* loss) {
    using nomad_host::fail;
    if (!c || !W || !H || !gW || !gH) return fail(NOMAD_EINVAL, "nomad_obj_eval_host: null argument");
    if (hipSetDevice(c->p.device) != hipSuccess) return fail(NOMAD_EHIP, "nomad_obj_eval_host: hipSetDevice");
    const int m = c->p.m, n = c->p.n, k = c->p.k, ldf = c->ldf;
    int rc;
    if (!c->hW) {
        if ((rc = c->alloc(reinterpret_cast<char**>(&c->hW), size_t(m) * ldf * c->esz)) ||
            (rc = c->alloc(reinterpret_cast<char**>(&c->hH), size_t(n) * ldf * c->esz)) ||
            (rc = c->alloc(reinterpret_cast<char**>(&c->hgW), size_t(m) * ldf * c->esz)) ||
            (rc = c->alloc(reinterpret_cast<char**>(&c->hgH), size_t(n) * ldf * c->esz)))
            return rc;
    }
    auto put = [&](void* dst, const double* src, int rows) -> int {
        std::vector<char> buf(size_t(rows) * ldf * c->esz, 0);
        for (int r = 0; r < rows; ++r)
            for (int d = 0; d < k; ++d) {
                const double v = src[size_t(r) * k + d];
                if (c->esz == 8)
                    reinterpret_cast<double*>(buf.data())[size_t(r) * ldf + d] = v;
                else
                    reinterpret_cast<float*>(buf.data())[size_t(r) * ldf + d] = float(v);
            }
        return hipMemcpy(dst, buf.data(), buf.size(), hipMemcpyHostToDevice) == hipSuccess
                   ? NOMAD_OK
                   : fail(NOMAD_EHIP, "nomad_obj_eval_host: copy to device failed");
    };
    auto get = [&](double* dst, const void* src, int rows) -> int {
        std::vector<char> buf(size_t(rows) * ldf * c->esz);
        if (hipMemcpy(buf.data(), src, buf.size(), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(NOMAD_EHIP, "nomad_obj_eval_host: copy to host failed");
        for (int r = 0; r < rows; ++r)
            for (int d = 0; d < k; ++d)
                dst[size_t(r) * k + d] =
                    c->esz == 8 ? reinterpret_cast<const double*>(buf.data())[size_t(r) * ldf + d]
                                : double(reinterpret_cast<const float*>(buf.data())[size_t(r) * ldf + d]);
        return NOMAD_OK;
    };
    if ((rc = put(c->hW, W, m)) || (rc = put(c->hH, H, n))) return rc;
    if ((rc = nomad_obj_eval(c, c->hW, c->hH, reg, c->hgW, c->hgH, loss, nullptr))) return rc;
    if ((rc = get(gW, c->hgW, m)) || (rc = get(gH, c->hgH, n))) return rc;
    return NOMAD_OK;
}
