// CCD++ rank-one sweeps on the GPU (the reference's competitor solver,
// Code/competitor/ccdr1-mpi-omp.cpp:334-558, at one process), FP64.
//
// The residual R^ = R - W H^T is kept once, in CSC order, for the training and
// the test ratings.  Per rank t (u = W[:,t], v = H[:,t]):
//   R^ += u v^T                                   (UpdateRating add, :444-447)
//   inner iterations: every column j  v_j = g_j / h_j with
//       g_j = sum_i u_i R^_ij,  h_j = lambda nnz_j + sum_i u_i^2     (RankOneUpdate,
//     then every row i with v the same way; fundec = sum h (old - new)^2  :334-358)
//     stop when fundec < fundec_max * eps (first inner iteration: count an
//     early stop; five in one outer iteration end it)              (:466-497)
//   R^ -= u v^T, loss = sum R^^2, reg += the rank's weighted norms   (:505-530)
// Column and row passes split every column (row) into tasks of at most
// kTask ratings, one wave per task; the partial sums of a column are added in
// task order, so every result is deterministic.  The row pass reads the
// residual's CSR-ordered copy, which every change of the residual updates as
// well (the reference keeps R and Rt alike, :444-447, 511-512): read through
// the CSR -> CSC position map instead, a row pass fetched 6.8x its bytes.
//
// The early stop is decided on the device: after each inner iteration one
// thread compares fundec with fundec_max * eps and, when the rank is done,
// sets a flag that makes the remaining queued iterations of the rank return
// at once.  The host enqueues a rank's inner iterations without waiting and
// reads back once per rank (the rank's objective and test RMSE, and the
// count of first-iteration stops that ends the outer iteration).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nomad_hip.h"

namespace nomad_host {
int fail(int code, const std::string& msg);
}

namespace {

constexpr int kTask = 2048;    // ratings per pass task
constexpr int kRedBlocks = 2048;  // blocks of the grid-stride passes and first reduction stage

// early-stop state of the outer iteration (:421-423, 466-497)
struct CcdState {
    double fundec_max;
    int32_t early_stop;  // ranks stopped at their first inner iteration
    int32_t stopped;     // the current rank's inner iterations are over
    int64_t inner_done;
};

struct CcdTask {
    int64_t e0;
    int32_t len;
    int32_t line;  // column (column pass) or row (row pass)
};

__device__ __forceinline__ double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// r[e] = val[e] - sum_t Wt[t][row] Ht[t][col]
__global__ void ccd_residual_init(int64_t nnz, int m, int n, int k, const int32_t* rowidx, const int32_t* ecol,
                                  const double* val, const double* Wt, const double* Ht, double* r) {
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < nnz; e += int64_t(gridDim.x) * blockDim.x) {
        const int i = rowidx[e], j = ecol[e];
        double d = 0;
        for (int t = 0; t < k; ++t) d += Wt[int64_t(t) * m + i] * Ht[int64_t(t) * n + j];
        r[e] = val[e] - d;
    }
}

// r[e] += sign u[row] v[col]; with part != nullptr also the sum of the new
// r[e]^2, one partial per block (threads add their elements in grid-stride
// order, the block adds its threads in a fixed tree)
__global__ void ccd_rank_update(int64_t nnz, const int32_t* rowidx, const int32_t* ecol, const double* u,
                                const double* v, double sign, double* r, double* part) {
    __shared__ double s[256];
    double acc = 0;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < nnz; e += int64_t(gridDim.x) * blockDim.x) {
        const double x = r[e] + sign * (u[rowidx[e]] * v[ecol[e]]);
        r[e] = x;
        acc += x * x;
    }
    if (!part) return;
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Ranks that stop at their first inner iteration before an outer iteration
// ends early (Code/competitor/ccdr1-mpi-omp.cpp:432).
constexpr int kEarlyStopRanks = 5;

// Rank t's term out and the next rank's back in, in one pass:
//   x = r[e] - u[row] v[col] ; r[e] = x + u_n[row] v_n[col]
// with the two roundings of two separate ccd_rank_update passes (each line
// is that kernel's expression), and with part != nullptr the sum of x^2 per
// block (the loss after rank t).  The next rank is t + 1 (u1, v1) while the
// device's early-stop count is below the limit, else rank 0 of the next
// outer iteration (u0, v0; null: none).
__global__ void ccd_rank_swap(int64_t nnz, const int32_t* rowidx, const int32_t* ecol, const double* u,
                              const double* v, const double* u1, const double* v1, const double* u0,
                              const double* v0, const CcdState* st, double* r, double* part) {
    __shared__ double s[256];
    const bool go1 = u1 && st->early_stop < kEarlyStopRanks;
    const double* un = go1 ? u1 : u0;
    const double* vn = go1 ? v1 : v0;
    const double sign = -1.0, sign_n = 1.0;
    double acc = 0;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < nnz; e += int64_t(gridDim.x) * blockDim.x) {
        const int32_t i = rowidx[e], j = ecol[e];
        const double x = r[e] + sign * (u[i] * v[j]);
        acc += x * x;
        r[e] = un ? x + sign_n * (un[i] * vn[j]) : x;
    }
    if (!part) return;
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// the CSR-ordered copy of the residual: rt[d] = r[perm[d]] (once, at set-up)
__global__ void ccd_gather(int64_t nnz, const int32_t* perm, const double* r, double* rt) {
    for (int64_t d = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; d < nnz; d += int64_t(gridDim.x) * blockDim.x)
        rt[d] = r[perm[d]];
}

// partial (g, h) of every task: one wave per task
__global__ void ccd_pass_partial(const CcdTask* tasks, int ntasks, const int32_t* idx, const double* r,
                                 const double* x, double2* part, const CcdState* st) {
    if (st->stopped) return;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int t = wave; t < ntasks; t += nwaves) {
        const CcdTask tk = tasks[t];
        double g = 0, h = 0;
        // kU of the lane's ratings in flight, added in the same order
        constexpr int kU = 8;
        int q = lane;
        for (; q + 64 * (kU - 1) < tk.len; q += 64 * kU) {
            double xv[kU], rv[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int64_t e = tk.e0 + q + 64 * u;
                xv[u] = x[idx[e]];
                rv[u] = r[e];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                g += xv[u] * rv[u];
                h += xv[u] * xv[u];
            }
        }
        for (; q < tk.len; q += 64) {
            const int64_t e = tk.e0 + q;
            const double xi = x[idx[e]];
            const double rv = r[e];
            g += xi * rv;
            h += xi * xi;
        }
        g = wave_sum(g);
        h = wave_sum(h);
        if (lane == 0) part[t] = make_double2(g, h);
    }
}

// new value of every line from its tasks' partials, in task order
__global__ void ccd_pass_finalize(int nlines, const int32_t* task_ptr, const int64_t* ptr, const double2* part,
                                  double lambda, double* y, double* fundec, const CcdState* st) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nlines || st->stopped) return;
    const int64_t cnt = ptr[c + 1] - ptr[c];
    if (cnt == 0) {  // RankOneUpdate returns 0 for an empty line
        y[c] = 0;
        fundec[c] = 0;
        return;
    }
    double g = 0, h = 0;
    for (int t = task_ptr[c]; t < task_ptr[c + 1]; ++t) {
        g += part[t].x;
        h += part[t].y;
    }
    h += lambda * double(cnt);
    const double nv = g / h;
    const double delta = y[c] - nv;
    fundec[c] = h * delta * delta;
    y[c] = nv;
}

// sum of a[i] (or a[i]^2 (ptr[i+1] - ptr[i]) with ptr) : stage 1, one partial per
// block (skipped while st->stopped, st may be null)
__global__ void ccd_sum_stage1(int64_t n, const double* a, const int64_t* ptr, double* part, const CcdState* st) {
    if (st && st->stopped) return;
    __shared__ double s[256];
    double acc = 0;
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        acc += ptr ? a[i] * a[i] * double(ptr[i + 1] - ptr[i]) : a[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// stage 2: one block adds the stage-1 partials; out[slot] = sum
__global__ void ccd_sum_stage2(int n, const double* part, double* out, int slot, const CcdState* st) {
    if (st && st->stopped) return;
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

// after inner iteration `iter` of rank t in outer iteration oiter: the
// reference's test on fundec = out[0] + out[1] (ccdr1-mpi-omp.cpp:490-497)
__global__ void ccd_inner_decide(CcdState* st, const double* out, int iter, int oiter, int t, double eps) {
    if (st->stopped) return;
    ++st->inner_done;
    const double cur = out[0] + out[1];
    if (cur < st->fundec_max * eps) {
        if (iter == 1) st->early_stop += 1;
        st->stopped = 1;
        return;
    }
    if (!(oiter == 1 && t == 0 && iter == 1)) st->fundec_max = fmax(st->fundec_max, cur);
}

// rank start (stopped = 0) or outer-iteration start (everything reset)
__global__ void ccd_state_reset(CcdState* st, int outer) {
    if (outer) {
        st->fundec_max = 0;
        st->early_stop = 0;
    }
    st->stopped = 0;
}

// ---------------------------------------------------------------------------
struct Buffers {
    std::vector<void*> ptrs;
    ~Buffers() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** p, size_t count) {
        *p = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess)
            return nomad_host::fail(NOMAD_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        ptrs.push_back(*p);
        return NOMAD_OK;
    }
    template <class T>
    int upload(T** p, const T* src, size_t count) {
        int rc = alloc(p, count);
        if (rc) return rc;
        if (count && hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return nomad_host::fail(NOMAD_EHIP, "hipMemcpy to device failed");
        return NOMAD_OK;
    }
};

// tasks of at most kTask entries over lines [0, nlines) of ptr
void make_tasks(int nlines, const int64_t* ptr, std::vector<CcdTask>& tasks, std::vector<int32_t>& task_ptr) {
    tasks.clear();
    task_ptr.assign(size_t(nlines) + 1, 0);
    for (int c = 0; c < nlines; ++c) {
        for (int64_t e = ptr[c]; e < ptr[c + 1]; e += kTask)
            tasks.push_back(CcdTask{e, int32_t(std::min<int64_t>(kTask, ptr[c + 1] - e)), c});
        task_ptr[size_t(c) + 1] = int32_t(tasks.size());
    }
}

struct Side {  // one of the two passes
    int nlines = 0;
    int ntasks = 0;
    const int64_t* ptr = nullptr;  // device
    const int32_t* idx = nullptr;
    const double* r = nullptr;  // residual in this side's order
    CcdTask* tasks = nullptr;
    int32_t* task_ptr = nullptr;
    double2* part = nullptr;
    double* fundec = nullptr;
};

}  // namespace

extern "C" int nomad_ccd(const nomad_ccd_params* p, const int64_t* colptr, const int32_t* rowidx, const double* val,
                         int64_t nnz, const int64_t* te_colptr, const int32_t* te_rowidx, const double* te_val,
                         int64_t te_nnz, double* W, double* H, double* obj, double* rmse, nomad_ccd_stats* st) {
    using nomad_host::fail;
    if (!p || !colptr || !rowidx || !val || !W || !H) return fail(NOMAD_EINVAL, "nomad_ccd: null argument");
    const int m = p->m, n = p->n, k = p->k;
    if (m <= 0 || n <= 0 || k <= 0 || p->iters < 0 || p->inner <= 0 || !(p->lambda >= 0) || !(p->eps >= 0))
        return fail(NOMAD_EINVAL, "nomad_ccd: m, n, k, inner must be positive, iters, lambda, eps >= 0");
    if (nnz < 0 || colptr[0] != 0 || colptr[n] != nnz || nnz >= (int64_t(1) << 31))
        return fail(NOMAD_EINVAL, "nomad_ccd: colptr must run from 0 to nnz < 2^31");
    const bool has_test = te_colptr && te_rowidx && te_val && te_nnz > 0;
    if (has_test && (te_colptr[0] != 0 || te_colptr[n] != te_nnz))
        return fail(NOMAD_EINVAL, "nomad_ccd: test colptr must run from 0 to te_nnz");
    for (int64_t e = 0; e < nnz; ++e)
        if (rowidx[e] < 0 || rowidx[e] >= m) return fail(NOMAD_EINVAL, "nomad_ccd: row index out of range");
    if (has_test)
        for (int64_t e = 0; e < te_nnz; ++e)
            if (te_rowidx[e] < 0 || te_rowidx[e] >= m)
                return fail(NOMAD_EINVAL, "nomad_ccd: test row index out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || p->device < 0 || p->device >= ndev)
        return fail(NOMAD_EHIP, "nomad_ccd: no such HIP device");
    if (hipSetDevice(p->device) != hipSuccess) return fail(NOMAD_EHIP, "nomad_ccd: hipSetDevice failed");
    const auto t_start = std::chrono::steady_clock::now();

    // host-side structure: column of every CSC entry, CSR with the position map
    std::vector<int32_t> ecol(static_cast<size_t>(nnz)), te_ecol(static_cast<size_t>(has_test ? te_nnz : 0));
    for (int j = 0; j < n; ++j)
        for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) ecol[size_t(e)] = j;
    if (has_test)
        for (int j = 0; j < n; ++j)
            for (int64_t e = te_colptr[j]; e < te_colptr[j + 1]; ++e) te_ecol[size_t(e)] = j;
    std::vector<int64_t> rowptr(size_t(m) + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) rowptr[size_t(rowidx[e]) + 1]++;
    for (int i = 0; i < m; ++i) rowptr[size_t(i) + 1] += rowptr[size_t(i)];
    std::vector<int32_t> colidx(static_cast<size_t>(nnz)), perm(static_cast<size_t>(nnz)), erow(static_cast<size_t>(nnz));
    for (int i = 0; i < m; ++i)
        for (int64_t d = rowptr[size_t(i)]; d < rowptr[size_t(i) + 1]; ++d) erow[size_t(d)] = i;
    {
        std::vector<int64_t> fill(rowptr.begin(), rowptr.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) {
                const int64_t d = fill[size_t(rowidx[e])]++;
                colidx[size_t(d)] = j;
                perm[size_t(d)] = int32_t(e);
            }
    }
    std::vector<CcdTask> ctasks, rtasks;
    std::vector<int32_t> ctask_ptr, rtask_ptr;
    make_tasks(n, colptr, ctasks, ctask_ptr);
    make_tasks(m, rowptr.data(), rtasks, rtask_ptr);
    // factors transposed: Wt[t][i], Ht[t][j]
    std::vector<double> Wt(size_t(k) * m), Ht(size_t(k) * n);
    for (int i = 0; i < m; ++i)
        for (int t = 0; t < k; ++t) Wt[size_t(t) * m + i] = W[size_t(i) * k + t];
    for (int j = 0; j < n; ++j)
        for (int t = 0; t < k; ++t) Ht[size_t(t) * n + j] = H[size_t(j) * k + t];

    Buffers b;
    int rc;
    int64_t *d_colptr, *d_rowptr;
    int32_t *d_rowidx, *d_ecol, *d_colidx, *d_erow, *d_perm, *d_te_rowidx = nullptr, *d_te_ecol = nullptr;
    double *d_val, *d_r, *d_rt, *d_te_val = nullptr, *d_te_r = nullptr, *d_Wt, *d_Ht, *d_part, *d_out;
    CcdState* d_state = nullptr;
    Side cs, rs;
    if ((rc = b.upload(&d_colptr, colptr, size_t(n) + 1)) || (rc = b.upload(&d_rowidx, rowidx, size_t(nnz))) ||
        (rc = b.upload(&d_ecol, ecol.data(), ecol.size())) || (rc = b.upload(&d_val, val, size_t(nnz))) ||
        (rc = b.alloc(&d_r, size_t(nnz))) || (rc = b.upload(&d_rowptr, rowptr.data(), rowptr.size())) ||
        (rc = b.upload(&d_colidx, colidx.data(), colidx.size())) || (rc = b.upload(&d_perm, perm.data(), perm.size())) ||
        (rc = b.upload(&d_erow, erow.data(), erow.size())) || (rc = b.alloc(&d_rt, size_t(nnz))) ||
        (rc = b.upload(&d_Wt, Wt.data(), Wt.size())) || (rc = b.upload(&d_Ht, Ht.data(), Ht.size())) ||
        (rc = b.alloc(&d_part, kRedBlocks)) || (rc = b.alloc(&d_out, 8)) || (rc = b.alloc(&d_state, 1)) ||
        (rc = b.upload(&cs.tasks, ctasks.data(), ctasks.size())) ||
        (rc = b.upload(&cs.task_ptr, ctask_ptr.data(), ctask_ptr.size())) ||
        (rc = b.alloc(&cs.part, ctasks.size())) || (rc = b.alloc(&cs.fundec, size_t(n))) ||
        (rc = b.upload(&rs.tasks, rtasks.data(), rtasks.size())) ||
        (rc = b.upload(&rs.task_ptr, rtask_ptr.data(), rtask_ptr.size())) ||
        (rc = b.alloc(&rs.part, rtasks.size())) || (rc = b.alloc(&rs.fundec, size_t(m))))
        return rc;
    if (has_test && ((rc = b.upload(&d_te_rowidx, te_rowidx, size_t(te_nnz))) ||
                     (rc = b.upload(&d_te_ecol, te_ecol.data(), te_ecol.size())) ||
                     (rc = b.upload(&d_te_val, te_val, size_t(te_nnz))) || (rc = b.alloc(&d_te_r, size_t(te_nnz)))))
        return rc;
    cs.nlines = n;
    cs.ntasks = int(ctasks.size());
    cs.ptr = d_colptr;
    cs.idx = d_rowidx;
    cs.r = d_r;
    rs.nlines = m;
    rs.ntasks = int(rtasks.size());
    rs.ptr = d_rowptr;
    rs.idx = d_colidx;
    rs.r = d_rt;

    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return fail(NOMAD_EHIP, "nomad_ccd: stream creation failed");
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    const int eb = 256;
    auto grid_for = [](int64_t count, int threads) {
        return unsigned(std::max<int64_t>(1, std::min<int64_t>(8192, (count + threads - 1) / threads)));
    };
    hipLaunchKernelGGL(ccd_residual_init, dim3(grid_for(nnz, eb)), dim3(eb), 0, s, nnz, m, n, k, d_rowidx, d_ecol,
                       d_val, d_Wt, d_Ht, d_r);
    if (has_test)
        hipLaunchKernelGGL(ccd_residual_init, dim3(grid_for(te_nnz, eb)), dim3(eb), 0, s, te_nnz, m, n, k,
                           d_te_rowidx, d_te_ecol, d_te_val, d_Wt, d_Ht, d_te_r);
    hipLaunchKernelGGL(ccd_gather, dim3(grid_for(nnz, eb)), dim3(eb), 0, s, int64_t(nnz), d_perm, d_r, d_rt);
    hipLaunchKernelGGL(ccd_state_reset, dim3(1), dim3(1), 0, s, d_state, 1);

    // out[slot] = sum; then read back all slots that were written
    auto sum_into = [&](int64_t count, const double* a, const int64_t* ptr, int slot,
                        const CcdState* skip = nullptr) {
        hipLaunchKernelGGL(ccd_sum_stage1, dim3(kRedBlocks), dim3(256), 0, s, count, a, ptr, d_part, skip);
        hipLaunchKernelGGL(ccd_sum_stage2, dim3(1), dim3(256), 0, s, kRedBlocks, d_part, d_out, slot, skip);
    };
    double host_out[8];
    CcdState host_st{};
    auto fetch = [&]() -> int {
        if (hipMemcpyAsync(host_out, d_out, sizeof(host_out), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(&host_st, d_state, sizeof(host_st), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail(NOMAD_EHIP, "nomad_ccd: kernel or copy failed");
        return NOMAD_OK;
    };
    auto rank_update = [&](int64_t count, const int32_t* ri, const int32_t* ci, const double* u, const double* v,
                           double sign, double* r, bool want_loss, int slot) {
        hipLaunchKernelGGL(ccd_rank_update, dim3(kRedBlocks), dim3(256), 0, s, count, ri, ci, u, v, sign, r,
                           want_loss ? d_part : nullptr);
        if (want_loss)
            hipLaunchKernelGGL(ccd_sum_stage2, dim3(1), dim3(256), 0, s, kRedBlocks, d_part, d_out, slot,
                               static_cast<const CcdState*>(nullptr));
    };
    auto rank_swap = [&](int64_t count, const int32_t* ri, const int32_t* ci, const double* u, const double* v,
                         const double* u1, const double* v1, const double* u0, const double* v0, double* r,
                         bool want_loss, int slot) {
        hipLaunchKernelGGL(ccd_rank_swap, dim3(kRedBlocks), dim3(256), 0, s, count, ri, ci, u, v, u1, v1, u0, v0,
                           static_cast<const CcdState*>(d_state), r, want_loss ? d_part : nullptr);
        if (want_loss)
            hipLaunchKernelGGL(ccd_sum_stage2, dim3(1), dim3(256), 0, s, kRedBlocks, d_part, d_out, slot,
                               static_cast<const CcdState*>(nullptr));
    };
    auto pass = [&](const Side& sd, const double* x, double* y) {
        hipLaunchKernelGGL(ccd_pass_partial, dim3(grid_for(int64_t(sd.ntasks) * 64, 256)), dim3(256), 0, s,
                           sd.tasks, sd.ntasks, sd.idx, sd.r, x, sd.part, d_state);
        hipLaunchKernelGGL(ccd_pass_finalize, dim3((sd.nlines + 255) / 256), dim3(256), 0, s, sd.nlines, sd.task_ptr,
                           sd.ptr, sd.part, p->lambda, y, sd.fundec, d_state);
    };

    // reg = sum_t sum_i W_it^2 nnz_i + sum_j H_jt^2 nnz_j
    double reg = 0;
    for (int t = 0; t < k; ++t) {
        sum_into(m, d_Wt + size_t(t) * m, d_rowptr, 0);
        sum_into(n, d_Ht + size_t(t) * n, d_colptr, 1);
        if ((rc = fetch())) return rc;
        reg += host_out[0] + host_out[1];
    }
    if (obj)
        for (int64_t q = 0; q < int64_t(p->iters) * k; ++q) obj[q] = std::nan("");
    if (rmse)
        for (int64_t q = 0; q < int64_t(p->iters) * k; ++q) rmse[q] = std::nan("");
    int64_t ranks_done = 0, inner_done = 0;
    if ((rc = fetch())) return rc;  // set-up work finished
    const auto t_solve = std::chrono::steady_clock::now();
    // The residual passes of consecutive ranks are fused: rank t's term is
    // taken out and the next rank's added back in one pass (ccd_rank_swap),
    // which decides on the device, from the same early-stop count the host
    // reads afterwards, which rank is next; `added` says that the current
    // rank's term is already back in the residuals.
    bool added = false;
    for (int oiter = 1; oiter <= p->iters; ++oiter) {
        // fundec_max and the early-stop count restart with every outer iteration
        hipLaunchKernelGGL(ccd_state_reset, dim3(1), dim3(1), 0, s, d_state, 1);
        host_st.early_stop = 0;
        for (int t = 0; t < k; ++t) {
            if (host_st.early_stop >= kEarlyStopRanks) break;
            double* u = d_Wt + size_t(t) * m;
            double* v = d_Ht + size_t(t) * n;
            hipLaunchKernelGGL(ccd_state_reset, dim3(1), dim3(1), 0, s, d_state, 0);
            if (!added) {
                rank_update(nnz, d_rowidx, d_ecol, u, v, 1.0, d_r, false, 0);
                rank_update(nnz, d_erow, d_colidx, u, v, 1.0, d_rt, false, 0);
                if (has_test) rank_update(te_nnz, d_te_rowidx, d_te_ecol, u, v, 1.0, d_te_r, false, 0);
            }
            sum_into(m, u, d_rowptr, 4);
            sum_into(n, v, d_colptr, 5);
            // all inner iterations queued at once; the device stops them
            for (int iter = 1; iter <= p->inner; ++iter) {
                pass(cs, u, v);
                pass(rs, v, u);
                sum_into(n, cs.fundec, nullptr, 0, d_state);
                sum_into(m, rs.fundec, nullptr, 1, d_state);
                hipLaunchKernelGGL(ccd_inner_decide, dim3(1), dim3(1), 0, s, d_state, d_out, iter, oiter, t, p->eps);
            }
            // next rank: t + 1 (if the early-stop count allows it), else rank 0
            // of the next outer iteration, else none
            const double* u1 = t + 1 < k ? d_Wt + size_t(t + 1) * m : nullptr;
            const double* v1 = t + 1 < k ? d_Ht + size_t(t + 1) * n : nullptr;
            const double* u0 = oiter < p->iters ? d_Wt : nullptr;
            const double* v0 = oiter < p->iters ? d_Ht : nullptr;
            rank_swap(nnz, d_rowidx, d_ecol, u, v, u1, v1, u0, v0, d_r, true, 2);
            rank_swap(nnz, d_erow, d_colidx, u, v, u1, v1, u0, v0, d_rt, false, 0);
            if (has_test) rank_swap(te_nnz, d_te_rowidx, d_te_ecol, u, v, u1, v1, u0, v0, d_te_r, true, 3);
            sum_into(m, u, d_rowptr, 0);
            sum_into(n, v, d_colptr, 1);
            if ((rc = fetch())) return rc;  // the one read-back of the rank
            // (the swap's choice, repeated: a next rank's term is back in)
            added = (t + 1 < k && host_st.early_stop < kEarlyStopRanks) || oiter < p->iters;
            inner_done = host_st.inner_done;
            reg -= host_out[4] + host_out[5];
            reg += host_out[0] + host_out[1];
            const int64_t q = int64_t(oiter - 1) * k + t;
            if (obj) obj[q] = host_out[2] + reg * p->lambda;
            if (rmse && has_test) rmse[q] = std::sqrt(host_out[3] / double(te_nnz));
            ++ranks_done;
        }
    }
    const double solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_solve).count();
    if (hipMemcpy(Wt.data(), d_Wt, Wt.size() * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(Ht.data(), d_Ht, Ht.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(NOMAD_EHIP, "nomad_ccd: copy of the factors failed");
    for (int i = 0; i < m; ++i)
        for (int t = 0; t < k; ++t) W[size_t(i) * k + t] = Wt[size_t(t) * m + i];
    for (int j = 0; j < n; ++j)
        for (int t = 0; t < k; ++t) H[size_t(j) * k + t] = Ht[size_t(t) * n + j];
    if (st) {
        st->ranks_done = ranks_done;
        st->inner_done = inner_done;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        st->solve_seconds = solve_s;
        // per inner iteration: both passes read index + residual of every
        // rating and write both vectors; per rank: one residual read + write
        // (the fused take-out and add-back), training (both orders) and test,
        // plus the first rank's add-back
        st->alg_bytes = double(inner_done) * (2.0 * double(nnz) * 12.0 + double(m + n) * 16.0) +
                        double(ranks_done + (ranks_done > 0 ? 1 : 0)) * double(2 * nnz + (has_test ? te_nnz : 0)) *
                            16.0;
    }
    return NOMAD_OK;
}
