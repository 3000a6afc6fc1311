/*
 * CPU oracle for the NOMAD SGD path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference's algorithm, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product (nomad_amd/) never links or calls it.
 *
 * Each function names the reference lines it follows.  Pinning (what checks
 * each function against something other than itself) is listed in the header
 * of nomad_oracle.cpp.
 *
 * Build: g++ -O2 -ffp-contract=off (no -march), so float/double arithmetic is
 * the IEEE sequence written in the source, like the reference's own build
 * (Code/nomad/CMakeLists.txt:9,40: no -march, so no FMA contraction).
 */
#ifndef NOMAD_ORACLE_H_
#define NOMAD_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* storage */
#define ORC_FP32 0
#define ORC_BF16 1
#define ORC_FP64 2
/* arithmetic order */
#define ORC_ARITH_REF 0 /* reference: sequential fold from -r, no fused ops   */
#define ORC_ARITH_GPU 1 /* the GPU kernels' 64-lane tree and fused updates    */
/* bf16 rounding */
#define ORC_ROUND_NEAREST 0
#define ORC_ROUND_STOCHASTIC 1

/* --- data file (Code/nomad/nb_data.cpp:19-176) ---------------------------- */
/* Sizes of part `part` of `nparts`; returns 0 or a negative error. */
int orc_dat_part_info(const char* path, int part, int nparts, int* nrows, int* ncols, int64_t* total_nnz,
                      int* min_row, int* local_rows, int64_t* local_nnz);
/* CSC of that part with local row ids: col_offset[ncols+1] (int, as the
 * reference), row_idx[local_nnz], row_val[local_nnz]. */
int orc_dat_part_read(const char* path, int part, int nparts, int* col_offset, int* row_idx, double* row_val);

/* --- initialisation (nb_main.cpp:70-98, 444-456; pool.cpp:32-38) ------------ */
int orc_init_factors(int m, int n, int k, int nthreads, uint64_t seed, double* W, double* H);
/* initial per-column permutations of 0..nthreads-1: perms[n * nthreads] */
int orc_init_perms(int n, int nthreads, uint64_t seed, int* perms);

/* --- step size (nb_main.cpp:514-515; nomad_body.hpp:313-320 for float) ---- */
double orc_step(int storage, double lrate, double drate, int count);

/* --- deterministic SGD replay --------------------------------------------- *
 * Global CSC (rows ascending in a column).  W (m x k), H (n x k) are double
 * on input/output and held in `storage` precision during the replay.
 * Schedule strat(B): for sub-epoch s, tiles (b, (b+s) % B) with row blocks of
 * ceil(m/B) rows, or rows [row_bounds[b], row_bounds[b+1]) when row_bounds
 * (B + 1 entries, 0 = first <= ... <= last = m) is given, and column blocks
 * of ceil(n/B) columns; inside a tile,
 * columns ascending, rows ascending.  B = 1 is the reference's one-thread
 * order (columns 0..n-1 cyclically).  The step of epoch e uses count
 * epoch0 + e.  Returns the number of updates, or a negative error. */
int64_t orc_sgd_strat(int storage, int arith, int rounding, uint32_t sr_seed, int m, int n, int k,
                      const int64_t* colptr, const int32_t* rowidx, const double* val, double* W, double* H,
                      int B, const int32_t* row_bounds, int epoch0, int nepochs, double lrate, double drate,
                      double reg);

/* --- one BF16 store of an element pair, as the kernels do it -------------- */
void orc_bf16_store_pair(float x0, float x1, int rounding, uint32_t seed, uint32_t a, uint32_t b, uint32_t lane,
                         float* out);

/* --- asynchronous NOMAD on CPU threads (the CPU baseline) ------------------ *
 * T threads, one row part each; column tokens in per-thread FIFO queues with
 * the reference's routing (nb_main.cpp:480-575).  Stops after every column
 * completes `epochs` passes (epochs > 0) or after `seconds` (seconds > 0).
 * Storage fp32, fp64, or bf16 with round to nearest (fp32 arithmetic; each
 * row rounded after its update, each column at the end of its visit);
 * reference arithmetic. */
typedef struct orc_async_stats {
    int64_t updates;
    int64_t failures;
    double seconds;
    int64_t min_cycles;
} orc_async_stats;
int orc_sgd_async(int storage, int T, int m, int n, int k, const int64_t* colptr, const int32_t* rowidx,
                  const double* val, double* W, double* H, uint64_t seed, double lrate, double drate,
                  double reg, double seconds, int64_t epochs, orc_async_stats* stats);

/* The same run as a state kept across timeout segments (nb_main.cpp:109-328):
 * each thread's W block and its visit count per column (col_update_counts,
 * nb_main.cpp:458-459, 544) persist; a segment stops after `seconds` (> 0),
 * or once `max_updates` (> 0) updates have been made (each thread completes
 * the visit it is in: at T = 1 the segment ends with the visit that reaches
 * the budget), or when every column has completed `epochs` (> 0) more route
 * cycles than the least advanced one had at the start; then every column is
 * gathered (queues in thread order, then parked columns) and given a fresh
 * route with set_perm from mt19937_64(seed + 139) (pool.cpp:32-38,
 * nb_init.cpp:219, nb_main.cpp:313-321).  orc_sgd_async is one segment of a
 * fresh state.  counts[t * n + j]: thread t's visits of column j. */
typedef struct orc_async orc_async;
orc_async* orc_async_create(int storage, int T, int m, int n, int k, const int64_t* colptr, const int32_t* rowidx,
                            const double* val, const double* W, const double* H, uint64_t seed, double lrate,
                            double drate, double reg);
int orc_async_segment(orc_async* s, double seconds, int64_t max_updates, int64_t epochs, orc_async_stats* stats);
int orc_async_get(const orc_async* s, double* W, double* H);
int orc_async_counts(const orc_async* s, int32_t* counts);
void orc_async_destroy(orc_async* s);

/* --- evaluation (nb_main.cpp:601-667) -------------------------------------- *
 * sum over ratings of (<h_j, w_i> - r)^2 in double, the dot in `storage`
 * precision with `arith` order. */
double orc_sse(int storage, int arith, int m, int n, int k, const int64_t* colptr, const int32_t* rowidx,
               const double* val, const double* W, const double* H);

/* --- objective and gradient in FP64 ----------------------------------------- *
 * L = sum_{(i,j)} (w_i.h_j - r)^2 + reg * (sum_i deg(i) |w_i|^2 + sum_j deg(j) |h_j|^2)
 * (Code/competitor/sgd-mpi-omp.cpp:122-164 calobj_mpi form); gW, gH may be
 * NULL. */
double orc_loss_grad(int m, int n, int k, const int64_t* colptr, const int32_t* rowidx, const double* val,
                     const double* W, const double* H, double reg, double* gW, double* gH);

/* --- CCD++ rank-one sweeps (Code/competitor/ccdr1-mpi-omp.cpp:334-558) ------ *
 * W (m x k) and H (n x k) row-major in double; `iters` outer iterations of
 * `inner` inner iterations per rank, early stop with `eps`; lambda is weighted
 * by the row/column nnz as in RankOneUpdate.  obj[(iter-1)*k + t] receives
 * the objective printed after rank t of outer iteration iter (NaN for ranks
 * skipped by the early stop); obj may be NULL.  With a test CSC (te_*),
 * rmse[(iter-1)*k + t] receives the test RMSE at the same points; te_* and
 * rmse may be NULL. */
int orc_ccd(int m, int n, int k, const int64_t* colptr, const int32_t* rowidx, const double* val,
            double* W, double* H, double lambda, int iters, int inner, double eps, double* obj,
            const int64_t* te_colptr, const int32_t* te_rowidx, const double* te_val, double* rmse);

#ifdef __cplusplus
}
#endif

#endif
