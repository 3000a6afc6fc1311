/*
 * nomad_hip.h -- C ABI of the MI355X NOMAD SGD library (libnomad_hip.so).
 *
 * The reference exposes no library interface for its hot path: the SGD loop is
 * the private member NomadBody::updater_func (Code/nomad/nomad_body.h:214,
 * Code/nomad/nb_main.cpp:415-710), driven by NomadBody::run
 * (Code/nomad/nomad_body.h:278, Code/nomad/nb_main.cpp:30-410).  This header is
 * the boundary that replaces the pieces of that member a host driver needs, as
 * plain C with pointers and sizes:
 *
 *   nomad_create        NomadBody::initial + Pool / job-queue setup
 *                       (nb_init.cpp:189-235, nb_main.cpp:55-98)
 *   nomad_upload_train  the per-thread CSC of Data (nomad_body.h:34-46) built by
 *   nomad_upload_test   read_data (nb_data.cpp:19-176); here one global CSC
 *   nomad_set_factors   W block init (nb_main.cpp:444-456) and H token values
 *   nomad_get_factors   (nb_main.cpp:70-98) / the --output dump (nb_main.cpp:
 *                       350-366, 683-705)
 *   nomad_train         one timeout segment of the token loop
 *                       (nb_main.cpp:109-134 + 480-575)
 *   nomad_eval          the test/train SSE pass (nb_main.cpp:601-676, 228-241)
 *   nomad_destroy       teardown (nb_main.cpp:368-408)
 *
 * Conventions: every function returns 0 on success and a negative NOMAD_E*
 * code on failure; nomad_last_error() then returns a message for the calling
 * host thread.  No C++ exception crosses the boundary.  Calls are blocking
 * with respect to the host.  One context is driven by one host thread.
 */
#ifndef NOMAD_HIP_H_
#define NOMAD_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NOMAD_ABI_VERSION 1

/* storage type of the factor matrices W (m x k) and H (n x k) */
#define NOMAD_STORE_FP32 0 /* nomad_float  (Code/nomad/nomad_body.hpp:326-353) */
#define NOMAD_STORE_BF16 1 /* bf16 storage, fp32 arithmetic                       */
#define NOMAD_STORE_FP64 2 /* nomad_double (Code/nomad/nb_main.cpp:444,517)      */

/* schedule */
#define NOMAD_MODE_ASYNC 0 /* token passing through device ring queues        */
#define NOMAD_MODE_STRAT 1 /* deterministic B x B stratified schedule          */

/* rounding of bf16 stores */
#define NOMAD_ROUND_NEAREST 0
#define NOMAD_ROUND_STOCHASTIC 1

/* nomad_params.flags */
#define NOMAD_FLAG_BALANCE 1 /* row blocks balanced by training ratings
                                (nomad_row_blocks) instead of ceil(m / P) rows:
                                async workers (with the last wave of a
                                workgroup at reduced capacity), strat strata
                                (capacities all 1; results stay bit-exact
                                against the oracle given the same bounds) */
#define NOMAD_FLAG_TIMING 2  /* async: collect the wall-clock counters [8..13]
                                of nomad_get_counters (slows the kernel)    */

/* error codes */
#define NOMAD_OK 0
#define NOMAD_EINVAL -1   /* bad argument                           */
#define NOMAD_EHIP -2     /* HIP runtime error                      */
#define NOMAD_ESTATE -3   /* call out of order (e.g. no data)       */
#define NOMAD_ENOMEM -4   /* device allocation failed               */
#define NOMAD_ETIMEOUT -5 /* device watchdog stopped a kernel       */

typedef struct nomad_ctx nomad_ctx;

typedef struct nomad_params {
    int32_t m, n, k;   /* rows (users), columns (items), latent dimension (--dim,
                          1..256; above 128 the kernels hold two element pairs
                          per lane)                                               */
    int32_t storage;   /* NOMAD_STORE_*                                          */
    int32_t mode;      /* NOMAD_MODE_*                                           */
    int32_t rounding;  /* NOMAD_ROUND_* (bf16 only)                              */
    int32_t nworkers;  /* async: workers = row blocks (0: 8 per compute unit);
                          strat: strata B (0: one per co-resident wave)          */
    int32_t ptoken;    /* README --ptoken (README.md:90, "number of tokens in the
                          pipeline"; commented out in nomad_option.cpp:43-44):
                          async mode holds at most this many columns resident
                          per workgroup (0: no cap).  The workgroup's LDS
                          slots bound it: 128 for fp32 and bf16 at 8 waves per
                          workgroup, 64 at 4 or 16 waves, 32 for fp64 (k >
                          128: 64 at 8 waves, 32 otherwise)                 */
    int32_t device;    /* HIP device ordinal                                     */
    int32_t flags;     /* NOMAD_FLAG_* bits; others must be 0                    */
    double lrate;      /* --lrate  */
    double drate;      /* --drate  */
    double reg;        /* --reg    */
    uint64_t seed;     /* --seed: route permutations use mt19937_64(seed + 137)
                          and re-permutations mt19937_64(seed + 139), as the
                          reference does at one rank (nb_main.cpp:79, nb_init.cpp:219) */
} nomad_params;

typedef struct nomad_train_stats {
    int64_t updates;      /* rating updates in this call (u=, nb_main.cpp:284)       */
    int64_t visits;       /* non-empty (column, worker) visits in this call          */
    int64_t failed_pops;  /* empty polls of a worker queue (f=, nb_main.cpp:284)     */
    int64_t epochs_done;  /* completed passes of every column after the call         */
    double train_seconds; /* device time of the training kernels                     */
    double alg_bytes;     /* algorithmic HBM bytes of the call: nnz-updates *
                             (4 + s_r + 2 k s) + visits * 2 k s                       */
} nomad_train_stats;

typedef struct nomad_eval_stats {
    double train_sse;
    int64_t train_count;
    double test_sse;
    int64_t test_count;
    double seconds;
} nomad_eval_stats;

/* ---- context lifecycle ---- */
int nomad_create(const nomad_params* params, nomad_ctx** out);
int nomad_destroy(nomad_ctx* ctx);
const char* nomad_last_error(void);
int nomad_abi_version(void);

/* Global CSC of the training / test ratings: colptr[n+1], rowidx[nnz] with
 * rows ascending inside each column, val[nnz].  Host buffers; the library
 * copies them.  Must be called before nomad_train / nomad_eval. */
int nomad_upload_train(nomad_ctx* ctx, const int64_t* colptr, const int32_t* rowidx,
                       const double* val, int64_t nnz);
int nomad_upload_test(nomad_ctx* ctx, const int64_t* colptr, const int32_t* rowidx,
                      const double* val, int64_t nnz);

/* Row-major host factors, W m x k and H n x k, in double; converted to the
 * storage type on upload (bf16: round to nearest even) and back on download.
 * Either pointer may be NULL to skip that matrix. */
int nomad_set_factors(nomad_ctx* ctx, const double* W, const double* H);
int nomad_get_factors(nomad_ctx* ctx, double* W, double* H);

/* One training segment.  Async mode runs until every column has completed
 * `epochs` more passes over all row blocks (epochs > 0) or until `seconds` of
 * device time have elapsed (seconds > 0), whichever comes first; at least one
 * of the two must be positive.  Strat mode runs `epochs` full B x B sweeps. */
int nomad_train(nomad_ctx* ctx, double seconds, int64_t epochs, nomad_train_stats* stats);

/* Batched sweep (no reference counterpart: the reference runs one problem
 * per process; this replaces a loop over separate nomad_double runs of a
 * hyper-parameter sweep, e.g. over --reg / --dim).  Runs `epochs` epochs of
 * every context's own stratified schedule in ONE persistent launch.  The
 * contexts must be in NOMAD_MODE_STRAT, on one device, with one storage
 * type and uploaded training data; each keeps its own ratings, factors,
 * rank, strata count and hyper-parameters, and its result is bit-identical
 * to nomad_train(ctx, 0, epochs) on that context alone.  count <= 256.
 * waves = persistent grid in wavefronts (0: one per tile of the widest
 * round, capped at what fits on the device; any positive value is valid).
 * stats (may be NULL) receives one record per context; train_seconds is
 * the device time of the whole launch. */
int nomad_train_batch(nomad_ctx* const* ctxs, int32_t count, int64_t epochs, int32_t waves,
                      nomad_train_stats* stats);

/* Autotuner (no reference counterpart; the reference's worker count is the
 * --nthreads option).  Times short training trials of candidate launch
 * configurations on the given problem and returns the fastest:
 *   async mode: workers (row blocks) per compute unit in {2, 4, 6, 8, 12,
 *               16} (the workgroup shape and the ring capacity follow from
 *               the worker count), with and without NOMAD_FLAG_BALANCE;
 *   strat mode: strata B in {1/4, 1/2, 1, 2, 4} x compute units (at most
 *               m), with and without NOMAD_FLAG_BALANCE, timed through
 *               nomad_train_batch with the default grid.
 * params gives the problem and hyper-parameters; its nworkers and the
 * BALANCE flag are ignored.  Each trial trains a scratch context from
 * factors drawn from mt19937_64(params->seed), U(0, 1/sqrt(k)): async
 * trials run two segments of trial_seconds (the second is timed), strat
 * trials one warm-up and one timed epoch.  trials (may be NULL) receives up
 * to max_trials records in the order run; *ntrials the number run.  A
 * trial whose context cannot be created records its error code and is
 * skipped; the call fails only if every trial fails. */
typedef struct nomad_tune_trial {
    int32_t nworkers;      /* workers (async) or strata B (strat) actually used */
    int32_t flags;         /* NOMAD_FLAG_BALANCE or 0                           */
    int32_t status;        /* NOMAD_OK or the trial's error code                */
    int32_t pad;
    double updates_per_s;  /* timed rate (0 for a failed trial)                 */
} nomad_tune_trial;

int nomad_autotune(const nomad_params* params, const int64_t* colptr, const int32_t* rowidx, const double* val,
                   int64_t nnz, double trial_seconds, nomad_tune_trial* best, nomad_tune_trial* trials,
                   int32_t max_trials, int32_t* ntrials);

/* Batch autotuner (no reference counterpart): picks the strata count (and,
 * with flags_out, the row balance) of each problem of a batched sweep by
 * timing whole batched epochs (nomad_train_batch), since the problems then
 * share the device and the best choice differs from each problem's own
 * optimum.  Stage 1: one strata count for every problem in {64, 128, 256,
 * 512, 1024} (each capped at the problem's rows).  Each trial builds
 * scratch strat contexts of the problems (params as given, nworkers
 * replaced; factors drawn from mt19937_64(params.seed), U(0, 1/sqrt(k)))
 * and times one batched epoch after a warm-up epoch.  Stage 2 (greedy):
 * each problem's own epoch is timed at the common count; the problems
 * within 40 % of the slowest, slowest first, try 1x, 2x and 4x that count
 * with NOMAD_FLAG_BALANCE as given and (only when flags_out is not NULL)
 * toggled, and keep the move that shortens the batched epoch of all
 * problems most, if by at least 1 %.  strata_out[i] receives problem i's
 * chosen count, flags_out[i] (may be NULL) its params.flags with the chosen
 * balance bit, *updates_per_s the batch's rate with them.  count <= 256. */
typedef struct nomad_batch_problem {
    nomad_params params;  /* NOMAD_MODE_STRAT; one storage type and device for all */
    const int64_t* colptr;
    const int32_t* rowidx;
    const double* val;
    int64_t nnz;
} nomad_batch_problem;

int nomad_autotune_batch(const nomad_batch_problem* probs, int32_t count, int32_t* strata_out,
                         int32_t* flags_out, double* updates_per_s);

/* Sum of squared errors over the training and test ratings, accumulated in
 * FP64 (the dot product runs in the storage's arithmetic type). */
int nomad_eval(nomad_ctx* ctx, nomad_eval_stats* stats);

/* nomad_eval without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result as nomad_eval, bit for bit; the per-wave partial
 * sums go to a fixed 8 MiB device array of the library (one copy per GPU,
 * held until the process ends) instead of a buffer the context allocates,
 * so the training and the test set may each hold at most 536,868,864
 * ratings ((2^20 - 4) x 512; larger sets fail with NOMAD_EINVAL).  Calls
 * that overlap on one GPU (several host threads) race on that array. */
int nomad_eval_noalloc(nomad_ctx* ctx, nomad_eval_stats* stats);

/* Row blocks of NOMAD_FLAG_BALANCE: bounds[0..P] (bounds[0] = 0, bounds[P] =
 * m) split rows 0..m-1 into P contiguous non-empty blocks whose largest
 * training-rating count relative to the block's capacity cap[b] (in (0, 1];
 * null: all 1) is the smallest any contiguous split allows (with all
 * capacities 1 exactly; otherwise unless a block had to be split to reach P
 * blocks).  Replaces the reference's equal-row split of the rows over
 * machines and threads (Code/nomad/nb_data.cpp:83-89) when the flag is set;
 * host code only (no device is touched).  Needs 0 < P <= m. */
int nomad_row_blocks(const int32_t* rowidx, int64_t nnz, int32_t m, int32_t P, const double* cap,
                     int32_t* bounds);

/* Number of workers / strata the context actually uses (after auto sizing). */
int nomad_get_nworkers(nomad_ctx* ctx, int32_t* nworkers);

/* Device counters of the last async nomad_train call (no reference
 * counterpart; diagnostics): out[0] updates, [1] (column, wave) visits,
 * [2] idle polls, [3] ring claim attempts, [4] columns taken from rings,
 * [5] columns leaving a workgroup, [6] sum over claim attempts of resident
 * columns, [7] claim attempts that found published entries queued behind an
 * unpublished one, [8] wall-clock ticks spent in visits, [9] ticks spent
 * handing columns to the next workgroup, [10] ticks spent claiming columns,
 * [11] ticks inside the SGD chains, [12] ticks between a column's
 * publication on a ring and its claim, [13] ticks slots waited in wave
 * inboxes, [14] sum over work items of the wave's inbox depth, [15] work
 * items (streams of pieces, or single long pieces) a wave ran.  Counters
 * [8..13] are collected only with NOMAD_FLAG_TIMING.  Ticks run at the
 * device's constant wall-clock rate.  When n exceeds 16, out[16 ..]
 * receives the same 16 counters for each worker (wave) in turn, as far as n
 * allows.  n is the capacity of out. */
int nomad_get_counters(nomad_ctx* ctx, int64_t* out, int32_t n);

/* Route state of an async context.  The step size of a visit follows from the
 * number of earlier visits of the column to the same worker, a count the
 * reference keeps per (thread, column) (col_update_counts,
 * Code/nomad/nb_main.cpp:458-459, 513-515, 544); here the count is kept per
 * (column, workgroup) pair, since a workgroup's waves see a column in one
 * hop.  For column j, entries vptr[j] .. vptr[j+1]-1 list the workgroups
 * holding ratings of j in ascending order: unit[e] the workgroup, nratings[e]
 * the column's ratings in its rows, count[e] the completed visits of j to
 * it, route_pos[e] its position in the column's
 * current route (the reference's perm_, Code/nomad/pool.h:11-37).
 * tok_pos[j] is the route position of column j's next hop and tok_cycle[j]
 * its completed route cycles (pos_ and the epoch count of the reference's
 * ColumnData).  Any output may be NULL; *nentries (may be NULL) receives the
 * number of pairs V.  vptr has n+1 entries, unit/count/route_pos V,
 * tok_pos/tok_cycle n.  Builds the routes if no launch has yet. */
int nomad_get_route_state(nomad_ctx* ctx, int64_t* vptr, int32_t* unit, int32_t* nratings, int32_t* count,
                          int32_t* route_pos, int32_t* tok_pos, int32_t* tok_cycle, int64_t* nentries);

/* Restores the visit counters of an async context (a checkpoint's counts,
 * in the entry order of nomad_get_route_state); nentries must equal V. */
int nomad_set_visit_counts(nomad_ctx* ctx, const int32_t* count, int64_t nentries);

/* Epoch progress of a context, for checkpoints: strat mode the completed
 * B x B sweeps (the step-size count of the next one); async mode the route
 * cycles every column has completed (what nomad_train(epochs) counts from).
 * set also gives every column of an async context that many completed
 * cycles.  The reference checkpoints no counter (nb_checkpoint.cpp:45-143);
 * a resumed run needs it to continue the same step schedule. */
int nomad_get_progress(nomad_ctx* ctx, int64_t* epochs);
int nomad_set_progress(nomad_ctx* ctx, int64_t epochs);

/* ---- CCD++ rank-one solver (the reference's competitor program) ----------
 * Replaces ccdr1_mpi at one process (Code/competitor/ccdr1-mpi-omp.cpp:398-558:
 * RankOneUpdate :334-358, UpdateRating :360-385, the rank / inner-iteration
 * loop with its fundec early stop :430-530), in FP64 on the GPU.  W (m x k)
 * and H (n x k) are row-major host arrays, read as the start point and
 * overwritten with the result (the reference starts from H = 0).  obj and
 * rmse (may be NULL) receive, at [(iter-1)*k + t], the objective
 * loss + lambda * sum(nnz-weighted norms) and the test RMSE after rank t of
 * outer iteration iter, NaN where the early stop skipped the rank.  The test
 * CSC may be NULL / 0.  nnz < 2^31. */
typedef struct nomad_ccd_params {
    int32_t m, n, k;  /* rows, columns, rank (-k)              */
    int32_t iters;    /* outer iterations (-t)                 */
    int32_t inner;    /* inner iterations per rank (-T)        */
    int32_t device;   /* HIP device ordinal                    */
    double lambda;    /* -l; weighted by each line's nnz       */
    double eps;       /* -e; inner stop fundec < max * eps     */
} nomad_ccd_params;

typedef struct nomad_ccd_stats {
    int64_t ranks_done;    /* ranks updated (not skipped)             */
    int64_t inner_done;    /* inner iterations run                    */
    double seconds;        /* wall time of the call                   */
    double alg_bytes;      /* algorithmic bytes: per inner iteration
                              2 nnz (4 + 8) + 16 (m + n); per rank
                              32 (2 nnz + test nnz) (residual updates
                              of the CSC and CSR copies and the test
                              residual)                               */
    double solve_seconds;  /* wall time of the rank / inner-iteration
                              loop alone (after the host builds the
                              CSR, uploads and the residual init)     */
} nomad_ccd_stats;

int nomad_ccd(const nomad_ccd_params* params, const int64_t* colptr, const int32_t* rowidx, const double* val,
              int64_t nnz, const int64_t* te_colptr, const int32_t* te_rowidx, const double* te_val,
              int64_t te_nnz, double* W, double* H, double* obj, double* rmse, nomad_ccd_stats* stats);

/* ---- objective and gradient ---------------------------------------------
 * L = sum (w_i.h_j - r_ij)^2 + reg (sum_i deg(i) |w_i|^2 + sum_j deg(j) |h_j|^2),
 * the objective the reference's competitors print (calobj_mpi,
 * Code/competitor/sgd-mpi-omp.cpp:122-164), and its gradient
 * dL/dw_i = 2 sum_j e_ij h_j + 2 reg deg(i) w_i (likewise for h_j), e = w.h - r.
 * No reference function computes the gradient; it is the closed form.
 * nomad_obj_create copies the ratings (host CSC) to the device once.
 * nomad_obj_eval reads W (m x ldf) and H (n x ldf), ldf = k rounded up to
 * even, row-major DEVICE arrays of the context's storage type (padding
 * columns zero), writes the gradients to device arrays of the same shapes,
 * runs on `stream` (a hipStream_t, NULL: the null stream) and returns L on
 * the host after synchronising that stream. */
typedef struct nomad_obj_params {
    int32_t m, n, k;
    int32_t storage;  /* NOMAD_STORE_FP64 or NOMAD_STORE_FP32 (FP32 arithmetic,
                         FP64 sums of squares)                                 */
    int32_t device;
} nomad_obj_params;

typedef struct nomad_obj_ctx nomad_obj_ctx;

int nomad_obj_create(const nomad_obj_params* params, const int64_t* colptr, const int32_t* rowidx,
                     const double* val, int64_t nnz, nomad_obj_ctx** out);
int nomad_obj_eval(nomad_obj_ctx* ctx, const void* W, const void* H, double reg, void* gW, void* gH,
                   double* loss, void* stream);
int nomad_obj_destroy(nomad_obj_ctx* ctx);
/* The same evaluation for host arrays: W (m x k), H (n x k) row-major double
 * (converted to the storage type on the device), gradients returned in the
 * same shapes. */
int nomad_obj_eval_host(nomad_obj_ctx* ctx, const double* W, const double* H, double reg, double* gW,
                        double* gH, double* loss);

#ifdef __cplusplus
}
#endif

#endif /* NOMAD_HIP_H_ */
