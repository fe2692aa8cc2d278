/* bnmf_hip.h — C ABI of the MI355X-native multiplicative BLD solver.
 *
 * Every entry point takes plain pointers and sizes; tensors are dense
 * row-major (x rows, y columns, z fastest), matching the reference's Eigen
 * RowMajor types (src/defs.hpp:26-52). Functions return a bnmf_status; the
 * message of the last failure on the calling thread is bnmf_last_error().
 *
 * Citations are to eozd/bnmf-algs, relative to the reference root.
 */
#ifndef BNMF_HIP_H
#define BNMF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bnmf_status {
    BNMF_OK = 0,
    BNMF_ERR_INVALID_ARGUMENT = 1, /* BNMF_ASSERT failures (src/defs.hpp:12-20) */
    BNMF_ERR_OUT_OF_MEMORY = 2,
    BNMF_ERR_DEVICE = 3,           /* a HIP runtime call failed */
    BNMF_ERR_NOT_READY = 4,        /* serving: result not finished yet */
    BNMF_ERR_NOT_FOUND = 5,        /* serving: unknown ticket */
    BNMF_ERR_REJECTED = 6,         /* serving: request exceeds the HBM budget */
    BNMF_ERR_TIMEOUT = 7           /* serving: wait timed out */
} bnmf_status;

typedef enum bnmf_dtype { BNMF_F32 = 0, BNMF_F64 = 1 } bnmf_dtype;

/* Solver of a batch: bld::bld_mult (multiplicative updates) or bld::bld_add
 * (additive gradient ascent, src/bld/bld_add.hpp:69-184). */
typedef enum bnmf_algo { BNMF_ALGO_MULT = 0, BNMF_ALGO_ADD = 1 } bnmf_algo;

/* Message of the last failure on this thread ("" if none). */
const char* bnmf_last_error(void);
/* ABI version of this header (major * 100 + minor). */
int bnmf_abi_version(void);
/* Selects the HIP device used by this thread's later calls
 * (role of cuda::init, src/cuda/util.hpp:15-18). */
bnmf_status bnmf_init(int device);

/* ------------------------------------------------------------------------
 * Drop-in solver: bld::bld_mult / bld::bld_mult_cuda
 * (src/bld/bld_mult/bld_mult.hpp:83-86 and :226-229).
 * X: x*y; alpha: x; beta: z (Params::alpha / Params::beta,
 * src/alloc_model/alloc_model_params.hpp:25-84); S_out: x*y*z.
 * S0 (x*y*z) is the initial tensor; when NULL the solver draws it as
 * init_S does (bld_mult_funcs.hpp:36-74): S0_ijk = X_ij * Dirichlet(1_z)_k,
 * from a counter-based stream keyed by `seed`.
 * Exactly max_iter iterations run (the reference has no convergence test).
 * Parameter errors (X < 0 or NaN, null pointers) return
 * BNMF_ERR_INVALID_ARGUMENT, as check_bld_params asserts
 * (src/bld/util_details.hpp:20-28). As there, z has no upper bound (fibers
 * deeper than 64 run a wavefront-per-fiber pass); z = 0, like x = 0 or
 * y = 0, is an empty instance. x, y and z must each be below 2^31.
 * ---------------------------------------------------------------------- */
/* Plain FP64 form: every step in double, exact digamma unless use_psi_appr. */
bnmf_status bnmf_bld_mult_f64(const double* X, size_t x, size_t y, size_t z,
                              const double* alpha, const double* beta,
                              size_t max_iter, int use_psi_appr, double eps,
                              const double* S0, uint64_t seed, double* S_out);
/* FP32-bulk form: S, the update and the axis sums in float with compensated
 * cross-tile sums; X/alpha/beta as given. */
bnmf_status bnmf_bld_mult_f32(const float* X, size_t x, size_t y, size_t z,
                              const double* alpha, const double* beta,
                              size_t max_iter, int use_psi_appr, double eps,
                              const float* S0, uint64_t seed, float* S_out);

/* ------------------------------------------------------------------------
 * No-allocation variant of bnmf_bld_mult_f32. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and the same kernels as bnmf_bld_mult_f32 (S
 * is bitwise identical), but X (x*y), alpha (x), beta (z), S0 (x*y*z, or
 * NULL: drawn from `seed` as bnmf_bld_mult_f32 draws it) and S_out (x*y*z)
 * are device buffers the caller owns (X, S0 and S_out 16-byte aligned), and
 * the solve is enqueued on `stream` (a hipStream_t; NULL: the default
 * stream) without waiting. The workspace (digamma rows, partial axis sums,
 * descriptors, work lists) is a fixed 32 MiB device array inside the
 * library, held by each GPU it runs on until the process ends; an instance
 * that needs more returns BNMF_ERR_INVALID_ARGUMENT (the partial axis sums
 * dominate: about x*y*z/16 + 4*y*z*ceil(x/128) bytes, so 2048 x 2048 x 64
 * needs about 25 MiB and fits). Nothing is
 * allocated: no hipMalloc, hipFree or allocating library call. Every call
 * on a GPU shares that array and leaves its contents behind; calls that
 * overlap on one GPU (other streams or threads) race. X is not checked on
 * the host (it is device memory): it must be nonnegative, as
 * check_bld_params requires (src/bld/util_details.hpp:20-28).
 * ---------------------------------------------------------------------- */
bnmf_status bnmf_bld_mult_f32_noalloc(const float* X, size_t x, size_t y, size_t z,
                                      const double* alpha, const double* beta, size_t max_iter,
                                      int use_psi_appr, double eps, const float* S0, uint64_t seed,
                                      float* S_out, void* stream);

/* ------------------------------------------------------------------------
 * bld::bld_add (src/bld/bld_add.hpp:69-184): additive gradient ascent on
 * Z = log S with S = X * softmax_k(Z), step 0.1 / (iteration + 1)^0.55.
 * Same arguments as bnmf_bld_mult_*; every digamma is psi_appr, as in the
 * reference (no use_psi_appr argument there either). The initial tensor is
 * drawn as in bld_mult (the reference draws it inline, :76-91).
 * ---------------------------------------------------------------------- */
bnmf_status bnmf_bld_add_f64(const double* X, size_t x, size_t y, size_t z,
                             const double* alpha, const double* beta, size_t max_iter,
                             double eps, const double* S0, uint64_t seed, double* S_out);
/* FP32 form: the step in float; exp(log S + t) is evaluated as S exp(t), and
 * a fiber whose exp(Z) row is all zero (X_ij = 0) updates to 0. */
bnmf_status bnmf_bld_add_f32(const float* X, size_t x, size_t y, size_t z,
                             const double* alpha, const double* beta, size_t max_iter,
                             double eps, const float* S0, uint64_t seed, float* S_out);

/* ------------------------------------------------------------------------
 * bld::bld_appr (src/bld/bld_appr.hpp:287-372): S kept in factored form
 * S_ijk = X_ij nu_ik mu_kj / (nu mu + eps)_ij, with multiplicative updates of
 * nu (x*z, rows sum to 1) and mu (z*y, columns sum to 1) in turn. nu0 / mu0
 * (NULL: rows of nu and columns of mu drawn from Dirichlet(1_z) with `seed`,
 * as the reference draws them, :302-318). Outputs S (x*y*z), nu, mu; any may
 * be NULL. Every digamma is psi_appr (FP64) or the FP32 digamma. An empty X
 * (x*y = 0) runs no iteration: nu and mu are returned as initialised.
 * ---------------------------------------------------------------------- */
bnmf_status bnmf_bld_appr_f64(const double* X, size_t x, size_t y, size_t z,
                              const double* alpha, const double* beta, size_t max_iter,
                              double eps, const double* nu0, const double* mu0, uint64_t seed,
                              double* S_out, double* nu_out, double* mu_out);
bnmf_status bnmf_bld_appr_f32(const float* X, size_t x, size_t y, size_t z,
                              const double* alpha, const double* beta, size_t max_iter,
                              double eps, const float* nu0, const float* mu0, uint64_t seed,
                              float* S_out, float* nu_out, float* mu_out);

/* Many bld_appr instances resident on the device (no reference counterpart):
 * the factors stay in HBM, each iteration is a replayed graph. */
typedef struct bnmf_appr_desc {
    size_t x, y, z;
    const void* X;        /* host, x*y, of the dtype */
    const double* alpha;  /* host, x */
    const double* beta;   /* host, z */
    const void* nu0;      /* host x*z or NULL */
    const void* mu0;      /* host z*y or NULL */
    uint64_t seed;        /* factor stream when nu0 / mu0 is NULL */
} bnmf_appr_desc;
typedef struct bnmf_appr bnmf_appr;
bnmf_status bnmf_appr_create(bnmf_appr** out, int dtype, const bnmf_appr_desc* descs, size_t n,
                             double eps);
/* Restarts every instance from nu0 / mu0 and runs max_iter iterations. */
bnmf_status bnmf_appr_solve(bnmf_appr* a, size_t max_iter);
/* S (formed from the current factors), nu, mu of instance idx; any may be NULL. */
bnmf_status bnmf_appr_get(bnmf_appr* a, size_t idx, void* S_out, void* nu_out, void* mu_out);
/* Mean milliseconds of one iteration (graph replays timed with HIP events);
 * advances the factors. */
bnmf_status bnmf_appr_time_iteration(bnmf_appr* a, int reps, double* iter_ms);
void bnmf_appr_destroy(bnmf_appr* a);

/* alloc_model::log_marginal_S (src/alloc_model/alloc_model_funcs.hpp:574-586),
 * evaluated on the GPU in FP64. S: x*y*z of the given dtype. */
bnmf_status bnmf_log_marginal_S(int dtype, const void* S, size_t x, size_t y, size_t z,
                                const double* alpha, const double* beta, double a,
                                double b, double* out);

/* bld::bld_fact (src/bld/bld_fact.hpp:36-80) of a host tensor S (x*y*z of
 * the given dtype): W (x*z), H (z*y), L (y) in double; axis sums and
 * normalisation on the GPU. */
bnmf_status bnmf_bld_fact(int dtype, const void* S, size_t x, size_t y, size_t z, const double* alpha,
                          const double* beta, double a, double b, double eps, double* W, double* H,
                          double* L);

/* cuda::tensor_sums (src/cuda/tensor_ops.hpp:40-42, cuda_src/tensor_ops.cu:10-55):
 * S_pjk (y*z) = sum over axis 0, S_ipk (x*z) = axis 1, S_ijp (x*y) = axis 2.
 * Any output may be NULL. */
bnmf_status bnmf_tensor_sums(int dtype, const void* S, size_t x, size_t y, size_t z,
                             void* S_pjk, void* S_ipk, void* S_ijp);

/* ------------------------------------------------------------------------
 * Batched solver: many independent instances of uneven (x, y, z) advance
 * together, one iteration per launch pair (prep + fused pass) over a device
 * work list; the per-iteration chain is replayed from a captured hipGraph.
 * Inputs are uploaded once; S stays resident until collected.
 * No counterpart in the reference (it solves one instance per call).
 * ---------------------------------------------------------------------- */
typedef struct bnmf_bld_desc {
    size_t x, y, z;
    const void* X;        /* host, x*y, of the batch dtype */
    const double* alpha;  /* host, x */
    const double* beta;   /* host, z */
    double a, b;          /* Gamma prior, used by the log-marginal only */
    const void* S0;       /* host, x*y*z of the batch dtype, or NULL */
    uint64_t seed;        /* S0 stream when S0 is NULL */
} bnmf_bld_desc;

typedef struct bnmf_batch bnmf_batch;

bnmf_status bnmf_batch_create(bnmf_batch** out, int dtype, const bnmf_bld_desc* descs,
                              size_t n, int use_psi_appr, double eps);
/* A batch of the given solver (bnmf_algo); bnmf_batch_create is
 * BNMF_ALGO_MULT. With BNMF_ALGO_ADD use_psi_appr is ignored. */
bnmf_status bnmf_batch_create_algo(bnmf_batch** out, int dtype, int algo,
                                   const bnmf_bld_desc* descs, size_t n, int use_psi_appr,
                                   double eps);
/* Restarts every instance from its S0 and runs max_iter iterations. */
bnmf_status bnmf_batch_solve(bnmf_batch* b, size_t max_iter);
/* Same, enqueued on the batch's stream without waiting. */
bnmf_status bnmf_batch_solve_async(bnmf_batch* b, size_t max_iter);
/* Optional convergence check (no reference counterpart: the reference runs
 * exactly max_iter iterations): restarts from S0 and runs at most max_iter
 * iterations; every check_every iterations the FP64 log_marginal_S of every
 * instance is evaluated and the solve stops once, for each instance,
 * |lm - lm_previous| <= tol * |lm|. *iters_run receives the iterations run
 * and lm_out (n values, may be NULL) the final log-marginals. A check costs
 * about two passes over S. With tol = 0 the result equals bnmf_batch_solve's. */
bnmf_status bnmf_batch_solve_tol(bnmf_batch* b, size_t max_iter, double tol, size_t check_every,
                                 size_t* iters_run, double* lm_out);
bnmf_status bnmf_batch_synchronize(bnmf_batch* b);
/* Copies instance idx's current S (x*y*z, batch dtype) to host memory. */
bnmf_status bnmf_batch_get_S(bnmf_batch* b, size_t idx, void* S_out);
/* FP64 log_marginal_S of every instance's current S; out has n entries. */
bnmf_status bnmf_batch_log_marginal(bnmf_batch* b, double* out);
/* Device-side timing: mean milliseconds of one update pass (all its kernel
 * launches) and of one whole iteration (prep + pass), each from a captured
 * graph of `reps` repetitions replayed as solves replay theirs and timed
 * with HIP events on the batch stream. Advances S by 4*reps passes; solve
 * again (from S0) to restore a result. */
bnmf_status bnmf_batch_time_pass(bnmf_batch* b, int reps, double* pass_ms, double* iter_ms);

/* bld::bld_fact (src/bld/bld_fact.hpp:36-80) of instance idx's current S:
 * W (x*z), H (z*y), L (y), in double, host arrays (any may be NULL). After
 * a solve the axis sums it needs are already in the workspace (the last
 * pass accumulates them), so no pass over S runs; otherwise one sums pass. */
bnmf_status bnmf_batch_fact(bnmf_batch* b, size_t idx, double eps, double* W, double* H, double* L);

typedef struct bnmf_batch_info {
    size_t n_instances;
    size_t elements;        /* sum of x*y*z */
    size_t fibers;          /* sum of x*y */
    size_t device_bytes;    /* HBM held by the batch */
    size_t pass_items;      /* tiles per fused pass */
    size_t prep_items;
    int n_buckets;          /* distinct z buckets (one pass launch each) */
} bnmf_batch_info;
bnmf_status bnmf_batch_get_info(bnmf_batch* b, bnmf_batch_info* info);
void bnmf_batch_destroy(bnmf_batch* b);

/* ------------------------------------------------------------------------
 * Tile table (written by the autotuner, bnmf_hip.autotune): rows per pass
 * tile for the shape class of an instance. z classes end at z = 4, 8, 12,
 * 16, 24, 32, 48, 64; size classes split fibers = x*y at 2^14, 2^16, 2^18,
 * 2^20. rows = 0 selects the built-in rule. The table is loaded on first use
 * from $BNMF_TUNE_FILE, else from tune/gfx950.txt next to the library's
 * directory. An instance is planned when its batch or request is created,
 * so a table change affects later solves only. No counterpart in the
 * reference (its kernels use fixed block shapes).
 * ---------------------------------------------------------------------- */
bnmf_status bnmf_tune_set(int dtype, size_t z, size_t fibers, unsigned rows);
bnmf_status bnmf_tune_get(int dtype, size_t z, size_t fibers, unsigned* rows);
bnmf_status bnmf_tune_load(const char* path);
bnmf_status bnmf_tune_save(const char* path);
/* Rows per tile an instance of this shape gets now (table or rule; 16 for
 * z > 64, whose wide pass has fixed 16 x 16 tiles); 0 for a shape with no
 * tiles (x, y or z = 0) or outside the solver's range (an unknown dtype, x,
 * y or z at or above 2^31). The table covers z in [1, 64]. */
unsigned bnmf_tile_rows(int dtype, size_t x, size_t y, size_t z);
/* Path of the loaded table file, "" if none (valid until the next load). */
const char* bnmf_tune_source(void);

/* ------------------------------------------------------------------------
 * Serving: a continuous-batching solve server. Requests are admitted in
 * arrival order while their HBM footprint (S plus the per-instance
 * workspace, bnmf_server_request_bytes) fits in the budget; a request larger
 * than the whole budget is rejected (BNMF_ERR_REJECTED from collect). Active
 * requests advance together; new ones join at the next scheduling step.
 * S stays in HBM until collected. When a request has no S0 it is drawn on
 * the GPU from the same stream as bnmf_init_S (seed). The request's host
 * buffers are copied at submit time. No counterpart in the reference.
 * ---------------------------------------------------------------------- */
typedef struct bnmf_server bnmf_server;

typedef struct bnmf_request_stats {
    double submit_s, admit_s, done_s; /* steady-clock seconds */
    size_t iterations;
    size_t device_bytes;
} bnmf_request_stats;

typedef struct bnmf_server_stats {
    size_t submitted, completed, rejected, active, queued, steps;
    size_t bytes_in_use, peak_bytes, peak_active;
    double busy_s; /* wall time the scheduler spent in device steps */
} bnmf_server_stats;

bnmf_status bnmf_server_create(bnmf_server** out, int dtype, size_t hbm_budget_bytes,
                               size_t max_active, int use_psi_appr, double eps);
bnmf_status bnmf_server_submit(bnmf_server* s, const bnmf_bld_desc* req, size_t max_iter,
                               int want_log_marginal, uint64_t* ticket);
/* *done = 1 once the request has finished (or failed). */
bnmf_status bnmf_server_poll(bnmf_server* s, uint64_t ticket, int* done);
/* Waits up to timeout_s (< 0: no limit); BNMF_ERR_TIMEOUT if still running;
 * otherwise the request's own status. */
bnmf_status bnmf_server_wait(bnmf_server* s, uint64_t ticket, double timeout_s);
/* Copies S (x*y*z) and the FP64 log-marginal (NaN unless requested) of a
 * finished request, frees its HBM and forgets the ticket. Any output may be
 * NULL. */
bnmf_status bnmf_server_collect(bnmf_server* s, uint64_t ticket, void* S_out,
                                double* log_marginal_out, bnmf_request_stats* stats);
bnmf_status bnmf_server_get_stats(bnmf_server* s, bnmf_server_stats* out);
/* HBM bytes a request of this shape holds while admitted (0 if x, y or z is
 * 2^31 or more). */
size_t bnmf_server_request_bytes(int dtype, size_t x, size_t y, size_t z);
void bnmf_server_destroy(bnmf_server* s);

/* ------------------------------------------------------------------------
 * Input synthesis (host): the allocation model of bnmf_priors / sample_S
 * (src/alloc_model/alloc_model_funcs.hpp:463-552) from a counter-based
 * Philox4x32-10 stream, so the output does not depend on the thread count.
 * W[:,k] ~ Dir(alpha), H[:,j] ~ Dir(beta), L_j ~ Gamma(a, scale b),
 * S_ijk ~ Poisson(W_ik H_kj L_j).
 * ---------------------------------------------------------------------- */
/* X = S_ij+ ~ Poisson(L_j * sum_k W_ik H_kj)  (the sum of the z independent
 * Poisson draws has exactly this law). X_out: x*y. */
bnmf_status bnmf_sample_X(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                          const double* beta, double a, double b, double* X_out);
/* The full tensor S (x*y*z) as sample_S draws it. */
bnmf_status bnmf_sample_S(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                          const double* beta, double a, double b, double* S_out);
/* The same draws with the Poisson variates drawn on the GPU (one thread per
 * variate, the host's Philox streams; the priors W, H, L on the host). Equal
 * to bnmf_sample_X / bnmf_sample_S unless a device exp / log / lgamma differs
 * from the host's in the last bit at an acceptance boundary. */
bnmf_status bnmf_sample_X_gpu(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                              const double* beta, double a, double b, double* X_out);
bnmf_status bnmf_sample_S_gpu(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                              const double* beta, double a, double b, double* S_out);
/* init_S (bld_mult_funcs.hpp:36-74): S0_ijk = X_ij * Dirichlet(1_z)_k. */
bnmf_status bnmf_init_S(int dtype, uint64_t seed, const void* X, size_t x, size_t y, size_t z,
                        void* S0_out);

/* Evaluates the device digamma forms on the GPU (kind 0: FP32 bulk psi,
 * 1: exact FP64, 2: util::psi_appr in FP64, src/util/util.hpp:440-514,
 * 3: FP32 psi(s+1) of the update passes evaluated at x = s+1, 4: FP64
 * psi(s+1) of the update passes evaluated at x = s+1).
 * x and out are host arrays of n doubles. */
bnmf_status bnmf_psi_eval(int kind, const double* x, double* out, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* BNMF_HIP_H */
