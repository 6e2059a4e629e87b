/*
 * mipx_pdlp.h -- C-ABI of the MI355X PDLP library (libmipx_pdlp.so).
 *
 * This is the drop-in boundary for mipx's GPU PDLP path.  In the reference
 * the GPU path is `PdlpSolver::solveGpu()` (src/lp/pdlp.cpp:1739-2792) plus
 * the kernels of src/lp/pdlp_kernels.cu / .cuh and cuSPARSE SpMV calls
 * (src/lp/pdlp.cpp:1899-1978).  Everything here uses plain C types: no HIP,
 * CUDA or C++ types cross the boundary and every function returns an int
 * error code (MPX_OK == 0); the last error text of the calling thread is
 * available from mpx_last_error().
 *
 * Numbers: values are IEEE double (the reference's `Real`, include/mipx/core.h:7),
 * indices are int32 (`Index`, include/mipx/core.h:9).  Status codes follow the
 * order of `mipx::Status` (include/mipx/core.h:11-19).
 */
#ifndef MIPX_PDLP_H
#define MIPX_PDLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define MPX_OK 0
#define MPX_ERR_INVALID 1      /* bad argument (sizes, null pointers, ...) */
#define MPX_ERR_DEVICE 2       /* a HIP runtime call failed */
#define MPX_ERR_OUT_OF_MEMORY 3
#define MPX_ERR_UNSUPPORTED 4  /* option combination not implemented */
#define MPX_ERR_BUSY 5         /* serving loop: no slot / budget free */

/* ---- solver status (same order as mipx::Status, core.h:11-19) ---------- */
#define MPX_STATUS_OPTIMAL 0
#define MPX_STATUS_INFEASIBLE 1
#define MPX_STATUS_UNBOUNDED 2
#define MPX_STATUS_ITER_LIMIT 3
#define MPX_STATUS_TIME_LIMIT 4
#define MPX_STATUS_NODE_LIMIT 5
#define MPX_STATUS_ERROR 6

/* ---- precision modes (MI355X extension; FP64 is the reference's) ------- */
#define MPX_PREC_FP64 0   /* every iteration in FP64 (the reference's GPU path) */
#define MPX_PREC_MIXED 1  /* FP32 iterations with FP32 accumulation; FP64 eval,
                             termination and a final FP64 refinement */

/*
 * An LP in the reference's ranged form (include/mipx/lp_problem.h:14-64):
 *   min/max  obj'x + obj_offset   s.t.  row_lower <= A x <= row_upper,
 *                                       col_lower <= x <= col_upper.
 * A is CSR with column indices sorted within each row
 * (include/mipx/sparse_matrix.h:36-42).  Infinite bounds are +-INFINITY.
 */
typedef struct mpx_lp {
    int32_t num_rows;
    int32_t num_cols;
    int32_t nnz;
    int32_t sense;                /* 0 = Minimize, 1 = Maximize (core.h:21-24) */
    const int32_t* row_starts;    /* [num_rows + 1] */
    const int32_t* col_indices;   /* [nnz] */
    const double* values;         /* [nnz] */
    const double* obj;            /* [num_cols] */
    const double* col_lower;      /* [num_cols] */
    const double* col_upper;      /* [num_cols] */
    const double* row_lower;      /* [num_rows] */
    const double* row_upper;      /* [num_rows] */
    double obj_offset;
} mpx_lp;

/* Cooperative cancellation: polled between iteration blocks, like the
 * reference's `stop_flag` (include/mipx/pdlp.h:86, src/lp/pdlp.cpp:2136-2140).
 * Return nonzero to stop; the solve then reports MPX_STATUS_ITER_LIMIT. */
typedef int (*mpx_stop_fn)(void* user);

/*
 * Mirrors `mipx::PdlpOptions` (include/mipx/pdlp.h:19-87) field by field,
 * same names and defaults (mpx_pdlp_default_options fills them).  Booleans
 * are int32 0/1.  The fields after `stop_user` are MI355X extensions.
 */
typedef struct mpx_pdlp_options {
    int32_t max_iter;                 /* 4000000 */
    double optimality_tol;            /* 5e-5 */
    double primal_tol;                /* 5e-5 */
    double dual_tol;                  /* 5e-5 */
    double max_solve_seconds;         /* -1 = no limit */
    int32_t optimality_norm;          /* 0 = L2, 1 = LInf */
    int32_t sv_max_iter;              /* 5000 */
    double sv_tol;                    /* 1e-4 */
    int32_t adaptive_step;            /* 0 */
    double adaptive_step_down;        /* 0.7 */
    double adaptive_step_up;          /* 1.0 */
    double adaptive_step_min;         /* 1e-6 */
    double adaptive_step_max;         /* 1e3 */
    double primal_weight;             /* 1.0 */
    int32_t update_primal_weight;     /* 1 */
    double pid_kp;                    /* 0.99 */
    double pid_ki;                    /* 0.01 */
    double pid_kd;                    /* 0.0 */
    double pid_i_smooth;              /* 0.3 */
    double restart_sufficient_decay;  /* 0.2 */
    double restart_necessary_decay;   /* 0.5 */
    double restart_artificial_fraction; /* 0.36 */
    int32_t termination_eval_frequency; /* 200 */
    double matrix_zero_tol;           /* 0.0 */
    int32_t enable_polishing;         /* 1 */
    double polish_primal_tol;         /* 1e-6 */
    double polish_dual_tol;           /* 1e-6 */
    int32_t polish_max_iter;          /* 500000 */
    int32_t do_ruiz_scaling;          /* 1 */
    int32_t ruiz_iterations;          /* 10 */
    int32_t do_pock_chambolle_scaling; /* 1 */
    int32_t do_bound_obj_rescaling;   /* 1 */
    int32_t preconditioner_refresh;   /* 0 */
    double preconditioner_refresh_ratio; /* 10.0 */
    int32_t anderson_acceleration;    /* 0 */
    int32_t anderson_window_size;     /* 5 */
    double anderson_regularization;   /* 1e-10 */
    int32_t use_gpu;                  /* 1 (this library always runs on the GPU) */
    int32_t gpu_min_rows;             /* 512 (gate evaluated by the caller) */
    int32_t gpu_min_nnz;              /* 10000 (gate evaluated by the caller) */
    int32_t verbose;                  /* 1 */
    mpx_stop_fn stop_fn;              /* NULL */
    void* stop_user;                  /* NULL */
    /* ---- MI355X extensions ---- */
    int32_t precision;                /* MPX_PREC_FP64 */
    double refine_tol;                /* mixed precision: after the handover
                                         (see mixed_handover) the refinement
                                         runs until primal, dual and gap are
                                         below refine_tol (<= 0: the solve's
                                         own tolerances) */
    int32_t spmv_group;               /* lanes per row/column in the SpMV
                                         kernels, 0 = autotuned.  1 gives the
                                         sequential per-row summation order of
                                         the reference's CPU SpMV. */
    int32_t device;                   /* HIP device ordinal, default 0 */
    double mixed_handover;            /* mixed precision: FP32 hands over to
                                         FP64 once max(primal/primal_tol,
                                         dual/dual_tol, gap/optimality_tol),
                                         evaluated in FP64, is at most this
                                         (default 3), or when FP32 stalls */
    int32_t spmv_chunk_cols;          /* single LP: the dual step's A x_r in
                                         column chunks of this many columns,
                                         each chunk's x_r slice served from
                                         L2 (partial sums added chunk by
                                         chunk); 0 = automatic (chunks of
                                         2^19 columns when n > 2^20), < 0 =
                                         never */
    int32_t packed_preset;            /* batched / served LPs in the packed
                                         kernels: lanes per column and per row
                                         (preset 0..7: (1,1) (2,2) (2,4) (4,4)
                                         (4,8) (8,8) (8,16) (16,16)); -1 = the
                                         autotuned table for the LP's shape
                                         (mean entries per column / row), the
                                         nearest-shape heuristic where the
                                         table has no entry */
    int32_t mixed_refine;             /* mixed precision: after the plain FP32
                                         phase (see mixed_handover), FP32
                                         iterations continue on the
                                         correction around an FP64 base point
                                         that is re-centred at every restart
                                         (FP64 iterative refinement) until
                                         refine_tol; 0 = FP64 iterations
                                         instead (default 1) */
} mpx_pdlp_options;

/* Mirrors `mipx::LpResult` (include/mipx/lp_solver.h:11-16) plus the final
 * termination quantities (reference definitions: src/lp/pdlp.cpp:1105-1202). */
typedef struct mpx_pdlp_result {
    int32_t status;          /* MPX_STATUS_* */
    int32_t iterations;      /* total PDHG iterations incl. polish */
    double objective;        /* offset + obj'x on the recovered x (pdlp.cpp:1698-1703) */
    double work_units;       /* iterations * (4 nnz + 2 (m + n)) (pdlp.cpp:1716-1719) */
    double rel_primal;       /* last evaluated relative primal residual */
    double rel_dual;         /* last evaluated relative dual residual */
    double rel_gap;          /* last evaluated relative gap */
    double primal_obj;       /* last evaluated primal objective (original units) */
    double dual_obj;         /* last evaluated dual objective (original units) */
    double step;             /* final step: 0.998 / ||A_scaled||_2, moved by adaptive_step */
    double primal_weight;    /* final primal weight */
    int32_t restarts;
    int32_t iterations_fp32; /* MPX_PREC_MIXED: iterations run in FP32 (plain
                                FP32 phase and refinement) */
    double seconds;          /* wall time inside the library */
    double setup_seconds;    /* upload + scaling + power iteration */
    int32_t iterations_refine; /* MPX_PREC_MIXED: of iterations_fp32, those on
                                  the correction around the FP64 base */
    int32_t pad;
} mpx_pdlp_result;

void mpx_pdlp_default_options(mpx_pdlp_options* opts);
const char* mpx_last_error(void);
const char* mpx_version(void);
int mpx_device_count(int32_t* count);
/* Waits for all work on the device (hipDeviceSynchronize). */
int mpx_device_synchronize(int32_t device);
/* Copies `bytes` bytes device-to-device with a plain streaming kernel
 * (read `bytes`, write `bytes`); used to calibrate HBM counters.  Returns the
 * kernel time in milliseconds in *ms. */
int mpx_copy_bench(int64_t bytes, double* ms);

/*
 * Solve one LP on the GPU.  Replaces `PdlpSolver::solveGpu()`
 * (src/lp/pdlp.cpp:1739-2792): upload, on-device Ruiz / Pock-Chambolle /
 * bound-objective scaling (reference host code pdlp.cpp:469-669), on-device
 * power iteration (pdlp.cpp:1988-2041), the restarted Halpern PDHG loop in
 * hipGraph-captured blocks of `termination_eval_frequency` iterations, the
 * polish phase and recovery to original space (pdlp.cpp:2678-2712).
 *
 * warm_x[num_cols] / warm_y[num_rows] may be NULL (reference setWarmStart,
 * pdlp.cpp:457-467,1812-1827).  x_out[num_cols] and y_out[num_rows] receive
 * the recovered primal and dual (on Infeasible: the Farkas ray in y_out; on
 * Unbounded: the primal ray in x_out; pdlp.cpp:2318-2340); either may be NULL.
 * The trivial n == 0 and m == 0 cases are handled in closed form
 * (pdlp.cpp:812-860).  The non-default options behave as in the GPU host
 * loop: adaptive_step adapts at each evaluation from x^ and A x^
 * (pdlp.cpp:2271-2307), anderson_acceleration extrapolates at main-loop
 * restarts (293-437, 2418-2466; window at most 64), matrix_zero_tol drops
 * |a| < tol before scaling (499-518; certificates use the LP as given).
 * Returns MPX_ERR_INVALID for invalid arguments and MPX_ERR_OUT_OF_MEMORY
 * (res->status MPX_STATUS_ERROR) when device memory cannot be allocated, so
 * that a caller can fall back to a CPU solve as the reference does
 * (pdlp.cpp:1767-1771); otherwise MPX_OK, with any other device failure
 * reported as MPX_STATUS_ERROR in `res`.
 */
int mpx_pdlp_solve(const mpx_lp* lp, const mpx_pdlp_options* opts,
                   const double* warm_x, const double* warm_y,
                   double* x_out, double* y_out, mpx_pdlp_result* res);

/*
 * Batched solve of `count` LPs that share one constraint matrix, objective
 * and row bounds and differ only in column bounds -- branch-and-bound node
 * relaxations (the reference's caller: mip_solver.cpp node LPs via
 * setColBounds, pdlp.cpp:2890-2898).  col_lower/col_upper are [count][num_cols]
 * row-major (NULL = use base->col_lower / col_upper for every instance).
 * All instances iterate in the same launches (lane-per-instance kernels that
 * read A once for 64 instances) with per-instance restart, primal-weight and
 * termination logic on the device.  x_out is [count][num_cols], y_out
 * [count][num_rows] (either may be NULL); results[count].  The batched and
 * serving entry points return MPX_ERR_UNSUPPORTED for adaptive_step,
 * anderson_acceleration and matrix_zero_tol (single-LP options only).
 */
int mpx_pdlp_solve_batch_shared(const mpx_lp* base, int32_t count,
                                const double* col_lower, const double* col_upper,
                                const mpx_pdlp_options* opts,
                                double* x_out, double* y_out,
                                mpx_pdlp_result* results);

/*
 * mpx_pdlp_solve_batch_shared with a warm start per instance: warm_x is
 * [count][num_cols], warm_y [count][num_rows] in original space (both NULL:
 * cold), mapped into the scaled space as setWarmStart does
 * (pdlp.cpp:915-930).  Instance i then solves like mpx_pdlp_solve of the base
 * LP with its column bounds and warm start (pdlp.cpp:2890-2898 setColBounds
 * followed by setWarmStart, without rebuilding the scaling).
 */
int mpx_pdlp_solve_batch_shared_warm(const mpx_lp* base, int32_t count,
                                     const double* col_lower, const double* col_upper,
                                     const double* warm_x, const double* warm_y,
                                     const mpx_pdlp_options* opts,
                                     double* x_out, double* y_out,
                                     mpx_pdlp_result* results);

/*
 * Batched solve of `count` independent LPs of uneven sizes in one persistent
 * launch.  x_out[i] / y_out[i] are per-LP output arrays (or NULL).
 */
int mpx_pdlp_solve_batch(const mpx_lp* lps, int32_t count,
                         const mpx_pdlp_options* opts,
                         double* const* x_out, double* const* y_out,
                         mpx_pdlp_result* results);

/* ---- resident batches ----------------------------------------------------
 * The same batched solves with the problem data kept in HBM: create uploads
 * once (the arrays passed in are only read during create, except base/lps
 * objective vectors, which are copied), run solves on the device only (it
 * can be called again: it re-runs scaling, power iteration and PDHG from
 * the uploaded data), results copies solutions back.
 */
typedef struct mpx_batch mpx_batch;

typedef struct mpx_batch_stats {
    double upload_ms;          /* host -> device at create */
    double setup_ms;           /* device scaling + power iteration + init (last run) */
    double run_ms;             /* wall time of the last mpx_batch_run (incl. setup) */
    int32_t rounds;            /* rounds of N iterations in the last run */
    int32_t interleaved;       /* 1: lane-per-instance kernels (shared matrix) */
    int64_t instance_iterations; /* sum of PDHG iterations over instances */
    double latency_p50_ms;     /* per-instance completion time from run start */
    double latency_p99_ms;
    double latency_max_ms;
    /* mpx_batch_profile: averages of individually timed kernel launches
       (HIP events on the batch's stream) at the start of a fresh run */
    double primal_kernel_us;
    double dual_kernel_us;
    int32_t profiled_slots;    /* instances active in the profiled launches */
    int32_t profiled_launches;
    int32_t columns_per_wave;  /* lane-per-instance kernels: autotuned columns
                                  per wave (primal), 0 before the first run */
    int32_t rows_per_wave;     /* ... rows per wave (dual) */
} mpx_batch_stats;

int mpx_batch_create_shared(const mpx_lp* base, int32_t count, const double* col_lower,
                            const double* col_upper, const mpx_pdlp_options* opts,
                            mpx_batch** out);
int mpx_batch_create(const mpx_lp* lps, int32_t count, const mpx_pdlp_options* opts,
                     mpx_batch** out);
int mpx_batch_run(mpx_batch* b);
int mpx_batch_results(mpx_batch* b, double* const* x_out, double* const* y_out,
                      mpx_pdlp_result* results);
int mpx_batch_get_stats(mpx_batch* b, mpx_batch_stats* stats);
/* Re-initialises the batch and times `launches` launches of the two
 * iteration kernels one by one (fills the primal/dual kernel fields). */
int mpx_batch_profile(mpx_batch* b, int32_t launches, mpx_batch_stats* stats);
/* Mixed precision: the same for the refinement's correction kernels, with
 * every instance first moved into the refinement phase around its initial
 * point (the correction kernels at full occupancy). */
int mpx_batch_profile_refine(mpx_batch* b, int32_t launches, mpx_batch_stats* stats);
int mpx_batch_destroy(mpx_batch* b);

/* ---- serving loop (continuous batching) ---------------------------------
 * No single counterpart in the reference: its GPU path allocates, solves and
 * frees one LP per PdlpSolver::solveGpu() call (src/lp/pdlp.cpp:1767-1897,
 * 2751-2784), and concurrent solves are separate threads with separate
 * streams (src/mip/mip_solver.cpp:4499-4501).  A server instead keeps one
 * device context, a pool of device memory of `hbm_budget_bytes` (<= 0: half
 * of the free device memory) and `max_slots` instance slots (<= 0: 1024).
 *
 * Requests are admitted first-in first-out into free slots while the pool
 * has room; a fresh LP's scaling and power iteration run on a side stream
 * while the running instances keep iterating; each mpx_server_step() runs
 * one round (termination_eval_frequency iterations, then the per-instance
 * evaluation / restart / termination logic) of every running instance in
 * one launch sequence, retires the finished ones and recycles their slots
 * and memory.  Fresh LPs iterate in the packed kernels, bounds variants of a
 * base in the lane-per-instance kernels of that base.  Every instance's
 * arithmetic depends only on its own data, so a fresh request's result is
 * bit-identical to mpx_pdlp_solve_batch() of the same LP alone, and a bounds
 * variant's to mpx_pdlp_solve_batch_shared() of its base with those column
 * bounds alone (same options), whatever else is in flight.
 *
 * Per request, max_solve_seconds (if >= 0) counts from its activation.
 * stop_fn is polled once per mpx_server_step(): while it returns nonzero,
 * every instance running at that step finishes with MPX_STATUS_ITER_LIMIT.
 * Calls on one server must not overlap.
 */
typedef struct mpx_server mpx_server;

typedef struct mpx_server_stats {
    int64_t submitted;
    int64_t finished;
    int32_t queued;            /* waiting for a slot / memory */
    int32_t preparing;         /* scaling + power iteration in flight */
    int32_t running;           /* in slots, iterating */
    int32_t slots;             /* max_slots */
    int64_t rounds;            /* rounds launched */
    double pool_bytes;         /* device memory pool (the HBM budget) */
    double pool_used_bytes;    /* in use now (bases + requests in flight) */
    double pool_peak_bytes;
    double iterate_ms;         /* host wall time spent in rounds (launch + wait) */
    int64_t instance_iterations; /* PDHG iterations of finished requests */
    double admit_ms;           /* host wall time in admission (uploads, setup launch) */
    double activate_ms;        /* ... in activation (incl. waiting for a setup) */
    double retire_ms;          /* ... in retirement (recovery, copies out) */
    double max_step_ms;        /* longest single mpx_server_step */
    double setup_query_ms;     /* ... of it in polling setup completion */
    double setup_wait_ms;      /* ... of it waiting for a setup (nothing running) */
} mpx_server_stats;

int mpx_server_create(const mpx_pdlp_options* opts, double hbm_budget_bytes,
                      int32_t max_slots, mpx_server** out);
/* Uploads a base LP whose scaled matrix, objective, row bounds and power
 * iteration are shared by the bounds variants submitted against it (the
 * branch-and-bound node relaxations of one MIP).  Kept until destroy.  A
 * base has at most 4,194,303 rows and columns (MPX_ERR_UNSUPPORTED beyond). */
int mpx_server_add_base(mpx_server* srv, const mpx_lp* base, int32_t* base_id);
/* Copies the LP; returns a ticket >= 0 in *ticket. */
int mpx_server_submit(mpx_server* srv, const mpx_lp* lp, int64_t* ticket);
/* A variant of base `base_id` with its own column bounds (copied). */
int mpx_server_submit_bounds(mpx_server* srv, int32_t base_id, const double* col_lower,
                             const double* col_upper, int64_t* ticket);
/* The same with an original-space warm start warm_x[num_cols] /
 * warm_y[num_rows] (copied), mapped into the scaled space as the reference's
 * setWarmStart (pdlp.cpp:915-930): a child node started from its parent's
 * solution.  The result equals mpx_pdlp_solve_batch_shared_warm of the
 * request alone.  No counterpart in the reference, which solves node LPs one
 * at a time. */
int mpx_server_submit_bounds_warm(mpx_server* srv, int32_t base_id, const double* col_lower,
                                  const double* col_upper, const double* warm_x,
                                  const double* warm_y, int64_t* ticket);
/* Runs one scheduling round: activate prepared requests, admit queued ones,
 * iterate one round, retire finished ones.  Returns the number of requests
 * still queued, preparing or running in *active. */
int mpx_server_step(mpx_server* srv, int32_t* active);
/* Fetches a finished result once (x_out[n] / y_out[m] may be NULL); the
 * server then forgets the ticket.  *done = 0 if it is not finished yet.
 * res->seconds is the time from submit to completion, res->setup_seconds
 * the part before the request started iterating (queue + scaling). */
int mpx_server_result(mpx_server* srv, int64_t ticket, int32_t* done,
                      double* x_out, double* y_out, mpx_pdlp_result* res);
/* Tickets finished since the last poll, in completion order: up to
 * `capacity` of them into tickets[], their number in *count. */
int mpx_server_poll(mpx_server* srv, int64_t* tickets, int32_t capacity, int32_t* count);
int mpx_server_get_stats(mpx_server* srv, mpx_server_stats* stats);
int mpx_server_destroy(mpx_server* srv);

/* ---- device-pointer kernels (lower-level entry points) -------------------
 * These operate on device buffers the caller owns (e.g. from hipMalloc or a
 * framework allocator) on the caller's HIP stream (`stream` is a hipStream_t
 * passed as void*, NULL = default stream).  They are the building blocks of
 * mpx_pdlp_solve and are used by the parity tests.
 */

/* y = A x (CSR, FP64).  Reference: SparseMatrix::multiply
 * (src/lp/sparse_matrix.cpp:181-193) / cusparseSpMV (pdlp.cpp:1969-1973).
 * group = lanes per row (1, 2, 4, ..., 64; 0 = auto).  group == 1 sums each
 * row sequentially in storage order, as the reference's CPU loop does. */
int mpx_spmv_csr_f64(int32_t rows, int32_t cols, const int32_t* row_starts,
                     const int32_t* col_indices, const double* values,
                     const double* x, double* y, int32_t group, void* stream);
int mpx_spmv_csr_f32(int32_t rows, int32_t cols, const int32_t* row_starts,
                     const int32_t* col_indices, const float* values,
                     const float* x, float* y, int32_t group, void* stream);

/*
 * Scaled problem, as produced on the device by the reference's
 * buildScaledProblem + buildTransposeCSR (src/lp/pdlp.cpp:469-707) and the
 * power iteration (pdlp.cpp:748-795).  Host buffers, filled by
 * mpx_pdlp_scale_problem (any pointer may be NULL to skip it).
 */
typedef struct mpx_scaled_problem {
    double* values;        /* [nnz] scaled A, CSR order */
    double* at_values;     /* [nnz] scaled A' (CSC of A, rows increasing) */
    int32_t* at_row_starts; /* [num_cols + 1] */
    int32_t* at_col_indices; /* [nnz] */
    double* cscaled;       /* [n] */
    double* col_lower;     /* [n] scaled, +-inf kept */
    double* col_upper;     /* [n] */
    double* row_lower;     /* [m] */
    double* row_upper;     /* [m] */
    double* row_scale;     /* [m] */
    double* col_scale;     /* [n] */
    double* sigma_base;    /* [m] */
    double* tau_base;      /* [n] */
    double objective_scale;
    double constraint_scale;
    double spectral_norm;  /* sqrt(max(sigma^2, 1e-12)) from the power iteration */
    int32_t power_iterations;
} mpx_scaled_problem;

int mpx_pdlp_scale_problem(const mpx_lp* lp, const mpx_pdlp_options* opts,
                           mpx_scaled_problem* out);

/* The same outputs from the batched engines' device setup (multi-workgroup
 * passes over a cohort of matrices, pdlp.cpp:469-669 and 748-795), which the
 * batched solves and the serving loop use. */
int mpx_pdlp_scale_problem_batched(const mpx_lp* lp, const mpx_pdlp_options* opts,
                                   mpx_scaled_problem* out);

/*
 * Runs `iters` Halpern PDHG iterations (reference loop body,
 * src/lp/pdlp.cpp:980-1072; GPU kernels pdlp_kernels.cu:46-157) on the
 * device from the given scaled state, with inner-iteration offset
 * `inner_count` (lambda = (inner+k)/(inner+k+1)), fixed step and primal
 * weight.  State arrays are host buffers updated in place:
 * current_x, initial_x, pdhg_x [n]; current_y, initial_y, pdhg_y [m].
 * `sp` provides values/at_* structure, cscaled, bounds, sigma/tau bases.
 * precision selects FP64 or FP32 iterations.
 */
int mpx_pdhg_iterate(const mpx_lp* lp, const mpx_scaled_problem* sp,
                     int32_t iters, int32_t inner_count, double step,
                     double primal_weight, int32_t precision, int32_t group,
                     double* current_x, double* initial_x, double* pdhg_x,
                     double* current_y, double* initial_y, double* pdhg_y);

/*
 * Relative KKT residuals of an original-space pair (x, y) on the device, in
 * original units: the termination quantities of the reference's evaluation
 * (src/lp/pdlp.cpp:1105-1202, GPU loop 2196-2269; norms of the original
 * problem, 53-81), computed from the LP as given so that any returned
 * solution can be certified where it lies.  `lp` is a host struct whose
 * array pointers are DEVICE pointers (CSR); col_starts[num_cols + 1],
 * row_indices[nnz], col_values[nnz] are A in CSC (device).  x[num_cols],
 * y[num_rows] device.  out (device, 8 doubles) receives rel_primal (L2),
 * rel_dual (L2: the part of sgn (c + A'y) that no finite bound absorbs),
 * rel_gap, primal objective, dual objective, max bound violation of x, max
 * row violation (absolute), max dual-sign violation of y.  Enqueued on
 * `stream` (hipStream_t as void*, NULL = default); the per-block partial
 * sums live in a stream-ordered allocation.  Deterministic (fixed-order
 * reductions, one thread per row / column).
 */
int mpx_kkt_check_device(const mpx_lp* lp, const int32_t* col_starts,
                         const int32_t* row_indices, const double* col_values,
                         const double* x, const double* y, double* out, void* stream);

/*
 * mpx_kkt_check_device without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The partial sums live in a fixed 1 MiB device array per GPU, held until
 * the process ends, so the LP may have at most 26,214 blocks of 256 rows
 * plus columns (about 6.7 million rows and columns together); larger LPs are
 * refused with MPX_ERR_INVALID before any device work.  Same arguments and
 * outputs, bit-identical results.  Calls that overlap on one GPU share the
 * array and race: serialise them (one stream, or an event between calls).
 */
int mpx_kkt_check_device_noalloc(const mpx_lp* lp, const int32_t* col_starts,
                                 const int32_t* row_indices, const double* col_values,
                                 const double* x, const double* y, double* out, void* stream);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* MIPX_PDLP_H */
