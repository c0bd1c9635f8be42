/*
 * C-ABI of the MI355X LP presolve/postsolve library (libgpu_presolver_amd.so).
 *
 * The reference exposes only a C++ API (cpp/include/gpu_presolver/presolve/
 * gpu_presolve.hpp:34-40, gpu_postsolve.hpp:16-27, rules/rule_NAME.hpp,
 * gpu_presolve_kernels.hpp:7-16).  Its std::vector members make that API a
 * C++ ABI, so this header is the flat, plain-pointer form of the same entry
 * points that an FFI (ctypes, cgo, JNI) binds.  The C++ API in
 * include/gpu_presolver/presolve/ is implemented on top of the same code.
 *
 * Conventions
 *  - Every array argument of a gp_* compute call is HIP device memory unless
 *    the comment says "host".
 *  - Calls are synchronous with respect to the host (like the reference,
 *    which returns after cudaDeviceSynchronize) and run on the library's
 *    internal stream.
 *  - Errors: a nonzero gp_status; gp_last_error() gives the message of the
 *    last failure on the calling thread.  LP outcomes (infeasible, unbounded)
 *    are not errors; they are flags in the results, as in the reference.
 */
#ifndef GP_PRESOLVE_CAPI_H
#define GP_PRESOLVE_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gp_status {
  GP_OK = 0,
  GP_ERR_INVALID_ARGUMENT = 1,
  GP_ERR_DEVICE = 2,      /* a HIP runtime call failed */
  GP_ERR_INTERNAL = 3,    /* std::exception inside the library */
  GP_ERR_NO_DEVICE = 4,   /* no usable gfx950 device */
  GP_ERR_UNSUPPORTED = 5, /* a structure-specific rule would act on the LP: those rules are not
                             implemented, so the presolve is refused (disable the rule or
                             enable_structure_specific_rules to presolve without it) */
} gp_status;

/* int32 zero-based CSR (reference DeviceCsrMatrix, presolve_structs.hpp:137-144). */
typedef struct gp_csr {
  int32_t rows;
  int32_t cols;
  int32_t nnz;
  int32_t* row_ptr;
  int32_t* col_idx;
  double* val;
} gp_csr;

/* The LP (reference LPInfoGpu, presolve_structs.hpp:146-155). */
typedef struct gp_lp {
  gp_csr A;
  gp_csr AT;
  double* c;
  double* AL;
  double* AU;
  double* l;
  double* u;
  double obj_constant;
} gp_lp;

/* PresolveParams (reference presolve_structs.hpp:21-134), every field in
 * declaration order, booleans as int32.  gp_params_default() fills the
 * reference defaults. */
typedef struct gp_params {
  int32_t max_iters;
  double max_time;
  int32_t verbose;
  int32_t record_postsolve_tape;
  int32_t record_postsolve_tape_cpu;
  double feasibility_tol;
  double bound_tol;
  double zero_tol;
  int32_t implied_variable_bounds_max_rounds;
  int32_t implied_variable_bounds_max_bound_only_rounds;
  int32_t implied_variable_bounds_preserve_free_zero_cost;
  int32_t doubleton_equations_single_batch_per_iter;
  int32_t doubleton_equations_max_fill_in_proxy;
  int32_t doubleton_equations_scan;
  int32_t doubleton_equations_min_selected_per_batch;
  double doubleton_equations_min_selected_ratio;
  int32_t doubleton_equations_max_batch_rounds;
  double doubleton_equations_max_time;
  int32_t enable_infeasible_fixed_variables;
  int32_t enable_empty_rows;
  int32_t enable_singleton_rows;
  int32_t enable_infeasible_redundant_rows;
  int32_t enable_implied_variable_bounds;
  int32_t enable_duplicate_rows;
  int32_t enable_empty_cols;
  int32_t enable_column_singletons_eq;
  int32_t enable_column_singletons_dual_infer;
  int32_t enable_doubleton_equations;
  int32_t enable_dual_fix;
  int32_t dual_fix_max_nonzero_cost_col_degree;
  int32_t enable_duplicate_columns;
  double duplicate_columns_max_abs_ratio;
  int32_t duplicate_columns_scale_guard_max_group_size;
  int32_t duplicate_columns_scale_guard_min_groups;
  int32_t enable_structure_specific_rules;
  int32_t enable_zero_cost_redundant_box_bounds;
  int32_t enable_redundant_bounds;
  int32_t use_tiered_scheduler;
  int32_t enable_tiered_bootstrap;
  int64_t implied_variable_bounds_bound_only_nnz_round_budget;
  /* The structure-specific rules' switches and the thresholds their
   * screens read (presolve_structs.hpp:75-115, in that order); appended after
   * the fields above. */
  int32_t enable_bounded_two_row_projection;
  int32_t bounded_two_row_projection_max_row_nnz;
  double bounded_two_row_projection_safety_tol;
  int32_t bounded_two_row_projection_min_candidates;
  double bounded_two_row_projection_min_candidate_ratio;
  int32_t enable_orphan_mccormick_projection;
  int32_t orphan_mccormick_projection_min_candidates;
  double orphan_mccormick_projection_min_candidate_ratio;
  int32_t enable_linf_components;
  int32_t linf_probe_pairs;
  int32_t linf_min_pairs;
  int32_t linf_max_tail_cols;
  int32_t enable_structural_l1_substitution;
  int32_t enable_antipodal_components;
  int32_t antipodal_probe_pairs;
  int32_t antipodal_min_pairs;
  int32_t enable_covering_cost_dominance;
  int32_t covering_cost_dominance_min_fixed_cols;
  double covering_cost_dominance_min_candidate_ratio;
} gp_params;

/* Summary scalars of one presolve (reference GpuPresolveSummary, gpu_presolve.hpp:9-29). */
typedef struct gp_summary {
  int32_t original_rows;
  int32_t original_cols;
  int32_t reduced_rows;
  int32_t reduced_cols;
  int32_t reduced_nnz;
  int32_t iterations;
  double elapsed_seconds;
  double obj_constant_delta;
  int32_t has_infeasible;
  int32_t has_unbounded;
} gp_summary;

/* Flat read-only view of a PresolveRecordGpu (reference presolve_structs.hpp:
 * 255-319), all pointers host memory.  A view returned by the library stays
 * valid until its owner is freed. */
typedef struct gp_record_view {
  int32_t m0, n0, m1, n1;
  const int32_t* row_red2org;   /* m1 */
  const int32_t* col_red2org;   /* n1 */
  int32_t fixed_count;
  const int32_t* fixed_idx;
  const double* fixed_val;
  double obj_constant_old;
  double obj_constant_new;
  int32_t tape_count;           /* number of records */
  const int32_t* tape_types;    /* tape_count */
  const int32_t* tape_index_starts; /* tape_count + 1 */
  const int32_t* tape_value_starts; /* tape_count + 1 */
  const uint8_t* tape_dual_modes;   /* tape_count */
  int32_t tape_index_count;
  const int32_t* tape_indices;
  int32_t tape_value_count;
  const double* tape_vals;
} gp_record_view;

/* Reduction plan of one phase (reference PresolvePlanGpu, presolve_structs.hpp:
 * 329-359) in plain form, for calling one rule directly as the reference's
 * unit tests do.  Any vector a rule does not read may be null. */
typedef struct gp_plan_view {
  uint8_t* keep_row;
  uint8_t* keep_col;
  double* c;
  double* l;
  double* u;
  double* AL;
  double* AU;
  double obj_constant_delta;
  int32_t has_change;
  int32_t has_row_action;
  int32_t has_col_action;
  int32_t has_infeasible;
  int32_t has_unbounded;
  int32_t has_new_A;   /* out: set when the rule rebuilt A (doubleton) */
  gp_csr new_A;        /* out: device CSR owned by the caller (gp_device_free) */
} gp_plan_view;

/* PresolveStatsGpu (reference presolve_structs.hpp:158-171). */
typedef struct gp_stats_view {
  int32_t* row_nnz;
  uint8_t* singleton_row_mask;
  int32_t* singleton_row_col;
  double* singleton_row_val;
  int32_t* col_nnz;
  uint8_t* empty_col_mask;
  uint8_t* column_singleton_mask;
  int32_t* column_singleton_row;
  double* column_singleton_val;
} gp_stats_view;

/* Rule identifiers for gp_rule_apply (one per reference rules/rule_*.hpp). */
typedef enum gp_rule_id {
  GP_RULE_EMPTY_ROWS = 0,
  GP_RULE_SINGLETON_ROWS = 1,
  GP_RULE_INFEASIBLE_REDUNDANT_ROWS = 2,
  GP_RULE_IMPLIED_VARIABLE_BOUNDS = 3,
  GP_RULE_DUPLICATE_ROWS = 4,
  GP_RULE_INFEASIBLE_FIXED_VARIABLES = 5,
  GP_RULE_EMPTY_COLS = 6,
  GP_RULE_COLUMN_SINGLETONS_DUAL_INFER = 7,
  GP_RULE_COLUMN_SINGLETONS_EQ = 8,
  GP_RULE_DOUBLETON_EQUATIONS = 9,
  GP_RULE_DUAL_FIX = 10,
  GP_RULE_DUPLICATE_COLUMNS = 11,
  GP_RULE_REDUNDANT_BOUNDS = 12,
} gp_rule_id;

typedef struct gp_presolve_handle gp_presolve_handle;
typedef struct gp_tape_handle gp_tape_handle;

/* ---- library ---------------------------------------------------------- */
const char* gp_last_error(void);
const char* gp_version(void);
void gp_params_default(gp_params* params);
/* hipFree of a pointer the library handed to the caller. */
gp_status gp_device_free(void* device_ptr);

/* ---- primitives (reference gpu_presolve_kernels.hpp:7-16) -------------- */
/* Replaces compute_row_activity_summary (reference gpu_presolve_kernels.hpp:7-14).
 * `stream` (a hipStream_t) may be NULL: the call then runs on the library's
 * stream and returns when the results are written.  With a stream the kernel
 * is only enqueued on it (no synchronization), for callers that order their
 * own work, time it with events on that stream, or capture it in a graph. */
gp_status gp_row_activity(const gp_csr* A, const double* l, const double* u, double zero_tol,
                          double* row_min_fin, double* row_max_fin,
                          int32_t* row_min_neg_inf_count, int32_t* row_max_pos_inf_count,
                          void* stream);
/* No counterpart in the reference.  gp_row_activity without any allocation,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc): the staging of the (l, u)
 * pairs uses a fixed 32 MiB device array of the library (one per GPU, held
 * until the process ends) instead of a pool block.  A->cols must be at most
 * GP_ROW_ACTIVITY_NOALLOC_MAX_COLS, else GP_ERR_INVALID_ARGUMENT.  `stream` NULL: the null stream, and the call
 * returns when the results are written; otherwise the kernels are only
 * enqueued.  Calls that overlap on one GPU share the array and race; the
 * results are bit-identical to gp_row_activity's. */
gp_status gp_row_activity_noalloc(const gp_csr* A, const double* l, const double* u, double zero_tol,
                                  double* row_min_fin, double* row_max_fin, int32_t* row_min_neg_inf_count,
                                  int32_t* row_max_pos_inf_count, void* stream);
#define GP_ROW_ACTIVITY_NOALLOC_MAX_COLS 2097152
/* Replaces compute_col_max_abs (reference gpu_presolve_kernels.hpp:16); `stream` as above. */
gp_status gp_col_max_abs(const gp_csr* AT, double* col_max_abs, void* stream);
/* Mask compaction of a CSR matrix and its re-transpose, the apply step of the
 * reference (gpu_presolve.cu:1881-1978 and :1804-1879).  keep masks are
 * device uint8; outputs are allocated by the library (gp_device_free). */
gp_status gp_compact_csr(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col,
                         gp_csr* out);
gp_status gp_transpose_csr(const gp_csr* A, gp_csr* out);

/* ---- rule operator API (reference rules/rule_*.hpp) -------------------- */
/* Runs one rule on `plan` exactly as apply_rule_<name> does.  `lp` and
 * `stats` are device views; `candidate_col_mask` is used by the dual-fix
 * rule only (may be null).  Tape records produced by the rule are appended
 * to *tape (created on first use; free with gp_tape_free). */
gp_status gp_rule_apply(int32_t rule_id, gp_plan_view* plan, const gp_lp* lp,
                        const gp_stats_view* stats, const gp_params* params,
                        const uint8_t* candidate_col_mask, gp_tape_handle** tape);
gp_status gp_tape_view(const gp_tape_handle* tape, gp_record_view* out);
void gp_tape_free(gp_tape_handle* tape);

/* ---- presolve (reference gpu_presolve.hpp:34-40) ----------------------- */
/* run_gpu_presolve_with_reduced_lp (keep_reduced_lp=1) or
 * run_gpu_presolve_with_record (keep_reduced_lp=0).  `lp` is device memory
 * owned by the caller and must outlive presolve and postsolve. */
gp_status gp_presolve(const gp_lp* lp, const gp_params* params, int32_t keep_reduced_lp,
                      gp_presolve_handle** out);
gp_status gp_presolve_summary(const gp_presolve_handle* h, gp_summary* out);
gp_status gp_presolve_record(const gp_presolve_handle* h, gp_record_view* out);
/* Device view of the reduced LP, owned by the handle (valid until gp_presolve_free). */
gp_status gp_presolve_reduced_lp(const gp_presolve_handle* h, gp_lp* out);
/* free_gpu_presolve_reduced_lp + record release (reference gpu_presolve.hpp:40). */
void gp_presolve_free(gp_presolve_handle* h);

/* Replaces _has_good_nnz_progress (reference gpu_presolve.hpp:32): 1 when
 * nnz_after < ratio * nnz_before (and nnz_before > 0), else 0.  Host only. */
int32_t gp_has_good_nnz_progress(int32_t nnz_before, int32_t nnz_after, double ratio);

/* Working-LP nnz from which each fast phase opens with the batched singleton
 * closure (column-singleton EQ, dual fix and singleton rows repeated on one
 * plan, then one compaction).  The reference fixes it at 10,000,000
 * (kMinNnzForBatchedClosure, gpu_presolve.cu:2506), which is the default;
 * lowering it exercises the closure on small LPs.  Process-wide, host only. */
void gp_set_batched_closure_min_nnz(int32_t nnz);
int32_t gp_batched_closure_min_nnz(void);

/* Testing hook (no counterpart in the reference): the calling thread's
 * count-th presolve record update from now fails with an internal error right
 * after it has queued the record's device-to-host copies; 0 disarms it.  Used
 * to check that a failed call leaves the thread's next call unaffected. */
void gp_test_fail_after_record_updates(int32_t count);
/* Testing hook (no counterpart in the reference): a phase whose predecessor
 * changed nothing reuses that phase's plan and statistics instead of building
 * them again.  With on != 0, the calling thread's presolves compare every
 * reused pair with a freshly built one, bit for bit, and fail with
 * GP_ERR_INTERNAL on any difference. */
void gp_test_check_plan_reuse(int32_t on);
/* Number of phases on the calling thread that reused their predecessor's plan. */
int64_t gp_test_plan_reuses(void);

/* ---- postsolve (reference gpu_postsolve.hpp:16-27) --------------------- */
/* postsolve_gpu with caller-provided outputs: x_org, z_org hold n0 doubles,
 * y_org m0 doubles (device).  x_red/y_red/z_red may be null.  `original`
 * (device view, may be null) enables the dual recoveries that need A. */
gp_status gp_postsolve(const gp_record_view* record, const double* x_red, const double* y_red,
                       const double* z_red, const gp_lp* original, double* x_org, double* y_org,
                       double* z_org);

/* ---- prepared postsolve (no counterpart in the reference) ---------------
 * gp_postsolve_prepare builds what a postsolve of one record needs apart from
 * the reduced point (device tape and maps, both replay schedules, the
 * protected-column mask); gp_postsolve_plan_run then postsolves any reduced
 * point of that record exactly as gp_postsolve does, without rebuilding them.
 * With use_graph != 0 and all three of x_red, y_red, z_red given, the run's
 * launches are captured into a hipGraph on first use and replayed afterwards
 * (the inputs and outputs pass through buffers the plan owns).  The record
 * is copied; the original LP must stay valid while the plan exists.  A plan
 * runs one postsolve at a time. */
typedef struct gp_postsolve_plan gp_postsolve_plan;
gp_status gp_postsolve_prepare(const gp_record_view* record, const gp_lp* original, int32_t use_graph,
                               gp_postsolve_plan** out);
gp_status gp_postsolve_plan_run(gp_postsolve_plan* plan, const double* x_red, const double* y_red,
                                const double* z_red, double* x_org, double* y_org, double* z_org);
/* Number of runs so far that replayed the captured graph. */
int64_t gp_postsolve_plan_graph_runs(const gp_postsolve_plan* plan);
/* Shape of a prepared postsolve: tape records, dependency levels of the
 * primal replay (over all its ranges) and of the dual replay, and the bytes
 * of the tape on the device. */
typedef struct gp_postsolve_shape {
  int64_t records;
  int32_t primal_levels;
  int32_t dual_levels;
  int64_t tape_bytes;
} gp_postsolve_shape;
gp_status gp_postsolve_plan_shape(const gp_postsolve_plan* plan, gp_postsolve_shape* out);
void gp_postsolve_plan_free(gp_postsolve_plan* plan);

/* ---- serving and batching (no counterpart in the reference) ------------
 * The reference presolves one LP per call on the legacy stream.  The serving
 * runtime keeps several jobs in flight, each on its own stream and execution
 * context, so their kernels overlap on the device.  A job presolves a device
 * LP the caller owns (it must stay valid until the job's result is
 * returned); with postsolve on, it then draws the seeded reduced-space point
 * gp_synthetic_point(seed ^ tag, reduced LP) in place of an LP solver,
 * postsolves it against the original LP and reports gp_checksum of the
 * postsolved x, y, z.  Jobs start in submission order when a worker is free
 * and 4x the LP's device bytes fit hbm_budget_bytes next to the jobs in
 * flight (a job alone is always admitted).  Every kernel is deterministic,
 * so a job's result is bit-identical alone and in any mix. */
typedef struct gp_serve_config {
  int32_t workers;         /* jobs in flight, >= 1 */
  int32_t postsolve;       /* 1: presolve + seeded point + postsolve; 0: presolve only */
  double hbm_budget_bytes; /* admission budget; <= 0: unlimited */
  uint64_t seed;           /* job seeds are seed ^ tag */
} gp_serve_config;

typedef struct gp_job_result {
  int64_t job; /* submission index */
  uint64_t tag;
  int32_t status; /* gp_status of the job */
  int32_t worker;
  int32_t reduced_rows, reduced_cols, reduced_nnz, iterations;
  int32_t has_infeasible, has_unbounded;
  uint64_t checksum;                 /* 0 without postsolve */
  double t_arrival, t_start, t_end;  /* seconds since the server was created */
  double t_ready; /* the job's LP resident: t_start, or the end of its generation (gp_lp_spec jobs) */
} gp_job_result;

/* A seeded synthetic LP, generated on the device (no counterpart in the
 * reference): the same spec gives the same LP bit for bit.  Dyadic values
 * (coefficients +-U[0.5,2] rounded to 1/8; c ~ N(0,1), u ~ U[1,10], the
 * feasible point x0 and the row-bound offsets U[0,1], rounded to 2^-10);
 * l = 0, u = +inf on 20% of the columns (only where c >= 0); 1% empty, 2%
 * fixed and 5% singleton columns; 5% singleton rows, 3% doubleton
 * equalities, 2% scaled duplicates of general rows; 30% equality and 10%
 * one-sided rows.  General rows have U[len_lo, len_hi] entries, or
 * max(min_len, Poisson(mean_len)) when mean_len > 0. */
typedef struct gp_lp_spec {
  uint64_t seed;
  int32_t m, n;
  int32_t len_lo, len_hi;
  double mean_len;
  int32_t min_len;
} gp_lp_spec;
/* Generates the LP into device memory owned by the library (A, AT rows
 * sorted, c, AL, AU, l, u); release it with gp_free_lp. */
gp_status gp_generate_lp(const gp_lp_spec* spec, gp_lp* out);
gp_status gp_free_lp(gp_lp* lp);

typedef struct gp_server gp_server;
gp_status gp_server_create(const gp_params* params, const gp_serve_config* config, gp_server** out);
/* Queues a job (t_arrival < 0: now, on the server's clock); *job gets its index. */
gp_status gp_server_submit(gp_server* server, const gp_lp* lp, uint64_t tag, double t_arrival, int64_t* job);
/* Moves up to max finished results to out, waiting up to timeout_s for the
 * first one; returns how many (host only). */
int32_t gp_server_wait(gp_server* server, gp_job_result* out, int32_t max, double timeout_s);
double gp_server_now(const gp_server* server);
/* Finishes every queued job, then stops the workers. */
void gp_server_destroy(gp_server* server);

/* Queues a job whose LP the worker first generates from `spec` on its own
 * stream (released after the job); t_ready marks the end of the generation. */
gp_status gp_server_submit_spec(gp_server* server, const gp_lp_spec* spec, uint64_t tag, double t_arrival,
                                int64_t* job);
/* The arrival trace form of gp_server_submit_spec: job k generates
 * specs[k], arrives at arrival_s[k] (NULL: all at once) and has tag k. */
gp_status gp_serve_spec_trace(const gp_lp_spec* specs, int32_t jobs, const double* arrival_s, const gp_params* params,
                              const gp_serve_config* config, gp_job_result* results);
/* A whole arrival trace on a fresh server: job k runs lps[lp_of_job[k]] and
 * arrives arrival_s[k] seconds after the start (non-decreasing; NULL: all at
 * once, the batched mode); its tag is k and results[k] receives its result. */
gp_status gp_serve_trace(const gp_lp* lps, int32_t lp_count, const int32_t* lp_of_job, const double* arrival_s,
                         int32_t jobs, const gp_params* params, const gp_serve_config* config,
                         gp_job_result* results);

/* One job exactly as a serving worker runs it (seed ^ tag for the reduced
 * point), but through the plain synchronous entry points on the calling
 * thread: the one-at-a-time baseline of the batched and serving modes, and
 * the reference result a served job is compared with.  out->worker is -1;
 * times are seconds on the host's steady clock around the call. */
gp_status gp_plain_job(const gp_lp* lp, const gp_params* params, uint64_t seed, uint64_t tag, int32_t postsolve,
                       gp_job_result* out);
/* gp_plain_job on the LP of `spec`, generated first on the calling thread. */
gp_status gp_plain_job_spec(const gp_lp_spec* spec, const gp_params* params, uint64_t seed, uint64_t tag,
                            int32_t postsolve, gp_job_result* out);

/* ---- MPS input (reference cpp/mps_reader, tools/gpu_presolver_tool_common.hpp)
 * Host-only reading of a continuous MPS model (free format, then the fixed
 * column layout), with the reference reader's decisions (number parsing,
 * duplicate summing, default bounds, RANGES, first RHS/RANGES/BOUNDS set).
 * gp_mps_view gives host pointers valid until gp_mps_free; gp_mps_upload
 * copies the LP to the device (A from the CSR, AT from the CSC) in arrays
 * released by gp_free_lp. */
typedef struct gp_mps_model gp_mps_model;
typedef struct gp_mps_view {
  const char* name;
  int32_t rows, cols, nnz;
  const int32_t* csr_row_ptr;
  const int32_t* csr_col_idx;
  const double* csr_val;
  const int32_t* csc_col_ptr;
  const int32_t* csc_row_idx;
  const double* csc_val;
  const double* c;
  const double* row_lower;
  const double* row_upper;
  const double* col_lower;
  const double* col_upper;
  double obj_constant;
} gp_mps_view;
gp_status gp_mps_read(const char* path, gp_mps_model** out);
gp_status gp_mps_parse(const char* text, size_t length, gp_mps_model** out);
gp_status gp_mps_get_view(const gp_mps_model* model, gp_mps_view* out);
gp_status gp_mps_upload(const gp_mps_model* model, gp_lp* out);
void gp_mps_free(gp_mps_model* model);

/* The seeded reduced-space point of a serving job: x in [l, u] on the 2^-10
 * grid, y and z uniform in [-1, 1); lp is the reduced LP (device). */
gp_status gp_synthetic_point(uint64_t seed, const gp_lp* lp, double* x, double* y, double* z);
/* Order-free checksum of the bit patterns of x (n), y (m) and z (n). */
gp_status gp_checksum(const double* x, int32_t n, const double* y, int32_t m, const double* z, uint64_t* out);

/* ---- solver hand-off (the external GPU solver's place in the reference's
 * pipeline, README.md:145-169) ------------------------------------------
 * gp_pdhg_solve solves a device LP (normally gp_presolve_reduced_lp's view,
 * so the reduced LP never leaves the GPU) with restarted PDHG (averaged or
 * reflected Halpern, by `method`): Ruiz
 * scaling, step sizes from a power-iteration estimate of ||A||_2, restarts
 * by KKT error and primal-weight updates, every vector in device memory, one
 * wait per check period.  x, z: device arrays of lp->A.cols doubles, y of
 * lp->A.rows, written with the returned iterate; z holds c - A^T y projected
 * on the reduced-cost cone of the column bounds (so c - A^T y - z is the dual
 * residual).  Tolerances are relative: ||row violation|| / (1 + ||b||),
 * ||dual residual|| / (1 + ||c||), |primal - dual objective| /
 * (1 + |primal| + |dual|), b the larger finite |AL|, |AU| of each row.
 * Deterministic: the same LP and parameters give the same bits. */
typedef struct gp_pdhg_params {
  int32_t max_iters;   /* default 200000 */
  double max_time;     /* seconds, default 3600 */
  double tol;          /* relative KKT tolerance, default 1e-4 */
  int32_t check_every; /* iterations between KKT checks, default 64 */
  int32_t ruiz_iters;  /* default 10 */
  int32_t power_iters; /* default 40 */
  int32_t use_graph;   /* 1: each check period is one graph launch (default) */
  int32_t verbose;     /* 1: one line per check on stderr */
  int32_t method;      /* 0: restarted averaged PDHG (default); 1: restarted reflected Halpern PDHG,
                          z <- (k+1)/(k+2) (2 T(z) - z) + 1/(k+2) z_restart, the iteration of the
                          Halpern Peaceman-Rachford family */
} gp_pdhg_params;
typedef struct gp_pdhg_result {
  int32_t status; /* 0 optimal within tol, 1 iteration limit, 2 time limit */
  int32_t iterations;
  int32_t restarts;
  int32_t checks;
  double primal_obj; /* c^T x + obj_constant */
  double dual_obj;
  double primal_res; /* relative, as above */
  double dual_res;
  double gap;
  double elapsed_seconds;
  double norm_estimate; /* ||A||_2 of the scaled A */
  double primal_weight;
  int64_t graph_launches;
} gp_pdhg_result;
void gp_pdhg_params_default(gp_pdhg_params* params);
gp_status gp_pdhg_solve(const gp_lp* lp, const gp_pdhg_params* params, double* x, double* y, double* z,
                        gp_pdhg_result* out);

/* Presolve, solve and postsolve of one LP in one call: gp_presolve, then
 * gp_pdhg_solve on the device-resident reduced LP while a second thread and
 * stream prepare the postsolve (gp_postsolve_prepare's work, which needs only
 * the record), then the prepared postsolve of the solution against `lp`.
 * x, z: device arrays of lp->A.cols doubles, y of lp->A.rows; they receive
 * exactly what those calls one after another give.  When presolve finds the
 * LP infeasible or unbounded nothing is solved: solve.status is -1 and x, y,
 * z are zero. */
typedef struct gp_pipeline_result {
  int32_t reduced_rows;
  int32_t reduced_cols;
  int32_t reduced_nnz;
  int32_t presolve_iterations;
  double obj_constant_delta;
  int32_t has_infeasible;
  int32_t has_unbounded;
  gp_pdhg_result solve;
  double presolve_seconds;
  double solve_seconds;
  double prepare_seconds;  /* postsolve preparation, overlapped with the solve */
  double postsolve_seconds;
  double total_seconds;
} gp_pipeline_result;
gp_status gp_presolve_solve(const gp_lp* lp, const gp_params* params, const gp_pdhg_params* solve_params,
                            double* x, double* y, double* z, gp_pipeline_result* out);

#ifdef __cplusplus
}
#endif

#endif /* GP_PRESOLVE_CAPI_H */
