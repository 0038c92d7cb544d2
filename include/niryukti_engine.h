/* MI355X PDHG LP engine: array-level C-ABI.
 *
 * niryukti_solve_lp() is the in-memory form of the reference's
 *   vantage::Result vantage::solve(const Model &, const Options &)   src/solver.cpp:55
 * for continuous LPs: the model arrives as CSR arrays instead of a Model
 * object, options and results as plain structs (same names, defaults and
 * meaning as include/vantage/vantage.hpp:40-123), and errors as a return code
 * plus niryukti_engine_last_error() instead of exceptions.  Statuses are the
 * reference's strings: OPTIMAL, INFEASIBLE, UNBOUNDED, ITERATION_LIMIT,
 * TIME_LIMIT, NUMERICAL_ERROR, INTERRUPTED, UNSUPPORTED.
 *
 * The niryukti_instance_* functions expose the seeded synthetic workload
 * generators (planted optimum known in closed form) used by the benchmark.
 */
#ifndef NIRYUKTI_ENGINE_H
#define NIRYUKTI_ENGINE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t rows, cols;
    const int64_t *ptr;   /* rows + 1 */
    const int64_t *index; /* nnz */
    const double *value;  /* nnz */
    const double *c, *lb, *ub; /* cols (canonical minimisation costs) */
    const double *rl, *ru;     /* rows */
    double offset, sense;      /* reported objective = sense * (c^T x + offset) */
} niryukti_lp;

typedef struct {
    const char *device;           /* "gpu" (default) | "hip" | "auto"; "cpu"/"cuda" -> UNSUPPORTED */
    const char *method;           /* "pdhg" | "halpern" | "rhpdhg" | "r2hpdhg" */
    const char *scaling;          /* "ruiz" | "combined" */
    const char *primal_weight;    /* "displacement" | "pid" */
    const char *matrix_precision; /* "fp64" | "mixed" (FP32 values) | "fp32" (FP32 SpMV arithmetic, FP64 refinement) */
    const char *gpu_indices;      /* "auto" | "32" | "64" */
    double tol, time_limit;
    int64_t iteration_limit;
    int check_every, scaling_passes, threads;
    int presolve, restart, adaptive, graphs, verbose;
    int device_checks;       /* -1 auto (>= 100,000 nonzeros), 0 host checks, 1 device checks */
    int device_prepare;      /* Ruiz scaling + transpose: -1 auto (>= 100,000 nonzeros, ruiz), 0 host, 1 device */
    int power_iterations;    /* 0..1000: power-iteration estimate of ||A||_2 for the first step (0 = off) */
    int polishing;           /* 1: feasibility polishing (src/solver.cpp:634-690), default 0 */
    const double *initial_x; /* NULL or cols */
    const double *initial_y; /* NULL or rows */
    /* Continuous checkpoints (src/solver.cpp:384-509, the reference's
     * --checkpoint-out / --resume / --checkpoint-nodes): NULL or a path. */
    const char *checkpoint_path, *resume_path;
    int64_t checkpoint_nodes; /* iterations between checkpoint writes, default 100 */
} niryukti_options;

typedef struct {
    char status[32];
    char message[256];
    double objective; /* sense-restored; NaN when not finite */
    double primal_residual, dual_residual, relative_gap, kkt_error, complementarity, dual_bound;
    double primal_absolute, dual_absolute;
    int finite;
    int64_t iterations, restarts, rejected_steps, weight_updates, host_candidate_checks, device_candidate_checks;
    int64_t removed_rows, removed_columns;
    double seconds, preprocess_seconds, transfer_seconds, iteration_seconds, verification_seconds;
    int index_bits;
    int device_prepare; /* 1 when scaling and transpose ran on the device */
    double operator_norm_estimate; /* power-iteration estimate (0 when not requested) */
    int64_t polishing_iterations, polishing_attempts;
    /* Decision trace (src/solver.cpp:299-333, 692-725): the first step and
     * primal weight, the iteration of the first restart and the primal weight
     * right after it (both 0 when no restart happened). */
    double initial_step, initial_weight, first_restart_weight;
    int64_t first_restart_iteration;
    int64_t fp32_iterations; /* matrix_precision "fp32": iterations before the FP64 refinement phase */
    double *x; /* NULL or cols: receives the primal solution */
    double *y; /* NULL or rows: receives the dual solution */
} niryukti_result;

void niryukti_default_options(niryukti_options *o);
/* 0 on success (any solver status), -1 on invalid input or a device error. */
int niryukti_solve_lp(const niryukti_lp *lp, const niryukti_options *o, niryukti_result *r);
const char *niryukti_engine_last_error(void);

/* The host prepare step on its own (src/preprocess.cpp:133-346: presolve,
 * scaling, objective scale), for tools and tests.  Output arrays are sized
 * by the ORIGINAL model: ptr rows+1, index/value nnz, c/lb/ub/column_scale/
 * col_map cols, rl/ru/row_scale/row_map rows.  Returns 0 (out_rows/out_cols
 * set), 1 when presolve proved the LP infeasible (message in
 * niryukti_engine_last_error), -1 on error. */
int niryukti_prepare_lp(const niryukti_lp *lp, const niryukti_options *o, int64_t *out_rows, int64_t *out_cols,
                        int64_t *ptr, int64_t *index, double *value, double *c, double *lb, double *ub, double *rl,
                        double *ru, double *column_scale, double *row_scale, int64_t *col_map, int64_t *row_map,
                        double *objective_scale, double *offset);

/* The original-space accuracy of a point (x: cols, y: rows) as the certifying
 * check computes it: vantage::verify (src/verify.cpp:241-244, Verifier::evaluate
 * at :103-240).  With safe_bound != 0 the gap includes the outward-rounded
 * Lagrangian bound (src/verify.cpp:12-88), as for a certificate.  0 or -1. */
typedef struct {
    double objective, primal, dual, gap, kkt, complementarity, lower_bound, primal_absolute, dual_absolute;
    int finite;
} niryukti_accuracy;
int niryukti_verify_lp(const niryukti_lp *lp, const double *x, const double *y, int safe_bound,
                       niryukti_accuracy *out);

/* Power-iteration estimate of ||A||_2 of the LP's matrix (src/research.cpp:6-50:
 * same seed, start vector and normalisation order as the reference); 0 when
 * iterations < 1, the matrix is empty or the input is invalid. */
double niryukti_power_norm(const niryukti_lp *lp, int iterations);

/* Models read by the reference's readers (native .json or .mps, src/io.cpp:84-140,
 * 299-585): exactly one of `path` and `json_text` is given.  NULL on error. */
typedef struct niryukti_model niryukti_model;
niryukti_model *niryukti_model_load(const char *path, const char *json_text);
void niryukti_model_destroy(niryukti_model *m);
/* Borrowed views valid until destroy. */
void niryukti_model_view(const niryukti_model *m, niryukti_lp *lp, const char **name, const char **fingerprint);
/* Writes the model as .mps or native .json by the path's extension, in the
 * reference's `convert` layout (src/io.cpp:765-798, 815-901).  0 or -1. */
int niryukti_model_write(const niryukti_model *m, const char *path);
/* Result document with the reference's keys (src/io.cpp:902-995) for a result
 * of this model; the string is valid until the next call on this thread. */
const char *niryukti_result_json(const niryukti_model *m, const niryukti_result *r, const char *backend);

/* Seeded synthetic instances.  config 0 (2,000 x 5,000 equality LP), 1
 * (1,000,000 x 2,000,000 power-law LP), 4 (block-angular); `scale` divides the
 * dimensions.  config 2 selects instance `index` of the batched workload,
 * config 3 request `index` of the serving stream (nonzeros log-uniform in
 * [1e3, 5e6], batch generator otherwise). */
typedef struct niryukti_instance niryukti_instance;
niryukti_instance *niryukti_instance_create(int config, uint64_t seed, int64_t scale, int64_t index);
void niryukti_instance_destroy(niryukti_instance *p);
/* Borrowed views valid until destroy; planted_x / planted_y are the certified optimum. */
void niryukti_instance_view(const niryukti_instance *p, niryukti_lp *lp, double *planted_objective,
                            const double **planted_x, const double **planted_y);

#ifdef __cplusplus
}
#endif
#endif
