/* CPU oracle for the PDHG LP path: a C-ABI over oracle/pdhg_oracle.cpp.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP engine is compared
 * against; only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  The product never links or calls it.
 *
 * All arrays are caller-owned.  Models are LPs in the reference's canonical
 * form: minimise sense*(c^T x + offset) subject to rl <= A x <= ru, lb <= x <= ub,
 * with A in CSR (int64 offsets and column indices, FP64 values).
 */
#ifndef NIRYUKTI_ORACLE_H
#define NIRYUKTI_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t rows, cols;
    const int64_t *ptr;   /* rows + 1 */
    const int64_t *index; /* nnz */
    const double *value;  /* nnz */
    const double *c, *lb, *ub; /* cols */
    const double *rl, *ru;     /* rows */
    double offset;
} oracle_lp;

typedef struct {
    double tol, time_limit;
    int64_t iteration_limit;
    int check_every, scaling_passes, threads;
    int presolve, restart, adaptive;
    const char *method;        /* "pdhg" | "halpern" | "rhpdhg" | "r2hpdhg" */
    const char *scaling;       /* "ruiz" | "combined" */
    const char *primal_weight; /* "displacement" | "pid" */
    const double *initial_x;   /* NULL or cols */
    const double *initial_y;   /* NULL or rows */
} oracle_options;

typedef struct {
    double objective, primal, dual, gap, kkt, lower_bound, complementarity, primal_absolute,
        dual_absolute;
    int finite;
} oracle_accuracy;

typedef struct {
    char status[32];
    oracle_accuracy accuracy;
    int64_t iterations, restarts, rejected_steps, weight_updates, host_candidate_checks,
        removed_rows, removed_columns;
    double seconds, iteration_seconds, preprocess_seconds, verification_seconds;
    /* decision trace: first step and primal weight, iteration of the first
     * restart and the primal weight right after it (0 when none) */
    double initial_step, initial_weight, first_restart_weight;
    int64_t first_restart_iteration;
    double *x; /* NULL or cols: receives the returned primal point */
    double *y; /* NULL or rows: receives the returned dual point */
} oracle_result;

void oracle_default_options(oracle_options *o);
/* Returns 0 on success, -1 on error (message in err). */
int oracle_solve(const oracle_lp *lp, const oracle_options *o, oracle_result *r, char *err, int errlen);
int oracle_verify(const oracle_lp *lp, const double *x, const double *y, int safe_bound,
                  oracle_accuracy *a, char *err, int errlen);
double oracle_infeasibility_bound(const oracle_lp *lp, const double *ray);
/* Presolve + scaling.  Buffers sized by the ORIGINAL cols/rows. */
int oracle_prepare(const oracle_lp *lp, const oracle_options *o, int64_t *out_rows, int64_t *out_cols,
                   double *column_scale, double *row_scale, double *objective_scale, int64_t *col_map,
                   int64_t *row_map, char *failure, int failure_len);

/* The prepared (presolved, scaled) model itself.  Buffers sized by the
 * ORIGINAL dimensions: ptr rows+1, index/value nnz, c/lb/ub cols, rl/ru rows. */
int oracle_prepared_model(const oracle_lp *lp, const oracle_options *o, int64_t *out_rows, int64_t *out_cols,
                          int64_t *ptr, int64_t *index, double *value, double *c, double *lb, double *ub,
                          double *rl, double *ru, double *offset, char *failure, int failure_len);

/* Raw iteration backend on an (already prepared) LP. */
void *oracle_backend_create(const oracle_lp *lp, const double *x0, const double *y0, int adaptive,
                            int halpern, double reflection, int residual_restarts, char *err, int errlen);
int oracle_backend_advance(void *b, int count, double tau, double sigma, char *err, int errlen);
void oracle_backend_candidates(void *b, double *x, double *y, double *xa, double *ya);
void oracle_backend_reset(void *b, const double *x, const double *y);
int64_t oracle_backend_rejected(void *b);
int oracle_backend_restart_requested(void *b);
void oracle_backend_destroy(void *b);

/* The PID primal-weight controller (src/research.cpp:52-70) on its own:
 * update() feeds one epoch displacement (dx = ||x - x_epoch||^2,
 * dy = ||y - y_epoch||^2) and returns the next weight. */
void *oracle_pid_create(void);
double oracle_pid_update(void *pid, double weight, double dx, double dy);
void oracle_pid_destroy(void *pid);

#ifdef __cplusplus
}
#endif
#endif
