/* MI355X PDHG iteration backend: C-ABI seam between the C++ driver and the
 * hand-written gfx950 HIP kernels.
 *
 * It replaces the reference's internal IterationBackend seam
 *   cuda_backend(model, x0, y0, adaptive, halpern, reflection, residual_restarts,
 *                graphs, indices, precision)          src/internal.hpp:129-134,
 *                                                      src/cuda.cu:1288-1303
 * and the virtual verbs of IterationBackend (src/internal.hpp:102-124) with
 * plain-pointer entry points: no C++ types or exceptions cross it.  Every call
 * returns 0 on success and a negative code on failure; npdhg_last_error()
 * then holds a message for the calling thread.
 *
 * The LP handed to npdhg_create is the presolved, scaled LP (the output of the
 * driver's prepare step): minimise c^T x s.t. rl <= A x <= ru, lb <= x <= ub.
 */
#ifndef NIRYUKTI_PDHG_H
#define NIRYUKTI_PDHG_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct npdhg_backend npdhg_backend;

typedef struct {
    int64_t rows, cols;
    const int64_t *ptr;   /* rows + 1 CSR offsets */
    const int64_t *index; /* nnz column indices */
    const double *value;  /* nnz values */
    const double *c, *lb, *ub; /* cols */
    const double *rl, *ru;     /* rows */
} npdhg_lp;

typedef struct {
    int adaptive;          /* adaptive step-size backtracking (src/cpu.cpp:151-163) */
    int halpern;           /* anchored Halpern iteration (src/cpu.cpp:174-202) */
    double reflection;     /* 0 = Halpern, 1 = reflected Halpern (r2hpdhg) */
    int residual_restarts; /* fixed-point residual restart test */
    int graphs;            /* replay chunks as hipGraphs */
    int index_bits;        /* 0 = auto, 32 or 64 (device CSR offsets) */
    int fp32_values;       /* 1 = FP32 matrix values, FP64 accumulation and iterates */
    int fp32_arithmetic;   /* 1 = FP32 values, gathered vectors and per-lane sums in both products
                              (iterates, steps and reductions FP64); FP64 values kept for the
                              refinement phase (npdhg_set_fp32_phase) */
} npdhg_config;

/* Status of the last advance. */
typedef struct {
    int64_t accepted_total, rejected_total;
    int restart_requested;
    int index_bits;
    double factor; /* current adaptive step factor */
    int fp32_phase; /* 1 while the attempts run in FP32 arithmetic */
    int pad;
} npdhg_status;

const char *npdhg_last_error(void);
int npdhg_device_count(void);
int npdhg_device_name(int device, char *name, int len);

/* Uploads the LP and the start point; x0 has cols entries, y0 rows entries.
 * Replaces cuda_backend(...) (src/cuda.cu:1288-1303). */
int npdhg_create(const npdhg_lp *lp, const double *x0, const double *y0, const npdhg_config *cfg,
                 npdhg_backend **out);
/* The same, with the explicit transpose of lp's matrix supplied by the caller
 * (cols + 1 offsets, nnz row indices and values in row order per column, as
 * Sparse::transpose builds it, src/model.cpp:39-58), e.g. from
 * npdhg_scale_lp (include/niryukti_prepare.h), instead of building it here. */
int npdhg_create_transposed(const npdhg_lp *lp, const int64_t *t_ptr, const int64_t *t_index, const double *t_value,
                            const double *x0, const double *y0, const npdhg_config *cfg, npdhg_backend **out);
void npdhg_destroy(npdhg_backend *b);

/* Performs `count` accepted PDHG steps with base steps tau = primal_step * factor,
 * sigma = dual_step * factor, rejecting and retrying unsafe trials on the device.
 * Returns the number of accepted steps in *accepted (fewer than count only when
 * a Halpern residual restart is requested).  Replaces IterationBackend::advance
 * (src/cpu.cpp:96-222 semantics: count counts accepted steps). */
int npdhg_advance(npdhg_backend *b, int count, double primal_step, double dual_step, int *accepted);
int npdhg_status_get(npdhg_backend *b, npdhg_status *s);

/* Copies out the current iterate and the running average (or, in Halpern mode,
 * the last T(z)), then rebases the activity A x on the device.
 * Replaces IterationBackend::candidates (src/cpu.cpp:223-231). */
int npdhg_candidates(npdhg_backend *b, double *x, double *y, double *xa, double *ya);
/* Restart from host vectors.  Replaces IterationBackend::reset (src/cpu.cpp:232-247). */
int npdhg_reset(npdhg_backend *b, const double *x, const double *y);
/* Restart from the resident current (average = 0) or averaged (average = 1)
 * candidate without an upload.  Replaces IterationBackend::reset_candidate
 * (src/cuda.cu:892-930). */
int npdhg_reset_candidate(npdhg_backend *b, int average);

/* Moves a backend created with fp32_arithmetic between its FP32 arithmetic
 * phase (on = 1, where it starts) and FP64 arithmetic (on = 0: the FP64
 * refinement phase); the iterates, averages, step factor and counters carry
 * on.  There is no counterpart in the reference, whose GPU path runs one
 * precision for a whole solve (src/cuda.cu:142-160). */
int npdhg_set_fp32_phase(npdhg_backend *b, int on);

/* Resumable backend state for checkpoints: the current iterate, averages,
 * activity, Halpern anchors, step factor, average mass, residual history and
 * counters, plus (once npdhg_attach_original has run) the epoch point and the
 * best restored candidate; npdhg_snapshot_size gives its length in doubles.
 * npdhg_restore puts a snapshot of the same backend configuration back, after
 * which advance continues exactly as from the snapshot's moment.  Replace
 * IterationBackend::snapshot / restore (src/internal.hpp:106-107,
 * src/cpu.cpp:49-89, src/cuda.cu:739-770). */
int npdhg_snapshot_size(npdhg_backend *b, int64_t *doubles);
int npdhg_snapshot(npdhg_backend *b, double *state);
int npdhg_restore(npdhg_backend *b, const double *state);

/* ---- Device-side candidate checks ----
 *
 * The reference checks both candidates on the host every chunk: restore to the
 * original space, Verifier::evaluate twice, Farkas rays, epoch displacement
 * (src/solver.cpp:550-633, 692-716; src/cuda.cu:867-876 copies four vectors
 * down for it).  These entry points run the same formulas on the device
 * against a resident copy of the original model, so a check moves a few
 * scalars instead of four vectors; the host certifies a candidate with its own
 * FP64 verifier and outward-rounded bound before reporting OPTIMAL. */

/* The original model and the maps of the prepare step (src/internal.hpp:16-25):
 * restore_x puts x[j] * column_scale[j] at cols_map[j] over `fixed`;
 * restore_y puts y[i] * row_scale[i] * objective_scale at upper_rows[i] when
 * y[i] >= 0, else at lower_rows[i]. */
typedef struct {
    int64_t rows, cols;                 /* original m0, n0 */
    const int64_t *ptr, *index;         /* original A, CSR */
    const double *value;
    const double *c, *lb, *ub, *rl, *ru; /* canonical (minimisation) data */
    double offset;
    const int64_t *cols_map;            /* n entries */
    const int64_t *lower_rows, *upper_rows; /* m entries */
    const double *fixed;                /* n0 entries */
    const double *column_scale, *row_scale; /* n, m entries */
    double objective_scale;
} npdhg_original;

/* Verifier::evaluate quantities without the safe bound (src/verify.cpp:103-240),
 * plus whether the candidate's dual could be a Farkas ray: 0 means a rigorous
 * upper bound of the ray's Lagrangian value (src/verify.cpp:12-88) is <= 0, so
 * the host test (src/solver.cpp:599-633) would reject it. */
typedef struct {
    double objective, primal, dual, gap, kkt;
    int finite, farkas_possible;
} npdhg_accuracy;

/* Uploads the original model; the epoch point becomes the current iterate.
 * Fails (and device checks stay off) when the row maps are not one-to-one. */
int npdhg_attach_original(npdhg_backend *b, const npdhg_original *orig);
/* Evaluates the current (out[0]) and averaged (out[1]) candidates in the
 * original space, keeps both restored points on the device and rebases the
 * activity A x as npdhg_candidates does. */
int npdhg_evaluate(npdhg_backend *b, npdhg_accuracy *current, npdhg_accuracy *average);
/* Best-so-far tracking: copies the restored candidate of the last evaluate. */
int npdhg_keep_best(npdhg_backend *b, int average);
/* Downloads a restored candidate of the last evaluate (which = 0 current,
 * 1 average) or the best (which = 2); x0 has n0 entries, y0 m0 (either NULL). */
int npdhg_original_point(npdhg_backend *b, int which, double *x0, double *y0);
/* Farkas pre-check of the restored epoch displacement of a candidate's dual
 * (average = 0 current, 1 averaged); with *possible != 0 and ray != NULL, the
 * restored displacement (m0 entries) is copied out for the host's exact test. */
int npdhg_farkas_displacement(npdhg_backend *b, int average, int *possible, double *ray);
/* Squared epoch displacement of a candidate in the scaled space, summed in
 * double-double (src/solver.cpp:702-716 sums in long double). */
int npdhg_epoch_displacement(npdhg_backend *b, int average, double *dx2, double *dy2);
/* Epoch := the current (average = 0) or averaged candidate. */
int npdhg_set_epoch(npdhg_backend *b, int average);

/* y = A x (or y = A^T x when transpose != 0) on the device with the backend's
 * SpMV kernels; host pointers in and out.  Used by tests of the SpMV alone. */
int npdhg_spmv(npdhg_backend *b, int transpose, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
