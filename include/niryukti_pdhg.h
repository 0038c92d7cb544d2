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
} npdhg_config;

/* Status of the last advance. */
typedef struct {
    int64_t accepted_total, rejected_total;
    int restart_requested;
    int index_bits;
    double factor; /* current adaptive step factor */
} npdhg_status;

const char *npdhg_last_error(void);
int npdhg_device_count(void);
int npdhg_device_name(int device, char *name, int len);

/* Uploads the LP and the start point; x0 has cols entries, y0 rows entries.
 * Replaces cuda_backend(...) (src/cuda.cu:1288-1303). */
int npdhg_create(const npdhg_lp *lp, const double *x0, const double *y0, const npdhg_config *cfg,
                 npdhg_backend **out);
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

/* y = A x (or y = A^T x when transpose != 0) on the device with the backend's
 * SpMV kernels; host pointers in and out.  Used by tests of the SpMV alone. */
int npdhg_spmv(npdhg_backend *b, int transpose, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
