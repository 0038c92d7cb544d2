/* MI355X device-side scaling and transpose of the prepare step: C-ABI.
 *
 * The reference prepares every LP on the host (src/preprocess.cpp:133-346):
 * presolve, then `scaling_passes` Ruiz passes (src/preprocess.cpp:306-324:
 * row factors 1/sqrt(max |a_ij|) clipped to [1e-3, 1e3], applied by
 * apply_scaling, src/preprocess.cpp:12-65, which skips any row or column whose
 * factor would erase a nonzero or make a finite value infinite), then the
 * objective scale max |c| (src/preprocess.cpp:336-345); the GPU backend then
 * builds the explicit transpose (Sparse::transpose, src/model.cpp:39-58).
 *
 * npdhg_scale_lp runs the Ruiz passes, the objective scale and the transpose
 * on the device for an LP that presolve has already reduced (the presolve
 * rules stay on the host: they use long-double guards).  Results are
 * bit-identical to the host passes: maxima do not depend on order, and the
 * factors, the safety tests and every product are the same IEEE FP64
 * operations.  The transpose keeps each column's entries in row order, as
 * Sparse::transpose does.  Only the "ruiz" scaling is offered here.
 *
 * Limits: nonzeros below 2^31, rows and columns below 2^31.  Returns 0, or
 * -1 with a message in npdhg_scale_last_error().
 */
#ifndef NIRYUKTI_PREPARE_H
#define NIRYUKTI_PREPARE_H
#include <stdint.h>
#include "niryukti_pdhg.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double *value;            /* nnz: scaled A values (the input's ptr/index order) */
    double *c, *lb, *ub;      /* cols */
    double *rl, *ru;          /* rows */
    double *row_scale;        /* rows: product of the applied row factors */
    double *column_scale;     /* cols: product of the applied column factors */
    double objective_scale;   /* max |c| after the passes (1 when c = 0 or passes = 0) */
    int64_t *t_ptr;           /* cols + 1: offsets of A^T */
    int64_t *t_index;         /* nnz: row index of each A^T entry */
    double *t_value;          /* nnz: A^T values (scaled) */
    double device_seconds;    /* HIP-event time of the kernels */
    double seconds;           /* wall time including the transfers */
} npdhg_scaled;

/* lp: the presolved, unscaled LP on the host.  passes: Ruiz passes (the
 * reference's scaling_passes, 5 by default; 0 = no scaling, objective_scale 1,
 * transpose only).  Every output array of `out` must be allocated by the caller. */
int npdhg_scale_lp(const npdhg_lp *lp, int passes, npdhg_scaled *out);
const char *npdhg_scale_last_error(void);

/* npdhg_scale_lp for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It has no counterpart in the reference.  Every array of `lp` and of `out`
 * is a DEVICE pointer (lp->rows / lp->cols and out->objective_scale are host
 * values); outputs are written in place.  `stream` is a hipStream_t (NULL for
 * the null stream); the call synchronises it before returning.  Same kernels
 * and bit-identical results as npdhg_scale_lp; device_seconds and seconds are
 * 0.  The work arrays live in one static device array of 128 MiB per device
 * (held until the process ends): an LP needing more than
 * 24 bytes x nonzeros + 8 bytes x columns (+ a few KiB) is refused with an
 * error.  It calls no hipMalloc or hipFree.  Every call on a device shares
 * that device's array, so calls that overlap on one device race. */
int npdhg_scale_lp_noalloc(const npdhg_lp *lp, int passes, npdhg_scaled *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
