/* MI355X LP serving loop: C-ABI.
 *
 * The reference serves POST /v1/solve with one synchronous CLI solve per
 * worker and answers 429 when every worker is busy
 * (python/niryukti/service.py:32-71).  This server queues instead and keeps
 * the GPU busy with continuous batching:
 *
 *  - LPs below `large_nonzeros` and `batch_max_size` rows + columns go to one
 *    batched engine (include/niryukti_batch.h):
 *    a scheduler thread submits them (their host prepare runs on the engine's
 *    prepare workers, off the launch path), steps the engine, and completes
 *    requests as their LPs are certified, while other LPs keep iterating;
 *  - larger LPs, and requests with options of their own, run on the grid-wide
 *    single-LP engine (niryukti_solve_lp) on `grid_workers` host threads, each
 *    on its own HIP stream, admitted against the rest of the HBM budget by the
 *    reference's footprint estimate (src/cuda.cu:1294-1295), doubled for the
 *    resident original model of the device checks.
 *
 * Both routes are deterministic per LP, so a result is the same alone and in
 * any mix.  There is no counterpart entry point in the reference.
 */
#ifndef NIRYUKTI_SERVE_H
#define NIRYUKTI_SERVE_H
#include <stdint.h>
#include "niryukti_engine.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct niryukti_server niryukti_server;

typedef struct {
    int64_t hbm_budget;     /* bytes for resident LPs, both routes together */
    double batch_fraction;  /* share of hbm_budget for the batched engine's arena (0, 1) */
    int64_t large_nonzeros; /* LPs with at least this many nonzeros take the grid route */
    int grid_workers;       /* concurrent grid-route solves, 1..16 */
    int checks_per_launch;  /* check periods per batched scheduling round */
    int64_t max_slots;      /* resident LPs of the batched engine */
    int64_t batch_max_size; /* LPs with more rows + columns than this take the grid route: above the
                             * batched engine's largest LDS class (8,000) an LP runs on one
                             * workgroup walking its CSR from memory, 2-80x slower per iteration
                             * than on the whole GPU */
} niryukti_server_limits;

typedef struct {
    int64_t submitted, finished, batch_routed, grid_routed, grid_running;
    int64_t grid_bytes_in_use, grid_budget_bytes;
    int64_t batch_rounds;
} niryukti_server_stats;

void niryukti_server_default_limits(niryukti_server_limits *l);
/* opts: the solve options of every request without options of its own (NULL = defaults). */
int niryukti_server_create(const niryukti_options *opts, const niryukti_server_limits *limits,
                           niryukti_server **out);
/* Finishes every submitted request, then frees the server. */
void niryukti_server_destroy(niryukti_server *s);
/* Queues one LP (arrays copied).  x0 / y0: optional original-space warm start
 * (cols / rows entries).  route: -1 by size, 0 batched, 1 grid.  opts: NULL, or
 * options for this request alone (the request then takes the grid route).
 * Invalid input (a nonfinite warm start, a malformed LP) is refused here
 * with -1 and leaves the server running.  *request receives the id. */
int niryukti_server_submit(niryukti_server *s, const niryukti_lp *lp, const double *x0, const double *y0, int route,
                           const niryukti_options *opts, int64_t *request);
/* Waits up to timeout_seconds (< 0: no limit) for a request.  Returns 1 with
 * *r filled (x / y copied when given and sized cols / rows) and *latency the
 * seconds from submit to completion, 0 on timeout, -1 when the request failed
 * (message in niryukti_server_last_error) or the id is unknown.  *route
 * (when given) receives 0 batched or 1 grid. */
int niryukti_server_wait(niryukti_server *s, int64_t request, double timeout_seconds, niryukti_result *r,
                         double *latency, int *route);
/* Forgets a finished request. */
int niryukti_server_release(niryukti_server *s, int64_t request);
int niryukti_server_stats_get(niryukti_server *s, niryukti_server_stats *out);
const char *niryukti_server_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
