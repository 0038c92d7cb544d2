/* MI355X batched / served PDHG engine: C-ABI.
 *
 * Many independent LPs share one resident device arena and one kernel launch
 * at a time: every admitted LP owns one workgroup that runs the complete
 * restarted-PDHG driver of the reference (src/solver.cpp:511-726: chunked
 * advance, original-space candidate checks of the current and averaged
 * iterates, best-candidate tracking, restarts and primal-weight updates) on
 * the device.  The host prepares each LP (presolve and scaling as the
 * reference's prepare(), src/preprocess.cpp:133), admits it into a free slot
 * when the arena has room (continuous batching: admission happens between
 * launches while other LPs keep iterating), and certifies every claimed
 * optimum with the reference's FP64 verifier and outward-rounded bound
 * (src/verify.cpp:12-240) before it is reported OPTIMAL.
 *
 * There is no counterpart entry point in the reference: its serving path
 * (python/niryukti/service.py) runs one CLI process per request.  Results are
 * reported in the same niryukti_result struct as niryukti_solve_lp().
 */
#ifndef NIRYUKTI_BATCH_H
#define NIRYUKTI_BATCH_H
#include <stdint.h>
#include "niryukti_engine.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct niryukti_batch niryukti_batch;

typedef struct {
    int64_t submitted, queued, active, finished;
    int64_t arena_bytes, arena_used_bytes;
    int64_t launches, host_certifications, host_farkas_checks;
    double device_seconds;  /* host wall time spent waiting on kernel launches */
    double advance_seconds; /* HIP-event time of the PDHG step kernel launches */
    double check_seconds;   /* HIP-event time of the candidate-check kernel launches */
    double advance_bytes;   /* algorithmic HBM bytes of all PDHG attempts run by the step kernel */
    int64_t attempts;       /* PDHG attempts (accepted + rejected) run by the step kernel */
    /* Per launch group: the classes of LPs with m + n <= 2048 / 4096 / 8000
     * whose y, dx and CSR offsets live in LDS, the global-memory class, and
     * the tail group (LDS classes with at most 256 LPs left, merged into one
     * 1024-thread launch): step-kernel event time, algorithmic bytes and
     * number of step-kernel launches.  Groups run on separate streams. */
    double class_seconds[5], class_bytes[5];
    int64_t class_launches[5];
    int64_t handoffs;         /* LPs moved to the grid-wide engine so far */
    int64_t handoffs_running; /* of those, still running there */
    /* Launch shape of each crowded LDS class's step kernel (threads per
     * workgroup, rows per thread): the autotuner's current choice, and 1 once
     * its trials are over.  Shapes never change results. */
    int32_t class_threads[3], class_rows[3], class_tuned[3], pad;
    /* Per step-kernel instantiation of the crowded LDS classes (256 x 1,
     * 256 x 2, 256 x 4, 1024 x 1 threads x rows), whichever class launched it:
     * HIP-event time, algorithmic bytes and launches (what a kernel trace
     * reports per kernel symbol). */
    double shape_seconds[4], shape_bytes[4];
    int64_t shape_launches[4];
    /* Host time of the scheduling thread outside the launches: waiting for
     * the previous round's certification, downloading and sorting the
     * stopped LPs of a round, admitting queued LPs into slots. */
    double certify_wait_seconds, settle_seconds, admit_seconds;
} niryukti_batch_stats;

/* opts: solver options shared by every LP of the batch (NULL = defaults).
 * arena_bytes: HBM budget for resident LPs.  max_slots: resident LP cap. */
int niryukti_batch_create(const niryukti_options *opts, int64_t arena_bytes, int64_t max_slots,
                          niryukti_batch **out);
void niryukti_batch_destroy(niryukti_batch *b);
/* Prepares and queues one LP (arrays are copied); optional warm start x0/y0
 * in original space (NULL for a cold start).  Returns its ticket.  May be
 * called from any thread, also while another thread is inside
 * niryukti_batch_step / _run; every other call on one batch must come from
 * one thread at a time. */
int niryukti_batch_submit(niryukti_batch *b, const niryukti_lp *lp, const double *x0, const double *y0,
                          int64_t *ticket);
/* One scheduling round: wait for the round launched by the previous call
 * and read back its LPs' states, admit queued LPs into free slots, launch the
 * next round -- `checks_per_launch` check periods per resident LP (LPs of a
 * class with at most 256 LPs left run in one fused launch that, when it has
 * the device to itself and no LP waits for admission, continues past that
 * number until about 32 ms have passed, so LPs of different speed do not wait
 * for each other) -- and return while it runs.  LPs that stopped in the
 * finished round are certified before the call returns when they are few (at
 * most 32); a larger set is certified on a host thread beside the new round
 * and is retired (or resumes) at the end of the next call.  Results do not
 * depend on how periods fall into rounds.  *running receives the number still
 * queued, resident or being certified, plus one while a round is in flight. */
int niryukti_batch_step(niryukti_batch *b, int checks_per_launch, int64_t *running);
/* Admits queued LPs into free slots (uploading them to the arena) without
 * launching; returns the number admitted in *admitted. */
int niryukti_batch_admit(niryukti_batch *b, int64_t *admitted);
/* Runs rounds until every submitted LP has finished. */
int niryukti_batch_run(niryukti_batch *b, int checks_per_launch);
/* 1 if the ticket finished (result filled), 0 if still running, <0 on error. */
int niryukti_batch_result(niryukti_batch *b, int64_t ticket, niryukti_result *r);
int niryukti_batch_stats_get(niryukti_batch *b, niryukti_batch_stats *s);
/* LPs still running after `iterations` iterations (a multiple of check_every;
 * 0, the default, disables it) leave the batch and continue on the grid-wide
 * single-LP engine, warm started from their current iterate, on
 * `workers` host threads (default 4).  The move depends only on the LP's own
 * iteration count, so results stay the same alone and in any mix. */
int niryukti_batch_set_handoff(niryukti_batch *b, int64_t iterations, int workers);
/* Up to `capacity` tickets finished since the last call, in finishing order
 * (each reported once); *count receives how many were written. */
int niryukti_batch_finished(niryukti_batch *b, int64_t *tickets, int64_t capacity, int64_t *count);
/* Frees the host copy of a finished LP once its result has been read; the
 * ticket is invalid afterwards.  Long-running servers call it per request. */
int niryukti_batch_release(niryukti_batch *b, int64_t ticket);
/* niryukti_batch_finished + niryukti_batch_result (+ niryukti_batch_release
 * when `release` is 1) for up to `capacity` finished tickets in one call:
 * tickets[k] and results[k] (every field but x and y, which are not copied;
 * their pointers are overwritten with NULL).  *count receives how many. */
int niryukti_batch_collect(niryukti_batch *b, int64_t *tickets, niryukti_result *results, int64_t capacity,
                           int release, int64_t *count);
const char *niryukti_batch_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
