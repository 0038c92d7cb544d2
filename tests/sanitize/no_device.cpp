// Test double for the HIP side of libniryukti, for the host-code sanitizer
// builds (tests/sanitize/Makefile): every device entry point reports that no
// device is visible, so the host paths that need no GPU (model I/O, prepare,
// verifier, session protocol, checkpoint files, the serving loop's threads)
// run under AddressSanitizer, UndefinedBehaviorSanitizer or ThreadSanitizer.
// Never linked into the product.
#include <cstdint>

#include "niryukti_batch.h"
#include "niryukti_pdhg.h"
#include "niryukti_prepare.h"

namespace {
const char *const kNoDevice = "No HIP device is visible (sanitizer build)";
}

extern "C" {
const char *npdhg_last_error(void) { return kNoDevice; }
int npdhg_device_count(void) { return 0; }
int npdhg_device_name(int, char *name, int len) {
    if (len > 0) name[0] = 0;
    return -1;
}
int npdhg_create(const npdhg_lp *, const double *, const double *, const npdhg_config *, npdhg_backend **) { return -1; }
int npdhg_create_transposed(const npdhg_lp *, const int64_t *, const int64_t *, const double *, const double *,
                            const double *, const npdhg_config *, npdhg_backend **) {
    return -1;
}
void npdhg_destroy(npdhg_backend *) {}
int npdhg_advance(npdhg_backend *, int, double, double, int *) { return -1; }
int npdhg_status_get(npdhg_backend *, npdhg_status *) { return -1; }
int npdhg_candidates(npdhg_backend *, double *, double *, double *, double *) { return -1; }
int npdhg_reset(npdhg_backend *, const double *, const double *) { return -1; }
int npdhg_reset_candidate(npdhg_backend *, int) { return -1; }
int npdhg_set_fp32_phase(npdhg_backend *, int) { return -1; }
int npdhg_snapshot_size(npdhg_backend *, int64_t *) { return -1; }
int npdhg_snapshot(npdhg_backend *, double *) { return -1; }
int npdhg_restore(npdhg_backend *, const double *) { return -1; }
int npdhg_attach_original(npdhg_backend *, const npdhg_original *) { return -1; }
int npdhg_evaluate(npdhg_backend *, npdhg_accuracy *, npdhg_accuracy *) { return -1; }
int npdhg_keep_best(npdhg_backend *, int) { return -1; }
int npdhg_original_point(npdhg_backend *, int, double *, double *) { return -1; }
int npdhg_farkas_displacement(npdhg_backend *, int, int *, double *) { return -1; }
int npdhg_epoch_displacement(npdhg_backend *, int, double *, double *) { return -1; }
int npdhg_set_epoch(npdhg_backend *, int) { return -1; }
int npdhg_spmv(npdhg_backend *, int, const double *, double *) { return -1; }
int npdhg_scale_lp(const npdhg_lp *, int, npdhg_scaled *) { return -1; }
const char *npdhg_scale_last_error(void) { return kNoDevice; }
int niryukti_batch_create(const niryukti_options *, int64_t, int64_t, niryukti_batch **) { return -1; }
void niryukti_batch_destroy(niryukti_batch *) {}
int niryukti_batch_submit(niryukti_batch *, const niryukti_lp *, const double *, const double *, int64_t *) {
    return -1;
}
int niryukti_batch_step(niryukti_batch *, int, int64_t *) { return -1; }
int niryukti_batch_admit(niryukti_batch *, int64_t *) { return -1; }
int niryukti_batch_run(niryukti_batch *, int) { return -1; }
int niryukti_batch_result(niryukti_batch *, int64_t, niryukti_result *) { return -1; }
int niryukti_batch_stats_get(niryukti_batch *, niryukti_batch_stats *) { return -1; }
int niryukti_batch_set_handoff(niryukti_batch *, int64_t, int) { return -1; }
int niryukti_batch_finished(niryukti_batch *, int64_t *, int64_t, int64_t *) { return -1; }
int niryukti_batch_release(niryukti_batch *, int64_t) { return -1; }
const char *niryukti_batch_last_error(void) { return kNoDevice; }
}
