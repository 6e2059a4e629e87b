// pdlp_batch.hip -- batched PDLP (many LPs in one persistent launch) and the
// serving loop.  Work in progress: the entry points are exported and report
// MPX_ERR_UNSUPPORTED until the persistent kernel lands.
#include <hip/hip_runtime.h>

#include "../../include/mipx_pdlp.h"
#include "device_common.h"

extern "C" {

int mpx_pdlp_solve_batch_shared(const mpx_lp*, int32_t, const double*, const double*,
                                const mpx_pdlp_options*, double*, double*, mpx_pdlp_result*) {
    mpx::setError("mpx_pdlp_solve_batch_shared: not implemented yet");
    return MPX_ERR_UNSUPPORTED;
}

int mpx_pdlp_solve_batch(const mpx_lp*, int32_t, const mpx_pdlp_options*, double* const*,
                         double* const*, mpx_pdlp_result*) {
    mpx::setError("mpx_pdlp_solve_batch: not implemented yet");
    return MPX_ERR_UNSUPPORTED;
}

int mpx_server_create(const mpx_pdlp_options*, double, int32_t, mpx_server** out) {
    if (out) *out = nullptr;
    mpx::setError("mpx_server_create: not implemented yet");
    return MPX_ERR_UNSUPPORTED;
}
int mpx_server_submit(mpx_server*, const mpx_lp*, int64_t*) { return MPX_ERR_UNSUPPORTED; }
int mpx_server_step(mpx_server*, int32_t*) { return MPX_ERR_UNSUPPORTED; }
int mpx_server_result(mpx_server*, int64_t, int32_t*, double*, double*, mpx_pdlp_result*) {
    return MPX_ERR_UNSUPPORTED;
}
int mpx_server_destroy(mpx_server*) { return MPX_OK; }

}  // extern "C"
