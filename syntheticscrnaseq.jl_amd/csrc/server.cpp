// Serving loop (continuous batching of simulate() requests).
#include "runtime_internal.h"

using namespace scrna;

extern "C" {

int scrna_server_create(scrna_ctx* ctx, int32_t max_slots, scrna_server** out) {
    (void)ctx; (void)max_slots;
    if (out) *out = nullptr;
    return set_error(SCRNA_ERR_UNSUPPORTED, "serving is not available in this build");
}
int scrna_submit(scrna_server*, const scrna_request*, int64_t*) {
    return set_error(SCRNA_ERR_UNSUPPORTED, "serving is not available in this build");
}
int scrna_wait(scrna_server*, int64_t, int64_t, scrna_output*) {
    return set_error(SCRNA_ERR_UNSUPPORTED, "serving is not available in this build");
}
int scrna_cancel(scrna_server*, int64_t) {
    return set_error(SCRNA_ERR_UNSUPPORTED, "serving is not available in this build");
}
int scrna_server_stats(scrna_server*, scrna_latency*) {
    return set_error(SCRNA_ERR_UNSUPPORTED, "serving is not available in this build");
}
int scrna_server_destroy(scrna_server*) { return SCRNA_OK; }

}
