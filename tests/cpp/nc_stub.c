/*
 * Stand-in for the C-ABI of libncorr_hip.so (include/ncorr_hip.h), used only
 * by the sanitizer build of the host libraries in the CPU suite, where no GPU
 * runtime is loaded.  Every count fails with the status in NC_STUB_STATUS
 * (default NC_ERR_NOMEM), so the drop-in class's error path can be driven:
 * an allocation failure must reach the caller as std::runtime_error, as the
 * reference's gpuErrchk does (gpuerrchk.h:9-14).
 */
#include <stdlib.h>

#include "ncorr_hip.h"

struct nc_context {
    int device;
};
static struct nc_context stub_ctx;

int nc_create(nc_context **ctx, int device) {
    if (!ctx) return NC_ERR_ARG;
    stub_ctx.device = device;
    *ctx = &stub_ctx;
    return NC_OK;
}
int nc_destroy(nc_context *ctx) { return ctx ? NC_OK : NC_ERR_ARG; }
const char *nc_last_error(const nc_context *ctx) { return ctx ? "stub: device allocation failed" : "null context"; }
int nc_count_host(nc_context *ctx, const nc_params *params, const double *p1, int64_t n1, const double *p2, int64_t n2,
                  const double *p3, int64_t n3, int path, double *hist, uint64_t *counts) {
    (void)params, (void)p1, (void)n1, (void)p2, (void)n2, (void)p3, (void)n3, (void)path, (void)hist, (void)counts;
    if (!ctx) return NC_ERR_ARG;
    const char *s = getenv("NC_STUB_STATUS");
    return s ? atoi(s) : NC_ERR_NOMEM;
}
int nc_last_timing(const nc_context *ctx, nc_timing *out) {
    (void)ctx, (void)out;
    return NC_ERR_ARG;
}
