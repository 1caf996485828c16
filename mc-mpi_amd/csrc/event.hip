/*
 * event.hip -- event-based transport kernel (not yet implemented: the
 * launcher reports an error instead of falling back to another path).
 */
#include "internal.h"

int mc_launch_event(mcgpu_context *ctx, const mc_transport_args &a, const mcgpu_options &o,
                    hipStream_t s) {
  (void)a;
  (void)o;
  (void)s;
  return mc_fail(ctx, MCGPU_ERR_ARG, "event kernel not available in this build");
}
