// Batched trace and fit-graph entry points (filled in by the batched kernels).
#include "iact_abi_internal.h"

using namespace iact;

extern "C" {

size_t iact_batch_workspace_size(const iact_scene*, const iact_batch*, const iact_launch*) {
  return 0;
}
int32_t iact_trace_batched_fwd(const iact_scene*, const iact_batch*, int32_t, const iact_launch*,
                               void*, size_t, float*, void*) {
  return set_error(IACT_ERR_UNSUPPORTED, "batched trace not built yet");
}
int32_t iact_trace_batched_bwd(const iact_scene*, const iact_batch*, int32_t, const float*,
                               const iact_launch*, void*, size_t, double*, void*) {
  return set_error(IACT_ERR_UNSUPPORTED, "batched trace not built yet");
}
int32_t iact_l2_loss_grad(const float*, const float*, int64_t, int32_t, float*, double*, void*) {
  return set_error(IACT_ERR_UNSUPPORTED, "not built yet");
}
int32_t iact_adam_step(float*, const double*, float*, float*, int32_t*, int64_t, float, float,
                       float, float, const int32_t*, void*) {
  return set_error(IACT_ERR_UNSUPPORTED, "not built yet");
}
int32_t iact_fit_graph_create(const iact_scene*, const iact_batch*, int32_t, const iact_launch*,
                              void*, size_t, float*, const float*, float*, double*, double*,
                              float*, float*, float*, int32_t*, float, void**) {
  return set_error(IACT_ERR_UNSUPPORTED, "not built yet");
}
int32_t iact_graph_launch(void*, void*) { return set_error(IACT_ERR_UNSUPPORTED, "not built yet"); }
int32_t iact_graph_destroy(void*) { return set_error(IACT_ERR_UNSUPPORTED, "not built yet"); }
}
