// Launch-shape autotuner: times candidate iact_launch settings for one
// forward render on the caller's own inputs and returns the fastest.  The
// reference has nothing to tune (XLA picks its own launch shapes); the
// candidates here are the knobs that measurably matter on gfx950 -- image
// layout (lanes = samples of a facet or lanes = sources), block size,
// blocks per CU, and the source tile / sample group.
#include <float.h>
#include <string.h>

#include "iact_abi_internal.h"

using namespace iact;

extern "C" {

int32_t iact_autotune(const iact_scene* scene, const double* sources, const float* values,
                      int32_t n_sources, int32_t source_type, int32_t mode,
                      const iact_launch* base, void* workspace, size_t workspace_bytes,
                      float* out, int32_t reps, iact_launch* best, float* best_ms,
                      void* stream) {
  if (!best) return set_error(IACT_ERR_INVALID, "null result launch");
  int rc = validate_scene(scene);
  if (rc) return rc;
  if (mode != IACT_MODE_IMAGE && mode != IACT_MODE_RESPONSE)
    return set_error(IACT_ERR_INVALID, "autotune covers image and response renders");
  if (!out) return set_error(IACT_ERR_INVALID, "null output");
  if (reps < 1) reps = 3;
  const iact_launch b0 = base ? *base : iact_launch{};
  *best = b0;
  if (best_ms) *best_ms = -1.f;
  // candidate grid (the deterministic flag and accumulate mode are kept)
  iact_launch cand[64];
  int nc = 0;
  const int layouts[2] = {1, 2};
  const int threads[2] = {256, 128};
  const int bpcs[2] = {16, 64};
  const int tiles[2] = {64, 256};
  for (int li = 0; li < (mode == IACT_MODE_IMAGE ? 2 : 1); ++li)
    for (int ti = 0; ti < 2; ++ti)
      for (int bi = 0; bi < 2; ++bi)
        for (int si = 0; si < 2; ++si) {
          iact_launch L = b0;
          L.layout = mode == IACT_MODE_IMAGE ? layouts[li] : 0;
          L.block_threads = threads[ti];
          L.blocks_per_cu = bpcs[bi];
          if (L.layout == 2) {
            L.sample_group = tiles[si];
            L.source_tile = 0;
          } else {
            L.source_tile = tiles[si];
            L.sample_group = 0;
          }
          cand[nc++] = L;
        }
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return hip_error(e, "hipEventCreate");
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_error(e, "hipEventCreate");
  }
  float best_t = FLT_MAX;
  int tried = 0;
  for (int i = 0; i < nc; ++i) {
    const iact_launch& L = cand[i];
    if (L.layout == 2 && (b0.deterministic || scene->n_stages > 0)) continue;
    if (iact_workspace_size(scene, n_sources, mode, &L) > workspace_bytes) continue;
    // warm-up launch; a candidate the library rejects is skipped
    if (iact_trace_fwd(scene, sources, values, n_sources, source_type, mode, &L, workspace,
                       workspace_bytes, out, nullptr, nullptr, stream) != IACT_OK)
      continue;
    if ((e = hipEventRecord(e0, st)) != hipSuccess) break;
    for (int r = 0; r < reps; ++r)
      (void)iact_trace_fwd(scene, sources, values, n_sources, source_type, mode, &L, workspace,
                           workspace_bytes, out, nullptr, nullptr, stream);
    if ((e = hipEventRecord(e1, st)) != hipSuccess) break;
    if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
    ms /= (float)reps;
    ++tried;
    if (ms < best_t) {
      best_t = ms;
      *best = L;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return hip_error(e, "autotune timing");
  if (tried == 0)
    return set_error(IACT_ERR_UNSUPPORTED, "no candidate launch fits the workspace given");
  if (best_ms) *best_ms = best_t;
  return IACT_OK;
}

}  // extern "C"
