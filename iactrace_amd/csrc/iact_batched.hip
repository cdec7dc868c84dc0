// Batched render of many telescope instances, the alignment-fit step helpers
// (L2 image loss and its gradient, Adam) and HIP-graph capture of one whole
// fit iteration.  None of these has a counterpart in the reference, which
// renders one telescope per jitted call (render.py:174-220) and leaves the
// fit loop to the user (optax over jax.grad).
//
// Instances share the stage-0 samples, positions, obstructions, stage>=1
// mirrors and sensor; each carries its own stage-0 rotations (the fitted
// tilts), its own source list and its own image.  The trace and adjoint
// kernels are the plain ones: their work list is extended to (instance,
// sample chunk, source tile) items through a per-instance item prefix that a
// one-thread kernel writes into the workspace, so nothing depends on host
// state once launched and the whole iteration can be captured in a graph.
#include <math.h>
#include <string.h>

#include <new>

#include "iact_abi_internal.h"

namespace iact {

// item_prefix[i] = sum_{j<i} items_j, items_j = chunks * ceil(ns_j / tile)
// rounded up to a multiple of `pad` (the waves of a block), so that a block's
// waves never straddle two instances; item_prefix[n] = total.  Also the
// per-instance sizes and decode multipliers.
__global__ void k_batch_prefix(const int64_t* src_prefix, int n_inst, long long nchunks,
                               int tile, int pad, uint32_t span, long long* item_prefix,
                               InstDesc* desc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  long long acc = 0;
  item_prefix[0] = 0;
  for (int i = 0; i < n_inst; ++i) {
    const long long ns = src_prefix[i + 1] - src_prefix[i];
    const long long nst = (ns + tile - 1) / tile;
    const long long items = nchunks * nst;
    acc += (items + pad - 1) / pad * pad;
    item_prefix[i + 1] = acc;
    InstDesc d;
    d.pb = make_fastdiv((uint32_t)(span * (nst > 0 ? nst : 1)));
    d.nst = (int32_t)nst;
    d.ns = (int32_t)ns;
    d.sbase = src_prefix[i];
    desc[i] = d;
  }
}

// One block per instance: g = 2 (image - target), loss = sum (image - target)^2
// (fp64 accumulation, fixed order: bitwise repeatable).
__global__ __launch_bounds__(IACT_MAX_BLOCK) void k_l2_loss_grad(const float* img,
                                                                const float* tgt, long long n,
                                                                float* g, double* loss) {
  __shared__ double red[IACT_MAX_BLOCK / 64];
  const size_t base = (size_t)blockIdx.x * n;
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    const float d = img[base + i] - tgt[base + i];
    if (g) g[base + i] = 2.f * d;
    s = fma((double)d, (double)d, s);
  }
  s = wave_sum_d(s);
  if (lane_id() == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0 && loss) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    loss[blockIdx.x] = t;
  }
}

// torch.optim.Adam (defaults: no weight decay, no amsgrad, not maximising)
// for one parameter tensor; the step about to be taken is *step + 1.
__global__ void k_adam(float* p, const double* grad, float* m, float* v, const int32_t* step,
                       long long n, float lr, float b1, float b2, float eps, int mask) {
  const double t = (double)(*step + 1);
  const double bc1 = 1.0 - pow((double)b1, t);
  const double bc2 = 1.0 - pow((double)b2, t);
  const float step_size = (float)((double)lr / bc1);
  const float bc2_sqrt = (float)sqrt(bc2);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    if (!((mask >> (int)(i % 3)) & 1)) continue;
    const float g = (float)grad[i];
    const float mi = fmaf(1.f - b1, g - m[i], m[i]);          // exp_avg.lerp_(g, 1 - b1)
    const float vi = fmaf((1.f - b2) * g, g, b2 * v[i]);       // mul_(b2).addcmul_(g, g, 1 - b2)
    m[i] = mi;
    v[i] = vi;
    const float denom = __builtin_sqrtf(vi) / bc2_sqrt + eps;
    p[i] = p[i] - step_size * (mi / denom);
  }
}

__global__ void k_step_inc(int32_t* step) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *step = *step + 1;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace of a batched render: item prefix, then the instance table.
static size_t prefix_bytes(int n) { return align256((size_t)(n + 1) * sizeof(long long)); }
static size_t batch_ws_bytes(int n) {
  return prefix_bytes(n) + align256((size_t)n * sizeof(InstDesc));
}

static int validate_batch(const iact_batch* b, long long& total, int& max_ns) {
  if (!b) return set_error(IACT_ERR_INVALID, "null batch descriptor");
  if (b->n_instances < 1) return set_error(IACT_ERR_INVALID, "n_instances %d", b->n_instances);
  if (!b->rotations || !b->source_offsets || !b->host_source_offsets)
    return set_error(IACT_ERR_INVALID, "batch needs rotations and source offsets");
  const int64_t* h = b->host_source_offsets;
  if (h[0] != 0) return set_error(IACT_ERR_INVALID, "source offsets must start at 0");
  max_ns = 0;
  for (int i = 0; i < b->n_instances; ++i) {
    const int64_t ns = h[i + 1] - h[i];
    if (ns < 0) return set_error(IACT_ERR_INVALID, "source offsets must not decrease");
    if (ns > 0x7fffffff) return set_error(IACT_ERR_UNSUPPORTED, "instance with >2^31 sources");
    if (ns > max_ns) max_ns = (int)ns;
  }
  total = h[b->n_instances];
  if (total > 0 && (!b->sources || !b->values))
    return set_error(IACT_ERR_INVALID, "null batch sources/values");
  return IACT_OK;
}

// Launch shape of a batch (no workspace needed): trace arguments, shape and
// the culling decision.  The facet cones of a batch cover the sources of
// every instance (a cone only has to contain the rays it bounds, and the
// instances' samples stay within the same distance of their facet
// positions), so one candidate table serves all instances.
static int batch_shape(const iact_scene* scene, const iact_batch* batch, int source_type,
                       const iact_launch* launch, TraceArgs& A, LaunchShape& shp,
                       iact_launch& L, DevInfo& dev, long long& total) {
  int rc = validate_scene(scene);
  if (rc) return rc;
  int max_ns;
  rc = validate_batch(batch, total, max_ns);
  if (rc) return rc;
  if (source_type != IACT_SOURCE_POINT && source_type != IACT_SOURCE_PARALLEL &&
      source_type != -1)
    return set_error(IACT_ERR_INVALID, "unknown source type %d", source_type);
  if (launch && launch->deterministic)
    return set_error(IACT_ERR_UNSUPPORTED, "batched renders have no deterministic mode");
  rc = device_info(dev);
  if (rc) return rc;
  fill_trace_args(A, *scene, batch->sources, batch->values, max_ns);
  // sample-major layout, wave-merged accumulation into a per-block LDS image
  // of the instance being traced (no run-length or deterministic variants)
  L = launch ? *launch : iact_launch{};
  L.accumulate = 0;
  L.layout = 1;
  rc = shape_trace(*scene, max_ns, IACT_MODE_IMAGE, &L, A, shp, dev, -1);
  if (rc) return rc;
  if (total >= (1ll << 31)) {
    if (L.culling == 1)
      return set_error(IACT_ERR_UNSUPPORTED, "culling a batch of 2^31 or more sources");
    A.cull_all = 1;  // the cone pass counts sources in 32 bits
  }
  return IACT_OK;
}

// The batched forward traces instance by instance with the source-major
// kernels (iact_trace_fwd on the instance's rotations and sources) wherever
// a plain render would take that layout -- primary-only scenes, not
// deterministic, fewer than 2^31 samples -- unless the launch asks for the
// sample-major layout.  The backward keeps the batched adjoint.
static bool batch_source_major(const iact_scene* scene, const iact_launch* launch) {
  if (scene->n_stages > 0) return false;
  if (launch && (launch->layout == 1 || launch->deterministic)) return false;
  return (long long)scene->facets.n_facets * scene->facets.n_samples < (1ll << 31);
}

// Workspace of one instance's source-major forward with ns sources: the
// record table and the culling tables after it (an upper bound for an
// instance that takes the sample-major layout, which has no records).
static size_t instance_need(const iact_scene* scene, long long ns, int words, int all) {
  const size_t recs = (size_t)IACT_REC_WORDS *
                      ((size_t)scene->facets.n_facets * scene->facets.n_samples + 2) *
                      sizeof(float);
  const long long total = (long long)scene->facets.n_facets * scene->facets.n_samples;
  const CullSpec c{ns, 1, nullptr, true, (total + 63) / 64};
  const size_t end = cull_layout(scene->facets.n_facets, words, all, recs, c).end;
  return image64_offset(end) + (size_t)scene->sensor.n_pixels * sizeof(double);
}

// Workspace end of the batched kernels' tables (prefix, instance table,
// candidate tables); the instances' fp64 image sums follow it.
static size_t batch_tables_end(const iact_scene* scene, const iact_batch* batch, int words,
                               int all, long long total) {
  if (total >= (1ll << 31)) all = 1;  // as batch_shape
  const CullSpec c{total, batch->n_instances, batch->source_offsets, true};
  return cull_layout(scene->facets.n_facets, words, all, batch_ws_bytes(batch->n_instances), c)
      .end;
}

// Workspace of a batch: prefix and instance table, then the candidate
// tables when the batch culls (known without a device); at least one
// instance's source-major forward when the forward takes that path.
static int batch_need(const iact_scene* scene, const iact_batch* batch, const iact_launch* launch,
                      size_t& need) {
  int rc = validate_scene(scene);
  if (rc) return rc;
  long long total;
  int max_ns;
  rc = validate_batch(batch, total, max_ns);
  if (rc) return rc;
  int words = 0, all = 0;
  rc = cull_words_for(*scene, launch, words, all);
  if (rc) return rc;
  need = batch_source_major(scene, launch) ? instance_need(scene, max_ns, words, all) : 0;
  if (bwd_source_major(*scene, max_ns, launch, false)) {
    // the backward instance by instance through iact_trace_bwd
    DevInfo dev;
    rc = device_info(dev);
    if (rc) return rc;
    TraceArgs A;
    LaunchShape shp;
    AdjSmLayout lay;
    rc = adj_sm_layout(*scene, nullptr, nullptr, max_ns, launch, dev, A, shp, lay);
    if (rc) return rc;
    if (lay.end > need) need = lay.end;
  }
  const size_t bat = image64_offset(batch_tables_end(scene, batch, words, all, total)) +
                     (size_t)batch->n_instances * scene->sensor.n_pixels * sizeof(double);
  if (bat > need) need = bat;
  return IACT_OK;
}

// Shared set-up of the batched forward and backward: trace arguments with
// the batch fields, the launch shape and the item count from the host
// offsets, and the obstruction candidate tables.
static int prepare_batch(const iact_scene* scene, const iact_batch* batch, int source_type,
                         const iact_launch* launch, void* workspace, size_t workspace_bytes,
                         TraceArgs& A, LaunchShape& shp, bool& empty,
                         const int64_t*& src_prefix, iact_launch& Lout, DevInfo& devout,
                         hipStream_t st, bool tables = true) {
  iact_launch L;
  DevInfo dev;
  long long total;
  if (source_type != IACT_SOURCE_POINT && source_type != IACT_SOURCE_PARALLEL)
    return set_error(IACT_ERR_INVALID, "unknown source type %d", source_type);
  if (launch && launch->deterministic)
    return set_error(IACT_ERR_UNSUPPORTED, "batched renders have no deterministic mode");
  size_t need = 0;
  int rc = batch_need(scene, batch, launch, need);
  if (rc) return rc;
  if (!workspace || workspace_bytes < need)
    return set_error(IACT_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need,
                     workspace_bytes);
  rc = batch_shape(scene, batch, source_type, launch, A, shp, L, dev, total);
  if (rc) return rc;
  if (tables && total > 0) {
    const CullSpec c{total, batch->n_instances, batch->source_offsets, true};
    rc = prepare_cull(A, workspace, batch_ws_bytes(batch->n_instances), source_type, st, c,
                      nullptr);
    if (rc) return rc;
  }
  const int wpb = shp.threads / 64;
  const long long nchunks = (long long)A.nf * A.cpf;
  long long items = 0;
  for (int i = 0; i < batch->n_instances; ++i) {
    const long long ns = batch->host_source_offsets[i + 1] - batch->host_source_offsets[i];
    const long long it = nchunks * ((ns + A.source_tile - 1) / A.source_tile);
    items += (it + wpb - 1) / wpb * wpb;
  }
  A.n_items = items;
  A.n_inst = batch->n_instances;
  A.item_prefix = (const long long*)workspace;
  A.inst_desc = (const InstDesc*)((char*)workspace + prefix_bytes(batch->n_instances));
  src_prefix = batch->source_offsets;
  A.rot_stride = 3ll * A.nf;
  A.rotations = batch->rotations;
  A.det = 0;
  empty = items == 0;
  Lout = L;
  devout = dev;
  // grid for the batched forward kernel; the adjoint regrids for its own
  return regrid(items, &L, dev, shp, A);
}

static hipError_t launch_prefix(const TraceArgs& A, const LaunchShape& shp,
                                const int64_t* src_prefix, void* workspace, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_prefix, dim3(1), dim3(64), 0, st, src_prefix, A.n_inst,
                     (long long)A.nchunks, A.source_tile, shp.threads / 64, A.span,
                     (long long*)workspace,
                     (InstDesc*)((char*)workspace + prefix_bytes(A.n_inst)));
  return hipGetLastError();
}

struct FitGraph {
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

}  // namespace iact

using namespace iact;

extern "C" {

size_t iact_batch_workspace_size(const iact_scene* scene, const iact_batch* batch,
                                 const iact_launch* launch) {
  size_t need = 0;
  return batch_need(scene, batch, launch, need) ? 0 : need;
}

int32_t iact_trace_batched_fwd(const iact_scene* scene, const iact_batch* batch,
                               int32_t source_type, const iact_launch* launch, void* workspace,
                               size_t workspace_bytes, float* images, void* stream) {
  if (scene && batch && batch_source_major(scene, launch)) {
    // instance by instance, sharing the workspace (one stream orders them);
    // checks in the order of the batched kernel's path
    if (source_type != IACT_SOURCE_POINT && source_type != IACT_SOURCE_PARALLEL)
      return set_error(IACT_ERR_INVALID, "unknown source type %d", source_type);
    size_t need = 0;
    int rc = batch_need(scene, batch, launch, need);  // validates scene and batch
    if (rc) return rc;
    if (!workspace || workspace_bytes < need)
      return set_error(IACT_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need,
                       workspace_bytes);
    if (!images) return set_error(IACT_ERR_INVALID, "null images");
    const long long P = scene->sensor.n_pixels;
    const int64_t* h = batch->host_source_offsets;
    for (int i = 0; i < batch->n_instances; ++i) {
      iact_scene si = *scene;
      si.facets.rotations = batch->rotations + (size_t)i * 3 * scene->facets.n_facets;
      rc = iact_trace_fwd(&si, batch->sources + 3 * h[i], batch->values + h[i],
                          (int32_t)(h[i + 1] - h[i]), source_type, IACT_MODE_IMAGE, launch,
                          workspace, workspace_bytes, images + (size_t)i * P, nullptr, nullptr,
                          stream);
      if (rc) return rc;
    }
    return IACT_OK;
  }
  TraceArgs A;
  LaunchShape shp;
  bool empty;
  const int64_t* sp;
  iact_launch L;
  DevInfo dev;
  int rc = prepare_batch(scene, batch, source_type, launch, workspace, workspace_bytes, A, shp,
                         empty, sp, L, dev, (hipStream_t)stream);
  if (rc) return rc;
  if (!images) return set_error(IACT_ERR_INVALID, "null images");
  hipStream_t st = (hipStream_t)stream;
  const long long P = scene->sensor.n_pixels;
  const long long nP = P * batch->n_instances;
  hipError_t e;
  if (empty) {
    e = launch_fill(images, nP, 0.f, st);
    return e == hipSuccess ? IACT_OK : hip_error(e, "clear images");
  }
  int words = 0, all = 0;
  rc = cull_words_for(*scene, launch, words, all);
  if (rc) return rc;
  long long total = batch->host_source_offsets[batch->n_instances];
  A.out64 = (double*)((char*)workspace +
                      image64_offset(batch_tables_end(scene, batch, words, all, total)));
  if ((e = launch_fill_d(A.out64, nP, 0.0, st)) != hipSuccess) return hip_error(e, "clear sums");
  if ((e = launch_prefix(A, shp, sp, workspace, st)) != hipSuccess) return hip_error(e, "k_batch_prefix");
  A.out = images;
  e = launch_trace(A, shp, source_type, IACT_MODE_IMAGE, st);
  if (e != hipSuccess) return hip_error(e, "k_trace(batched)");
  e = launch_f64_to_f32(A.out64, images, nP, st);
  return e == hipSuccess ? IACT_OK : hip_error(e, "k_f64_to_f32");
}

int32_t iact_trace_batched_bwd(const iact_scene* scene, const iact_batch* batch,
                               int32_t source_type, const float* grad_images,
                               const iact_launch* launch, void* workspace,
                               size_t workspace_bytes, double* grad_rotations, void* stream) {
  if (scene && batch && batch_source_major(scene, launch) && scene->n_stages == 0) {
    // instance by instance with the plain backward (the source-major
    // adjoint wherever it applies), sharing the workspace
    if (source_type != IACT_SOURCE_POINT && source_type != IACT_SOURCE_PARALLEL)
      return set_error(IACT_ERR_INVALID, "unknown source type %d", source_type);
    size_t need = 0;
    int rc = batch_need(scene, batch, launch, need);  // validates scene and batch
    if (rc) return rc;
    if (!workspace || workspace_bytes < need)
      return set_error(IACT_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need,
                       workspace_bytes);
    if (!grad_images || !grad_rotations)
      return set_error(IACT_ERR_INVALID, "null image or rotation gradients");
    const long long P = scene->sensor.n_pixels;
    const int nf = scene->facets.n_facets;
    const int64_t* h = batch->host_source_offsets;
    for (int i = 0; i < batch->n_instances; ++i) {
      iact_scene si = *scene;
      si.facets.rotations = batch->rotations + (size_t)i * 3 * nf;
      iact_grads gi;
      memset(&gi, 0, sizeof(gi));
      gi.rotations = grad_rotations + (size_t)i * 3 * nf;
      rc = iact_trace_bwd(&si, batch->sources + 3 * h[i], batch->values + h[i],
                          (int32_t)(h[i + 1] - h[i]), source_type, grad_images + (size_t)i * P,
                          launch, workspace, workspace_bytes, &gi, stream);
      if (rc) return rc;
    }
    return IACT_OK;
  }
  TraceArgs A;
  LaunchShape shp;
  bool empty;
  const int64_t* sp;
  iact_launch L;
  DevInfo dev;
  int rc = prepare_batch(scene, batch, source_type, launch, workspace, workspace_bytes, A, shp,
                         empty, sp, L, dev, (hipStream_t)stream);
  if (rc) return rc;
  if (!grad_images || !grad_rotations)
    return set_error(IACT_ERR_INVALID, "null image or rotation gradients");
  if (scene->n_stages > 0)
    return set_error(IACT_ERR_UNSUPPORTED,
                     "the adjoint kernel covers primary-only scenes (stage>=1 mirrors present)");
  // the adjoint stages one instance's image gradient (+ hex table) in LDS
  const size_t tab = hex_table_bytes(scene->sensor);
  shp.lds = (size_t)scene->sensor.n_pixels * sizeof(float) + tab;
  if (shp.lds > 65536)
    return set_error(IACT_ERR_UNSUPPORTED, "image gradient + tables need %zu bytes of LDS",
                     shp.lds);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e =
      launch_fill_d(grad_rotations, 3ll * A.nf * batch->n_instances, 0.0, st);
  if (e != hipSuccess) return hip_error(e, "clear grad");
  if (empty) return IACT_OK;
  AdjArgs B;
  memset(&B, 0, sizeof(B));
  B.T = A;
  B.grad_image = grad_images;
  B.g_rot = grad_rotations;
  rc = regrid(A.n_items, &L, dev, shp, B.T);
  if (rc) return rc;
  e = launch_prefix(B.T, shp, sp, workspace, st);
  if (e != hipSuccess) return hip_error(e, "k_batch_prefix");
  e = launch_adjoint(B, shp, source_type, st);
  return e == hipSuccess ? IACT_OK : hip_error(e, "k_trace_adj(batched)");
}

int32_t iact_l2_loss_grad(const float* images, const float* target, int64_t n_per_instance,
                          int32_t n_instances, float* grad_images, double* loss, void* stream) {
  if (n_per_instance < 0 || n_instances < 0)
    return set_error(IACT_ERR_INVALID, "negative sizes");
  if (n_instances == 0) return IACT_OK;
  if (!images || !target) return set_error(IACT_ERR_INVALID, "null images/target");
  if (!grad_images && !loss) return set_error(IACT_ERR_INVALID, "nothing to compute");
  hipLaunchKernelGGL(k_l2_loss_grad, dim3(n_instances), dim3(256), 0, (hipStream_t)stream,
                     images, target, (long long)n_per_instance, grad_images, loss);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? IACT_OK : hip_error(e, "k_l2_loss_grad");
}

int32_t iact_adam_step(float* params, const double* grads, float* exp_avg, float* exp_avg_sq,
                       int32_t* step, int64_t n, float lr, float beta1, float beta2, float eps,
                       const int32_t* mask3, void* stream) {
  if (n < 0) return set_error(IACT_ERR_INVALID, "negative size");
  if (!step) return set_error(IACT_ERR_INVALID, "null step counter");
  if (n > 0 && (!params || !grads || !exp_avg || !exp_avg_sq))
    return set_error(IACT_ERR_INVALID, "null Adam buffer");
  if (!(beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f && eps >= 0.f))
    return set_error(IACT_ERR_INVALID, "Adam betas must be in [0, 1), eps >= 0");
  const int mask = mask3 ? ((mask3[0] ? 1 : 0) | (mask3[1] ? 2 : 0) | (mask3[2] ? 4 : 0)) : 7;
  hipStream_t st = (hipStream_t)stream;
  long long g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_adam, dim3((int)g), dim3(256), 0, st, params, grads, exp_avg, exp_avg_sq,
                     (const int32_t*)step, (long long)n, lr, beta1, beta2, eps, mask);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, "k_adam");
  hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, st, step);
  e = hipGetLastError();
  return e == hipSuccess ? IACT_OK : hip_error(e, "k_step_inc");
}

int32_t iact_fit_graph_create(const iact_scene* scene, const iact_batch* batch,
                              int32_t source_type, const iact_launch* launch, void* workspace,
                              size_t workspace_bytes, float* images, const float* target,
                              float* grad_images, double* loss, double* grad_rotations,
                              float* rotations, float* exp_avg, float* exp_avg_sq,
                              int32_t* step, float lr, void** graph_out) {
  if (!graph_out) return set_error(IACT_ERR_INVALID, "null graph handle");
  *graph_out = nullptr;
  // validate everything before capture starts
  TraceArgs A;
  LaunchShape shp;
  bool empty;
  const int64_t* sp;
  iact_launch L;
  DevInfo dev;
  int rc = prepare_batch(scene, batch, source_type, launch, workspace, workspace_bytes, A, shp,
                         empty, sp, L, dev, nullptr, false);
  if (rc) return rc;
  if (scene->n_stages > 0)
    return set_error(IACT_ERR_UNSUPPORTED, "fit graphs cover primary-only scenes");
  if (!images || !target || !grad_images || !grad_rotations || !rotations || !exp_avg ||
      !exp_avg_sq || !step)
    return set_error(IACT_ERR_INVALID, "null fit buffer");
  if (rotations != batch->rotations)
    return set_error(IACT_ERR_INVALID, "the fitted rotations must be the batch rotations");
  if (!(lr > 0.f)) return set_error(IACT_ERR_INVALID, "learning rate must be > 0");
  FitGraph* G = new (std::nothrow) FitGraph;
  if (!G) return set_error(IACT_ERR_INVALID, "out of host memory");
  hipError_t e = hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete G;
    return hip_error(e, "hipStreamCreateWithFlags");
  }
  e = hipStreamBeginCapture(G->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(G->stream);
    delete G;
    return hip_error(e, "hipStreamBeginCapture");
  }
  const int n_inst = batch->n_instances;
  const long long P = scene->sensor.n_pixels;
  rc = iact_trace_batched_fwd(scene, batch, source_type, launch, workspace, workspace_bytes,
                              images, G->stream);
  if (!rc) rc = iact_l2_loss_grad(images, target, P, n_inst, grad_images, loss, G->stream);
  if (!rc)
    rc = iact_trace_batched_bwd(scene, batch, source_type, grad_images, launch, workspace,
                                workspace_bytes, grad_rotations, G->stream);
  if (!rc)
    rc = iact_adam_step(rotations, grad_rotations, exp_avg, exp_avg_sq, step,
                        3ll * scene->facets.n_facets * n_inst, lr, 0.9f, 0.999f, 1e-8f, nullptr,
                        G->stream);
  e = hipStreamEndCapture(G->stream, &G->graph);
  if (rc || e != hipSuccess) {
    if (G->graph) (void)hipGraphDestroy(G->graph);
    (void)hipStreamDestroy(G->stream);
    delete G;
    return rc ? rc : hip_error(e, "hipStreamEndCapture");
  }
  e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(G->graph);
    (void)hipStreamDestroy(G->stream);
    delete G;
    return hip_error(e, "hipGraphInstantiate");
  }
  *graph_out = G;
  return IACT_OK;
}

int32_t iact_graph_launch(void* graph, void* stream) {
  if (!graph) return set_error(IACT_ERR_INVALID, "null graph");
  FitGraph* G = (FitGraph*)graph;
  const hipError_t e = hipGraphLaunch(G->exec, (hipStream_t)stream);
  return e == hipSuccess ? IACT_OK : hip_error(e, "hipGraphLaunch");
}

int32_t iact_graph_destroy(void* graph) {
  if (!graph) return IACT_OK;
  FitGraph* G = (FitGraph*)graph;
  hipError_t e = hipGraphExecDestroy(G->exec);
  const hipError_t e2 = hipGraphDestroy(G->graph);
  const hipError_t e3 = hipStreamDestroy(G->stream);
  delete G;
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = e3;
  return e == hipSuccess ? IACT_OK : hip_error(e, "graph destroy");
}

}  // extern "C"
