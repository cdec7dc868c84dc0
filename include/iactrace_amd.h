/*
 * iactrace_amd -- C ABI of the MI355X (gfx950) ray-tracing hot path.
 *
 * The reference (GerritRo/iactrace, pure JAX) has no FFI: its boundary is the
 * jitted Python functions render / render_debug / render_response_matrix
 * (iactrace/core/render.py:174,223,271), the MC pupil sampler
 * (iactrace/core/integrators.py:58-188) and the per-group world transform
 * (iactrace/telescope/mirrors.py:64-79).  Each entry point below replaces one
 * of those; the Python package iactrace_amd binds them with ctypes.
 *
 * Conventions
 *  - Every buffer pointer is a DEVICE pointer unless the field says "host".
 *    The library never allocates, frees or retains caller memory; scratch is
 *    passed in as `workspace` (size from iact_workspace_size).
 *  - Every call returns IACT_OK (0) or a negative status; the text of the
 *    last error on the calling thread is returned by iact_last_error().
 *    Nothing is thrown across the ABI.  Arguments are validated before any
 *    HIP call is made, so invalid descriptors fail without touching a GPU.
 *  - Launches are asynchronous on `stream` (a hipStream_t; NULL = default).
 *  - Output buffers are overwritten (the library clears them first).
 *  - Floats are IEEE fp32; source coordinates are fp64 (point sources at
 *    ~1e10 m need it for the direction computation).
 */
#ifndef IACTRACE_AMD_H
#define IACTRACE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IACT_ABI_VERSION 1
#define IACT_MAX_ASPHERIC 8      /* polynomial terms per surface            */
#define IACT_MAX_POLY_VERTS 8    /* polygon aperture vertices               */
#define IACT_MAX_SURFACES 16     /* distinct surfaces per scene             */
#define IACT_MIRROR_STRIDE 40    /* floats per stage>=1 mirror table row    */

enum iact_status {
  IACT_OK = 0,
  IACT_ERR_INVALID = -1,      /* bad descriptor / argument                  */
  IACT_ERR_UNSUPPORTED = -2,  /* valid request this build cannot serve      */
  IACT_ERR_HIP = -3,          /* a HIP runtime call failed                  */
  IACT_ERR_WORKSPACE = -4     /* workspace missing or too small             */
};

enum iact_source_type { IACT_SOURCE_POINT = 0, IACT_SOURCE_PARALLEL = 1 };
enum iact_mode { IACT_MODE_IMAGE = 0, IACT_MODE_RESPONSE = 1, IACT_MODE_DEBUG = 2 };
enum iact_sensor_kind { IACT_SENSOR_SQUARE = 0, IACT_SENSOR_HEX = 1 };
enum iact_aperture { IACT_APERTURE_DISK = 0, IACT_APERTURE_POLYGON = 1 };

/* Conic + even polynomial surface (reference AsphericSurface,
 * iactrace/core/surfaces.py:8-39): sag(r2) = c r2 / (1 + sqrt(1-(1+k)c^2 r2))
 * + sum_i aspheric[i] * r2^(2(i+1)).                                          */
typedef struct iact_surface {
  float curvature;
  float conic;
  int32_t n_aspheric;
  float aspheric[IACT_MAX_ASPHERIC];
} iact_surface;

/* Stage-0 (primary) facets and their MC samples in the facet frame
 * (reference MirrorGroup fields, iactrace/telescope/mirrors.py:46-62).        */
typedef struct iact_facets {
  int32_t n_facets;
  int32_t n_samples;
  const float* points;             /* (n_facets, n_samples, 3)                */
  const float* normals;            /* (n_facets, n_samples, 3)                */
  const float* delta;              /* (n_facets, n_samples, 3) unit roughness */
  const float* weights;            /* (n_facets, n_samples)                   */
  const float* positions;          /* (n_facets, 3)                           */
  const float* rotations;          /* (n_facets, 3) Euler degrees (tip,tilt,rot) */
  const float* perturbation_scale; /* (n_facets) radians                      */
  const int32_t* surface_index;    /* (n_facets) -> scene surfaces[]; NULL ok
                                      unless aspheric gradients are requested */
  const float* offsets;            /* (n_facets, 2) offset on the parent
                                      surface; NULL ok as above              */
  const float* area_over_m;        /* (n_facets) aperture area / n_samples;
                                      NULL ok as above                       */
  float max_sample_radius;         /* an upper bound of |point| over every
                                      sample, metres; 0 = not given.  A point
                                      source farther than 2^30 times
                                      max(this, 1 mm) from a facet takes one
                                      direction per (facet, source), formed
                                      from the facet position (the samples'
                                      own differ from it by < 2^-30 rad);
                                      otherwise, and when 0, one per ray    */
} iact_facets;

/* Obstruction tables, one per primitive type, in the reference's group order
 * (iactrace/core/obstructions.py:258-277).  Row layouts (fp32):
 *   cylinders       (n, 8):  p1[3], unit axis[3], height, radius
 *   boxes           (n, 6):  min[3], max[3]
 *   spheres         (n, 4):  center[3], radius
 *   oriented_boxes  (n, 15): center[3], half_extents[3], rotation[9] row-major
 *   triangles       (n, 9):  v0[3], v1[3], v2[3]                              */
typedef struct iact_obstructions {
  int32_t n_cylinders;     const float* cylinders;
  int32_t n_boxes;         const float* boxes;
  int32_t n_spheres;       const float* spheres;
  int32_t n_oriented_boxes; const float* oriented_boxes;
  int32_t n_triangles;     const float* triangles;
} iact_obstructions;

/* Sensor (reference SquareSensor sensors/square.py:28-91, HexagonalSensor
 * sensors/hexagonal.py:104-194 and their Gaussian-splat variants).  All
 * derived grid numbers are passed as the fp32 values the reference computes
 * them to; `rotation` is euler_to_matrix(sensor.rotation), row-major, whose
 * columns are the frame (u1, u2, normal).                                     */
typedef struct iact_sensor {
  int32_t kind;            /* iact_sensor_kind                               */
  int32_t splat;           /* 0 = hard binning, 1 = Gaussian splat           */
  int32_t n_pixels;
  float position[3];
  float rotation[9];
  /* square */
  int32_t width, height;
  float x0, y0, dx, dy, edge_width;
  /* hexagonal */
  float hex_size, hex_inradius, grid_rotation;   /* grid_rotation radians    */
  float grid_offset[2];
  float edge_threshold;    /* fp32(1 - edge_width / inradius)                */
  int32_t q_min, r_min, table_q, table_r;
  const int32_t* lookup_table;   /* (table_q, table_r), -1 = empty          */
  /* splat */
  float sigma;
  int32_t kernel_size;
} iact_sensor;

/* Everything one render reads.  Stage>=1 mirrors are rows of `mirrors`
 * (IACT_MIRROR_STRIDE floats each):
 *   [0..2] position, [3..5] Euler degrees, [6..14] rotation row-major
 *   (local -> world), [15..16] offset, [17] aperture (0 disk, 1 polygon),
 *   [18] disk radius, [19] n_vertices, [20..35] polygon vertices (x,y) CCW,
 *   [36] optical stage (>=1), [37] group ordinal, [38] surface index.        */
typedef struct iact_scene {
  iact_facets facets;
  int32_t n_surfaces;
  iact_surface surfaces[IACT_MAX_SURFACES];
  int32_t n_mirrors;
  const float* mirrors;
  int32_t n_stages;          /* distinct stages >= 1 present in `mirrors`   */
  int32_t stages[8];         /* host: those stages in ascending order       */
  iact_obstructions obstructions;
  iact_sensor sensor;
} iact_scene;

/* Launch shape (filled by the autotuner; zeros select defaults).            */
typedef struct iact_launch {
  int32_t block_threads;   /* multiple of 64, <= 256                         */
  int32_t source_tile;     /* sources per work item                          */
  int32_t blocks_per_cu;   /* persistent grid = CUs * blocks_per_cu          */
  int32_t deterministic;   /* 1: fixed-order reductions, bitwise repeatable  */
  int32_t accumulate;      /* pixel accumulation: 0 wave-merged (default),
                              1 per-lane run-length, 2 direct per-ray atomics;
                              deterministic mode always uses 0               */
  int32_t chunk_block;     /* work order: sample chunks per block, inside
                              which the source tile is the slower index;
                              0 = one grid sweep of waves (default),
                              -1 = all chunks (source tile slowest)          */
  int32_t block_order;     /* 0 = XCD-major block numbering (default),
                              1 = hardware block numbering                  */
  int32_t layout;          /* image renders: 0 = automatic, 1 = lanes are
                              the samples of one facet (wave-merged pixel
                              sums; best when a wave's rays share pixels),
                              2 = lanes are sources, samples read from a
                              per-call world-space record table in the
                              workspace (one LDS add per ray; best for many
                              sources spread over the sensor)               */
  int32_t sample_group;    /* layout 2: samples per work item (0 = 64)     */
  int32_t culling;         /* obstruction culling: a prepass bounds, per
                              facet, the bundle of source-facing rays and
                              keeps the obstructions it can reach (a bit
                              mask per facet in the workspace); shadows of
                              the other obstructions are not tested.  0 =
                              automatic (on with any obstruction), 1 = on,
                              -1 = off.  Results are identical either way. */
} iact_launch;

/* Gradient outputs (fp64, overwritten).  NULL = not requested.               */
typedef struct iact_grads {
  double* rotations;       /* (n_facets, 3) d/d(Euler degree)                 */
  double* positions;       /* (n_facets, 3)                                   */
  double* sources;         /* (n_sources, 3)                                  */
  double* values;          /* (n_sources)                                     */
  double* aspheric;        /* (n_surfaces, IACT_MAX_ASPHERIC)                 */
} iact_grads;

/* Pupil-sampling request for one mirror group (replaces
 * MCIntegrator._sample_disk_group / _sample_polygon_group,
 * iactrace/core/integrators.py:97-188).  Streams are Philox4x32-10 with
 * counter (sample, facet, group, stream) and key (seed lo, seed hi).        */
typedef struct iact_sample_desc {
  uint64_t seed;
  int32_t group;
  int32_t n_facets;
  int32_t n_samples;
  int32_t aperture;          /* iact_aperture                                 */
  int32_t n_vertices;
  const float* radii;        /* (n_facets) for disks                          */
  const float* vertices;     /* (n_facets, n_vertices, 2) for polygons        */
  const float* offsets;      /* (n_facets, 2)                                 */
  iact_surface surface;
} iact_sample_desc;

/* Batched render of many telescope instances that share one facet set but
 * carry their own stage-0 rotations and uneven source lists.                 */
typedef struct iact_batch {
  int32_t n_instances;
  const float* rotations;          /* (n_instances, n_facets, 3)            */
  const int64_t* source_offsets;   /* device (n_instances+1) prefix offsets */
  const int64_t* host_source_offsets; /* HOST copy of the same offsets      */
  const double* sources;           /* (total, 3)                            */
  const float* values;             /* (total)                               */
} iact_batch;

/* ---- library ------------------------------------------------------------ */
int32_t iact_abi_version(void);
const char* iact_last_error(void);
/* sizeof of the ABI structs, for binding self-checks: 0 surface, 1 facets,
 * 2 obstructions, 3 sensor, 4 scene, 5 launch, 6 grads, 7 sample_desc,
 * 8 batch; -1 for anything else.                                            */
int64_t iact_struct_size(int32_t which);
/* Device properties used for launch shaping (CU count, max LDS per block). */
int32_t iact_device_info(int32_t device, int32_t* n_cu, int32_t* lds_per_block,
                         char* arch, size_t arch_len);

/* ---- sampling (integrators.py:58-188, sampling.py:8-67, reflection.py:22-50)
 * Writes local-frame points, unit normals, unit roughness deltas and
 * weights cos_z * n_samples / area for every (facet, sample).               */
int32_t iact_sample_facets(const iact_sample_desc* desc, float* points, float* normals,
                           float* delta, float* weights, void* stream);

/* ---- world transform (mirrors.py:64-79) --------------------------------
 * out: SoA (7, n_facets * n_samples): px, py, pz, nx, ny, nz, weight.       */
int32_t iact_facet_world(const iact_facets* facets, float* out, void* stream);

/* ---- sensor binning alone (SquareSensor.accumulate square.py:66-91,
 * HexagonalSensor.accumulate hexagonal.py:174-194 and the splat variants) --
 * out (n_pixels) = segment-sum of v at the pixel of each (x, y); pixel
 * (nullable) receives the hard pixel id or -1 per point.                     */
int32_t iact_sensor_accumulate(const iact_sensor* sensor, const float* x, const float* y,
                               const float* v, int64_t n, float* out, int32_t* pixel,
                               void* stream);

/* ---- forward render (render.py:174-324) --------------------------------
 * mode IMAGE:    out = image (n_pixels)            == render
 * mode RESPONSE: out = (n_sources, n_pixels)        == render_response_matrix
 * mode DEBUG:    out = values (n_facets*n_sources*n_samples), debug_xy the
 *                sensor-plane (x, y) pairs, debug_pixel (nullable) the hard
 *                pixel id or -1, facet-major       == render_debug            */
size_t iact_workspace_size(const iact_scene* scene, int32_t n_sources, int32_t mode,
                           const iact_launch* launch);
int32_t iact_trace_fwd(const iact_scene* scene, const double* sources, const float* values,
                       int32_t n_sources, int32_t source_type, int32_t mode,
                       const iact_launch* launch, void* workspace, size_t workspace_bytes,
                       float* out, float* debug_xy, int32_t* debug_pixel, void* stream);

/* ---- forward render without allocation -----------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): iact_trace_fwd with the workspace taken from a
 * library-owned static device array of IACT_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends) instead of a caller buffer.  Arguments,
 * outputs and errors as iact_trace_fwd; an image render whose default
 * layout needs more scratch than that (the source-major layout's record
 * table, 64 B per facet sample) runs the sample-major layout instead, and a
 * render that still does not fit fails with IACT_ERR_UNSUPPORTED.  Calls no
 * allocating routine.  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (on different streams or host threads) race and
 * must be ordered by the caller.                                            */
#define IACT_NOALLOC_SCRATCH_BYTES (32u << 20)
int32_t iact_trace_fwd_noalloc(const iact_scene* scene, const double* sources,
                               const float* values, int32_t n_sources, int32_t source_type,
                               int32_t mode, const iact_launch* launch, float* out,
                               float* debug_xy, int32_t* debug_pixel, void* stream);

/* ---- adjoint (reverse-mode AD of render, which the reference gets from
 * jax.grad through render.py:174-220) --------------------------------------
 * Given dL/dimage, writes dL/d(stage-0 rotations, positions), dL/d(sources,
 * values) and dL/d(aspheric coefficients of every surface).  Recomputes every
 * ray; stores no per-ray state.                                              */
size_t iact_bwd_workspace_size(const iact_scene* scene, int32_t n_sources,
                               const iact_launch* launch);
int32_t iact_trace_bwd(const iact_scene* scene, const double* sources, const float* values,
                       int32_t n_sources, int32_t source_type, const float* grad_image,
                       const iact_launch* launch, void* workspace, size_t workspace_bytes,
                       const iact_grads* grads, void* stream);

/* ---- batched (no reference counterpart: the reference renders one
 * telescope per call).  Many instances of one telescope, each with its own
 * stage-0 rotations and source list; images (n_instances, n_pixels) and, for
 * the backward, per-instance rotation gradients (n_instances, n_facets, 3).
 * Primary-only scenes (the default launch) run the plain source-major
 * forward and adjoint instance by instance on one stream, sharing the
 * workspace; with layout 1 or deterministic launches, and for scenes with
 * stage >= 1 mirrors, one persistent sample-major launch covers every
 * (instance, facet chunk, source tile) work item.  The batched backward
 * covers primary-only scenes.                                               */
size_t iact_batch_workspace_size(const iact_scene* scene, const iact_batch* batch,
                                 const iact_launch* launch);
int32_t iact_trace_batched_fwd(const iact_scene* scene, const iact_batch* batch,
                               int32_t source_type, const iact_launch* launch,
                               void* workspace, size_t workspace_bytes, float* images,
                               void* stream);
int32_t iact_trace_batched_bwd(const iact_scene* scene, const iact_batch* batch,
                               int32_t source_type, const float* grad_images,
                               const iact_launch* launch, void* workspace,
                               size_t workspace_bytes, double* grad_rotations,
                               void* stream);

/* ---- alignment-fit step helpers (no reference counterpart) -------------
 * loss = sum((image - target)^2) per instance; grad_image = 2 (image-target).
 * Adam update of float rotations from fp64 gradients (torch.optim.Adam
 * semantics, bias-corrected; `step` is a device int32 counter advanced by
 * the call so a captured graph can be replayed).                            */
int32_t iact_l2_loss_grad(const float* images, const float* target, int64_t n_per_instance,
                          int32_t n_instances, float* grad_images, double* loss, void* stream);
int32_t iact_adam_step(float* params, const double* grads, float* exp_avg, float* exp_avg_sq,
                       int32_t* step, int64_t n, float lr, float beta1, float beta2,
                       float eps, const int32_t* mask3, void* stream);

/* ---- launch-shape autotuner (no reference counterpart: XLA chooses its own
 * launch shapes) ------------------------------------------------------------
 * Times candidate iact_launch settings -- image layout, block size, blocks
 * per CU, source tile / sample group -- for one IMAGE or RESPONSE render of
 * the given inputs (`reps` timed renders each, after a warm-up) and writes
 * the fastest to *best (and its time to *best_ms, nullable).  `base` supplies
 * the fields that are kept (deterministic, accumulate; nullable).  Candidates
 * needing more workspace than given are skipped.  Synchronises `stream`;
 * `out` receives the last render.                                            */
int32_t iact_autotune(const iact_scene* scene, const double* sources, const float* values,
                      int32_t n_sources, int32_t source_type, int32_t mode,
                      const iact_launch* base, void* workspace, size_t workspace_bytes,
                      float* out, int32_t reps, iact_launch* best, float* best_ms,
                      void* stream);

/* ---- HIP graphs ---------------------------------------------------------
 * Capture a fit iteration (batched forward, loss gradient, batched backward,
 * Adam) once and replay it.  The graph handle owns its executable graph; all
 * buffers named at capture must outlive it.                                  */
int32_t iact_fit_graph_create(const iact_scene* scene, const iact_batch* batch,
                              int32_t source_type, const iact_launch* launch,
                              void* workspace, size_t workspace_bytes, float* images,
                              const float* target, float* grad_images, double* loss,
                              double* grad_rotations, float* rotations, float* exp_avg,
                              float* exp_avg_sq, int32_t* step, float lr, void** graph_out);
int32_t iact_graph_launch(void* graph, void* stream);
int32_t iact_graph_destroy(void* graph);

#ifdef __cplusplus
}
#endif
#endif /* IACTRACE_AMD_H */
