// C ABI of libiactrace_amd: argument validation, launch shaping and the
// extern "C" entry points declared in include/iactrace_amd.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>

#include "iact_abi_internal.h"

namespace iact {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_error(hipError_t e, const char* where) {
  return set_error(IACT_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

// ------------------------------------------------------------- device info
static std::mutex g_dev_mu;
static DevInfo g_dev[64];
static bool g_dev_ok[64];

int device_info(DevInfo& out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_error(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) return set_error(IACT_ERR_UNSUPPORTED, "device ordinal %d", dev);
  std::lock_guard<std::mutex> lk(g_dev_mu);
  if (!g_dev_ok[dev]) {
    hipDeviceProp_t p;
    e = hipGetDeviceProperties(&p, dev);
    if (e != hipSuccess) return hip_error(e, "hipGetDeviceProperties");
    g_dev[dev].n_cu = p.multiProcessorCount;
    int lds = 0;
    e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return hip_error(e, "hipDeviceGetAttribute");
    g_dev[dev].lds_per_block = lds;
    strncpy(g_dev[dev].arch, p.gcnArchName, sizeof(g_dev[dev].arch) - 1);
    g_dev_ok[dev] = true;
  }
  out = g_dev[dev];
  return IACT_OK;
}

// ------------------------------------------------------------- validation
static bool finite3(const float* v) {
  return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
}

int validate_facets(const iact_facets& F) {
  if (F.n_facets < 0 || F.n_samples < 0)
    return set_error(IACT_ERR_INVALID, "negative facet/sample count (%d, %d)", F.n_facets,
                     F.n_samples);
  if (F.n_facets > 0 && F.n_samples == 0)
    return set_error(IACT_ERR_INVALID, "facets have no samples (resample the mirrors)");
  if (F.n_facets == 0) return IACT_OK;
  if (!F.points || !F.normals || !F.delta || !F.weights || !F.positions || !F.rotations ||
      !F.perturbation_scale)
    return set_error(IACT_ERR_INVALID, "null facet buffer");
  if ((long long)F.n_facets * F.n_samples > (1ll << 40))
    return set_error(IACT_ERR_UNSUPPORTED, "facet sample count too large");
  return IACT_OK;
}

int validate_sensor(const iact_sensor& S) {
  if (S.kind != IACT_SENSOR_SQUARE && S.kind != IACT_SENSOR_HEX)
    return set_error(IACT_ERR_INVALID, "unknown sensor kind %d", S.kind);
  if (S.n_pixels <= 0) return set_error(IACT_ERR_INVALID, "sensor has no pixels");
  if (S.n_pixels >= (1 << 24))
    return set_error(IACT_ERR_UNSUPPORTED, "sensor with %d pixels (max 2^24 - 1)", S.n_pixels);
  if (!finite3(S.position)) return set_error(IACT_ERR_INVALID, "non-finite sensor position");
  for (int i = 0; i < 9; ++i)
    if (!isfinite(S.rotation[i])) return set_error(IACT_ERR_INVALID, "non-finite sensor frame");
  if (S.kind == IACT_SENSOR_SQUARE) {
    if (S.width <= 0 || S.height <= 0 || (long long)S.width * S.height != S.n_pixels)
      return set_error(IACT_ERR_INVALID, "square sensor %dx%d does not match %d pixels", S.width,
                       S.height, S.n_pixels);
    if (!(S.dx > 0.f) || !(S.dy > 0.f))
      return set_error(IACT_ERR_INVALID, "square sensor pixel pitch must be > 0");
  } else {
    if (!S.lookup_table) return set_error(IACT_ERR_INVALID, "hex sensor without lookup table");
    if (S.table_q <= 0 || S.table_r <= 0)
      return set_error(IACT_ERR_INVALID, "hex lookup table shape %dx%d", S.table_q, S.table_r);
    if (!(S.hex_size > 0.f) || !(S.hex_inradius > 0.f))
      return set_error(IACT_ERR_INVALID, "hex pixel size must be > 0");
  }
  if (S.splat) {
    if (!(S.sigma > 0.f)) return set_error(IACT_ERR_INVALID, "splat sigma must be > 0");
    if (S.kernel_size < 0 || S.kernel_size > 4)
      return set_error(IACT_ERR_UNSUPPORTED, "splat kernel_size %d (0..4)", S.kernel_size);
  }
  return IACT_OK;
}

int validate_scene(const iact_scene* sc) {
  if (!sc) return set_error(IACT_ERR_INVALID, "null scene");
  int rc = validate_facets(sc->facets);
  if (rc) return rc;
  if (sc->n_surfaces < 0 || sc->n_surfaces > IACT_MAX_SURFACES)
    return set_error(IACT_ERR_INVALID, "n_surfaces %d (max %d)", sc->n_surfaces,
                     IACT_MAX_SURFACES);
  for (int i = 0; i < sc->n_surfaces; ++i) {
    const iact_surface& s = sc->surfaces[i];
    if (s.n_aspheric < 0 || s.n_aspheric > IACT_MAX_ASPHERIC)
      return set_error(IACT_ERR_UNSUPPORTED, "surface %d has %d aspheric terms (max %d)", i,
                       s.n_aspheric, IACT_MAX_ASPHERIC);
  }
  if (sc->n_mirrors < 0) return set_error(IACT_ERR_INVALID, "negative mirror count");
  if (sc->n_mirrors > 0 && !sc->mirrors) return set_error(IACT_ERR_INVALID, "null mirror table");
  if (sc->n_mirrors > 0 && sc->n_surfaces == 0)
    return set_error(IACT_ERR_INVALID, "stage>=1 mirrors but no surfaces");
  if (sc->n_stages < 0 || sc->n_stages > 8)
    return set_error(IACT_ERR_UNSUPPORTED, "n_stages %d (max 8)", sc->n_stages);
  if (sc->n_stages > 0 && sc->n_mirrors == 0)
    return set_error(IACT_ERR_INVALID, "stages listed without mirrors");
  for (int i = 0; i < sc->n_stages; ++i)
    if (sc->stages[i] < 1 || (i && sc->stages[i] <= sc->stages[i - 1]))
      return set_error(IACT_ERR_INVALID, "stage list must be ascending and >= 1");
  const iact_obstructions& o = sc->obstructions;
  const int counts[5] = {o.n_cylinders, o.n_boxes, o.n_spheres, o.n_oriented_boxes,
                         o.n_triangles};
  const float* ptrs[5] = {o.cylinders, o.boxes, o.spheres, o.oriented_boxes, o.triangles};
  for (int i = 0; i < 5; ++i) {
    if (counts[i] < 0) return set_error(IACT_ERR_INVALID, "negative obstruction count");
    if (counts[i] > 0 && !ptrs[i]) return set_error(IACT_ERR_INVALID, "null obstruction table");
  }
  return validate_sensor(sc->sensor);
}

// ------------------------------------------------------------- derived data
DevSensor make_dev_sensor(const iact_sensor& s) {
  DevSensor S;
  memset(&S, 0, sizeof(S));
  S.kind = s.kind;
  S.splat = s.splat;
  S.n_pixels = s.n_pixels;
  S.pos = f3{s.position[0], s.position[1], s.position[2]};
  const float* R = s.rotation;  // row-major; columns are u1, u2, n
  S.u1 = f3{R[0], R[3], R[6]};
  S.u2 = f3{R[1], R[4], R[7]};
  S.nrm = f3{R[2], R[5], R[8]};
  S.ndotp = S.nrm.x * S.pos.x + S.nrm.y * S.pos.y + S.nrm.z * S.pos.z;
  S.axis = (R[0] == 1.f && R[1] == 0.f && R[2] == 0.f && R[3] == 0.f && R[4] == 1.f &&
            R[5] == 0.f && R[6] == 0.f && R[7] == 0.f && R[8] == 1.f)
               ? 1
               : 0;
  S.width = s.width;
  S.height = s.height;
  S.x0 = s.x0;
  S.y0 = s.y0;
  S.dx = s.dx;
  S.dy = s.dy;
  S.inv_dx = s.dx > 0.f ? (float)(1.0 / (double)s.dx) : 0.f;
  S.inv_dy = s.dy > 0.f ? (float)(1.0 / (double)s.dy) : 0.f;
  S.nx0 = s.dx > 0.f ? (float)(-(double)s.x0 / (double)s.dx) : 0.f;
  S.ny0 = s.dy > 0.f ? (float)(-(double)s.y0 / (double)s.dy) : 0.f;
  S.v_scale = S.inv_dx > 0.f ? S.inv_dy / S.inv_dx : 1.f;
  S.wf = (float)s.width;
  S.hf = (float)s.height;
  S.edge_width = s.edge_width;
  // an integer flag: the per-ray edge branch is then a scalar compare
  S.edge = s.kind == IACT_SENSOR_SQUARE ? (s.edge_width > 0.f) : (s.edge_threshold < 1.f);
  S.size = s.hex_size;
  S.inv_size = s.hex_size > 0.f ? 1.f / s.hex_size : 0.f;
  S.size_s3 = (float)((double)s.hex_size * 1.7320508075688772);
  S.size_15 = (float)((double)s.hex_size * 1.5);
  S.inv_inradius = s.hex_inradius > 0.f ? 1.f / s.hex_inradius : 0.f;
  S.gcos = cosf(-s.grid_rotation);
  S.gsin = sinf(-s.grid_rotation);
  S.gox = s.grid_offset[0];
  S.goy = s.grid_offset[1];
  if (s.hex_size > 0.f) {
    // q = (x_g / sqrt3 - y_g / 3) / size, r = (2 y_g / 3) / size with
    // (x_g, y_g) = rotate(x - ox, y - oy, -grid_rotation)
    const double gc = cos(-(double)s.grid_rotation), gs = sin(-(double)s.grid_rotation);
    const double sz = s.hex_size, s33 = 0.5773502691896257;
    const double aqx = (s33 * gc - gs / 3.0) / sz, aqy = (-s33 * gs - gc / 3.0) / sz;
    const double arx = (2.0 / 3.0) * gs / sz, ary = (2.0 / 3.0) * gc / sz;
    const double ox = s.grid_offset[0], oy = s.grid_offset[1];
    S.aqx = (float)aqx;
    S.aqy = (float)aqy;
    // axial coordinates on the device are relative to the lookup table's
    // origin (q - q_min, r - r_min): rounding commutes with the integer
    // shift, and the table index needs no subtraction per ray
    S.aq0 = (float)(-(aqx * ox + aqy * oy) - (double)s.q_min);
    S.arx = (float)arx;
    S.ary = (float)ary;
    S.ar0 = (float)(-(arx * ox + ary * oy) - (double)s.r_min);
  }
  S.An = S.nrm;
  if (s.kind == IACT_SENSOR_SQUARE) {
    S.Au = S.u1;
    S.Av = f3{(float)((double)S.v_scale * R[1]), (float)((double)S.v_scale * R[4]),
              (float)((double)S.v_scale * R[7])};
  } else {
    const double a = S.aqx, b = S.aqy, c = S.arx, d = S.ary;
    S.Au = f3{(float)(a * R[0] + b * R[1]), (float)(a * R[3] + b * R[4]),
              (float)(a * R[6] + b * R[7])};
    S.Av = f3{(float)(c * R[0] + d * R[1]), (float)(c * R[3] + d * R[4]),
              (float)(c * R[6] + d * R[7])};
  }
  // bounds for plane_test_redundant (grid units): the smallest stretch of an
  // in-plane direction by the grid map, and the grid's extent
  if (s.kind == IACT_SENSOR_SQUARE) {
    S.gscale = fminf(1.f, S.v_scale);
    S.gext = (float)s.width + (float)s.height + 4.f;
  } else {
    const double a = S.aqx, b = S.aqy, c = S.arx, d = S.ary;
    const double F = a * a + b * b + c * c + d * d, D = fabs(a * d - b * c);
    const double smin2 = 0.5 * (F - sqrt(fmax(F * F - 4.0 * D * D, 0.0)));
    S.gscale = (float)(0.999 * sqrt(fmax(smin2, 0.0)));
    S.gext = (float)s.table_q + (float)s.table_r + 4.f;
  }
  S.edge_thr = s.edge_threshold;
  S.q_min = s.q_min;
  S.r_min = s.r_min;
  S.tq = s.table_q;
  S.tr = s.table_r;
  S.qmin_f = 0.f;
  S.rmin_f = 0.f;
  S.tq_f = (float)s.table_q;
  S.tr_f = (float)s.table_r;
  S.table = s.lookup_table;
  S.neg_half_inv_sigma2 = s.sigma > 0.f ? -0.5f / (s.sigma * s.sigma) : 0.f;
  S.ks = s.kernel_size;
  return S;
}

void fill_trace_args(TraceArgs& A, const iact_scene& sc, const double* sources,
                     const float* values, int ns) {
  memset(&A, 0, sizeof(A));
  const iact_facets& F = sc.facets;
  A.nf = F.n_facets;
  A.M = F.n_samples;
  A.cpf = (F.n_samples + 63) / 64;
  A.points = F.points;
  A.normals = F.normals;
  A.delta = F.delta;
  A.weights = F.weights;
  A.positions = F.positions;
  A.rotations = F.rotations;
  A.pscale = F.perturbation_scale;
  A.sources = sources;
  A.values = values;
  A.ns = ns;
  // far point sources: |e| >= 2^30 max(R, 1 mm), R the samples' radius bound
  const double rf = (double)fmaxf(F.max_sample_radius, 1e-3f) * 1073741824.0;
  A.far_e2 = F.max_sample_radius > 0.f ? (float)(rf * rf) : INFINITY;
  A.n_mirrors = sc.n_mirrors;
  A.n_stages = sc.n_stages;
  for (int i = 0; i < sc.n_stages; ++i) A.stages[i] = sc.stages[i];
  A.mirrors = sc.mirrors;
  A.n_surfaces = sc.n_surfaces;
  for (int i = 0; i < sc.n_surfaces; ++i) A.surfaces[i] = sc.surfaces[i];
  A.ob = sc.obstructions;
  A.S = make_dev_sensor(sc.sensor);
  A.max_rounds = 16;
  A.n_inst = 1;
}

// Work order of the trace kernels: chunk-block size and the multipliers of
// the item decode; `inst0` describes a plain render's single instance.
int work_order(const iact_launch* L, long long nchunks, int grid, int wpb, TraceArgs& A) {
  const int cb = L ? L->chunk_block : 0;
  const int bo = L ? L->block_order : 0;
  if (cb < -1) return set_error(IACT_ERR_INVALID, "chunk_block %d", cb);
  if (bo < 0 || bo > 1) return set_error(IACT_ERR_INVALID, "block_order %d", bo);
  if (nchunks >= (1ll << 31))
    return set_error(IACT_ERR_UNSUPPORTED, "%lld sample chunks (max 2^31)", nchunks);
  long long span = cb == 0 ? (long long)grid * wpb : (cb < 0 ? nchunks : cb);
  if (span > nchunks) span = nchunks;
  if (span < 1) span = 1;
  const long long nc = nchunks > 0 ? nchunks : 1;
  const long long n_cb = (nc + span - 1) / span;
  A.span = (uint32_t)span;
  A.nchunks = (uint32_t)nchunks;
  A.d_span = make_fastdiv((uint32_t)span);
  A.d_last = make_fastdiv((uint32_t)(nc - (n_cb - 1) * span));
  A.d_cpf = make_fastdiv((uint32_t)(A.cpf > 0 ? A.cpf : 1));
  A.inst0.nst = A.n_stiles;
  A.inst0.ns = A.ns;
  A.inst0.sbase = 0;
  A.inst0.pb = make_fastdiv((uint32_t)(span * (A.n_stiles > 0 ? A.n_stiles : 1)));
  A.xcd_order = bo == 0;
  return IACT_OK;
}

int shape_trace(const iact_scene& sc, int ns, int mode, const iact_launch* L, TraceArgs& A,
                LaunchShape& out, const DevInfo& dev, int source_type) {
  const int threads = (L && L->block_threads) ? L->block_threads : 256;
  if (threads % 64 || threads < 64 || threads > IACT_MAX_BLOCK)
    return set_error(IACT_ERR_INVALID, "block_threads %d must be a multiple of 64 in [64,%d]",
                     threads, IACT_MAX_BLOCK);
  const int wpb = threads / 64;
  int bpc = (L && L->blocks_per_cu) ? L->blocks_per_cu : 0;
  if (bpc < 0 || bpc > 64) return set_error(IACT_ERR_INVALID, "blocks_per_cu %d", bpc);
  int ts = (L && L->source_tile) ? L->source_tile : 256;
  if (ts < 1) return set_error(IACT_ERR_INVALID, "source_tile %d", ts);
  if (ts > ns) ts = ns > 0 ? ns : 1;
  // items of one instance (chunks x source tiles) are decoded in 32 bits
  const long long nch = (long long)A.nf * A.cpf;
  if (!(L && L->source_tile) && nch > 0 && ns > 0) {
    // a scene of few sample chunks (few facets): smaller source tiles (down
    // to 16 sources), so that the items give every SIMD eight waves
    const long long want = 8ll * 4 * (dev.n_cu > 0 ? dev.n_cu : 256);
    const long long tiles = (want + nch - 1) / nch;
    if (tiles > (ns + ts - 1) / ts) {
      const long long t = (ns + tiles - 1) / tiles;
      ts = (int)(t < 16 ? (ns < 16 ? ns : 16) : t);
    }
  }
  if (nch > 0 && ns > 0) {
    const long long max_tiles = ((1ll << 31) - 1) / nch;
    if (max_tiles < 1)
      return set_error(IACT_ERR_UNSUPPORTED, "%lld sample chunks (max 2^31)", nch);
    if ((ns + ts - 1) / ts > max_tiles) ts = (int)((ns + max_tiles - 1) / max_tiles);
  }
  A.source_tile = ts;
  A.n_stiles = ns > 0 ? (ns + ts - 1) / ts : 0;
  A.det = (L && L->deterministic) ? 1 : 0;
  A.acc = (L && !A.det) ? L->accumulate : 0;
  if (A.acc < 0 || A.acc > 2) return set_error(IACT_ERR_INVALID, "accumulate mode %d", A.acc);
  const int P = sc.sensor.n_pixels;
  const size_t tab = hex_table_bytes(sc.sensor);
  size_t budget = (size_t)(dev.lds_per_block > 0 ? dev.lds_per_block : 65536);
  if (budget > 65536) budget = 65536;
  if (tab > budget / 2)
    return set_error(IACT_ERR_UNSUPPORTED, "hex lookup table (%zu bytes) exceeds LDS budget", tab);
  A.lds_table = tab > 0;
  size_t img = 0;
  if (mode == IACT_MODE_IMAGE || mode == IACT_MODE_RESPONSE) {
    const size_t per = (size_t)P * sizeof(float);
    const size_t copies = (mode == IACT_MODE_IMAGE && A.det) ? wpb : 1;
    // (+ one spare float: the source-major trace's entry for pixel -1; that
    // trace is never deterministic)
    const size_t spare = A.det ? 0 : sizeof(float);
    if (per * copies + spare + tab <= budget) {
      img = per * copies + spare;
      A.lds_image = 1;
    } else if (A.det) {
      return set_error(IACT_ERR_UNSUPPORTED,
                       "deterministic mode needs %zu bytes of LDS per block (have %zu)",
                       per * copies + tab, budget);
    }
  }
  out.threads = threads;
  out.lds = img + tab;
  const long long nchunks = (long long)A.nf * A.cpf;
  A.n_items = nchunks * A.n_stiles;
  // layout of image renders
  const int lay = L ? L->layout : 0;
  if (lay < 0 || lay > 2) return set_error(IACT_ERR_INVALID, "layout %d", lay);
  A.layout = 1;
  // the source-major trace indexes samples in 31 bits
  const bool sm_fits = (long long)A.nf * A.M < (1ll << 31);
  if (mode == IACT_MODE_IMAGE && lay == 2 && (A.det || sc.n_stages > 0 || !sm_fits))
    return set_error(IACT_ERR_UNSUPPORTED,
                     "the source-major layout covers primary-only, non-deterministic renders "
                     "of fewer than 2^31 samples");
  if (mode == IACT_MODE_IMAGE && !A.det && sc.n_stages == 0 && sm_fits &&
      (lay == 2 || (lay == 0 && ns >= 64)))
    A.layout = 2;
  // obstruction culling
  A.cull = nullptr;
  int rc = cull_words_for(sc, L, A.cull_words, A.cull_all);
  if (rc) return rc;
  A.d_m = make_fastdiv((uint32_t)(A.M > 0 ? A.M : 1));
  if (bpc == 0) bpc = default_blocks_per_cu(A.det);
  const long long cap = (long long)dev.n_cu * bpc;
  long long grid;
  if (A.layout == 2) {
    const long long total = (long long)A.nf * A.M;
    long long sg = (L && L->sample_group) ? L->sample_group : 64;
    if (sg < 1) return set_error(IACT_ERR_INVALID, "sample_group %lld", sg);
    // two sources per lane from 128 sources: a sample's record, loop control
    // and facet bookkeeping then serve two rays per lane
    A.sm_spl = ns >= 128 ? 2 : 1;
    const long long cw = 64ll * A.sm_spl;
    const long long n_sc = ns > 0 ? (ns + cw - 1) / cw : 1;
    // items are decoded in 32 bits
    while (((total + sg - 1) / sg) * n_sc >= (1ll << 31)) sg *= 2;
    if (sg > total) sg = total > 0 ? total : 1;
    A.sample_group = (int)sg;
    A.n_sc = (uint32_t)n_sc;
    A.d_sc = make_fastdiv((uint32_t)n_sc);
    A.sm_items = ((total + sg - 1) / sg) * n_sc;
    grid = (A.sm_items + wpb - 1) / wpb;
    if (!(L && L->blocks_per_cu)) {
      // every block ends with a flush of its image (an fp64 atomic per lit
      // pixel) or starts by loading the image gradient: by default a block
      // gets at least 16 items per wave (measured best or within noise from
      // 256 to 16384 C3 sources, forward and adjoint), from 4 to 64 blocks
      // per CU
      const long long per = (A.sm_items + 16ll * wpb - 1) / (16ll * wpb);
      const long long lo = 4ll * dev.n_cu;
      grid = per < lo ? (grid < lo ? grid : lo) : per;
    }
  } else if (mode == IACT_MODE_RESPONSE) {
    grid = ns;
  } else {
    grid = (A.n_items + wpb - 1) / wpb;
  }
  if (grid > cap) grid = cap;
  if (grid < 1) grid = 1;
  out.grid = (int)grid;
  return work_order(L, nchunks, (int)grid, wpb, A);
}

// Default persistent grid: 64 blocks per CU for the trace kernels -- several
// rounds of blocks, which the hardware hands to CUs as earlier ones finish,
// balance the uneven cost of work items (measured faster than a grid of
// resident blocks only); deterministic renders keep 8, their workspace
// holding one partial image per block.
int default_blocks_per_cu(int det) { return det ? 8 : 64; }

int regrid(long long items, const iact_launch* L, const DevInfo& dev, LaunchShape& shp,
           TraceArgs& A) {
  const int bpc = (L && L->blocks_per_cu) ? L->blocks_per_cu : default_blocks_per_cu(0);
  const int wpb = shp.threads / 64;
  long long grid = (items + wpb - 1) / wpb;
  const long long cap = (long long)dev.n_cu * bpc;
  if (grid > cap) grid = cap;
  if (grid < 1) grid = 1;
  shp.grid = (int)grid;
  return work_order(L, (long long)A.nf * A.cpf, shp.grid, wpb, A);
}

static size_t base_workspace_bytes(const iact_scene& sc, int mode, const LaunchShape& shp,
                                   const TraceArgs& A) {
  if (mode == IACT_MODE_RESPONSE)
    return (size_t)7 * sc.facets.n_facets * sc.facets.n_samples * sizeof(float);
  if (mode == IACT_MODE_IMAGE && A.layout == 2)
    // two records of padding: the trace prefetches up to two records ahead
    return (size_t)IACT_REC_WORDS * ((size_t)sc.facets.n_facets * sc.facets.n_samples + 2) *
           sizeof(float);
  if (mode == IACT_MODE_IMAGE && A.det) return (size_t)shp.grid * sc.sensor.n_pixels * sizeof(float);
  return 0;
}

// Culling tables follow the mode's own workspace (256-B aligned pieces):
// 8 floats of facet cone and cull_words mask words per facet; with chunk
// culling the masks per (source chunk, facet); with source sorting the sort
// scratch, the permutation and the sorted sources and values.
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int cull_words_for(const iact_scene& sc, const iact_launch* L, int& words, int& all) {
  const int cm = L ? L->culling : 0;
  if (cm < -1 || cm > 1) return set_error(IACT_ERR_INVALID, "culling mode %d", cm);
  const int nob = n_obstructions(sc.obstructions);
  words = nob > 0 ? (nob + 31) / 32 : 0;
  all = !(cm == 1 || (cm == 0 && nob >= kAutoCullObstructions));
  return IACT_OK;
}

CullLayout cull_layout(int nf, int words, int all, size_t base, const CullSpec& c) {
  CullLayout L;
  memset(&L, 0, sizeof(L));
  if (words == 0) {
    L.end = base;
    return L;
  }
  size_t o = align256(base);
  L.cones = o;
  o += (size_t)nf * 8 * sizeof(float);
  L.fmasks = o;
  o = align256(o + (size_t)nf * words * sizeof(uint32_t));
  L.chunks = !all && c.n_src >= kChunkCullMinSources && c.n_src < (1ll << 31);
  if (L.chunks) {
    const size_t nch = (size_t)((c.n_src + 63) / 64);
    L.cmasks = o;
    o = align256(o + nch * nf * words * sizeof(uint32_t));
    L.ccones = o;
    o = align256(o + nch * nf * 4 * sizeof(float));
    if (c.n_groups > 0) {
      L.gbounds = o;
      o = align256(o + (size_t)c.n_groups * 4 * sizeof(float));
    }
  }
  L.sort = L.chunks && c.sort;
  if (L.sort) {
    L.sws = o;
    o = align256(o + sort_workspace_bytes(c.n_src, c.n_inst));
    L.perm = o;
    o = align256(o + (size_t)c.n_src * sizeof(int32_t));
    L.src = o;
    o = align256(o + (size_t)c.n_src * 3 * sizeof(double));
    L.val = o;
    o = align256(o + (size_t)c.n_src * sizeof(float));
  }
  L.end = o;
  return L;
}

// The fp64 image sums of a non-deterministic image render follow the
// culling tables (256-B aligned): n_images * n_pixels doubles.
size_t image64_offset(size_t end) { return align256(end); }

size_t workspace_bytes_for(const iact_scene& sc, int ns, int mode, const LaunchShape& shp,
                           const TraceArgs& A) {
  const size_t base = base_workspace_bytes(sc, mode, shp, A);
  const CullSpec c{ns, 1, nullptr, mode == IACT_MODE_IMAGE && !A.det, refine_groups(A, mode)};
  const size_t end = cull_layout(A.nf, A.cull_words, A.cull_all, base, c).end;
  if (mode == IACT_MODE_IMAGE && !A.det)
    return image64_offset(end) + (size_t)sc.sensor.n_pixels * sizeof(double);
  return end;
}

int prepare_cull(TraceArgs& A, void* workspace, size_t base, int source_type, hipStream_t st,
                 const CullSpec& c, const int32_t** perm) {
  if (perm) *perm = nullptr;
  A.cull_chunks = 0;
  if (A.cull_words == 0 || A.nf == 0 || c.n_src == 0) return IACT_OK;
  const CullLayout L = cull_layout(A.nf, A.cull_words, A.cull_all, base, c);
  char* ws = (char*)workspace;
  hipError_t e;
  if (L.sort) {
    int32_t* pm = (int32_t*)(ws + L.perm);
    double* so = (double*)(ws + L.src);
    float* vo = (float*)(ws + L.val);
    e = launch_sort_sources(A.sources, A.values, c.n_src, c.offsets, c.n_inst, source_type,
                            ws + L.sws, pm, so, vo, st);
    if (e != hipSuccess) return hip_error(e, "source sort");
    A.sources = so;
    A.values = vo;
    if (perm) *perm = pm;
  }
  float* cones = (float*)(ws + L.cones);
  uint32_t* fmasks = (uint32_t*)(ws + L.fmasks);
  // the cones cover every source of the trace (every instance of a batch)
  TraceArgs C = A;
  C.ns = (int)c.n_src;
  e = launch_cull(C, source_type, A.cull_all, cones, fmasks, st);
  if (e != hipSuccess) return hip_error(e, "k_facet_cones/k_cull_masks");
  A.cull = fmasks;
  if (L.chunks) {
    uint32_t* cm = (uint32_t*)(ws + L.cmasks);
    float* cc = (float*)(ws + L.ccones);
    e = launch_chunk_masks(C, source_type, c.n_src, cones, fmasks, cm, cc, st);
    if (e != hipSuccess) return hip_error(e, "k_chunk_masks");
    A.cull = cm;
    A.cull_chunks = 1;
    A.ccones = cc;
    // filled by the caller once the sample records exist
    if (c.n_groups > 0) A.gbounds = (const float*)(ws + L.gbounds);
  }
  return IACT_OK;
}

// The source-major adjoint (k_adj_sm) traces from the image layout's sample
// records: its workspace is the source-major image trace's (records, then
// the culling tables with the source sort and the block bounds) followed by
// the per-facet torques.
bool bwd_source_major(const iact_scene& sc, int ns, const iact_launch* L, bool want_asph) {
  if (sc.n_stages > 0 || sc.sensor.splat || want_asph) return false;
  if (L && (L->layout == 1 || L->deterministic)) return false;
  if ((long long)sc.facets.n_facets * sc.facets.n_samples >= (1ll << 31)) return false;
  return ns >= 64 || (L && L->layout == 2);
}

int adj_sm_layout(const iact_scene& sc, const double* sources, const float* values, int ns,
                  const iact_launch* L, const DevInfo& dev, TraceArgs& A, LaunchShape& shp,
                  AdjSmLayout& out) {
  fill_trace_args(A, sc, sources, values, ns);
  iact_launch l = L ? *L : iact_launch{};
  l.layout = 2;
  int rc = shape_trace(sc, ns, IACT_MODE_IMAGE, &l, A, shp, dev, -1);
  if (rc) return rc;
  out.recs_end = base_workspace_bytes(sc, IACT_MODE_IMAGE, shp, A);
  const CullSpec c{ns, 1, nullptr, true, refine_groups(A, IACT_MODE_IMAGE)};
  const size_t end = cull_layout(A.nf, A.cull_words, A.cull_all, out.recs_end, c).end;
  out.torque = align256(end);
  out.end = out.torque + (size_t)3 * sc.facets.n_facets * sizeof(double);
  return IACT_OK;
}

long long refine_groups(const TraceArgs& A, int mode) {
  if (mode != IACT_MODE_IMAGE || A.layout != 2 || !has_obstructions(A.ob)) return 0;
  return ((long long)A.nf * A.M + 63) / 64;
}

}  // namespace iact

using namespace iact;

extern "C" {

int32_t iact_abi_version(void) { return IACT_ABI_VERSION; }

int64_t iact_struct_size(int32_t which) {
  switch (which) {
    case 0: return (int64_t)sizeof(iact_surface);
    case 1: return (int64_t)sizeof(iact_facets);
    case 2: return (int64_t)sizeof(iact_obstructions);
    case 3: return (int64_t)sizeof(iact_sensor);
    case 4: return (int64_t)sizeof(iact_scene);
    case 5: return (int64_t)sizeof(iact_launch);
    case 6: return (int64_t)sizeof(iact_grads);
    case 7: return (int64_t)sizeof(iact_sample_desc);
    case 8: return (int64_t)sizeof(iact_batch);
    default: return -1;
  }
}

int32_t iact_sensor_accumulate(const iact_sensor* sensor, const float* x, const float* y,
                               const float* v, int64_t n, float* out, int32_t* pixel,
                               void* stream) {
  if (!sensor) return set_error(IACT_ERR_INVALID, "null sensor");
  int rc = validate_sensor(*sensor);
  if (rc) return rc;
  if (n < 0) return set_error(IACT_ERR_INVALID, "negative point count");
  if (!out) return set_error(IACT_ERR_INVALID, "null output");
  if (n > 0 && (!x || !y || !v)) return set_error(IACT_ERR_INVALID, "null point buffers");
  if (hex_table_bytes(*sensor) > 32768)
    return set_error(IACT_ERR_UNSUPPORTED, "hex lookup table exceeds 32 KiB of LDS");
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = launch_fill(out, sensor->n_pixels, 0.f, st);
  if (e != hipSuccess) return hip_error(e, "clear image");
  if (n == 0) return IACT_OK;
  const DevSensor S = make_dev_sensor(*sensor);
  e = launch_accumulate(S, x, y, v, n, out, pixel, st);
  return e == hipSuccess ? IACT_OK : hip_error(e, "k_accumulate");
}

const char* iact_last_error(void) { return g_last_error.c_str(); }

int32_t iact_device_info(int32_t device, int32_t* n_cu, int32_t* lds_per_block, char* arch,
                         size_t arch_len) {
  if (device < 0) return set_error(IACT_ERR_INVALID, "device %d", device);
  int cur = 0;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return hip_error(e, "hipGetDevice");
  if (cur != device) {
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_error(e, "hipSetDevice");
  }
  DevInfo d;
  const int rc = device_info(d);
  if (cur != device) (void)hipSetDevice(cur);
  if (rc) return rc;
  if (n_cu) *n_cu = d.n_cu;
  if (lds_per_block) *lds_per_block = d.lds_per_block;
  if (arch && arch_len) {
    strncpy(arch, d.arch, arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  return IACT_OK;
}

int32_t iact_sample_facets(const iact_sample_desc* D, float* points, float* normals, float* delta,
                           float* weights, void* stream) {
  if (!D) return set_error(IACT_ERR_INVALID, "null sample descriptor");
  if (D->n_facets < 0 || D->n_samples < 0)
    return set_error(IACT_ERR_INVALID, "negative facet/sample count");
  if (D->n_facets == 0 || D->n_samples == 0) return IACT_OK;
  if (!points || !normals || !delta || !weights || !D->offsets)
    return set_error(IACT_ERR_INVALID, "null sampling buffer");
  if (D->aperture == IACT_APERTURE_DISK) {
    if (!D->radii) return set_error(IACT_ERR_INVALID, "disk group without radii");
  } else if (D->aperture == IACT_APERTURE_POLYGON) {
    if (!D->vertices) return set_error(IACT_ERR_INVALID, "polygon group without vertices");
    if (D->n_vertices < 3 || D->n_vertices > IACT_MAX_POLY_VERTS)
      return set_error(IACT_ERR_UNSUPPORTED, "polygon with %d vertices (3..%d)", D->n_vertices,
                       IACT_MAX_POLY_VERTS);
  } else {
    return set_error(IACT_ERR_INVALID, "unknown aperture type %d", D->aperture);
  }
  if (D->surface.n_aspheric < 0 || D->surface.n_aspheric > IACT_MAX_ASPHERIC)
    return set_error(IACT_ERR_UNSUPPORTED, "%d aspheric terms (max %d)", D->surface.n_aspheric,
                     IACT_MAX_ASPHERIC);
  const hipError_t e = launch_sample(*D, points, normals, delta, weights, (hipStream_t)stream);
  return e == hipSuccess ? IACT_OK : hip_error(e, "k_sample_facets");
}

int32_t iact_facet_world(const iact_facets* F, float* out, void* stream) {
  if (!F) return set_error(IACT_ERR_INVALID, "null facets");
  int rc = validate_facets(*F);
  if (rc) return rc;
  if (F->n_facets == 0) return IACT_OK;
  if (!out) return set_error(IACT_ERR_INVALID, "null output");
  DevSensor S;
  memset(&S, 0, sizeof(S));
  const hipError_t e = launch_world(*F, S, out, (hipStream_t)stream);
  return e == hipSuccess ? IACT_OK : hip_error(e, "k_facet_world");
}

size_t iact_workspace_size(const iact_scene* scene, int32_t n_sources, int32_t mode,
                           const iact_launch* launch) {
  if (validate_scene(scene) || n_sources < 0) return 0;
  DevInfo dev;
  if (device_info(dev)) return 0;
  TraceArgs A;
  fill_trace_args(A, *scene, nullptr, nullptr, n_sources);
  LaunchShape shp;
  if (shape_trace(*scene, n_sources, mode, launch, A, shp, dev, -1)) return 0;
  return workspace_bytes_for(*scene, n_sources, mode, shp, A);
}

int32_t iact_trace_fwd(const iact_scene* scene, const double* sources, const float* values,
                       int32_t n_sources, int32_t source_type, int32_t mode,
                       const iact_launch* launch, void* workspace, size_t workspace_bytes,
                       float* out, float* debug_xy, int32_t* debug_pixel, void* stream) {
  int rc = validate_scene(scene);
  if (rc) return rc;
  if (n_sources < 0) return set_error(IACT_ERR_INVALID, "negative source count");
  if (n_sources > 0 && (!sources || !values))
    return set_error(IACT_ERR_INVALID, "null sources/values");
  if (source_type != IACT_SOURCE_POINT && source_type != IACT_SOURCE_PARALLEL)
    return set_error(IACT_ERR_INVALID, "unknown source type %d", source_type);
  if (mode != IACT_MODE_IMAGE && mode != IACT_MODE_RESPONSE && mode != IACT_MODE_DEBUG)
    return set_error(IACT_ERR_INVALID, "unknown mode %d", mode);
  const iact_scene& sc = *scene;
  const long long P = sc.sensor.n_pixels;
  const long long nrays = (long long)sc.facets.n_facets * n_sources * sc.facets.n_samples;
  if (mode == IACT_MODE_DEBUG && nrays > 0 && !debug_xy)
    return set_error(IACT_ERR_INVALID, "debug mode needs debug_xy");
  const bool writes = mode == IACT_MODE_IMAGE || (mode == IACT_MODE_RESPONSE && n_sources > 0) ||
                      (mode == IACT_MODE_DEBUG && nrays > 0);
  if (!out && writes) return set_error(IACT_ERR_INVALID, "null output");
  hipStream_t st = (hipStream_t)stream;
  DevInfo dev;
  rc = device_info(dev);
  if (rc) return rc;
  TraceArgs A;
  fill_trace_args(A, sc, sources, values, n_sources);
  LaunchShape shp;
  rc = shape_trace(sc, n_sources, mode, launch, A, shp, dev, source_type);
  if (rc) return rc;
  const size_t need = workspace_bytes_for(sc, n_sources, mode, shp, A);
  if (need > 0 && (!workspace || workspace_bytes < need))
    return set_error(IACT_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need,
                     workspace_bytes);
  A.out = out;
  A.dbg_xy = debug_xy;
  A.dbg_pix = debug_pixel;
  hipError_t e = hipSuccess;
  const bool empty = (sc.facets.n_facets == 0 || n_sources == 0);
  if (!empty) {
    const CullSpec c{n_sources, 1, nullptr, mode == IACT_MODE_IMAGE && !A.det,
                     refine_groups(A, mode)};
    rc = prepare_cull(A, workspace, base_workspace_bytes(sc, mode, shp, A), source_type, st, c,
                      nullptr);
    if (rc) return rc;
  }
  if (mode == IACT_MODE_IMAGE) {
    if (empty || A.det) {
      if (!(A.det && A.lds_image && !empty)) {
        e = launch_fill(out, P, 0.f, st);
        if (e != hipSuccess) return hip_error(e, "clear image");
      }
      if (empty) return IACT_OK;
    } else {
      // fp64 sums, converted into `out` after the trace
      const CullSpec c{n_sources, 1, nullptr, true, refine_groups(A, mode)};
      const size_t end = cull_layout(A.nf, A.cull_words, A.cull_all,
                                     base_workspace_bytes(sc, mode, shp, A), c).end;
      A.out64 = (double*)((char*)workspace + image64_offset(end));
      e = launch_fill_d(A.out64, P, 0.0, st);
      if (e != hipSuccess) return hip_error(e, "clear image sums");
    }
    if (A.layout == 2) {
      A.recs = (const float*)workspace;
      e = launch_records(A, (float*)workspace, st);
      if (e != hipSuccess) return hip_error(e, "k_sample_records");
      if (A.gbounds) {
        e = launch_group_bounds(A, (float*)A.gbounds, st);
        if (e != hipSuccess) return hip_error(e, "k_group_bounds");
      }
      e = launch_trace_sm(A, shp, source_type, st);
      if (e != hipSuccess) return hip_error(e, "k_trace_sm");
    } else {
      A.partials = (float*)workspace;
      e = launch_trace(A, shp, source_type, mode, st);
      if (e != hipSuccess) return hip_error(e, "k_trace");
    }
    if (A.det && A.lds_image) {
      e = launch_reduce(A.partials, shp.grid, (int)P, out, st);
      if (e != hipSuccess) return hip_error(e, "k_reduce_partials");
    } else {
      e = launch_f64_to_f32(A.out64, out, P, st);
      if (e != hipSuccess) return hip_error(e, "k_f64_to_f32");
    }
    return IACT_OK;
  }
  if (mode == IACT_MODE_RESPONSE) {
    if (n_sources > 0 && (empty || !A.lds_image)) {
      e = launch_fill(out, (long long)n_sources * P, 0.f, st);
      if (e != hipSuccess) return hip_error(e, "clear response matrix");
    }
    if (empty) return IACT_OK;
    float* world = (float*)workspace;
    e = launch_world(sc.facets, A.S, world, st);
    if (e != hipSuccess) return hip_error(e, "k_facet_world");
    A.world = world;
    e = launch_trace(A, shp, source_type, mode, st);
    return e == hipSuccess ? IACT_OK : hip_error(e, "k_trace_response");
  }
  if (empty) return IACT_OK;
  e = launch_trace(A, shp, source_type, mode, st);
  return e == hipSuccess ? IACT_OK : hip_error(e, "k_trace(debug)");
}

// ------------------------------------------------------- no-allocation entry
// Scratch of iact_trace_fwd_noalloc: a file-scope device array, one copy per
// GPU (each device's code object holds its own), kept until the process
// ends.  Host code uses only the per-device address from hipGetSymbolAddress
// (the array's name is a host-side stand-in, not a device address).
static __device__ __align__(256) unsigned char scratch[IACT_NOALLOC_SCRATCH_BYTES];

namespace iact {
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

// Device address of the current device's scratch array, looked up on the
// first call on that device (with it current) and kept.
static int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_error(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) return set_error(IACT_ERR_UNSUPPORTED, "device ordinal %d", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(IACT_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return IACT_OK;
}
}  // namespace iact

int32_t iact_trace_fwd_noalloc(const iact_scene* scene, const double* sources,
                               const float* values, int32_t n_sources, int32_t source_type,
                               int32_t mode, const iact_launch* launch, float* out,
                               float* debug_xy, int32_t* debug_pixel, void* stream) {
  int rc = validate_scene(scene);
  if (rc) return rc;
  if (n_sources < 0) return set_error(IACT_ERR_INVALID, "negative source count");
  if (mode != IACT_MODE_IMAGE && mode != IACT_MODE_RESPONSE && mode != IACT_MODE_DEBUG)
    return set_error(IACT_ERR_INVALID, "unknown mode %d", mode);
  // the workspace the render needs; an image render whose default layout
  // needs more than the scratch holds (the source-major record table, 64 B
  // per sample) takes the sample-major layout, which has none
  iact_launch L = launch ? *launch : iact_launch{};
  // (a size of 0 from an invalid launch: iact_trace_fwd reports the error)
  size_t need = iact_workspace_size(scene, n_sources, mode, &L);
  if (need > (size_t)IACT_NOALLOC_SCRATCH_BYTES && mode == IACT_MODE_IMAGE && L.layout == 0) {
    L.layout = 1;
    need = iact_workspace_size(scene, n_sources, mode, &L);
  }
  if (need > (size_t)IACT_NOALLOC_SCRATCH_BYTES)
    return set_error(IACT_ERR_UNSUPPORTED,
                     "render needs %zu bytes of scratch; iact_trace_fwd_noalloc holds %zu "
                     "(use iact_trace_fwd with a caller workspace)",
                     need, (size_t)IACT_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  rc = scratch_address(&ws);
  if (rc) return rc;
  return iact_trace_fwd(scene, sources, values, n_sources, source_type, mode, &L, ws,
                        (size_t)IACT_NOALLOC_SCRATCH_BYTES, out, debug_xy, debug_pixel, stream);
}

}  // extern "C"
