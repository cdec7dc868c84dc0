// Forward kernels: MC pupil sampling, world transform, image / debug trace
// and response-matrix trace.  Host launchers at the bottom are called by the
// C ABI in iact_abi.hip.
#include "iact_kernels.h"
#include <type_traits>

namespace iact {

// ------------------------------------------------------------------ utility
__global__ void k_fill_f32(float* p, long long n, float v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_f64(double* p, long long n, double v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// Ordered sum of per-block partial images: out[i] = sum_b partials[b][i].
__global__ void k_reduce_partials(const float* partials, int nblocks, int P, float* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += (double)partials[(size_t)b * P + i];
    out[i] = (float)s;
  }
}

// fp64 image sums -> the fp32 image (round to nearest)
__global__ void k_f64_to_f32(const double* in, float* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

// ------------------------------------------------------------------ sampling
// MCIntegrator.sample_group for one group (integrators.py:97-188) with the
// build's Philox streams: stream 0 places the sample on the aperture
// (sampling.py:8-67), stream 1 gives the two N(0,1) roughness draws
// (reflection.py:22-50, Box-Muller).
__global__ void k_sample_facets(iact_sample_desc D, float* points, float* normals, float* delta,
                                float* weights) {
  const long long total = (long long)D.n_facets * D.n_samples;
  const SurfT<float> s = load_surface<float>(D.surface);
  const uint32_t k0 = (uint32_t)D.seed, k1 = (uint32_t)(D.seed >> 32);
  for (long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x; gi < total;
       gi += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(gi / D.n_samples), m = (int)(gi % D.n_samples);
    const u4 a = philox4x32_10(u4{(uint32_t)m, (uint32_t)f, (uint32_t)D.group, 0u}, k0, k1);
    const u4 b = philox4x32_10(u4{(uint32_t)m, (uint32_t)f, (uint32_t)D.group, 1u}, k0, k1);
    float x, y, area;
    if (D.aperture == IACT_APERTURE_DISK) {
      const float rad = D.radii[f];
      const float r = __builtin_sqrtf(u01(a.x));
      const float th = u01(a.y) * 2.f * 3.14159265358979323846f;
      float sn, cs;
      sincosf(th, &sn, &cs);
      x = r * cs * rad;
      y = r * sn * rad;
      area = 3.14159265358979323846f * (rad * rad);
    } else {
      const float* v = D.vertices + (size_t)f * D.n_vertices * 2;
      const int nt = D.n_vertices - 2;
      // fan triangles from v0, area-weighted pick of the triangle
      float cum[IACT_MAX_POLY_VERTS];
      float acc = 0.f;
      for (int i = 0; i < nt; ++i) {
        const float ax = v[0], ay = v[1], bx = v[2 * (i + 1)], by = v[2 * (i + 1) + 1];
        const float cx = v[2 * (i + 2)], cy = v[2 * (i + 2) + 1];
        acc += fabsf((bx - ax) * (cy - ay) - (cx - ax) * (by - ay)) / 2.f;
        cum[i] = acc;
      }
      const float target = u01(a.x) * cum[nt - 1];
      int tri = 0;
      for (int i = 0; i < nt - 1; ++i) tri += (target >= cum[i]) ? 1 : 0;
      const float u = __builtin_sqrtf(u01(a.y)), vv = u01(a.z);
      const float A = 1.f - u, B = u * (1.f - vv), C = u * vv;
      x = A * v[0] + B * v[2 * (tri + 1)] + C * v[2 * (tri + 2)];
      y = A * v[1] + B * v[2 * (tri + 1) + 1] + C * v[2 * (tri + 2) + 1];
      float sh = 0.f;
      for (int i = 0; i < D.n_vertices; ++i) {
        const int j = (i + 1 == D.n_vertices) ? 0 : i + 1;
        sh += v[2 * i] * v[2 * j + 1] - v[2 * j] * v[2 * i + 1];
      }
      area = 0.5f * fabsf(sh);
    }
    const float x0 = D.offsets[2 * f], y0 = D.offsets[2 * f + 1];
    const float X = x + x0, Y = y + y0;
    const float z0 = sag_raw<float>(s, x0, y0);
    float z, g;
    sag_and_dr2<float>(s, X * X + Y * Y, z, g);
    const float zx = 2.f * X * g, zy = 2.f * Y * g;
    const float inv = 1.f / __builtin_sqrtf(zx * zx + zy * zy + 1.f);
    const f3 n = mk3(-zx * inv, -zy * inv, inv);
    points[3 * gi] = x;
    points[3 * gi + 1] = y;
    points[3 * gi + 2] = z - z0;
    normals[3 * gi] = n.x;
    normals[3 * gi + 1] = n.y;
    normals[3 * gi + 2] = n.z;
    // roughness basis: ref axis x if |n.z| > 0.9 else z
    const f3 ref = (fabsf(n.z) > 0.9f) ? mk3(1.f, 0.f, 0.f) : mk3(0.f, 0.f, 1.f);
    f3 t1 = cross(n, ref);
    t1 = (1.f / __builtin_sqrtf(dot(t1, t1))) * t1;
    const f3 t2 = cross(n, t1);
    const float ua = u01_open(b.x), ub = u01(b.y);
    const float rr = __builtin_sqrtf(-2.f * logf(ua));
    float sb, cb;
    sincosf(6.283185307179586f * ub, &sb, &cb);
    const float th1 = rr * cb, th2 = rr * sb;
    delta[3 * gi] = th1 * t1.x + th2 * t2.x;
    delta[3 * gi + 1] = th1 * t1.y + th2 * t2.y;
    delta[3 * gi + 2] = th1 * t1.z + th2 * t2.z;
    weights[gi] = n.z / area * (float)D.n_samples;
  }
}

// ------------------------------------------------------------------ world
__global__ void k_facet_world(iact_facets F, DevSensor S, float* out) {
  const long long total = (long long)F.n_facets * F.n_samples;
  for (long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x; gi < total;
       gi += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(gi / F.n_samples);
    const m3 R = facet_rotation(F.rotations, f);
    const f3 pos = mk3(F.positions[3 * f], F.positions[3 * f + 1], F.positions[3 * f + 2]);
    f3 p, n;
    world_sample(F.points, F.normals, F.delta, R, pos, F.perturbation_scale[f], (size_t)gi, p, n);
    out[gi] = p.x;
    out[total + gi] = p.y;
    out[2 * total + gi] = p.z;
    out[3 * total + gi] = n.x;
    out[4 * total + gi] = n.y;
    out[5 * total + gi] = n.z;
    out[6 * total + gi] = F.weights[gi];
  }
}

// ------------------------------------------------------------------ culling
// Unit direction from a facet centre c towards point source sp and the
// source's distance, for the cone bounds: the difference in fp64 (sources
// may be 1e10 m away), its length by one fp64 square root, the rest in fp32
// (the cones carry 2e-4 rad of slack, far above fp32 rounding).
IACT_DEV f3 cone_source_dir(const double* sp, f3 c, double& l) {
  const double ux = sp[0] - (double)c.x, uy = sp[1] - (double)c.y, uz = sp[2] - (double)c.z;
  l = sqrt(ux * ux + uy * uy + uz * uz);
  const float il = frcp((float)l);
  return mk3((float)ux * il, (float)uy * il, (float)uz * il);
}

// Per facet, the bundle of primary-leg shadow rays: every ray starts inside
// the ball of radius rho around the facet position (rho = the largest sample
// distance in the facet frame) and points inside the cone of half angle
// theta around axis a, the directions towards the sources (for point
// sources widened by asin(rho / distance)).  cones[8 f ..]: position,
// rho, a, tan(theta); tan < 0 disables culling for the facet.
__global__ __launch_bounds__(256) void k_facet_cones(TraceArgs A, int source_type, float* cones) {
  __shared__ float red[4][256];
  const int t = threadIdx.x;
  for (int f = blockIdx.x; f < A.nf; f += gridDim.x) {
    const f3 c = mk3(A.positions[3 * f], A.positions[3 * f + 1], A.positions[3 * f + 2]);
    float rho = 0.f;
    for (int m = t; m < A.M; m += blockDim.x) {
      const size_t gi = (size_t)f * A.M + m;
      const f3 pl = mk3(A.points[3 * gi], A.points[3 * gi + 1], A.points[3 * gi + 2]);
      rho = fmaxf(rho, sqrtf(dot(pl, pl)));
    }
    // axis: mean of the unit directions towards the sources
    f3 sum = mk3(0.f, 0.f, 0.f);
    for (int s = t; s < A.ns; s += blockDim.x) {
      f3 u;
      if (source_type == IACT_SOURCE_POINT) {
        double l;
        u = cone_source_dir(A.sources + 3 * (size_t)s, c, l);
      } else {
        u = mk3(-(float)A.sources[3 * s], -(float)A.sources[3 * s + 1],
                -(float)A.sources[3 * s + 2]);
        u = frsq(dot(u, u)) * u;
      }
      sum = sum + u;
    }
    red[0][t] = sum.x;
    red[1][t] = sum.y;
    red[2][t] = sum.z;
    red[3][t] = rho;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
      if (t < h) {
        red[0][t] += red[0][t + h];
        red[1][t] += red[1][t + h];
        red[2][t] += red[2][t + h];
        red[3][t] = fmaxf(red[3][t], red[3][t + h]);
      }
      __syncthreads();
    }
    const f3 sa = mk3(red[0][0], red[1][0], red[2][0]);
    const float rho_f = red[3][0];
    const float la = sqrtf(dot(sa, sa));
    const f3 a = la > 0.f ? (1.f / la) * sa : mk3(0.f, 0.f, 1.f);
    __syncthreads();
    // half angle, including the parallax of point sources
    float ang = 0.f;
    for (int s = t; s < A.ns; s += blockDim.x) {
      float par = 0.f;
      f3 u;
      if (source_type == IACT_SOURCE_POINT) {
        double l;
        u = cone_source_dir(A.sources + 3 * (size_t)s, c, l);
        par = l > (double)rho_f ? asinf(rho_f * frcp((float)l)) : 3.1415927f;
      } else {
        u = mk3(-(float)A.sources[3 * s], -(float)A.sources[3 * s + 1],
                -(float)A.sources[3 * s + 2]);
        u = frsq(dot(u, u)) * u;
      }
      ang = fmaxf(ang, acosf(fminf(fmaxf(dot(u, a), -1.f), 1.f)) + par);
    }
    red[0][t] = ang;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
      if (t < h) red[0][t] = fmaxf(red[0][t], red[0][t + h]);
      __syncthreads();
    }
    if (t == 0) {
      // 2e-4 rad of slack for the fp32 directions of the trace
      const float th = red[0][0] + 2e-4f;
      float* o = cones + 8 * (size_t)f;
      o[0] = c.x;
      o[1] = c.y;
      o[2] = c.z;
      o[3] = rho_f;
      o[4] = a.x;
      o[5] = a.y;
      o[6] = a.z;
      o[7] = th < 1.4f ? tanf(th) : -1.f;  // beyond ~80 degrees: no culling
    }
    __syncthreads();
  }
}

// Candidate bits: one thread per (facet, mask word).
__global__ void k_cull_masks(TraceArgs A, const float* cones, uint32_t* masks, int all) {
  const iact_obstructions& ob = A.ob;
  const int words = A.cull_words;
  const int n_ob = n_obstructions(ob);
  const long long total = (long long)A.nf * words;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(id / words), w = (int)(id - (long long)f * words);
    const float* cn = cones + 8 * (size_t)f;
    const f3 c = all ? mk3(0.f, 0.f, 0.f) : mk3(cn[0], cn[1], cn[2]);
    const float rho = all ? 0.f : cn[3], tn = all ? -1.f : cn[7];
    const f3 a = all ? mk3(0.f, 0.f, 1.f) : mk3(cn[4], cn[5], cn[6]);
    uint32_t bits = 0u;
    for (int j = 32 * w; j < min(32 * w + 32, n_ob); ++j)
      if (all || tn < 0.f || obstruction_reaches(ob, j, c, rho, a, tn))
        bits |= 1u << (j - 32 * w);
    masks[id] = bits;
  }
}

// Candidate bits per (source chunk, facet): the bundle of one facet's rays
// towards the 64 sources of one chunk (consecutive sources, which the host
// sorts into compact direction cells) is much narrower than the facet's
// bundle towards all sources.  One thread per (chunk, facet); the facet's
// own mask is the starting list, its cone gives the sample radius rho.
// masks_out[(chunk * nf + f) * words + w].
__global__ void k_chunk_masks(TraceArgs A, int source_type, long long n_sources,
                              const float* cones, const uint32_t* facet_masks,
                              uint32_t* masks_out, float* chunk_cones) {
  const iact_obstructions& ob = A.ob;
  const int words = A.cull_words;
  const long long n_chunks = (n_sources + 63) / 64;
  const long long total = n_chunks * A.nf;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const long long ch = id / A.nf;
    const int f = (int)(id - ch * A.nf);
    const float* cn = cones + 8 * (size_t)f;
    const f3 c = mk3(cn[0], cn[1], cn[2]);
    const float rho = cn[3];
    const long long s0 = ch * 64, s1 = min(s0 + 64, n_sources);
    // axis: mean unit direction towards the chunk's sources; then the half
    // angle with the parallax of point sources (as k_facet_cones)
    f3 sum = mk3(0.f, 0.f, 0.f);
    for (long long s = s0; s < s1; ++s) {
      f3 u;
      if (source_type == IACT_SOURCE_POINT) {
        double l;
        u = cone_source_dir(A.sources + 3 * (size_t)s, c, l);
      } else {
        u = mk3(-(float)A.sources[3 * s], -(float)A.sources[3 * s + 1],
                -(float)A.sources[3 * s + 2]);
        u = frsq(dot(u, u)) * u;
      }
      sum = sum + u;
    }
    const float la = sqrtf(dot(sum, sum));
    const f3 a = la > 0.f ? (1.f / la) * sum : mk3(0.f, 0.f, 1.f);
    float ang = 0.f;
    for (long long s = s0; s < s1; ++s) {
      float par = 0.f;
      f3 u;
      if (source_type == IACT_SOURCE_POINT) {
        double l;
        u = cone_source_dir(A.sources + 3 * (size_t)s, c, l);
        par = l > (double)rho ? asinf(rho * frcp((float)l)) : 3.1415927f;
      } else {
        u = mk3(-(float)A.sources[3 * s], -(float)A.sources[3 * s + 1],
                -(float)A.sources[3 * s + 2]);
        u = frsq(dot(u, u)) * u;
      }
      ang = fmaxf(ang, acosf(fminf(fmaxf(dot(u, a), -1.f), 1.f)) + par);
    }
    const float th = ang + 2e-4f;
    const float tn = th < 1.4f ? tanf(th) : -1.f;
    float* cc = chunk_cones + 4 * (size_t)id;
    cc[0] = a.x;
    cc[1] = a.y;
    cc[2] = a.z;
    cc[3] = tn;
    const uint32_t* fm = facet_masks + (size_t)f * words;
    uint32_t* out = masks_out + (size_t)id * words;
    for (int w = 0; w < words; ++w) {
      uint32_t bits = fm[w], keep = 0u;
      if (tn < 0.f) {
        keep = bits;
      } else {
        while (bits) {
          const int b = __builtin_ctz(bits);
          bits &= bits - 1u;
          if (obstruction_reaches(ob, 32 * w + b, c, rho, a, tn)) keep |= 1u << b;
        }
      }
      out[w] = keep;
    }
  }
}

// ------------------------------------------------------------------ binning
// Sensor.accumulate on given sensor-plane coordinates.
template <int KIND, int SPLAT>
__global__ __launch_bounds__(IACT_MAX_BLOCK) void k_accumulate(DevSensor S, const float* x,
                                                               const float* y, const float* v,
                                                               long long n, float* out,
                                                               int32_t* pixel) {
  extern __shared__ float smem[];
  const int32_t* tab = (const int32_t*)smem;
  if (KIND == IACT_SENSOR_HEX) {
    tab = stage_hex_table(S, (int32_t*)smem);
    __syncthreads();
  }
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long nround = (n + stride - 1) / stride * stride;  // uniform trip count
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nround; i += stride) {
    const bool active = i < n;
    const float xi = active ? x[i] : 0.f, yi = active ? y[i] : 0.f, vi = active ? v[i] : 0.f;
    if (!SPLAT) {
      const int pix = active ? hard_pixel<KIND>(S, tab, xi, yi) : -1;
      if (pixel && active) pixel[i] = pix;
      wave_accumulate<true>(out, pix, vi, 16);
    } else {
      if (pixel && active) pixel[i] = hard_pixel<KIND>(S, tab, xi, yi);
      float gu, gv;
      to_grid<KIND>(S, xi, yi, gu, gv);
      splat_taps<KIND>(S, tab, gu, gv, active,
                       [&](int pix, float w) { wave_accumulate<true>(out, pix, vi * w, 16); });
    }
  }
}

template <int KIND, int SPLAT>
static hipError_t launch_acc_t(const DevSensor& S, const float* x, const float* y,
                               const float* v, long long n, float* out, int32_t* pixel,
                               hipStream_t st) {
  const size_t lds = KIND == IACT_SENSOR_HEX ? (size_t)hex_table_words(S) * sizeof(int32_t) : 0;
  long long g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  hipLaunchKernelGGL((k_accumulate<KIND, SPLAT>), dim3((int)g), dim3(256), lds, st, S, x, y, v, n,
                     out, pixel);
  return hipGetLastError();
}

hipError_t launch_accumulate(const DevSensor& S, const float* x, const float* y, const float* v,
                             long long n, float* out, int32_t* pixel, hipStream_t st) {
  if (S.kind == IACT_SENSOR_SQUARE)
    return S.splat ? launch_acc_t<IACT_SENSOR_SQUARE, 1>(S, x, y, v, n, out, pixel, st)
                   : launch_acc_t<IACT_SENSOR_SQUARE, 0>(S, x, y, v, n, out, pixel, st);
  return S.splat ? launch_acc_t<IACT_SENSOR_HEX, 1>(S, x, y, v, n, out, pixel, st)
                 : launch_acc_t<IACT_SENSOR_HEX, 0>(S, x, y, v, n, out, pixel, st);
}

// ------------------------------------------------------------------ trace
// Image / debug mode.  Work item = (source tile, chunk of 64 samples of one
// facet); items are distributed over the waves of a persistent grid.
// LEAN: the launch accumulates into one LDS image per block with wave-merged
// adds (the batched forward and the default sample-major image); the other
// accumulation paths are compiled out, which keeps scalar registers free.
template <int SRC, int KIND, int SPLAT, int MODE, bool STAGES, bool AXIS, bool LEAN = false>
__global__ __launch_bounds__(IACT_MAX_BLOCK) void k_trace(TraceArgs A) {
  extern __shared__ float smem[];
  const bool lds_image = LEAN || A.lds_image;
  const bool det = !LEAN && A.det;
  const int acc = LEAN ? 0 : A.acc;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), wpb = block_waves();
  const int lane = lane_id();
  const int P = A.S.n_pixels;
  float* img_lds = smem;
  int32_t* tab_lds = (int32_t*)(smem + (lds_image ? (det ? wpb : 1) * P : 0));
  if (MODE == IACT_MODE_IMAGE && lds_image) {
    const int n = (det ? wpb : 1) * P;
    for (int i = threadIdx.x; i < n; i += blockDim.x) img_lds[i] = 0.f;
  }
  // hex tables are always staged in LDS
  const int32_t* table =
      KIND == IACT_SENSOR_HEX && A.lds_table ? stage_hex_table(A.S, tab_lds) : tab_lds;
  __syncthreads();
  float* img = img_lds + (det ? wave * P : 0);

  int run_pix = -1;
  float run_v = 0.f;
  const long long n_items = total_items(A);
  const long long n_bitems = (n_items + wpb - 1) / wpb;
  int cur = 0;  // instance whose image the LDS holds
  // the waves of a block take consecutive items, which belong to one
  // instance; everything down to `continue` is block-uniform
  for (long long bi = xcd_block(A.xcd_order); bi < n_bitems; bi += grid_blocks()) {
    const int inst = item_instance(A, bi * wpb);
    if (MODE == IACT_MODE_IMAGE && lds_image && inst != cur) {
      __syncthreads();
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const float v = img_lds[i];
        if (v != 0.f) unsafeAtomicAdd(A.out64 + (size_t)cur * P + i, (double)v);
        img_lds[i] = 0.f;
      }
      __syncthreads();
    }
    cur = inst;
    const long long item = bi * wpb + wave;
    // initialised: an item that is skipped then carries no values over from
    // the previous one, which would join with lane-dependent control flow
    // and make the uniform source loop look divergent to the compiler
    WorkItem w = {0, 0, 0, 0, 0, 0};
    if (item >= n_items || !decode_item(A, item, inst, lane, w)) continue;
    const int f = w.f, m = w.m;
    const bool active = m < A.M;
    const size_t gi = (size_t)f * A.M + (active ? m : A.M - 1);
    const m3 R = facet_rotation(A.rotations + w.inst * A.rot_stride, f);
    const f3 pos = mk3(A.positions[3 * f], A.positions[3 * f + 1], A.positions[3 * f + 2]);
    f3 pw, nw;
    world_sample(A.points, A.normals, A.delta, R, pos, A.pscale[f], gi, pw, nw);
    Sample smp = make_sample<KIND>(A.S, pw, nw, A.weights[gi], pos);
    const int s0 = w.s0, s1 = w.s1;
    double* const gout = A.out64 + (size_t)w.inst * P;
    const double* src = A.sources + 3 * w.sbase;
    const cptr<float> vals = as_const(A.values + w.sbase);
    // the facet's source offsets, 64 sources at a time (lane k holds source
    // s + k), read by the wave one by one; the candidate mask of the facet
    // and the current source chunk
    f3 e_lane = mk3(0.f, 0.f, 0.f);
    long long chunk = -1;
    float sv = vals[s0];
    for (int s = s0; s < s1; ++s) {
      const int k = (s - s0) & 63;
      if (k == 0) {
        const int sl = min(s + lane, s1 - 1);
        e_lane = source_offset<SRC>(src[3 * sl], src[3 * sl + 1], src[3 * sl + 2], pos);
      }
      if (((w.sbase + s) >> 6) != chunk) {
        chunk = (w.sbase + s) >> 6;
        set_facet(smp, facet_cull(A, f, chunk));
      }
      const f3 e = lane_value(e_lane, k);
      // prefetch the next value while this source is traced
      const float nv = vals[s + 1 < s1 ? s + 1 : s];
      float gu, gv, v, x, y;
      const bool ok = trace_ray<SRC, KIND, STAGES, AXIS, 2, MODE == IACT_MODE_DEBUG>(
          A, smp, active, e, source_qs(A, e), sv, gu, gv, v, x, y);
      sv = nv;
      if (MODE == IACT_MODE_DEBUG) {
        if (active) {
          const size_t idx = ((size_t)f * A.ns + s) * A.M + m;
          A.out[idx] = v;
          A.dbg_xy[2 * idx] = ok ? x : 1e10f;
          A.dbg_xy[2 * idx + 1] = ok ? y : 1e10f;
          if (A.dbg_pix) A.dbg_pix[idx] = grid_pixel<KIND>(A.S, table, gu, gv, ok);
        }
      } else if (!SPLAT) {
        const int pix = grid_pixel<KIND>(A.S, table, gu, gv, ok);
        if (lds_image && acc == 1) {
          // per-lane run of equal pixels, flushed when the pixel changes
          if (pix != run_pix) {
            if (run_pix >= 0 && run_v != 0.f) atomicAdd(img + run_pix, run_v);
            run_pix = pix;
            run_v = 0.f;
          }
          run_v += v;  // a run of pix = -1 is never added
        } else if (lds_image && acc == 2) {
          if (pix >= 0 && v != 0.f) atomicAdd(img + pix, v);
        } else if (lds_image) {
          wave_accumulate<false>(img, pix, v, A.max_rounds);
        } else {
          wave_accumulate<true>(gout, pix, v, A.max_rounds);
        }
      } else {
        splat_taps<KIND>(A.S, table, gu, gv, ok, [&](int pix, float wt) {
          if (lds_image)
            wave_accumulate<false>(img, pix, v * wt, A.max_rounds);
          else
            wave_accumulate<true>(gout, pix, v * wt, A.max_rounds);
        });
      }
    }
  }

  if (MODE == IACT_MODE_IMAGE && lds_image) {
    if (!SPLAT && acc == 1 && run_pix >= 0 && run_v != 0.f) atomicAdd(img + run_pix, run_v);
    __syncthreads();
    if (det) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        float s = 0.f;
        for (int w = 0; w < wpb; ++w) s += img_lds[w * P + i];
        A.partials[(size_t)blockIdx.x * P + i] = s;
      }
    } else {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const float v = img_lds[i];
        if (v != 0.f) unsafeAtomicAdd(A.out64 + (size_t)cur * P + i, (double)v);
      }
    }
  }
}

// World-space records of every (facet, sample) for the source-major trace,
// one block per facet.  Within a facet the records are written in the order
// of a 16 x 16 Morton grid over the samples' aperture positions (a counting
// sort in LDS): consecutive samples then lie close together on the facet,
// their rays towards one source land on the same pixel for longer, and the
// trace's per-lane runs of equal pixels are added to the image less often.
// The order inside a cell is the order of the LDS atomics (an image sum does
// not depend on it).
IACT_DEV uint32_t spread4(uint32_t v) {  // 4 bits -> even positions
  v &= 0xfu;
  v = (v | (v << 2)) & 0x33u;
  v = (v | (v << 1)) & 0x55u;
  return v;
}

// World-space records of every sample, per facet sorted by a 16 x 16 Morton
// cell of the aperture position; *flag (zeroed by the launcher) becomes
// non-zero when some sample needs the parallel-ray test.
template <int KIND>
__global__ __launch_bounds__(256) void k_sample_records(TraceArgs A, float* recs,
                                                        uint32_t* flag) {
  __shared__ uint32_t cnt[256];
  __shared__ float ext[256];
  const int t = threadIdx.x;
  for (int f = blockIdx.x; f < A.nf; f += gridDim.x) {
    const size_t g0 = (size_t)f * A.M;
    // extent of the facet's aperture positions
    float e = 0.f;
    for (int m = t; m < A.M; m += blockDim.x)
      e = fmaxf(e, fmaxf(fabsf(A.points[3 * (g0 + m)]), fabsf(A.points[3 * (g0 + m) + 1])));
    ext[t] = e;
    cnt[t] = 0u;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
      if (t < h) ext[t] = fmaxf(ext[t], ext[t + h]);
      __syncthreads();
    }
    const float sc = 8.f / fmaxf(ext[0], 1e-30f);  // 16 cells over [-ext, ext]
    auto key = [&](int m) {
      const float x = A.points[3 * (g0 + m)], y = A.points[3 * (g0 + m) + 1];
      const int cx = (int)fminf(fmaxf(floorf(fmaf(x, sc, 8.f)), 0.f), 15.f);
      const int cy = (int)fminf(fmaxf(floorf(fmaf(y, sc, 8.f)), 0.f), 15.f);
      return spread4((uint32_t)cx) | (spread4((uint32_t)cy) << 1);
    };
    for (int m = t; m < A.M; m += blockDim.x) atomicAdd(cnt + key(m), 1u);
    __syncthreads();
    if (t == 0) {  // exclusive scan of 256 counts
      uint32_t run = 0u;
      for (int i = 0; i < 256; ++i) {
        const uint32_t c = cnt[i];
        cnt[i] = run;
        run += c;
      }
    }
    __syncthreads();
    const m3 R = facet_rotation(A.rotations, f);
    const f3 pos = mk3(A.positions[3 * f], A.positions[3 * f + 1], A.positions[3 * f + 2]);
    for (int m = t; m < A.M; m += blockDim.x) {
      const uint32_t slot = atomicAdd(cnt + key(m), 1u);
      f3 pw, nw;
      world_sample(A.points, A.normals, A.delta, R, pos, A.pscale[f], g0 + m, pw, nw);
      const Sample smp = make_sample<KIND>(A.S, pw, nw, A.weights[g0 + m], pos);
      store_record(recs + IACT_REC_WORDS * (g0 + slot), smp);
      // a sample whose rays need the parallel-ray test marks the table
      if (!plane_test_redundant(A.S, smp.G)) atomicOr(flag, 1u);
    }
    __syncthreads();
  }
}

// Candidates of the SPL chunks (c SPL + k) of facet f for the rays of sample
// block `blk`: each chunk's list is spread one obstruction per lane and each
// lane re-tests its obstruction against the bundle of the block's ball and
// the chunk's cone; rbits[k] selects the lanes of rlist[k] that may block.
// False (no refinement) when a chunk has more than 64 candidates.
template <int SPL>
IACT_DEV bool refine_candidates(const TraceArgs& A, int f, int c, int blk, int lane,
                                const FacetCull* fc, int* rlist, uint64_t* rbits) {
  const cptr<float> gb = as_const(A.gbounds) + 4 * (size_t)blk;  // scalar loads
  const f3 gc = mk3(gb[0], gb[1], gb[2]);
  const float grho = gb[3];
#pragma unroll
  for (int k = 0; k < SPL; ++k) {
    int n = 0, list = 0;
    if (fc[k].any) {
      const cptr<uint32_t> m = as_const(A.cull) + fc[k].row;
      for (int w = 0; w < A.cull_words; ++w) {
        uint32_t bits = w == 0 ? fc[k].first : m[w];
        while (bits != 0u) {
          if (n == 64) return false;
          list = lane == n ? 32 * w + __builtin_ctz(bits) : list;
          bits &= bits - 1u;
          ++n;
        }
      }
    }
    const long long chunk = min(c * SPL + k, (A.ns - 1) >> 6);
    const cptr<float> cc = as_const(A.ccones) + 4 * (size_t)(chunk * A.nf + f);
    const f3 ax = mk3(cc[0], cc[1], cc[2]);
    const float tn = cc[3];
    const bool keep =
        lane < n && (tn < 0.f || obstruction_reaches(A.ob, list, gc, grho, ax, tn));
    rlist[k] = list;
    rbits[k] = __ballot(keep);
  }
  return true;
}

// Image trace, source-major layout: lane = SPL sources (64 apart), a work
// item is a group of `sample_group` consecutive samples (facet-major) traced
// against a chunk of 64 SPL sources; the sample's world-space record is
// wave-uniform (scalar loads, prefetched one sample ahead) and serves all
// SPL sources of a lane, which halves the per-ray share of the record loads,
// the loop control and the facet bookkeeping for SPL = 2.  Rays of a wave
// belong to different sources and land on different pixels; with hard
// binning each lane sums, per source, its consecutive rays that hit the same
// pixel (the samples of one facet put a source's rays on one or two pixels)
// and adds the run to the LDS image when the pixel changes; splat taps are
// added one by one.  The per-ray arithmetic is trace_ray's, as in the
// sample-major kernel.  In obstructed scenes (OBST) each facet run of an
// item takes one of two paths, uniform over the wave: runs without an
// obstruction candidate for any of the lane's chunks (most of them when
// culling) are traced without shadow tests, runs with one take the shadow
// path (refined candidates, far_run_shadow2 bits or per-ray tests).  The
// kernel is held to 80 VGPRs (6 waves per SIMD; unhinted it takes 94, 5
// waves): on the alignment fit's instances one such launch was 7.6 % faster
// than the two launches of an earlier split by path (each over every item,
// each with its own register allocation), and it is 5 % faster than at 5
// waves and 8 % faster than at 7, whose register limit moves spills into the
// per-ray loop.
template <int SRC, int KIND, int SPLAT, bool STAGES, bool AXIS, int OBST, int SPL>
__global__ __launch_bounds__(IACT_MAX_BLOCK)
__attribute__((amdgpu_waves_per_eu(OBST ? 6 : 1))) void k_trace_sm(TraceArgs A) {
  extern __shared__ float smem[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), wpb = block_waves();
  const int lane = lane_id();
  const int P = A.S.n_pixels;
  // the LDS image behind one spare entry, img[-1]: a run of pixel -1 (no
  // pixel yet, or dropped rays) is added there, without a test
  float* img = smem + 1;
  int32_t* tab_lds = (int32_t*)(smem + (A.lds_image ? P + 1 : 0));
  if (A.lds_image)
    for (int i = threadIdx.x; i < P; i += blockDim.x) img[i] = 0.f;
  const int32_t* table =
      KIND == IACT_SENSOR_HEX && A.lds_table ? stage_hex_table(A.S, tab_lds) : tab_lds;
  __syncthreads();
  // a run's add goes to the LDS image or, when it does not fit, to global
  // memory (pixel -1 and zero runs skipped) -- two separate pointers, so each
  // add is a plain LDS or global atomic (a pointer that may be either would
  // make it a flat atomic, which also stalls the record prefetch)
  auto add_run = [&](int pix, float v) {
    if (A.lds_image)
      lds_add(img + pix, v);
    else if (pix >= 0 && v != 0.f)
      unsafeAtomicAdd(A.out64 + pix, (double)v);
  };
  const long long total = (long long)A.nf * A.M;
  // no sample needs the parallel-ray test (the flag after the records), the
  // sensor has no edge test and square pixels: the lean loop serves the launch
  const bool lean = as_const((const uint32_t*)(A.recs + IACT_REC_WORDS * total))[0] == 0u &&
                    !A.S.edge && (KIND != IACT_SENSOR_SQUARE || A.S.v_scale == 1.f);
  for (long long item = xcd_block(A.xcd_order) * wpb + wave; item < A.sm_items;
       item += (long long)grid_blocks() * wpb) {
    // source chunk fastest: consecutive waves share the sample group
    const uint32_t g = fdiv((uint32_t)item, A.d_sc);
    const int c = (int)((uint32_t)item - g * A.n_sc);
    // lanes past the last source trace its rays with value 0 (their runs
    // are never added), which keeps every lane of the trace active; the
    // source's position is read again at each facet run (cached), where it
    // is used, rather than held across the item's loops
    int sl[SPL];
    float value[SPL];
#pragma unroll
    for (int k = 0; k < SPL; ++k) {
      const int s = (c * SPL + k) * 64 + lane;
      sl[k] = s < A.ns ? s : A.ns - 1;
      value[k] = s < A.ns ? A.values[sl[k]] : 0.f;
    }
    // sample indices fit in 31 bits (checked on the host): scalar compares
    const int j0 = (int)g * A.sample_group;
    const int j1 = (int)min((long long)j0 + A.sample_group, total);
    const float4* recs = (const float4*)A.recs;
    int run_pix[SPL];
    float run_v[SPL];
    // far runs (FR) sum cu nrwh over the facet (run_l) and take the facet's
    // 1/|u| (run_s) once, when the run is added or the facet ends
    float run_l[SPL], run_s[SPL];
    // the current facet's obstruction candidates (per chunk of 64 sources),
    // their refinement for the sample block, and each lane source's offset
    FacetCull fc[SPL];
    f3 e[SPL];
    float qs[SPL];
    bool refined = false;
    int rlist[SPL];
    uint64_t rbits[SPL];
    uint64_t blk[SPL];  // shadow bits of a far run (far_run_shadow)
#pragma unroll
    for (int k = 0; k < SPL; ++k) {
      run_pix[k] = -1;
      run_v[k] = 0.f;
      run_l[k] = 0.f;
      run_s[k] = 1.f;
      blk[k] = 0ull;
    }
    // one sample's rays, with shadow tests (SH) or without; the shadow
    // kernel forms the reflection from n (its records' m words stay unread:
    // its scalar registers are the scarcer resource).  GEN (generic) false:
    // no parallel-ray test, no edge test and square pixels (the launch needs
    // neither test and has such pixels).  FR: every lane's sources are far
    // (source_qs 0), u = e.
    auto trace_one = [&](auto sh, auto gen, auto far, auto sbits, const RecRaw& rec, int bi) {
      constexpr int OB = decltype(sh)::value ? OBST : 0;
      constexpr bool GEN = decltype(gen)::value;
      constexpr bool FR = decltype(far)::value;
      constexpr bool SB = decltype(sbits)::value;
      const Sample smp0 = unpack_record(rec);
#pragma unroll
      for (int k = 0; k < SPL; ++k) {
        Sample smp = smp0;
        set_facet(smp, fc[k]);
        if (OB) {
          smp.rmode = refined;
          smp.rlist = rlist[k];
          smp.rbits = rbits[k];
        }
        float gu, gv, v, x, y, vf, cu;
        // a ray's value is linear in its source's: hard binning traces unit
        // values and scales a run once, when it is added
        const bool ok = trace_ray<SRC, KIND, STAGES, AXIS, OB, false, GEN, FR>(
            A, smp, true, e[k], qs[k], SPLAT ? value[k] : 1.f, gu, gv, v, x, y, &vf, &cu);
        if (SB) {
          // the run's shadow bits (far_run_shadow) in place of the test
          const float sb = ((blk[k] >> bi) & 1ull) ? 0.f : 1.f;
          v *= sb;
          vf *= sb;
          cu *= sb;
        }
        if (!SPLAT) {
          // run of equal pixels per lane and source (consecutive samples of
          // a facet put a source's rays on the same one or two pixels), added
          // when it ends
          const int pix = grid_pixel<KIND, GEN>(A.S, table, gu, gv, ok);
          if (pix != run_pix[k]) {
            add_run(run_pix[k], value[k] * fmaf(run_s[k], run_l[k], run_v[k]));
            run_pix[k] = pix;
            run_v[k] = 0.f;
            run_l[k] = 0.f;
          }
          // a run of pix = -1 is never added; without shadow tests and
          // stages v = vf * nrwh, summed in one fma (far runs: v = cu nrwh
          // run_s)
          if (OB || STAGES)
            run_v[k] += v;
          else if (SRC == IACT_SOURCE_POINT && FR)
            run_l[k] = fmaf(cu, smp.nrwh, run_l[k]);
          else
            run_v[k] = fmaf(vf, smp.nrwh, run_v[k]);
        } else {
          splat_taps<KIND>(A.S, table, gu, gv, ok, [&](int pix, float w) {
            const float vw = v * w;
            if (pix >= 0 && vw != 0.f) add_run(pix, vw);
          });
        }
      }
    };
    for (int j = j0; j < j1;) {
      // the run of the group's samples on one facet
      const int f = (int)fdiv((uint32_t)j, A.d_m);
      const int jend = min((f + 1) * A.M, j1);
      int any = 0;
#pragma unroll
      for (int k = 0; k < SPL; ++k) {
        // (a chunk past the last source has value 0 everywhere: it reads
        // the last chunk's row, which exists)
        fc[k] = OBST ? facet_cull(A, f, min(c * SPL + k, (A.ns - 1) >> 6)) : FacetCull{f, 0u, 0, 0};
        any |= fc[k].any;
      }
      const bool shadow = OBST && any;
      // the shadow path re-tests the candidates of a run inside one block
      // of 64 samples against the block's bundle (the block's ball, the
      // chunk's cone), one candidate per lane; lists of more than 64 are not
      // refined
      refined = false;
      if (shadow && A.gbounds && (j >> 6) == ((jend - 1) >> 6))
        refined = refine_candidates<SPL>(A, f, c, j >> 6, lane, fc, rlist, rbits);
      const cptr<float> fp = as_const(A.positions) + 3 * f;  // scalar loads
      const f3 pos = mk3(fp[0], fp[1], fp[2]);
      int near = 0;
#pragma unroll
      for (int k = 0; k < SPL; ++k) {
        const double* sp = A.sources + 3 * (size_t)sl[k];
        e[k] = source_offset<SRC>(sp[0], sp[1], sp[2], pos);
        qs[k] = source_qs(A, e[k]);
        near |= qs[k] != 0.f;
        run_s[k] = frsq(dot(e[k], e[k]));  // trace_ray's 1/|u| of a far ray
      }
      const bool all_far = SRC == IACT_SOURCE_POINT && __ballot(near) == 0ull;
      // the next record is loaded while one is traced (the record table has
      // padding at its end, so the prefetch needs no clamping)
      const float4* rp = recs + 4 * (size_t)j;
      const int jb = j;
      RecRaw cur = fetch_record(rp);
      auto run = [&](auto sh, auto gen, auto far, auto sbits) {
        for (; j < jend; ++j) {
          rp += 4;
          const RecRaw nxt = fetch_record(rp);
          trace_one(sh, gen, far, sbits, cur, j - jb);
          cur = nxt;
        }
        // the facet's far sums join the run (which may go on into the next
        // facet)
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
          run_v[k] = fmaf(run_s[k], run_l[k], run_v[k]);
          run_l[k] = 0.f;
        }
      };
      using T_ = std::true_type;
      using F_ = std::false_type;
      if (shadow) {
        // (the runs with per-ray shadow tests have no lean copy: their tests
        // dominate; the shadow-bit runs do, since they trace like the
        // shadow-free runs)
        if (all_far && refined) {
          // far sources over a refined run: the candidates' shadow bits
          // first, then the run without shadow tests
          f3 dsh[SPL];
#pragma unroll
          for (int k = 0; k < SPL; ++k) {
            const f3 u = e[k];
            dsh[k] = -(frsq(dot(u, u)) * u);  // trace_ray's -d of a far ray
          }
          const float* rec0 = (const float*)(recs + 4 * (size_t)jb);
          const cptr<float> gb = as_const(A.gbounds) + 4 * (size_t)(jb >> 6);  // the run's ball
          const f3 gc = mk3(gb[0], gb[1], gb[2]);
          if constexpr (SPL == 2) {
            far_run_shadow2(A, rec0, jend - jb, rlist, rbits, dsh, gc, gb[3], blk);
          } else {
#pragma unroll
            for (int k = 0; k < SPL; ++k)
              blk[k] = far_run_shadow(A, rec0, jend - jb, rlist[k], rbits[k], dsh[k], gc, gb[3]);
          }
          if (lean)
            run(F_{}, F_{}, T_{}, T_{});
          else
            run(F_{}, T_{}, T_{}, T_{});
        } else if (all_far) {
          run(T_{}, T_{}, T_{}, F_{});
        } else {
          run(T_{}, T_{}, F_{}, F_{});
        }
      } else if (lean) {
        if (all_far)
          run(F_{}, F_{}, T_{}, F_{});
        else
          run(F_{}, F_{}, F_{}, F_{});
      } else {
        run(F_{}, T_{}, F_{}, F_{});
      }
      j = jend;  // (run has traced up to the run's end)
    }
    if (!SPLAT) {
#pragma unroll
      for (int k = 0; k < SPL; ++k) {
        add_run(run_pix[k], value[k] * fmaf(run_s[k], run_l[k], run_v[k]));
      }
    }
  }
  if (A.lds_image) {
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
      const float v = img[i];
      if (v != 0.f) unsafeAtomicAdd(A.out64 + i, (double)v);
    }
  }
}

// Adjoint of the source-major image trace for primary-only, hard-binned
// renders (the reference gets it from jax.grad through render.py:174-220).
// Lanes, items, records, the two paths of obstructed scenes and the per-ray
// arithmetic up to the pixel are k_trace_sm's, so every ray lands on the
// pixel the forward gave it; the pixel's image gradient G is read from LDS.
// With hard binning only the ray value v = value (-d.n) / w * shadow
// (render.py:141) carries a derivative: dc = dL/d(d.n) = G value shadow nrw,
// dL/dn = dc d and, for point sources (d = u / |u|, u = p - src),
// dL/dp = dc / |u| (n - (d.n) d).  Per sample each lane sums over its SPL
// sources
//   An = sum dc d,  Tn = sum dc / |u|,  Wd = sum dc / |u| (d.n) d,
// so that the lane's share of dL/dp is Tn n - Wd.  A rotation moves p - pos
// and n by w x (.) with w the angle's rotation axis (mirrors.py:64-79:
// n = R n_l / |n_l|, so the normalisation drops out), hence
// dL/dangle = w . sum_m (q x dL/dp + n x dL/dn) over the facet's samples.
// That torque is linear in the per-lane sums, so each lane keeps its own share
// over a facet run and the wave reduces it (and dL/dpos) once per run; the
// fp64 per-facet sums go to k_torque_to_euler, which dots them with the axes.
// Source and value gradients are summed per lane (per source) over an item's
// samples.
// Held to 80 VGPRs (6 waves per SIMD; unhinted the rotation-only kernel
// takes 92, the other 104): on the alignment fit's instances 14 % faster than
// the two launches of an earlier split by path at their natural allocations,
// and 3 % faster than the rotation-only kernel at 5 waves without spills.
// GX false: the launch wants rotation gradients only (no value, source or
// position gradients), and the kernel has no code or registers for the rest.
template <int SRC, int KIND, bool AXIS, int OBST, int SPL, bool GX>
__global__ __launch_bounds__(IACT_MAX_BLOCK)
__attribute__((amdgpu_waves_per_eu(6))) void k_adj_sm(AdjArgs B) {
  const TraceArgs& A = B.T;
  extern __shared__ float smem[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), wpb = block_waves();
  const int lane = lane_id();
  const int P = A.S.n_pixels;
  // the image gradient behind one zero entry: a dropped ray (pixel -1)
  // reads G[-1] = 0
  float* G = smem + 1;
  int32_t* tab_lds = (int32_t*)(smem + 1 + P);
  if (threadIdx.x == 0) smem[0] = 0.f;
  for (int i = threadIdx.x; i < P; i += blockDim.x) G[i] = B.grad_image[i];
  const int32_t* table = KIND == IACT_SENSOR_HEX ? stage_hex_table(A.S, tab_lds) : tab_lds;
  __syncthreads();
  const long long total = (long long)A.nf * A.M;
  const bool lean = as_const((const uint32_t*)(A.recs + IACT_REC_WORDS * total))[0] == 0u &&
                    !A.S.edge && (KIND != IACT_SENSOR_SQUARE || A.S.v_scale == 1.f);
  const bool want_src = GX && B.g_src != nullptr, want_val = GX && B.g_val != nullptr;
  const bool want_pos = GX && B.g_pos != nullptr;
  for (long long item = xcd_block(A.xcd_order) * wpb + wave; item < A.sm_items;
       item += (long long)grid_blocks() * wpb) {
    const uint32_t g = fdiv((uint32_t)item, A.d_sc);
    const int c = (int)((uint32_t)item - g * A.n_sc);
    int sl[SPL];  // the source, read where it is used (k_trace_sm)
    float value[SPL], gval[SPL];
    f3 gsrc[SPL];
#pragma unroll
    for (int k = 0; k < SPL; ++k) {
      const int s = (c * SPL + k) * 64 + lane;
      sl[k] = s < A.ns ? s : A.ns - 1;
      value[k] = s < A.ns ? A.values[sl[k]] : 0.f;
      gval[k] = 0.f;
      gsrc[k] = mk3(0.f, 0.f, 0.f);
    }
    const int j0 = (int)g * A.sample_group;
    const int j1 = (int)min((long long)j0 + A.sample_group, total);
    const float4* recs = (const float4*)A.recs;
    FacetCull fc[SPL];
    f3 e[SPL];
    float qs[SPL];
    bool refined = false;
    int rlist[SPL];
    uint64_t rbits[SPL];
    // far sources (source_qs 0 on every lane): each source's u = e, d and
    // 1/|u| are fixed over a facet run, and the lane sums per source
    //   N = sum b n,  C = sum b c2     (b = dc, c2 = -2 d.n)
    // over the run, from which its dL/dp is inv (N + C d / 2) and its torque
    // share N x d.  The torque's other terms, inv (sum b q x n + sum b c2 q x
    // d / 2), are below 2^-30 * 1.5 * sum |b| (|q| inv <= 2^-30, |c2| <= 2):
    // some 40 times below the fp32 rounding of the sum N x d that they join
    // (also per ray, in the reference's fp32 gradient), and are left out.
    // With b = a nrw = -2 a nrwh and n = -m / 2 the lane sums a nrwh m, which
    // is b n exactly (scalings by powers of two), from the record's m and
    // nrwh words, and fC holds -C / 2; both leave out the source's value
    // (a = G shadow value), which multiplies them at the run's end.
    f3 fd[SPL], fN[SPL];
    float finv[SPL], fC[SPL];
    uint64_t blk[SPL];  // shadow bits of a far run (far_run_shadow)
    const bool want_pc = want_pos || want_src;
    // one sample's rays (SPL per lane): the forward's trace up to the pixel,
    // then the lane's share of the sample's torque and dL/dp (tq, gp), or
    // (FR) the far sums.
    auto adjoint_one = [&](auto sh, auto gen, auto far, auto sbits, const RecRaw& rec, int bi,
                           f3& tq, f3& gp) {
      constexpr int OB = decltype(sh)::value ? OBST : 0;
      constexpr bool GEN = decltype(gen)::value;
      constexpr bool FR = decltype(far)::value;
      constexpr bool SB = decltype(sbits)::value;
      const Sample smp0 = unpack_record(rec);
      float san_x = 0.f, san_y = 0.f, san_z = 0.f, st = 0.f;
      float sw_x = 0.f, sw_y = 0.f, sw_z = 0.f;
#pragma unroll
      for (int k = 0; k < SPL; ++k) {
        Sample smp = smp0;
        set_facet(smp, fc[k]);
        if (OB) {
          smp.rmode = refined;
          smp.rlist = rlist[k];
          smp.rbits = rbits[k];
        }
        // the forward's arithmetic (trace_ray): the same pixel
        float inv, c2u, rs;
        f3 u, d;
        if (SRC == IACT_SOURCE_POINT) {
          if (FR) {
            u = e[k];
            inv = finv[k];
            d = fd[k];
          } else {
            u = fma3(qs[k], smp.q, e[k]);
            inv = frsq(dot(u, u));
            d = inv * u;
          }
          rs = inv;
        } else {
          d = primary_direction<SRC>(smp, e[k], 1.f, inv);
          rs = 1.f;
          u = d;
        }
        c2u = dot(u, smp.m);
        const float shadow = SB ? (((blk[k] >> bi) & 1ull) ? 0.f : 1.f)
                                : primary_shadow<OB>(A, smp, d, true);
        float gu, gv;
        const bool ok = grid_hit<GEN>(smp.G, smp.N, sensor_dirs(A.S, u), c2u, gu, gv, rs);
        const int pix = grid_pixel<KIND, GEN>(A.S, table, gu, gv, ok);
        const float gp_pix = G[pix];  // 0 for pix = -1
        // a = G value shadow; dc = a nrw is applied per sample
        const float gs = (OB || SB) ? gp_pix * shadow : gp_pix;
        const float a = gs * value[k];
        if (want_val) {
          // dL/dvalue = G v / value = G shadow (d.m) nrwh
          gval[k] = fmaf(gs * (c2u * inv), smp.nrwh, gval[k]);
        }
        if (SRC == IACT_SOURCE_POINT && FR) {
          // b n = bm m; the source's value scales the run's sums once
          const float bm = gs * smp.nrwh;
          fN[k] = fma3(bm, smp.m, fN[k]);
          if (want_pc) fC[k] = fmaf(bm, c2u * inv, fC[k]);
        } else if (SRC == IACT_SOURCE_POINT) {
          const float k1 = a * inv;  // dc / |u| before nrw
          san_x = fmaf(k1, u.x, san_x);
          san_y = fmaf(k1, u.y, san_y);
          san_z = fmaf(k1, u.z, san_z);
          st += k1;
          // (d.n) d = -c2u / 2 inv^2 u: the factor -1/2 is applied per sample
          const float k2 = k1 * c2u * (inv * inv);
          sw_x = fmaf(k2, u.x, sw_x);
          sw_y = fmaf(k2, u.y, sw_y);
          sw_z = fmaf(k2, u.z, sw_z);
          if (want_src) {
            // dL/dsrc = -dL/du = -(dc / |u|) (n - (d.n) d)
            const float tsrc = k1 * smp.nrw;
            const float cdn = -0.5f * c2u * inv;
            const f3 gu3 = tsrc * fma3(-cdn, d, smp.n);
            gsrc[k] = gsrc[k] - gu3;
          }
        } else {
          san_x = fmaf(a, d.x, san_x);
          san_y = fmaf(a, d.y, san_y);
          san_z = fmaf(a, d.z, san_z);
          if (want_src) {
            // d = the source vector as given: dL/dsrc = dc n
            gsrc[k] = fma3(a * smp.nrw, smp.n, gsrc[k]);
          }
        }
      }
      if (FR) return;
      // the lane's share of the sample's torque: linear in the lane's sums,
      // so the wave reduces only once per facet run
      const float nrw = smp0.nrw;
      const f3 An = nrw * mk3(san_x, san_y, san_z);
      if (SRC == IACT_SOURCE_POINT) {
        const float Tn = nrw * st;
        const f3 Wd = (-0.5f * nrw) * mk3(sw_x, sw_y, sw_z);
        gp = fma3(Tn, smp0.n, mk3(-Wd.x, -Wd.y, -Wd.z));
        tq = cross(smp0.q, gp) + cross(smp0.n, An);
      } else {
        gp = mk3(0.f, 0.f, 0.f);
        tq = cross(smp0.n, An);
      }
    };
    for (int j = j0; j < j1;) {
      const int f = (int)fdiv((uint32_t)j, A.d_m);
      const int jend = min((f + 1) * A.M, j1);
      int any = 0;
#pragma unroll
      for (int k = 0; k < SPL; ++k) {
        fc[k] = OBST ? facet_cull(A, f, min(c * SPL + k, (A.ns - 1) >> 6)) : FacetCull{f, 0u, 0, 0};
        any |= fc[k].any;
      }
      const bool shadow = OBST && any;
      refined = false;
      if (shadow && A.gbounds && (j >> 6) == ((jend - 1) >> 6))
        refined = refine_candidates<SPL>(A, f, c, j >> 6, lane, fc, rlist, rbits);
      const cptr<float> fp = as_const(A.positions) + 3 * f;
      const f3 pos = mk3(fp[0], fp[1], fp[2]);
      int near = 0;
#pragma unroll
      for (int k = 0; k < SPL; ++k) {
        const double* sp = A.sources + 3 * (size_t)sl[k];
        e[k] = source_offset<SRC>(sp[0], sp[1], sp[2], pos);
        qs[k] = source_qs(A, e[k]);
        near |= qs[k] != 0.f;
        // u = e: trace_ray's 1/|u| and d of every far sample
        finv[k] = frsq(dot(e[k], e[k]));
        fd[k] = finv[k] * e[k];
        fN[k] = mk3(0.f, 0.f, 0.f);
        fC[k] = 0.f;
      }
      const bool all_far = SRC == IACT_SOURCE_POINT && __ballot(near) == 0ull;
      const float4* rp = recs + 4 * (size_t)j;
      RecRaw cur = fetch_record(rp);
      f3 T = mk3(0.f, 0.f, 0.f), Gp = mk3(0.f, 0.f, 0.f);
      const int jb = j;
      auto run = [&](auto sh, auto gen, auto far, auto sbits) {
        for (; j < jend; ++j) {
          rp += 4;
          const RecRaw nxt = fetch_record(rp);
          f3 tq, gp;
          adjoint_one(sh, gen, far, sbits, cur, j - jb, tq, gp);
          if (!decltype(far)::value) {
            T = T + tq;
            Gp = Gp + gp;
          }
          cur = nxt;
        }
      };
      using T_ = std::true_type;
      using F_ = std::false_type;
      if (shadow) {
        if (all_far && refined) {
          // the forward's shadow bits of the run (far_run_shadow), then the
          // run without shadow tests
          f3 dsh[SPL];
#pragma unroll
          for (int k = 0; k < SPL; ++k) dsh[k] = -fd[k];
          const float* rec0 = (const float*)(recs + 4 * (size_t)jb);
          const cptr<float> gb = as_const(A.gbounds) + 4 * (size_t)(jb >> 6);  // the run's ball
          const f3 gc = mk3(gb[0], gb[1], gb[2]);
          if constexpr (SPL == 2) {
            far_run_shadow2(A, rec0, jend - jb, rlist, rbits, dsh, gc, gb[3], blk);
          } else {
#pragma unroll
            for (int k = 0; k < SPL; ++k)
              blk[k] = far_run_shadow(A, rec0, jend - jb, rlist[k], rbits[k], dsh[k], gc, gb[3]);
          }
          if (lean)
            run(F_{}, F_{}, T_{}, T_{});
          else
            run(F_{}, T_{}, T_{}, T_{});
        } else if (all_far) {
          run(T_{}, T_{}, T_{}, F_{});
        } else {
          run(T_{}, T_{}, F_{}, F_{});
        }
      } else if (lean) {
        if (all_far)
          run(F_{}, F_{}, T_{}, F_{});
        else
          run(F_{}, F_{}, F_{}, F_{});
      } else {
        run(F_{}, T_{}, F_{}, F_{});
      }
      j = jend;  // (run has traced up to the run's end)
      if (all_far) {
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
          const f3 Nk = value[k] * fN[k];
          T = T + cross(Nk, fd[k]);
          if (want_pc) {
            const f3 gpk = finv[k] * fma3(-value[k] * fC[k], fd[k], Nk);  // fC = -C / 2
            Gp = Gp + gpk;
            if (want_src) gsrc[k] = gsrc[k] - gpk;
          }
        }
      }
      T = mk3(wave_sum(T.x), wave_sum(T.y), wave_sum(T.z));
      if (want_pos) Gp = mk3(wave_sum(Gp.x), wave_sum(Gp.y), wave_sum(Gp.z));
      if (lane == 0) {
        unsafeAtomicAdd(B.torque + 3 * f, (double)T.x);
        unsafeAtomicAdd(B.torque + 3 * f + 1, (double)T.y);
        unsafeAtomicAdd(B.torque + 3 * f + 2, (double)T.z);
        if (want_pos) {
          unsafeAtomicAdd(B.g_pos + 3 * f, (double)Gp.x);
          unsafeAtomicAdd(B.g_pos + 3 * f + 1, (double)Gp.y);
          unsafeAtomicAdd(B.g_pos + 3 * f + 2, (double)Gp.z);
        }
      }
    }
    // the lanes' sources: gradients to the caller's source order
#pragma unroll
    for (int k = 0; k < SPL; ++k) {
      const int s = (c * SPL + k) * 64 + lane;
      if (s < A.ns && (want_val || want_src)) {
        const long long so = B.src_perm ? (long long)as_const(B.src_perm)[s] : (long long)s;
        if (want_val && gval[k] != 0.f) unsafeAtomicAdd(B.g_val + so, (double)gval[k]);
        if (want_src) {
          unsafeAtomicAdd(B.g_src + 3 * so, (double)gsrc[k].x);
          unsafeAtomicAdd(B.g_src + 3 * so + 1, (double)gsrc[k].y);
          unsafeAtomicAdd(B.g_src + 3 * so + 2, (double)gsrc[k].z);
        }
      }
    }
  }
}

// dL/d(Euler degrees) of each facet from its torque: R = Rz(c) Ry(b) Rx(a),
// dR/da R^T = [Rz Ry e_x]x, dR/db R^T = [Rz e_y]x, dR/dc R^T = [e_z]x.
__global__ void k_torque_to_euler(const double* torque, const float* rotations, int nf,
                                  double* g_rot) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double k = 3.14159265358979323846 / 180.0;
  const double b = (double)rotations[3 * f + 1] * k, cz = (double)rotations[3 * f + 2] * k;
  const double sb = sin(b), cb = cos(b), sc = sin(cz), cc = cos(cz);
  const double tx = torque[3 * f], ty = torque[3 * f + 1], tz = torque[3 * f + 2];
  g_rot[3 * f] = k * (cc * cb * tx + sc * cb * ty - sb * tz);
  g_rot[3 * f + 1] = k * (-sc * tx + cc * ty);
  g_rot[3 * f + 2] = k * tz;
}

// Response matrix: each block owns one source at a time, accumulates its
// image row in LDS over every (facet, sample), then writes the full row
// (render.py:271-324).  World samples come from the precomputed SoA.
template <int SRC, int KIND, int SPLAT>
__global__ __launch_bounds__(IACT_MAX_BLOCK) void k_trace_response(TraceArgs A) {
  extern __shared__ float smem[];
  const int P = A.S.n_pixels;
  float* row = smem;
  int32_t* tab_lds = (int32_t*)(smem + (A.lds_image ? P : 0));
  const int32_t* table =
      KIND == IACT_SENSOR_HEX && A.lds_table ? stage_hex_table(A.S, tab_lds) : tab_lds;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), wpb = block_waves();
  const int lane = lane_id();
  const long long total = (long long)A.nf * A.M;
  const long long nchunks = (long long)A.nf * A.cpf;
  for (int s = blockIdx.x; s < A.ns; s += grid_blocks()) {
    if (A.lds_image)
      for (int i = threadIdx.x; i < P; i += blockDim.x) row[i] = 0.f;
    __syncthreads();
    float* gdst = A.out + (size_t)s * P;
    const double sx = A.sources[3 * s], sy = A.sources[3 * s + 1], sz = A.sources[3 * s + 2];
    const float val = A.values[s];
    int fe = -1;
    f3 e = mk3(0.f, 0.f, 0.f);
    float qs = 1.f;
    for (long long c = wave; c < nchunks; c += wpb) {
      const int f = (int)fdiv((uint32_t)c, A.d_cpf);
      const int m = ((int)c - f * A.cpf) * 64 + lane;
      const bool active = m < A.M;
      const size_t gi = (size_t)f * A.M + (active ? m : A.M - 1);
      const f3 p = mk3(A.world[gi], A.world[total + gi], A.world[2 * total + gi]);
      const f3 n = mk3(A.world[3 * total + gi], A.world[4 * total + gi], A.world[5 * total + gi]);
      const f3 pos = mk3(A.positions[3 * f], A.positions[3 * f + 1], A.positions[3 * f + 2]);
      if (f != fe) {
        e = source_offset<SRC>(sx, sy, sz, pos);
        qs = source_qs(A, e);
        fe = f;
      }
      Sample smp = make_sample<KIND>(A.S, p, n, A.world[6 * total + gi], pos);
      set_facet(smp, facet_cull(A, f, s >> 6));
      float gu, gv, v, x, y;
      const bool ok =
          trace_ray<SRC, KIND, true, false>(A, smp, active, e, qs, val, gu, gv, v, x, y);
      if (!SPLAT) {
        const int pix = grid_pixel<KIND>(A.S, table, gu, gv, ok);
        if (A.lds_image)
          wave_accumulate<false>(row, pix, v, A.max_rounds);
        else
          wave_accumulate<true>(gdst, pix, v, A.max_rounds);
      } else {
        splat_taps<KIND>(A.S, table, gu, gv, ok, [&](int pix, float w) {
          if (A.lds_image)
            wave_accumulate<false>(row, pix, v * w, A.max_rounds);
          else
            wave_accumulate<true>(gdst, pix, v * w, A.max_rounds);
        });
      }
    }
    __syncthreads();
    if (A.lds_image)
      for (int i = threadIdx.x; i < P; i += blockDim.x) A.out[(size_t)s * P + i] = row[i];
    __syncthreads();
  }
}

// ------------------------------------------------------------------ launchers
static int grid_for(long long n, int threads, int cap) {
  long long g = (n + threads - 1) / threads;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_fill(float* p, long long n, float v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_d(double* p, long long n, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_sample(const iact_sample_desc& D, float* p, float* n, float* d, float* w,
                         hipStream_t st) {
  const long long total = (long long)D.n_facets * D.n_samples;
  hipLaunchKernelGGL(k_sample_facets, dim3(grid_for(total, 256, 16384)), dim3(256), 0, st, D, p,
                     n, d, w);
  return hipGetLastError();
}

hipError_t launch_world(const iact_facets& F, const DevSensor& S, float* out, hipStream_t st) {
  const long long total = (long long)F.n_facets * F.n_samples;
  hipLaunchKernelGGL(k_facet_world, dim3(grid_for(total, 256, 16384)), dim3(256), 0, st, F, S,
                     out);
  return hipGetLastError();
}

// ---- kernel selection.  Every trace launch and every occupancy query goes
// through pick_trace_kernel, so the grid is sized for the kernel that runs.
template <int SRC, int KIND, int SPLAT, int MODE>
static TraceKernel pick_trace_t(const TraceArgs& A) {
  // image mode is specialised for primary-only scenes and identity sensor
  // frames; both specialisations compute bit-identical results to the
  // generic kernel, they only skip instructions
  if (MODE == IACT_MODE_IMAGE && A.n_stages == 0) {
    if (A.lds_image && A.acc == 0 && !A.det)
      return A.S.axis ? k_trace<SRC, KIND, SPLAT, MODE, false, true, true>
                      : k_trace<SRC, KIND, SPLAT, MODE, false, false, true>;
    return A.S.axis ? k_trace<SRC, KIND, SPLAT, MODE, false, true>
                    : k_trace<SRC, KIND, SPLAT, MODE, false, false>;
  }
  return k_trace<SRC, KIND, SPLAT, MODE, true, false>;
}

// source-major kernels exist for primary-only scenes (the host selects that
// layout only for those); obstructed scenes take the kernel with the shadow
// path
template <int SRC, int KIND, int SPLAT, int SPL>
static TraceKernel pick_sm_spl(const TraceArgs& A) {
  if (has_obstructions(A.ob))
    return A.S.axis ? k_trace_sm<SRC, KIND, SPLAT, false, true, 1, SPL>
                    : k_trace_sm<SRC, KIND, SPLAT, false, false, 1, SPL>;
  return A.S.axis ? k_trace_sm<SRC, KIND, SPLAT, false, true, 0, SPL>
                  : k_trace_sm<SRC, KIND, SPLAT, false, false, 0, SPL>;
}
template <int SRC, int KIND, int SPLAT>
static TraceKernel pick_sm_t(const TraceArgs& A) {
  return A.sm_spl == 2 ? pick_sm_spl<SRC, KIND, SPLAT, 2>(A) : pick_sm_spl<SRC, KIND, SPLAT, 1>(A);
}

template <int SRC>
static TraceKernel pick_src(const TraceArgs& A, int which) {
  const bool sq = A.S.kind == IACT_SENSOR_SQUARE;
  const bool sp = A.S.splat != 0;
  switch (which) {
    case TRACE_RESPONSE:
      if (sq) return sp ? k_trace_response<SRC, IACT_SENSOR_SQUARE, 1>
                        : k_trace_response<SRC, IACT_SENSOR_SQUARE, 0>;
      return sp ? k_trace_response<SRC, IACT_SENSOR_HEX, 1>
                : k_trace_response<SRC, IACT_SENSOR_HEX, 0>;
    case TRACE_DEBUG:
      return sq ? pick_trace_t<SRC, IACT_SENSOR_SQUARE, 0, IACT_MODE_DEBUG>(A)
                : pick_trace_t<SRC, IACT_SENSOR_HEX, 0, IACT_MODE_DEBUG>(A);
    case TRACE_SOURCE_MAJOR:
      if (sq) return sp ? pick_sm_t<SRC, IACT_SENSOR_SQUARE, 1>(A)
                        : pick_sm_t<SRC, IACT_SENSOR_SQUARE, 0>(A);
      return sp ? pick_sm_t<SRC, IACT_SENSOR_HEX, 1>(A) : pick_sm_t<SRC, IACT_SENSOR_HEX, 0>(A);
    default:
      if (sq) return sp ? pick_trace_t<SRC, IACT_SENSOR_SQUARE, 1, IACT_MODE_IMAGE>(A)
                        : pick_trace_t<SRC, IACT_SENSOR_SQUARE, 0, IACT_MODE_IMAGE>(A);
      return sp ? pick_trace_t<SRC, IACT_SENSOR_HEX, 1, IACT_MODE_IMAGE>(A)
                : pick_trace_t<SRC, IACT_SENSOR_HEX, 0, IACT_MODE_IMAGE>(A);
  }
}

TraceKernel pick_trace_kernel(const TraceArgs& A, int source_type, int which) {
  return source_type == IACT_SOURCE_POINT ? pick_src<IACT_SOURCE_POINT>(A, which)
                                          : pick_src<IACT_SOURCE_PARALLEL>(A, which);
}

hipError_t launch_trace_kernel(TraceKernel k, const TraceArgs& A, const LaunchShape& L,
                               hipStream_t st) {
  // the primary-leg shadow reads the candidate masks of any obstructed scene
  if (has_obstructions(A.ob) && !A.cull) return hipErrorInvalidValue;
  TraceArgs a = A;
  void* args[] = {&a};
  return hipLaunchKernel((const void*)k, dim3(L.grid), dim3(L.threads), args, L.lds, st);
}

hipError_t launch_trace(const TraceArgs& A, const LaunchShape& L, int source_type, int mode,
                        hipStream_t st) {
  const int which = mode == IACT_MODE_RESPONSE ? TRACE_RESPONSE
                    : mode == IACT_MODE_DEBUG  ? TRACE_DEBUG
                                               : TRACE_IMAGE;
  return launch_trace_kernel(pick_trace_kernel(A, source_type, which), A, L, st);
}

hipError_t launch_trace_sm(const TraceArgs& A, const LaunchShape& L, int source_type,
                           hipStream_t st) {
  return launch_trace_kernel(pick_trace_kernel(A, source_type, TRACE_SOURCE_MAJOR), A, L, st);
}

template <int SRC, int KIND, bool AXIS, int SPL>
static const void* pick_adj_sm_spl(const AdjArgs& B) {
  const bool gx = B.g_val || B.g_src || B.g_pos;
  if (has_obstructions(B.T.ob))
    return gx ? (const void*)k_adj_sm<SRC, KIND, AXIS, 1, SPL, true>
              : (const void*)k_adj_sm<SRC, KIND, AXIS, 1, SPL, false>;
  return gx ? (const void*)k_adj_sm<SRC, KIND, AXIS, 0, SPL, true>
            : (const void*)k_adj_sm<SRC, KIND, AXIS, 0, SPL, false>;
}
template <int SRC, int KIND>
static const void* pick_adj_sm_k(const AdjArgs& B) {
  const bool ax = B.T.S.axis != 0;
  if (B.T.sm_spl == 2)
    return ax ? pick_adj_sm_spl<SRC, KIND, true, 2>(B) : pick_adj_sm_spl<SRC, KIND, false, 2>(B);
  return ax ? pick_adj_sm_spl<SRC, KIND, true, 1>(B) : pick_adj_sm_spl<SRC, KIND, false, 1>(B);
}
template <int SRC>
static const void* pick_adj_sm(const AdjArgs& B) {
  return B.T.S.kind == IACT_SENSOR_SQUARE ? pick_adj_sm_k<SRC, IACT_SENSOR_SQUARE>(B)
                                          : pick_adj_sm_k<SRC, IACT_SENSOR_HEX>(B);
}

hipError_t launch_adjoint_sm(const AdjArgs& B, const LaunchShape& L, int source_type,
                             hipStream_t st) {
  if (has_obstructions(B.T.ob) && !B.T.cull) return hipErrorInvalidValue;
  if (B.T.S.splat || B.T.n_stages > 0 || B.T.layout != 2 || !B.T.recs || !B.torque)
    return hipErrorInvalidValue;
  AdjArgs b = B;
  void* args[] = {&b};
  const void* k = source_type == IACT_SOURCE_POINT ? pick_adj_sm<IACT_SOURCE_POINT>(B)
                                                   : pick_adj_sm<IACT_SOURCE_PARALLEL>(B);
  return hipLaunchKernel(k, dim3(L.grid), dim3(L.threads), args, L.lds, st);
}

hipError_t launch_torque_to_euler(const double* torque, const float* rotations, int nf,
                                  double* g_rot, hipStream_t st) {
  if (nf <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_torque_to_euler, dim3((nf + 255) / 256), dim3(256), 0, st, torque,
                     rotations, nf, g_rot);
  return hipGetLastError();
}

hipError_t launch_cull(const TraceArgs& A, int source_type, int all, float* cones,
                       uint32_t* masks, hipStream_t st) {
  if (!all) {
    hipLaunchKernelGGL(k_facet_cones, dim3(grid_for(A.nf, 1, 4096)), dim3(256), 0, st, A,
                       source_type, cones);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const long long total = (long long)A.nf * A.cull_words;
  hipLaunchKernelGGL(k_cull_masks, dim3(grid_for(total, 256, 4096)), dim3(256), 0, st, A, cones,
                     masks, all);
  return hipGetLastError();
}

hipError_t launch_chunk_masks(const TraceArgs& A, int source_type, long long n_sources,
                              const float* cones, const uint32_t* facet_masks, uint32_t* out,
                              float* chunk_cones, hipStream_t st) {
  const long long total = (n_sources + 63) / 64 * A.nf;
  hipLaunchKernelGGL(k_chunk_masks, dim3(grid_for(total, 256, 16384)), dim3(256), 0, st, A,
                     source_type, n_sources, cones, facet_masks, out, chunk_cones);
  return hipGetLastError();
}

// One thread per block of 64 consecutive records: the centre of the block's
// bounding box of sample positions and the largest distance from it.
__global__ void k_group_bounds(const float* recs, long long total, float* bounds) {
  // one wave per block of 64 records, lane = record: the extremes and the
  // largest distance from the centre are wave reductions (min and max are
  // exact, so the ball is that of a serial pass)
  const long long nb = (total + 63) / 64;
  const int lane = lane_id(), wpb = block_waves();
  const long long wave = (long long)blockIdx.x * wpb + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (long long b = wave; b < nb; b += (long long)gridDim.x * wpb) {
    const long long i = 64 * b + lane;
    const bool in = i < total;  // (a lane past the end repeats the block's first record)
    const float* p = recs + IACT_REC_WORDS * (in ? i : 64 * b) + 13;
    const f3 q = mk3(p[0], p[1], p[2]);
    f3 lo = q, hi = q;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      lo = mk3(fminf(lo.x, __shfl_xor(lo.x, off, 64)), fminf(lo.y, __shfl_xor(lo.y, off, 64)),
               fminf(lo.z, __shfl_xor(lo.z, off, 64)));
      hi = mk3(fmaxf(hi.x, __shfl_xor(hi.x, off, 64)), fmaxf(hi.y, __shfl_xor(hi.y, off, 64)),
               fmaxf(hi.z, __shfl_xor(hi.z, off, 64)));
    }
    const f3 c = 0.5f * (lo + hi);
    const f3 dv = q - c;
    float r2 = dot(dv, dv);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) r2 = fmaxf(r2, __shfl_xor(r2, off, 64));
    // (1e-6 relative + 1 um of slack for the rounding of centre and radius)
    if (lane == 0) {
      float* o = bounds + 4 * b;
      o[0] = c.x;
      o[1] = c.y;
      o[2] = c.z;
      o[3] = fmaf(sqrtf(r2), 1.000001f, 1e-6f);
    }
  }
}

hipError_t launch_group_bounds(const TraceArgs& A, float* bounds, hipStream_t st) {
  const long long total = (long long)A.nf * A.M;
  hipLaunchKernelGGL(k_group_bounds, dim3(grid_for((total + 63) / 64, 4, 4096)), dim3(256), 0,
                     st, A.recs, total, bounds);
  return hipGetLastError();
}

hipError_t launch_records(const TraceArgs& A, float* recs, hipStream_t st) {
  const long long total = (long long)A.nf * A.M;
  // the flag word is the first word of the padding after the records
  float* flag = recs + IACT_REC_WORDS * total;
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, st, flag, 1ll, 0.f);
  if (A.S.kind == IACT_SENSOR_SQUARE)
    hipLaunchKernelGGL(k_sample_records<IACT_SENSOR_SQUARE>, dim3(grid_for(A.nf, 1, 16384)),
                       dim3(256), 0, st, A, recs, (uint32_t*)flag);
  else
    hipLaunchKernelGGL(k_sample_records<IACT_SENSOR_HEX>, dim3(grid_for(A.nf, 1, 16384)),
                       dim3(256), 0, st, A, recs, (uint32_t*)flag);
  return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double* in, float* out, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(n, 256, 1024)), dim3(256), 0, st, in, out, n);
  return hipGetLastError();
}

hipError_t launch_reduce(const float* partials, int nblocks, int P, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(grid_for(P, 256, 1024)), dim3(256), 0, st, partials,
                     nblocks, P, out);
  return hipGetLastError();
}

}  // namespace iact
