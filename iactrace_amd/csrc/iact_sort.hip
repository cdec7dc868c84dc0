// Source ordering for obstruction culling per source chunk.
//
// The trace culls obstructions per (source chunk, facet): the 64 consecutive
// sources of a chunk bound one narrow bundle of rays per facet.  Sources in
// caller order (a random star field, say) make every chunk span the whole
// field, so the image and gradient traces first sort each instance's
// sources into direction cells: a counting sort on the Morton index of the
// cell of the source's direction in the gnomonic projection about the mean
// direction, on a G x G grid over the field's extent.  The order inside a
// cell is the order in which atomics land (a sum does not care); results
// that are per source (gradients) are written back through the permutation.
#include "iact_kernels.h"

namespace iact {

// Cells per side: 64 (4096 cells per instance), fewer for many instances so
// that the bin array stays at most 2^20 entries.
static int grid_side(int n_inst) {
  int g = 64;
  while (g > 4 && (long long)n_inst * g * g > (1ll << 20)) g >>= 1;
  return g;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t sort_workspace_bytes(long long n_sources, int n_inst) {
  (void)n_sources;
  const int g = grid_side(n_inst);
  return align256(16 * sizeof(float)) + align256((size_t)n_inst * g * g * sizeof(uint32_t));
}

// unit direction towards a source: from the origin for point sources (the
// order only needs to be coherent), opposite the propagation direction for
// parallel ones
IACT_DEV f3 source_dir(const double* src, long long s, int source_type) {
  if (source_type == IACT_SOURCE_POINT) {
    const double x = src[3 * s], y = src[3 * s + 1], z = src[3 * s + 2];
    const double il = 1.0 / sqrt(x * x + y * y + z * z);
    return mk3((float)(x * il), (float)(y * il), (float)(z * il));
  }
  f3 u = mk3(-(float)src[3 * s], -(float)src[3 * s + 1], -(float)src[3 * s + 2]);
  return frsq(dot(u, u)) * u;
}

// One block: mean direction a, a frame (e1, e2) across it and the largest
// gnomonic coordinate |u.e / u.a| over the sources (directions more than
// ~84 degrees off the mean are clamped to the border).  frame: a, e1, e2,
// 1 / cell size, 16 floats.
__global__ __launch_bounds__(1024) void k_src_frame(const double* src, long long n,
                                                     int source_type, int g, float* frame) {
  __shared__ float red[3][1024];
  const int t = threadIdx.x;
  f3 sum = mk3(0.f, 0.f, 0.f);
  for (long long s = t; s < n; s += blockDim.x) sum = sum + source_dir(src, s, source_type);
  red[0][t] = sum.x;
  red[1][t] = sum.y;
  red[2][t] = sum.z;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (t < h) {
      red[0][t] += red[0][t + h];
      red[1][t] += red[1][t + h];
      red[2][t] += red[2][t + h];
    }
    __syncthreads();
  }
  f3 a = mk3(red[0][0], red[1][0], red[2][0]);
  const float la = sqrtf(dot(a, a));
  a = la > 0.f ? (1.f / la) * a : mk3(0.f, 0.f, 1.f);
  const f3 ref = fabsf(a.z) < 0.9f ? mk3(0.f, 0.f, 1.f) : mk3(1.f, 0.f, 0.f);
  f3 e1 = cross(ref, a);
  e1 = frsq(dot(e1, e1)) * e1;
  const f3 e2 = cross(a, e1);
  __syncthreads();
  float ext = 0.f;
  for (long long s = t; s < n; s += blockDim.x) {
    const f3 u = source_dir(src, s, source_type);
    const float w = fmaxf(dot(u, a), 0.1f);
    ext = fmaxf(ext, fmaxf(fabsf(dot(u, e1)), fabsf(dot(u, e2))) / w);
  }
  red[0][t] = ext;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (t < h) red[0][t] = fmaxf(red[0][t], red[0][t + h]);
    __syncthreads();
  }
  if (t == 0) {
    const float e = fmaxf(red[0][0], 1e-7f) * 1.0001f;
    frame[0] = a.x;
    frame[1] = a.y;
    frame[2] = a.z;
    frame[3] = e1.x;
    frame[4] = e1.y;
    frame[5] = e1.z;
    frame[6] = e2.x;
    frame[7] = e2.y;
    frame[8] = e2.z;
    frame[9] = (float)g / (2.f * e);  // cells per unit of gnomonic coordinate
  }
}

IACT_DEV uint32_t spread_bits(uint32_t v) {  // 6 bits -> even positions
  v &= 0x3fu;
  v = (v | (v << 4)) & 0x30fu;
  v = (v | (v << 2)) & 0x333u;
  v = (v | (v << 1)) & 0x555u;
  return v;
}

// Instance of source s: offsets is a (n_inst + 1) prefix.
IACT_DEV int instance_of(const int64_t* offsets, int n_inst, long long s) {
  if (!offsets) return 0;
  int lo = 0, hi = n_inst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (offsets[mid] <= s) lo = mid; else hi = mid - 1;
  }
  return lo;
}

IACT_DEV uint32_t source_key(const double* src, long long s, int source_type, const float* fr,
                             int g, int inst) {
  const f3 u = source_dir(src, s, source_type);
  const f3 a = mk3(fr[0], fr[1], fr[2]), e1 = mk3(fr[3], fr[4], fr[5]), e2 = mk3(fr[6], fr[7], fr[8]);
  const float w = fmaxf(dot(u, a), 0.1f);
  const float x = dot(u, e1) / w, y = dot(u, e2) / w;
  const float hg = 0.5f * (float)g;
  const int cx = (int)fminf(fmaxf(floorf(fmaf(x, fr[9], hg)), 0.f), (float)(g - 1));
  const int cy = (int)fminf(fmaxf(floorf(fmaf(y, fr[9], hg)), 0.f), (float)(g - 1));
  const uint32_t morton = spread_bits((uint32_t)cx) | (spread_bits((uint32_t)cy) << 1);
  return (uint32_t)inst * (uint32_t)(g * g) + morton;
}

__global__ void k_src_count(const double* src, long long n, int source_type, const float* frame,
                            int g, const int64_t* offsets, int n_inst, uint32_t* bins) {
  for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (long long)gridDim.x * blockDim.x) {
    const int inst = instance_of(offsets, n_inst, s);
    atomicAdd(bins + source_key(src, s, source_type, frame, g, inst), 1u);
  }
}

// Exclusive scan of the bin counts in place (one block).
__global__ __launch_bounds__(1024) void k_bin_scan(uint32_t* bins, int n) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x, T = blockDim.x;
  const int per = (n + T - 1) / T;
  const int b0 = min(t * per, n), b1 = min(b0 + per, n);
  uint32_t sum = 0u;
  for (int i = b0; i < b1; ++i) sum += bins[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {  // inclusive Hillis-Steele scan
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (int i = b0; i < b1; ++i) {
    const uint32_t c = bins[i];
    bins[i] = run;
    run += c;
  }
}

__global__ void k_src_scatter(const double* src, const float* val, long long n, int source_type,
                              const float* frame, int g, const int64_t* offsets, int n_inst,
                              uint32_t* cursor, int32_t* perm, double* src_out, float* val_out) {
  for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (long long)gridDim.x * blockDim.x) {
    const int inst = instance_of(offsets, n_inst, s);
    const uint32_t pos = atomicAdd(cursor + source_key(src, s, source_type, frame, g, inst), 1u);
    perm[pos] = (int32_t)s;
    src_out[3 * (size_t)pos] = src[3 * s];
    src_out[3 * (size_t)pos + 1] = src[3 * s + 1];
    src_out[3 * (size_t)pos + 2] = src[3 * s + 2];
    val_out[pos] = val[s];
  }
}

__global__ void k_zero_u32(uint32_t* p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0u;
}

static int blocks_for(long long n) {
  long long g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_sort_sources(const double* src, const float* val, long long n_sources,
                               const int64_t* offsets, int n_inst, int source_type, void* ws,
                               int32_t* perm, double* src_out, float* val_out, hipStream_t st) {
  const int g = grid_side(n_inst);
  float* frame = (float*)ws;
  uint32_t* bins = (uint32_t*)((char*)ws + align256(16 * sizeof(float)));
  const long long n_bins = (long long)n_inst * g * g;
  hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(n_bins)), dim3(256), 0, st, bins, n_bins);
  hipLaunchKernelGGL(k_src_frame, dim3(1), dim3(1024), 0, st, src, n_sources, source_type, g,
                     frame);
  hipLaunchKernelGGL(k_src_count, dim3(blocks_for(n_sources)), dim3(256), 0, st, src, n_sources,
                     source_type, frame, g, offsets, n_inst, bins);
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, st, bins, (int)n_bins);
  hipLaunchKernelGGL(k_src_scatter, dim3(blocks_for(n_sources)), dim3(256), 0, st, src, val,
                     n_sources, source_type, frame, g, offsets, n_inst, bins, perm, src_out,
                     val_out);
  return hipGetLastError();
}

}  // namespace iact
