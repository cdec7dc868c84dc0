// Launchers shared between the kernel translation units and the C ABI.
#pragma once

#include "iact_trace.h"

#define IACT_MAX_BLOCK 256

namespace iact {

struct LaunchShape {
  int grid;
  int threads;
  size_t lds;
};

hipError_t launch_fill(float* p, long long n, float v, hipStream_t st);
hipError_t launch_fill_d(double* p, long long n, double v, hipStream_t st);
hipError_t launch_sample(const iact_sample_desc& D, float* p, float* n, float* d, float* w,
                         hipStream_t st);
hipError_t launch_world(const iact_facets& F, const DevSensor& S, float* out, hipStream_t st);
// The trace kernels, by role (TRACE_SOURCE_MAJOR: the source-major image
// trace).
enum TraceRole { TRACE_IMAGE = 0, TRACE_DEBUG = 1, TRACE_RESPONSE = 2, TRACE_SOURCE_MAJOR = 3 };
using TraceKernel = void (*)(TraceArgs);
TraceKernel pick_trace_kernel(const TraceArgs& A, int source_type, int which);
hipError_t launch_trace_kernel(TraceKernel k, const TraceArgs& A, const LaunchShape& L,
                               hipStream_t st);
hipError_t launch_trace(const TraceArgs& A, const LaunchShape& L, int source_type, int mode,
                        hipStream_t st);
hipError_t launch_accumulate(const DevSensor& S, const float* x, const float* y, const float* v,
                             long long n, float* out, int32_t* pixel, hipStream_t st);
hipError_t launch_trace_sm(const TraceArgs& A, const LaunchShape& L, int source_type,
                           hipStream_t st);
hipError_t launch_records(const TraceArgs& A, float* recs, hipStream_t st);
hipError_t launch_cull(const TraceArgs& A, int source_type, int all, float* cones,
                       uint32_t* masks, hipStream_t st);
hipError_t launch_chunk_masks(const TraceArgs& A, int source_type, long long n_sources,
                              const float* cones, const uint32_t* facet_masks, uint32_t* out,
                              float* chunk_cones, hipStream_t st);
// Bounding balls of the blocks of 64 consecutive sample records (A.recs).
hipError_t launch_group_bounds(const TraceArgs& A, float* bounds, hipStream_t st);
// Source ordering for chunk culling (iact_sort.hip): sorts the sources of
// each instance (offsets: n_inst + 1 device prefix, or NULL for one) into
// direction cells; writes perm (sorted -> original), sorted sources/values.
size_t sort_workspace_bytes(long long n_sources, int n_inst);
hipError_t launch_sort_sources(const double* src, const float* val, long long n_sources,
                               const int64_t* offsets, int n_inst, int source_type,
                               void* ws, int32_t* perm, double* src_out, float* val_out,
                               hipStream_t st);
hipError_t launch_reduce(const float* partials, int nblocks, int P, float* out, hipStream_t st);
// out[i] = (float)in[i], i < n (the fp64 image sums of a render)
hipError_t launch_f64_to_f32(const double* in, float* out, long long n, hipStream_t st);

// Arguments of the adjoint kernel (iact_adjoint.hip).
struct AdjArgs {
  TraceArgs T;
  const float* grad_image;  // (P), or (n_inst, P) for batched renders
  double* g_rot;            // (nf, 3), or (n_inst, nf, 3) for batched renders
  double* g_pos;
  double* g_src;
  double* g_val;
  double* g_asph;
  const int32_t* surface_index;
  const float* offsets;
  const int32_t* src_perm;  // sorted -> caller source index (NULL: not sorted)
  double* torque;           // source-major adjoint: (nf, 3) per-facet sums of
                            // q x dL/dp + n x dL/dn (workspace)
};
hipError_t launch_adjoint(const AdjArgs& B, const LaunchShape& L, int source_type,
                          hipStream_t st);
// Source-major adjoint of a primary-only hard-binned image render (sample
// records in B.T.recs, laid out as the source-major forward's): per-facet
// torques into B.torque, positions / source / value gradients when set.
hipError_t launch_adjoint_sm(const AdjArgs& B, const LaunchShape& L, int source_type,
                             hipStream_t st);
// g_rot[f] = (pi / 180) (w_a, w_b, w_c) . torque[f]: the rotation gradient of
// each facet (Euler degrees) from its torque, w_k the rotation axis of angle k.
hipError_t launch_torque_to_euler(const double* torque, const float* rotations, int nf,
                                  double* g_rot, hipStream_t st);
const void* pick_adjoint_kernel(const AdjArgs& B, int source_type);

}  // namespace iact
