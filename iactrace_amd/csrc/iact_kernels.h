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
hipError_t launch_trace(const TraceArgs& A, const LaunchShape& L, int source_type, int mode,
                        hipStream_t st);
hipError_t launch_accumulate(const DevSensor& S, const float* x, const float* y, const float* v,
                             long long n, float* out, int32_t* pixel, hipStream_t st);
hipError_t launch_reduce(const float* partials, int nblocks, int P, float* out, hipStream_t st);

}  // namespace iact
