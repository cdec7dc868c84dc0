// Helpers shared by the C ABI translation units.
#pragma once

#include "iact_kernels.h"

namespace iact {

struct DevInfo {
  int n_cu = 0;
  int lds_per_block = 0;
  char arch[64] = {0};
};
int device_info(DevInfo& out);
int shape_trace(const iact_scene& sc, int ns, int mode, const iact_launch* L, TraceArgs& A,
                LaunchShape& out, const DevInfo& dev);
int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* where);
int validate_scene(const iact_scene* sc);
int validate_facets(const iact_facets& F);
int validate_sensor(const iact_sensor& S);
DevSensor make_dev_sensor(const iact_sensor& s);
void fill_trace_args(TraceArgs& A, const iact_scene& sc, const double* sources,
                     const float* values, int ns);

}  // namespace iact
