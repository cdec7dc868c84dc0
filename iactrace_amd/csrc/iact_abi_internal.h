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
int work_order(const iact_launch* L, long long nchunks, int grid, int wpb, TraceArgs& A);
int shape_trace(const iact_scene& sc, int ns, int mode, const iact_launch* L, TraceArgs& A,
                LaunchShape& out, const DevInfo& dev, int source_type);
int default_blocks_per_cu(int det);
// Persistent grid for `items` wave-sized work items at the shape's block
// size; updates the work order that depends on the grid.
int regrid(long long items, const iact_launch* L, const DevInfo& dev, LaunchShape& shp,
           TraceArgs& A);
// Obstruction culling: automatic from this many obstructions (a dish with a
// camera box and four masts already traces 2.3x faster culled).  Every scene
// with obstructions gets candidate masks (all bits set when not culling);
// from kChunkCullMinSources sources the masks are per (source chunk of 64,
// facet), and traces that may reorder their sources sort them into compact
// direction cells first.  The tables follow a mode's own workspace of `base`
// bytes (cull_layout) and are filled by prepare_cull, which points A.cull at
// the masks, A.sources/A.values at the sorted copies and *perm at the
// permutation (sorted -> caller index) when it sorts.
constexpr int kAutoCullObstructions = 1;
constexpr long long kChunkCullMinSources = 128;
struct CullSpec {
  long long n_src;         // sources of the trace (all instances)
  int n_inst;              // instances (1 for a plain trace)
  const int64_t* offsets;  // device (n_inst + 1) source prefix, or NULL
  bool sort;               // the trace may reorder its sources
  long long n_groups = 0;  // blocks of 64 samples whose bounds refine the
                           // chunk masks (source-major image traces), or 0
};
struct CullLayout {
  // byte offsets: facet cones and masks; per (chunk, facet) masks and cone
  // axes; bounds of the sample blocks; source sort
  size_t cones, fmasks, cmasks, ccones, gbounds, sws, perm, src, val, end;
  bool chunks, sort;
};
// Sample blocks of a trace whose chunk masks the trace refines per block.
long long refine_groups(const TraceArgs& A, int mode);
int cull_words_for(const iact_scene& sc, const iact_launch* L, int& words, int& all);
// Offset of the fp64 image sums after a workspace of `end` bytes.
size_t image64_offset(size_t end);
// Source-major adjoint: whether a backward takes it, and its workspace
// layout (record table, culling tables from recs_end, per-facet torques at
// `torque`, `end` bytes in all) with the trace arguments and launch shape.
struct AdjSmLayout {
  size_t recs_end, torque, end;
};
bool bwd_source_major(const iact_scene& sc, int ns, const iact_launch* L, bool want_asph);
int adj_sm_layout(const iact_scene& sc, const double* sources, const float* values, int ns,
                  const iact_launch* L, const DevInfo& dev, TraceArgs& A, LaunchShape& shp,
                  AdjSmLayout& out);
CullLayout cull_layout(int nf, int words, int all, size_t base, const CullSpec& c);
int prepare_cull(TraceArgs& A, void* workspace, size_t base, int source_type, hipStream_t st,
                 const CullSpec& c, const int32_t** perm);
int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* where);
int validate_scene(const iact_scene* sc);
int validate_facets(const iact_facets& F);
int validate_sensor(const iact_sensor& S);
DevSensor make_dev_sensor(const iact_sensor& s);
void fill_trace_args(TraceArgs& A, const iact_scene& sc, const double* sources,
                     const float* values, int ns);
// LDS bytes of a hexagonal sensor's staged lookup table (the table with a
// border entry on every side, stage_hex_table); 0 for square sensors.
inline size_t hex_table_bytes(const iact_sensor& s) {
  return s.kind == IACT_SENSOR_HEX ? (size_t)(s.table_q + 2) * (s.table_r + 2) * sizeof(int32_t)
                                   : 0;
}

}  // namespace iact
