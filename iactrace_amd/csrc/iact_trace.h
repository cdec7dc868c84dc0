// Per-ray forward trace shared by the forward, adjoint and batched kernels.
//
// Work decomposition (MI355X): one lane = one MC sample of one primary facet;
// a wave holds 64 samples of the same facet, so the facet pose, the source
// being traced, the obstruction tables and the stage>=1 mirror table are all
// wave-uniform (scalar loads).  A wave loops over a tile of sources with the
// sample's world-space point/normal kept in registers; the hit pixels of the
// wave are merged before touching the LDS image.
#pragma once

#include "iact_device.h"

namespace iact {

// Unsigned 32-bit division by an invariant divisor d >= 1 (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI
// 1994, fig. 4.1): q = (t + ((n - t) >> s1)) >> s2 with t = mulhi(m, n),
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, s1 = min(l, 1),
// s2 = max(l - 1, 0).  Exact for every 32-bit n.
struct FastDiv {
  uint32_t d, m, s1, s2;
};

__host__ __device__ inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv D;
  D.d = d;
  uint32_t l = 0;
  while (l < 32 && (1ull << l) < d) ++l;
  const unsigned __int128 num = ((unsigned __int128)1 << 32) * ((1ull << l) - d);
  D.m = (uint32_t)(num / d + 1);
  D.s1 = l > 0 ? 1u : 0u;
  D.s2 = l > 0 ? l - 1 : 0u;
  return D;
}

IACT_DEV uint32_t fdiv(uint32_t n, const FastDiv& D) {
  const uint32_t t = __umulhi(D.m, n);
  return (t + ((n - t) >> D.s1)) >> D.s2;
}

// Sizes of one instance of a render (items = nchunks * nst, padded).
struct InstDesc {
  FastDiv pb;          // divisor span * nst: items per full chunk block
  int32_t nst, ns;     // source tiles, sources
  long long sbase;     // first source in the packed source arrays
};
static_assert(sizeof(InstDesc) == 32 && offsetof(InstDesc, sbase) == 24,
              "decode_item reads InstDesc as 8 words");

struct TraceArgs {
  // stage-0 facets (local frame samples + pose)
  int nf, M, cpf;  // cpf = chunks of 64 samples per facet
  const float* points;
  const float* normals;
  const float* delta;
  const float* weights;
  const float* positions;
  const float* rotations;
  const float* pscale;
  // sources
  const double* sources;
  const float* values;
  int ns;
  // far point sources: |e|^2 from which a (facet, source) pair forms its
  // direction from e alone (+inf: never)
  float far_e2;
  int source_tile, n_stiles;
  long long n_items;
  // stage >= 1
  int n_mirrors, n_stages;
  int stages[8];
  const float* mirrors;
  int n_surfaces;
  iact_surface surfaces[IACT_MAX_SURFACES];
  iact_obstructions ob;
  DevSensor S;
  // outputs / scratch
  float* out;
  // image renders (not deterministic): the image is summed here, in fp64,
  // and converted to `out` after the trace -- an image is a sum of ~10^7
  // rays per pixel, and fp32 atomics from thousands of blocks would round
  // it well above 1e-5
  double* out64;
  float* dbg_xy;
  int32_t* dbg_pix;
  float* partials;
  const float* world;  // SoA (7, nf*M) when precomputed
  int lds_image;       // image accumulated in LDS
  int lds_table;       // hex table staged in LDS
  int det;             // deterministic (wave-private images, ordered sums)
  int max_rounds;
  int acc;             // 0 wave-merged, 1 per-lane run-length, 2 direct atomics
  // work order (see decode_item)
  uint32_t span;       // chunks per block of the order
  uint32_t nchunks;    // nf * cpf
  FastDiv d_span, d_last, d_cpf;  // divide by span, by the last block's size, by cpf
  InstDesc inst0;      // the single instance of a plain render
  // batched renders: n_inst telescope instances share the facet samples and
  // carry their own stage-0 rotations, source lists and images
  int n_inst;
  const long long* item_prefix;  // (n_inst + 1) work-item prefix; NULL = one instance
  const InstDesc* inst_desc;     // (n_inst) per-instance sizes
  long long rot_stride;          // floats between two instances' rotations
  int xcd_order;                 // 1: XCD-major block numbering
  // source-major layout: lanes are sources, samples come from `recs`
  int layout;                    // 1 sample-major, 2 source-major
  const float* recs;             // (nf*M, IACT_REC_WORDS) world-space sample records
  int sample_group;              // samples per work item
  int sm_spl;                    // sources per lane (1 or 2)
  uint32_t n_sc;                 // source chunks of 64 * sm_spl
  FastDiv d_sc;                  // divide by n_sc
  FastDiv d_m;                   // divide by M (sample -> facet)
  long long sm_items;            // sample groups * source chunks
  // per-facet obstruction candidate bit masks for the primary leg's shadow
  // rays; set whenever the scene has obstructions (culling off = every bit)
  const uint32_t* cull;
  int cull_words;
  int cull_all;                  // masks with every obstruction (culling off)
  int cull_chunks;               // masks per (source chunk of 64, facet)
  // with chunk masks: per (chunk, facet) the axis and tan(half angle) of the
  // chunk's ray bundle (tan < 0: not culled); for source-major image traces
  // the bounding ball (centre, radius) of each block of 64 sample records,
  // with which a facet run inside one block refines its chunk's candidates
  const float* ccones;
  const float* gbounds;
};

// World-space sample record of the source-major trace (16 words, read with
// scalar loads): q = p - facet position, N = (n.An, n.Au, n.Av), 1/(2 weight),
// the sensor-grid constants of p, m = -2 n, p -- the fields of the Sample
// that make_sample builds which the image trace uses (n = -m/2 exactly).
#define IACT_REC_WORDS 16

// Position of a block in an XCD-major order.  Workgroups are dispatched to
// the 8 XCDs round-robin (block b runs on XCD b % 8), so consecutive logical
// blocks share one XCD's L2.  A permutation of [0, gridDim.x).
// Grid size in blocks and block size in waves as values the compiler knows
// to be uniform: the dispatch-packet loads behind gridDim / blockDim can
// otherwise reach a loop through a join after lane-dependent loops (the LDS
// fills at a kernel's start), which marks everything derived from them --
// and the loops they bound -- divergent.
IACT_DEV int grid_blocks() { return __builtin_amdgcn_readfirstlane((int)gridDim.x); }
IACT_DEV int block_waves() { return __builtin_amdgcn_readfirstlane((int)(blockDim.x >> 6)); }

IACT_DEV long long xcd_block(int on) {
  const int G = grid_blocks(), b = blockIdx.x;
  if (!on) return b;
  const int x = b & 7, per = G >> 3, rem = G & 7;
  return (long long)x * per + min(x, rem) + (b >> 3);
}

// One work item: 64 consecutive samples of facet f (lane -> sample m) traced
// against sources [s0, s1) of instance inst.
//
// Order within an instance: chunks come in blocks of `span` (by default the
// waves of one grid sweep); inside a block the source tile is the slower
// index and the chunk the faster.  One sweep of the persistent grid therefore
// traces the chunks of a block against one source tile -- all waves read the
// same sources -- and the next sweeps revisit the same chunks, which each XCD
// keeps in its L2 (span / 8 chunks of 2.5 KB).  Instances are padded to a
// multiple of the block's wave count so that the waves of a block always work
// on the same instance.  Every division of the decode is by a divisor fixed
// per launch or per instance, done with a precomputed multiplier: integer
// division has no scalar hardware and costs a long vector sequence.
struct WorkItem {
  int inst, f, m, s0, s1;
  long long sbase;       // first source of the instance in the source arrays
};

IACT_DEV long long total_items(const TraceArgs& A) {
  return A.item_prefix ? as_const(A.item_prefix)[A.n_inst] : A.n_items;
}

// Instance of an item (0 for single renders).
IACT_DEV int item_instance(const TraceArgs& A, long long item) {
  if (!A.item_prefix) return 0;
  const cptr<long long> ip = as_const(A.item_prefix);
  int lo = 0, hi = A.n_inst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ip[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// false for the padding items at the end of an instance.
IACT_DEV bool decode_item(const TraceArgs& A, long long item, int inst, int lane, WorkItem& w) {
  long long local64 = item;
  InstDesc I;
  if (A.item_prefix) {
    local64 = item - as_const(A.item_prefix)[inst];
    // scalar loads of the 8 words of the instance's descriptor
    const cptr<uint32_t> u = (cptr<uint32_t>)as_const(A.inst_desc + inst);
    I.pb.d = u[0];
    I.pb.m = u[1];
    I.pb.s1 = u[2];
    I.pb.s2 = u[3];
    I.nst = (int32_t)u[4];
    I.ns = (int32_t)u[5];
    I.sbase = (long long)((uint64_t)u[6] | ((uint64_t)u[7] << 32));
  } else {
    I.pb.d = A.inst0.pb.d;
    I.pb.m = A.inst0.pb.m;
    I.pb.s1 = A.inst0.pb.s1;
    I.pb.s2 = A.inst0.pb.s2;
    I.nst = A.inst0.nst;
    I.ns = A.inst0.ns;
    I.sbase = A.inst0.sbase;
  }
  const uint32_t local = (uint32_t)local64;
  if (local64 >= (long long)A.nchunks * I.nst) return false;
  const uint32_t cb = fdiv(local, I.pb);
  const uint32_t rem = local - cb * I.pb.d;
  const uint32_t first = cb * A.span;
  const bool full = A.nchunks - first >= A.span;
  // the divisor picked field by field (a selected struct would go through
  // the stack)
  FastDiv Db;
  Db.d = full ? A.d_span.d : A.d_last.d;
  Db.m = full ? A.d_span.m : A.d_last.m;
  Db.s1 = full ? A.d_span.s1 : A.d_last.s1;
  Db.s2 = full ? A.d_span.s2 : A.d_last.s2;
  const uint32_t stile = fdiv(rem, Db);
  const uint32_t cbs = Db.d;
  const uint32_t chunk = first + (rem - stile * cbs);
  const uint32_t f = fdiv(chunk, A.d_cpf);
  // everything but the lane's sample is wave-uniform; readfirstlane states it
  // for the compiler, which otherwise may keep these in vector registers in
  // large kernels (then the source loop and its loads turn vector too)
  w.inst = __builtin_amdgcn_readfirstlane(inst);
  w.sbase = (long long)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(
                             (int)(uint32_t)((uint64_t)I.sbase >> 32))
                         << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)I.sbase));
  w.f = __builtin_amdgcn_readfirstlane((int)f);
  w.m = (int)(chunk - f * A.d_cpf.d) * 64 + lane;
  w.s0 = __builtin_amdgcn_readfirstlane((int)stile * A.source_tile);
  w.s1 = __builtin_amdgcn_readfirstlane(min((int)stile * A.source_tile + A.source_tile, I.ns));
  return true;
}

// One sample in world space, ready to be traced against many sources.
struct Sample {
  f3 p, n;
  f3 m;            // -2 n: the reflection's scaled cosine is one dot product
  f3 N;            // (n.An, n.Au, n.Av): the normal's part of the hit (grid_hit)
  f3 q;            // p - facet position, in fp32
  float nrw;       // -1 / weight (the reflected value is value * cos / w, cos = -d.n)
  float nrwh;      // -nrw / 2 (value * (d.m) * nrwh is value * (d.n) * nrw exactly)
  PlaneK K;        // sensor-plane constants of the primary origin
  GridK G;         // the same in sensor-grid coordinates
  int f;           // facet (selects its obstruction candidate list)
  uint32_t cand0;  // the facet's first candidate-mask word
  int cand_any;    // the facet has an obstruction candidate
  size_t cand_row; // offset of the facet's candidate mask
  // refined candidates (source-major traces): rmode set, the candidates are
  // the lanes of rlist that rbits selects (shadow_factor_list)
  int rmode;
  int rlist;
  uint64_t rbits;
};

// Offset of mask row `row` (facet, or chunk * nf + facet), wave-uniform.
IACT_DEV size_t mask_offset(const TraceArgs& A, long long row) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)row);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)row >> 32));
  return (size_t)(((uint64_t)hi << 32) | lo) * (size_t)A.cull_words;
}

// A facet's obstruction candidates as the trace needs them per ray (read
// once per facet, or per facet and source chunk, with scalar loads; both
// are wave-uniform).
struct FacetCull {
  int f;
  uint32_t first;
  int any;
  size_t row;
};
IACT_DEV FacetCull facet_cull(const TraceArgs& A, int f, long long chunk) {
  FacetCull c{f, 0u, 0, 0};
  if (A.cull) {
    const long long row = A.cull_chunks ? chunk * A.nf + f : f;
    c.row = mask_offset(A, row);
    const cptr<uint32_t> m = as_const(A.cull) + c.row;
    uint32_t any = m[0];
    c.first = any;
    for (int w = 1; w < A.cull_words; ++w) any |= m[w];
    c.any = any != 0u;
  }
  return c;
}
IACT_DEV void set_facet(Sample& s, const FacetCull& c) {
  s.f = c.f;
  s.cand0 = c.first;
  s.cand_any = c.any;
  s.cand_row = c.row;
}

template <int KIND>
IACT_DEV Sample make_sample(const DevSensor& S, f3 p, f3 n, float w, f3 facet_pos) {
  Sample s;
  s.p = p;
  s.n = n;
  s.m = -2.f * n;
  s.q = p - facet_pos;
  s.nrw = -frcp(w);
  s.nrwh = -0.5f * s.nrw;
  s.K = plane_consts(S, p);
  s.G = grid_consts<KIND>(S, s.K);
  s.N = sensor_dirs(S, n);
  s.f = 0;
  s.cand0 = 0u;
  s.cand_any = 0;
  s.cand_row = 0;
  s.rmode = 0;
  s.rlist = 0;
  s.rbits = 0ull;
  return s;
}

// Primary direction of a ray.  Point sources sit up to ~1e10 m away, so the
// offset between a facet and a source is formed in fp64 and rounded once,
// e = fp32(facet position - source), per (facet, source); the ray from a
// sample p of the facet has direction normalise(q + e), q = p - facet
// position.  Parallel sources carry their direction, used as given.
template <int SRC>
IACT_DEV f3 source_offset(double sx, double sy, double sz, f3 facet_pos) {
  if (SRC == IACT_SOURCE_POINT)
    return mk3((float)((double)facet_pos.x - sx), (float)((double)facet_pos.y - sy),
               (float)((double)facet_pos.z - sz));
  return mk3((float)sx, (float)sy, (float)sz);
}

// A point source so far from a facet that the directions from its samples
// differ from the one from the facet position by < 2^-30 rad (|q| <= R for
// every sample and |e| >= 2^30 R; fp32 resolves 2^-24) takes u = e for every
// sample: one direction per (facet, source), as the reference's fp32
// p - source (render.py:130), which loses p at such distances, has it.
// Returns qs, the multiplier of q in u = qs q + e: 0 far, 1 otherwise.  The
// same test in every kernel keeps their rays alike.
IACT_DEV float source_qs(const TraceArgs& A, f3 e) { return dot(e, e) >= A.far_e2 ? 0.f : 1.f; }

template <int SRC>
IACT_DEV f3 primary_direction(const Sample& s, f3 e, float qs, float& inv) {
  if (SRC == IACT_SOURCE_POINT) {
    const f3 u = fma3(qs, s.q, e);
    inv = frsq(fmaf(u.x, u.x, fmaf(u.y, u.y, u.z * u.z)));
    return mk3(u.x * inv, u.y * inv, u.z * inv);
  }
  inv = 1.f;
  return e;
}

// Wave-uniform read of a lane's value (v_readlane).
IACT_DEV float lane_value(float x, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}
IACT_DEV f3 lane_value(f3 x, int k) {
  return mk3(lane_value(x.x, k), lane_value(x.y, k), lane_value(x.z, k));
}

// Record of a sample (layout documented at IACT_REC_WORDS).
IACT_DEV void store_record(float* r, const Sample& s) {
  r[0] = s.q.x;
  r[1] = s.q.y;
  r[2] = s.q.z;
  r[3] = s.N.x;
  r[4] = s.N.y;
  r[5] = s.N.z;
  r[6] = s.nrwh;
  r[7] = s.G.kt;
  r[8] = s.G.ku;
  r[9] = s.G.kv;
  r[10] = s.m.x;
  r[11] = s.m.y;
  r[12] = s.m.z;
  r[13] = s.p.x;
  r[14] = s.p.y;
  r[15] = s.p.z;
}

// Record read with scalar loads (the address is wave-uniform) into SGPRs;
// words a trace does not use (p without obstructions) are not loaded.
// Vector loads of a uniform address would return the same bytes to each of
// the 64 lanes, and that data return, not the arithmetic, bounds the trace;
// SGPR operands cost a move only where an instruction needs two.
struct RecRaw {
  float4 a, b, c, d;
};
IACT_DEV RecRaw fetch_record(const float4* r) {
  const cptr<float> q = as_const((const float*)r);
  return RecRaw{make_float4(q[0], q[1], q[2], q[3]), make_float4(q[4], q[5], q[6], q[7]),
                make_float4(q[8], q[9], q[10], q[11]), make_float4(q[12], q[13], q[14], q[15])};
}
IACT_DEV Sample unpack_record(const RecRaw& q) {
  Sample s;
  s.q = mk3(q.a.x, q.a.y, q.a.z);
  s.N = mk3(q.a.w, q.b.x, q.b.y);
  s.nrwh = q.b.z;
  s.nrw = -2.f * s.nrwh;
  s.G = GridK{q.b.w, q.c.x, q.c.y};
  s.m = mk3(q.c.z, q.c.w, q.d.x);
  s.n = -0.5f * s.m;  // exact: m = -2 n
  s.p = mk3(q.d.y, q.d.z, q.d.w);
  s.K = PlaneK{0.f, 0.f, 0.f};  // world hit coordinates are not traced from records
  s.f = 0;
  s.cand0 = 0u;
  s.cand_any = 0;
  s.cand_row = 0;
  s.rmode = 0;
  s.rlist = 0;
  s.rbits = 0ull;
  return s;
}

// MirrorGroup.transform_to_world (mirrors.py:64-79) for one sample.
IACT_DEV void world_sample(const float* points, const float* normals, const float* delta,
                           const m3& R, f3 pos, float scale, size_t gi, f3& pw, f3& nw) {
  const f3 pl = mk3(points[3 * gi], points[3 * gi + 1], points[3 * gi + 2]);
  const f3 nl = mk3(normals[3 * gi], normals[3 * gi + 1], normals[3 * gi + 2]);
  const f3 dl = mk3(delta[3 * gi], delta[3 * gi + 1], delta[3 * gi + 2]);
  pw = mul(R, pl) + pos;
  const f3 nr = mul(R, nl);
  const f3 dr = mul(R, dl);
  const f3 pert = nr + scale * dr;
  nw = frsq(dot(pert, pert)) * pert;
}

IACT_DEV m3 facet_rotation(const float* rotations, int f) {
  return euler_to_matrix(rotations[3 * f], rotations[3 * f + 1], rotations[3 * f + 2]);
}

// Aperture test of a stage>=1 mirror (mirrors.py:147-149 disk,
// 209-220 convex CCW polygon).
IACT_DEV bool in_aperture(const float* M, float x, float y) {
  if (M[17] == 0.f) return x * x + y * y <= M[18] * M[18];
  const int n = (int)M[19];
  bool ok = true;
  for (int i = 0; i < n; ++i) {
    const int j = (i + 1 == n) ? 0 : i + 1;
    const float ax = M[20 + 2 * i], ay = M[21 + 2 * i];
    const float bx = M[20 + 2 * j], by = M[21 + 2 * j];
    ok = ok && ((bx - ax) * (y - ay) - (by - ay) * (x - ax) >= 0.f);
  }
  return ok;
}

// Hit of one stage>=1 mirror in its local frame (surfaces.py:67-108):
// float64 conic guess + Newton, then fp32 point/normal at the rounded (x, y).
struct MirrorHit {
  float t;
  f3 pl, nl;
};

IACT_DEV MirrorHit mirror_hit(const TraceArgs& A, const float* M, f3 ol, f3 dl) {
  const iact_surface& su = A.surfaces[(int)M[38]];
  const SurfT<double> s64 = load_surface<double>(su);
  const double x0 = (double)M[15], y0 = (double)M[16];
  const double z0 = sag_raw<double>(s64, x0, y0);
  const double t0 = intersect_conic((double)ol.x + x0, (double)ol.y + y0, (double)ol.z + z0,
                                    (double)dl.x, (double)dl.y, (double)dl.z, s64.c, s64.k);
  bool ok;
  const double t = newton_t(s64, ol.x, ol.y, ol.z, dl.x, dl.y, dl.z, x0, y0, z0, t0, ok);
  MirrorHit h;
  const float xh = (float)((double)ol.x + t * (double)dl.x);
  const float yh = (float)((double)ol.y + t * (double)dl.y);
  const SurfT<float> s32 = load_surface<float>(su);
  const float z0f = sag_raw<float>(s32, M[15], M[16]);
  surface_point_normal(s32, xh, yh, M[15], M[16], z0f, h.pl, h.nl);
  h.t = (ok && in_aperture(M, xh, yh)) ? (float)t : __builtin_inff();
  return h;
}

IACT_DEV m3 mirror_rotation(const float* M) {
  m3 R;
#pragma unroll
  for (int i = 0; i < 9; ++i) R.m[i] = M[6 + i];
  return R;
}

// What the adjoint needs to know about one stage>=1 reflection of a ray.
struct StageRec {
  f3 o, r;        // ray entering the stage
  float v;        // value entering the stage
  float t;        // hit parameter along the local ray
  float xh, yh;   // local hit (x, y)
  float c2;       // cos between the ray and the hit normal
  float hs;       // hit * shadow (0 or 1)
  int mi;         // mirror row, -1 when no mirror was hit
};

// _reflect_at_stage (render.py:44-79): nearest hit over the stage's mirrors
// (first minimum wins, table is ordered by group then index), reflect, and
// value *= hit * shadow * |cos|.  A miss keeps point 0 / normal 0.  Fills rec
// when REC.
template <bool REC>
IACT_DEV void stage_forward(const TraceArgs& A, int stage, f3& o, f3& d, float& v, bool active,
                            StageRec& rec) {
  float best = __builtin_inff();
  f3 bp = mk3(0.f, 0.f, 0.f), bn = mk3(0.f, 0.f, 0.f);
  int bi = -1;
  float bx = 0.f, by = 0.f;
  for (int i = 0; i < A.n_mirrors; ++i) {
    const float* M = A.mirrors + (size_t)IACT_MIRROR_STRIDE * i;
    if ((int)M[36] != stage) continue;
    const m3 R = mirror_rotation(M);
    const f3 pos = mk3(M[0], M[1], M[2]);
    const MirrorHit h = mirror_hit(A, M, mulT(R, o - pos), mulT(R, d));
    if (h.t < best) {
      best = h.t;
      bp = mul(R, h.pl) + pos;
      bn = mul(R, h.nl);
      if (REC) {
        bi = i;
        bx = h.pl.x;
        by = h.pl.y;
      }
    }
  }
  const float c = dot(d, bn);
  const f3 r = fma3(-2.f * c, bn, d);
  const float hit = best < 1e10f ? 1.f : 0.f;
  const float shadow = has_obstructions(A.ob) ? shadow_factor(A.ob, o, d, active) : 1.f;
  if (REC) {
    rec.o = o;
    rec.r = d;
    rec.v = v;
    rec.t = best;
    rec.xh = bx;
    rec.yh = by;
    rec.c2 = c;
    rec.hs = hit * shadow;
    rec.mi = hit > 0.f ? bi : -1;
  }
  v = v * hit * shadow * fabsf(c);
  o = bp;
  d = r;
}

IACT_DEV void reflect_at_stage(const TraceArgs& A, int stage, f3& o, f3& d, float& v,
                               bool active) {
  StageRec unused;
  stage_forward<false>(A, stage, o, d, v, active, unused);
}

// Shadow factor of the primary leg: obstructions between the sample and
// the source (render.py:118-128), tested from the facet's candidate mask
// (every obstruction when the launch does not cull; the host builds the
// masks before any trace of a scene with obstructions).  OBST: 0 = the scene
// has no obstructions, 1 = it has, 2 = test at run time.
template <int OBST>
IACT_DEV float primary_shadow(const TraceArgs& A, const Sample& s, f3 d, bool active) {
  if (!(OBST == 1 || (OBST == 2 && has_obstructions(A.ob)))) return 1.f;
  if (s.rmode) return shadow_factor_list(A.ob, s.rlist, s.rbits, s.p, -d, active);
  return shadow_factor_culled(A.ob, as_const(A.cull) + s.cand_row, A.cull_words, s.cand0,
                              s.cand_any, s.p, -d, active);
}

// Can a ray of the bundle (centre c, radius rho, axis a, tan theta) pass
// within R of b?  Along the axis the ray point lies in [-rho + t cos, rho + t],
// across it within rho + t sin of the axis, so it is necessary that
// s + rho + R >= 0 and e <= rho + R + tan(theta) (s + rho + R), with s, e the
// axial and radial coordinates of b.  Radii get 0.1% + 1 mm of slack.
IACT_DEV bool bundle_reaches(f3 c, float rho, f3 a, float tn, f3 b, float R) {
  const f3 w = b - c;
  const float s = dot(w, a);
  const f3 perp = fma3(-s, a, w);
  const float e = sqrtf(dot(perp, perp));
  const float rr = fmaf(R, 1.001f, 1e-3f) + rho;
  // written so that NaN geometry keeps the candidate
  return !(s + rr < 0.f) && !(e > fmaf(tn, s + rr, rr));
}

// Can a ray of the bundle reach obstruction j (mask numbering: cylinders,
// boxes, spheres, oriented boxes, triangles)?  Cylinders are tested as the
// capsule of their axis segment (cylinder_reaches), the other kinds through
// a bounding sphere.
IACT_DEV bool cylinder_reaches(const float* q, f3 c, float rho, f3 a, float tn);
IACT_DEV bool obstruction_reaches(const iact_obstructions& ob, int j, f3 c, float rho, f3 a,
                                  float tn) {
  const int c0 = ob.n_cylinders, c1 = c0 + ob.n_boxes, c2 = c1 + ob.n_spheres;
  const int c3 = c2 + ob.n_oriented_boxes;
  if (j < c0) return cylinder_reaches(ob.cylinders + 8 * j, c, rho, a, tn);
  if (j < c1) {
    const float* q = ob.boxes + 6 * (j - c0);
    const f3 lo = mk3(q[0], q[1], q[2]), hi = mk3(q[3], q[4], q[5]);
    const f3 dd = hi - lo;
    return bundle_reaches(c, rho, a, tn, 0.5f * (lo + hi), 0.5f * sqrtf(dot(dd, dd)));
  }
  if (j < c2) {
    const float* q = ob.spheres + 4 * (j - c1);
    return bundle_reaches(c, rho, a, tn, mk3(q[0], q[1], q[2]), q[3]);
  }
  if (j < c3) {
    const float* q = ob.oriented_boxes + 15 * (j - c2);
    const f3 he = mk3(q[3], q[4], q[5]);
    return bundle_reaches(c, rho, a, tn, mk3(q[0], q[1], q[2]), sqrtf(dot(he, he)));
  }
  const float* q = ob.triangles + 9 * (j - c3);
  const f3 v0 = mk3(q[0], q[1], q[2]), v1 = mk3(q[3], q[4], q[5]);
  const f3 v2 = mk3(q[6], q[7], q[8]);
  const f3 ctr = (1.f / 3.f) * (v0 + v1 + v2);
  const f3 e0 = v0 - ctr, e1 = v1 - ctr, e2 = v2 - ctr;
  const float R = sqrtf(fmaxf(fmaxf(dot(e0, e0), dot(e1, e1)), dot(e2, e2)));
  return bundle_reaches(c, rho, a, tn, ctr, R);
}

// Can a ray of the bundle reach cylinder q (p1, unit axis, height, radius)?
// The cylinder lies in the capsule of its axis segment, so it is necessary
// that bundle_reaches holds for some axis point b(t) = p1 + t axis, t in
// [0, h], with R = the radius: along the segment the axial coordinate s(t)
// and the bound g(t) = rr + tan(theta) (s(t) + rr) are linear and the radial
// distance squared e(t)^2 quadratic, so the test is whether
// Q(t) = e(t)^2 - g(t)^2 <= 0 somewhere on the interval where s(t) >= -rr
// and g(t) >= 0 (ends and the vertex of Q).  Q gets a tolerance of 1e-5 of
// its terms' magnitude; NaN geometry keeps the candidate.
IACT_DEV bool cylinder_reaches(const float* q, f3 c, float rho, f3 a, float tn) {
  const f3 p1 = mk3(q[0], q[1], q[2]), ax = mk3(q[3], q[4], q[5]);
  const float h = q[6], r = q[7];
  const float rr = fmaf(r, 1.001f, 1e-3f) + rho;
  const f3 w0 = p1 - c;
  const float s0 = dot(w0, a), s1 = dot(ax, a);
  const f3 P0 = fma3(-s0, a, w0), P1 = fma3(-s1, a, ax);
  const float G0 = fmaf(tn, s0 + rr, rr), G1 = tn * s1;
  // interval of t with s(t) + rr >= 0 and g(t) >= 0
  float lo = 0.f, hi = h;
  auto clip = [&](float k0, float k1) {  // k0 + k1 t >= 0
    if (k1 > 0.f) lo = fmaxf(lo, -k0 / k1);
    else if (k1 < 0.f) hi = fminf(hi, -k0 / k1);
    else if (k0 < 0.f) hi = -1.f;
  };
  clip(s0 + rr, s1);
  clip(G0, G1);
  const float qa = dot(P1, P1) - G1 * G1;
  const float qb = 2.f * (dot(P0, P1) - G0 * G1);
  const float qc = dot(P0, P0) - G0 * G0;
  auto Q = [&](float t) { return fmaf(fmaf(qa, t, qb), t, qc); };
  float m = fminf(Q(lo), Q(hi));
  if (qa > 0.f) m = fminf(m, Q(fminf(fmaxf(-qb / (2.f * qa), lo), hi)));
  const float hh = fmaxf(fabsf(lo), fabsf(hi));
  const float tol =
      1e-5f * (dot(P0, P0) + G0 * G0 + hh * hh * (dot(P1, P1) + G1 * G1) + 1e-6f);
  if (!(lo <= hi)) return lo != lo || hi != hi;  // empty interval (NaN: keep)
  return !(m > tol);
}

// Shadow bits of a facet run of far sources (source_qs 0): bit i is set when
// the shadow ray of the run's sample i towards the lane's source -- the
// direction dsh = -d, fixed over the run -- meets one of the candidates that
// shadow_factor_list would test (list, bits: one obstruction per lane, the
// wave-uniform selection).  Candidates are the outer loop, so a cylinder's
// and a box's direction terms are formed once per run; the tests are the
// per-ray kernels' own expressions (hit_cylinder = hit_cylinder_dir of
// cyl_dir, and so on), so every ray is decided as they decide it.  The run's
// n <= 64 sample points come one per lane (vector loads of the records' p
// words) and are handed out with v_readlane.  Before its samples a
// candidate is tested per lane with the lane's own direction against the
// ball (gc, grho) that holds the run's samples (obstruction_reaches, the
// culling test with a bundle of zero angle): a lane it cannot reach counts
// as done for it, and a candidate that reaches no lane is skipped -- a ray
// it cannot reach it does not block, so the bits are unchanged.
IACT_DEV uint64_t far_run_shadow(const TraceArgs& A, const float* rec0, int n, int list,
                                 uint64_t bits, f3 dsh, f3 gc, float grho) {
  const iact_obstructions& ob = A.ob;
  const int lane = lane_id();
  const float* pr = rec0 + (size_t)IACT_REC_WORDS * (lane < n ? lane : 0) + 13;
  const f3 myp = mk3(pr[0], pr[1], pr[2]);
  const int c0 = ob.n_cylinders, c1 = c0 + ob.n_boxes;
  uint64_t blocked = 0ull;
  while (bits != 0ull) {
    const int pos = __builtin_ctzll(bits);
    bits &= bits - 1ull;
    const int j = __builtin_amdgcn_readlane(list, pos);
    const bool out = !obstruction_reaches(ob, j, gc, grho, dsh, 0.f);
    if (__all(out)) continue;
    if (j < c0) {
      const cptr<float> c = as_const(ob.cylinders) + 8 * j;
      const CylDir D = cyl_dir(dsh, c);
      for (int i = 0; i < n; ++i) {
        if (__all(out || ((blocked >> i) & 1ull))) continue;
        if (hit_cylinder_dir(lane_value(myp, i), D, c) < 1e10f) blocked |= 1ull << i;
      }
    } else if (j < c1) {
      const cptr<float> b = as_const(ob.boxes) + 6 * (j - c0);
      const BoxDir D = box_dir(dsh);
      for (int i = 0; i < n; ++i) {
        if (__all(out || ((blocked >> i) & 1ull))) continue;
        if (hit_box_dir(lane_value(myp, i), D, b) < 1e10f) blocked |= 1ull << i;
      }
    } else {
      for (int i = 0; i < n; ++i) {
        if (__all(out || ((blocked >> i) & 1ull))) continue;
        if (hit_obstruction(ob, j, lane_value(myp, i), dsh) < 1e10f) blocked |= 1ull << i;
      }
    }
  }
  return blocked;
}

// far_run_shadow for a lane's two sources (two source chunks) at once: the
// union of both chunks' candidates is walked once and each candidate is
// tested along both directions, sharing the sample's terms (the origin's
// part of the cylinder test, the box's slab offsets).  A candidate tested for
// the other chunk's rays is an obstruction like any other: a ray it blocks
// is blocked, so the bits are those of two far_run_shadow calls.
IACT_DEV void far_run_shadow2(const TraceArgs& A, const float* rec0, int n, const int* list,
                              const uint64_t* bits, const f3* dsh, f3 gc, float grho,
                              uint64_t* blk) {
  const iact_obstructions& ob = A.ob;
  const int lane = lane_id();
  const float* pr = rec0 + (size_t)IACT_REC_WORDS * (lane < n ? lane : 0) + 13;
  const f3 myp = mk3(pr[0], pr[1], pr[2]);
  const int c0 = ob.n_cylinders, c1 = c0 + ob.n_boxes;
  const bool in0 = (bits[0] >> lane) & 1ull;
  uint64_t b0 = 0ull, b1 = 0ull;
#pragma unroll
  for (int k0 = 0; k0 < 2; ++k0) {
    uint64_t bb = bits[k0];
    while (bb != 0ull) {
      const int pos = __builtin_ctzll(bb);
      bb &= bb - 1ull;
      const int j = __builtin_amdgcn_readlane(list[k0], pos);
      // (a candidate of both chunks is walked with the first)
      if (k0 == 1 && __ballot(in0 && list[0] == j) != 0ull) continue;
      // directions the candidate cannot reach count as done (far_run_shadow)
      const uint64_t o0 = obstruction_reaches(ob, j, gc, grho, dsh[0], 0.f) ? 0ull : ~0ull;
      const uint64_t o1 = obstruction_reaches(ob, j, gc, grho, dsh[1], 0.f) ? 0ull : ~0ull;
      if (__all((o0 & o1) != 0ull)) continue;
      if (j < c0) {
        const cptr<float> c = as_const(ob.cylinders) + 8 * j;
        const CylDir D0 = cyl_dir(dsh[0], c), D1 = cyl_dir(dsh[1], c);
        for (int i = 0; i < n; ++i) {
          if (__all((((b0 | o0) >> i) & ((b1 | o1) >> i) & 1ull) != 0ull)) continue;
          float t0, t1;
          hit_cylinder_dir2(lane_value(myp, i), D0, D1, c, t0, t1);
          if (t0 < 1e10f) b0 |= 1ull << i;
          if (t1 < 1e10f) b1 |= 1ull << i;
        }
      } else if (j < c1) {
        const cptr<float> b = as_const(ob.boxes) + 6 * (j - c0);
        const BoxDir D0 = box_dir(dsh[0]), D1 = box_dir(dsh[1]);
        for (int i = 0; i < n; ++i) {
          if (__all((((b0 | o0) >> i) & ((b1 | o1) >> i) & 1ull) != 0ull)) continue;
          const f3 o = lane_value(myp, i);
          if (hit_box_dir(o, D0, b) < 1e10f) b0 |= 1ull << i;
          if (hit_box_dir(o, D1, b) < 1e10f) b1 |= 1ull << i;
        }
      } else {
        for (int i = 0; i < n; ++i) {
          if (__all((((b0 | o0) >> i) & ((b1 | o1) >> i) & 1ull) != 0ull)) continue;
          const f3 o = lane_value(myp, i);
          if (hit_obstruction(ob, j, o, dsh[0]) < 1e10f) b0 |= 1ull << i;
          if (hit_obstruction(ob, j, o, dsh[1]) < 1e10f) b1 |= 1ull << i;
        }
      }
    }
  }
  blk[0] = b0;
  blk[1] = b1;
}

// Trace one (source, sample) ray to the sensor plane (render.py:118-157).
// GEN false (image traces of launches whose samples all satisfy
// plane_test_redundant, on square pixels) leaves out the parallel-ray test
// and the pixel aspect ratio; ok may then hold for a ray that lands far off
// the grid, which binning drops.
// e is the facet's source offset (source_offset), qs its source_qs.
// Returns false when the ray misses the plane (the reference's 1e10
// sentinel) or the lane is inactive; v is 0 for inactive lanes.  (gu, gv) is
// the hit in sensor-grid coordinates (what the binning reads); XY also gives
// the hit (x, y) in the sensor frame.
template <int SRC, int KIND, bool STAGES, bool AXIS, int OBST = 2, bool XY = false,
          bool GEN = true, bool FAR = false>
IACT_DEV bool trace_ray(const TraceArgs& A, const Sample& s, bool active, f3 e, float qs,
                        float value, float& gu, float& gv, float& v, float& x, float& y,
                        float* vfac = nullptr, float* cup = nullptr) {
  // With m = -2 n, d.m = -2 (d.n) and value * (d.m) * nrwh = value * (d.n) *
  // nrw, both exactly (scalings by powers of two).
  float inv, c2, cu, rs = 1.f;
  f3 d, u;
  if (SRC == IACT_SOURCE_POINT && (!STAGES || A.n_stages == 0)) {
    // The reflection of u = q + e, r = u + (u.m) n, is |u| times the unit
    // reflection, and the sensor hit depends only on its direction: u is not
    // normalised, 1/|u| enters the cosine and the parallel-ray test only
    // (the unit direction d is formed for the shadow test alone).  FAR: the
    // caller knows qs = 0 for every lane (u = e, loop-invariant).
    u = FAR ? e : fma3(qs, s.q, e);
    inv = frsq(dot(u, u));
    cu = dot(u, s.m);
    c2 = cu * inv;
    rs = inv;
    d = inv * u;
  } else {
    d = primary_direction<SRC>(s, e, qs, inv);
    cu = dot(d, s.m);
    c2 = cu;
    u = d;
  }
  // vfac: the value before the weight, v = vfac * nrwh without obstructions
  // or further stages
  if (vfac) *vfac = value * c2;
  if (cup) *cup = cu;  // vfac = value cu rs for point sources
  v = value * c2 * s.nrwh;
  v *= primary_shadow<OBST>(A, s, d, active);
  bool ok;
  if (!STAGES || A.n_stages == 0) {
    // the hit of r = cu n + u from the sample's N and the direction's U
    // (grid_hit; U is loop-invariant for far and parallel sources); every
    // kernel bins a primary ray this way, so their pixel decisions agree
    ok = grid_hit<GEN>(s.G, s.N, sensor_dirs(A.S, u), cu, gu, gv, rs);
    if (XY) plane_xy<AXIS>(A.S, s.K, fma3(cu, s.n, u), x, y, rs);
  } else {
    f3 r = fma3(cu, s.n, u);
    f3 o = s.p;
    for (int k = 0; k < A.n_stages; ++k) reflect_at_stage(A, A.stages[k], o, r, v, active);
    const PlaneK K = plane_consts(A.S, o);
    const GridK G = grid_consts<KIND>(A.S, K);
    ok = plane_grid<KIND, AXIS, true>(A.S, G, r, gu, gv, 1.f);
    if (XY) plane_xy<AXIS>(A.S, K, r, x, y, 1.f);
  }
  if (!active) v = 0.f;
  return ok && active;
}

// Hard pixel of a traced ray's hit in sensor-grid coordinates; -1 when
// dropped or when the trace rejected the ray (ok false).
// EDGE false: the caller knows that the sensor has no edge test.
template <int KIND, bool EDGE = true>
IACT_DEV int grid_pixel(const DevSensor& S, const int32_t* table, float gu, float gv, bool ok) {
  if (KIND == IACT_SENSOR_SQUARE) return square_pixel_hit<EDGE>(S, gu, gv, ok);
  return hex_pixel_hit<EDGE>(S, table, gu, gv, ok);
}

// Hard pixel of a hit (x, y) in the sensor frame.
template <int KIND>
IACT_DEV int hard_pixel(const DevSensor& S, const int32_t* table, float x, float y) {
  if (KIND == IACT_SENSOR_SQUARE) return square_pixel(S, x, y);
  return hex_pixel(S, table, x, y);
}

// Sensor-frame (x, y) -> sensor-grid coordinates.
template <int KIND>
IACT_DEV void to_grid(const DevSensor& S, float x, float y, float& gu, float& gv) {
  if (KIND == IACT_SENSOR_SQUARE) {
    gu = fmaf(x, S.inv_dx, S.nx0);
    gv = fmaf(y, S.inv_dy, S.ny0);
  } else {
    hex_axial(S, x, y, gu, gv);
  }
}

// Splat taps of one ray (square.py:144-173, hexagonal.py:264-315).  Calls
// fn(pixel, weight) for every tap; weights are normalised to sum 1.  Rays
// that miss the plane or whose normaliser underflows contribute nothing.
// The hit is given in sensor-grid coordinates (gu, gv).
template <int KIND, typename Fn>
IACT_DEV void splat_taps(const DevSensor& S, const int32_t* table, float gu, float gv,
                         bool ok_in, Fn&& fn) {
  const int ks = S.ks;
  if (KIND == IACT_SENSOR_SQUARE) {
    const float xp = gu, yp = gv;
    const float xb = floorf(xp), yb = floorf(yp);
    const float xf = xp - xb, yf = yp - yb;
    float tot = 0.f;
    for (int oy = -ks; oy <= ks; ++oy)
      for (int ox = -ks; ox <= ks; ++ox) {
        const float ex = xf - (float)ox, ey = yf - (float)oy;
        tot += __expf(S.neg_half_inv_sigma2 * fmaf(ex, ex, ey * ey));
      }
    const bool ok = ok_in && tot > 0.f && fabsf(xb) < 1e8f && fabsf(yb) < 1e8f;
    const float it = ok ? frcp(tot) : 0.f;
    for (int oy = -ks; oy <= ks; ++oy)
      for (int ox = -ks; ox <= ks; ++ox) {
        const float ex = xf - (float)ox, ey = yf - (float)oy;
        const float w = __expf(S.neg_half_inv_sigma2 * fmaf(ex, ex, ey * ey)) * it;
        const float xi = xb + (float)ox, yi = yb + (float)oy;
        const bool in = ok && xi >= 0.f && xi < S.wf && yi >= 0.f && yi < S.hf;
        fn(in ? (int)fmaf(yi, S.wf, xi) : -1, w);
      }
  } else {
    const float q = gu, r = gv;
    float qb, rb;
    hex_round(q, r, qb, rb);
    const float dq = q - qb, dr = r - rb;
    float tot = 0.f;
    for (int oq = -ks; oq <= ks; ++oq)
      for (int orr = -ks; orr <= ks; ++orr) {
        if (max(max(abs(oq), abs(orr)), abs(oq + orr)) > ks) continue;
        const float hd = hex_norm_axial(dq - (float)oq, dr - (float)orr);
        tot += __expf(S.neg_half_inv_sigma2 * hd * hd);
      }
    const bool ok = ok_in && tot > 0.f && fabsf(qb) < 1e8f && fabsf(rb) < 1e8f;
    const float it = ok ? frcp(tot) : 0.f;
    for (int oq = -ks; oq <= ks; ++oq)
      for (int orr = -ks; orr <= ks; ++orr) {
        if (max(max(abs(oq), abs(orr)), abs(oq + orr)) > ks) continue;
        const float hd = hex_norm_axial(dq - (float)oq, dr - (float)orr);
        const float w = __expf(S.neg_half_inv_sigma2 * hd * hd) * it;
        const int pid = ok ? hex_lookup(S, table, qb + (float)oq, rb + (float)orr) : -1;
        fn(pid, w);
      }
  }
}

}  // namespace iact
