// Adjoint (reverse-mode) kernel of the image render for primary-only scenes.
//
// The reference obtains gradients with jax.grad through render.py:174-220
// (no custom VJP).  Here every ray is recomputed exactly as the forward
// kernel traces it (same device functions, so the same pixel decisions) and
// its adjoint is propagated by hand:
//   image -> ray value (hard binning) or ray value + sensor-plane hit (splat)
//   -> reflection -> source direction -> world sample (point, normal, weight).
// A lane owns one MC sample and loops over a tile of sources; the adjoints of
// the sample's world point, world normal and 1/weight are summed over the
// sources in registers and pushed through the world transform (rotation
// angles, position) and, when requested, the sample's surface (aspheric
// coefficients) once per sample.  Source adjoints are wave-reduced per source.
// Nothing per ray is stored.
#include <math.h>
#include <string.h>

#include "iact_abi_internal.h"

namespace iact {

// dR/d(angle) of R = Rz(c) Ry(b) Rx(a) for each of the three angles (radians).
IACT_DEV void euler_derivs(float tip, float tilt, float rot, m3& Da, m3& Db, m3& Dc) {
  const float k = 0.017453292519943295f;
  float sa, ca, sb, cb, sc, cc;
  sincosf(tip * k, &sa, &ca);
  sincosf(tilt * k, &sb, &cb);
  sincosf(rot * k, &sc, &cc);
  // Rz Ry = [[cc cb, -sc, cc sb], [sc cb, cc, sc sb], [-sb, 0, cb]]
  // dRx/da = [[0,0,0],[0,-sa,-ca],[0,ca,-sa]]
  const float A[9] = {cc * cb, -sc, cc * sb, sc * cb, cc, sc * sb, -sb, 0.f, cb};
  for (int i = 0; i < 3; ++i) {
    Da.m[3 * i + 0] = 0.f;
    Da.m[3 * i + 1] = A[3 * i + 1] * (-sa) + A[3 * i + 2] * ca;
    Da.m[3 * i + 2] = A[3 * i + 1] * (-ca) + A[3 * i + 2] * (-sa);
  }
  // dR/db = Rz dRy/db Rx, dRy/db = [[-sb,0,cb],[0,0,0],[-cb,0,-sb]]
  // Rz dRy = [[-cc sb, 0, cc cb], [-sc sb, 0, sc cb], [-cb, 0, -sb]]
  const float B[9] = {-cc * sb, 0.f, cc * cb, -sc * sb, 0.f, sc * cb, -cb, 0.f, -sb};
  for (int i = 0; i < 3; ++i) {
    Db.m[3 * i + 0] = B[3 * i + 0];
    Db.m[3 * i + 1] = B[3 * i + 1] * ca + B[3 * i + 2] * sa;
    Db.m[3 * i + 2] = -B[3 * i + 1] * sa + B[3 * i + 2] * ca;
  }
  // dR/dc = dRz/dc Ry Rx, dRz/dc Ry = [[-sc cb, -cc, -sc sb], [cc cb, -sc, cc sb], [0,0,0]]
  const float C[9] = {-sc * cb, -cc, -sc * sb, cc * cb, -sc, cc * sb, 0.f, 0.f, 0.f};
  for (int i = 0; i < 3; ++i) {
    Dc.m[3 * i + 0] = C[3 * i + 0];
    Dc.m[3 * i + 1] = C[3 * i + 1] * ca + C[3 * i + 2] * sa;
    Dc.m[3 * i + 2] = -C[3 * i + 1] * sa + C[3 * i + 2] * ca;
  }
}

IACT_DEV float frob(const m3& A, const float* B) {
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) s = fmaf(A.m[i], B[i], s);
  return s;
}

// Adjoint of the sensor-plane hit: given dL/dx, dL/dy, accumulate dL/do and
// dL/dr for x = kx + t (r.u1), y = ky + t (r.u2), t = kt / (r.n).
IACT_DEV void plane_adjoint(const DevSensor& S, f3 r, float t, float gx, float gy, f3& go,
                            f3& gr) {
  const float nd = dot(r, S.nrm);
  const float ind = frcp(nd);
  const float ru = dot(r, S.u1), rv = dot(r, S.u2);
  // P1 = u1 - (r.u1 / r.n) n, P2 = u2 - (r.u2 / r.n) n
  const f3 P1 = fma3(-ru * ind, S.nrm, S.u1);
  const f3 P2 = fma3(-rv * ind, S.nrm, S.u2);
  const f3 g = fma3(gx, P1, gy * P2);
  go = go + g;
  gr = fma3(t, g, gr);
}

// Splat tap weights and their gradient w.r.t. the sensor-plane hit:
// returns sum_k G_k w_k and writes dL/dx, dL/dy for a ray of value v.
template <int KIND>
IACT_DEV float splat_adjoint(const DevSensor& S, const int32_t* table, const float* G, float x,
                             float y, bool ok_in, float v, float& gx, float& gy) {
  const int ks = S.ks;
  const float h2 = 2.f * S.neg_half_inv_sigma2;
  gx = 0.f;
  gy = 0.f;
  if (KIND == IACT_SENSOR_SQUARE) {
    const float xp = fmaf(x, S.inv_dx, S.nx0), yp = fmaf(y, S.inv_dy, S.ny0);
    const float xb = floorf(xp), yb = floorf(yp);
    const float xf = xp - xb, yf = yp - yb;
    float tot = 0.f, sx = 0.f, sy = 0.f;
    for (int oy = -ks; oy <= ks; ++oy)
      for (int ox = -ks; ox <= ks; ++ox) {
        const float ex = xf - (float)ox, ey = yf - (float)oy;
        const float e = __expf(S.neg_half_inv_sigma2 * fmaf(ex, ex, ey * ey));
        tot += e;
        sx = fmaf(e, ex, sx);
        sy = fmaf(e, ey, sy);
      }
    const bool ok = ok_in && tot > 0.f && fabsf(xb) < 1e8f && fabsf(yb) < 1e8f;
    if (!ok) return 0.f;
    const float it = frcp(tot);
    const float mx = sx * it, my = sy * it;
    float gw = 0.f, ax = 0.f, ay = 0.f;
    for (int oy = -ks; oy <= ks; ++oy)
      for (int ox = -ks; ox <= ks; ++ox) {
        const float ex = xf - (float)ox, ey = yf - (float)oy;
        const float w = __expf(S.neg_half_inv_sigma2 * fmaf(ex, ex, ey * ey)) * it;
        const float xi = xb + (float)ox, yi = yb + (float)oy;
        const bool in = xi >= 0.f && xi < S.wf && yi >= 0.f && yi < S.hf;
        const float g = in ? G[(int)yi * S.width + (int)xi] : 0.f;
        gw = fmaf(g, w, gw);
        ax = fmaf(g * w, ex - mx, ax);
        ay = fmaf(g * w, ey - my, ay);
      }
    gx = v * h2 * ax * S.inv_dx;
    gy = v * h2 * ay * S.inv_dy;
    return gw;
  } else {
    float q, r, qb, rb;
    hex_axial(S, x, y, q, r);
    hex_round(q, r, qb, rb);
    const float dq = q - qb, dr = r - rb;
    float tot = 0.f, sq = 0.f, sr = 0.f;
    // d ln e_k / d(dq, dr) = h2 hd_k (d hd_k / d(dq, dr))
    for (int oq = -ks; oq <= ks; ++oq)
      for (int orr = -ks; orr <= ks; ++orr) {
        if (max(max(abs(oq), abs(orr)), abs(oq + orr)) > ks) continue;
        const float a0 = dq - (float)oq, b0 = dr - (float)orr;
        const float hd = hex_norm_axial(a0, b0);
        const float e = __expf(S.neg_half_inv_sigma2 * hd * hd);
        const float ee = fmaf(0.5f, b0, a0);
        const float se = ee > 0.f ? 1.f : (ee < 0.f ? -1.f : 0.f);
        const float sb = b0 > 0.f ? 1.f : (b0 < 0.f ? -1.f : 0.f);
        const bool first = 2.f * fabsf(ee) >= fmaf(1.5f, fabsf(b0), fabsf(ee));
        const float da = first ? 2.f * se : se;
        const float db = first ? se : fmaf(0.5f, se, 1.5f * sb);
        tot += e;
        sq = fmaf(e, h2 * hd * da, sq);
        sr = fmaf(e, h2 * hd * db, sr);
      }
    const bool ok = ok_in && tot > 0.f && fabsf(qb) < 1e8f && fabsf(rb) < 1e8f;
    if (!ok) return 0.f;
    const float it = frcp(tot);
    const float mq = sq * it, mr = sr * it;
    float gw = 0.f, aq = 0.f, ar = 0.f;
    for (int oq = -ks; oq <= ks; ++oq)
      for (int orr = -ks; orr <= ks; ++orr) {
        if (max(max(abs(oq), abs(orr)), abs(oq + orr)) > ks) continue;
        const float a0 = dq - (float)oq, b0 = dr - (float)orr;
        const float hd = hex_norm_axial(a0, b0);
        const float w = __expf(S.neg_half_inv_sigma2 * hd * hd) * it;
        const float ee = fmaf(0.5f, b0, a0);
        const float se = ee > 0.f ? 1.f : (ee < 0.f ? -1.f : 0.f);
        const float sb = b0 > 0.f ? 1.f : (b0 < 0.f ? -1.f : 0.f);
        const bool first = 2.f * fabsf(ee) >= fmaf(1.5f, fabsf(b0), fabsf(ee));
        const float da = first ? 2.f * se : se;
        const float db = first ? se : fmaf(0.5f, se, 1.5f * sb);
        const int pid = hex_lookup(S, table, qb + (float)oq, rb + (float)orr);
        const float g = pid >= 0 ? G[pid] : 0.f;
        gw = fmaf(g, w, gw);
        aq = fmaf(g * w, h2 * hd * da - mq, aq);
        ar = fmaf(g * w, h2 * hd * db - mr, ar);
      }
    const float gq = v * aq, gr = v * ar;
    gx = fmaf(gq, S.aqx, gr * S.arx);
    gy = fmaf(gq, S.aqy, gr * S.ary);
    return gw;
  }
}

// Stages >= 1 whose reflections the adjoint follows (the host rejects more).
#define IACT_ADJ_MAX_STAGES 2

// Add per-lane fp64 values into table[key] for lanes with key >= 0: lanes
// sharing a key are summed across the wave first (one atomic per key).
IACT_DEV void wave_add_keyed(double* table, int key, double v) {
  int k = key;
  for (int round = 0; round < 64; ++round) {
    const unsigned long long live = __ballot(k >= 0);
    if (live == 0ull) break;
    const int leader = __ffsll((long long)live) - 1;
    const int kl = __shfl(k, leader);
    const bool mine = k == kl;
    const double tot = wave_sum_d(mine ? v : 0.0);
    if (lane_id() == leader) unsafeAtomicAdd(table + kl, tot);
    if (mine) k = -1;
  }
}

// Adjoint of one stage>=1 reflection (stage_forward).  In: the record and the
// adjoints (go, gr, gv) of the outgoing origin, direction and value; out, in
// place: the adjoints of the incoming ones.  The hit is the converged root of
// F(t) = ol.z + t dl.z - S(ol.x + t dl.x, ol.y + t dl.y) in the mirror's
// frame, so by the implicit function theorem dt = -(m.dol + t m.ddl -
// dS/da da) / (m.dl) with m = (-Sx, -Sy, 1); the hit point is (x, y, S(x, y))
// and the normal m / |m|, whose change with (x, y) needs the sag's Hessian.
// Gradients of the hit surface's aspheric coefficients go to g_asph.
IACT_DEV void stage_adjoint(const AdjArgs& B, const StageRec& rec, f3& go, f3& gr, float& gv,
                            bool want_asph, double* asph_table) {
  const TraceArgs& A = B.T;
  const f3 zero = mk3(0.f, 0.f, 0.f);
  const bool live = rec.mi >= 0 && rec.hs != 0.f;
  // a stage that zeroes the value zeroes every adjoint downstream of it
  const float* M = A.mirrors + (size_t)IACT_MIRROR_STRIDE * (live ? rec.mi : 0);
  const m3 R = mirror_rotation(M);
  const f3 pos = mk3(M[0], M[1], M[2]);
  const int si = (int)M[38];
  const SurfT<float> su = load_surface<float>(A.surfaces[si]);
  const float x0 = M[15], y0 = M[16];
  const float z0 = sag_raw<float>(su, x0, y0);
  const SurfD D = surface_derivs(su, rec.xh, rec.yh, x0, y0, z0);
  const f3 m = mk3(-D.Sx, -D.Sy, 1.f);
  const float imn = frsq(dot(m, m));
  const f3 nl = imn * m;
  const f3 bn = mul(R, nl);
  const float c2 = rec.c2;
  // value: v' = v hs |c2|; direction: r' = r - 2 c2 bn, c2 = r.bn
  const float sgn = c2 > 0.f ? 1.f : (c2 < 0.f ? -1.f : 0.f);
  const float gc2 = fmaf(gv * rec.v, sgn, -2.f * dot(bn, gr));
  const f3 gbn = fma3(gc2, rec.r, (-2.f * c2) * gr);
  f3 gr_in = fma3(gc2, bn, gr);
  const float gv_in = gv * fabsf(c2);
  // hit point and normal in the mirror frame
  const f3 gpl = mulT(R, go);
  const f3 gnl = mulT(R, gbn);
  const f3 gm = imn * fma3(-dot(nl, gnl), nl, gnl);
  const float gSx = -gm.x, gSy = -gm.y, gS = gpl.z;
  const float gxh = gpl.x + gS * D.Sx + gSx * D.Sxx + gSy * D.Sxy;
  const float gyh = gpl.y + gS * D.Sy + gSx * D.Sxy + gSy * D.Syy;
  const f3 dl = mulT(R, rec.r);
  const float md = dot(m, dl);
  const float k = (gxh * dl.x + gyh * dl.y) * frcp(md);
  const f3 gol = mk3(gxh - k * m.x, gyh - k * m.y, -k * m.z);
  const f3 go_in = mul(R, gol);
  gr_in = fma3(rec.t, mul(R, gol), gr_in);  // gdl = t gol
  if (want_asph) {
    const float r20 = fmaf(x0, x0, y0 * y0);
    const float r4 = D.r2 * D.r2, r40 = r20 * r20;
    float pw = r4, pw0 = r40, q = 1.f;
#pragma unroll
    for (int i = 0; i < IACT_MAX_ASPHERIC; ++i) {
      const bool use = live && i < su.na;
      const float dS = pw - pw0;                      // d S / d a_i
      const float dg = (float)(2 * (i + 1)) * D.r2 * q;  // d (dz/dr2) / d a_i
      const float ga = fmaf(gS + k, dS, 2.f * dg * (gSx * D.X + gSy * D.Y));
      wave_add_keyed(asph_table, use ? IACT_MAX_ASPHERIC * si + i : -1, (double)ga);
      pw *= r4;
      pw0 *= r40;
      q *= r4;
    }
  }
  go = live ? go_in : zero;
  gr = live ? gr_in : zero;
  gv = live ? gv_in : 0.f;
}

// BATCH: many instances (TraceArgs batch fields); the LDS holds the image
// gradient of the instance the block is working on and is reloaded when the
// block moves on to the next instance; only rotation gradients, per
// instance, are produced.
// AXIS: the sensor frame is the world frame (the binning then reads the
// direction's components instead of three dot products), as in the forward.
template <int SRC, int KIND, int SPLAT, bool BATCH, bool STAGES, bool AXIS>
__global__ __launch_bounds__(IACT_MAX_BLOCK) void k_trace_adj(AdjArgs B) {
  const TraceArgs& A = B.T;
  extern __shared__ float smem[];
  const int P = A.S.n_pixels;
  float* G = smem;
  for (int i = threadIdx.x; i < P; i += blockDim.x) G[i] = B.grad_image[i];
  const int32_t* tab = KIND == IACT_SENSOR_HEX ? stage_hex_table(A.S, (int32_t*)(smem + P))
                                               : (const int32_t*)(smem + P);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), wpb = block_waves();
  const int lane = lane_id();
  const bool want_src = !BATCH && B.g_src != nullptr, want_val = !BATCH && B.g_val != nullptr;
  const bool want_asph = !BATCH && B.g_asph != nullptr && B.surface_index != nullptr &&
                         B.offsets != nullptr;
  const bool want_stage_asph = STAGES && B.g_asph != nullptr;
  // stage mirrors' aspheric gradients: summed per block in LDS (every wave
  // of every block adds to the same few coefficients, which as global
  // atomics serialise) and added to g_asph once per block at the end
  __shared__ double s_gasph[STAGES ? IACT_MAX_SURFACES * IACT_MAX_ASPHERIC : 1];
  if (want_stage_asph) {
    for (int i = threadIdx.x; i < IACT_MAX_SURFACES * IACT_MAX_ASPHERIC; i += blockDim.x)
      s_gasph[i] = 0.0;
    __syncthreads();
  }
  const long long n_items = total_items(A);
  const long long n_bitems = (n_items + wpb - 1) / wpb;
  int cur = 0;  // instance whose image gradient the LDS holds

  for (long long bi = xcd_block(A.xcd_order); bi < n_bitems; bi += grid_blocks()) {
    const int inst = item_instance(A, bi * wpb);  // block-uniform
    if (BATCH && inst != cur) {
      __syncthreads();
      for (int i = threadIdx.x; i < P; i += blockDim.x)
        G[i] = B.grad_image[(size_t)inst * P + i];
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
    const float* rot = A.rotations + w.inst * A.rot_stride;
    double* const g_rot = B.g_rot ? B.g_rot + (size_t)w.inst * 3 * A.nf : nullptr;
    const f3 pos = mk3(A.positions[3 * f], A.positions[3 * f + 1], A.positions[3 * f + 2]);
    f3 nw;
    Sample smp;
    {
      const m3 R = facet_rotation(rot, f);
      f3 pw;
      world_sample(A.points, A.normals, A.delta, R, pos, A.pscale[f], gi, pw, nw);
      smp = make_sample<KIND>(A.S, pw, nw, A.weights[gi], pos);
    }
    f3 Ap = mk3(0.f, 0.f, 0.f), An = mk3(0.f, 0.f, 0.f);
    float Anrw = 0.f;
    // hard binning through the primary leg only: the point-source term of
    // the sample's adjoint is sum_s t_s (n - c_s d_s), t_s = dc_s / |u_s|,
    // with n fixed in the source loop -- summed as sum t_s and sum t_s c_s d_s
    constexpr bool SPLIT_AP = SRC == IACT_SOURCE_POINT && !SPLAT && !STAGES;
    float Tn = 0.f;
    f3 Wd = mk3(0.f, 0.f, 0.f);
    const int s0 = w.s0, s1 = w.s1;
    const double* src = A.sources + 3 * w.sbase;
    const cptr<float> vals = as_const(A.values + w.sbase);
    f3 e_lane = mk3(0.f, 0.f, 0.f);  // source offsets, 64 at a time (see k_trace)
    long long chunk = -1;            // candidate mask per source chunk (see k_trace)
    for (int si = s0; si < s1; ++si) {
      // the source index as an opaque scalar: derived lane values (s + lane)
      // cannot turn the loop's induction variable into a vector one
      const int s = __builtin_amdgcn_readfirstlane(si);
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
      const float qs = source_qs(A, e);
      const float value = vals[s];
      // ---- forward recompute (the operations of trace_ray)
      float inv;
      const f3 d = primary_direction<SRC>(smp, e, qs, inv);
      const float c = dot(d, smp.n);
      const f3 r = fma3(-2.f * c, smp.n, d);
      const float shadow = primary_shadow<2>(A, smp, d, active);
      const float vunit = active ? c * smp.nrw * shadow : 0.f;  // ray value / source value
      // stages >= 1 (render.py:44-79), recorded for their adjoint
      f3 of = smp.p, rf = r;
      float vf = value * vunit;
      StageRec rec[IACT_ADJ_MAX_STAGES];
      PlaneK K = smp.K;
      GridK Kg = smp.G;
      if (STAGES) {
#pragma unroll
        for (int k = 0; k < IACT_ADJ_MAX_STAGES; ++k)
          if (k < A.n_stages) stage_forward<true>(A, A.stages[k], of, rf, vf, active, rec[k]);
        K = plane_consts(A.S, of);
        Kg = grid_consts<KIND>(A.S, K);
      }
      // ---- image adjoint -> ray adjoints (pixels picked as the forward does)
      float gv, gx = 0.f, gy = 0.f;
      if (!SPLAT) {
        float gu_, gv_;
        bool ok;
        if (!STAGES || A.n_stages == 0) {
          // the hit as trace_ray forms it (grid_hit of r = (u.m) n + u with
          // u = q + e not normalised for point sources, u = d for parallel
          // ones), so that the pixel is the forward's
          const f3 u = SRC == IACT_SOURCE_POINT ? fma3(qs, smp.q, e) : d;
          ok = grid_hit<true>(Kg, smp.N, sensor_dirs(A.S, u), dot(u, smp.m), gu_, gv_,
                              SRC == IACT_SOURCE_POINT ? inv : 1.f) &&
               active;
        } else {
          ok = plane_grid<KIND, AXIS>(A.S, Kg, rf, gu_, gv_) && active;
        }
        const int pix = grid_pixel<KIND>(A.S, tab, gu_, gv_, ok);
        gv = pix >= 0 ? G[pix] : 0.f;
      } else {
        float x, y;
        const bool ok = plane_xy<false>(A.S, K, rf, x, y) && active;
        gv = splat_adjoint<KIND>(A.S, tab, G, x, y, ok, vf, gx, gy);
      }
      // the sensor-plane hit depends on the last origin and direction
      f3 go = mk3(0.f, 0.f, 0.f), gr = mk3(0.f, 0.f, 0.f);
      if (SPLAT) {
        const float t = K.kt * frcp(dot(rf, A.S.nrm));
        plane_adjoint(A.S, rf, t, gx, gy, go, gr);
      }
      if (STAGES) {
#pragma unroll
        for (int k = IACT_ADJ_MAX_STAGES - 1; k >= 0; --k)
          if (k < A.n_stages) stage_adjoint(B, rec[k], go, gr, gv, want_stage_asph, s_gasph);
      }
      // now (go, gr, gv) are the adjoints of the primary leg's origin (the
      // sample), reflected direction and value = value_s * c * nrw * shadow
      const float gvs = gv * value * shadow;
      const float dc = gvs * smp.nrw;
      if (want_asph) Anrw = fmaf(gvs, c, Anrw);
      An = fma3(dc, d, An);
      float gsx = 0.f, gsy = 0.f, gsz = 0.f;
      if (SPLIT_AP) {
        // d = u / |u|, u = p - src, dL/dd = dc n: dL/dp = t (n - c d),
        // t = dc / |u|, and dL/dsrc = -dL/dp
        const float t = dc * inv;
        Tn += t;
        Wd = fma3(t * c, d, Wd);
        if (want_src) {
          const f3 gu = t * fma3(-c, d, smp.n);
          gsx = -gu.x;
          gsy = -gu.y;
          gsz = -gu.z;
        }
      } else {
        Ap = Ap + go;
        f3 gd = dc * smp.n;
        if (SPLAT || STAGES) {
          // r = d - 2 (d.n) n
          const float ng = dot(smp.n, gr);
          gd = gd + fma3(-2.f * ng, smp.n, gr);
          An = An + fma3(-2.f * ng, d, (-2.f * c) * gr);
        }
        if (SRC == IACT_SOURCE_POINT) {
          // d = u / |u|, u = p - src
          const float dd = dot(d, gd);
          const f3 gu = inv * fma3(-dd, d, gd);
          Ap = Ap + gu;
          gsx = -gu.x;
          gsy = -gu.y;
          gsz = -gu.z;
        } else {
          gsx = gd.x;
          gsy = gd.y;
          gsz = gd.z;
        }
      }
      if (want_src || want_val) {
        const float a = want_src ? wave_sum(active ? gsx : 0.f) : 0.f;
        const float b = want_src ? wave_sum(active ? gsy : 0.f) : 0.f;
        const float cz = want_src ? wave_sum(active ? gsz : 0.f) : 0.f;
        const float e = want_val ? wave_sum(gv * vunit) : 0.f;
        if (lane == 0) {
          // sources the host sorted report to the caller's index
          const long long so = B.src_perm ? (long long)as_const(B.src_perm)[s] : (long long)s;
          if (want_src) {
            unsafeAtomicAdd(B.g_src + 3 * so, (double)a);
            unsafeAtomicAdd(B.g_src + 3 * so + 1, (double)b);
            unsafeAtomicAdd(B.g_src + 3 * so + 2, (double)cz);
          }
          if (want_val) unsafeAtomicAdd(B.g_val + so, (double)e);
        }
      }
    }
    if (SPLIT_AP) Ap = fma3(Tn, smp.n, mk3(-Wd.x, -Wd.y, -Wd.z));
    if (!active) {
      Ap = mk3(0.f, 0.f, 0.f);
      An = mk3(0.f, 0.f, 0.f);
      Anrw = 0.f;
    }
    // the sample's facet-frame data and pose are read again rather than kept
    // in registers through the source loop (the barrier stops the compiler
    // from reusing the first reads)
    asm volatile("" ::: "memory");
    const m3 R = facet_rotation(rot, f);
    const f3 pl = mk3(A.points[3 * gi], A.points[3 * gi + 1], A.points[3 * gi + 2]);
    const f3 nl = mk3(A.normals[3 * gi], A.normals[3 * gi + 1], A.normals[3 * gi + 2]);
    const f3 dl = mk3(A.delta[3 * gi], A.delta[3 * gi + 1], A.delta[3 * gi + 2]);
    const float scale = A.pscale[f];
    const float wgt = A.weights[gi];
    // ---- per-sample adjoint through the world transform (mirrors.py:64-79)
    // n = pert / |pert|, pert = R (nl + s dl);  p = R pl + pos
    const f3 nls = fma3(scale, dl, nl);
    const f3 pert = mul(R, nls);
    const float ipn = frsq(dot(pert, pert));
    const f3 gpert = ipn * fma3(-dot(nw, An), nw, An);
    if (g_rot) {
      float dR[9];
      dR[0] = fmaf(Ap.x, pl.x, gpert.x * nls.x);
      dR[1] = fmaf(Ap.x, pl.y, gpert.x * nls.y);
      dR[2] = fmaf(Ap.x, pl.z, gpert.x * nls.z);
      dR[3] = fmaf(Ap.y, pl.x, gpert.y * nls.x);
      dR[4] = fmaf(Ap.y, pl.y, gpert.y * nls.y);
      dR[5] = fmaf(Ap.y, pl.z, gpert.y * nls.z);
      dR[6] = fmaf(Ap.z, pl.x, gpert.z * nls.x);
      dR[7] = fmaf(Ap.z, pl.y, gpert.z * nls.y);
      dR[8] = fmaf(Ap.z, pl.z, gpert.z * nls.z);
      m3 Da, Db, Dc;
      euler_derivs(rot[3 * f], rot[3 * f + 1], rot[3 * f + 2], Da, Db, Dc);
      const float k = 0.017453292519943295f;
      const float ga = wave_sum(frob(Da, dR) * k);
      const float gb = wave_sum(frob(Db, dR) * k);
      const float gc = wave_sum(frob(Dc, dR) * k);
      if (lane == 0) {
        unsafeAtomicAdd(g_rot + 3 * f, (double)ga);
        unsafeAtomicAdd(g_rot + 3 * f + 1, (double)gb);
        unsafeAtomicAdd(g_rot + 3 * f + 2, (double)gc);
      }
    }
    if (!BATCH && B.g_pos) {
      const float a = wave_sum(Ap.x), b = wave_sum(Ap.y), cz = wave_sum(Ap.z);
      if (lane == 0) {
        unsafeAtomicAdd(B.g_pos + 3 * f, (double)a);
        unsafeAtomicAdd(B.g_pos + 3 * f + 1, (double)b);
        unsafeAtomicAdd(B.g_pos + 3 * f + 2, (double)cz);
      }
    }
    if (want_asph) {
      // the sample re-derived from its aperture position (x, y) on the
      // surface: pl.z = sag(x+x0, y+y0) - sag(x0, y0), nl = normalize(-zx, -zy, 1),
      // weight = nl.z * M / area (integrators.py:97-188); delta held fixed.
      const int si = B.surface_index[f];
      const iact_surface& su = A.surfaces[si];
      const float x0 = B.offsets[2 * f], y0 = B.offsets[2 * f + 1];
      const f3 gnl_w = mulT(R, gpert);   // dL/d nl (nl + s dl is rotated)
      const f3 gpl = mulT(R, Ap);        // dL/d pl
      // weight: nrw = -1/w, dL/dw = Anrw / w^2; w = nl.z * const
      const float dLdw = Anrw * frcp(wgt * wgt);
      f3 gnl = gnl_w;
      gnl.z = fmaf(dLdw, wgt * frcp(nl.z), gnl.z);
      // nl = mvec / |mvec|, mvec = (-zx, -zy, 1), |mvec| = 1 / nl.z
      const f3 gm = nl.z * fma3(-dot(nl, gnl), nl, gnl);
      const float X = pl.x + x0, Y = pl.y + y0;
      const float r2 = fmaf(X, X, Y * Y), r20 = fmaf(x0, x0, y0 * y0);
      const float r4 = r2 * r2, r40 = r20 * r20;
      float pw_ = r4, pw0 = r40, q = 1.f;
      for (int i = 0; i < IACT_MAX_ASPHERIC; ++i) {
        if (i >= su.n_aspheric) break;
        const float dz = pw_ - pw0;                         // d pl.z / d a_i
        const float dg = (float)(2 * (i + 1)) * r2 * q;     // d (dz/dr2) / d a_i
        const float dzx = 2.f * X * dg, dzy = 2.f * Y * dg;  // d (zx, zy) / d a_i
        const float contrib = active ? fmaf(gpl.z, dz, -(gm.x * dzx + gm.y * dzy)) : 0.f;
        const float tot = wave_sum(contrib);
        if (lane == 0) unsafeAtomicAdd(B.g_asph + IACT_MAX_ASPHERIC * si + i, (double)tot);
        pw_ *= r4;
        pw0 *= r40;
        q *= r4;
      }
    }
  }
  if (want_stage_asph) {
    __syncthreads();
    for (int i = threadIdx.x; i < IACT_MAX_SURFACES * IACT_MAX_ASPHERIC; i += blockDim.x) {
      const double v = s_gasph[i];
      if (v != 0.0) unsafeAtomicAdd(B.g_asph + i, v);
    }
  }
}

using AdjKernel = void (*)(AdjArgs);

template <int SRC, int KIND, int SPLAT, bool AXIS>
static AdjKernel pick_adj_ax(const AdjArgs& B) {
  if (B.T.item_prefix) return k_trace_adj<SRC, KIND, SPLAT, true, false, AXIS>;
  return B.T.n_stages > 0 ? k_trace_adj<SRC, KIND, SPLAT, false, true, AXIS>
                          : k_trace_adj<SRC, KIND, SPLAT, false, false, AXIS>;
}

template <int SRC, int KIND, int SPLAT>
static AdjKernel pick_adj_t(const AdjArgs& B) {
  // splat sensors bin through plane_xy, which has no axis variant
  return (B.T.S.axis && !SPLAT) ? pick_adj_ax<SRC, KIND, SPLAT, true>(B)
                                : pick_adj_ax<SRC, KIND, SPLAT, false>(B);
}

template <int SRC>
static AdjKernel pick_adj_src(const AdjArgs& B) {
  const DevSensor& S = B.T.S;
  if (S.kind == IACT_SENSOR_SQUARE)
    return S.splat ? pick_adj_t<SRC, IACT_SENSOR_SQUARE, 1>(B)
                   : pick_adj_t<SRC, IACT_SENSOR_SQUARE, 0>(B);
  return S.splat ? pick_adj_t<SRC, IACT_SENSOR_HEX, 1>(B) : pick_adj_t<SRC, IACT_SENSOR_HEX, 0>(B);
}

const void* pick_adjoint_kernel(const AdjArgs& B, int source_type) {
  return (const void*)(source_type == IACT_SOURCE_POINT ? pick_adj_src<IACT_SOURCE_POINT>(B)
                                                        : pick_adj_src<IACT_SOURCE_PARALLEL>(B));
}

hipError_t launch_adjoint(const AdjArgs& B, const LaunchShape& L, int source_type,
                          hipStream_t st) {
  if (has_obstructions(B.T.ob) && !B.T.cull) return hipErrorInvalidValue;
  AdjArgs b = B;
  void* args[] = {&b};
  return hipLaunchKernel(pick_adjoint_kernel(B, source_type), dim3(L.grid), dim3(L.threads),
                         args, L.lds, st);
}

}  // namespace iact

using namespace iact;

extern "C" {

size_t iact_bwd_workspace_size(const iact_scene* scene, int32_t n_sources,
                               const iact_launch* launch) {
  if (validate_scene(scene) || n_sources < 0) return 0;
  DevInfo dev;
  if (device_info(dev)) return 0;
  TraceArgs A;
  fill_trace_args(A, *scene, nullptr, nullptr, n_sources);
  LaunchShape shp;
  if (shape_trace(*scene, n_sources, IACT_MODE_DEBUG, launch, A, shp, dev, -1)) return 0;
  const CullSpec c{n_sources, 1, nullptr, true};
  size_t need = cull_layout(A.nf, A.cull_words, A.cull_all, 0, c).end;
  // (the source-major adjoint is not taken when aspheric gradients are
  // requested; the size covers both)
  if (bwd_source_major(*scene, n_sources, launch, false)) {
    TraceArgs B;
    AdjSmLayout lay;
    if (adj_sm_layout(*scene, nullptr, nullptr, n_sources, launch, dev, B, shp, lay)) return 0;
    if (lay.end > need) need = lay.end;
  }
  return need;
}

int32_t iact_trace_bwd(const iact_scene* scene, const double* sources, const float* values,
                       int32_t n_sources, int32_t source_type, const float* grad_image,
                       const iact_launch* launch, void* workspace, size_t workspace_bytes,
                       const iact_grads* grads, void* stream) {
  int rc = validate_scene(scene);
  if (rc) return rc;
  if (!grads) return set_error(IACT_ERR_INVALID, "null gradient descriptor");
  if (n_sources < 0) return set_error(IACT_ERR_INVALID, "negative source count");
  if (n_sources > 0 && (!sources || !values))
    return set_error(IACT_ERR_INVALID, "null sources/values");
  if (!grad_image) return set_error(IACT_ERR_INVALID, "null image gradient");
  if (source_type != IACT_SOURCE_POINT && source_type != IACT_SOURCE_PARALLEL)
    return set_error(IACT_ERR_INVALID, "unknown source type %d", source_type);
  const iact_scene& sc = *scene;
  if (sc.n_stages > IACT_ADJ_MAX_STAGES)
    return set_error(IACT_ERR_UNSUPPORTED, "the adjoint follows at most %d stages >= 1 (%d given)",
                     IACT_ADJ_MAX_STAGES, sc.n_stages);
  if (grads->aspheric && (!sc.facets.surface_index || !sc.facets.offsets))
    return set_error(IACT_ERR_INVALID, "aspheric gradients need facet surface_index/offsets");
  hipStream_t st = (hipStream_t)stream;
  const int nf = sc.facets.n_facets;
  hipError_t e;
  if (grads->rotations && (e = launch_fill_d(grads->rotations, 3ll * nf, 0.0, st)) != hipSuccess)
    return hip_error(e, "clear grad");
  if (grads->positions && (e = launch_fill_d(grads->positions, 3ll * nf, 0.0, st)) != hipSuccess)
    return hip_error(e, "clear grad");
  if (grads->sources &&
      (e = launch_fill_d(grads->sources, 3ll * n_sources, 0.0, st)) != hipSuccess)
    return hip_error(e, "clear grad");
  if (grads->values && (e = launch_fill_d(grads->values, n_sources, 0.0, st)) != hipSuccess)
    return hip_error(e, "clear grad");
  if (grads->aspheric &&
      (e = launch_fill_d(grads->aspheric, (long long)IACT_MAX_SURFACES * IACT_MAX_ASPHERIC, 0.0,
                         st)) != hipSuccess)
    return hip_error(e, "clear grad");
  if (nf == 0 || n_sources == 0) return IACT_OK;
  DevInfo dev;
  rc = device_info(dev);
  if (rc) return rc;
  AdjArgs B;
  memset(&B, 0, sizeof(B));
  if (bwd_source_major(sc, n_sources, launch, grads->aspheric != nullptr)) {
    // lanes = sources over the image trace's sample records (k_adj_sm)
    LaunchShape shp;
    AdjSmLayout lay;
    TraceArgs& A = B.T;
    rc = adj_sm_layout(sc, sources, values, n_sources, launch, dev, A, shp, lay);
    if (rc) return rc;
    if (lay.end > 0 && (!workspace || workspace_bytes < lay.end))
      return set_error(IACT_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", lay.end,
                       workspace_bytes);
    const size_t tab = hex_table_bytes(sc.sensor);
    shp.lds = (size_t)(sc.sensor.n_pixels + 1) * sizeof(float) + tab;  // G[-1] = 0 + G
    if (shp.lds > 65536)
      return set_error(IACT_ERR_UNSUPPORTED, "image gradient + tables need %zu bytes of LDS",
                       shp.lds);
    const CullSpec c{n_sources, 1, nullptr, true, refine_groups(A, IACT_MODE_IMAGE)};
    rc = prepare_cull(A, workspace, lay.recs_end, source_type, st, c, &B.src_perm);
    if (rc) return rc;
    A.recs = (const float*)workspace;
    e = launch_records(A, (float*)workspace, st);
    if (e != hipSuccess) return hip_error(e, "k_sample_records");
    if (A.gbounds) {
      e = launch_group_bounds(A, (float*)A.gbounds, st);
      if (e != hipSuccess) return hip_error(e, "k_group_bounds");
    }
    B.torque = (double*)((char*)workspace + lay.torque);
    e = launch_fill_d(B.torque, 3ll * nf, 0.0, st);
    if (e != hipSuccess) return hip_error(e, "clear torques");
    B.grad_image = grad_image;
    B.g_pos = grads->positions;
    B.g_src = grads->sources;
    B.g_val = grads->values;
    e = launch_adjoint_sm(B, shp, source_type, st);
    if (e != hipSuccess) return hip_error(e, "k_adj_sm");
    if (grads->rotations) {
      e = launch_torque_to_euler(B.torque, sc.facets.rotations, nf, grads->rotations, st);
      if (e != hipSuccess) return hip_error(e, "k_torque_to_euler");
    }
    return IACT_OK;
  }
  fill_trace_args(B.T, sc, sources, values, n_sources);
  LaunchShape shp;
  rc = shape_trace(sc, n_sources, IACT_MODE_DEBUG, launch, B.T, shp, dev, -1);
  if (rc) return rc;
  const CullSpec cspec{n_sources, 1, nullptr, true};
  const size_t need = cull_layout(B.T.nf, B.T.cull_words, B.T.cull_all, 0, cspec).end;
  if (need > 0 && (!workspace || workspace_bytes < need))
    return set_error(IACT_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need,
                     workspace_bytes);
  const size_t lds = (size_t)sc.sensor.n_pixels * sizeof(float) + shp.lds;
  // (kernels following stages >= 1 also hold the aspheric-gradient table)
  const size_t lds_static =
      sc.n_stages > 0 ? (size_t)IACT_MAX_SURFACES * IACT_MAX_ASPHERIC * sizeof(double) : 0;
  if (lds + lds_static > 65536)
    return set_error(IACT_ERR_UNSUPPORTED, "image gradient + tables need %zu bytes of LDS",
                     lds + lds_static);
  shp.lds = lds;
  B.grad_image = grad_image;
  B.g_rot = grads->rotations;
  B.g_pos = grads->positions;
  B.g_src = grads->sources;
  B.g_val = grads->values;
  B.g_asph = grads->aspheric;
  B.surface_index = sc.facets.surface_index;
  B.offsets = sc.facets.offsets;
  rc = regrid(B.T.n_items, launch, dev, shp, B.T);
  if (rc) return rc;
  rc = prepare_cull(B.T, workspace, 0, source_type, st, cspec, &B.src_perm);
  if (rc) return rc;
  e = launch_adjoint(B, shp, source_type, st);
  return e == hipSuccess ? IACT_OK : hip_error(e, "k_trace_adj");
}

}  // extern "C"
