// Device-side building blocks of the gfx950 ray tracer: vector helpers,
// Philox RNG, obstruction primitives, conic/asphere root refinement, sensor
// binning and wave-level pixel aggregation.  Every function operates on one
// ray held by one lane; tables are read with wave-uniform indices so they
// come through the scalar cache (or LDS for per-lane gathers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/iactrace_amd.h"

#define IACT_DEV __device__ __forceinline__

namespace iact {

// Read-only view in the constant address space: wave-uniform loads through
// it become scalar (s_load) loads served by the scalar cache.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
IACT_DEV cptr<T> as_const(const T* p) {
  return (cptr<T>)p;
}

// ---------------------------------------------------------------- vectors
struct f3 {
  float x, y, z;
};
IACT_DEV f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
IACT_DEV f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
IACT_DEV f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
IACT_DEV f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
IACT_DEV f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
IACT_DEV float dot(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
IACT_DEV f3 fma3(float s, f3 a, f3 b) {  // s * a + b
  return f3{fmaf(s, a.x, b.x), fmaf(s, a.y, b.y), fmaf(s, a.z, b.z)};
}
IACT_DEV f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// 3x3 row-major matrix
struct m3 {
  float m[9];
};
IACT_DEV f3 mul(const m3& R, f3 v) {  // R v
  return f3{fmaf(R.m[0], v.x, fmaf(R.m[1], v.y, R.m[2] * v.z)),
            fmaf(R.m[3], v.x, fmaf(R.m[4], v.y, R.m[5] * v.z)),
            fmaf(R.m[6], v.x, fmaf(R.m[7], v.y, R.m[8] * v.z))};
}
IACT_DEV f3 mulT(const m3& R, f3 v) {  // R^T v
  return f3{fmaf(R.m[0], v.x, fmaf(R.m[3], v.y, R.m[6] * v.z)),
            fmaf(R.m[1], v.x, fmaf(R.m[4], v.y, R.m[7] * v.z)),
            fmaf(R.m[2], v.x, fmaf(R.m[5], v.y, R.m[8] * v.z))};
}

// Fast hardware approximations (about 1 ulp).  The reference divides and
// takes square roots with IEEE rounding; these differ by at most a few ulp,
// which only moves rays that lie within the reported edge band.
IACT_DEV float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
IACT_DEV float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
IACT_DEV float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Euler (tip, tilt, rotation) in degrees -> R = Rz Ry Rx (transforms.py:72-107)
IACT_DEV m3 euler_to_matrix(float tip, float tilt, float rot) {
  const float k = 0.017453292519943295f;
  float sx, cx, sy, cy, sz, cz;
  sincosf(tip * k, &sx, &cx);
  sincosf(tilt * k, &sy, &cy);
  sincosf(rot * k, &sz, &cz);
  m3 R;
  R.m[0] = cz * cy;
  R.m[1] = cz * sy * sx - sz * cx;
  R.m[2] = cz * sy * cx + sz * sx;
  R.m[3] = sz * cy;
  R.m[4] = sz * sy * sx + cz * cx;
  R.m[5] = sz * sy * cx - cz * sx;
  R.m[6] = -sy;
  R.m[7] = cy * sx;
  R.m[8] = cy * cx;
  return R;
}

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11): 10 rounds of two 32x32->64 multiplies, Weyl key schedule.
struct u4 {
  uint32_t x, y, z, w;
};
IACT_DEV u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
IACT_DEV float u01(uint32_t u) { return (float)(u >> 8) * 5.9604644775390625e-08f; }
IACT_DEV float u01_open(uint32_t u) { return (float)((u >> 8) + 1u) * 5.9604644775390625e-08f; }

// ---------------------------------------------------------------- surfaces
// sag_raw and its slope (surfaces.py:25-58) in float or double.
template <typename T>
struct SurfT {
  T c, k, K;  // K = (1+k) c^2
  int na;
  T a[IACT_MAX_ASPHERIC];
};

template <typename T>
IACT_DEV SurfT<T> load_surface(const iact_surface& s) {
  SurfT<T> o;
  o.c = (T)s.curvature;
  o.k = (T)s.conic;
  o.K = ((T)1 + o.k) * o.c * o.c;
  o.na = s.n_aspheric;
#pragma unroll
  for (int i = 0; i < IACT_MAX_ASPHERIC; ++i) o.a[i] = (T)s.aspheric[i];
  return o;
}

template <typename T>
IACT_DEV T sqrt_t(T x);
template <>
IACT_DEV float sqrt_t<float>(float x) { return fsqrt(x); }
template <>
IACT_DEV double sqrt_t<double>(double x) { return __builtin_sqrt(x); }

template <typename T>
IACT_DEV T sag_raw(const SurfT<T>& s, T X, T Y) {
  const T r2 = X * X + Y * Y;
  const T den = (T)1 + sqrt_t<T>((T)1 - s.K * r2);
  T z = r2 * s.c / den;
  const T r4 = r2 * r2;
  T p = r4;
#pragma unroll
  for (int i = 0; i < IACT_MAX_ASPHERIC; ++i) {
    if (i < s.na) z += s.a[i] * p;
    p *= r4;
  }
  return z;
}

// sag and d sag / d r2 at the same point (shares the square root).
template <typename T>
IACT_DEV void sag_and_dr2(const SurfT<T>& s, T r2, T& z, T& g) {
  const T sq = sqrt_t<T>((T)1 - s.K * r2);
  const T den = (T)1 + sq;
  const T inv_den = (T)1 / den;
  z = r2 * s.c * inv_den;
  g = s.c * inv_den + (r2 * s.c) * s.K / ((T)2 * sq * den * den);
  const T r4 = r2 * r2;
  T p = r4, q = (T)1;
#pragma unroll
  for (int i = 0; i < IACT_MAX_ASPHERIC; ++i) {
    if (i < s.na) {
      z += s.a[i] * p;
      g += s.a[i] * (T)(2 * (i + 1)) * r2 * q;
    }
    p *= r4;
    q *= r4;
  }
}

// Local point (x, y, sag) and unit normal at (x, y) of a surface used with
// offset (x0, y0) (surfaces.py:41-65).
IACT_DEV void surface_point_normal(const SurfT<float>& s, float x, float y, float x0, float y0,
                                   float z0, f3& p, f3& n) {
  const float X = x + x0, Y = y + y0;
  float z, g;
  sag_and_dr2<float>(s, X * X + Y * Y, z, g);
  p = mk3(x, y, z - z0);
  const float zx = 2.f * X * g, zy = 2.f * Y * g;
  const float inv = frsq(zx * zx + zy * zy + 1.f);
  n = mk3(-zx * inv, -zy * inv, inv);
}

// S(x, y) = sag(x + x0, y + y0) - z0 with its gradient and Hessian at a
// local point (fp32; used by the adjoint of stage>=1 hits).  The gradient is
// the one surface_point_normal uses; the second derivative of the sag in r2
// is c K / (4 (1 - K r2)^(3/2)) + sum_i a_i p_i (p_i - 1) r2^(p_i - 2),
// p_i = 2(i+1).
struct SurfD {
  float S, Sx, Sy, Sxx, Sxy, Syy;
  float X, Y, r2;
};
IACT_DEV SurfD surface_derivs(const SurfT<float>& s, float x, float y, float x0, float y0,
                              float z0) {
  SurfD D;
  D.X = x + x0;
  D.Y = y + y0;
  D.r2 = D.X * D.X + D.Y * D.Y;
  float z, g;
  sag_and_dr2<float>(s, D.r2, z, g);
  const float sq = fsqrt(1.f - s.K * D.r2);
  float gp = s.c * s.K * 0.25f * frcp(sq * sq * sq);
  float q = 1.f;  // r2^(p_i - 2) = r2^(2i)
#pragma unroll
  for (int i = 0; i < IACT_MAX_ASPHERIC; ++i) {
    const float p = (float)(2 * (i + 1));
    if (i < s.na) gp = fmaf(s.a[i] * p * (p - 1.f), q, gp);
    q *= D.r2 * D.r2;
  }
  D.S = z - z0;
  D.Sx = 2.f * D.X * g;
  D.Sy = 2.f * D.Y * g;
  D.Sxx = fmaf(4.f * D.X * D.X, gp, 2.f * g);
  D.Sxy = 4.f * D.X * D.Y * gp;
  D.Syy = fmaf(4.f * D.Y * D.Y, gp, 2.f * g);
  return D;
}

// Closed-form conic root (intersections.py:229-285), float64.
IACT_DEV double intersect_conic(double ox, double oy, double oz, double dx, double dy, double dz,
                                double c, double k) {
  const double ok = 1.0 + k;
  const double A = c * (dx * dx + dy * dy + ok * dz * dz);
  const double B = 2.0 * (c * (ox * dx + oy * dy + ok * oz * dz) - dz);
  const double C = c * (ox * ox + oy * oy + ok * oz * oz) - 2.0 * oz;
  if (fabs(c) < 1e-12) return (fabs(dz) > 1e-10) ? -oz / dz : __builtin_inf();
  const double disc = B * B - 4.0 * A * C;
  if (disc < 0.0) return __builtin_inf();
  const double sq = __builtin_sqrt(disc);
  const double inv = 1.0 / (2.0 * A + 1e-30);
  const double t1 = (-B - sq) * inv, t2 = (-B + sq) * inv;
  const bool v1 = t1 > 1e-8, v2 = t2 > 1e-8;
  return (v1 && v2) ? fmin(t1, t2) : (v1 ? t1 : (v2 ? t2 : __builtin_inf()));
}

// Fixed-count Newton (intersections.py:290-367) in float64 with the
// reference's freeze rule: a lane stops updating one iteration after |g| < tol.
// Iterations after every lane has frozen (or gone non-finite, which can never
// become valid again) change nothing, so the loop leaves early when the whole
// wave is done -- the result is identical to running all ten.
IACT_DEV double newton_t(const SurfT<double>& s, double ox, double oy, double oz, double dx,
                         double dy, double dz, double x0, double y0, double z0, double t,
                         bool& valid) {
  bool conv = false;
  for (int it = 0; it < 10; ++it) {
    const double X = ox + t * dx + x0, Y = oy + t * dy + y0;
    double z, g;
    sag_and_dr2<double>(s, X * X + Y * Y, z, g);
    const double gv = (oz + t * dz) - (z - z0);
    double gp = dz - (2.0 * X * g * dx + 2.0 * Y * g * dy);
    if (!(fabs(gp) > 1e-12)) gp = 1e-12;
    const double tn = t - gv / gp;
    const bool newc = conv || (fabs(gv) < 1e-8);
    if (!conv) t = tn;
    conv = newc;
    const bool done = conv || !__builtin_isfinite(t);
    if (__all(done)) break;
  }
  const double X = ox + t * dx + x0, Y = oy + t * dy + y0;
  const double res = fabs((oz + t * dz) - (sag_raw<double>(s, X, Y) - z0));
  valid = (t > 1e-8) && (res < 1e-6);
  return t;
}

// ---------------------------------------------------------------- obstructions
// Each returns the ray parameter of the nearest hit or +inf
// (intersections.py:44-226).  The shadow test only asks t < 1e10.
// The cylinder test in two parts: the terms of the ray direction alone
// (cyl_dir), which a run of rays with one direction forms once, and the
// rest (hit_cylinder_dir).  hit_cylinder is their composition: every caller
// evaluates the same expressions.
struct CylDir {
  f3 w, rdp;
  float ww, rda, a, ia, ir;
};
IACT_DEV CylDir cyl_dir(f3 d, cptr<float> c) {
  const float eps = 1e-8f;
  const f3 ax = mk3(c[3], c[4], c[5]);
  CylDir D;
  D.w = cross(d, ax);
  D.ww = dot(D.w, D.w);
  D.rda = dot(d, ax);
  D.rdp = d - D.rda * ax;
  D.a = dot(D.rdp, D.rdp);
  D.ia = frcp(2.f * D.a + eps);
  D.ir = frcp(D.rda + eps);
  return D;
}
// Early out.  With w = d x ax, (oc.w)^2 / |w|^2 is the squared distance of
// the ray's line from the axis; when it exceeds r^2 by 2e-4 r^2 plus
// 1e-5 |oc|^2 (a bound on the rounding of this estimate and of the
// discriminant below, which has the same sign as r^2 minus that distance) no
// side or cap hit is possible.  Lines parallel to the axis (w = 0) and NaN
// geometry never take it.  The per-lane result is the miss for such a lane
// whatever its wave, and the rest of the test is skipped when it holds for
// the whole wave.
IACT_DEV bool cyl_far(f3 oc, const CylDir& D, float r) {
  const float ow = dot(oc, D.w);
  return ow * ow > D.ww * fmaf(1e-5f, dot(oc, oc), 1.0002f * r * r);
}
// The side and cap hits from the origin's axial split (oca = oc.ax,
// ocp = oc - oca ax, cc = |ocp|^2 - r^2) and the direction terms.
IACT_DEV float cyl_rest(f3 ocp, float oca, float cc, const CylDir& D, float h, float r,
                        bool far) {
  const float eps = 1e-8f;
  const float inf = __builtin_inff();
  const float b = 2.f * dot(ocp, D.rdp);
  const float disc = b * b - 4.f * D.a * cc;
  const float sq = fsqrt(fmaxf(disc, 0.f));
  float t1 = (-b - sq) * D.ia, t2 = (-b + sq) * D.ia;
  const float y1 = oca + t1 * D.rda, y2 = oca + t2 * D.rda;
  t1 = (t1 > eps && y1 >= 0.f && y1 <= h && disc >= 0.f) ? t1 : inf;
  t2 = (t2 > eps && y2 >= 0.f && y2 <= h && disc >= 0.f) ? t2 : inf;
  float tb = -oca * D.ir, tt = (h - oca) * D.ir;
  const f3 pb = ocp + tb * D.rdp, pt = ocp + tt * D.rdp;
  tb = (tb > eps && dot(pb, pb) <= r * r) ? tb : inf;
  tt = (tt > eps && dot(pt, pt) <= r * r) ? tt : inf;
  return far ? inf : fminf(fminf(t1, t2), fminf(tb, tt));
}
IACT_DEV float hit_cylinder_dir(f3 o, const CylDir& D, cptr<float> c) {
  const f3 p1 = mk3(c[0], c[1], c[2]);
  const f3 ax = mk3(c[3], c[4], c[5]);
  const float h = c[6], r = c[7];
  const f3 oc = o - p1;
  const bool far = cyl_far(oc, D, r);
  if (__all(far)) return __builtin_inff();
  const float oca = dot(oc, ax);
  const f3 ocp = oc - oca * ax;
  const float cc = dot(ocp, ocp) - r * r;
  return cyl_rest(ocp, oca, cc, D, h, r, far);
}
// hit_cylinder_dir for two directions from one origin: the origin's terms
// are formed once; each direction's result is hit_cylinder_dir's.
IACT_DEV void hit_cylinder_dir2(f3 o, const CylDir& D0, const CylDir& D1, cptr<float> c,
                                float& t0, float& t1) {
  const f3 p1 = mk3(c[0], c[1], c[2]);
  const f3 ax = mk3(c[3], c[4], c[5]);
  const float h = c[6], r = c[7];
  const f3 oc = o - p1;
  const bool far0 = cyl_far(oc, D0, r), far1 = cyl_far(oc, D1, r);
  t0 = t1 = __builtin_inff();
  if (__all(far0 && far1)) return;
  const float oca = dot(oc, ax);
  const f3 ocp = oc - oca * ax;
  const float cc = dot(ocp, ocp) - r * r;
  if (!__all(far0)) t0 = cyl_rest(ocp, oca, cc, D0, h, r, far0);
  if (!__all(far1)) t1 = cyl_rest(ocp, oca, cc, D1, h, r, far1);
}
IACT_DEV float hit_cylinder(f3 o, f3 d, cptr<float> c) {
  return hit_cylinder_dir(o, cyl_dir(d, c), c);
}

// The axis-aligned box test in two parts, as the cylinder's.
struct BoxDir {
  float ix, iy, iz;
};
IACT_DEV BoxDir box_dir(f3 d) {
  const float eps = 1e-8f;
  return BoxDir{frcp(d.x + eps), frcp(d.y + eps), frcp(d.z + eps)};
}
IACT_DEV float hit_box_dir(f3 o, BoxDir D, cptr<float> b) {
  const float eps = 1e-8f;
  const float tx1 = (b[0] - o.x) * D.ix, tx2 = (b[3] - o.x) * D.ix;
  const float ty1 = (b[1] - o.y) * D.iy, ty2 = (b[4] - o.y) * D.iy;
  const float tz1 = (b[2] - o.z) * D.iz, tz2 = (b[5] - o.z) * D.iz;
  const float tmin = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2));
  const float tmax = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2));
  const bool hit = (tmax >= tmin) && (tmax > eps);
  return hit ? (tmin > eps ? tmin : tmax) : __builtin_inff();
}
IACT_DEV float hit_box(f3 o, f3 d, cptr<float> b) { return hit_box_dir(o, box_dir(d), b); }

IACT_DEV float hit_sphere(f3 o, f3 d, cptr<float> s) {
  const float eps = 1e-8f;
  const f3 oc = o - mk3(s[0], s[1], s[2]);
  const float a = dot(d, d), b = 2.f * dot(oc, d), c = dot(oc, oc) - s[3] * s[3];
  const float disc = b * b - 4.f * a * c;
  const float sq = fsqrt(fmaxf(disc, 0.f));
  const float ia = frcp(2.f * a + eps);
  const float t1 = (-b - sq) * ia, t2 = (-b + sq) * ia;
  const float inf = __builtin_inff();
  return fminf((t1 > eps && disc >= 0.f) ? t1 : inf, (t2 > eps && disc >= 0.f) ? t2 : inf);
}

IACT_DEV float signf1(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }

IACT_DEV float hit_oriented_box(f3 o, f3 d, cptr<float> q) {
  const float eps = 1e-8f;
  m3 R;
#pragma unroll
  for (int i = 0; i < 9; ++i) R.m[i] = q[6 + i];
  const f3 lo = mulT(R, o - mk3(q[0], q[1], q[2]));
  const f3 ld = mulT(R, d);
  const float ix = frcp(ld.x + eps * signf1(ld.x + eps));
  const float iy = frcp(ld.y + eps * signf1(ld.y + eps));
  const float iz = frcp(ld.z + eps * signf1(ld.z + eps));
  const float tx1 = (-q[3] - lo.x) * ix, tx2 = (q[3] - lo.x) * ix;
  const float ty1 = (-q[4] - lo.y) * iy, ty2 = (q[4] - lo.y) * iy;
  const float tz1 = (-q[5] - lo.z) * iz, tz2 = (q[5] - lo.z) * iz;
  const float tmin = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2));
  const float tmax = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2));
  const bool hit = (tmax >= tmin) && (tmax > eps);
  const float tr = tmin > eps ? tmin : tmax;
  return (hit && tr > eps) ? tr : __builtin_inff();
}

IACT_DEV float hit_triangle(f3 o, f3 d, cptr<float> v) {
  const float eps = 1e-8f;
  const f3 v0 = mk3(v[0], v[1], v[2]);
  const f3 e1 = mk3(v[3], v[4], v[5]) - v0;
  const f3 e2 = mk3(v[6], v[7], v[8]) - v0;
  const f3 h = cross(d, e2);
  const float a = dot(e1, h);
  const bool parallel = fabsf(a) < eps;
  const float f = frcp(a + eps * signf1(a + eps));
  const f3 s = o - v0;
  const float u = f * dot(s, h);
  const f3 q = cross(s, e1);
  const float vv = f * dot(d, q);
  const float t = f * dot(e2, q);
  const bool ok = !parallel && u >= 0.f && u <= 1.f && vv >= 0.f && u + vv <= 1.f && t > eps;
  return ok ? t : __builtin_inff();
}

// 0 if any obstruction has t < 1e10 along the unbounded ray (render.py:21-41).
// Once every lane of the wave is blocked the remaining primitives cannot
// change any result, so the test returns early for the whole wave.
// Obstruction rows are read through the constant address space: the
// primitive index is wave-uniform, so they become scalar loads.
IACT_DEV float shadow_factor(const iact_obstructions& ob, f3 o, f3 d, bool active) {
  const float big = 1e10f;
  bool blocked = !active;
  for (int i = 0; i < ob.n_cylinders; ++i) {
    if (__all(blocked)) break;
    blocked |= hit_cylinder(o, d, as_const(ob.cylinders) + 8 * i) < big;
  }
  for (int i = 0; i < ob.n_boxes; ++i) {
    if (__all(blocked)) break;
    blocked |= hit_box(o, d, as_const(ob.boxes) + 6 * i) < big;
  }
  for (int i = 0; i < ob.n_spheres; ++i) {
    if (__all(blocked)) break;
    blocked |= hit_sphere(o, d, as_const(ob.spheres) + 4 * i) < big;
  }
  for (int i = 0; i < ob.n_oriented_boxes; ++i) {
    if (__all(blocked)) break;
    blocked |= hit_oriented_box(o, d, as_const(ob.oriented_boxes) + 15 * i) < big;
  }
  for (int i = 0; i < ob.n_triangles; ++i) {
    if (__all(blocked)) break;
    blocked |= hit_triangle(o, d, as_const(ob.triangles) + 9 * i) < big;
  }
  return (blocked && active) ? 0.f : 1.f;
}

// A copy of v that the compiler cannot trace back to v: arithmetic on it
// stays in the branch that makes the copy, where it would otherwise be
// hoisted out of the candidate loop and run for every ray with candidates,
// whichever obstruction kinds they are.
IACT_DEV f3 pinned(f3 v) {
  asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z));
  return v;
}

// Ray parameter of the hit of obstruction j (mask numbering: cylinders,
// boxes, spheres, oriented boxes, triangles), +inf for a miss.
IACT_DEV float hit_obstruction(const iact_obstructions& ob, int j, f3 o, f3 d) {
  const int c0 = ob.n_cylinders, c1 = c0 + ob.n_boxes, c2 = c1 + ob.n_spheres;
  const int c3 = c2 + ob.n_oriented_boxes;
  if (j < c0) return hit_cylinder(o, d, as_const(ob.cylinders) + 8 * j);
  if (j < c1) return hit_box(o, pinned(d), as_const(ob.boxes) + 6 * (j - c0));
  if (j < c2) return hit_sphere(o, pinned(d), as_const(ob.spheres) + 4 * (j - c1));
  if (j < c3) return hit_oriented_box(o, pinned(d), as_const(ob.oriented_boxes) + 15 * (j - c2));
  return hit_triangle(o, pinned(d), as_const(ob.triangles) + 9 * (j - c3));
}

// shadow_factor over a candidate list held one per lane: bit i of `bits`
// selects the obstruction index in lane i of `list` (wave-uniform bits).
IACT_DEV float shadow_factor_list(const iact_obstructions& ob, int list, uint64_t bits, f3 o,
                                  f3 d, bool active) {
  bool blocked = !active;
  while (bits != 0ull) {
    if (__all(blocked)) break;
    const int pos = __builtin_ctzll(bits);
    bits &= bits - 1ull;
    const int j = __builtin_amdgcn_readlane(list, pos);
    blocked |= hit_obstruction(ob, j, o, d) < 1e10f;
  }
  return (blocked && active) ? 0.f : 1.f;
}

// shadow_factor over a candidate list: bit j of mask word w selects
// obstruction 32 w + j in the order cylinders, boxes, spheres, oriented
// boxes, triangles.  The list is a superset of the obstructions the rays can
// hit, so the result is shadow_factor's.
// The facet's candidate mask is `mask` (words words); the caller has read
// its first word `first` and its "any bit set" flag `any` once per facet, so
// a ray of a facet without candidates (most facets of a dish) leaves at
// once, before the per-ray set-up that the compiler hoists out of the tests,
// and without a scalar-memory wait (which would also wait for the caller's
// prefetches).
IACT_DEV float shadow_factor_culled(const iact_obstructions& ob, cptr<uint32_t> mask, int words,
                                    uint32_t first, int any, f3 o, f3 d, bool active) {
  const float big = 1e10f;
  const int c0 = ob.n_cylinders, c1 = c0 + ob.n_boxes, c2 = c1 + ob.n_spheres;
  const int c3 = c2 + ob.n_oriented_boxes;
  if (!any) return 1.f;
  bool blocked = !active;
  for (int w = 0; w < words; ++w) {
    uint32_t bits = w == 0 ? first : mask[w];
    while (bits != 0u) {
      if (__all(blocked)) return (blocked && active) ? 0.f : 1.f;
      const int j = 32 * w + __builtin_ctz(bits);
      bits &= bits - 1u;
      float t;
      if (j < c0)
        t = hit_cylinder(o, d, as_const(ob.cylinders) + 8 * j);
      else if (j < c1)
        t = hit_box(o, pinned(d), as_const(ob.boxes) + 6 * (j - c0));
      else if (j < c2)
        t = hit_sphere(o, pinned(d), as_const(ob.spheres) + 4 * (j - c1));
      else if (j < c3)
        t = hit_oriented_box(o, pinned(d), as_const(ob.oriented_boxes) + 15 * (j - c2));
      else
        t = hit_triangle(o, pinned(d), as_const(ob.triangles) + 9 * (j - c3));
      blocked |= t < big;
    }
  }
  return (blocked && active) ? 0.f : 1.f;
}

__host__ __device__ inline int n_obstructions(const iact_obstructions& ob) {
  return ob.n_cylinders + ob.n_boxes + ob.n_spheres + ob.n_oriented_boxes + ob.n_triangles;
}

__host__ __device__ inline bool has_obstructions(const iact_obstructions& ob) {
  return (ob.n_cylinders | ob.n_boxes | ob.n_spheres | ob.n_oriented_boxes | ob.n_triangles) != 0;
}

// ---------------------------------------------------------------- sensors
// Derived, launch-constant sensor numbers.
struct DevSensor {
  int kind, splat, n_pixels;
  int axis;        // frame is the identity (u1 = x, u2 = y, normal = z)
  f3 pos, u1, u2, nrm;
  // a direction r enters the hit only through r.An (= r.nrm) and its grid
  // components r.Au, r.Av (square: u1 and v_scale u2; hexagonal: the axial
  // rows of the grid map over u1, u2)
  f3 An, Au, Av;
  float ndotp;
  float nx0, ny0;  // -x0/dx, -y0/dy: pixel coordinate = fma(x, 1/dx, -x0/dx)
  float v_scale;   // (1/dy) / (1/dx): 1 for square pixels
  float aqx, aqy, aq0, arx, ary, ar0;  // sensor (x, y) -> fractional axial (q, r)
  // square
  int width, height;
  float x0, y0, dx, dy, inv_dx, inv_dy, wf, hf, edge_width;
  // hex
  float size, inv_size, size_s3, size_15, inv_inradius, gcos, gsin, gox, goy, edge_thr;
  float qmin_f, rmin_f, tq_f, tr_f;  // qmin_f = rmin_f = 0: axial is table-relative
  int q_min, r_min, tq, tr;
  int edge;  // the sensor has an edge test (square edge_width > 0, hex edge_thr < 1)
  // A ray from origin constants G reaching the plane at t = G.kt / r_n lands
  // at least |G.kt| gscale |r_inplane| / |r_n| (grid units) from (G.ku, G.kv);
  // any hit on the grid's table lies within |ku| + |kv| + gext of it.
  float gscale, gext;
  const int32_t* table;
  // splat
  float neg_half_inv_sigma2;
  int ks;
};

// Sensor-plane intersection (intersections.py:6-41) split into per-origin
// constants -- kt = n.(c - o), kx = (o - c).u1, ky = (o - c).u2 -- and the per-
// ray part x = kx + t (r.u1), y = ky + t (r.u2), t = kt / (r.n).  For the
// primary leg the origin is the facet sample, so the constants are computed
// once per sample and reused for every source.
struct PlaneK {
  float kt, kx, ky;
};
IACT_DEV PlaneK plane_consts(const DevSensor& S, f3 o) {
  const f3 q = o - S.pos;
  return PlaneK{-dot(q, S.nrm), dot(q, S.u1), dot(q, S.u2)};
}
// Returns false for rays parallel to the plane or hitting it at t <= 0, the
// reference's 1e10 sentinel cases.
template <bool AXIS>
// rs: 1/|r| when r is not a unit vector (the hit depends only on r's
// direction; the parallel-ray test is on the unit vector's normal component).
IACT_DEV bool plane_xy(const DevSensor& S, const PlaneK& K, f3 r, float& x, float& y,
                       float rs = 1.f) {
  float nd, ru, rv;
  if (AXIS) {
    nd = r.z;
    ru = r.x;
    rv = r.y;
  } else {
    nd = dot(r, S.nrm);
    ru = dot(r, S.u1);
    rv = dot(r, S.u2);
  }
  const float t = K.kt * frcp(nd);
  x = fmaf(t, ru, K.kx);
  y = fmaf(t, rv, K.ky);
  return !(fabsf(nd * rs) < 1e-10f) && t > 0.f;
}

// Sensor-grid constants of an origin.  The plane hit (x, y) = (kx + t ru,
// ky + t rv) enters the binning only through an affine map -- pixel units
// ((x - x0)/dx, (y - y0)/dy) for square sensors, fractional axial (q, r)
// for hexagonal ones -- so the origin's part folds into (ku, kv) and a ray
// is left with t times the grid components of its direction.  Square
// sensors also fold 1/dx into kt (the sign of t is unchanged).
struct GridK {
  float kt, ku, kv;
};
template <int KIND>
IACT_DEV GridK grid_consts(const DevSensor& S, const PlaneK& K) {
  if (KIND == IACT_SENSOR_SQUARE)
    return GridK{K.kt * S.inv_dx, fmaf(K.kx, S.inv_dx, S.nx0), fmaf(K.ky, S.inv_dy, S.ny0)};
  return GridK{K.kt, fmaf(K.kx, S.aqx, fmaf(K.ky, S.aqy, S.aq0)),
               fmaf(K.kx, S.arx, fmaf(K.ky, S.ary, S.ar0))};
}

// plane_xy in grid coordinates (gu, gv); the same miss conditions.
// GEN (generic) false: the launch needs no parallel-ray test (a ray that
// fails it lands off the grid anyway, plane_test_redundant) and the pixels
// are square (v_scale = 1).
template <int KIND, bool AXIS, bool GEN = true>
IACT_DEV bool plane_grid(const DevSensor& S, const GridK& G, f3 r, float& gu, float& gv,
                         float rs = 1.f) {
  float nd, ru, rv;
  if (AXIS) {
    nd = r.z;
    ru = r.x;
    rv = r.y;
  } else {
    nd = dot(r, S.nrm);
    ru = dot(r, S.u1);
    rv = dot(r, S.u2);
  }
  const float t = G.kt * frcp(nd);
  if (KIND == IACT_SENSOR_SQUARE) {
    gu = fmaf(t, ru, G.ku);
    gv = fmaf(t, GEN ? rv * S.v_scale : rv, G.kv);
  } else {
    gu = fmaf(t, fmaf(ru, S.aqx, rv * S.aqy), G.ku);
    gv = fmaf(t, fmaf(ru, S.arx, rv * S.ary), G.kv);
  }
  return (!GEN || !(fabsf(nd * rs) < 1e-10f)) && t > 0.f;
}

// The hit of a primary ray in grid coordinates from its reflection written
// as r = c n + u (c = u.m, m = -2 n): r.A = c (n.A) + u.A for each of the
// sensor's directions A, so with N = (n.An, n.Au, n.Av) per sample and
// U = (u.An, u.Au, u.Av) -- per (facet, source) when u is -- the ray needs
// three FMAs for its normal and grid components.  The same miss conditions
// as plane_grid (rs: 1/|u|).
IACT_DEV f3 sensor_dirs(const DevSensor& S, f3 u) {
  return mk3(dot(u, S.An), dot(u, S.Au), dot(u, S.Av));
}
template <bool GEN = true>
IACT_DEV bool grid_hit(const GridK& G, f3 N, f3 U, float c, float& gu, float& gv,
                       float rs = 1.f) {
  const float nd = fmaf(c, N.x, U.x);
  const float ru = fmaf(c, N.y, U.y);
  const float rv = fmaf(c, N.z, U.z);
  const float t = G.kt * frcp(nd);
  gu = fmaf(t, ru, G.ku);
  gv = fmaf(t, rv, G.kv);
  return (!GEN || !(fabsf(nd * rs) < 1e-10f)) && t > 0.f;
}

// True when the parallel-ray test (|r_n| < 1e-10 for a unit r) cannot change
// a binning decision for rays from origin constants G: such a ray lands at
// least 0.99e10 |G.kt| gscale from (ku, kv), and 1e9 |G.kt| gscale already
// exceeds the farthest grid position; NaN constants keep the test.
IACT_DEV bool plane_test_redundant(const DevSensor& S, const GridK& G) {
  return fabsf(G.kt) * S.gscale * 1e9f > fabsf(G.ku) + fabsf(G.kv) + S.gext;
}

// Square binning (square.py:66-91) of pixel coordinates (xc, yc): floor,
// range, edge distance.  The edge test, when the sensor has one, acts on the
// pixel id inside a launch-uniform branch.
IACT_DEV int square_pixel_grid(const DevSensor& S, float xc, float yc) {
  const float xf = floorf(xc), yf = floorf(yc);
  // bitwise ands: four compares and no branch
  const bool ok = (xf >= 0.f) & (xf < S.wf) & (yf >= 0.f) & (yf < S.hf);
  // row-major index in fp32: exact, the sensor having fewer than 2^24 pixels
  int pix = ok ? (int)fmaf(yf, S.wf, xf) : -1;
  if (S.edge) {
    const float xr = xc - xf, yr = yc - yf;
    const float dist = fminf(fminf(xr, 1.f - xr) * S.dx, fminf(yr, 1.f - yr) * S.dy);
    pix = dist < S.edge_width ? -1 : pix;
  }
  return pix;
}

// floor(x) converted to int32 in one instruction (v_cvt_flr_i32_f32; the
// compiler forms it from floor + convert only when NaNs are ruled out for the
// whole build).  Out-of-range values saturate and NaN converts to 0, so it
// serves coordinates of rays that the caller has accepted or selects away.
IACT_DEV int floor_i32(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// Square binning of the hit of a traced ray: the pixel of square_pixel_grid
// for finite coordinates (a ray whose hit is not finite is rejected by the
// trace itself), with floor and conversion in one instruction, the range
// test as two unsigned compares and a 24-bit multiply-add for the index.
// `valid` (the trace's own acceptance) joins the range test, so a rejected
// ray costs no extra select.
// EDGE false: the caller knows that the sensor has no edge test.
template <bool EDGE = true>
IACT_DEV int square_pixel_hit(const DevSensor& S, float xc, float yc, bool valid) {
  const int ix = floor_i32(xc), iy = floor_i32(yc);
  const bool ok =
      valid & ((unsigned)ix < (unsigned)S.width) & ((unsigned)iy < (unsigned)S.height);
  int pix = ok ? (int)(__umul24((unsigned)iy, (unsigned)S.width) + (unsigned)ix) : -1;
  if (EDGE && S.edge) {
    const float xr = xc - floorf(xc), yr = yc - floorf(yc);
    const float dist = fminf(fminf(xr, 1.f - xr) * S.dx, fminf(yr, 1.f - yr) * S.dy);
    pix = dist < S.edge_width ? -1 : pix;
  }
  return pix;
}

IACT_DEV int square_pixel(const DevSensor& S, float x, float y) {
  return square_pixel_grid(S, fmaf(x, S.inv_dx, S.nx0), fmaf(y, S.inv_dy, S.ny0));
}

// Sensor (x, y) -> fractional axial (q, r): the grid offset, the grid
// rotation and the pointy-top axial map (hexagonal.py:22-24, 146-150) folded
// into one affine map.
IACT_DEV void hex_axial(const DevSensor& S, float x, float y, float& q, float& r) {
  q = fmaf(x, S.aqx, fmaf(y, S.aqy, S.aq0));
  r = fmaf(x, S.arx, fmaf(y, S.ary, S.ar0));
}

// Cube rounding, half-to-even (hexagonal.py:32-39).
IACT_DEV void hex_round(float q, float r, float& qi_out, float& ri_out) {
  const float s = -q - r;
  const float qi = rintf(q), ri = rintf(r), si = rintf(s);
  const float dq = fabsf(qi - q), dr = fabsf(ri - r), ds = fabsf(si - s);
  const bool fix_q = dq > dr && dq > ds;
  const bool fix_r = !fix_q && dr > dq && dr > ds;  // the q fix-up cannot also hold
  qi_out = fix_q ? -ri - si : qi;
  ri_out = fix_r ? -qi - si : ri;
}

// The hexagonal lookup table as every kernel reads it: a copy in LDS with a
// border of -1 entries around the sensor's tq x tr entries (rows of tr + 2),
// so that a cell index clamped into [-1, tq] x [-1, tr] reads -1 off the
// table, and with every entry below -1 stored as -1 (the reference's empty
// marker): cell (q, r) sits at row q + 1, column r + 1.  The block copies it
// from S.table and must synchronise before reading.
__host__ __device__ inline int hex_table_words(const DevSensor& S) {
  return (S.tq + 2) * (S.tr + 2);
}
IACT_DEV const int32_t* stage_hex_table(const DevSensor& S, int32_t* lds) {
  const int w = S.tr + 2, n = hex_table_words(S);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int a = i / w - 1, b = i % w - 1;
    lds[i] = (a >= 0 && a < S.tq && b >= 0 && b < S.tr) ? max(S.table[a * S.tr + b], -1) : -1;
  }
  return lds;
}

// Table lookup with bounds (hexagonal.py:155-172); -1 when off-table/empty.
// table: the staged copy (stage_hex_table).
IACT_DEV int hex_lookup(const DevSensor& S, const int32_t* table, float qi, float ri) {
  const float qa = qi - S.qmin_f, ra = ri - S.rmin_f;
  if (!(qa >= 0.f && qa < S.tq_f && ra >= 0.f && ra < S.tr_f)) return -1;
  return table[((int)qa + 1) * (S.tr + 2) + (int)ra + 1];
}

// Hexagonal norm (hexagonal.py:42-47) of the offset from the cell centre,
// in units of the inradius, written in axial offsets (dq, dr):
// x/inradius = 2 (dq + dr/2), y/inradius = sqrt(3) dr.
IACT_DEV float hex_norm_axial(float dq, float dr) {
  const float a = fabsf(fmaf(0.5f, dr, dq)), b = fabsf(dr);
  return fmaxf(2.f * a, fmaf(1.5f, b, a));
}

// Hard hex binning (hexagonal.py:174-194) of fractional axial (q, r).
IACT_DEV int hex_pixel_grid(const DevSensor& S, const int32_t* table, float q, float r) {
  float qi, ri;
  hex_round(q, r, qi, ri);
  const int pid = hex_lookup(S, table, qi, ri);
  const float hd = hex_norm_axial(q - qi, r - ri);
  return (pid >= 0 && !(hd > S.edge_thr)) ? pid : -1;
}


// Hard hex binning of a traced ray's hit (table-relative fractional axial
// q, r), -1 when the trace rejected the ray (ok false): the pixel of
// hex_pixel_grid (up to points within rounding of a cell boundary, see
// hex_cell_staged), the staged table read for every lane and then selected,
// and the edge test only for sensors with an edge.  Without one
// (edge_thr = 1) the reference's test can only drop a point whose hex norm
// rounds above 1, i.e. a point within rounding of a cell boundary.
// Nearest hex centre from three floors: with cube coordinates x = q,
// z = r, y = -q - r, every cell boundary lies on an integer line of
// a = x - y = 2q + r, b = y - z = -q - 2r or c = z - x = r - q, and on each
// triangle those lines cut out the centre is ((A - C + 1) div 3,
// (C - B + 1) div 3) for A, B, C the floors (checked against cube rounding on
// the CPU).  k div 3 is floor(k/3 + 1/2) for an integer k: 1/6 clear of
// every integer, so the fp32 product is safe, and so is floor(k/3 + 3/2),
// the staged table's row (cell + 1).  Clamped into [0, tq + 1] x [0, tr + 1]
// before the conversion (a med3 each; NaN converts to 0), every cell off the
// sensor's table lands on the staged border (-1), so no range test is needed.
// The cell of hex_round except for points within rounding of a cell
// boundary; 12 vector instructions with the clamps and conversions.
IACT_DEV void hex_cell_staged(const DevSensor& S, float q, float r, int& row, int& col) {
  const float A = floorf(fmaf(2.f, q, r));
  const float B = floorf(fmaf(-2.f, r, -q));
  const float C = floorf(r - q);
  row = floor_i32(__builtin_amdgcn_fmed3f(fmaf(A - C, 1.f / 3.f, 1.5f), 0.f, S.tq_f + 1.f));
  col = floor_i32(__builtin_amdgcn_fmed3f(fmaf(C - B, 1.f / 3.f, 1.5f), 0.f, S.tr_f + 1.f));
}

template <bool EDGE = true>
IACT_DEV int hex_pixel_hit(const DevSensor& S, const int32_t* table, float q, float r, bool ok) {
  int row, col;
  hex_cell_staged(S, q, r, row, col);
  // the entries are -1 or pixel ids; the index is kept whole (an empty asm)
  // so that it stays one multiply-add, which the address then shifts
  unsigned idx = __umul24((unsigned)row, (unsigned)(S.tr + 2)) + (unsigned)col;
  asm volatile("" : "+v"(idx));
  const int pid = table[idx];
  bool good = ok;
  // (a clamped cell reads -1, whatever the edge test says)
  if (EDGE && S.edge)
    good = good & !(hex_norm_axial(q - (float)(row - 1), r - (float)(col - 1)) > S.edge_thr);
  return good ? pid : -1;
}

IACT_DEV int hex_pixel(const DevSensor& S, const int32_t* table, float x, float y) {
  float q, r;
  hex_axial(S, x, y, q, r);
  return hex_pixel_grid(S, table, q, r);
}

// Add to a block's LDS image: relaxed, workgroup scope (the image is private
// to the block; a distinct scope also keeps the compiler from merging it with
// a global add into one flat atomic).
IACT_DEV void lds_add(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---------------------------------------------------------------- wave ops
IACT_DEV int lane_id() { return __lane_id(); }

// Cross-lane sum of a wave64 with DPP row operations (no LDS round trips):
// pairs, quads, half rows, rows, then row broadcasts; lane 63 ends up with
// the total, which is returned to every lane as a scalar.
template <int CTRL, int ROW_MASK = 0xf>
IACT_DEV float dpp_f(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xf, true));
}
IACT_DEV float wave_sum(float v) {
  v += dpp_f<0xB1>(v);         // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);         // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);        // row_half_mirror
  v += dpp_f<0x140>(v);        // row_mirror
  v += dpp_f<0x142, 0xa>(v);   // row_bcast:15 into rows 1, 3
  v += dpp_f<0x143, 0xc>(v);   // row_bcast:31 into rows 2, 3
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}
IACT_DEV double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Add every lane's value into img[pix] (pix < 0: nothing).  Lanes that share
// a pixel are summed across the wave first so the LDS (or global) image sees
// one atomic per distinct pixel; after `max_rounds` distinct pixels the rest
// add directly.  Must be called by the whole wave in uniform control flow.
template <bool GLOBAL, class T>
IACT_DEV void wave_accumulate(T* img, int pix, float v, int max_rounds) {
  uint64_t rem = __ballot(pix >= 0 && v != 0.f);
  int rounds = 0;
  const int lane = lane_id();
  while (rem) {
    if (rounds == max_rounds) {
      if ((rem >> lane) & 1ull) {
        if (GLOBAL)
          unsafeAtomicAdd(img + pix, (T)v);
        else
          atomicAdd(img + pix, (T)v);
      }
      break;
    }
    const int leader = __builtin_ctzll(rem);
    const int key = __builtin_amdgcn_readlane(pix, leader);
    const bool mine = ((rem >> lane) & 1ull) && pix == key;
    const uint64_t m = __ballot(mine);
    const float s = wave_sum(mine ? v : 0.f);
    if (lane == leader) {
      if (GLOBAL)
        unsafeAtomicAdd(img + key, (T)s);
      else
        atomicAdd(img + key, (T)s);
    }
    rem &= ~m;
    ++rounds;
  }
}

}  // namespace iact
