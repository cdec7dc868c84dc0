/*
 * cpu_ref -- C++/OpenMP restatement of the reference renderer's image path.
 * TEST INFRASTRUCTURE ONLY: the CPU baseline timed by bench.py, a second,
 * independent restatement checked against oracle/iact_oracle.py in tests, and
 * the FP64 run that full-size GPU images are compared with.
 *
 * Follows iactrace/core/render.py:118-220 in order: for every primary facet
 * (lax.scan, render.py:210-219) and every (source, sample) ray:
 * source direction (render.py:129-133), occlusion toward the source
 * (render.py:138, 21-41), reflect (reflection.py:5-19), value = value * cos / w
 * * shadow (render.py:141), stage>=1 mirrors (render.py:44-115; conic guess +
 * Newton, intersections.py:229-367), sensor plane (intersections.py:6-41) and
 * hard binning (square.py:66-91, hexagonal.py:174-194).
 *
 * One template, two arithmetics:
 *  - R = float: the reference's fp32 arithmetic (IEEE division and square
 *    root, no contraction) with the two sanctioned FP64 steps (point-source
 *    direction, Newton refinement) -- the CPU baseline;
 *  - R = double: every step in double -- the FP64 oracle run.
 * Both accumulate the image in double (per-thread images, summed in thread
 * order), so the summation error is far below fp32 resolution.
 *
 * The FP64 run can also measure the edge band: a ray whose sensor-grid
 * coordinate lies within `band` (pixels for square sensors, hex-norm units
 * for hexagonal ones) of a binning boundary -- a cell edge or the edge-mask
 * threshold -- is one that an fp32 evaluation may bin differently.  Its |value|
 * is added to band_image at its own pixel and at the pixel across the nearest
 * cell edge, so |image_fp32(p) - image_fp64(p)| <= band_image(p) + (summation
 * error) whenever every fp32 coordinate is within `band` of the fp64 one.
 *
 * It reads the scene through the same descriptor layout as the HIP library
 * (include/iactrace_amd.h) but with HOST pointers.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/iactrace_amd.h"

namespace {

template <class R>
struct v3 {
  R x, y, z;
};
template <class R>
inline v3<R> V(R x, R y, R z) {
  return v3<R>{x, y, z};
}
template <class R>
inline v3<R> add(v3<R> a, v3<R> b) {
  return V(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class R>
inline v3<R> sub(v3<R> a, v3<R> b) {
  return V(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class R>
inline v3<R> scl(R s, v3<R> a) {
  return V(s * a.x, s * a.y, s * a.z);
}
template <class R>
inline R dot(v3<R> a, v3<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
inline v3<R> cross(v3<R> a, v3<R> b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R>
inline v3<R> mul(const R* M, v3<R> v) {  // row-major 3x3 times v
  return V(M[0] * v.x + M[1] * v.y + M[2] * v.z, M[3] * v.x + M[4] * v.y + M[5] * v.z,
           M[6] * v.x + M[7] * v.y + M[8] * v.z);
}
template <class R>
inline v3<R> mulT(const R* M, v3<R> v) {
  return V(M[0] * v.x + M[3] * v.y + M[6] * v.z, M[1] * v.x + M[4] * v.y + M[7] * v.z,
           M[2] * v.x + M[5] * v.y + M[8] * v.z);
}
template <class R>
inline v3<R> load3(const float* p) {
  return V((R)p[0], (R)p[1], (R)p[2]);
}
template <class R>
inline R inf() {
  return (R)INFINITY;
}

/* transforms.py:72-107: R = Rz Ry Rx of Euler degrees */
template <class R>
void euler(const float* a, R* M) {
  const R k = (R)(3.14159265358979323846 / 180.0);
  const R rx = (R)a[0] * k, ry = (R)a[1] * k, rz = (R)a[2] * k;
  const R cx = std::cos(rx), sx = std::sin(rx), cy = std::cos(ry), sy = std::sin(ry);
  const R cz = std::cos(rz), sz = std::sin(rz);
  const R A[9] = {cz * cy, -sz, cz * sy, sz * cy, cz, sz * sy, -sy, (R)0, cy}; /* Rz Ry */
  M[0] = A[0];
  M[1] = A[1] * cx + A[2] * sx;
  M[2] = -A[1] * sx + A[2] * cx;
  M[3] = A[3];
  M[4] = A[4] * cx + A[5] * sx;
  M[5] = -A[4] * sx + A[5] * cx;
  M[6] = A[6];
  M[7] = A[7] * cx + A[8] * sx;
  M[8] = -A[7] * sx + A[8] * cx;
}

/* ---------------------------------------------------------------- surfaces */
/* surfaces.py:25-39 (aspheric exponent 2(i+1) on r2) */
template <class T>
T sag(const iact_surface* s, T X, T Y) {
  const T c = (T)s->curvature, k = (T)s->conic, r2 = X * X + Y * Y;
  T z = r2 * c / ((T)1 + std::sqrt((T)1 - ((T)1 + k) * c * c * r2));
  const T r4 = r2 * r2;
  T p = r4;
  for (int i = 0; i < s->n_aspheric; ++i) {
    z += (T)s->aspheric[i] * p;
    p *= r4;
  }
  return z;
}
/* d sag / d r2 (the analytic form of surfaces.py:51-58's autodiff) */
template <class T>
T dsag(const iact_surface* s, T r2) {
  const T c = (T)s->curvature, k = (T)s->conic, K = ((T)1 + k) * c * c;
  const T sq = std::sqrt((T)1 - K * r2), den = (T)1 + sq;
  T g = c / den + (r2 * c) * K / ((T)2 * sq * den * den);
  const T r4 = r2 * r2;
  T q = (T)1;
  for (int i = 0; i < s->n_aspheric; ++i) {
    g += (T)s->aspheric[i] * (T)(2 * (i + 1)) * r2 * q;
    q *= r4;
  }
  return g;
}

/* intersections.py:229-285 (FP64) */
double conic_root(double ox, double oy, double oz, double dx, double dy, double dz, double c,
                  double k) {
  const double ok = 1.0 + k;
  const double A = c * (dx * dx + dy * dy + ok * dz * dz);
  const double B = 2.0 * (c * (ox * dx + oy * dy + ok * oz * dz) - dz);
  const double C = c * (ox * ox + oy * oy + ok * oz * oz) - 2.0 * oz;
  if (fabs(c) < 1e-12) return fabs(dz) > 1e-10 ? -oz / dz : INFINITY;
  const double disc = B * B - 4.0 * A * C;
  if (disc < 0) return INFINITY;
  const double sq = sqrt(disc);
  const double t1 = (-B - sq) / (2.0 * A + 1e-30), t2 = (-B + sq) / (2.0 * A + 1e-30);
  const int v1 = t1 > 1e-8, v2 = t2 > 1e-8;
  return (v1 && v2) ? fmin(t1, t2) : (v1 ? t1 : (v2 ? t2 : INFINITY));
}

/* surfaces.py:67-108 + intersections.py:290-367: FP64 conic guess and Newton
 * (10 iterations, the step of the iteration that first converges is taken),
 * hit point and normal in R, aperture test mirrors.py:147-149 / 209-220 */
template <class R>
R mirror_hit(const iact_surface* s, const float* M, v3<R> ol, v3<R> dl, v3<R>* pl, v3<R>* nl) {
  const double x0 = M[15], y0 = M[16];
  const double z0 = sag<double>(s, x0, y0);
  double t = conic_root((double)ol.x + x0, (double)ol.y + y0, (double)ol.z + z0, dl.x, dl.y,
                        dl.z, s->curvature, s->conic);
  int conv = 0;
  for (int it = 0; it < 10; ++it) {
    const double X = (double)ol.x + t * dl.x + x0, Y = (double)ol.y + t * dl.y + y0;
    const double r2 = X * X + Y * Y;
    const double gv = ((double)ol.z + t * dl.z) - (sag<double>(s, X, Y) - z0);
    const double g = dsag<double>(s, r2);
    double gp = dl.z - (2.0 * X * g * dl.x + 2.0 * Y * g * dl.y);
    if (!(fabs(gp) > 1e-12)) gp = 1e-12;
    const double tn = t - gv / gp;
    const int nc = conv || fabs(gv) < 1e-8;
    if (!conv) t = tn;
    conv = nc;
  }
  const double X = (double)ol.x + t * dl.x + x0, Y = (double)ol.y + t * dl.y + y0;
  const double res = fabs(((double)ol.z + t * dl.z) - (sag<double>(s, X, Y) - z0));
  const int valid = (t > 1e-8) && (res < 1e-6);
  const R xh = (R)((double)ol.x + t * dl.x), yh = (R)((double)ol.y + t * dl.y);
  const R fx0 = (R)M[15], fy0 = (R)M[16];
  const R Xs = xh + fx0, Ys = yh + fy0;
  const R zs = sag<R>(s, Xs, Ys) - sag<R>(s, fx0, fy0);
  const R gs = dsag<R>(s, Xs * Xs + Ys * Ys);
  const R zx = (R)2 * Xs * gs, zy = (R)2 * Ys * gs;
  const R nn = std::sqrt(zx * zx + zy * zy + (R)1);
  *pl = V(xh, yh, zs);
  *nl = V(-zx / nn, -zy / nn, (R)1 / nn);
  int inside = 1;
  if (M[17] == 0.f) {
    inside = xh * xh + yh * yh <= (R)M[18] * (R)M[18];
  } else {
    const int n = (int)M[19];
    for (int i = 0; i < n; ++i) {
      const int j = (i + 1) % n;
      const R ax = M[20 + 2 * i], ay = M[21 + 2 * i], bx = M[20 + 2 * j], by = M[21 + 2 * j];
      inside = inside && ((bx - ax) * (yh - ay) - (by - ay) * (xh - ax) >= (R)0);
    }
  }
  return (valid && inside) ? (R)t : inf<R>();
}

/* ------------------------------------------------------------ obstructions */
/* intersections.py:44-81 */
template <class R>
R t_cyl(v3<R> o, v3<R> d, const float* c) {
  const R eps = (R)1e-8;
  const v3<R> p1 = load3<R>(c), ax = load3<R>(c + 3);
  const R h = c[6], r = c[7];
  const v3<R> oc = sub(o, p1);
  const R oca = dot(oc, ax), rda = dot(d, ax);
  const v3<R> ocp = sub(oc, scl(oca, ax)), rdp = sub(d, scl(rda, ax));
  const R a = dot(rdp, rdp), b = (R)2 * dot(ocp, rdp), cc = dot(ocp, ocp) - r * r;
  const R disc = b * b - (R)4 * a * cc;
  const R sq = std::sqrt(std::fmax(disc, (R)0));
  R t1 = (-b - sq) / ((R)2 * a + eps), t2 = (-b + sq) / ((R)2 * a + eps);
  const R y1 = oca + t1 * rda, y2 = oca + t2 * rda;
  t1 = (t1 > eps && y1 >= 0 && y1 <= h && disc >= 0) ? t1 : inf<R>();
  t2 = (t2 > eps && y2 >= 0 && y2 <= h && disc >= 0) ? t2 : inf<R>();
  R tb = -oca / (rda + eps), tt = (h - oca) / (rda + eps);
  const v3<R> pb = add(ocp, scl(tb, rdp)), pt = add(ocp, scl(tt, rdp));
  tb = (tb > eps && dot(pb, pb) <= r * r) ? tb : inf<R>();
  tt = (tt > eps && dot(pt, pt) <= r * r) ? tt : inf<R>();
  return std::fmin(std::fmin(t1, t2), std::fmin(tb, tt));
}
/* intersections.py:84-110 */
template <class R>
R t_box(v3<R> o, v3<R> d, const float* b) {
  const R eps = (R)1e-8;
  const R ix = (R)1 / (d.x + eps), iy = (R)1 / (d.y + eps), iz = (R)1 / (d.z + eps);
  const R tx1 = ((R)b[0] - o.x) * ix, tx2 = ((R)b[3] - o.x) * ix;
  const R ty1 = ((R)b[1] - o.y) * iy, ty2 = ((R)b[4] - o.y) * iy;
  const R tz1 = ((R)b[2] - o.z) * iz, tz2 = ((R)b[5] - o.z) * iz;
  const R tmin = std::fmax(std::fmax(std::fmin(tx1, tx2), std::fmin(ty1, ty2)), std::fmin(tz1, tz2));
  const R tmax = std::fmin(std::fmin(std::fmax(tx1, tx2), std::fmax(ty1, ty2)), std::fmax(tz1, tz2));
  return ((tmax >= tmin) && (tmax > eps)) ? (tmin > eps ? tmin : tmax) : inf<R>();
}
/* intersections.py:195-226 */
template <class R>
R t_sph(v3<R> o, v3<R> d, const float* s) {
  const R eps = (R)1e-8;
  const v3<R> oc = sub(o, load3<R>(s));
  const R a = dot(d, d), b = (R)2 * dot(oc, d), c = dot(oc, oc) - (R)s[3] * (R)s[3];
  const R disc = b * b - (R)4 * a * c, sq = std::sqrt(std::fmax(disc, (R)0));
  const R t1 = (-b - sq) / ((R)2 * a + eps), t2 = (-b + sq) / ((R)2 * a + eps);
  return std::fmin((t1 > eps && disc >= 0) ? t1 : inf<R>(), (t2 > eps && disc >= 0) ? t2 : inf<R>());
}
template <class R>
R sgn(R x) {
  return x > 0 ? (R)1 : (x < 0 ? (R)-1 : (R)0);
}
/* intersections.py:113-150 */
template <class R>
R t_obox(v3<R> o, v3<R> d, const float* q) {
  const R eps = (R)1e-8;
  R Mq[9];
  for (int i = 0; i < 9; ++i) Mq[i] = q[6 + i];
  const v3<R> lo = mulT(Mq, sub(o, load3<R>(q))), ld = mulT(Mq, d);
  const R ix = (R)1 / (ld.x + eps * sgn(ld.x + eps)), iy = (R)1 / (ld.y + eps * sgn(ld.y + eps));
  const R iz = (R)1 / (ld.z + eps * sgn(ld.z + eps));
  const R tx1 = (-(R)q[3] - lo.x) * ix, tx2 = ((R)q[3] - lo.x) * ix;
  const R ty1 = (-(R)q[4] - lo.y) * iy, ty2 = ((R)q[4] - lo.y) * iy;
  const R tz1 = (-(R)q[5] - lo.z) * iz, tz2 = ((R)q[5] - lo.z) * iz;
  const R tmin = std::fmax(std::fmax(std::fmin(tx1, tx2), std::fmin(ty1, ty2)), std::fmin(tz1, tz2));
  const R tmax = std::fmin(std::fmin(std::fmax(tx1, tx2), std::fmax(ty1, ty2)), std::fmax(tz1, tz2));
  const R tr = tmin > eps ? tmin : tmax;
  return ((tmax >= tmin) && (tmax > eps) && tr > eps) ? tr : inf<R>();
}
/* intersections.py:153-192 (Moller-Trumbore) */
template <class R>
R t_tri(v3<R> o, v3<R> d, const float* v) {
  const R eps = (R)1e-8;
  const v3<R> v0 = load3<R>(v), e1 = sub(load3<R>(v + 3), v0), e2 = sub(load3<R>(v + 6), v0);
  const v3<R> h = cross(d, e2);
  const R a = dot(e1, h);
  const R f = (R)1 / (a + eps * sgn(a + eps));
  const v3<R> s = sub(o, v0);
  const R u = f * dot(s, h);
  const v3<R> q = cross(s, e1);
  const R w = f * dot(d, q), t = f * dot(e2, q);
  const int ok = !(std::fabs(a) < eps) && u >= 0 && u <= 1 && w >= 0 && u + w <= 1 && t > eps;
  return ok ? t : inf<R>();
}
/* render.py:21-41: shadowed when any primitive has a hit with t < 1e10 */
template <class R>
R shadow(const iact_obstructions* ob, v3<R> o, v3<R> d) {
  for (int i = 0; i < ob->n_cylinders; ++i)
    if (t_cyl(o, d, ob->cylinders + 8 * i) < (R)1e10) return (R)0;
  for (int i = 0; i < ob->n_boxes; ++i)
    if (t_box(o, d, ob->boxes + 6 * i) < (R)1e10) return (R)0;
  for (int i = 0; i < ob->n_spheres; ++i)
    if (t_sph(o, d, ob->spheres + 4 * i) < (R)1e10) return (R)0;
  for (int i = 0; i < ob->n_oriented_boxes; ++i)
    if (t_obox(o, d, ob->oriented_boxes + 15 * i) < (R)1e10) return (R)0;
  for (int i = 0; i < ob->n_triangles; ++i)
    if (t_tri(o, d, ob->triangles + 9 * i) < (R)1e10) return (R)0;
  return (R)1;
}

/* ----------------------------------------------------------------- sensors */
template <class R>
struct Frame {
  v3<R> pos, u1, u2, n;
  R ndotp, gc, gs;
};

/* A ray's binning: pixel id (-1 = dropped) and, for the FP64 run's edge band,
 * its distance to the nearest binning boundary and the pixel across the
 * nearest cell edge (-1 = none or outside). */
struct Bin {
  int pid;
  double margin;
  int other;
};

/* square.py:11-25, 66-91; BAND: also the edge-band fields */
template <class R, bool BAND>
Bin bin_square(const iact_sensor* S, R x, R y) {
  const R xc = (x - (R)S->x0) / (R)S->dx, yc = (y - (R)S->y0) / (R)S->dy;
  const R xf = std::floor(xc), yf = std::floor(yc);
  const bool in = xf >= 0 && xf < (R)S->width && yf >= 0 && yf < (R)S->height;
  const R xr = xc - xf, yr = yc - yf;
  const R ex = std::fmin(xr, (R)1 - xr), ey = std::fmin(yr, (R)1 - yr);
  const R dist = std::fmin(ex * (R)S->dx, ey * (R)S->dy);
  const bool keep = in && !(dist < (R)S->edge_width);
  Bin b{keep ? (int)yf * S->width + (int)xf : -1, INFINITY, -1};
  if (!BAND || !std::isfinite((double)xc) || !std::isfinite((double)yc)) return b;
  // nearest cell edge (pixel units) and the pixel beyond it
  int nx = (int)xf, ny = (int)yf;
  if (ex <= ey)
    nx += xr < (R)0.5 ? -1 : 1;
  else
    ny += yr < (R)0.5 ? -1 : 1;
  b.margin = (double)std::fmin(ex, ey);
  if (S->edge_width > 0.f) {  // the edge-mask threshold is a boundary too
    const double em = fabs((double)dist - (double)S->edge_width) /
                      (double)std::fmin(S->dx, S->dy);
    b.margin = fmin(b.margin, em);
  }
  if (nx >= 0 && nx < S->width && ny >= 0 && ny < S->height) b.other = ny * S->width + nx;
  return b;
}

/* hexagonal.py:32-47 (half-even cube rounding), 155-194 (lookup, edge mask) */
template <class R>
int hex_lookup(const iact_sensor* S, R qi, R ri) {
  const R qa = qi - (R)S->q_min, ra = ri - (R)S->r_min;
  if (!(qa >= 0 && qa < (R)S->table_q && ra >= 0 && ra < (R)S->table_r)) return -1;
  return S->lookup_table[(int)qa * S->table_r + (int)ra];
}
template <class R>
void hex_round(R q, R r, R* qo, R* ro) {
  const R s = -q - r;
  R qi = std::nearbyint(q), ri = std::nearbyint(r);
  const R si = std::nearbyint(s);
  const R dq = std::fabs(qi - q), dr = std::fabs(ri - r), ds = std::fabs(si - s);
  if (dq > dr && dq > ds)
    qi = -ri - si;
  else if (dr > dq && dr > ds)
    ri = -qi - si;
  *qo = qi;
  *ro = ri;
}
template <class R, bool BAND>
Bin bin_hex(const iact_sensor* S, const Frame<R>* F, R x, R y) {
  const R hs = (R)S->hex_size;
  const R ux = x - (R)S->grid_offset[0], uy = y - (R)S->grid_offset[1];
  const R xg = F->gc * ux - F->gs * uy, yg = F->gs * ux + F->gc * uy;
  const R q = ((R)0.5773502691896257 * xg - yg / (R)3) / hs;
  const R r = ((R)2 * yg / (R)3) / hs;
  Bin b{-1, INFINITY, -1};
  if (!(std::fabs(q) < (R)1e6 && std::fabs(r) < (R)1e6)) return b;
  R qi, ri;
  hex_round(q, r, &qi, &ri);
  const int pid = hex_lookup(S, qi, ri);
  const R cx = (R)((double)S->hex_size * 1.7320508075688772) * (qi + ri / (R)2);
  const R cy = (R)((double)S->hex_size * 1.5) * ri;
  const R ax = std::fabs(xg - cx), ay = std::fabs(yg - cy);
  const R hd = std::fmax(ax, (R)0.5 * ax + (R)0.8660254037844386 * ay) / (R)S->hex_inradius;
  b.pid = (pid >= 0 && !(hd > (R)S->edge_threshold)) ? pid : -1;
  if (!BAND) return b;
  // nearest cell edge: the edge normals are at 0, +-60, 180, +-120 degrees
  // in the grid frame; the neighbour is the centre two inradii along it
  const double ox = (double)xg - (double)cx, oy = (double)yg - (double)cy;
  double best = -INFINITY, bx = 0, by = 0;
  for (int k = 0; k < 6; ++k) {
    const double a = k * (3.14159265358979323846 / 3.0);
    const double nx = cos(a), ny = sin(a);
    const double p = ox * nx + oy * ny;
    if (p > best) {
      best = p;
      bx = nx;
      by = ny;
    }
  }
  const double ir = S->hex_inradius;
  b.margin = 1.0 - best / ir;
  if (S->edge_threshold < 1.f) b.margin = fmin(b.margin, fabs((double)hd - S->edge_threshold));
  const double qx = (double)cx + 2.0 * ir * bx, qy = (double)cy + 2.0 * ir * by;
  const double qq = (0.5773502691896257 * qx - qy / 3.0) / (double)S->hex_size;
  const double rr = (2.0 * qy / 3.0) / (double)S->hex_size;
  double qn, rn;
  hex_round(qq, rr, &qn, &rn);
  b.other = hex_lookup(S, qn, rn);
  return b;
}

/* ------------------------------------------------------------------ render */
template <class R, bool BAND>
int render_t(const iact_scene* sc, const double* sources, const float* values, int ns,
             int source_type, double* image, int nthreads, double band, double* band_image,
             long long* stats) {
  const iact_facets* Fc = &sc->facets;
  const iact_sensor* S = &sc->sensor;
  const int P = S->n_pixels, nf = Fc->n_facets, M = Fc->n_samples;
  memset(image, 0, sizeof(double) * P);
  if (band_image) memset(band_image, 0, sizeof(double) * P);
  if (stats) stats[0] = stats[1] = stats[2] = 0;
  if (nf == 0 || ns == 0) return 0;
  Frame<R> F;
  F.pos = load3<R>(S->position);
  F.u1 = V((R)S->rotation[0], (R)S->rotation[3], (R)S->rotation[6]);
  F.u2 = V((R)S->rotation[1], (R)S->rotation[4], (R)S->rotation[7]);
  F.n = V((R)S->rotation[2], (R)S->rotation[5], (R)S->rotation[8]);
  F.ndotp = dot(F.n, F.pos);
  F.gc = std::cos(-(R)S->grid_rotation);
  F.gs = std::sin(-(R)S->grid_rotation);
  /* world samples (mirrors.py:64-79) */
  std::vector<R> W((size_t)7 * nf * M);
  for (int f = 0; f < nf; ++f) {
    R Rm[9];
    euler<R>(Fc->rotations + 3 * f, Rm);
    const v3<R> pos = load3<R>(Fc->positions + 3 * f);
    for (int m = 0; m < M; ++m) {
      const size_t g = (size_t)f * M + m;
      const v3<R> p = add(mul(Rm, load3<R>(Fc->points + 3 * g)), pos);
      const v3<R> n = mul(Rm, load3<R>(Fc->normals + 3 * g));
      const v3<R> dl = mul(Rm, load3<R>(Fc->delta + 3 * g));
      const v3<R> pe = add(n, scl((R)Fc->perturbation_scale[f], dl));
      const R nn = std::sqrt(dot(pe, pe));
      R* w = W.data() + 7 * g;
      w[0] = p.x;
      w[1] = p.y;
      w[2] = p.z;
      w[3] = pe.x / nn;
      w[4] = pe.y / nn;
      w[5] = pe.z / nn;
      w[6] = (R)Fc->weights[g];
    }
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  const bool want_band = BAND;
  std::vector<double> priv((size_t)nt * P, 0.0), bpriv(want_band ? (size_t)nt * P : 0, 0.0);
  std::vector<long long> cnt((size_t)nt * 3, 0);
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* img = priv.data() + (size_t)tid * P;
    double* bimg = want_band ? bpriv.data() + (size_t)tid * P : nullptr;
    // counters in registers (threads' slots in one array would share cache
    // lines and every ray would pay for it)
    long long c3[3] = {0, 0, 0};
#pragma omp for schedule(dynamic, 1) collapse(2)
    for (int f = 0; f < nf; ++f)
      for (int s = 0; s < ns; ++s) {
        for (int m = 0; m < M; ++m) {
          const R* w = W.data() + 7 * ((size_t)f * M + m);
          const v3<R> p = V(w[0], w[1], w[2]), n = V(w[3], w[4], w[5]);
          v3<R> d;
          if (source_type == IACT_SOURCE_POINT) {
            const double dx = (double)p.x - sources[3 * s], dy = (double)p.y - sources[3 * s + 1];
            const double dz = (double)p.z - sources[3 * s + 2];
            const double nr = sqrt(dx * dx + dy * dy + dz * dz);
            d = V((R)(dx / nr), (R)(dy / nr), (R)(dz / nr));
          } else {
            d = V((R)sources[3 * s], (R)sources[3 * s + 1], (R)sources[3 * s + 2]);
          }
          const R sh = shadow<R>(&sc->obstructions, p, scl((R)-1, d));
          const R c = dot(d, n);
          v3<R> r = sub(d, scl((R)2 * c, n));
          R v = (R)values[s] * (-c) / w[6] * sh;
          v3<R> o = p;
          for (int k = 0; k < sc->n_stages; ++k) {
            R best = inf<R>();
            v3<R> bp = V((R)0, (R)0, (R)0), bn = V((R)0, (R)0, (R)0);
            for (int i = 0; i < sc->n_mirrors; ++i) {
              const float* Mr = sc->mirrors + (size_t)IACT_MIRROR_STRIDE * i;
              if ((int)Mr[36] != sc->stages[k]) continue;
              R Rm[9];
              for (int j = 0; j < 9; ++j) Rm[j] = Mr[6 + j];
              const v3<R> pos = load3<R>(Mr);
              v3<R> pl, nl;
              const R t = mirror_hit<R>(&sc->surfaces[(int)Mr[38]], Mr, mulT(Rm, sub(o, pos)),
                                        mulT(Rm, r), &pl, &nl);
              if (t < best) {  // rows sorted by (stage, group): first index wins ties
                best = t;
                bp = add(mul(Rm, pl), pos);
                bn = mul(Rm, nl);
              }
            }
            const R cc = dot(r, bn);
            const v3<R> rr = sub(r, scl((R)2 * cc, bn));
            v = v * (best < (R)1e10 ? (R)1 : (R)0) * shadow<R>(&sc->obstructions, o, r) *
                std::fabs(cc);
            o = bp;
            r = rr;
          }
          ++c3[2];
          const R ndotd = dot(r, F.n);
          const bool par = std::fabs(ndotd) < (R)1e-10;
          const R t = (F.ndotp - dot(o, F.n)) / (par ? (R)1 : ndotd);
          if (par || !(t > (R)0)) continue;
          const v3<R> op = sub(add(o, scl(t, r)), F.pos);
          const R x = dot(op, F.u1), y = dot(op, F.u2);
          const Bin b = S->kind == IACT_SENSOR_SQUARE ? bin_square<R, BAND>(S, x, y)
                                                      : bin_hex<R, BAND>(S, &F, x, y);
          if (b.pid >= 0) {
            img[b.pid] += (double)v;
            ++c3[0];
          }
          if (BAND && b.margin < band && v != (R)0) {
            ++c3[1];
            const double av = fabs((double)v);
            if (b.pid >= 0) bimg[b.pid] += av;
            if (b.other >= 0 && b.other != b.pid) bimg[b.other] += av;
          }
        }
      }
    for (int k = 0; k < 3; ++k) cnt[3 * tid + k] = c3[k];
  }
  for (int t = 0; t < nt; ++t)
    for (int i = 0; i < P; ++i) image[i] += priv[(size_t)t * P + i];
  if (want_band)
    for (int t = 0; t < nt; ++t)
      for (int i = 0; i < P; ++i) band_image[i] += bpriv[(size_t)t * P + i];
  if (stats)
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 3; ++k) stats[k] += cnt[3 * t + k];
  return 0;
}

/* ----------------------------------------------------------------- adjoint
 * Gradient of L = sum_p G[p] image[p] (hard binning, primary-only scenes)
 * with respect to the stage-0 rotations (Euler degrees) and the source
 * values -- what jax.grad of render (render.py:174-220) gives for them:
 * the pixel of a ray is piecewise constant, so only the ray value
 * v = value (-d.n) / w * shadow (render.py:141) carries a derivative, through
 * the world normal n = normalize(R (nl + s delta)) and, for point sources,
 * the direction d = normalize(R pl + pos - src) (mirrors.py:64-79).  fp32
 * arithmetic as the forward; facets are spread over the threads (each
 * facet's gradient has one writer), per-sample adjoints are summed over the
 * sources first and pushed through dR/dangle once per sample. */
template <class R>
void euler_grads(const float* a, R* D /* 3 x 9: dR/d(angle k), per radian */) {
  const R k = (R)(3.14159265358979323846 / 180.0);
  const R ra = (R)a[0] * k, rb = (R)a[1] * k, rc = (R)a[2] * k;
  const R sa = std::sin(ra), ca = std::cos(ra), sb = std::sin(rb), cb = std::cos(rb);
  const R sc = std::sin(rc), cc = std::cos(rc);
  // R = Rz(c) Ry(b) Rx(a); columns of Rx: e0, (0, ca, sa), (0, -sa, ca)
  const R Rzy[9] = {cc * cb, -sc, cc * sb, sc * cb, cc, sc * sb, -sb, (R)0, cb};
  const R dRzy_b[9] = {-cc * sb, (R)0, cc * cb, -sc * sb, (R)0, sc * cb, -cb, (R)0, -sb};
  const R dRzy_c[9] = {-sc * cb, -cc, -sc * sb, cc * cb, -sc, cc * sb, (R)0, (R)0, (R)0};
  for (int i = 0; i < 3; ++i) {
    // d/da: Rzy dRx/da, dRx/da columns: 0, (0, -sa, ca), (0, -ca, -sa)
    D[0 * 9 + 3 * i + 0] = (R)0;
    D[0 * 9 + 3 * i + 1] = -Rzy[3 * i + 1] * sa + Rzy[3 * i + 2] * ca;
    D[0 * 9 + 3 * i + 2] = -Rzy[3 * i + 1] * ca - Rzy[3 * i + 2] * sa;
    for (int j = 0; j < 2; ++j) {
      const R* M = j == 0 ? dRzy_b : dRzy_c;
      R* Dj = D + (1 + j) * 9;
      Dj[3 * i + 0] = M[3 * i + 0];
      Dj[3 * i + 1] = M[3 * i + 1] * ca + M[3 * i + 2] * sa;
      Dj[3 * i + 2] = -M[3 * i + 1] * sa + M[3 * i + 2] * ca;
    }
  }
}

template <class R>
int grad_t(const iact_scene* sc, const double* sources, const float* values, int ns,
           int source_type, const double* G, double* g_rot, double* g_val, int nthreads) {
  const iact_facets* Fc = &sc->facets;
  const iact_sensor* S = &sc->sensor;
  const int nf = Fc->n_facets, M = Fc->n_samples;
  if (g_rot) memset(g_rot, 0, sizeof(double) * 3 * nf);
  if (g_val) memset(g_val, 0, sizeof(double) * ns);
  if (nf == 0 || ns == 0) return 0;
  if (sc->n_stages > 0 || S->splat) return -2;
  Frame<R> F;
  F.pos = load3<R>(S->position);
  F.u1 = V((R)S->rotation[0], (R)S->rotation[3], (R)S->rotation[6]);
  F.u2 = V((R)S->rotation[1], (R)S->rotation[4], (R)S->rotation[7]);
  F.n = V((R)S->rotation[2], (R)S->rotation[5], (R)S->rotation[8]);
  F.ndotp = dot(F.n, F.pos);
  F.gc = std::cos(-(R)S->grid_rotation);
  F.gs = std::sin(-(R)S->grid_rotation);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  std::vector<double> gv_priv(g_val ? (size_t)nt * ns : 0, 0.0);
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    std::vector<R> An((size_t)3 * M), Ap((size_t)3 * M), W((size_t)7 * M);
    double* gvp = g_val ? gv_priv.data() + (size_t)tid * ns : nullptr;
#pragma omp for schedule(dynamic, 1)
    for (int f = 0; f < nf; ++f) {
      R Rm[9];
      euler<R>(Fc->rotations + 3 * f, Rm);
      const v3<R> pos = load3<R>(Fc->positions + 3 * f);
      for (int m = 0; m < M; ++m) {  // world samples (mirrors.py:64-79)
        const size_t g = (size_t)f * M + m;
        const v3<R> p = add(mul(Rm, load3<R>(Fc->points + 3 * g)), pos);
        const v3<R> n = mul(Rm, load3<R>(Fc->normals + 3 * g));
        const v3<R> dl = mul(Rm, load3<R>(Fc->delta + 3 * g));
        const v3<R> pe = add(n, scl((R)Fc->perturbation_scale[f], dl));
        const R nn = std::sqrt(dot(pe, pe));
        R* w = W.data() + 7 * m;
        w[0] = p.x, w[1] = p.y, w[2] = p.z;
        w[3] = pe.x / nn, w[4] = pe.y / nn, w[5] = pe.z / nn;
        w[6] = (R)Fc->weights[g];
      }
      std::fill(An.begin(), An.end(), (R)0);
      std::fill(Ap.begin(), Ap.end(), (R)0);
      for (int s = 0; s < ns; ++s) {
        double gval = 0.0;
        for (int m = 0; m < M; ++m) {
          const R* w = W.data() + 7 * m;
          const v3<R> p = V(w[0], w[1], w[2]), n = V(w[3], w[4], w[5]);
          v3<R> d;
          R inv_u = (R)0;  // 1 / |p - src| (point sources)
          if (source_type == IACT_SOURCE_POINT) {
            const double dx = (double)p.x - sources[3 * s], dy = (double)p.y - sources[3 * s + 1];
            const double dz = (double)p.z - sources[3 * s + 2];
            const double nr = sqrt(dx * dx + dy * dy + dz * dz);
            d = V((R)(dx / nr), (R)(dy / nr), (R)(dz / nr));
            inv_u = (R)(1.0 / nr);
          } else {
            d = V((R)sources[3 * s], (R)sources[3 * s + 1], (R)sources[3 * s + 2]);
          }
          const R sh = shadow<R>(&sc->obstructions, p, scl((R)-1, d));
          if (sh == (R)0) continue;
          const R c = dot(d, n);
          const v3<R> r = sub(d, scl((R)2 * c, n));
          const R ndotd = dot(r, F.n);
          const bool par = std::fabs(ndotd) < (R)1e-10;
          const R t = (F.ndotp - dot(p, F.n)) / (par ? (R)1 : ndotd);
          if (par || !(t > (R)0)) continue;
          const v3<R> op = sub(add(p, scl(t, r)), F.pos);
          const R x = dot(op, F.u1), y = dot(op, F.u2);
          const Bin b = S->kind == IACT_SENSOR_SQUARE ? bin_square<R, false>(S, x, y)
                                                      : bin_hex<R, false>(S, &F, x, y);
          if (b.pid < 0) continue;
          // v = value * (-c) / w * shadow; dL/dv = G[pix]
          const R gpix = (R)G[b.pid];
          if (gvp) gval += (double)(gpix * (-c) / w[6] * sh);
          const R k = gpix * (R)values[s] * sh / w[6];  // dL/d(-c)
          // dL/dn = -k d; dL/dd = -k n, through d = normalize(p - src)
          R* an = An.data() + 3 * m;
          an[0] -= k * d.x, an[1] -= k * d.y, an[2] -= k * d.z;
          if (source_type == IACT_SOURCE_POINT) {
            const R dn = dot(d, n);
            R* ap = Ap.data() + 3 * m;
            ap[0] -= k * inv_u * (n.x - dn * d.x);
            ap[1] -= k * inv_u * (n.y - dn * d.y);
            ap[2] -= k * inv_u * (n.z - dn * d.z);
          }
        }
        if (gvp) gvp[s] += gval;
      }
      if (!g_rot) continue;
      // per sample: n = pe / |pe|, pe = R (nl + s delta); p = R pl + pos
      R D[27];
      euler_grads<R>(Fc->rotations + 3 * f, D);
      double acc[3] = {0.0, 0.0, 0.0};
      for (int m = 0; m < M; ++m) {
        const size_t g = (size_t)f * M + m;
        const R* w = W.data() + 7 * m;
        const v3<R> n = V(w[3], w[4], w[5]);
        const v3<R> nls = add(load3<R>(Fc->normals + 3 * g),
                              scl((R)Fc->perturbation_scale[f], load3<R>(Fc->delta + 3 * g)));
        const v3<R> pl = load3<R>(Fc->points + 3 * g);
        const v3<R> pe = mul(Rm, nls);
        const R ipn = (R)1 / std::sqrt(dot(pe, pe));
        const v3<R> an = V(An[3 * m], An[3 * m + 1], An[3 * m + 2]);
        const v3<R> ap = V(Ap[3 * m], Ap[3 * m + 1], Ap[3 * m + 2]);
        // dL/dpe = (an - (an.n) n) / |pe|
        const v3<R> gpe = scl(ipn, sub(an, scl(dot(an, n), n)));
        for (int kk = 0; kk < 3; ++kk) {
          const v3<R> dpe = mul(D + 9 * kk, nls), dp = mul(D + 9 * kk, pl);
          acc[kk] += (double)(dot(gpe, dpe) + dot(ap, dp));
        }
      }
      const double k = 3.14159265358979323846 / 180.0;
      for (int kk = 0; kk < 3; ++kk) g_rot[3 * f + kk] = acc[kk] * k;
    }
  }
  if (g_val)
    for (int t = 0; t < nt; ++t)
      for (int i = 0; i < ns; ++i) g_val[i] += gv_priv[(size_t)t * ns + i];
  return 0;
}

}  // namespace

extern "C" {

/* dL/d(stage-0 rotations) (n_facets, 3) and dL/d(values) (n_sources) of
 * L = sum_p G[p] image[p], hard binning, primary-only scenes; fp32 arithmetic
 * with FP64 source directions.  Either output may be NULL.  -2: stage >= 1
 * mirrors or a splat sensor (not covered). */
int cpu_ref_grad(const iact_scene* sc, const double* sources, const float* values, int ns,
                 int source_type, const double* grad_image, double* grad_rotations,
                 double* grad_values, int nthreads) {
  if (!sc || !grad_image || (ns > 0 && (!sources || !values))) return -1;
  return grad_t<float>(sc, sources, values, ns, source_type, grad_image, grad_rotations,
                       grad_values, nthreads);
}

/* precision 32: the reference's fp32 arithmetic (with FP64 source directions
 * and Newton); 64: FP64 throughout.  image: (n_pixels) double.  band_image
 * and stats (binned rays, band rays, traced rays) may be NULL. */
int cpu_ref_render(const iact_scene* sc, const double* sources, const float* values, int ns,
                   int source_type, double* image, int nthreads, int precision, double band,
                   double* band_image, long long* stats) {
  if (!sc || !image || (ns > 0 && (!sources || !values))) return -1;
  const bool want_band = band > 0.0 && band_image;
  if (precision == 64)
    return want_band ? render_t<double, true>(sc, sources, values, ns, source_type, image,
                                              nthreads, band, band_image, stats)
                     : render_t<double, false>(sc, sources, values, ns, source_type, image,
                                               nthreads, band, band_image, stats);
  if (precision == 32)
    return want_band ? render_t<float, true>(sc, sources, values, ns, source_type, image,
                                             nthreads, band, band_image, stats)
                     : render_t<float, false>(sc, sources, values, ns, source_type, image,
                                              nthreads, band, band_image, stats);
  return -1;
}

int cpu_ref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
