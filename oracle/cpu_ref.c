/*
 * cpu_ref -- C/OpenMP restatement of the reference renderer's image path.
 * TEST INFRASTRUCTURE ONLY: the CPU baseline timed by bench.py and a second,
 * independent restatement checked against oracle/iact_oracle.py in tests.
 *
 * Follows iactrace/core/render.py:118-220 in order: for every primary facet
 * (lax.scan, render.py:210-219) and every (source, sample) ray:
 * source direction (fp64 for point sources, render.py:129-133), occlusion
 * toward the source (render.py:138, 21-41), reflect (reflection.py:5-19),
 * value = value * cos / w * shadow (render.py:141), stage>=1 mirrors
 * (render.py:44-115; fp64 conic guess + Newton, intersections.py:229-367),
 * sensor plane (intersections.py:6-41) and hard binning (square.py:66-91,
 * hexagonal.py:174-194).  fp32 arithmetic with IEEE division and square root
 * as XLA-CPU does it; images are summed per thread then added.
 *
 * It reads the scene through the same descriptor layout as the HIP library
 * (include/iactrace_amd.h) but with HOST pointers.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/iactrace_amd.h"

typedef struct {
  float x, y, z;
} v3;

static inline v3 V(float x, float y, float z) {
  v3 r = {x, y, z};
  return r;
}
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(float s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 mul(const float* R, v3 v) {
  return V(R[0] * v.x + R[1] * v.y + R[2] * v.z, R[3] * v.x + R[4] * v.y + R[5] * v.z,
           R[6] * v.x + R[7] * v.y + R[8] * v.z);
}
static inline v3 mulT(const float* R, v3 v) {
  return V(R[0] * v.x + R[3] * v.y + R[6] * v.z, R[1] * v.x + R[4] * v.y + R[7] * v.z,
           R[2] * v.x + R[5] * v.y + R[8] * v.z);
}

/* transforms.py:72-107 */
static void euler(const float* a, float* R) {
  const float k = (float)(3.14159265358979323846 / 180.0);
  float rx = a[0] * k, ry = a[1] * k, rz = a[2] * k;
  float cx = cosf(rx), sx = sinf(rx), cy = cosf(ry), sy = sinf(ry), cz = cosf(rz), sz = sinf(rz);
  float A[9] = {cz * cy, -sz, cz * sy, sz * cy, cz, sz * sy, -sy, 0.f, cy}; /* Rz Ry */
  R[0] = A[0];
  R[1] = A[1] * cx + A[2] * sx;
  R[2] = -A[1] * sx + A[2] * cx;
  R[3] = A[3];
  R[4] = A[4] * cx + A[5] * sx;
  R[5] = -A[4] * sx + A[5] * cx;
  R[6] = A[6];
  R[7] = A[7] * cx + A[8] * sx;
  R[8] = -A[7] * sx + A[8] * cx;
}

/* ---------------------------------------------------------------- surfaces */
static double sag64(const iact_surface* s, double X, double Y) {
  double c = s->curvature, k = s->conic, r2 = X * X + Y * Y;
  double z = r2 * c / (1.0 + sqrt(1.0 - (1.0 + k) * c * c * r2));
  double r4 = r2 * r2, p = r4;
  for (int i = 0; i < s->n_aspheric; ++i) {
    z += (double)s->aspheric[i] * p;
    p *= r4;
  }
  return z;
}
static double dsag64(const iact_surface* s, double r2) { /* d sag / d r2 */
  double c = s->curvature, k = s->conic, K = (1.0 + k) * c * c;
  double sq = sqrt(1.0 - K * r2), den = 1.0 + sq;
  double g = c / den + (r2 * c) * K / (2.0 * sq * den * den);
  double r4 = r2 * r2, q = 1.0;
  for (int i = 0; i < s->n_aspheric; ++i) {
    g += (double)s->aspheric[i] * (double)(2 * (i + 1)) * r2 * q;
    q *= r4;
  }
  return g;
}
static float sag32(const iact_surface* s, float X, float Y) {
  float c = s->curvature, k = s->conic, r2 = X * X + Y * Y;
  float z = r2 * c / (1.f + sqrtf(1.f - (1.f + k) * c * c * r2));
  float r4 = r2 * r2, p = r4;
  for (int i = 0; i < s->n_aspheric; ++i) {
    z += s->aspheric[i] * p;
    p *= r4;
  }
  return z;
}
static float dsag32(const iact_surface* s, float r2) {
  float c = s->curvature, k = s->conic, K = (1.f + k) * c * c;
  float sq = sqrtf(1.f - K * r2), den = 1.f + sq;
  float g = c / den + (r2 * c) * K / (2.f * sq * den * den);
  float r4 = r2 * r2, q = 1.f;
  for (int i = 0; i < s->n_aspheric; ++i) {
    g += s->aspheric[i] * (float)(2 * (i + 1)) * r2 * q;
    q *= r4;
  }
  return g;
}

static double conic_root(double ox, double oy, double oz, double dx, double dy, double dz,
                         double c, double k) {
  double ok = 1.0 + k;
  double A = c * (dx * dx + dy * dy + ok * dz * dz);
  double B = 2.0 * (c * (ox * dx + oy * dy + ok * oz * dz) - dz);
  double C = c * (ox * ox + oy * oy + ok * oz * oz) - 2.0 * oz;
  if (fabs(c) < 1e-12) return fabs(dz) > 1e-10 ? -oz / dz : INFINITY;
  double disc = B * B - 4.0 * A * C;
  if (disc < 0) return INFINITY;
  double sq = sqrt(disc);
  double t1 = (-B - sq) / (2.0 * A + 1e-30), t2 = (-B + sq) / (2.0 * A + 1e-30);
  int v1 = t1 > 1e-8, v2 = t2 > 1e-8;
  return (v1 && v2) ? fmin(t1, t2) : (v1 ? t1 : (v2 ? t2 : INFINITY));
}

/* surfaces.py:67-108 + intersections.py:290-367 (fp64 refinement) */
static float mirror_hit(const iact_surface* s, const float* M, v3 ol, v3 dl, v3* pl, v3* nl) {
  double x0 = M[15], y0 = M[16];
  double z0 = sag64(s, x0, y0);
  double t = conic_root(ol.x + x0, ol.y + y0, ol.z + z0, dl.x, dl.y, dl.z, s->curvature,
                        s->conic);
  int conv = 0;
  for (int it = 0; it < 10; ++it) {
    double X = ol.x + t * dl.x + x0, Y = ol.y + t * dl.y + y0;
    double r2 = X * X + Y * Y;
    double gv = (ol.z + t * dl.z) - (sag64(s, X, Y) - z0);
    double g = dsag64(s, r2);
    double gp = dl.z - (2.0 * X * g * dl.x + 2.0 * Y * g * dl.y);
    if (!(fabs(gp) > 1e-12)) gp = 1e-12;
    double tn = t - gv / gp;
    int nc = conv || fabs(gv) < 1e-8;
    if (!conv) t = tn;
    conv = nc;
  }
  double X = ol.x + t * dl.x + x0, Y = ol.y + t * dl.y + y0;
  double res = fabs((ol.z + t * dl.z) - (sag64(s, X, Y) - z0));
  int valid = (t > 1e-8) && (res < 1e-6);
  float xh = (float)(ol.x + t * dl.x), yh = (float)(ol.y + t * dl.y);
  float fx0 = M[15], fy0 = M[16];
  float Xs = xh + fx0, Ys = yh + fy0;
  float zs = sag32(s, Xs, Ys) - sag32(s, fx0, fy0);
  float gs = dsag32(s, Xs * Xs + Ys * Ys);
  float zx = 2.f * Xs * gs, zy = 2.f * Ys * gs;
  float nn = sqrtf(zx * zx + zy * zy + 1.f);
  *pl = V(xh, yh, zs);
  *nl = V(-zx / nn, -zy / nn, 1.f / nn);
  int inside = 1;
  if (M[17] == 0.f) {
    inside = xh * xh + yh * yh <= M[18] * M[18];
  } else {
    int n = (int)M[19];
    for (int i = 0; i < n; ++i) {
      int j = (i + 1) % n;
      float ax = M[20 + 2 * i], ay = M[21 + 2 * i], bx = M[20 + 2 * j], by = M[21 + 2 * j];
      inside = inside && ((bx - ax) * (yh - ay) - (by - ay) * (xh - ax) >= 0.f);
    }
  }
  return (valid && inside) ? (float)t : INFINITY;
}

/* ---------------------------------------------------------------- obstructions */
static float t_cyl(v3 o, v3 d, const float* c) {
  const float eps = 1e-8f;
  v3 p1 = V(c[0], c[1], c[2]), ax = V(c[3], c[4], c[5]);
  float h = c[6], r = c[7];
  v3 oc = sub(o, p1);
  float oca = dot(oc, ax), rda = dot(d, ax);
  v3 ocp = sub(oc, scl(oca, ax)), rdp = sub(d, scl(rda, ax));
  float a = dot(rdp, rdp), b = 2.f * dot(ocp, rdp), cc = dot(ocp, ocp) - r * r;
  float disc = b * b - 4.f * a * cc;
  float sq = sqrtf(fmaxf(disc, 0.f));
  float t1 = (-b - sq) / (2.f * a + eps), t2 = (-b + sq) / (2.f * a + eps);
  float y1 = oca + t1 * rda, y2 = oca + t2 * rda;
  t1 = (t1 > eps && y1 >= 0 && y1 <= h && disc >= 0) ? t1 : INFINITY;
  t2 = (t2 > eps && y2 >= 0 && y2 <= h && disc >= 0) ? t2 : INFINITY;
  float tb = -oca / (rda + eps), tt = (h - oca) / (rda + eps);
  v3 pb = add(ocp, scl(tb, rdp)), pt = add(ocp, scl(tt, rdp));
  tb = (tb > eps && dot(pb, pb) <= r * r) ? tb : INFINITY;
  tt = (tt > eps && dot(pt, pt) <= r * r) ? tt : INFINITY;
  return fminf(fminf(t1, t2), fminf(tb, tt));
}
static float t_box(v3 o, v3 d, const float* b) {
  const float eps = 1e-8f;
  float tx1 = (b[0] - o.x) * (1.f / (d.x + eps)), tx2 = (b[3] - o.x) * (1.f / (d.x + eps));
  float ty1 = (b[1] - o.y) * (1.f / (d.y + eps)), ty2 = (b[4] - o.y) * (1.f / (d.y + eps));
  float tz1 = (b[2] - o.z) * (1.f / (d.z + eps)), tz2 = (b[5] - o.z) * (1.f / (d.z + eps));
  float tmin = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2));
  float tmax = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2));
  return ((tmax >= tmin) && (tmax > eps)) ? (tmin > eps ? tmin : tmax) : INFINITY;
}
static float t_sph(v3 o, v3 d, const float* s) {
  const float eps = 1e-8f;
  v3 oc = sub(o, V(s[0], s[1], s[2]));
  float a = dot(d, d), b = 2.f * dot(oc, d), c = dot(oc, oc) - s[3] * s[3];
  float disc = b * b - 4.f * a * c, sq = sqrtf(fmaxf(disc, 0.f));
  float t1 = (-b - sq) / (2.f * a + eps), t2 = (-b + sq) / (2.f * a + eps);
  return fminf((t1 > eps && disc >= 0) ? t1 : INFINITY, (t2 > eps && disc >= 0) ? t2 : INFINITY);
}
static float sgn(float x) { return x > 0 ? 1.f : (x < 0 ? -1.f : 0.f); }
static float t_obox(v3 o, v3 d, const float* q) {
  const float eps = 1e-8f;
  v3 lo = mulT(q + 6, sub(o, V(q[0], q[1], q[2]))), ld = mulT(q + 6, d);
  float ix = 1.f / (ld.x + eps * sgn(ld.x + eps)), iy = 1.f / (ld.y + eps * sgn(ld.y + eps));
  float iz = 1.f / (ld.z + eps * sgn(ld.z + eps));
  float tx1 = (-q[3] - lo.x) * ix, tx2 = (q[3] - lo.x) * ix;
  float ty1 = (-q[4] - lo.y) * iy, ty2 = (q[4] - lo.y) * iy;
  float tz1 = (-q[5] - lo.z) * iz, tz2 = (q[5] - lo.z) * iz;
  float tmin = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2));
  float tmax = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2));
  float tr = tmin > eps ? tmin : tmax;
  return ((tmax >= tmin) && (tmax > eps) && tr > eps) ? tr : INFINITY;
}
static float t_tri(v3 o, v3 d, const float* v) {
  const float eps = 1e-8f;
  v3 v0 = V(v[0], v[1], v[2]), e1 = sub(V(v[3], v[4], v[5]), v0), e2 = sub(V(v[6], v[7], v[8]), v0);
  v3 h = cross(d, e2);
  float a = dot(e1, h);
  float f = 1.f / (a + eps * sgn(a + eps));
  v3 s = sub(o, v0);
  float u = f * dot(s, h);
  v3 q = cross(s, e1);
  float w = f * dot(d, q), t = f * dot(e2, q);
  int ok = !(fabsf(a) < eps) && u >= 0 && u <= 1 && w >= 0 && u + w <= 1 && t > eps;
  return ok ? t : INFINITY;
}
static float shadow(const iact_obstructions* ob, v3 o, v3 d) {
  for (int i = 0; i < ob->n_cylinders; ++i)
    if (t_cyl(o, d, ob->cylinders + 8 * i) < 1e10f) return 0.f;
  for (int i = 0; i < ob->n_boxes; ++i)
    if (t_box(o, d, ob->boxes + 6 * i) < 1e10f) return 0.f;
  for (int i = 0; i < ob->n_spheres; ++i)
    if (t_sph(o, d, ob->spheres + 4 * i) < 1e10f) return 0.f;
  for (int i = 0; i < ob->n_oriented_boxes; ++i)
    if (t_obox(o, d, ob->oriented_boxes + 15 * i) < 1e10f) return 0.f;
  for (int i = 0; i < ob->n_triangles; ++i)
    if (t_tri(o, d, ob->triangles + 9 * i) < 1e10f) return 0.f;
  return 1.f;
}

/* ---------------------------------------------------------------- sensors */
typedef struct {
  v3 pos, u1, u2, n;
  float ndotp, gc, gs;
} frame_t;

static int bin_square(const iact_sensor* S, float x, float y) {
  float xc = (x - S->x0) / S->dx, yc = (y - S->y0) / S->dy;
  float xf = floorf(xc), yf = floorf(yc);
  int ok = xf >= 0 && xf < (float)S->width && yf >= 0 && yf < (float)S->height;
  float xr = xc - xf, yr = yc - yf;
  float dist = fminf(fminf(xr, 1.f - xr) * S->dx, fminf(yr, 1.f - yr) * S->dy);
  ok = ok && !(dist < S->edge_width);
  return ok ? (int)yf * S->width + (int)xf : -1;
}
static int bin_hex(const iact_sensor* S, const frame_t* F, float x, float y) {
  float ux = x - S->grid_offset[0], uy = y - S->grid_offset[1];
  float xg = F->gc * ux - F->gs * uy, yg = F->gs * ux + F->gc * uy;
  float q = (0.5773502691896257f * xg - yg / 3.f) / S->hex_size;
  float r = (2.f * yg / 3.f) / S->hex_size;
  if (!(fabsf(q) < 1e6f && fabsf(r) < 1e6f)) return -1;
  float s = -q - r, qi = nearbyintf(q), ri = nearbyintf(r), si = nearbyintf(s);
  float dq = fabsf(qi - q), dr = fabsf(ri - r), ds = fabsf(si - s);
  if (dq > dr && dq > ds)
    qi = -ri - si;
  else if (dr > dq && dr > ds)
    ri = -qi - si;
  float qa = qi - (float)S->q_min, ra = ri - (float)S->r_min;
  if (!(qa >= 0 && qa < (float)S->table_q && ra >= 0 && ra < (float)S->table_r)) return -1;
  int pid = S->lookup_table[(int)qa * S->table_r + (int)ra];
  if (pid < 0) return -1;
  float cx = (float)((double)S->hex_size * 1.7320508075688772) * (qi + ri / 2.f);
  float cy = (float)((double)S->hex_size * 1.5) * ri;
  float ax = fabsf(xg - cx);
  float hd = fmaxf(ax, 0.5f * ax + 0.8660254037844386f * fabsf(yg - cy)) / S->hex_inradius;
  return hd > S->edge_threshold ? -1 : pid;
}

/* ---------------------------------------------------------------- render */
int cpu_ref_render(const iact_scene* sc, const double* sources, const float* values, int ns,
                   int source_type, float* image, int nthreads) {
  const iact_facets* Fc = &sc->facets;
  const iact_sensor* S = &sc->sensor;
  const int P = S->n_pixels, nf = Fc->n_facets, M = Fc->n_samples;
  memset(image, 0, sizeof(float) * P);
  if (nf == 0 || ns == 0) return 0;
  frame_t F;
  F.pos = V(S->position[0], S->position[1], S->position[2]);
  F.u1 = V(S->rotation[0], S->rotation[3], S->rotation[6]);
  F.u2 = V(S->rotation[1], S->rotation[4], S->rotation[7]);
  F.n = V(S->rotation[2], S->rotation[5], S->rotation[8]);
  F.ndotp = F.n.x * F.pos.x + F.n.y * F.pos.y + F.n.z * F.pos.z;
  F.gc = cosf(-S->grid_rotation);
  F.gs = sinf(-S->grid_rotation);
  /* world samples (mirrors.py:64-79) */
  float* W = (float*)malloc(sizeof(float) * 7 * (size_t)nf * M);
  if (!W) return -1;
  for (int f = 0; f < nf; ++f) {
    float R[9];
    euler(Fc->rotations + 3 * f, R);
    v3 pos = V(Fc->positions[3 * f], Fc->positions[3 * f + 1], Fc->positions[3 * f + 2]);
    for (int m = 0; m < M; ++m) {
      size_t g = (size_t)f * M + m;
      v3 p = add(mul(R, V(Fc->points[3 * g], Fc->points[3 * g + 1], Fc->points[3 * g + 2])), pos);
      v3 n = mul(R, V(Fc->normals[3 * g], Fc->normals[3 * g + 1], Fc->normals[3 * g + 2]));
      v3 dl = mul(R, V(Fc->delta[3 * g], Fc->delta[3 * g + 1], Fc->delta[3 * g + 2]));
      v3 pe = add(n, scl(Fc->perturbation_scale[f], dl));
      float nn = sqrtf(dot(pe, pe));
      float* w = W + 7 * g;
      w[0] = p.x; w[1] = p.y; w[2] = p.z;
      w[3] = pe.x / nn; w[4] = pe.y / nn; w[5] = pe.z / nn;
      w[6] = Fc->weights[g];
    }
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  float* priv = (float*)calloc((size_t)nt * P, sizeof(float));
  if (!priv) {
    free(W);
    return -1;
  }
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    float* img = priv + (size_t)tid * P;
#pragma omp for schedule(dynamic, 1) collapse(2)
    for (int f = 0; f < nf; ++f)
      for (int s = 0; s < ns; ++s) {
        for (int m = 0; m < M; ++m) {
          const float* w = W + 7 * ((size_t)f * M + m);
          v3 p = V(w[0], w[1], w[2]), n = V(w[3], w[4], w[5]);
          v3 d;
          if (source_type == IACT_SOURCE_POINT) {
            double dx = (double)p.x - sources[3 * s], dy = (double)p.y - sources[3 * s + 1];
            double dz = (double)p.z - sources[3 * s + 2];
            double nr = sqrt(dx * dx + dy * dy + dz * dz);
            d = V((float)(dx / nr), (float)(dy / nr), (float)(dz / nr));
          } else {
            d = V((float)sources[3 * s], (float)sources[3 * s + 1], (float)sources[3 * s + 2]);
          }
          float sh = shadow(&sc->obstructions, p, scl(-1.f, d));
          float c = dot(d, n);
          v3 r = sub(d, scl(2.f * c, n));
          float v = values[s] * (-c) / w[6] * sh;
          v3 o = p;
          for (int k = 0; k < sc->n_stages; ++k) {
            float best = INFINITY;
            v3 bp = V(0, 0, 0), bn = V(0, 0, 0);
            for (int i = 0; i < sc->n_mirrors; ++i) {
              const float* Mr = sc->mirrors + (size_t)IACT_MIRROR_STRIDE * i;
              if ((int)Mr[36] != sc->stages[k]) continue;
              const float* R = Mr + 6;
              v3 pos = V(Mr[0], Mr[1], Mr[2]);
              v3 pl, nl;
              float t = mirror_hit(&sc->surfaces[(int)Mr[38]], Mr, mulT(R, sub(o, pos)),
                                   mulT(R, r), &pl, &nl);
              if (t < best) {
                best = t;
                bp = add(mul(R, pl), pos);
                bn = mul(R, nl);
              }
            }
            float cc = dot(r, bn);
            v3 rr = sub(r, scl(2.f * cc, bn));
            v = v * (best < 1e10f ? 1.f : 0.f) * shadow(&sc->obstructions, o, r) * fabsf(cc);
            o = bp;
            r = rr;
          }
          float ndotd = dot(r, F.n);
          int par = fabsf(ndotd) < 1e-10f;
          float t = (F.ndotp - dot(o, F.n)) / (par ? 1.f : ndotd);
          if (par || !(t > 0.f)) continue;
          v3 op = sub(add(o, scl(t, r)), F.pos);
          float x = dot(op, F.u1), y = dot(op, F.u2);
          int pid = S->kind == IACT_SENSOR_SQUARE ? bin_square(S, x, y) : bin_hex(S, &F, x, y);
          if (pid >= 0) img[pid] += v;
        }
      }
  }
  for (int t = 0; t < nt; ++t)
    for (int i = 0; i < P; ++i) image[i] += priv[(size_t)t * P + i];
  free(priv);
  free(W);
  return 0;
}

int cpu_ref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
