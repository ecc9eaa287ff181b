/*
 * traverse.c -- oracle restatement of vitrum's f32 ray traversal.
 * TEST INFRASTRUCTURE ONLY (see vt_oracle.h).
 *
 * Built with -ffp-contract=off on x86-64 (SSE: every float operation rounds
 * once to binary32), which reproduces the `Math.fround`-per-operation
 * discipline of packages/shared-bvh/src/opticalWatertightTriangle.ts.
 */
#include "vt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

int orc_fail(const char *fmt, ...);

#define LEAF_FLAG 0xFFFF0000u
#define BVH_STACK_DEPTH 60  /* strides.ts:110 */
#define TLAS_STACK_DEPTH 64 /* strides.ts:112 */

typedef struct {
  float x, y, z;
} v3;

static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static float fabsf_(float x) { return x < 0 ? -x : (x == 0 ? 0.0f : x); }
static float fmax2(float a, float b) { return a > b ? a : b; }
static float fmin2(float a, float b) { return a < b ? a : b; }
/* Math.max semantics: NaN-propagating (the TS mirror's behaviour). */
static float jmax(float a, float b) {
  if (isnan(a) || isnan(b)) return NAN;
  return a > b ? a : b;
}
static float max3abs(v3 v) { return jmax(fabsf_(v.x), jmax(fabsf_(v.y), fabsf_(v.z))); }
static v3 sub(v3 a, v3 b) { return (v3){a.x - b.x, a.y - b.y, a.z - b.z}; }
static float comp(v3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
static float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* 3.402823e38 and 3.402823466e38 as binary32 (the WGSL literals). */
static float k_safe_max(void) { return u2f(0x7F7FFFFDu); }
static float k_infinity(void) { return u2f(0x7F7FFFFFu); }

/* ---- opticalWatertightTriangle.ts:15-96 -------------------------------- */

typedef struct {
  float x, y;
} v2;

/* edgeFunction (opticalWatertightTriangle.ts:19-26). */
static float edge_fn(v2 a, v2 b) {
  if (a.x == b.x && a.y == b.y) return 0.0f;
  int ordered = a.x < b.x || (a.x == b.x && a.y < b.y);
  v2 f = ordered ? a : b, s = ordered ? b : a;
  float canonical = f.x * s.y - f.y * s.x;
  return ordered ? canonical : -canonical;
}

/* rayParameter (opticalWatertightTriangle.ts:28-51). */
static float ray_param(v3 o, v3 dir, v3 p) {
  v3 delta = sub(p, o);
  float ds = max3abs(delta);
  float rs = max3abs(dir);
  if (!(rs > 0) || !isfinite(rs) || !isfinite(ds)) return -1.0f;
  if (!(ds > 0)) return 0.0f;
  v3 d = {dir.x / rs, dir.y / rs, dir.z / rs};
  v3 q = {delta.x / ds, delta.y / ds, delta.z / ds};
  float num = q.x * d.x + q.y * d.y + q.z * d.z;
  float den = d.x * d.x + d.y * d.y + d.z * d.z;
  return (num / den) * (ds / rs);
}

static int lex_before(v3 a, v3 b) {
  return a.x < b.x || (a.x == b.x && (a.y < b.y || (a.y == b.y && a.z < b.z)));
}

/* edgeRayParameter (opticalWatertightTriangle.ts:59-87). */
static float edge_ray_param(v3 o, v3 dir, v3 fi, v3 si, v2 fpi, v2 spi, v3 *point) {
  int ordered = lex_before(fi, si);
  v3 f = ordered ? fi : si, s = ordered ? si : fi;
  v2 fp = ordered ? fpi : spi, sp = ordered ? spi : fpi;
  v2 delta = {sp.x - fp.x, sp.y - fp.y};
  float interp;
  if (fabsf_(delta.x) >= fabsf_(delta.y)) {
    if (delta.x == 0) {
      *point = (v3){0, 0, 0};
      return -1.0f;
    }
    interp = -fp.x / delta.x;
  } else {
    if (delta.y == 0) {
      *point = (v3){0, 0, 0};
      return -1.0f;
    }
    interp = -fp.y / delta.y;
  }
  *point = (v3){f.x + interp * (s.x - f.x), f.y + interp * (s.y - f.y), f.z + interp * (s.z - f.z)};
  return ray_param(o, dir, *point);
}

typedef struct {
  int hit;
  float t;
  v3 point;
  v3 bary;
  float side;
  uint32_t zmask;
} wt_hit;

/* intersectOpticalTriangleWatertightF32 (opticalWatertightTriangle.ts:98-220).
 * A miss carries t = +Infinity, as in the reference's MISS record. */
static wt_hit watertight(v3 o, v3 dir, v3 a, v3 b, v3 c, float exclusive_min_t) {
  wt_hit miss = {0, INFINITY, {0, 0, 0}, {0, 0, 0}, 0, 0};
  float ds = max3abs(dir);
  if (!(ds > 0) || !isfinite(ds) || isnan(exclusive_min_t)) return miss;
  v3 d = {dir.x / ds, dir.y / ds, dir.z / ds};
  int kz = fabsf_(d.y) > fabsf_(d.x) ? 1 : 0;
  if (fabsf_(d.z) > fabsf_(comp(d, kz))) kz = 2;
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (comp(d, kz) < 0) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  float dz = comp(d, kz);
  float sx = comp(d, kx) / dz, sy = comp(d, ky) / dz, sz = 1.0f / dz;
  v3 pa = sub(a, o), pb = sub(b, o), pc = sub(c, o);
  v2 ap = {comp(pa, kx) - sx * comp(pa, kz), comp(pa, ky) - sy * comp(pa, kz)};
  v2 bp = {comp(pb, kx) - sx * comp(pb, kz), comp(pb, ky) - sy * comp(pb, kz)};
  v2 cp = {comp(pc, kx) - sx * comp(pc, kz), comp(pc, ky) - sy * comp(pc, kz)};
  float e0 = edge_fn(bp, cp), e1 = edge_fn(cp, ap), e2 = edge_fn(ap, bp);
  float det = (e0 + e1) + e2;
  if (det == 0 || !isfinite(det)) return miss;
  int positive = det > 0;
  float n0 = positive ? e0 : -e0, n1 = positive ? e1 : -e1, n2 = positive ? e2 : -e2;
  if (n0 < 0 || n1 < 0 || n2 < 0) return miss;
  float az = comp(pa, kz) * sz, bz = comp(pb, kz) * sz, cz = comp(pc, kz) * sz;
  float scaled_t = (e0 * az + e1 * bz) + e2 * cz;
  float t = (scaled_t / det) / ds;
  v3 point = {0, 0, 0};
  int zc = (n0 == 0) + (n1 == 0) + (n2 == 0);
  if (zc >= 2) {
    if (n0 != 0) point = a, t = ray_param(o, dir, point);
    if (n1 != 0) point = b, t = ray_param(o, dir, point);
    if (n2 != 0) point = c, t = ray_param(o, dir, point);
  } else if (n0 == 0) {
    t = edge_ray_param(o, dir, b, c, bp, cp, &point);
  } else if (n1 == 0) {
    t = edge_ray_param(o, dir, c, a, cp, ap, &point);
  } else if (n2 == 0) {
    t = edge_ray_param(o, dir, a, b, ap, bp, &point);
  }
  float lo = exclusive_min_t > 0 ? exclusive_min_t : 0.0f; /* Math.max(exclusiveMinT, 0) */
  if (!(t > lo) || !isfinite(t)) return miss;
  v3 e1r = sub(b, a), e2r = sub(c, a);
  float es = jmax(max3abs(e1r), max3abs(e2r));
  if (!(es > 0) || !isfinite(es)) return miss;
  v3 s1 = {e1r.x / es, e1r.y / es, e1r.z / es};
  v3 s2 = {e2r.x / es, e2r.y / es, e2r.z / es};
  v3 nr = {s1.y * s2.z - s1.z * s2.y, s1.z * s2.x - s1.x * s2.z, s1.x * s2.y - s1.y * s2.x};
  float ns = max3abs(nr);
  if (!(ns > 0) || !isfinite(ns)) return miss;
  v3 nd = {nr.x / ns, nr.y / ns, nr.z / ns};
  float nl2 = (nd.x * nd.x + nd.y * nd.y) + nd.z * nd.z;
  float nl = sqrtf(nl2);
  if (!(nl > 0) || !isfinite(nl)) return miss;
  v3 n = {nd.x / nl, nd.y / nl, nd.z / nl};
  float side_det = -((d.x * n.x + d.y * n.y) + d.z * n.z);
  if (side_det == 0 || !isfinite(side_det)) return miss;
  v3 bary = {e0 / det, e1 / det, e2 / det};
  if (zc == 0) {
    point = (v3){a.x + (bary.y * (b.x - a.x) + bary.z * (c.x - a.x)),
                 a.y + (bary.y * (b.y - a.y) + bary.z * (c.y - a.y)),
                 a.z + (bary.y * (b.z - a.z) + bary.z * (c.z - a.z))};
  }
  wt_hit h;
  h.hit = 1;
  h.t = t;
  h.point = point;
  h.bary = bary;
  h.side = side_det > 0 ? 1.0f : -1.0f;
  h.zmask = (n0 == 0 ? 1u : 0u) | (n1 == 0 ? 2u : 0u) | (n2 == 0 ? 4u : 0u);
  return h;
}

void orc_watertight(const float *origin, const float *direction, const float *a, const float *b,
                    const float *c, float exclusive_min_t, float *out) {
  v3 o = {origin[0], origin[1], origin[2]}, d = {direction[0], direction[1], direction[2]};
  v3 A = {a[0], a[1], a[2]}, B = {b[0], b[1], b[2]}, C = {c[0], c[1], c[2]};
  wt_hit h = watertight(o, d, A, B, C, exclusive_min_t);
  out[0] = (float)h.hit;
  out[1] = h.t;
  out[2] = h.point.x;
  out[3] = h.point.y;
  out[4] = h.point.z;
  out[5] = h.bary.x;
  out[6] = h.bary.y;
  out[7] = h.bary.z;
  out[8] = h.side;
  out[9] = (float)h.zmask;
}

/* ---- bvhIntersect.wgsl.ts:86-111 safeInvDir; intersectionCore.wgsl.ts:24-31 */

static float safe_inv_comp(float v) {
  float m = k_safe_max();
  if (v == 0.0f) return v >= 0.0f ? m : -m;
  float r = 1.0f / v;
  return r < -m ? -m : (r > m ? m : r);
}

typedef struct {
  v3 o, d, inv;
} ray3;

static int slab(const ray3 *r, const uint32_t *node, float tmin, float tmax, float *t_near) {
  v3 bmin = {u2f(node[0]), u2f(node[1]), u2f(node[2])};
  v3 bmax = {u2f(node[3]), u2f(node[4]), u2f(node[5])};
  v3 t1 = {(bmin.x - r->o.x) * r->inv.x, (bmin.y - r->o.y) * r->inv.y, (bmin.z - r->o.z) * r->inv.z};
  v3 t2 = {(bmax.x - r->o.x) * r->inv.x, (bmax.y - r->o.y) * r->inv.y, (bmax.z - r->o.z) * r->inv.z};
  float tn = fmax2(fmax2(fmin2(t1.x, t2.x), fmin2(t1.y, t2.y)), fmin2(t1.z, t2.z));
  float tf = fmin2(fmin2(fmax2(t1.x, t2.x), fmax2(t1.y, t2.y)), fmax2(t1.z, t2.z));
  if (t_near) *t_near = tn;
  return !(tn > tf || tf < tmin || tn > tmax);
}

/* ---- common.wgsl.ts safe_normalize; intersectionCore.wgsl.ts:68-83 ------ */

static v3 safe_normalize(v3 v) {
  float s = max3abs(v);
  if (!(s > 0) || s > k_safe_max()) return (v3){0, 1, 0};
  v3 q = {v.x / s, v.y / s, v.z / s};
  float len = sqrtf((q.x * q.x + q.y * q.y) + q.z * q.z);
  return (v3){q.x / len, q.y / len, q.z / len};
}

static v3 xform_point_cols(const float *m, v3 p) {
  float raw[4];
  for (int k = 0; k < 4; ++k) raw[k] = ((m[k] * p.x + m[4 + k] * p.y) + m[8 + k] * p.z) + m[12 + k];
  float s = fmax2(fmax2(fabsf_(raw[0]), fabsf_(raw[1])), fmax2(fabsf_(raw[2]), fabsf_(raw[3])));
  if (!(s > 0) || s > k_safe_max()) return (v3){0, 0, 0};
  float r[4] = {raw[0] / s, raw[1] / s, raw[2] / s, raw[3] / s};
  if (r[3] == 0) return (v3){0, 0, 0};
  v3 q = {r[0] / r[3], r[1] / r[3], r[2] / r[3]};
  if (isnan(q.x) || isnan(q.y) || isnan(q.z) || fabsf_(q.x) > k_safe_max() ||
      fabsf_(q.y) > k_safe_max() || fabsf_(q.z) > k_safe_max())
    return (v3){0, 0, 0};
  return q;
}

static v3 xform_dir_cols(const float *m, v3 d) {
  v3 v = {(m[0] * d.x + m[4] * d.y) + m[8] * d.z, (m[1] * d.x + m[5] * d.y) + m[9] * d.z,
          (m[2] * d.x + m[6] * d.y) + m[10] * d.z};
  return safe_normalize(v);
}

/* ---- traceMeshBvh: intersectionCore.wgsl.ts:461-596 --------------------- */

typedef struct {
  int did_hit;
  float dist;
  v3 position;
  uint32_t tri;
  float u, v;
  int front; /* triHit.side > 0 (intersectionCore.wgsl.ts:547) */
} local_hit;

typedef struct {
  uint32_t blas_nodes, tris, tlas_nodes, instances;
} counters;

static v3 vload(const orc_scene *s, uint32_t vi) {
  const float *p = s->positions + 4 * (size_t)vi;
  return (v3){p[0], p[1], p[2]};
}

static int trace_mesh_bvh(const orc_scene *s, const ray3 *r, float tmin, float tmax_bound,
                          int closest, local_hit *hit, uint32_t root, counters *cnt) {
  hit->did_hit = 0;
  hit->dist = tmax_bound;
  hit->position = (v3){0, 0, 0};
  hit->tri = 0;
  hit->u = hit->v = 0;
  hit->front = 0;
  if (s->bvh_node_count == 0 || root >= s->bvh_node_count) return 0;
  uint32_t stack[BVH_STACK_DEPTH];
  uint32_t sp = 0;
  stack[sp++] = root;
  while (sp > 0) {
    uint32_t ni = stack[--sp];
    if (ni >= s->bvh_node_count) continue;
    const uint32_t *node = s->bvh_nodes + 8 * (size_t)ni;
    float far_bound = closest ? hit->dist : tmax_bound;
    if (cnt) cnt->blas_nodes++;
    if (!slab(r, node, tmin, far_bound, NULL)) continue;
    uint32_t soc = node[7];
    if ((soc & LEAF_FLAG) == LEAF_FLAG) {
      uint32_t count = soc & 0xFFFFu, start = node[6];
      for (uint32_t i = 0; i < count; ++i) {
        uint32_t t = start + i;
        if (t >= s->triangle_count) continue;
        const uint32_t *tri = s->indices + 4 * (size_t)t;
        if (tri[0] >= s->vertex_count || tri[1] >= s->vertex_count || tri[2] >= s->vertex_count)
          continue;
        if (cnt) cnt->tris++;
        wt_hit th = watertight(r->o, r->d, vload(s, tri[0]), vload(s, tri[1]), vload(s, tri[2]), tmin);
        float ht = th.t;
        if (ht > tmin && ht < (closest ? hit->dist : tmax_bound)) {
          if (!closest) return 1;
          hit->did_hit = 1;
          hit->dist = ht;
          hit->position = th.point;
          hit->tri = t;
          hit->u = th.bary.y;
          hit->v = th.bary.z;
          hit->front = th.side > 0.0f;
        }
      }
    } else {
      uint32_t left = ni + 1, right = ni + node[6];
      if (sp + 2 <= BVH_STACK_DEPTH) {
        stack[sp++] = right;
        stack[sp++] = left;
      } else {
        return closest ? hit->did_hit : 1;
      }
    }
  }
  return closest ? hit->did_hit : 0;
}

/* ---- tlasSceneHitTraversal.wgsl.ts:92-269, pt-webgpu patch ------------- */

static ray3 make_ray(v3 o, v3 d) {
  ray3 r;
  r.o = o;
  r.d = d;
  r.inv = (v3){safe_inv_comp(d.x), safe_inv_comp(d.y), safe_inv_comp(d.z)};
  return r;
}

typedef struct {
  int did_hit;
  float dist;
  uint32_t tri, inst;
  float u, v;
  v3 pos; /* world hit position, L2W * local hit point (f32) */
  int front;
} world_hit;

/* transformLinearOrientationSign (intersectionCore.wgsl.ts:96-112): the sign
 * of the scaled 3x3 determinant of L2W; an orientation-reversing instance
 * flips the local front face (intersection.wgsl.ts:45). */
static float orientation_sign(const float *m) {
  const float s0 = fmax2(fabsf_(m[0]), fmax2(fabsf_(m[1]), fabsf_(m[2])));
  const float s1 = fmax2(fabsf_(m[4]), fmax2(fabsf_(m[5]), fabsf_(m[6])));
  const float s2 = fmax2(fabsf_(m[8]), fmax2(fabsf_(m[9]), fabsf_(m[10])));
  if (!(s0 > 0.0f) || s0 > 3.402823e38f || !(s1 > 0.0f) || s1 > 3.402823e38f || !(s2 > 0.0f) ||
      s2 > 3.402823e38f)
    return 1.0f;
  const v3 a = {m[0] / s0, m[1] / s0, m[2] / s0}, b = {m[4] / s1, m[5] / s1, m[6] / s1};
  const v3 c = {m[8] / s2, m[9] / s2, m[10] / s2};
  const v3 x = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  const float det = (x.x * c.x + x.y * c.y) + x.z * c.z;
  return det >= 0.0f ? 1.0f : -1.0f;
}

int orc_orientation_sign(const float *l2w) { return orientation_sign(l2w) > 0.0f ? 1 : -1; }

static void tlas_instance_closest(const orc_scene *s, uint32_t inst, const ray3 *r, float tmin,
                                  world_hit *hit, counters *cnt) {
  if (inst >= s->instance_count) return;
  const float *w2l = s->w2l + 16 * (size_t)inst;
  const float *l2w = s->l2w + 16 * (size_t)inst;
  if (cnt) cnt->instances++;
  v3 lo = xform_point_cols(w2l, r->o);
  v3 ld = xform_dir_cols(w2l, r->d);
  ray3 lr = make_ray(lo, ld);
  v3 start = {r->o.x + r->d.x * tmin, r->o.y + r->d.y * tmin, r->o.z + r->d.z * tmin};
  v3 lstart = xform_point_cols(w2l, start);
  float ltmin = fmax2(dot3(sub(lstart, lo), ld), 0.0f);
  float ltmax = k_infinity();
  if (hit->dist < k_infinity() * 0.5f) {
    v3 end = {r->o.x + r->d.x * hit->dist, r->o.y + r->d.y * hit->dist, r->o.z + r->d.z * hit->dist};
    v3 lend = xform_point_cols(w2l, end);
    ltmax = fmax2(dot3(sub(lend, lo), ld), ltmin);
  }
  local_hit lh;
  uint32_t root = s->blas_roots[inst];
  trace_mesh_bvh(s, &lr, ltmin, ltmax, 1, &lh, root, cnt);
  if (lh.did_hit && lh.dist > ltmin) {
    v3 wpos = xform_point_cols(l2w, lh.position);
    float wdist = dot3(sub(wpos, r->o), r->d);
    if (wdist > tmin && wdist < hit->dist) {
      hit->did_hit = 1;
      hit->dist = wdist;
      hit->tri = lh.tri;
      hit->inst = inst;
      hit->u = lh.u;
      hit->v = lh.v;
      hit->pos = wpos;
      hit->front = orientation_sign(l2w) > 0.0f ? lh.front : !lh.front;
    }
  }
}

static void trace_tlas_closest(const orc_scene *s, const ray3 *r, float tmin, float tmax,
                               world_hit *hit, counters *cnt) {
  hit->did_hit = 0;
  hit->dist = tmax;
  hit->tri = 0;
  hit->inst = 0xFFFFFFFFu;
  hit->u = hit->v = 0;
  hit->front = 0;
  if (s->tlas_node_count == 0 || s->instance_index_count == 0) return; /* fail closed */
  uint32_t stack[TLAS_STACK_DEPTH];
  uint32_t sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    uint32_t ni = stack[--sp];
    if (ni >= s->tlas_node_count) continue;
    const uint32_t *node = s->tlas_nodes + 8 * (size_t)ni;
    if (cnt) cnt->tlas_nodes++;
    if (!slab(r, node, tmin, hit->dist, NULL)) continue;
    uint32_t soc = node[7];
    if ((soc & LEAF_FLAG) == LEAF_FLAG) {
      uint32_t count = soc & 0xFFFFu, start = node[6];
      for (uint32_t i = 0; i < count; ++i) {
        uint32_t pi = start + i;
        if (pi >= s->instance_index_count) continue;
        tlas_instance_closest(s, s->instance_indices[pi], r, tmin, hit, cnt);
      }
    } else {
      uint32_t left = ni + 1, right = ni + node[6];
      if (sp + 2 <= TLAS_STACK_DEPTH) {
        stack[sp++] = right;
        stack[sp++] = left;
      } else {
        /* tlasClosestFallback: every instance in permutation order. */
        hit->did_hit = 0;
        hit->dist = tmax;
        hit->tri = 0;
        hit->inst = 0xFFFFFFFFu;
        for (uint32_t pi = 0; pi < s->instance_index_count; ++pi)
          tlas_instance_closest(s, s->instance_indices[pi], r, tmin, hit, cnt);
        return;
      }
    }
  }
}

static int tlas_instance_any(const orc_scene *s, uint32_t inst, const ray3 *r, float tmin,
                             float tmax) {
  if (inst >= s->instance_count) return 0;
  const float *w2l = s->w2l + 16 * (size_t)inst;
  v3 lo = xform_point_cols(w2l, r->o);
  v3 ld = xform_dir_cols(w2l, r->d);
  ray3 lr = make_ray(lo, ld);
  v3 start = {r->o.x + r->d.x * tmin, r->o.y + r->d.y * tmin, r->o.z + r->d.z * tmin};
  v3 lstart = xform_point_cols(w2l, start);
  float ltmin = fmax2(dot3(sub(lstart, lo), ld), 0.0f);
  float ltmax = tmax;
  if (tmax < k_infinity() * 0.5f) {
    v3 end = {r->o.x + r->d.x * tmax, r->o.y + r->d.y * tmax, r->o.z + r->d.z * tmax};
    v3 lend = xform_point_cols(w2l, end);
    ltmax = fmax2(dot3(sub(lend, lo), ld), ltmin);
  }
  local_hit lh;
  return trace_mesh_bvh(s, &lr, ltmin, ltmax, 0, &lh, s->blas_roots[inst], NULL);
}

static int trace_tlas_any(const orc_scene *s, const ray3 *r, float tmin, float tmax) {
  if (s->tlas_node_count == 0 || s->instance_index_count == 0) return 0;
  uint32_t stack[TLAS_STACK_DEPTH];
  uint32_t sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    uint32_t ni = stack[--sp];
    if (ni >= s->tlas_node_count) continue;
    const uint32_t *node = s->tlas_nodes + 8 * (size_t)ni;
    if (!slab(r, node, tmin, tmax, NULL)) continue;
    uint32_t soc = node[7];
    if ((soc & LEAF_FLAG) == LEAF_FLAG) {
      uint32_t count = soc & 0xFFFFu, start = node[6];
      for (uint32_t i = 0; i < count; ++i) {
        uint32_t pi = start + i;
        if (pi >= s->instance_index_count) continue;
        if (tlas_instance_any(s, s->instance_indices[pi], r, tmin, tmax)) return 1;
      }
    } else {
      uint32_t left = ni + 1, right = ni + node[6];
      if (sp + 2 <= TLAS_STACK_DEPTH) {
        stack[sp++] = right;
        stack[sp++] = left;
      } else {
        for (uint32_t pi = 0; pi < s->instance_index_count; ++pi)
          if (tlas_instance_any(s, s->instance_indices[pi], r, tmin, tmax)) return 1;
        return 0;
      }
    }
  }
  return 0;
}

/* ---- batch drivers ------------------------------------------------------ */

typedef struct {
  const orc_scene *s;
  const orc_ray *rays;
  uint32_t begin, end;
  orc_hit *hits;
  uint32_t *counts;
  uint32_t *any;
  float *pos;
} job_t;

static void *closest_job(void *arg) {
  job_t *j = (job_t *)arg;
  for (uint32_t i = j->begin; i < j->end; ++i) {
    const orc_ray *q = &j->rays[i];
    ray3 r = make_ray((v3){q->o[0], q->o[1], q->o[2]}, (v3){q->d[0], q->d[1], q->d[2]});
    world_hit h;
    counters c = {0, 0, 0, 0};
    trace_tlas_closest(j->s, &r, q->tmin, q->tmax, &h, j->counts ? &c : NULL);
    orc_hit *o = &j->hits[i];
    if (h.did_hit) {
      o->prim = h.tri;
      o->inst = h.inst;
      o->t = h.dist;
      o->u = h.u;
      o->v = h.v;
    } else {
      o->prim = 0xFFFFFFFFu;
      o->inst = 0xFFFFFFFFu;
      o->t = INFINITY;
      o->u = o->v = 0;
    }
    if (j->pos) {
      float *pp = j->pos + 3 * (size_t)i;
      pp[0] = h.did_hit ? h.pos.x : 0.0f;
      pp[1] = h.did_hit ? h.pos.y : 0.0f;
      pp[2] = h.did_hit ? h.pos.z : 0.0f;
    }
    if (j->counts) {
      uint32_t *cc = j->counts + 4 * (size_t)i;
      cc[0] = c.blas_nodes;
      cc[1] = c.tris;
      cc[2] = c.tlas_nodes;
      cc[3] = c.instances;
    }
  }
  return NULL;
}

static void *any_job(void *arg) {
  job_t *j = (job_t *)arg;
  for (uint32_t i = j->begin; i < j->end; ++i) {
    const orc_ray *q = &j->rays[i];
    ray3 r = make_ray((v3){q->o[0], q->o[1], q->o[2]}, (v3){q->d[0], q->d[1], q->d[2]});
    j->any[i] = (uint32_t)trace_tlas_any(j->s, &r, q->tmin, q->tmax);
  }
  return NULL;
}

static int run_jobs(void *(*fn)(void *), job_t proto, uint32_t n, uint32_t threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t tid[256];
  job_t jobs[256];
  uint32_t per = (n + threads - 1) / threads;
  uint32_t started = 0;
  for (uint32_t t = 0; t < threads; ++t) {
    jobs[t] = proto;
    jobs[t].begin = t * per < n ? t * per : n;
    jobs[t].end = (t + 1) * per < n ? (t + 1) * per : n;
    if (threads == 1) {
      fn(&jobs[t]);
    } else if (pthread_create(&tid[t], NULL, fn, &jobs[t]) == 0) {
      ++started;
    } else {
      fn(&jobs[t]);
      tid[t] = 0;
    }
  }
  if (threads > 1)
    for (uint32_t t = 0; t < threads; ++t)
      if (tid[t]) pthread_join(tid[t], NULL);
  (void)started;
  return 0;
}

int orc_trace_closest(const orc_scene *s, const orc_ray *rays, uint32_t n, orc_hit *out,
                      uint32_t *counts, uint32_t threads) {
  job_t proto = {s, rays, 0, 0, out, counts, NULL, NULL};
  return run_jobs(closest_job, proto, n, threads);
}

int orc_trace_closest_pos(const orc_scene *s, const orc_ray *rays, uint32_t n, orc_hit *out, float *pos,
                          uint32_t threads) {
  job_t proto = {s, rays, 0, 0, out, NULL, NULL, pos};
  return run_jobs(closest_job, proto, n, threads);
}

int orc_trace_any(const orc_scene *s, const orc_ray *rays, uint32_t n, uint32_t *out,
                  uint32_t threads) {
  job_t proto = {s, rays, 0, 0, NULL, NULL, out, NULL};
  return run_jobs(any_job, proto, n, threads);
}

/* ---- closest-hit and visibility wrappers (a8) ----------------------------
 * traceClosest (pt-webgpu pathTrace/intersection.wgsl.ts:462-490): a hit on
 * the back face of a single-sided opaque material is not a surface
 * (materialAcceptsSidedHit, material.wgsl.ts:3259-3263); the trace resumes
 * with the hit distance as the exclusive lower bound, and a cursor that does
 * not advance, or reaches tMax, ends it as a miss.  traceSurfaceVisibility
 * (connect.wgsl.ts:37-103), opaque materials: the first accepted surface
 * blocks; skipped surfaces count against the surface-hit limit. */

static void closest_sided(const orc_scene *s, const uint8_t *single_sided, const ray3 *r, float tmin,
                          float tmax, world_hit *h) {
  float cursor = tmin;
  for (;;) {
    trace_tlas_closest(s, r, cursor, tmax, h, NULL);
    if (!h->did_hit) return;
    if (h->front || !single_sided || h->tri >= s->triangle_count || !single_sided[h->tri]) return;
    const float next = h->dist;
    if (!(next > cursor) || !(next < tmax)) {
      h->did_hit = 0;
      return;
    }
    cursor = next;
  }
}

static int visible_sided(const orc_scene *s, const uint8_t *single_sided, uint32_t hit_limit, const ray3 *r,
                         float tmin, float tmax) {
  float cursor = tmin;
  uint32_t count = 0;
  for (;;) {
    world_hit h;
    trace_tlas_closest(s, r, cursor, tmax, &h, NULL);
    if (!h.did_hit) return 1;
    if (count >= hit_limit) return 0;
    ++count;
    const int accepted =
        h.front || !single_sided || h.tri >= s->triangle_count || !single_sided[h.tri];
    if (accepted) return 0;
    if (!(h.dist > cursor)) return 0;
    cursor = h.dist;
  }
}

typedef struct {
  job_t base;
  const uint8_t *single_sided;
  uint32_t hit_limit;
} sided_job_t;

static void *closest_sided_job(void *arg) {
  sided_job_t *sj = (sided_job_t *)arg;
  job_t *j = &sj->base;
  for (uint32_t i = j->begin; i < j->end; ++i) {
    const orc_ray *q = &j->rays[i];
    ray3 r = make_ray((v3){q->o[0], q->o[1], q->o[2]}, (v3){q->d[0], q->d[1], q->d[2]});
    world_hit h;
    closest_sided(j->s, sj->single_sided, &r, q->tmin, q->tmax, &h);
    orc_hit *o = &j->hits[i];
    o->prim = h.did_hit ? h.tri : 0xFFFFFFFFu;
    o->inst = h.did_hit ? h.inst : 0xFFFFFFFFu;
    o->t = h.did_hit ? h.dist : INFINITY;
    o->u = h.did_hit ? h.u : 0.0f;
    o->v = h.did_hit ? h.v : 0.0f;
    if (j->pos) {
      float *pp = j->pos + 3 * (size_t)i;
      pp[0] = h.did_hit ? h.pos.x : 0.0f;
      pp[1] = h.did_hit ? h.pos.y : 0.0f;
      pp[2] = h.did_hit ? h.pos.z : 0.0f;
    }
  }
  return NULL;
}

static void *visible_sided_job(void *arg) {
  sided_job_t *sj = (sided_job_t *)arg;
  job_t *j = &sj->base;
  for (uint32_t i = j->begin; i < j->end; ++i) {
    const orc_ray *q = &j->rays[i];
    ray3 r = make_ray((v3){q->o[0], q->o[1], q->o[2]}, (v3){q->d[0], q->d[1], q->d[2]});
    j->any[i] = visible_sided(j->s, sj->single_sided, sj->hit_limit, &r, q->tmin, q->tmax) ? 0u : 1u;
  }
  return NULL;
}

static int run_sided(void *(*fn)(void *), sided_job_t proto, uint32_t n, uint32_t threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t tid[256];
  sided_job_t jobs[256];
  const uint32_t per = (n + threads - 1) / threads;
  for (uint32_t t = 0; t < threads; ++t) {
    jobs[t] = proto;
    jobs[t].base.begin = t * per < n ? t * per : n;
    jobs[t].base.end = (t + 1) * per < n ? (t + 1) * per : n;
    if (pthread_create(&tid[t], NULL, fn, &jobs[t]) != 0) {
      fn(&jobs[t]);
      tid[t] = 0;
    }
  }
  for (uint32_t t = 0; t < threads; ++t)
    if (tid[t]) pthread_join(tid[t], NULL);
  return 0;
}

int orc_trace_closest_sided(const orc_scene *s, const uint8_t *single_sided, const orc_ray *rays, uint32_t n,
                            orc_hit *out, float *pos, uint32_t threads) {
  sided_job_t proto = {{s, rays, 0, 0, out, NULL, NULL, pos}, single_sided, 0};
  return run_sided(closest_sided_job, proto, n, threads);
}

int orc_trace_occluded_sided(const orc_scene *s, const uint8_t *single_sided, uint32_t hit_limit,
                             const orc_ray *rays, uint32_t n, uint32_t *occluded, uint32_t threads) {
  sided_job_t proto = {{s, rays, 0, 0, NULL, NULL, occluded, NULL}, single_sided, hit_limit};
  return run_sided(visible_sided_job, proto, n, threads);
}

/* ---- FP64 closed-form check --------------------------------------------- */

int orc_fp64_hit_params(const orc_scene *s, const orc_ray *rays, const orc_hit *hits, uint32_t n,
                        double *out) {
  for (uint32_t i = 0; i < n; ++i) {
    double *o = out + 3 * (size_t)i;
    o[0] = o[1] = o[2] = NAN;
    const orc_hit *h = &hits[i];
    if (h->inst == 0xFFFFFFFFu || h->prim >= s->triangle_count || h->inst >= s->instance_count)
      continue;
    const float *m = s->l2w + 16 * (size_t)h->inst;
    const uint32_t *tri = s->indices + 4 * (size_t)h->prim;
    double P[3][3];
    for (int v = 0; v < 3; ++v) {
      const float *p = s->positions + 4 * (size_t)tri[v];
      double x = p[0], y = p[1], z = p[2];
      double w = (double)m[3] * x + (double)m[7] * y + (double)m[11] * z + (double)m[15];
      for (int k = 0; k < 3; ++k)
        P[v][k] = ((double)m[k] * x + (double)m[4 + k] * y + (double)m[8 + k] * z + (double)m[12 + k]) / w;
    }
    const orc_ray *r = &rays[i];
    double O[3] = {r->o[0], r->o[1], r->o[2]}, D[3] = {r->d[0], r->d[1], r->d[2]};
    double e1[3], e2[3], pv[3], tv[3], qv[3];
    for (int k = 0; k < 3; ++k) e1[k] = P[1][k] - P[0][k], e2[k] = P[2][k] - P[0][k];
    pv[0] = D[1] * e2[2] - D[2] * e2[1];
    pv[1] = D[2] * e2[0] - D[0] * e2[2];
    pv[2] = D[0] * e2[1] - D[1] * e2[0];
    double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det == 0.0) continue;
    double inv = 1.0 / det;
    for (int k = 0; k < 3; ++k) tv[k] = O[k] - P[0][k];
    double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
    qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
    qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
    qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
    double v = (D[0] * qv[0] + D[1] * qv[1] + D[2] * qv[2]) * inv;
    double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
    /* The traversal reports dot(P - O, D) for its f32 unit direction D. */
    o[0] = t * (D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    o[1] = u;
    o[2] = v;
  }
  return 0;
}

/* ---- PCG32: packages/shared-samplers/src/wgsl/pcg.wgsl.ts:5-22 --------- */

uint32_t orc_pcg_init(uint32_t px, uint32_t py, uint32_t frame_seed) {
  uint32_t state = px * 1664525u + py * 1013904223u + frame_seed * 22695477u;
  state ^= state >> 17;
  state ^= state << 31;
  state ^= state >> 11;
  return state;
}

uint32_t orc_pcg_next(uint32_t *state) {
  *state = *state * 747796405u + 2891336453u;
  uint32_t word = ((*state >> ((*state >> 28) + 4u)) ^ *state) * 277803737u;
  return (word >> 22) ^ word;
}

float orc_rand_f32(uint32_t *state) { return (float)(orc_pcg_next(state) >> 8) / 16777216.0f; }

/* ---- pt-webgpu's opt-in `cwbvh-closest` traversal (binary32) -------------
 * cwbvhIntersectFirstHitRangeFromRoot (shared-bvh/src/wgsl/
 * cwbvhIntersect.wgsl.ts:197-336) over the 8-wide forest, mollerTrumboreCore
 * (bvhIntersect.wgsl.ts:160-238), and pt-webgpu's traceMeshCwbvhClosest /
 * traceTlasClosestCwbvh (pathTrace/intersection.wgsl.ts:576-838): the wide
 * walk proposes a Moller-Trumbore hit, the watertight test re-verifies it
 * (bit-identical t required), and any disagreement, corrupt layout or stack
 * overflow falls back to the binary traceMeshBvh for that BLAS. */

#define CW_STACK 64
#define CW_INF 3.402823e38f

typedef struct {
  int status; /* 0 complete, 1 overflow, 2 invalid layout */
  int did_hit;
  float dist;
  uint32_t tri;
} cw_hit;

static float cw_decode(uint32_t q, float mn, float mx) {
  if (!(mx > mn)) return mn;
  if (q == 0) return mn;
  if (q >= 65535u) return mx;
  const float t = (float)q / 65535.0f;
  return mn * (1.0f - t) + mx * t;
}

static int cw_bounds_valid(v3 mn, v3 mx) {
  return mn.x <= mx.x && mn.y <= mx.y && mn.z <= mx.z && fabsf_(mn.x) <= CW_INF && fabsf_(mn.y) <= CW_INF &&
         fabsf_(mn.z) <= CW_INF && fabsf_(mx.x) <= CW_INF && fabsf_(mx.y) <= CW_INF && fabsf_(mx.z) <= CW_INF;
}

static int cw_aabb_entry(v3 o, v3 inv, v3 bmin, v3 bmax, float tmax) {
  v3 t0 = {(bmin.x - o.x) * inv.x, (bmin.y - o.y) * inv.y, (bmin.z - o.z) * inv.z};
  v3 t1 = {(bmax.x - o.x) * inv.x, (bmax.y - o.y) * inv.y, (bmax.z - o.z) * inv.z};
  float tn = fmax2(fmax2(fmin2(t0.x, t1.x), fmin2(t0.y, t1.y)), fmin2(t0.z, t1.z));
  float tf = fmin2(fmin2(fmax2(t0.x, t1.x), fmax2(t0.y, t1.y)), fmax2(t0.z, t1.z));
  return !(tn > tf || tf < 0.0f || tn > tmax);
}

static v3 cross3(v3 a, v3 b) {
  return (v3){a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static float dot3l(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

/* mollerTrumboreCore: equilibrated edges and direction, angular determinant
 * rejection at 1e-7, barycentric tolerance 1e-6. */
static int mt_core(v3 o, v3 dir, v3 a, v3 b, v3 c, float tmin, float *t_out) {
  v3 e1r = sub(b, a), e2r = sub(c, a);
  float es = fmax2(fmax2(fmax2(fabsf_(e1r.x), fabsf_(e1r.y)), fabsf_(e1r.z)),
                   fmax2(fmax2(fabsf_(e2r.x), fabsf_(e2r.y)), fabsf_(e2r.z)));
  float ds = fmax2(fabsf_(dir.x), fmax2(fabsf_(dir.y), fabsf_(dir.z)));
  if (!(es > 0.0f) || es > CW_INF || !(ds > 0.0f) || ds > CW_INF) return 0;
  v3 e1 = {e1r.x / es, e1r.y / es, e1r.z / es};
  v3 e2 = {e2r.x / es, e2r.y / es, e2r.z / es};
  v3 dn = {dir.x / ds, dir.y / ds, dir.z / ds};
  v3 ao0 = sub(o, a);
  v3 ao = {ao0.x / es, ao0.y / es, ao0.z / es};
  v3 n = cross3(e1, e2);
  float ns = fmax2(fabsf_(n.x), fmax2(fabsf_(n.y), fabsf_(n.z)));
  if (!(ns > 0.0f) || ns > CW_INF) return 0;
  v3 nd = {n.x / ns, n.y / ns, n.z / ns};
  float nl = sqrtf(dot3l(nd, nd));
  float dl = sqrtf(dot3l(dn, dn));
  v3 rd = {dn.x / dl, dn.y / dl, dn.z / dl};
  float det = -dot3l(dn, n);
  v3 nn = {nd.x / nl, nd.y / nl, nd.z / nl};
  float ang = fabsf_(dot3l(rd, nn));
  if (!(ang > 1e-7f) || det == 0.0f) return 0;
  float inv_det = 1.0f / det;
  v3 dao = cross3(ao, dn);
  float u = dot3l(e2, dao) * inv_det;
  float v = -dot3l(e1, dao) * inv_det;
  float st = dot3l(ao, n) * inv_det;
  float t = (st * es) / ds;
  float w = (1.0f - u) - v;
  if (isnan(u) || isnan(v) || isnan(w) || isnan(t) || fabsf_(u) > CW_INF || fabsf_(v) > CW_INF ||
      fabsf_(w) > CW_INF || fabsf_(t) > CW_INF || u < -1e-6f || v < -1e-6f || w < -1e-6f ||
      t < fmax2(tmin, 0.0f))
    return 0;
  *t_out = t;
  return 1;
}

static cw_hit cw_walk(const orc_scene *s, const orc_cwbvh_view *cw, v3 o, v3 d, float tmin, float tmax,
                      uint32_t root) {
  cw_hit best = {0, 0, tmax, 0};
  const uint32_t nc = cw->node_count;
  if (nc == 0) return best;
  if (root >= nc) {
    best.status = 2;
    return best;
  }
  uint32_t stack[CW_STACK];
  uint32_t sp = 0;
  stack[sp++] = root;
  const v3 inv = {safe_inv_comp(d.x), safe_inv_comp(d.y), safe_inv_comp(d.z)};
  while (sp > 0) {
    const uint32_t node = stack[--sp];
    if (node >= nc) {
      best.status = 2;
      return best;
    }
    const uint32_t count = cw->child_count[node];
    if (count > 8) {
      best.status = 2;
      return best;
    }
    const float *pb = cw->node_bounds + 6 * (size_t)node;
    const v3 pmn = {pb[0], pb[1], pb[2]}, pmx = {pb[3], pb[4], pb[5]};
    if (!cw_bounds_valid(pmn, pmx)) {
      best.status = 2;
      return best;
    }
    for (uint32_t slot = 0; slot < count; ++slot) {
      const uint32_t *m = cw->child_meta + ((size_t)node * 8 + slot) * 3;
      if (m[0] == 0) {
        best.status = 2;
        return best;
      }
      const uint16_t *q = cw->child_bounds + ((size_t)node * 8 + slot) * 6;
      const v3 bmn = {cw_decode(q[0], pmn.x, pmx.x), cw_decode(q[1], pmn.y, pmx.y), cw_decode(q[2], pmn.z, pmx.z)};
      const v3 bmx = {cw_decode(q[3], pmn.x, pmx.x), cw_decode(q[4], pmn.y, pmx.y), cw_decode(q[5], pmn.z, pmx.z)};
      if (!cw_bounds_valid(bmn, bmx)) {
        best.status = 2;
        return best;
      }
      if (!cw_aabb_entry(o, inv, bmn, bmx, best.dist)) continue;
      if (m[0] == 1) {
        if (m[1] >= nc) {
          best.status = 2;
          return best;
        }
        if (sp >= CW_STACK) {
          best.status = 1;
          return best;
        }
        stack[sp++] = m[1];
      } else if (m[0] == 2) {
        if (m[2] == 0 || m[1] > s->triangle_count || m[2] > s->triangle_count - m[1]) {
          best.status = 2;
          return best;
        }
        for (uint32_t i = 0; i < m[2]; ++i) {
          const uint32_t tri = m[1] + i;
          const uint32_t *ix = s->indices + 4 * (size_t)tri;
          if (ix[0] >= s->vertex_count || ix[1] >= s->vertex_count || ix[2] >= s->vertex_count) {
            best.status = 2;
            return best;
          }
          float t;
          if (mt_core(o, d, vload(s, ix[0]), vload(s, ix[1]), vload(s, ix[2]), tmin, &t) && t > tmin &&
              t < best.dist) {
            best.did_hit = 1;
            best.dist = t;
            best.tri = tri;
          }
        }
      } else {
        best.status = 2;
        return best;
      }
    }
  }
  return best;
}

/* traceMeshCwbvhClosest */
static int trace_mesh_cwbvh(const orc_scene *s, const orc_cwbvh_view *cw, const ray3 *r, float tmin, float tmax,
                            local_hit *hit, uint32_t wide_root, uint32_t binary_root, counters *cnt) {
  hit->did_hit = 0;
  hit->dist = tmax;
  hit->position = (v3){0, 0, 0};
  hit->tri = 0;
  hit->u = hit->v = 0;
  hit->front = 0;
  if (cw->node_count == 0 || wide_root >= cw->node_count)
    return trace_mesh_bvh(s, r, tmin, tmax, 1, hit, binary_root, cnt);
  const cw_hit c = cw_walk(s, cw, r->o, r->d, tmin, tmax, wide_root);
  if (c.status != 0) return trace_mesh_bvh(s, r, tmin, tmax, 1, hit, binary_root, cnt);
  if (!c.did_hit || c.dist <= tmin || c.dist >= tmax) return 0;
  if (c.tri >= s->triangle_count) return trace_mesh_bvh(s, r, tmin, tmax, 1, hit, binary_root, cnt);
  const uint32_t *ix = s->indices + 4 * (size_t)c.tri;
  const wt_hit e = watertight(r->o, r->d, vload(s, ix[0]), vload(s, ix[1]), vload(s, ix[2]), tmin);
  uint32_t eb, cb;
  memcpy(&eb, &e.t, 4);
  memcpy(&cb, &c.dist, 4);
  if (!e.hit || eb != cb) return trace_mesh_bvh(s, r, tmin, tmax, 1, hit, binary_root, cnt);
  hit->did_hit = 1;
  hit->dist = c.dist;
  hit->position = e.point;
  hit->tri = c.tri;
  hit->u = e.bary.y;
  hit->v = e.bary.z;
  hit->front = e.side > 0.0f;
  return 1;
}

/* traceTlasClosestCwbvh: the reference TLAS order, cwbvh BLAS walks. */
static void trace_tlas_closest_cwbvh(const orc_scene *s, const orc_cwbvh_view *cw, const ray3 *r, float tmin,
                                     float tmax, world_hit *hit) {
  hit->did_hit = 0;
  hit->dist = tmax;
  hit->tri = 0;
  hit->inst = 0xFFFFFFFFu;
  hit->u = hit->v = 0;
  hit->front = 0;
  if (s->tlas_node_count == 0 || s->instance_index_count == 0) {
    trace_tlas_closest(s, r, tmin, tmax, hit, NULL);
    return;
  }
  uint32_t stack[TLAS_STACK_DEPTH];
  uint32_t sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const uint32_t ni = stack[--sp];
    if (ni >= s->tlas_node_count) continue;
    const uint32_t *node = s->tlas_nodes + 8 * (size_t)ni;
    if (!slab(r, node, tmin, hit->dist, NULL)) continue;
    const uint32_t soc = node[7];
    if ((soc & LEAF_FLAG) == LEAF_FLAG) {
      const uint32_t count = soc & 0xFFFFu, start = node[6];
      for (uint32_t i = 0; i < count; ++i) {
        const uint32_t pi = start + i;
        if (pi >= s->instance_index_count) continue;
        const uint32_t inst = s->instance_indices[pi];
        if (inst >= s->instance_count) continue;
        const float *w2l = s->w2l + 16 * (size_t)inst;
        const float *l2w = s->l2w + 16 * (size_t)inst;
        const v3 lo = xform_point_cols(w2l, r->o);
        const v3 ld = xform_dir_cols(w2l, r->d);
        const ray3 lr = make_ray(lo, ld);
        const v3 start_p = {r->o.x + r->d.x * tmin, r->o.y + r->d.y * tmin, r->o.z + r->d.z * tmin};
        const float ltmin = fmax2(dot3(sub(xform_point_cols(w2l, start_p), lo), ld), 0.0f);
        float ltmax = k_infinity();
        if (hit->dist < k_infinity() * 0.5f) {
          const v3 end = {r->o.x + r->d.x * hit->dist, r->o.y + r->d.y * hit->dist, r->o.z + r->d.z * hit->dist};
          ltmax = fmax2(dot3(sub(xform_point_cols(w2l, end), lo), ld), ltmin);
        }
        local_hit lh;
        trace_mesh_cwbvh(s, cw, &lr, ltmin, ltmax, &lh, cw->inst_root[inst], s->blas_roots[inst], NULL);
        if (lh.did_hit && lh.dist > ltmin) {
          const v3 wpos = xform_point_cols(l2w, lh.position);
          const float wdist = dot3(sub(wpos, r->o), r->d);
          if (wdist <= tmin || wdist >= hit->dist) continue;
          hit->did_hit = 1;
          hit->dist = wdist;
          hit->tri = lh.tri;
          hit->inst = inst;
          hit->u = lh.u;
          hit->v = lh.v;
          hit->pos = wpos;
          hit->front = orientation_sign(l2w) > 0.0f ? lh.front : !lh.front;
        }
      }
    } else {
      const uint32_t left = ni + 1, right = ni + node[6];
      if (sp + 2 <= TLAS_STACK_DEPTH) {
        stack[sp++] = right;
        stack[sp++] = left;
      } else {
        trace_tlas_closest(s, r, tmin, tmax, hit, NULL);
        return;
      }
    }
  }
}

typedef struct {
  const orc_scene *s;
  const orc_cwbvh_view *cw;
  const orc_ray *rays;
  orc_hit *out;
  uint32_t begin, end;
  const uint8_t *single_sided;
} cw_job;

static void *cw_job_fn(void *arg) {
  cw_job *j = arg;
  for (uint32_t i = j->begin; i < j->end; ++i) {
    const orc_ray *q = &j->rays[i];
    const ray3 r = make_ray((v3){q->o[0], q->o[1], q->o[2]}, (v3){q->d[0], q->d[1], q->d[2]});
    world_hit h;
    /* traceClosest's sided loop over traceClosestRaw in this mode */
    float cursor = q->tmin;
    for (;;) {
      trace_tlas_closest_cwbvh(j->s, j->cw, &r, cursor, q->tmax, &h);
      if (!j->single_sided || !h.did_hit || h.front || h.tri >= j->s->triangle_count ||
          !j->single_sided[h.tri])
        break;
      if (!(h.dist > cursor) || !(h.dist < q->tmax)) {
        h.did_hit = 0;
        break;
      }
      cursor = h.dist;
    }
    orc_hit *o = &j->out[i];
    o->prim = h.did_hit ? h.tri : 0xFFFFFFFFu;
    o->inst = h.did_hit ? h.inst : 0xFFFFFFFFu;
    o->t = h.did_hit ? h.dist : INFINITY;
    o->u = h.did_hit ? h.u : 0.0f;
    o->v = h.did_hit ? h.v : 0.0f;
  }
  return NULL;
}

int orc_trace_closest_cwbvh_sided(const orc_scene *s, const orc_cwbvh_view *cw, const uint8_t *single_sided,
                                  const orc_ray *rays, uint32_t n, orc_hit *out, uint32_t threads);

int orc_trace_closest_cwbvh(const orc_scene *s, const orc_cwbvh_view *cw, const orc_ray *rays, uint32_t n,
                            orc_hit *out, uint32_t threads) {
  return orc_trace_closest_cwbvh_sided(s, cw, NULL, rays, n, out, threads);
}

int orc_trace_closest_cwbvh_sided(const orc_scene *s, const orc_cwbvh_view *cw, const uint8_t *single_sided,
                                  const orc_ray *rays, uint32_t n, orc_hit *out, uint32_t threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t tid[256];
  cw_job jobs[256];
  const uint32_t per = (n + threads - 1) / threads;
  for (uint32_t t = 0; t < threads; ++t) {
    jobs[t] = (cw_job){s, cw, rays, out, t * per < n ? t * per : n, (t + 1) * per < n ? (t + 1) * per : n,
                       single_sided};
    if (pthread_create(&tid[t], NULL, cw_job_fn, &jobs[t]) != 0) {
      cw_job_fn(&jobs[t]);
      tid[t] = 0;
    }
  }
  for (uint32_t t = 0; t < threads; ++t)
    if (tid[t]) pthread_join(tid[t], NULL);
  return 0;
}
