/*
 * light_tree.c -- the spatial light tree that pt-webgpu selects NEE lights
 * with by default (lightTreeImportanceSampling, full trace tier, two or more
 * lights).  TEST INFRASTRUCTURE ONLY (see vt_oracle.h).
 *
 *   build    packages/shared-samplers/src/lightTree.ts:350-448 (buildSubtree:
 *            median split on the longest centroid axis, power-sorted median
 *            for co-located centroids), :250-307 (mergeCones), :492-593
 *   pack     :611-737 (packLightTreeForGPU: 16 f32 per node, represented
 *            power = total / max node power, outward-rounded cone cosines
 *            :183-212, subtree leaf counts)
 *   sample   :938-1100 (packed f32 importance), :1177-1204 (bucket split),
 *            :1237-1319 (samplePackedLightTreeCPU), :1321-1414 (pdf)
 *
 * Builder arithmetic is binary64 (JavaScript numbers); the packed importance
 * and the bucket split are binary32 with one rounding per operation, as the
 * reference's Math.fround mirror.  log2 and exp2 are evaluated in binary64
 * and rounded once (the mirror's f32(Math.log2(f32(x)))).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "vt_oracle.h"

int orc_fail(const char *fmt, ...);

#define LT_FLOATS 16
#define LT_BUCKETS 0x01000000u
static const float F32_MIN_NORMAL = 1.175494351e-38f;
static const float F32_MAXV = 3.402823466e38f;

typedef struct {
  double axis[3], theta_o, theta_e;
} cone_t;

typedef struct {
  int emitter;
  double power, mn[3], mx[3];
  int left, right;
  cone_t cone;
} lt_node;

typedef struct {
  int emitter;
  double power, c[3], mn[3], mx[3];
  cone_t cone;
} lt_item;

static double hypot3_js(double a, double b, double c) {
  /* Math.hypot (V8): max-scaled compensated sum */
  const double v[3] = {fabs(a), fabs(b), fabs(c)};
  double mx = 0;
  for (int i = 0; i < 3; ++i) {
    if (isinf(v[i])) return INFINITY;
    if (v[i] > mx) mx = v[i];
  }
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
  for (int i = 0; i < 3; ++i) {
    const double n = v[i] / mx, s = n * n - comp, pre = sum + s;
    comp = (pre - sum) - s;
    sum = pre;
  }
  return sqrt(sum) * mx;
}

/* vlen (lightTree.ts:237-242) */
static double vlen(const double *v) {
  const double s = fmax(fmax(fabs(v[0]), fabs(v[1])), fabs(v[2]));
  if (s == 0) return 0;
  return s * hypot3_js(v[0] / s, v[1] / s, v[2] / s);
}

static const cone_t FULL_SPHERE = {{0, 0, 0}, M_PI, M_PI};

/* mergeCones (lightTree.ts:250-307) */
static cone_t merge_cones(cone_t a, cone_t b) {
  const double la = vlen(a.axis), lb = vlen(b.axis);
  if (la == 0 || lb == 0 || a.theta_o >= M_PI || b.theta_o >= M_PI) {
    cone_t r = FULL_SPHERE;
    r.theta_e = fmin(M_PI, fmax(a.theta_e, b.theta_e));
    return r;
  }
  cone_t A = a, B = b;
  if (B.theta_o > A.theta_o) {
    A = b;
    B = a;
  }
  const double lA = vlen(A.axis), lB = vlen(B.axis);
  const double ax[3] = {A.axis[0] / lA, A.axis[1] / lA, A.axis[2] / lA};
  const double bx[3] = {B.axis[0] / lB, B.axis[1] / lB, B.axis[2] / lB};
  const double cos_d = fmax(-1, fmin(1, ax[0] * bx[0] + ax[1] * bx[1] + ax[2] * bx[2]));
  const double d_theta = acos(cos_d);
  const double theta_e = fmin(M_PI, fmax(A.theta_e, B.theta_e));
  if (fmin(d_theta + B.theta_o, M_PI) <= A.theta_o) {
    cone_t r = {{ax[0], ax[1], ax[2]}, A.theta_o, theta_e};
    return r;
  }
  const double theta_o = (A.theta_o + d_theta + B.theta_o) * 0.5;
  if (theta_o >= M_PI) {
    cone_t r = FULL_SPHERE;
    r.theta_e = theta_e;
    return r;
  }
  const double rot = theta_o - A.theta_o;
  double t[3] = {bx[0] - cos_d * ax[0], bx[1] - cos_d * ax[1], bx[2] - cos_d * ax[2]};
  const double tl = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  if (tl < 1e-8) {
    cone_t r = FULL_SPHERE;
    r.theta_e = theta_e;
    return r;
  }
  for (int k = 0; k < 3; ++k) t[k] /= tl;
  const double cr = cos(rot), sr = sin(rot);
  cone_t r = {{ax[0] * cr + t[0] * sr, ax[1] * cr + t[1] * sr, ax[2] * cr + t[2] * sr}, theta_o, theta_e};
  return r;
}

static int cmp_axis_k;
static int cmp_axis(const void *pa, const void *pb) {
  const lt_item *a = pa, *b = pb;
  const double d = a->c[cmp_axis_k] - b->c[cmp_axis_k];
  if (d < 0) return -1;
  if (d > 0) return 1;
  return (a->emitter > b->emitter) - (a->emitter < b->emitter);
}
static int cmp_power(const void *pa, const void *pb) {
  const lt_item *a = pa, *b = pb;
  const double d = b->power - a->power;
  if (d < 0) return -1;
  if (d > 0) return 1;
  return (a->emitter > b->emitter) - (a->emitter < b->emitter);
}

/* buildSubtree (lightTree.ts:350-448): preorder, left subtree first. */
static int build_subtree(lt_item *items, uint32_t n, lt_node *nodes, uint32_t *count) {
  const int idx = (int)(*count)++;
  if (n == 1) {
    lt_node *L = &nodes[idx];
    L->emitter = items[0].emitter;
    L->power = items[0].power;
    memcpy(L->mn, items[0].mn, sizeof L->mn);
    memcpy(L->mx, items[0].mx, sizeof L->mx);
    L->left = L->right = -1;
    L->cone = items[0].cone;
    return idx;
  }
  double cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      cmn[k] = fmin(cmn[k], items[i].c[k]);
      cmx[k] = fmax(cmx[k], items[i].c[k]);
    }
  const double sx = cmx[0] - cmn[0], sy = cmx[1] - cmn[1], sz = cmx[2] - cmn[2];
  /* union AABB and total power of the item set in the order received
   * (sumPower :331-338 adds in that order) */
  lt_node self;
  self.emitter = -1;
  self.power = 0;
  for (int k = 0; k < 3; ++k) {
    self.mn[k] = INFINITY;
    self.mx[k] = -INFINITY;
  }
  for (uint32_t i = 0; i < n; ++i) {
    self.power += items[i].power;
    for (int k = 0; k < 3; ++k) {
      self.mn[k] = fmin(self.mn[k], items[i].mn[k]);
      self.mx[k] = fmax(self.mx[k], items[i].mx[k]);
    }
  }
  nodes[idx] = self;
  if (fmax(fmax(sx, sy), sz) < 1e-4) {
    qsort(items, n, sizeof *items, cmp_power);
  } else {
    cmp_axis_k = (sx >= sy && sx >= sz) ? 0 : (sy >= sz ? 1 : 2);
    qsort(items, n, sizeof *items, cmp_axis);
  }
  const uint32_t mid = n / 2;
  const int l = build_subtree(items, mid, nodes, count);
  const int r = build_subtree(items + mid, n - mid, nodes, count);
  nodes[idx].left = l;
  nodes[idx].right = r;
  nodes[idx].cone = merge_cones(nodes[l].cone, nodes[r].cone);
  return idx;
}

static float next_down_f32(float v) {
  if (isnan(v) || v == -INFINITY) return v;
  if (v == 0) return -0x1p-149f;
  uint32_t b;
  memcpy(&b, &v, 4);
  b = v > 0 ? b - 1 : b + 1;
  memcpy(&v, &b, 4);
  return v;
}

/* conservativeConeCosineThreshold (lightTree.ts:193-212) */
static float cone_cos_threshold(double angle) {
  const float rounded = (float)cos(fmin(M_PI, angle));
  float t = rounded;
  for (int i = 0; i < 64; ++i) t = next_down_f32(t);
  const float abs_t = (float)((double)rounded - 0x1p-18);
  const float m = t < abs_t ? t : abs_t;
  return m > -1.0f ? m : -1.0f;
}

/* representedLightTreePower (lightTree.ts:214-224) */
static float represented_power(double total, double max_power) {
  if (!(total > 0) || !(max_power > 0)) return 0;
  const double ratio = total / max_power;
  if (!(ratio > 0)) return F32_MIN_NORMAL;
  const float r = (float)ratio;
  return r > F32_MIN_NORMAL ? r : F32_MIN_NORMAL;
}

/* packLightTreeForGPU (lightTree.ts:611-737) over explicit nodes. */
static int pack_nodes(const lt_node *nodes, uint32_t count, float *out) {
  double max_power = 0;
  for (uint32_t i = 0; i < count; ++i) max_power = fmax(max_power, nodes[i].power);
  for (uint32_t i = 0; i < count; ++i) {
    const lt_node *n = &nodes[i];
    float *o = out + (size_t)i * LT_FLOATS;
    const double al = vlen(n->cone.axis);
    if (al == 0 && n->cone.theta_o < M_PI) return orc_fail("light tree: zero cone axis needs a full sphere");
    o[0] = (float)n->emitter;
    o[1] = represented_power(n->power, max_power);
    o[2] = (float)n->left;
    o[3] = (float)n->right;
    for (int k = 0; k < 3; ++k) {
      o[4 + k] = (float)n->mn[k];
      o[7 + k] = (float)n->mx[k];
      o[10 + k] = al > 0 ? (float)(n->cone.axis[k] / al) : 0.0f;
    }
    o[13] = cone_cos_threshold(n->cone.theta_o);
    o[14] = cone_cos_threshold(n->cone.theta_o + n->cone.theta_e);
    o[15] = 0;
  }
  for (int i = (int)count - 1; i >= 0; --i) {
    const lt_node *n = &nodes[i];
    uint32_t leaves = 1;
    if (n->left >= 0 && n->right >= 0)
      leaves = (uint32_t)out[(size_t)n->left * LT_FLOATS + 15] + (uint32_t)out[(size_t)n->right * LT_FLOATS + 15];
    if (leaves > LT_BUCKETS) return orc_fail("light tree: more leaves than the 24-bit domain");
    out[(size_t)i * LT_FLOATS + 15] = (float)leaves;
  }
  return 0;
}

int orc_light_tree_pack_nodes(uint32_t count, const int32_t *emitter, const double *power, const int32_t *left,
                              const int32_t *right, const double *aabb_min, const double *aabb_max,
                              const double *axis, const double *theta_o, const double *theta_e, float *out) {
  lt_node *nodes = calloc(count ? count : 1, sizeof *nodes);
  if (!nodes) return orc_fail("light tree: out of memory");
  for (uint32_t i = 0; i < count; ++i) {
    nodes[i].emitter = emitter[i];
    nodes[i].power = power[i];
    nodes[i].left = left[i];
    nodes[i].right = right[i];
    for (int k = 0; k < 3; ++k) {
      nodes[i].mn[k] = aabb_min[3 * i + k];
      nodes[i].mx[k] = aabb_max[3 * i + k];
      nodes[i].cone.axis[k] = axis[3 * i + k];
    }
    nodes[i].cone.theta_o = theta_o[i];
    nodes[i].cone.theta_e = theta_e[i];
  }
  const int rc = pack_nodes(nodes, count, out);
  free(nodes);
  return rc;
}

/* buildLightTree (lightTree.ts:492-593) + packLightTreeForGPU.  has_cone[i]
 * = 0 gives the full-sphere cone; otherwise axis (normalised here), theta_o,
 * theta_e.  out capacity: (2n - 1) * 16 floats. */
int orc_light_tree_build(uint32_t n, const double *power, const double *centroid, const double *aabb_min,
                         const double *aabb_max, const uint8_t *has_cone, const double *axis,
                         const double *theta_o, const double *theta_e, float *out, uint32_t *out_nodes) {
  if (n == 0) return orc_fail("light tree: at least one emitter required");
  lt_item *items = calloc(n, sizeof *items);
  lt_node *nodes = calloc(2 * (size_t)n, sizeof *nodes);
  if (!items || !nodes) {
    free(items);
    free(nodes);
    return orc_fail("light tree: out of memory");
  }
  for (uint32_t i = 0; i < n; ++i) {
    lt_item *it = &items[i];
    it->emitter = (int)i;
    it->power = power[i];
    for (int k = 0; k < 3; ++k) {
      it->c[k] = centroid[3 * i + k];
      it->mn[k] = aabb_min[3 * i + k];
      it->mx[k] = aabb_max[3 * i + k];
    }
    if (!(power[i] >= 0)) {
      free(items);
      free(nodes);
      return orc_fail("light tree: negative or NaN power");
    }
    const double l = has_cone && has_cone[i] ? vlen(axis + 3 * i) : 0;
    if (l == 0) {
      it->cone = FULL_SPHERE;
    } else {
      for (int k = 0; k < 3; ++k) it->cone.axis[k] = axis[3 * i + k] / l;
      it->cone.theta_o = theta_o[i];
      it->cone.theta_e = fmin(M_PI, theta_e[i]);
    }
  }
  uint32_t count = 0;
  build_subtree(items, n, nodes, &count);
  const int rc = pack_nodes(nodes, count, out);
  *out_nodes = count;
  free(items);
  free(nodes);
  return rc;
}

/* ---- packed importance, binary32 (lightTree.ts:938-1151) ------------------ */

static float f32_log2(float x) { return (float)log2((double)x); }
static float f32_exp2(float x) { return (float)exp2((double)x); }
static float fmaxf_(float a, float b) { return a > b ? a : b; }
static float fminf_(float a, float b) { return a < b ? a : b; }
static float dot3f(float ax, float ay, float az, float bx, float by, float bz) {
  return (ax * bx + ay * by) + az * bz;
}

static float dist2_to_aabb(const float *p, const float *mn, const float *mx) {
  float h[3];
  for (int k = 0; k < 3; ++k) h[k] = fmaxf_(fmaxf_(0.5f * mn[k] - 0.5f * p[k], 0.0f), 0.5f * p[k] - 0.5f * mx[k]);
  const float hs = fmaxf_(fmaxf_(fabsf(h[0]), fabsf(h[1])), fabsf(h[2]));
  if (!(hs >= F32_MIN_NORMAL)) return 0;
  if (hs > 0.5f * F32_MAXV) return F32_MAXV;
  const float sx = h[0] / hs, sy = h[1] / hs, sz = h[2] / hs;
  const float sq = dot3f(sx, sy, sz, sx, sy, sz);
  const float scale = 2.0f * hs;
  const float limit = sqrtf(F32_MAXV / sq);
  if (scale >= limit) return F32_MAXV;
  return fminf_(F32_MAXV, (scale * scale) * sq);
}

static float cone_factor(const float *axis, float cos_o, float cos_oe, const float *p, const float *mn,
                         const float *mx) {
  const float as = fmaxf_(fmaxf_(fabsf(axis[0]), fabsf(axis[1])), fabsf(axis[2]));
  if (!(as > 0) || as > F32_MAXV) return 1;
  const float sa[3] = {axis[0] / as, axis[1] / as, axis[2] / as};
  const float al = sqrtf(dot3f(sa[0], sa[1], sa[2], sa[0], sa[1], sa[2]));
  float dv[3], rv[3];
  for (int k = 0; k < 3; ++k) {
    dv[k] = (0.5f * p[k] - 0.25f * mn[k]) - 0.25f * mx[k];
    rv[k] = 0.25f * mx[k] - 0.25f * mn[k];
  }
  const float ds = fmaxf_(fmaxf_(fabsf(dv[0]), fabsf(dv[1])), fabsf(dv[2]));
  if (!(ds >= F32_MIN_NORMAL) || ds > F32_MAXV) return 1;
  const float sd[3] = {dv[0] / ds, dv[1] / ds, dv[2] / ds};
  const float sdl = sqrtf(dot3f(sd[0], sd[1], sd[2], sd[0], sd[1], sd[2]));
  const float rs = fmaxf_(fmaxf_(fabsf(rv[0]), fabsf(rv[1])), fabsf(rv[2]));
  float r_over_d = 0;
  if (rs >= F32_MIN_NORMAL) {
    const float sr[3] = {rv[0] / rs, rv[1] / rs, rv[2] / rs};
    const float srl = sqrtf(dot3f(sr[0], sr[1], sr[2], sr[0], sr[1], sr[2]));
    const float lr = ((f32_log2(rs) + f32_log2(srl)) - f32_log2(ds)) - f32_log2(sdl);
    if (lr >= 0) return 1;
    r_over_d = f32_exp2(fmaxf_(lr, -126.0f));
  }
  const float d[3] = {sd[0] / sdl, sd[1] / sdl, sd[2] / sdl};
  const float a[3] = {sa[0] / al, sa[1] / al, sa[2] / al};
  const float cos_t = fmaxf_(-1.0f, fminf_(1.0f, dot3f(a[0], a[1], a[2], d[0], d[1], d[2])));
  const float sin_u = fmaxf_(0.0f, fminf_(1.0f, r_over_d));
  const float cos_u = sqrtf(fmaxf_(0.0f, 1.0f - sin_u * sin_u));
  float cos_adj = 1;
  if (cos_t < cos_u) {
    const float sin_t = sqrtf(fmaxf_(0.0f, 1.0f - cos_t * cos_t));
    cos_adj = fmaxf_(-1.0f, fminf_(1.0f, cos_t * cos_u + sin_t * sin_u));
  }
  if (cos_adj < cos_oe) return 0;
  if (cos_adj >= cos_o) return 1;
  const float span = cos_o - cos_oe;
  if (!(span > 0)) return 1;
  return fmaxf_(F32_MIN_NORMAL, fminf_(1.0f, (cos_adj - cos_oe) / span));
}

float orc_light_tree_importance(const float *packed, uint32_t node, const float *p, float floor_d2) {
  const float *b = packed + (size_t)node * LT_FLOATS;
  const float power = b[1];
  if (!(power > 0)) return 0;
  const float d2 = fmaxf_(fmaxf_(dist2_to_aabb(p, b + 4, b + 7), floor_d2), F32_MIN_NORMAL);
  const float c = cone_factor(b + 10, b[13], b[14], p, b + 4, b + 7);
  if (!(c > 0)) return 0;
  const float li = (f32_log2(power) - f32_log2(d2)) + f32_log2(fmaxf_(c, F32_MIN_NORMAL));
  if (li >= 127.99999f) return F32_MAXV;
  if (li <= -126.0f) return F32_MIN_NORMAL;
  return fmaxf_(F32_MIN_NORMAL, fminf_(F32_MAXV, f32_exp2(li)));
}

/* representedPairFirst + allocateLeftBuckets (lightTree.ts:1177-1204) */
static uint32_t left_buckets(uint32_t cur, uint32_t ll, uint32_t rl, float li, float ri) {
  const uint32_t remaining = cur - (ll + rl);
  float p_left;
  if (li == 0 && ri == 0) {
    p_left = 0.5f;
  } else {
    const float s = fmaxf_(li, ri);
    const float l = li / s, r = ri / s;
    p_left = l / (l + r);
  }
  const float ideal = (float)remaining * p_left;
  const float rounded = floorf(ideal + 0.5f);
  const float extra = fmaxf_(0.0f, rounded);
  uint32_t e = (uint32_t)extra;
  if (e > remaining) e = remaining;
  return ll + e;
}

/* samplePackedLightTreeCPU (:1237-1319).  `bucket` is the 24-bit root draw
 * (used only when the root is interior).  Returns the emitter or -1. */
int32_t orc_light_tree_sample(const float *packed, uint32_t nodes, const float *p, float floor_d2, uint32_t bucket,
                              float *pdf) {
  *pdf = 0;
  if (nodes == 0) return -1;
  uint32_t node = 0, cur = LT_BUCKETS, local = 0;
  if ((int)packed[2] >= 0 && (int)packed[3] >= 0) local = bucket < LT_BUCKETS ? bucket : LT_BUCKETS - 1;
  for (uint32_t guard = 0; guard < nodes + 1; ++guard) {
    const float *b = packed + (size_t)node * LT_FLOATS;
    const int l = (int)b[2], r = (int)b[3];
    if (l < 0 || r < 0) {
      *pdf = (float)cur / (float)LT_BUCKETS;
      return (int32_t)b[0];
    }
    const uint32_t lb = left_buckets(cur, (uint32_t)packed[(size_t)l * LT_FLOATS + 15],
                                     (uint32_t)packed[(size_t)r * LT_FLOATS + 15],
                                     orc_light_tree_importance(packed, (uint32_t)l, p, floor_d2),
                                     orc_light_tree_importance(packed, (uint32_t)r, p, floor_d2));
    if (local < lb) {
      cur = lb;
      node = (uint32_t)l;
    } else {
      local -= lb;
      cur -= lb;
      node = (uint32_t)r;
    }
  }
  return -1;
}

/* packedLightTreePdfCPU (:1321-1414) */
float orc_light_tree_pdf(const float *packed, uint32_t nodes, const float *p, float floor_d2, int32_t emitter) {
  if (nodes == 0 || emitter < 0) return 0;
  uint8_t *contains = calloc(nodes, 1);
  if (!contains) return 0;
  for (int i = (int)nodes - 1; i >= 0; --i) {
    const float *b = packed + (size_t)i * LT_FLOATS;
    const int l = (int)b[2], r = (int)b[3];
    contains[i] = l < 0 ? (b[0] == (float)emitter) : (contains[l] || contains[r]);
  }
  float out = 0;
  if (contains[0]) {
    uint32_t node = 0, cur = LT_BUCKETS;
    for (uint32_t guard = 0; guard < nodes + 1; ++guard) {
      const float *b = packed + (size_t)node * LT_FLOATS;
      const int l = (int)b[2], r = (int)b[3];
      if (l < 0 || r < 0) {
        out = b[0] == (float)emitter ? (float)cur / (float)LT_BUCKETS : 0.0f;
        break;
      }
      const uint32_t lb = left_buckets(cur, (uint32_t)packed[(size_t)l * LT_FLOATS + 15],
                                       (uint32_t)packed[(size_t)r * LT_FLOATS + 15],
                                       orc_light_tree_importance(packed, (uint32_t)l, p, floor_d2),
                                       orc_light_tree_importance(packed, (uint32_t)r, p, floor_d2));
      if (contains[l]) {
        cur = lb;
        node = (uint32_t)l;
      } else {
        cur -= lb;
        node = (uint32_t)r;
      }
    }
  }
  free(contains);
  return out;
}
