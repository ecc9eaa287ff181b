/*
 * cwbvh.c -- the reference's CPU closest-hit traversal: the 8-wide
 * compressed BVH collapse and its binary64 first-hit walker.  TEST
 * INFRASTRUCTURE ONLY (see vt_oracle.h).
 *
 *   collapse      packages/shared-bvh/src/compressedWideBvh.ts:215-418
 *                 (collectWideChildren, buildCompressedWideBvhFromArrayBvh,
 *                 quantizeBound / stableUnitOffset :180-204,
 *                 requiredCwbvhTraversalStackEntries :60-92)
 *   child decode  cwbvhChildBounds :481-510, dequantizeBound :206-213
 *   first hit     intersectCompressedWideBvhFirstHit :754-860 with
 *                 intersectTriangle :551-614 and intersectAabb :616-645
 *   TLAS          tlasIntersect (packages/shared-bvh/src/tlas.ts:1223-1310)
 *
 * All arithmetic is IEEE binary64, as the reference's JavaScript numbers;
 * Math.hypot follows V8's scaled, compensated sum (the only JS builtin here
 * whose result is implementation-defined; it only feeds the 1e-7 grazing
 * threshold).
 *
 * Scenes with instances: the reference has no single CPU routine for a TLAS
 * closest hit, so the composition used here is the one its own tests and the
 * renderer imply -- tlasIntersect yields candidate instances (stack order,
 * right child popped first), each candidate's BLAS is walked as a wide root
 * of the concatenated forest in that instance's local frame (local ray =
 * world-to-local applied to the origin as a point and to the direction as a
 * vector, unnormalised, so the local ray parameter is the world one), and the
 * smallest t wins with a strict `<` (first candidate wins ties).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "vt_oracle.h"

int orc_fail(const char *fmt, ...);

enum { KIDS = 8, BOUNDS_U16 = 6, META = 3, STACK_CAP = 64 };
enum { K_EMPTY = 0, K_NODE = 1, K_LEAF = 2 };
#define COUNT_INVALID 0xFFFFFFFFu

struct orc_cwbvh {
  uint32_t node_count, cap;
  float *node_bounds;     /* 6 per node */
  uint16_t *child_bounds; /* 48 per node */
  uint32_t *child_meta;   /* 24 per node */
  uint32_t *child_count;  /* 1 per node */
  uint32_t root_count;
  uint32_t *roots;        /* wide root per binary root */
  uint32_t *status;       /* per root: 0 wide, 1 empty, 2 binary-fallback */
  double *stack_need;     /* per root */
};

static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

typedef struct {
  double mn[3], mx[3];
  uint32_t w6, w7;
} bnode;

static bnode read_node(const uint32_t *w, uint32_t i) {
  bnode n;
  for (int k = 0; k < 3; ++k) {
    n.mn[k] = u2f(w[8 * i + k]);
    n.mx[k] = u2f(w[8 * i + 3 + k]);
  }
  n.w6 = w[8 * i + 6];
  n.w7 = w[8 * i + 7];
  return n;
}

static int is_leaf(uint32_t w7) { return (w7 >> 16) == 0xFFFFu; }

static double js_max(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == 0 && b == 0) return signbit(a) ? b : a;
  return a > b ? a : b;
}
static double js_min(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == 0 && b == 0) return signbit(a) ? a : b;
  return a < b ? a : b;
}

/* V8 Math.hypot: max-scaled, Kahan-compensated sum of squares. */
static double js_hypot3(double a, double b, double c) {
  const double v[3] = {fabs(a), fabs(b), fabs(c)};
  double mx = 0;
  int nan = 0;
  for (int i = 0; i < 3; ++i) {
    if (isinf(v[i])) return INFINITY;
    if (isnan(v[i])) nan = 1;
    else if (v[i] > mx) mx = v[i];
  }
  if (nan) return NAN;
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
  for (int i = 0; i < 3; ++i) {
    const double n = v[i] / mx;
    const double summand = n * n - comp;
    const double pre = sum + summand;
    comp = (pre - sum) - summand;
    sum = pre;
  }
  return sqrt(sum) * mx;
}

/* stableUnitOffset (compressedWideBvh.ts:185-195); returns 0 for null. */
static int unit_offset(double value, double mn, double mx, double *out) {
  if (!isfinite(value) || !isfinite(mn) || !isfinite(mx) || mn > mx) return 0;
  if (mx == mn) {
    *out = 0;
    return 1;
  }
  const double extent = mx - mn;
  if (isfinite(extent)) {
    *out = (value - mn) / extent;
    return 1;
  }
  const double half = mx * 0.5 - mn * 0.5;
  if (!(half > 0) || !isfinite(half)) return 0;
  *out = (value * 0.5 - mn * 0.5) / half;
  return 1;
}

static uint16_t clamp_u16(double v) {
  if (!isfinite(v)) return 0;
  /* Math.round of an integer-valued double is the identity */
  const double r = floor(v + 0.5);
  return (uint16_t)(r < 0 ? 0 : (r > 65535 ? 65535 : r));
}

/* quantizeBound (:197-204) */
static uint16_t quantize(double value, double mn, double mx, int ceil_mode) {
  double u;
  if (!unit_offset(value, mn, mx, &u)) return 0;
  if (mx == mn) return ceil_mode ? 65535 : 0;
  const double q = u * 65535;
  return clamp_u16(ceil_mode ? ceil(q) + 1 : floor(q) - 1);
}

/* dequantizeBound (:206-213) */
static double dequantize(double q, double mn, double mx) {
  if (!isfinite(mn) || !isfinite(mx) || mn > mx) return NAN;
  if (mx == mn) return mn;
  if (q <= 0) return mn;
  if (q >= 65535) return mx;
  const double t = q / 65535;
  return mn * (1 - t) + mx * t;
}

static double surface_area(const bnode *n) {
  const double dx = js_max(0, n->mx[0] - n->mn[0]);
  const double dy = js_max(0, n->mx[1] - n->mn[1]);
  const double dz = js_max(0, n->mx[2] - n->mn[2]);
  return dx * dy + dy * dz + dz * dx;
}

/* collectWideChildren (:215-246): repeatedly open the interior candidate of
 * largest surface area (first one on ties) in place, up to 8 candidates. */
static uint32_t collect_children(const uint32_t *w, uint32_t root, uint32_t *kids) {
  const bnode r = read_node(w, root);
  if (is_leaf(r.w7)) {
    kids[0] = root;
    return 1;
  }
  uint32_t n = 2;
  kids[0] = root + 1;
  kids[1] = root + r.w6;
  while (n < KIDS) {
    int at = -1;
    double best = -1;
    for (uint32_t i = 0; i < n; ++i) {
      const bnode c = read_node(w, kids[i]);
      if (is_leaf(c.w7)) continue;
      const double a = surface_area(&c);
      if (a > best) {
        best = a;
        at = (int)i;
      }
    }
    if (at < 0) break;
    const uint32_t open = kids[at];
    const bnode o = read_node(w, open);
    memmove(kids + at + 2, kids + at + 1, (n - (uint32_t)at - 1) * sizeof(uint32_t));
    kids[at] = open + 1;
    kids[at + 1] = open + o.w6;
    ++n;
  }
  return n;
}

static int grow(struct orc_cwbvh *c) {
  if (c->node_count < c->cap) return 0;
  const uint32_t cap = c->cap ? c->cap * 2 : 64;
  float *nb = realloc(c->node_bounds, (size_t)cap * 6 * sizeof(float));
  if (nb) c->node_bounds = nb;
  uint16_t *cb = realloc(c->child_bounds, (size_t)cap * KIDS * BOUNDS_U16 * sizeof(uint16_t));
  if (cb) c->child_bounds = cb;
  uint32_t *cm = realloc(c->child_meta, (size_t)cap * KIDS * META * sizeof(uint32_t));
  if (cm) c->child_meta = cm;
  uint32_t *cc = realloc(c->child_count, (size_t)cap * sizeof(uint32_t));
  if (cc) c->child_count = cc;
  if (!nb || !cb || !cm || !cc) return -1;
  c->cap = cap;
  return 0;
}

static uint32_t alloc_node(struct orc_cwbvh *c, const uint32_t *w, uint32_t bin) {
  if (grow(c)) return COUNT_INVALID;
  const uint32_t i = c->node_count++;
  for (int k = 0; k < 6; ++k) c->node_bounds[6 * i + k] = u2f(w[8 * bin + k]);
  c->child_count[i] = 0;
  memset(c->child_bounds + (size_t)i * KIDS * BOUNDS_U16, 0, KIDS * BOUNDS_U16 * sizeof(uint16_t));
  memset(c->child_meta + (size_t)i * KIDS * META, 0, KIDS * META * sizeof(uint32_t));
  return i;
}

/* buildWide (:352-374), preorder allocation, slot order of the candidates. */
static uint32_t build_wide(struct orc_cwbvh *c, const uint32_t *w, uint32_t bin, int *valid, int depth) {
  if (depth > 4096) return COUNT_INVALID;
  const bnode parent = read_node(w, bin);
  const uint32_t wide = alloc_node(c, w, bin);
  if (wide == COUNT_INVALID) return wide;
  uint32_t kids[KIDS];
  const uint32_t n = collect_children(w, bin, kids);
  c->child_count[wide] = n;
  for (uint32_t s = 0; s < n; ++s) {
    const bnode ch = read_node(w, kids[s]);
    for (int a = 0; a < 3; ++a) {
      double tmp;
      if (ch.mn[a] < parent.mn[a] || ch.mx[a] > parent.mx[a] || ch.mn[a] > ch.mx[a] ||
          !unit_offset(ch.mn[a], parent.mn[a], parent.mx[a], &tmp) ||
          !unit_offset(ch.mx[a], parent.mn[a], parent.mx[a], &tmp))
        *valid = 0;
    }
    uint16_t *b = c->child_bounds + (size_t)wide * KIDS * BOUNDS_U16 + s * BOUNDS_U16;
    for (int a = 0; a < 3; ++a) {
      b[a] = quantize(ch.mn[a], parent.mn[a], parent.mx[a], 0);
      b[3 + a] = quantize(ch.mx[a], parent.mn[a], parent.mx[a], 1);
    }
    uint32_t meta0, meta1, meta2;
    if (is_leaf(ch.w7)) {
      meta0 = K_LEAF;
      meta1 = ch.w6;
      meta2 = ch.w7 & 0xFFFFu;
    } else {
      const uint32_t sub = build_wide(c, w, kids[s], valid, depth + 1);
      if (sub == COUNT_INVALID) return sub;
      meta0 = K_NODE;
      meta1 = sub;
      meta2 = 0;
    }
    uint32_t *m = c->child_meta + (size_t)wide * KIDS * META + s * META;  /* after recursion: realloc */
    m[0] = meta0;
    m[1] = meta1;
    m[2] = meta2;
  }
  return wide;
}

/* requiredCwbvhTraversalStackEntries (:60-92) over the tree at `root`. */
static double stack_need(const struct orc_cwbvh *c, uint32_t root, uint32_t first, uint32_t end) {
  if (end == first) return 0;
  if (root < first || root >= end) return INFINITY;
  const uint32_t span = end - first;
  uint32_t *stack = malloc((size_t)(span * KIDS + 1) * sizeof(uint32_t));
  uint8_t *seen = calloc(span, 1);
  if (!stack || !seen) {
    free(stack);
    free(seen);
    return INFINITY;
  }
  uint32_t sp = 0;
  double maximum = 1;
  stack[sp++] = root;
  double result = 0;
  while (sp > 0) {
    const uint32_t node = stack[--sp];
    if (seen[node - first]) {
      result = INFINITY;
      break;
    }
    seen[node - first] = 1;
    const uint32_t count = c->child_count[node];
    if (count > KIDS) {
      result = INFINITY;
      break;
    }
    for (uint32_t s = 0; s < count; ++s) {
      const uint32_t *m = c->child_meta + (size_t)node * KIDS * META + s * META;
      if (m[0] != K_NODE) continue;
      if (m[1] < first || m[1] >= end) {
        result = INFINITY;
        break;
      }
      stack[sp++] = m[1];
    }
    if (result == INFINITY) break;
    if ((double)sp > maximum) maximum = sp;
  }
  free(stack);
  free(seen);
  return result == INFINITY ? INFINITY : maximum;
}

/* Collapses the binary trees at `roots` (a packed forest or a single BLAS at
 * node 0) into one concatenated wide forest.  tri_count = 0 gives the
 * reference's empty wide root. */
int orc_cwbvh_build(const uint32_t *bvh_words, uint32_t node_count, const uint32_t *roots,
                    uint32_t root_count, uint32_t tri_count, orc_cwbvh **out) {
  if (!out || (root_count && !roots) || (node_count && !bvh_words)) return orc_fail("orc_cwbvh_build: NULL argument");
  struct orc_cwbvh *c = calloc(1, sizeof *c);
  if (!c) return orc_fail("orc_cwbvh_build: out of memory");
  c->root_count = root_count;
  c->roots = calloc(root_count ? root_count : 1, sizeof(uint32_t));
  c->status = calloc(root_count ? root_count : 1, sizeof(uint32_t));
  c->stack_need = calloc(root_count ? root_count : 1, sizeof(double));
  if (!c->roots || !c->status || !c->stack_need) {
    orc_cwbvh_free(c);
    return orc_fail("orc_cwbvh_build: out of memory");
  }
  /* layout gate over every binary node (:284-293) */
  int layout_ok = 1;
  for (uint32_t i = 0; i < node_count && layout_ok; ++i) {
    const bnode n = read_node(bvh_words, i);
    for (int a = 0; a < 3; ++a) {
      double tmp;
      if (!unit_offset(n.mn[a], n.mn[a], n.mx[a], &tmp)) layout_ok = 0;
    }
  }
  for (uint32_t r = 0; r < root_count; ++r) {
    if (roots[r] >= node_count) {
      orc_cwbvh_free(c);
      return orc_fail("orc_cwbvh_build: root %u outside %u nodes", roots[r], node_count);
    }
    const uint32_t first = c->node_count;
    const bnode root = read_node(bvh_words, roots[r]);
    c->roots[r] = first;
    if (tri_count == 0 || !layout_ok || (is_leaf(root.w7) && (root.w7 & 0xFFFFu) == 0)) {
      /* emptyWideRoot (:249-265): one node, no children (or the invalid
       * sentinel count for a corrupt layout) */
      const uint32_t i = alloc_node(c, bvh_words, roots[r]);
      if (i == COUNT_INVALID) {
        orc_cwbvh_free(c);
        return orc_fail("orc_cwbvh_build: out of memory");
      }
      for (int k = 0; k < 6; ++k) c->node_bounds[6 * i + k] = 0.0f;
      c->child_count[i] = layout_ok ? 0 : COUNT_INVALID;
      c->status[r] = layout_ok ? 1 : 2;
      c->stack_need[r] = layout_ok ? 0 : INFINITY;
      continue;
    }
    int valid = 1;
    if (build_wide(c, bvh_words, roots[r], &valid, 0) == COUNT_INVALID) {
      orc_cwbvh_free(c);
      return orc_fail("orc_cwbvh_build: out of memory or cyclic binary tree");
    }
    const double need = valid ? stack_need(c, first, first, c->node_count) : INFINITY;
    c->stack_need[r] = need;
    if (!valid || !(need <= STACK_CAP)) {
      for (uint32_t i = first; i < c->node_count; ++i) c->child_count[i] = COUNT_INVALID;
      c->status[r] = 2;
    }
  }
  *out = c;
  return 0;
}

void orc_cwbvh_free(orc_cwbvh *c) {
  if (!c) return;
  free(c->node_bounds);
  free(c->child_bounds);
  free(c->child_meta);
  free(c->child_count);
  free(c->roots);
  free(c->status);
  free(c->stack_need);
  free(c);
}

uint32_t orc_cwbvh_node_count(const orc_cwbvh *c) { return c ? c->node_count : 0; }

int orc_cwbvh_copy(const orc_cwbvh *c, float *node_bounds, uint16_t *child_bounds, uint32_t *child_meta,
                   uint32_t *child_count, uint32_t *roots, uint32_t *status, double *stack_entries) {
  if (!c) return orc_fail("orc_cwbvh_copy: NULL");
  const size_t n = c->node_count;
  if (node_bounds) memcpy(node_bounds, c->node_bounds, n * 6 * sizeof(float));
  if (child_bounds) memcpy(child_bounds, c->child_bounds, n * KIDS * BOUNDS_U16 * sizeof(uint16_t));
  if (child_meta) memcpy(child_meta, c->child_meta, n * KIDS * META * sizeof(uint32_t));
  if (child_count) memcpy(child_count, c->child_count, n * sizeof(uint32_t));
  if (roots) memcpy(roots, c->roots, c->root_count * sizeof(uint32_t));
  if (status) memcpy(status, c->status, c->root_count * sizeof(uint32_t));
  if (stack_entries) memcpy(stack_entries, c->stack_need, c->root_count * sizeof(double));
  return 0;
}

/* Stack requirement of an arbitrary wide topology (for the capacity KATs). */
double orc_cwbvh_stack_entries(const uint32_t *child_meta, const uint32_t *child_count, uint32_t node_count,
                               uint32_t root) {
  if (node_count == 0) return 0;
  struct orc_cwbvh c;
  memset(&c, 0, sizeof c);
  c.child_meta = (uint32_t *)child_meta;
  c.child_count = (uint32_t *)child_count;
  c.node_count = node_count;
  return stack_need(&c, root, 0, node_count);
}

/* ---- binary64 first hit --------------------------------------------------- */

typedef struct {
  double o[3], d[3];
} dray;

static void sub3(const double *a, const double *b, double *r) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}
static void scale3(const double *a, double s, double *r) {
  r[0] = a[0] * s;
  r[1] = a[1] * s;
  r[2] = a[2] * s;
}
static void cross3(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3d(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* intersectTriangle (:551-614): scale-normalised Moller-Trumbore. */
static int tri_hit(const dray *r, const double *a, const double *b, const double *c, double tmin, double *t_out,
                   double *bary) {
  double e1r[3], e2r[3];
  sub3(b, a, e1r);
  sub3(c, a, e2r);
  double es = fabs(e1r[0]);
  for (int k = 1; k < 3; ++k) es = js_max(es, fabs(e1r[k]));
  for (int k = 0; k < 3; ++k) es = js_max(es, fabs(e2r[k]));
  const double ds = js_max(js_max(fabs(r->d[0]), fabs(r->d[1])), fabs(r->d[2]));
  if (!(es > 0) || !isfinite(es) || !(ds > 0) || !isfinite(ds)) return 0;
  double e1[3], e2[3], dir[3], ao0[3], ao[3], n[3], dao[3];
  scale3(e1r, 1 / es, e1);
  scale3(e2r, 1 / es, e2);
  scale3(r->d, 1 / ds, dir);
  sub3(r->o, a, ao0);
  scale3(ao0, 1 / es, ao);
  cross3(e1, e2, n);
  const double det = -dot3d(dir, n);
  const double nl = js_hypot3(n[0], n[1], n[2]);
  const double dl = js_hypot3(dir[0], dir[1], dir[2]);
  if (!(nl > 0) || !isfinite(nl) || !(dl > 0) || !isfinite(dl) || fabs(det) / (nl * dl) <= 1e-7) return 0;
  const double inv = 1 / det;
  cross3(ao, dir, dao);
  const double u = dot3d(e2, dao) * inv;
  const double v = -dot3d(e1, dao) * inv;
  const double t = (dot3d(ao, n) * inv * es) / ds;
  const double w = 1 - u - v;
  const double eps = 1e-6;
  if (!isfinite(u) || !isfinite(v) || !isfinite(w) || !isfinite(t) || u < -eps || v < -eps || w < -eps ||
      t < js_max(tmin, 0))
    return 0;
  *t_out = t;
  bary[0] = w;
  bary[1] = u;
  bary[2] = v;
  return 1;
}

/* intersectAabb (:616-645): returns 1 and the entry distance on a hit. */
static int aabb_hit(const dray *r, const double *mn, const double *mx, double tmin_in, double tmax_in) {
  double tmin = tmin_in, tmax = tmax_in;
  for (int a = 0; a < 3; ++a) {
    const double o = r->o[a], d = r->d[a];
    if (d == 0) {
      if (o < mn[a] || o > mx[a]) return 0;
      continue;
    }
    const double inv = 1 / d;
    double t0 = (mn[a] - o) * inv, t1 = (mx[a] - o) * inv;
    if (t0 > t1) {
      const double tmp = t0;
      t0 = t1;
      t1 = tmp;
    }
    tmin = js_max(tmin, t0);
    tmax = js_min(tmax, t1);
    if (tmax < tmin) return 0;
  }
  return 1;
}

typedef struct {
  int hit;
  double t;
  uint32_t tri;
  double bary[3];
} dhit;

static void load_pos(const orc_scene *s, uint32_t vi, double *p) {
  p[0] = s->positions[4 * (size_t)vi];
  p[1] = s->positions[4 * (size_t)vi + 1];
  p[2] = s->positions[4 * (size_t)vi + 2];
}

/* intersectCompressedWideBvhFirstHit (:754-860) from wide root `root`;
 * `closest` is the running upper bound (tMax).  Returns -1 on a layout the
 * reference would throw on. */
static int first_hit(const struct orc_cwbvh *c, const orc_scene *s, const dray *r, uint32_t root, double tmin,
                     double *closest, dhit *h, uint32_t *stack) {
  if (c->child_count[root] > KIDS) return -1;
  const float *nb = c->node_bounds + 6 * (size_t)root;
  const double rmn[3] = {nb[0], nb[1], nb[2]}, rmx[3] = {nb[3], nb[4], nb[5]};
  if (!aabb_hit(r, rmn, rmx, tmin, *closest)) return 0;
  uint32_t sp = 0;
  stack[sp++] = root;
  while (sp > 0) {
    const uint32_t node = stack[--sp];
    const uint32_t count = c->child_count[node];
    if (count > KIDS) return -1;
    const float *pb = c->node_bounds + 6 * (size_t)node;
    for (uint32_t slot = 0; slot < count; ++slot) {
      const uint32_t *m = c->child_meta + (size_t)node * KIDS * META + slot * META;
      const uint16_t *q = c->child_bounds + (size_t)node * KIDS * BOUNDS_U16 + slot * BOUNDS_U16;
      double mn[3], mx[3];
      for (int a = 0; a < 3; ++a) {
        mn[a] = dequantize(q[a], pb[a], pb[3 + a]);
        mx[a] = dequantize(q[3 + a], pb[a], pb[3 + a]);
      }
      if (!aabb_hit(r, mn, mx, tmin, *closest)) continue;
      if (m[0] == K_NODE) {
        if (m[1] >= c->node_count || sp >= c->node_count) return -1;
        stack[sp++] = m[1];
        continue;
      }
      if (m[0] != K_LEAF) return -1;
      if (m[2] == 0 || m[1] > s->triangle_count || m[2] > s->triangle_count - m[1]) return -1;
      for (uint32_t i = 0; i < m[2]; ++i) {
        const uint32_t tri = m[1] + i;
        const uint32_t *ix = s->indices + 4 * (size_t)tri;
        double a[3], b[3], cc[3], t, bary[3];
        load_pos(s, ix[0], a);
        load_pos(s, ix[1], b);
        load_pos(s, ix[2], cc);
        if (tri_hit(r, a, b, cc, tmin, &t, bary) && t > tmin && t < *closest) {
          *closest = t;
          h->hit = 1;
          h->t = t;
          h->tri = tri;
          memcpy(h->bary, bary, sizeof bary);
        }
      }
    }
  }
  return 0;
}

/* tlasIntersect (tlas.ts:1223-1310): candidate instances (original indices)
 * in stack order.  Returns the candidate count or -1 on a corrupt tree. */
static int tlas_candidates(const orc_scene *s, const double *o, const double *d, double tmax, uint32_t *out,
                           uint32_t cap, uint32_t *stack) {
  int n = 0;
  if (s->tlas_node_count == 0) return 0;
  if (!isfinite(o[0]) || !isfinite(o[1]) || !isfinite(o[2]) || !isfinite(d[0]) || !isfinite(d[1]) ||
      !isfinite(d[2]) || js_hypot3(d[0], d[1], d[2]) == 0 || isnan(tmax) || tmax < 0)
    return 0;
  uint32_t sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const uint32_t idx = stack[--sp];
    if (idx >= s->tlas_node_count) return -1;
    const uint32_t *w = s->tlas_nodes + 8 * (size_t)idx;
    double mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
      mn[a] = u2f(w[a]);
      mx[a] = u2f(w[3 + a]);
    }
    double t_enter = 0, t_exit = tmax;
    int pass = 1;
    for (int a = 0; a < 3 && pass; ++a) {
      if (fabs(d[a]) < 1e-30) {
        pass = o[a] >= mn[a] && o[a] <= mx[a];
        continue;
      }
      const double t0 = (mn[a] - o[a]) / d[a], t1 = (mx[a] - o[a]) / d[a];
      t_enter = js_max(t_enter, js_min(t0, t1));
      t_exit = js_min(t_exit, js_max(t0, t1));
      pass = t_enter <= t_exit;
    }
    if (!pass || !(t_enter <= t_exit && t_exit >= 0)) continue;
    const uint32_t split = w[7];
    if ((split >> 16) == 0xFFFFu) {
      const uint32_t off = w[6], count = split & 0xFFFFu;
      if (count == 0 || off + count > s->instance_index_count) return -1;
      for (uint32_t k = 0; k < count; ++k) {
        if ((uint32_t)n >= cap) return -1;
        out[n++] = s->instance_indices[off + k];
      }
    } else {
      const uint32_t right = idx + w[6];
      if (w[6] <= 1 || idx + 1 >= s->tlas_node_count || right >= s->tlas_node_count) return -1;
      stack[sp++] = idx + 1;
      stack[sp++] = right;
    }
  }
  return n;
}

typedef struct {
  const orc_scene *s;
  const struct orc_cwbvh *c;
  const uint32_t *blas_wide_root; /* per instance */
  const orc_ray *rays;
  uint32_t n;
  orc_hit64 *out;
  uint32_t next;
  pthread_mutex_t mu;
  int error;
} job;

static void trace_one(job *J, const orc_ray *ry, orc_hit64 *o, uint32_t *cand, uint32_t *stack) {
  const orc_scene *s = J->s;
  const double wo[3] = {ry->o[0], ry->o[1], ry->o[2]}, wd[3] = {ry->d[0], ry->d[1], ry->d[2]};
  double closest = ry->tmax;
  const double tmin = ry->tmin;
  o->hit = 0;
  o->prim = 0xFFFFFFFFu;
  o->inst = 0xFFFFFFFFu;
  o->t = closest;
  o->bary[0] = o->bary[1] = o->bary[2] = 0;
  if (!(tmin >= 0) || isnan(closest) || closest < tmin) return;
  const int nc = tlas_candidates(s, wo, wd, closest, cand, s->instance_count + 1, stack);
  if (nc < 0) {
    J->error = 1;
    return;
  }
  for (int k = 0; k < nc; ++k) {
    const uint32_t inst = cand[k];
    const float *m = s->w2l + 16 * (size_t)inst;
    dray lr;
    for (int i = 0; i < 3; ++i) {
      lr.o[i] = (double)m[i] * wo[0] + (double)m[4 + i] * wo[1] + (double)m[8 + i] * wo[2] + (double)m[12 + i];
      lr.d[i] = (double)m[i] * wd[0] + (double)m[4 + i] * wd[1] + (double)m[8 + i] * wd[2];
    }
    const double w = (double)m[3] * wo[0] + (double)m[7] * wo[1] + (double)m[11] * wo[2] + (double)m[15];
    if (fabs(w) > 1e-8 && w != 1.0)
      for (int i = 0; i < 3; ++i) lr.o[i] /= w;
    dhit h = {0, 0, 0, {0, 0, 0}};
    const int rc = first_hit(J->c, s, &lr, J->blas_wide_root[inst], tmin, &closest, &h, stack);
    if (rc < 0) {
      J->error = 1;
      return;
    }
    if (h.hit) {
      o->hit = 1;
      o->prim = h.tri;
      o->inst = inst;
      o->t = h.t;
      memcpy(o->bary, h.bary, sizeof h.bary);
    }
  }
}

static void *worker(void *arg) {
  job *J = arg;
  const uint32_t stack_words = J->c->node_count + J->s->tlas_node_count + 64;
  uint32_t *stack = malloc((size_t)stack_words * sizeof(uint32_t));
  uint32_t *cand = malloc((size_t)(J->s->instance_count + 1) * sizeof(uint32_t));
  if (!stack || !cand) {
    J->error = 1;
    free(stack);
    free(cand);
    return NULL;
  }
  for (;;) {
    pthread_mutex_lock(&J->mu);
    const uint32_t b = J->next;
    J->next = b + 256 < J->n ? b + 256 : J->n;
    const uint32_t e = J->next;
    pthread_mutex_unlock(&J->mu);
    if (b >= e) break;
    for (uint32_t i = b; i < e; ++i) trace_one(J, J->rays + i, J->out + i, cand, stack);
  }
  free(stack);
  free(cand);
  return NULL;
}

int orc_cwbvh_trace_closest(const orc_scene *s, const orc_cwbvh *c, const uint32_t *blas_wide_root,
                            const orc_ray *rays, uint32_t n, orc_hit64 *out, uint32_t threads) {
  if (!s || !c || !blas_wide_root || (n && (!rays || !out))) return orc_fail("orc_cwbvh_trace_closest: NULL");
  for (uint32_t i = 0; i < s->instance_count; ++i)
    if (blas_wide_root[i] >= c->node_count) return orc_fail("orc_cwbvh_trace_closest: bad wide root");
  job J;
  memset(&J, 0, sizeof J);
  J.s = s;
  J.c = c;
  J.blas_wide_root = blas_wide_root;
  J.rays = rays;
  J.n = n;
  J.out = out;
  pthread_mutex_init(&J.mu, NULL);
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t th[256];
  uint32_t started = 0;
  for (uint32_t t = 1; t < threads; ++t)
    if (pthread_create(&th[started], NULL, worker, &J) == 0) ++started;
  worker(&J);
  for (uint32_t t = 0; t < started; ++t) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&J.mu);
  if (J.error) return orc_fail("orc_cwbvh_trace_closest: corrupt layout");
  return 0;
}
