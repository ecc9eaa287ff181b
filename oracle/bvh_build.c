/*
 * bvh_build.c -- oracle restatement of vitrum's BLAS/TLAS builders and the
 * scene packer.  TEST INFRASTRUCTURE ONLY (see vt_oracle.h).
 *
 * Decisions are made in binary64 exactly as the reference's JavaScript does;
 * `js_min` / `js_max` reproduce Math.min / Math.max including signed zeros,
 * which reach the output through Float32Array stores.
 */
#include "vt_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

const char *orc_last_error(void) { return g_err; }

int orc_fail(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return -1;
}

#define LEAF_FLAG 0xFFFF0000u
/* packages/shared-bvh/src/strides.ts:110-125 */
#define BVH_STACK_DEPTH 60
#define TLAS_STACK_DEPTH 64
#define MAX_BUILD_DEPTH 58

static uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* Math.min / Math.max for non-NaN operands. */
static double js_min(double a, double b) {
  if (a < b) return a;
  if (b < a) return b;
  if (a == 0.0 && (signbit(a) || signbit(b))) return -0.0;
  return a;
}
static double js_max(double a, double b) {
  if (a > b) return a;
  if (b > a) return b;
  if (a == 0.0 && !(signbit(a) && signbit(b))) return 0.0;
  return a;
}

/* ======================================================================== */
/* BLAS: packages/shared-bvh/src/buildArrayBvh.ts                            */
/* ======================================================================== */

typedef struct {
  uint32_t tri;
  double mn[3], mx[3], c[3];
} tri_rec;

typedef struct {
  double mn[3], mx[3];
  uint32_t right;
  uint32_t split;
} bnode;

typedef struct {
  double mn[3], mx[3];
  uint32_t count;
} bin_t;

typedef struct {
  const tri_rec *recs;
  bnode *nodes;
  uint32_t node_count;
  uint32_t *ordered;
  uint32_t ordered_count;
  uint32_t max_leaf, num_bins, max_depth;
  bin_t *bins;
  float *pmn[3], *pmx[3], *smn[3], *smx[3];
  int32_t *pcnt, *scnt;
  int failed;
} bctx;

/* surfaceArea (buildArrayBvh.ts:223-231): half area with clamped extents. */
static double blas_sa(double x0, double y0, double z0, double x1, double y1,
                      double z1) {
  double dx = js_max(0.0, x1 - x0);
  double dy = js_max(0.0, y1 - y0);
  double dz = js_max(0.0, z1 - z0);
  return dx * dy + dy * dz + dz * dx;
}

static int g_sort_axis;
static const tri_rec *g_sort_recs;
static int cmp_tri_median(const void *pa, const void *pb) {
  const tri_rec *a = &g_sort_recs[*(const uint32_t *)pa];
  const tri_rec *b = &g_sort_recs[*(const uint32_t *)pb];
  double d = a->c[g_sort_axis] - b->c[g_sort_axis];
  if (d != 0.0) return d < 0.0 ? -1 : 1;
  return a->tri < b->tri ? -1 : (a->tri > b->tri ? 1 : 0);
}

/* makeLeaf (buildArrayBvh.ts:463-480). */
static uint32_t blas_leaf(bctx *x, bnode *node, const uint32_t *subset,
                          uint32_t n, uint32_t index) {
  if (n > x->max_leaf) {
    x->failed = orc_fail("leaf triangle count %u exceeds maxLeafTriangles=%u", n,
                         x->max_leaf);
    return index;
  }
  node->right = x->ordered_count;
  for (uint32_t i = 0; i < n; ++i) x->ordered[x->ordered_count++] = x->recs[subset[i]].tri;
  node->split = LEAF_FLAG | n;
  return index;
}

/* build (buildArrayBvh.ts:511-706). subset holds record indices. */
static uint32_t blas_build_rec(bctx *x, uint32_t *subset, uint32_t n,
                               uint32_t depth) {
  const tri_rec *R = x->recs;
  uint32_t index = x->node_count++;
  bnode *node = &x->nodes[index];
  for (int k = 0; k < 3; ++k) {
    node->mn[k] = INFINITY;
    node->mx[k] = -INFINITY;
  }
  node->right = 0;
  node->split = 0;
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      node->mn[k] = js_min(node->mn[k], R[subset[i]].mn[k]);
      node->mx[k] = js_max(node->mx[k], R[subset[i]].mx[k]);
    }
  if (n <= x->max_leaf) return blas_leaf(x, node, subset, n, index);
  if (depth >= x->max_depth) {
    /* buildArrayBvh.ts:537-546, with the 16-bit leaf-count suffix */
    x->failed = orc_fail("maxDepth=%u prevents enforcing maxLeafTriangles=%u: depth %u "
                         "still contains %u triangles%s.",
                         x->max_depth, x->max_leaf, depth, n,
                         n > 0xFFFFu ? " and the packed 16-bit limit (0xFFFF = 65535)" : "");
    return index;
  }
  double cmin[3] = {INFINITY, INFINITY, INFINITY};
  double cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      cmin[k] = js_min(cmin[k], R[subset[i]].c[k]);
      cmax[k] = js_max(cmax[k], R[subset[i]].c[k]);
    }
  double parent_sa =
      blas_sa(node->mn[0], node->mn[1], node->mn[2], node->mx[0], node->mx[1], node->mx[2]);
  double leaf_cost = (double)n;
  double best_cost = INFINITY;
  int best_axis = 0;
  uint32_t best_split = 0;
  const uint32_t K = x->num_bins;
  for (int axis = 0; axis < 3; ++axis) {
    double span = cmax[axis] - cmin[axis];
    if (span <= 1e-9) continue;
    for (uint32_t b = 0; b < K; ++b) {
      for (int k = 0; k < 3; ++k) {
        x->bins[b].mn[k] = INFINITY;
        x->bins[b].mx[k] = -INFINITY;
      }
      x->bins[b].count = 0;
    }
    for (uint32_t i = 0; i < n; ++i) {
      const tri_rec *r = &R[subset[i]];
      double t = (r->c[axis] - cmin[axis]) / span;
      double bf = floor(t * (double)K);
      uint32_t bi = bf < (double)(K - 1) ? (uint32_t)bf : K - 1;
      bin_t *bin = &x->bins[bi];
      for (int k = 0; k < 3; ++k) {
        bin->mn[k] = js_min(bin->mn[k], r->mn[k]);
        bin->mx[k] = js_max(bin->mx[k], r->mx[k]);
      }
      bin->count += 1;
    }
    {
      double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
      int32_t cnt = 0;
      for (uint32_t b = 0; b < K; ++b) {
        for (int k = 0; k < 3; ++k) {
          mn[k] = js_min(mn[k], x->bins[b].mn[k]);
          mx[k] = js_max(mx[k], x->bins[b].mx[k]);
          x->pmn[k][b] = (float)mn[k];
          x->pmx[k][b] = (float)mx[k];
        }
        cnt += (int32_t)x->bins[b].count;
        x->pcnt[b] = cnt;
      }
    }
    {
      double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
      int32_t cnt = 0;
      for (int32_t b = (int32_t)K - 1; b >= 0; --b) {
        for (int k = 0; k < 3; ++k) {
          mn[k] = js_min(mn[k], x->bins[b].mn[k]);
          mx[k] = js_max(mx[k], x->bins[b].mx[k]);
          x->smn[k][b] = (float)mn[k];
          x->smx[k][b] = (float)mx[k];
        }
        cnt += (int32_t)x->bins[b].count;
        x->scnt[b] = cnt;
      }
    }
    for (uint32_t s = 0; s + 1 < K; ++s) {
      int32_t lc = x->pcnt[s], rc = x->scnt[s + 1];
      if (lc == 0 || rc == 0) continue;
      double lsa = blas_sa(x->pmn[0][s], x->pmn[1][s], x->pmn[2][s], x->pmx[0][s],
                           x->pmx[1][s], x->pmx[2][s]);
      double rsa = blas_sa(x->smn[0][s + 1], x->smn[1][s + 1], x->smn[2][s + 1],
                           x->smx[0][s + 1], x->smx[1][s + 1], x->smx[2][s + 1]);
      double cost = parent_sa > 0.0 ? (lsa * (double)lc + rsa * (double)rc) / parent_sa
                                    : lsa * (double)lc + rsa * (double)rc;
      if (cost < best_cost) {
        best_cost = cost;
        best_axis = axis;
        best_split = s;
      }
    }
  }
  uint32_t split_axis = (uint32_t)best_axis;
  uint32_t *left = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint32_t *right = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint32_t nl = 0, nr = 0;
  if (best_cost < leaf_cost && best_cost != INFINITY) {
    double span = cmax[best_axis] - cmin[best_axis];
    for (uint32_t i = 0; i < n; ++i) {
      const tri_rec *r = &R[subset[i]];
      double t = (r->c[best_axis] - cmin[best_axis]) / span;
      double bf = floor(t * (double)K);
      uint32_t bi = bf < (double)(K - 1) ? (uint32_t)bf : K - 1;
      if (bi <= best_split)
        left[nl++] = subset[i];
      else
        right[nr++] = subset[i];
    }
  }
  if (nl == 0 || nr == 0) {
    /* Median fallback (buildArrayBvh.ts:677-694). */
    split_axis = 0;
    double widest = cmax[0] - cmin[0];
    for (int axis = 1; axis < 3; ++axis) {
      double s = cmax[axis] - cmin[axis];
      if (s > widest) {
        split_axis = (uint32_t)axis;
        widest = s;
      }
    }
    uint32_t *ordered = (uint32_t *)malloc(sizeof(uint32_t) * n);
    memcpy(ordered, subset, sizeof(uint32_t) * n);
    g_sort_axis = (int)split_axis;
    g_sort_recs = R;
    qsort(ordered, n, sizeof(uint32_t), cmp_tri_median);
    uint32_t mid = n / 2;
    nl = mid;
    nr = n - mid;
    memcpy(left, ordered, sizeof(uint32_t) * nl);
    memcpy(right, ordered + mid, sizeof(uint32_t) * nr);
    free(ordered);
  }
  blas_build_rec(x, left, nl, depth + 1);
  uint32_t rchild = x->failed ? 0 : blas_build_rec(x, right, nr, depth + 1);
  free(left);
  free(right);
  node = &x->nodes[index];
  node->right = rchild - index;
  node->split = split_axis;
  return index;
}

int orc_build_blas(const float *positions, uint32_t position_floats,
                   uint32_t position_stride, const uint32_t *indices,
                   uint32_t index_words, uint32_t index_stride,
                   const uint32_t *tri_material_ids, uint32_t max_leaf,
                   uint32_t num_bins, uint32_t max_depth,
                   uint32_t *out_node_words, uint32_t *out_node_count,
                   uint32_t *out_reordered_indices, uint32_t *out_reordered_mat,
                   uint32_t *out_reordered_to_source) {
  g_err[0] = 0;
  if (position_stride != 3 && position_stride != 4)
    return orc_fail("positionStride must be exactly 3 or 4.");
  if (index_stride != 3 && index_stride != 4)
    return orc_fail("indexStride must be exactly 3 or 4.");
  if (max_leaf < 1 || max_leaf > 0xFFFF) return orc_fail("maxLeafTriangles out of range");
  if (num_bins < 2 || num_bins > 256) return orc_fail("binCount out of range");
  if (max_depth > MAX_BUILD_DEPTH) return orc_fail("maxDepth out of range");
  if (position_floats % position_stride != 0)
    return orc_fail("positions.length %u must be divisible by positionStride %u",
                    position_floats, position_stride);
  if (index_words % index_stride != 0)
    return orc_fail("indices.length %u must be divisible by indexStride %u", index_words,
                    index_stride);
  for (uint32_t i = 0; i < position_floats; ++i)
    if (!isfinite(positions[i])) return orc_fail("positions[%u] must be finite", i);
  uint32_t tri_count = index_words / index_stride;
  if (tri_count == 0) {
    memset(out_node_words, 0, 32);
    out_node_words[7] = LEAF_FLAG;
    *out_node_count = 1;
    if (index_words) memcpy(out_reordered_indices, indices, 4 * (size_t)index_words);
    return 0;
  }
  uint32_t vertex_count = position_floats / position_stride;
  tri_rec *recs = (tri_rec *)malloc(sizeof(tri_rec) * tri_count);
  for (uint32_t t = 0; t < tri_count; ++t) {
    uint32_t i0 = indices[t * index_stride], i1 = indices[t * index_stride + 1],
             i2 = indices[t * index_stride + 2];
    if (i0 >= vertex_count || i1 >= vertex_count || i2 >= vertex_count) {
      free(recs);
      return orc_fail("triangle %u references an out-of-range vertex", t);
    }
    const float *a = positions + (size_t)i0 * position_stride;
    const float *b = positions + (size_t)i1 * position_stride;
    const float *c = positions + (size_t)i2 * position_stride;
    recs[t].tri = t;
    for (int k = 0; k < 3; ++k) {
      recs[t].mn[k] = js_min(js_min(a[k], b[k]), c[k]);
      recs[t].mx[k] = js_max(js_max(a[k], b[k]), c[k]);
      recs[t].c[k] = 0.5 * (recs[t].mn[k] + recs[t].mx[k]);
    }
  }
  bctx x;
  memset(&x, 0, sizeof x);
  x.recs = recs;
  x.nodes = (bnode *)malloc(sizeof(bnode) * (2 * (size_t)tri_count));
  x.ordered = (uint32_t *)malloc(sizeof(uint32_t) * tri_count);
  x.max_leaf = max_leaf;
  x.num_bins = num_bins;
  x.max_depth = max_depth;
  x.bins = (bin_t *)malloc(sizeof(bin_t) * num_bins);
  for (int k = 0; k < 3; ++k) {
    x.pmn[k] = (float *)malloc(4 * num_bins);
    x.pmx[k] = (float *)malloc(4 * num_bins);
    x.smn[k] = (float *)malloc(4 * num_bins);
    x.smx[k] = (float *)malloc(4 * num_bins);
  }
  x.pcnt = (int32_t *)malloc(4 * num_bins);
  x.scnt = (int32_t *)malloc(4 * num_bins);
  uint32_t *all = (uint32_t *)malloc(sizeof(uint32_t) * tri_count);
  for (uint32_t t = 0; t < tri_count; ++t) all[t] = t;
  blas_build_rec(&x, all, tri_count, 0);
  free(all);
  int rc = 0;
  if (x.failed) {
    rc = -1;
  } else {
    for (uint32_t i = 0; i < x.node_count; ++i) {
      uint32_t *w = out_node_words + 8 * (size_t)i;
      for (int k = 0; k < 3; ++k) {
        w[k] = f2u((float)x.nodes[i].mn[k]);
        w[3 + k] = f2u((float)x.nodes[i].mx[k]);
      }
      w[6] = x.nodes[i].right;
      w[7] = x.nodes[i].split;
    }
    *out_node_count = x.node_count;
    for (uint32_t nt = 0; nt < x.ordered_count; ++nt) {
      uint32_t ot = x.ordered[nt];
      if (out_reordered_to_source) out_reordered_to_source[nt] = ot;
      for (uint32_t k = 0; k < 3; ++k)
        out_reordered_indices[nt * index_stride + k] = indices[ot * index_stride + k];
      for (uint32_t k = 3; k < index_stride; ++k) out_reordered_indices[nt * index_stride + k] = 0;
      if (out_reordered_mat) out_reordered_mat[nt] = tri_material_ids ? tri_material_ids[ot] : 0;
    }
    rc = orc_validate_bvh(out_node_words, x.node_count, NULL, -1, x.ordered_count,
                          BVH_STACK_DEPTH, NULL);
  }
  free(recs);
  free(x.nodes);
  free(x.ordered);
  free(x.bins);
  for (int k = 0; k < 3; ++k) {
    free(x.pmn[k]);
    free(x.pmx[k]);
    free(x.smn[k]);
    free(x.smx[k]);
  }
  free(x.pcnt);
  free(x.scnt);
  return rc;
}

/* ======================================================================== */
/* validateBvhEncoding.ts:55                                                 */
/* ======================================================================== */

typedef struct {
  uint32_t start, count;
} range_t;
static int cmp_range(const void *a, const void *b) {
  uint32_t x = ((const range_t *)a)->start, y = ((const range_t *)b)->start;
  return x < y ? -1 : (x > y ? 1 : 0);
}

int orc_validate_bvh(const uint32_t *w, uint32_t total, const uint32_t *roots_in,
                     int32_t root_count, int64_t triangle_count, uint32_t cap,
                     uint32_t *out_proof) {
  if (cap < 1) return orc_fail("traversalStackCapacity must be a positive safe integer.");
  uint32_t default_root = 0;
  const uint32_t *roots = roots_in;
  uint32_t nroots;
  if (root_count < 0) {
    roots = &default_root;
    nroots = total == 0 ? 0 : 1;
  } else {
    nroots = (uint32_t)root_count;
  }
  if (total == 0) {
    if (nroots) return orc_fail("an empty node buffer cannot declare roots.");
    if (triangle_count > 0) return orc_fail("empty node buffer cannot cover triangles.");
    if (out_proof) memset(out_proof, 0, 7 * 4), out_proof[6] = cap;
    return 0;
  }
  if (nroots == 0) return orc_fail("non-empty node buffer must declare at least one root.");
  uint8_t *visited = (uint8_t *)calloc(total, 1);
  uint8_t *root_seen = (uint8_t *)calloc(total, 1);
  range_t *ranges = (range_t *)malloc(sizeof(range_t) * (total + 1));
  uint32_t nranges = 0, interior = 0, leaves = 0, max_depth = 0, max_stack = 0;
  uint32_t *stk_node = (uint32_t *)malloc(sizeof(uint32_t) * (2 * (size_t)total + 2));
  uint32_t *stk_depth = (uint32_t *)malloc(sizeof(uint32_t) * (2 * (size_t)total + 2));
  int rc = 0;
  for (uint32_t ri = 0; ri < nroots && !rc; ++ri) {
    uint32_t root = roots[ri];
    if (root >= total) {
      rc = orc_fail("root %u is outside [0, %u].", root, total - 1);
      break;
    }
    if (root_seen[root]) {
      rc = orc_fail("root %u is declared more than once.", root);
      break;
    }
    root_seen[root] = 1;
    size_t sp = 0;
    stk_node[sp] = root;
    stk_depth[sp] = 0;
    ++sp;
    while (sp > 0) {
      --sp;
      uint32_t node = stk_node[sp], depth = stk_depth[sp];
      if (visited[node]) {
        rc = orc_fail("node %u is reachable more than once.", node);
        break;
      }
      visited[node] = 1;
      if (depth > max_depth) max_depth = depth;
      if (depth + 1 > max_stack) max_stack = depth + 1;
      if (max_stack > cap) {
        rc = orc_fail("traversal requires %u live stack entries, exceeding shader capacity %u.",
                      max_stack, cap);
        break;
      }
      const uint32_t *n = w + 8 * (size_t)node;
      float mn[3], mx[3];
      for (int k = 0; k < 3; ++k) mn[k] = u2f(n[k]), mx[k] = u2f(n[3 + k]);
      if (!isfinite(mn[0]) || !isfinite(mn[1]) || !isfinite(mn[2]) || !isfinite(mx[0]) ||
          !isfinite(mx[1]) || !isfinite(mx[2])) {
        rc = orc_fail("node %u has non-finite bounds.", node);
        break;
      }
      if (mn[0] > mx[0] || mn[1] > mx[1] || mn[2] > mx[2]) {
        rc = orc_fail("node %u has inverted bounds.", node);
        break;
      }
      uint32_t soc = n[7];
      if ((soc >> 16) == 0xFFFFu) {
        ++leaves;
        uint32_t count = soc & 0xFFFFu, off = n[6];
        if (triangle_count >= 0) {
          if ((int64_t)off > triangle_count || (int64_t)count > triangle_count - off) {
            rc = orc_fail("leaf node %u triangle range exceeds triangle capacity.", node);
            break;
          }
          ranges[nranges].start = off;
          ranges[nranges].count = count;
          ++nranges;
        }
        continue;
      }
      ++interior;
      if (soc > 2) {
        rc = orc_fail("interior node %u has invalid split axis %u.", node, soc);
        break;
      }
      uint32_t ro = n[6];
      uint64_t lc = (uint64_t)node + 1, rch = (uint64_t)node + ro;
      if (ro <= 1 || lc >= total || rch >= total) {
        rc = orc_fail("interior node %u has invalid relative right-child offset %u.", node, ro);
        break;
      }
      uint64_t ch[2] = {lc, rch};
      for (int c = 0; c < 2 && !rc; ++c) {
        const uint32_t *q = w + 8 * ch[c];
        if (u2f(q[0]) < mn[0] || u2f(q[1]) < mn[1] || u2f(q[2]) < mn[2] || u2f(q[3]) > mx[0] ||
            u2f(q[4]) > mx[1] || u2f(q[5]) > mx[2])
          rc = orc_fail("interior node %u bounds do not enclose child %u.", node, (uint32_t)ch[c]);
      }
      if (rc) break;
      stk_node[sp] = (uint32_t)rch;
      stk_depth[sp] = depth + 1;
      ++sp;
      stk_node[sp] = (uint32_t)lc;
      stk_depth[sp] = depth + 1;
      ++sp;
    }
  }
  if (!rc) {
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < total; ++i) cnt += visited[i];
    if (cnt != total) rc = orc_fail("%u node(s) are unreachable.", total - cnt);
  }
  if (!rc && triangle_count >= 0) {
    qsort(ranges, nranges, sizeof(range_t), cmp_range);
    uint64_t cursor = 0;
    for (uint32_t i = 0; i < nranges && !rc; ++i) {
      if (ranges[i].start != cursor)
        rc = orc_fail("leaf triangle range [%u, %u) %s triangle %u.", ranges[i].start,
                      ranges[i].start + ranges[i].count,
                      ranges[i].start < cursor ? "overlaps" : "leaves a gap after",
                      (uint32_t)cursor);
      cursor += ranges[i].count;
    }
    if (!rc && (int64_t)cursor != triangle_count)
      rc = orc_fail("leaf ranges cover %u triangle(s); expected %lld.", (uint32_t)cursor,
                    (long long)triangle_count);
  }
  if (!rc && out_proof) {
    out_proof[0] = total;
    out_proof[1] = nroots;
    out_proof[2] = interior;
    out_proof[3] = leaves;
    out_proof[4] = max_depth;
    out_proof[5] = max_stack;
    out_proof[6] = cap;
  }
  free(visited);
  free(root_seen);
  free(ranges);
  free(stk_node);
  free(stk_depth);
  return rc;
}

/* ======================================================================== */
/* TLAS: packages/shared-bvh/src/tlas.ts                                     */
/* ======================================================================== */

typedef struct {
  uint32_t orig;
  double c[3], mn[3], mx[3];
} inst_rec;

typedef struct {
  double mn[3], mx[3];
  uint32_t right, split;
} tnode;

typedef struct {
  tnode *nodes;
  uint32_t node_count;
  uint32_t *perm;
  uint32_t perm_count;
  uint32_t max_leaf, num_bins, max_depth;
  uint32_t fallbacks[4];
  int failed;
} tctx;

/* aabbSurfaceArea (tlas.ts:257-265). */
static double tlas_sa(const double *mn, const double *mx) {
  double dx = js_max(0.0, mx[0] - mn[0]);
  double dy = js_max(0.0, mx[1] - mn[1]);
  double dz = js_max(0.0, mx[2] - mn[2]);
  return 2.0 * (dx * dy + dy * dz + dz * dx);
}

/* minimumBalancedDepth (tlas.ts:416-419): ceil(log2(ceil(count/maxLeaf))). */
static uint32_t min_balanced_depth(uint32_t count, uint32_t max_leaf) {
  uint64_t leaves = ((uint64_t)count + max_leaf - 1) / max_leaf;
  if (leaves <= 1) return 0;
  uint32_t d = 0;
  while ((1ull << d) < leaves) ++d;
  return d;
}

static int split_fits_depth(uint32_t l, uint32_t r, uint32_t child_depth, uint32_t max_depth,
                            uint32_t max_leaf) {
  int64_t remaining = (int64_t)max_depth - (int64_t)child_depth;
  return (int64_t)min_balanced_depth(l, max_leaf) <= remaining &&
         (int64_t)min_balanced_depth(r, max_leaf) <= remaining;
}

static const inst_rec *g_inst_sort;
static int cmp_inst_median(const void *pa, const void *pb) {
  const inst_rec *a = &g_inst_sort[*(const uint32_t *)pa];
  const inst_rec *b = &g_inst_sort[*(const uint32_t *)pb];
  double d = a->c[g_sort_axis] - b->c[g_sort_axis];
  if (d != 0.0) return d < 0.0 ? -1 : 1;
  return a->orig < b->orig ? -1 : (a->orig > b->orig ? 1 : 0);
}

static uint32_t tlas_build_rec(tctx *x, const inst_rec *R, uint32_t *recs, uint32_t n,
                               uint32_t depth) {
  uint32_t this_idx = x->node_count;
  if (n <= x->max_leaf) {
    /* buildLeaf (tlas.ts:286-309): strict comparisons. */
    tnode *leaf = &x->nodes[x->node_count++];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        if (R[recs[i]].mn[k] < mn[k]) mn[k] = R[recs[i]].mn[k];
        if (R[recs[i]].mx[k] > mx[k]) mx[k] = R[recs[i]].mx[k];
      }
    memcpy(leaf->mn, mn, sizeof mn);
    memcpy(leaf->mx, mx, sizeof mx);
    leaf->split = LEAF_FLAG | n;
    leaf->right = x->perm_count;
    for (uint32_t i = 0; i < n; ++i) x->perm[x->perm_count++] = R[recs[i]].orig;
    return this_idx;
  }
  if (depth >= x->max_depth || min_balanced_depth(n, x->max_leaf) > x->max_depth - depth) {
    x->failed = orc_fail("Cannot encode %u instances at depth %u within the traversal-safe "
                         "maximum depth %u.",
                         n, depth, x->max_depth);
    return this_idx;
  }
  /* computeCentroidBounds (tlas.ts:316-328). */
  double cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      if (R[recs[i]].c[k] < cmin[k]) cmin[k] = R[recs[i]].c[k];
      if (R[recs[i]].c[k] > cmax[k]) cmax[k] = R[recs[i]].c[k];
    }
  /* pickSplit (tlas.ts:330-393). */
  int have = 0, s_axis = 0;
  uint32_t s_bin = 0;
  double s_cost = 0;
  double ext[3] = {cmax[0] - cmin[0], cmax[1] - cmin[1], cmax[2] - cmin[2]};
  const uint32_t K = x->num_bins;
  int degenerate = ext[0] <= 0 && ext[1] <= 0 && ext[2] <= 0;
  if (!degenerate) {
    bin_t *bins = (bin_t *)malloc(sizeof(bin_t) * K);
    double *lsa = (double *)malloc(sizeof(double) * K), *rsa = (double *)malloc(sizeof(double) * K);
    uint32_t *lcn = (uint32_t *)malloc(4 * K), *rcn = (uint32_t *)malloc(4 * K);
    for (int axis = 0; axis < 3; ++axis) {
      if (ext[axis] <= 0) continue;
      for (uint32_t b = 0; b < K; ++b) {
        for (int k = 0; k < 3; ++k) bins[b].mn[k] = INFINITY, bins[b].mx[k] = -INFINITY;
        bins[b].count = 0;
      }
      for (uint32_t i = 0; i < n; ++i) {
        const inst_rec *r = &R[recs[i]];
        double tb = floor((r->c[axis] - cmin[axis]) / ext[axis] * (double)K);
        double cl = tb < (double)(K - 1) ? tb : (double)(K - 1);
        if (cl < 0) cl = 0;
        bin_t *bin = &bins[(uint32_t)cl];
        for (int k = 0; k < 3; ++k) {
          if (r->mn[k] < bin->mn[k]) bin->mn[k] = r->mn[k];
          if (r->mx[k] > bin->mx[k]) bin->mx[k] = r->mx[k];
        }
        bin->count += 1;
      }
      bin_t acc;
      for (int k = 0; k < 3; ++k) acc.mn[k] = INFINITY, acc.mx[k] = -INFINITY;
      acc.count = 0;
      for (uint32_t b = 0; b < K; ++b) {
        if (bins[b].count) { /* unionBinInto (tlas.ts:275-284) */
          for (int k = 0; k < 3; ++k) {
            if (bins[b].mn[k] < acc.mn[k]) acc.mn[k] = bins[b].mn[k];
            if (bins[b].mx[k] > acc.mx[k]) acc.mx[k] = bins[b].mx[k];
          }
          acc.count += bins[b].count;
        }
        lsa[b] = tlas_sa(acc.mn, acc.mx);
        lcn[b] = acc.count;
      }
      for (int k = 0; k < 3; ++k) acc.mn[k] = INFINITY, acc.mx[k] = -INFINITY;
      acc.count = 0;
      for (int32_t b = (int32_t)K - 1; b >= 0; --b) {
        if (bins[b].count) {
          for (int k = 0; k < 3; ++k) {
            if (bins[b].mn[k] < acc.mn[k]) acc.mn[k] = bins[b].mn[k];
            if (bins[b].mx[k] > acc.mx[k]) acc.mx[k] = bins[b].mx[k];
          }
          acc.count += bins[b].count;
        }
        rsa[b] = tlas_sa(acc.mn, acc.mx);
        rcn[b] = acc.count;
      }
      double total_sa = lsa[K - 1];
      if (total_sa <= 0) continue;
      for (uint32_t b = 0; b + 1 < K; ++b) {
        uint32_t lc = lcn[b], rc = rcn[b + 1];
        if (lc == 0 || rc == 0) continue;
        double sah = ((double)lc * lsa[b] + (double)rc * rsa[b + 1]) / total_sa;
        if (!have || sah < s_cost) {
          have = 1;
          s_axis = axis;
          s_bin = b;
          s_cost = sah;
        }
      }
    }
    free(bins);
    free(lsa);
    free(rsa);
    free(lcn);
    free(rcn);
  }
  uint32_t axis = have ? (uint32_t)s_axis : 0;
  uint32_t *left = (uint32_t *)malloc(4 * (size_t)n), *right = (uint32_t *)malloc(4 * (size_t)n);
  uint32_t nl = 0, nr = 0;
  int reason = -1;
  if (!have) {
    reason = 0;
  } else if (s_cost >= (double)n) {
    reason = 1;
  } else {
    /* partitionByBin (tlas.ts:395-414). */
    double e = cmax[s_axis] - cmin[s_axis];
    for (uint32_t i = 0; i < n; ++i) {
      double tb = floor((R[recs[i]].c[s_axis] - cmin[s_axis]) / e * (double)K);
      double cl = tb < (double)(K - 1) ? tb : (double)(K - 1);
      if (cl < 0) cl = 0;
      if ((uint32_t)cl <= s_bin)
        left[nl++] = recs[i];
      else
        right[nr++] = recs[i];
    }
    if (nl == 0 || nr == 0)
      reason = 2;
    else if (!split_fits_depth(nl, nr, depth + 1, x->max_depth, x->max_leaf))
      reason = 3;
  }
  if (reason >= 0) {
    /* balancedPartition (tlas.ts:429-451). */
    axis = 0;
    if (ext[1] > ext[axis]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    uint32_t *ordered = (uint32_t *)malloc(4 * (size_t)n);
    memcpy(ordered, recs, 4 * (size_t)n);
    g_sort_axis = (int)axis;
    g_inst_sort = R;
    qsort(ordered, n, 4, cmp_inst_median);
    nl = n / 2;
    nr = n - nl;
    memcpy(left, ordered, 4 * (size_t)nl);
    memcpy(right, ordered + nl, 4 * (size_t)nr);
    free(ordered);
    x->fallbacks[reason] += 1;
  }
  if (nl == 0 || nr == 0 || !split_fits_depth(nl, nr, depth + 1, x->max_depth, x->max_leaf)) {
    free(left);
    free(right);
    x->failed = orc_fail("Internal balanced fallback failed to make a non-empty partition.");
    return this_idx;
  }
  tnode *node = &x->nodes[x->node_count++];
  memset(node, 0, sizeof *node);
  node->split = axis;
  tlas_build_rec(x, R, left, nl, depth + 1);
  uint32_t ridx = x->failed ? 0 : tlas_build_rec(x, R, right, nr, depth + 1);
  free(left);
  free(right);
  if (x->failed) return this_idx;
  node = &x->nodes[this_idx];
  node->right = ridx - this_idx;
  const tnode *ln = &x->nodes[this_idx + 1], *rn = &x->nodes[ridx];
  for (int k = 0; k < 3; ++k) {
    node->mn[k] = js_min(ln->mn[k], rn->mn[k]);
    node->mx[k] = js_max(ln->mx[k], rn->mx[k]);
  }
  return this_idx;
}

/* validateTlasBuild (tlas.ts:630-850). */
static int validate_tlas(const uint32_t *nodes, uint32_t node_count, const uint32_t *perm,
                         uint32_t perm_count, uint32_t ninst, const float *xf,
                         uint32_t max_leaf, uint32_t max_depth) {
  if (node_count == 0) return orc_fail("validateTlasBuild: nodeCount must be positive.");
  uint8_t *vn = (uint8_t *)calloc(node_count, 1), *vp = (uint8_t *)calloc(perm_count + 1, 1),
          *vi = (uint8_t *)calloc(ninst + 1, 1);
  uint32_t *sn = (uint32_t *)malloc(4 * (2 * (size_t)node_count + 2)),
           *sd = (uint32_t *)malloc(4 * (2 * (size_t)node_count + 2));
  size_t sp = 0;
  sn[sp] = 0;
  sd[sp] = 0;
  ++sp;
  uint32_t visited = 0, max_stack = 1;
  int rc = 0;
  while (sp > 0 && !rc) {
    --sp;
    uint32_t ni = sn[sp], depth = sd[sp];
    if (ni >= node_count) {
      rc = orc_fail("validateTlasBuild: node reference %u is out of range.", ni);
      break;
    }
    if (vn[ni]) {
      rc = orc_fail("validateTlasBuild: node %u is referenced more than once.", ni);
      break;
    }
    if (depth > max_depth) {
      rc = orc_fail("validateTlasBuild: node %u depth %u exceeds maximum %u.", ni, depth,
                    max_depth);
      break;
    }
    vn[ni] = 1;
    ++visited;
    if (depth + 1 > max_stack) max_stack = depth + 1;
    if (max_stack > TLAS_STACK_DEPTH) {
      rc = orc_fail("validateTlasBuild: traversal requires %u stack entries.", max_stack);
      break;
    }
    const uint32_t *n = nodes + 8 * (size_t)ni;
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k) mn[k] = u2f(n[k]), mx[k] = u2f(n[3 + k]);
    int bad = 0;
    for (int k = 0; k < 3; ++k) bad |= !isfinite(mn[k]) || !isfinite(mx[k]) || mn[k] > mx[k];
    if (bad) {
      rc = orc_fail("validateTlasBuild: node %u has invalid bounds.", ni);
      break;
    }
    uint32_t soc = n[7];
    if ((soc >> 16) == 0xFFFFu) {
      uint32_t count = soc & 0xFFFFu, off = n[6];
      if (count == 0 || count > max_leaf || off > perm_count || count > perm_count - off) {
        rc = orc_fail("validateTlasBuild: leaf %u has invalid range.", ni);
        break;
      }
      for (uint32_t i = 0; i < count && !rc; ++i) {
        uint32_t pi = off + i;
        if (vp[pi]) {
          rc = orc_fail("validateTlasBuild: permutation slot %u referenced twice.", pi);
          break;
        }
        vp[pi] = 1;
        uint32_t ii = perm[pi];
        if (ii >= ninst) {
          rc = orc_fail("validateTlasBuild: instance reference %u is out of range.", ii);
          break;
        }
        if (vi[ii]) {
          rc = orc_fail("validateTlasBuild: original instance %u referenced twice.", ii);
          break;
        }
        vi[ii] = 1;
        for (int l = 0; l < 16; ++l)
          if (!isfinite(xf[16 * (size_t)ii + l])) {
            rc = orc_fail("validateTlasBuild: instance %u has a non-finite transform.", ii);
            break;
          }
      }
      continue;
    }
    if (depth >= max_depth) {
      rc = orc_fail("validateTlasBuild: interior node %u reaches maximum depth.", ni);
      break;
    }
    if (soc > 2) {
      rc = orc_fail("validateTlasBuild: interior node %u has invalid split axis.", ni);
      break;
    }
    uint32_t ro = n[6];
    uint64_t lc = (uint64_t)ni + 1, rch = (uint64_t)ni + ro;
    if (ro <= 1 || lc >= node_count || rch >= node_count) {
      rc = orc_fail("validateTlasBuild: interior node %u has invalid child references.", ni);
      break;
    }
    uint64_t ch[2] = {lc, rch};
    for (int c = 0; c < 2; ++c) {
      const uint32_t *q = nodes + 8 * ch[c];
      if (u2f(q[0]) < mn[0] || u2f(q[1]) < mn[1] || u2f(q[2]) < mn[2] || u2f(q[3]) > mx[0] ||
          u2f(q[4]) > mx[1] || u2f(q[5]) > mx[2]) {
        rc = orc_fail("validateTlasBuild: interior node %u bounds do not enclose child.", ni);
        break;
      }
    }
    if (rc) break;
    sn[sp] = (uint32_t)rch;
    sd[sp] = depth + 1;
    ++sp;
    sn[sp] = (uint32_t)lc;
    sd[sp] = depth + 1;
    ++sp;
  }
  if (!rc && visited != node_count)
    rc = orc_fail("validateTlasBuild: %u node(s) are unreachable from root.", node_count - visited);
  for (uint32_t i = 0; !rc && i < perm_count; ++i)
    if (!vp[i]) rc = orc_fail("validateTlasBuild: permutation slot %u is not referenced.", i);
  free(vn);
  free(vp);
  free(vi);
  free(sn);
  free(sd);
  return rc;
}

int orc_build_tlas(uint32_t n, const uint32_t *blas_ids, const double *aabb_min,
                   const double *aabb_max, const float *w2l, uint32_t max_leaf,
                   uint32_t num_bins, uint32_t max_depth, uint32_t *out_nodes,
                   uint32_t *out_node_count, uint32_t *out_instance_indices,
                   uint32_t *out_blas_roots, float *out_transforms, uint32_t *out_fallbacks) {
  g_err[0] = 0;
  if (n == 0) return orc_fail("buildTlas: instances list is empty.");
  if (max_leaf < 1 || max_leaf > 0xFFFF) return orc_fail("buildTlas: maxLeafInstances range");
  if (num_bins < 2 || num_bins > 256) return orc_fail("buildTlas: numBins range");
  if (max_depth > MAX_BUILD_DEPTH) return orc_fail("buildTlas: maxDepth range");
  inst_rec *R = (inst_rec *)malloc(sizeof(inst_rec) * n);
  for (uint32_t i = 0; i < n; ++i) {
    const double *mn = aabb_min + 3 * (size_t)i, *mx = aabb_max + 3 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      if (!isfinite(mn[k]) || !isfinite(mx[k]) || !isfinite((float)mn[k]) ||
          !isfinite((float)mx[k])) {
        free(R);
        return orc_fail("buildTlas: instance %u aabb must be finite and representable as "
                        "float32.",
                        i);
      }
    }
    if (mx[0] < mn[0] || mx[1] < mn[1] || mx[2] < mn[2]) {
      free(R);
      return orc_fail("buildTlas: instance %u has inverted AABB.", i);
    }
    for (int l = 0; l < 16; ++l)
      if (!isfinite(w2l[16 * (size_t)i + l])) {
        free(R);
        return orc_fail("buildTlas: instance %u has non-finite AABB or worldToLocal data.", i);
      }
    R[i].orig = i;
    for (int k = 0; k < 3; ++k) {
      R[i].mn[k] = mn[k];
      R[i].mx[k] = mx[k];
      R[i].c[k] = (mn[k] + mx[k]) * 0.5;
    }
  }
  if (min_balanced_depth(n, max_leaf) > max_depth) {
    free(R);
    return orc_fail("buildTlas: %u valid instances cannot fit maxLeafInstances=%u within depth "
                    "%u.",
                    n, max_leaf, max_depth);
  }
  tctx x;
  memset(&x, 0, sizeof x);
  x.nodes = (tnode *)malloc(sizeof(tnode) * 2 * (size_t)n);
  x.perm = (uint32_t *)malloc(4 * (size_t)n);
  x.max_leaf = max_leaf;
  x.num_bins = num_bins;
  x.max_depth = max_depth;
  uint32_t *all = (uint32_t *)malloc(4 * (size_t)n);
  for (uint32_t i = 0; i < n; ++i) all[i] = i;
  tlas_build_rec(&x, R, all, n, 0);
  free(all);
  int rc = x.failed ? -1 : 0;
  if (!rc) {
    for (uint32_t i = 0; i < x.node_count; ++i) {
      uint32_t *w = out_nodes + 8 * (size_t)i;
      for (int k = 0; k < 3; ++k) {
        w[k] = f2u((float)x.nodes[i].mn[k]);
        w[3 + k] = f2u((float)x.nodes[i].mx[k]);
      }
      w[6] = x.nodes[i].right;
      w[7] = x.nodes[i].split;
    }
    *out_node_count = x.node_count;
    memcpy(out_instance_indices, x.perm, 4 * (size_t)x.perm_count);
    for (uint32_t i = 0; i < n; ++i) out_blas_roots[i] = blas_ids[i];
    memcpy(out_transforms, w2l, 64 * (size_t)n);
    if (out_fallbacks) memcpy(out_fallbacks, x.fallbacks, sizeof x.fallbacks);
    rc = validate_tlas(out_nodes, x.node_count, x.perm, x.perm_count, n, w2l, max_leaf,
                       max_depth);
  }
  free(R);
  free(x.nodes);
  free(x.perm);
  return rc;
}

/* ======================================================================== */
/* mathUtils.ts:21 invertMat4                                                */
/* ======================================================================== */

int orc_invert_mat4(const float *mf, float *out) {
  double m[16];
  for (int i = 0; i < 16; ++i) {
    m[i] = mf[i];
    if (!isfinite(m[i])) return 0;
  }
  double a00 = m[0], a01 = m[1], a02 = m[2], a03 = m[3];
  double a10 = m[4], a11 = m[5], a12 = m[6], a13 = m[7];
  double a20 = m[8], a21 = m[9], a22 = m[10], a23 = m[11];
  double a30 = m[12], a31 = m[13], a32 = m[14], a33 = m[15];
  double b00 = a00 * a11 - a01 * a10;
  double b01 = a00 * a12 - a02 * a10;
  double b02 = a00 * a13 - a03 * a10;
  double b03 = a01 * a12 - a02 * a11;
  double b04 = a01 * a13 - a03 * a11;
  double b05 = a02 * a13 - a03 * a12;
  double b06 = a20 * a31 - a21 * a30;
  double b07 = a20 * a32 - a22 * a30;
  double b08 = a20 * a33 - a23 * a30;
  double b09 = a21 * a32 - a22 * a31;
  double b10 = a21 * a33 - a23 * a31;
  double b11 = a22 * a33 - a23 * a32;
  double det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
  if (!isfinite(det) || det == 0.0) return 0;
  double inv = 1.0 / det;
  float o[16];
  o[0] = (float)((a11 * b11 - a12 * b10 + a13 * b09) * inv);
  o[1] = (float)((-a01 * b11 + a02 * b10 - a03 * b09) * inv);
  o[2] = (float)((a31 * b05 - a32 * b04 + a33 * b03) * inv);
  o[3] = (float)((-a21 * b05 + a22 * b04 - a23 * b03) * inv);
  o[4] = (float)((-a10 * b11 + a12 * b08 - a13 * b07) * inv);
  o[5] = (float)((a00 * b11 - a02 * b08 + a03 * b07) * inv);
  o[6] = (float)((-a30 * b05 + a32 * b02 - a33 * b01) * inv);
  o[7] = (float)((a20 * b05 - a22 * b02 + a23 * b01) * inv);
  o[8] = (float)((a10 * b10 - a11 * b08 + a13 * b06) * inv);
  o[9] = (float)((-a00 * b10 + a01 * b08 - a03 * b06) * inv);
  o[10] = (float)((a30 * b04 - a31 * b02 + a33 * b00) * inv);
  o[11] = (float)((-a20 * b04 + a21 * b02 - a23 * b00) * inv);
  o[12] = (float)((-a10 * b09 + a11 * b07 - a12 * b06) * inv);
  o[13] = (float)((a00 * b09 - a01 * b07 + a02 * b06) * inv);
  o[14] = (float)((-a30 * b03 + a31 * b01 - a32 * b00) * inv);
  o[15] = (float)((a20 * b03 - a21 * b01 + a22 * b00) * inv);
  for (int i = 0; i < 16; ++i)
    if (!isfinite(o[i])) return 0;
  /* Reciprocal check in both orders (mathUtils.ts:83-103). */
  for (int pass = 0; pass < 2; ++pass) {
    for (int col = 0; col < 4; ++col)
      for (int row = 0; row < 4; ++row) {
        double prod = 0, abs_sum = 0;
        for (int in = 0; in < 4; ++in) {
          double l = pass == 0 ? m[in * 4 + row] : (double)o[in * 4 + row];
          double r = pass == 0 ? (double)o[col * 4 + in] : m[col * 4 + in];
          double term = l * r;
          prod += term;
          abs_sum += fabs(term);
        }
        double expected = row == col ? 1.0 : 0.0;
        double tol = 1e-5 * (abs_sum > 1.0 ? abs_sum : 1.0);
        if (!isfinite(prod) || fabs(prod - expected) > tol) return 0;
      }
  }
  memcpy(out, o, sizeof o);
  return 1;
}

/* ======================================================================== */
/* scenePack.ts:1104 packSceneFromCore                                       */
/* ======================================================================== */

typedef struct {
  void *data;
  uint64_t words, cap;
} vec32;

static void v_push(vec32 *v, const void *src, uint64_t words) {
  if (v->words + words > v->cap) {
    uint64_t nc = v->cap ? v->cap * 2 : 1024;
    while (nc < v->words + words) nc *= 2;
    v->data = realloc(v->data, nc * 4);
    v->cap = nc;
  }
  memcpy((uint32_t *)v->data + v->words, src, words * 4);
  v->words += words;
}
static void v_push_u32(vec32 *v, uint32_t x) { v_push(v, &x, 1); }

struct orc_pack {
  vec32 f[ORC_FIELD_COUNT];
  uint32_t warnings;
};

/* transformPoint (scenePack.ts:161-170) in binary64. */
static void xform_point(const float *m, const double *p, double *out) {
  double x = p[0], y = p[1], z = p[2];
  double tx = (double)m[0] * x + (double)m[4] * y + (double)m[8] * z + (double)m[12];
  double ty = (double)m[1] * x + (double)m[5] * y + (double)m[9] * z + (double)m[13];
  double tz = (double)m[2] * x + (double)m[6] * y + (double)m[10] * z + (double)m[14];
  double tw = (double)m[3] * x + (double)m[7] * y + (double)m[11] * z + (double)m[15];
  if (fabs(tw) > 1e-8) {
    out[0] = tx / tw;
    out[1] = ty / tw;
    out[2] = tz / tw;
  } else {
    out[0] = tx;
    out[1] = ty;
    out[2] = tz;
  }
}

static const float k_identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

int orc_pack_scene(const orc_primitive *prims, uint32_t prim_count, orc_pack **out) {
  g_err[0] = 0;
  orc_pack *p = (orc_pack *)calloc(1, sizeof(orc_pack));
  vec32 inst_mn = {0}, inst_mx = {0}, inst_w2l = {0}, inst_l2w = {0}, inst_root = {0};
  vec32 roots = {0};
  int rc = 0;
  for (uint32_t pi = 0; pi < prim_count && !rc; ++pi) {
    const orc_primitive *pr = &prims[pi];
    /* packOneMeshLikePrimitive (scenePack.ts:338-503). */
    uint32_t vcount = pr->position_floats / 3;
    if (vcount < 3) {
      p->warnings++;
      continue;
    }
    uint32_t idx_len = pr->indices ? pr->index_count : vcount;
    uint32_t tcount = idx_len / 3;
    if (tcount == 0) {
      p->warnings++;
      continue;
    }
    float *lpos = (float *)calloc((size_t)vcount * 4, 4);
    float *lnrm = (float *)calloc((size_t)vcount * 4, 4);
    for (uint32_t i = 0; i < vcount; ++i) {
      for (int k = 0; k < 3; ++k) lpos[i * 4 + k] = pr->positions[i * 3 + k];
      for (int k = 0; k < 3; ++k) {
        uint32_t o = i * 3 + k;
        lnrm[i * 4 + k] = o < pr->normal_floats ? pr->normals[o] : (k == 1 ? 1.0f : 0.0f);
      }
    }
    uint32_t *lidx = (uint32_t *)calloc((size_t)tcount * 4, 4);
    for (uint32_t t = 0; t < tcount; ++t)
      for (int k = 0; k < 3; ++k) lidx[t * 4 + k] = pr->indices ? pr->indices[t * 3 + k] : t * 3 + k;
    uint32_t *lmat = (uint32_t *)malloc(4 * (size_t)tcount);
    for (uint32_t t = 0; t < tcount; ++t) lmat[t] = pr->material_id;
    uint32_t *nodes = (uint32_t *)malloc(32 * (2 * (size_t)tcount));
    uint32_t node_count = 0;
    uint32_t *ridx = (uint32_t *)malloc(16 * (size_t)tcount);
    uint32_t *rmat = (uint32_t *)malloc(4 * (size_t)tcount);
    uint32_t *rsrc = (uint32_t *)malloc(4 * (size_t)tcount);
    rc = orc_build_blas(lpos, vcount * 4, 4, lidx, tcount * 4, 4, lmat, 4, 16, MAX_BUILD_DEPTH,
                        nodes, &node_count, ridx, rmat, rsrc);
    double lmin[3] = {INFINITY, INFINITY, INFINITY}, lmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < vcount; ++i) /* computeLocalAabb (scenePack.ts:172-195) */
      for (int k = 0; k < 3; ++k) {
        lmin[k] = js_min(lmin[k], lpos[i * 4 + k]);
        lmax[k] = js_max(lmax[k], lpos[i * 4 + k]);
      }
    int zero_instances = pr->kind == 1 && pr->transform_count == 0;
    uint32_t ntx = pr->kind == 1 ? pr->transform_count : 1;
    uint32_t vertex_base = (uint32_t)(p->f[ORC_POSITIONS].words / 4);
    uint32_t tri_base = (uint32_t)p->f[ORC_TRI_MATERIAL].words;
    uint32_t node_base = (uint32_t)(p->f[ORC_BVH_NODES].words / 8);
    uint32_t survivors = 0;
    if (!rc && !zero_instances) {
      for (uint32_t s = 0; s < ntx; ++s) {
        const float *l2w =
            (pr->kind == 1 || (pr->transforms && pr->transform_count)) ? pr->transforms + 16 * (size_t)s
                                                                     : k_identity;
        float w2l[16];
        if (!orc_invert_mat4(l2w, w2l)) { /* resolveOneTransform (scenePack.ts:532) */
          p->warnings++;
          continue;
        }
        double wmn[3] = {INFINITY, INFINITY, INFINITY}, wmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = 0; c < 8; ++c) { /* transformAabb (scenePack.ts:197-223) */
          double corner[3] = {(c & 1) ? lmax[0] : lmin[0], (c & 2) ? lmax[1] : lmin[1],
                              (c & 4) ? lmax[2] : lmin[2]};
          double q[3];
          xform_point(l2w, corner, q);
          for (int k = 0; k < 3; ++k) {
            wmn[k] = js_min(wmn[k], q[k]);
            wmx[k] = js_max(wmx[k], q[k]);
          }
        }
        v_push(&inst_mn, wmn, 6);
        v_push(&inst_mx, wmx, 6);
        v_push(&inst_w2l, w2l, 16);
        v_push(&inst_l2w, l2w, 16);
        v_push_u32(&inst_root, node_base);
        v_push_u32(&p->f[ORC_INSTANCE_PRIMITIVE], pi);
        v_push_u32(&p->f[ORC_INSTANCE_SOURCE], s);
        ++survivors;
      }
      if (survivors == 0) p->warnings++;
    } else if (zero_instances) {
      p->warnings++;
    }
    if (!rc && survivors > 0) {
      v_push(&p->f[ORC_POSITIONS], lpos, (uint64_t)vcount * 4);
      v_push(&p->f[ORC_NORMALS], lnrm, (uint64_t)vcount * 4);
      for (uint32_t t = 0; t < tcount; ++t) {
        uint32_t q[4] = {ridx[t * 4] + vertex_base, ridx[t * 4 + 1] + vertex_base,
                         ridx[t * 4 + 2] + vertex_base, ridx[t * 4 + 3]};
        v_push(&p->f[ORC_INDICES], q, 4);
        v_push_u32(&p->f[ORC_TRI_MATERIAL], rmat[t]);
        v_push_u32(&p->f[ORC_TRI_SOURCE], rsrc[t]);
        v_push_u32(&p->f[ORC_TRI_PRIMITIVE], pi);
      }
      for (uint32_t nd = 0; nd < node_count; ++nd) {
        uint32_t w[8];
        memcpy(w, nodes + 8 * (size_t)nd, 32);
        if ((w[7] >> 16) == 0xFFFFu) w[6] += tri_base;
        v_push(&p->f[ORC_BVH_NODES], w, 8);
      }
      uint32_t b[8] = {node_base, survivors, vertex_base, vcount, tri_base, tcount, pi, pr->kind};
      v_push(&p->f[ORC_BINDINGS], b, 8);
      float ab[6] = {(float)lmin[0], (float)lmin[1], (float)lmin[2],
                     (float)lmax[0], (float)lmax[1], (float)lmax[2]};
      v_push(&p->f[ORC_BINDING_AABB], ab, 6);
      v_push_u32(&roots, node_base);
    }
    free(lpos);
    free(lnrm);
    free(lidx);
    free(lmat);
    free(nodes);
    free(ridx);
    free(rmat);
    free(rsrc);
  }
  uint32_t ninst = (uint32_t)inst_root.words;
  if (!rc && ninst > 0) {
    uint32_t *tn = (uint32_t *)malloc(32 * (2 * (size_t)ninst));
    uint32_t tcount = 0;
    uint32_t *perm = (uint32_t *)malloc(4 * (size_t)ninst);
    uint32_t *broots = (uint32_t *)malloc(4 * (size_t)ninst);
    float *xf = (float *)malloc(64 * (size_t)ninst);
    rc = orc_build_tlas(ninst, (const uint32_t *)inst_root.data, (const double *)inst_mn.data,
                        (const double *)inst_mx.data, (const float *)inst_w2l.data, 1, 16,
                        MAX_BUILD_DEPTH, tn, &tcount, perm, broots, xf, NULL);
    if (!rc) {
      v_push(&p->f[ORC_TLAS_NODES], tn, (uint64_t)tcount * 8);
      v_push(&p->f[ORC_TLAS_INSTANCE_INDICES], perm, ninst);
      v_push(&p->f[ORC_TLAS_BLAS_ROOTS], broots, ninst);
      v_push(&p->f[ORC_TLAS_W2L], xf, (uint64_t)ninst * 16);
      v_push(&p->f[ORC_TLAS_L2W], inst_l2w.data, (uint64_t)ninst * 16);
    }
    free(tn);
    free(perm);
    free(broots);
    free(xf);
  }
  if (!rc && roots.words > 0) {
    rc = orc_validate_bvh((const uint32_t *)p->f[ORC_BVH_NODES].data,
                          (uint32_t)(p->f[ORC_BVH_NODES].words / 8), (const uint32_t *)roots.data,
                          (int32_t)roots.words, (int64_t)p->f[ORC_TRI_MATERIAL].words,
                          BVH_STACK_DEPTH, NULL);
  }
  free(inst_mn.data);
  free(inst_mx.data);
  free(inst_w2l.data);
  free(inst_l2w.data);
  free(inst_root.data);
  free(roots.data);
  if (rc) {
    orc_pack_free(p);
    *out = NULL;
    return rc;
  }
  *out = p;
  return 0;
}

int64_t orc_pack_field_size(const orc_pack *p, int field) {
  if (!p || field < 0 || field >= ORC_FIELD_COUNT) return -1;
  return (int64_t)p->f[field].words;
}

int orc_pack_copy_field(const orc_pack *p, int field, void *dst, uint64_t dst_words) {
  if (!p || field < 0 || field >= ORC_FIELD_COUNT) return orc_fail("bad field");
  if (dst_words < p->f[field].words) return orc_fail("destination too small");
  if (p->f[field].words) memcpy(dst, p->f[field].data, p->f[field].words * 4);
  return 0;
}

uint32_t orc_pack_warning_count(const orc_pack *p) { return p ? p->warnings : 0; }

void orc_pack_free(orc_pack *p) {
  if (!p) return;
  for (int i = 0; i < ORC_FIELD_COUNT; ++i) free(p->f[i].data);
  free(p);
}
