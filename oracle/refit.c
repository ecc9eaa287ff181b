/*
 * refit.c -- BVH and TLAS refit, restated for the tests.  TEST INFRASTRUCTURE
 * ONLY (see vt_oracle.h): the checker of the product's device refit.
 *
 *   orc_refit_bvh_bounds      refitBvhBounds + preflightRefitInputs
 *                             (packages/shared-bvh/src/refitBvhBounds.ts:3-179)
 *   orc_refit_tlas_instances  refitTlasInstances + tlasRefitNodeIndices +
 *                             validateTlasRefitAabbs
 *                             (packages/shared-bvh/src/tlas.ts:1012-1209)
 *
 * Bounds are accumulated with the reference's strict comparisons in its
 * visiting order (`if (x < mn) mn = x`), which fixes the sign of a zero
 * coordinate; the TLAS interior union is Math.min / Math.max of the two
 * children's binary32 values.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "vt_oracle.h"

int orc_fail(const char *fmt, ...);

#define LEAF(split) (((split) >> 16) == 0xFFFFu)

static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

/* Math.min / Math.max of non-NaN binary32 values (signed zeros included). */
static float js_minf(float a, float b) {
  if (a < b) return a;
  if (b < a) return b;
  if (a == 0.0f) return (signbit(a) || signbit(b)) ? -0.0f : 0.0f;
  return a;
}
static float js_maxf(float a, float b) {
  if (a > b) return a;
  if (b > a) return b;
  if (a == 0.0f) return (signbit(a) && signbit(b)) ? -0.0f : 0.0f;
  return a;
}

/* preflightRefitInputs (refitBvhBounds.ts:3-63): an explicit depth-first walk
 * from node 0 that proves every node is reached exactly once and every leaf
 * names valid triangles with finite vertices.  -2: stride (RangeError);
 * -1: any other rejection (Error). */
static int refit_preflight(const uint32_t *w, uint32_t total, const uint32_t *indices,
                           uint32_t index_words, uint32_t index_stride, const float *positions,
                           uint32_t position_floats, uint32_t position_stride) {
  if (position_stride != 3 && position_stride != 4) {
    orc_fail("position stride must be 3 or 4.");
    return -2;
  }
  if (index_stride != 3 && index_stride != 4) {
    orc_fail("index stride must be 3 or 4.");
    return -2;
  }
  if (position_floats % position_stride || index_words % index_stride)
    return orc_fail("geometry buffer is not stride-aligned.");
  if (total == 0) return 0;
  const uint64_t tris = index_words / index_stride, verts = position_floats / position_stride;
  uint8_t *seen = (uint8_t *)calloc(total, 1);
  /* every push follows the first visit of an interior node */
  uint32_t *pending = (uint32_t *)malloc(sizeof(uint32_t) * (2 * (size_t)total + 1));
  size_t sp = 0;
  uint64_t visited = 0;
  int rc = 0;
  pending[sp++] = 0;
  while (sp > 0 && !rc) {
    const uint32_t node = pending[--sp];
    if (node >= total || seen[node]) {
      rc = orc_fail("invalid or repeated topology node %u.", node);
      break;
    }
    seen[node] = 1;
    ++visited;
    const uint32_t *q = w + 8 * (size_t)node;
    if (!LEAF(q[7])) {
      const uint64_t left = (uint64_t)node + 1, right = (uint64_t)node + q[6];
      if (q[7] > 2 || left >= total || right <= node || right >= total) {
        rc = orc_fail("interior node %u is corrupt.", node);
        break;
      }
      pending[sp++] = (uint32_t)right;
      pending[sp++] = (uint32_t)left;
      continue;
    }
    const uint64_t count = q[7] & 0xFFFFu, off = q[6];
    if (off > tris || count > tris - off) {
      rc = orc_fail("leaf %u triangle range is corrupt.", node);
      break;
    }
    for (uint64_t t = off; t < off + count && !rc; ++t)
      for (int lane = 0; lane < 3 && !rc; ++lane) {
        const uint32_t v = indices[t * index_stride + lane];
        if (v >= verts) {
          rc = orc_fail("triangle %llu has an invalid vertex.", (unsigned long long)t);
          break;
        }
        const float *p = positions + (size_t)v * position_stride;
        if (!isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2]))
          rc = orc_fail("vertex %u is non-finite.", v);
      }
  }
  if (!rc && visited != total) rc = orc_fail("topology contains unreachable nodes.");
  free(seen);
  free(pending);
  return rc;
}

int orc_refit_bvh_bounds(uint32_t *w, uint32_t total, const uint32_t *indices, uint32_t index_words,
                         uint32_t index_stride, const float *positions, uint32_t position_floats,
                         uint32_t position_stride) {
  int rc = refit_preflight(w, total, indices, index_words, index_stride, positions, position_floats,
                           position_stride);
  if (rc || total == 0) return rc;
  /* Post-order (refitBvhBounds.ts:94-114): a node is emitted after both of
   * its children, left subtree first. */
  uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * total);
  uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)total);
  size_t n = 0, sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const uint32_t entry = stack[--sp];
    const uint32_t node = entry & 0x7FFFFFFFu;
    if (entry & 0x80000000u) {
      order[n++] = node;
      continue;
    }
    const uint32_t *q = w + 8 * (size_t)node;
    if (LEAF(q[7])) {
      order[n++] = node;
      continue;
    }
    stack[sp++] = node | 0x80000000u;
    stack[sp++] = node + q[6];
    stack[sp++] = node + 1;
  }
  for (size_t i = 0; i < n; ++i) {
    uint32_t *q = w + 8 * (size_t)order[i];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (LEAF(q[7])) {
      const uint32_t count = q[7] & 0xFFFFu, off = q[6];
      if (count == 0) /* the canonical empty BVH keeps finite point bounds */
        for (int k = 0; k < 3; ++k) mn[k] = 0.0f, mx[k] = 0.0f;
      for (uint32_t t = 0; t < count; ++t)
        for (int lane = 0; lane < 3; ++lane) {
          const float *p = positions + (size_t)indices[(size_t)(off + t) * index_stride + lane] * position_stride;
          for (int k = 0; k < 3; ++k) {
            if (p[k] < mn[k]) mn[k] = p[k];
            if (p[k] > mx[k]) mx[k] = p[k];
          }
        }
    } else {
      const uint32_t *c[2] = {w + 8 * ((size_t)order[i] + 1), w + 8 * ((size_t)order[i] + q[6])};
      for (int s = 0; s < 2; ++s)
        for (int k = 0; k < 3; ++k) {
          const float lo = u2f(c[s][k]), hi = u2f(c[s][3 + k]);
          if (lo < mn[k]) mn[k] = lo;
          if (hi > mx[k]) mx[k] = hi;
        }
    }
    for (int k = 0; k < 3; ++k) q[k] = f2u(mn[k]), q[3 + k] = f2u(mx[k]);
  }
  free(order);
  free(stack);
  return 0;
}

int orc_refit_tlas_instances(uint32_t *w, uint32_t node_count, const uint32_t *perm, uint32_t perm_count,
                             uint32_t instance_count, const double *aabb_min, const double *aabb_max,
                             const uint32_t *changed_list, uint32_t changed_count, uint32_t *out_affected,
                             uint32_t *out_affected_count) {
  /* validateTlasRefitAabbs (tlas.ts:1012-1039): every referenced instance's
   * box is ordered and finite. */
  uint8_t *included = (uint8_t *)calloc(instance_count ? instance_count : 1, 1);
  int rc = 0;
  for (uint32_t k = 0; k < perm_count && !rc; ++k) {
    const uint32_t i = perm[k];
    if (i >= instance_count) {
      rc = orc_fail("refitTlas: instance reference %u is out of range.", i);
      break;
    }
    if (included[i]) continue;
    included[i] = 1;
    const double *a = aabb_min + 3 * (size_t)i, *b = aabb_max + 3 * (size_t)i;
    if (b[0] < a[0] || b[1] < a[1] || b[2] < a[2]) rc = orc_fail("refitTlas: instance %u has inverted AABB.", i);
    else if (!isfinite(a[0]) || !isfinite(a[1]) || !isfinite(a[2]) || !isfinite(b[0]) || !isfinite(b[1]) ||
             !isfinite(b[2]))
      rc = orc_fail("refitTlas: instance %u has non-finite AABB.", i);
  }
  free(included);
  if (rc) return rc;
  /* tlasRefitNodeIndices (tlas.ts:1046-1133): leaves holding a changed
   * instance and all their ancestors, children before parents. */
  *out_affected_count = 0;
  if (changed_count == 0) return 0;
  uint8_t *changed = (uint8_t *)calloc(instance_count ? instance_count : 1, 1);
  for (uint32_t k = 0; k < changed_count; ++k) {
    if (changed_list[k] >= instance_count) {
      free(changed);
      return orc_fail("tlasRefitNodeIndices: instance index %u is out of bounds.", changed_list[k]);
    }
    changed[changed_list[k]] = 1;
  }
  int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (node_count ? node_count : 1));
  uint8_t *direct = (uint8_t *)calloc(node_count ? node_count : 1, 1);
  uint8_t *referenced = (uint8_t *)calloc(instance_count ? instance_count : 1, 1);
  uint8_t *affected = (uint8_t *)calloc(node_count ? node_count : 1, 1);
  for (uint32_t i = 0; i < node_count; ++i) parent[i] = -1;
  for (uint32_t node = 0; node < node_count && !rc; ++node) {
    const uint32_t *q = w + 8 * (size_t)node;
    if (LEAF(q[7])) {
      const uint32_t off = q[6], count = q[7] & 0xFFFFu;
      if (count == 0 || (uint64_t)off + count > perm_count) {
        rc = orc_fail("tlasRefitNodeIndices: leaf %u has invalid instance span.", node);
        break;
      }
      for (uint32_t k = 0; k < count; ++k) {
        const uint32_t orig = perm[off + k];
        if (orig >= instance_count) {
          rc = orc_fail("tlasRefitNodeIndices: leaf %u references invalid instance %u.", node, orig);
          break;
        }
        referenced[orig] = 1;
        if (changed[orig]) direct[node] = 1;
      }
      continue;
    }
    const uint64_t left = (uint64_t)node + 1, right = (uint64_t)node + q[6];
    if (left >= node_count || right <= node || right >= node_count || parent[left] != -1 ||
        parent[right] != -1) {
      rc = orc_fail("tlasRefitNodeIndices: interior node %u has invalid children.", node);
      break;
    }
    parent[left] = (int32_t)node;
    parent[right] = (int32_t)node;
  }
  for (uint32_t i = 0; i < instance_count && !rc; ++i)
    if (changed[i] && !referenced[i]) rc = orc_fail("tlasRefitNodeIndices: changed instance %u has no TLAS leaf.", i);
  if (!rc) {
    for (uint32_t node = 0; node < node_count; ++node) {
      if (!direct[node]) continue;
      for (int32_t c = (int32_t)node; c >= 0; c = parent[c]) {
        if (affected[c]) break;
        affected[c] = 1;
      }
    }
    uint32_t n = 0;
    for (int64_t node = (int64_t)node_count - 1; node >= 0; --node)
      if (affected[node]) out_affected[n++] = (uint32_t)node;
    if (n == 0) rc = orc_fail("tlasRefitNodeIndices: changed instance has no TLAS leaf.");
    *out_affected_count = n;
    /* refitTlasInstances (tlas.ts:1141-1209): leaves take their instances'
     * binary64 boxes through strict comparisons, stored as binary32;
     * interior nodes take Math.min / Math.max of their children. */
    for (uint32_t k = 0; k < n && !rc; ++k) {
      const uint32_t i = out_affected[k];
      uint32_t *q = w + 8 * (size_t)i;
      if (LEAF(q[7])) {
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        const uint32_t off = q[6], count = q[7] & 0xFFFFu;
        for (uint32_t j = 0; j < count; ++j) {
          const uint32_t orig = perm[off + j];
          for (int a = 0; a < 3; ++a) {
            if (aabb_min[3 * (size_t)orig + a] < mn[a]) mn[a] = aabb_min[3 * (size_t)orig + a];
            if (aabb_max[3 * (size_t)orig + a] > mx[a]) mx[a] = aabb_max[3 * (size_t)orig + a];
          }
        }
        for (int a = 0; a < 3; ++a) q[a] = f2u((float)mn[a]), q[3 + a] = f2u((float)mx[a]);
      } else {
        const uint32_t *l = w + 8 * ((size_t)i + 1), *r = w + 8 * ((size_t)i + q[6]);
        for (int a = 0; a < 3; ++a) {
          q[a] = f2u(js_minf(u2f(l[a]), u2f(r[a])));
          q[3 + a] = f2u(js_maxf(u2f(l[3 + a]), u2f(r[3 + a])));
        }
      }
    }
  }
  free(changed);
  free(parent);
  free(direct);
  free(referenced);
  free(affected);
  return rc;
}
