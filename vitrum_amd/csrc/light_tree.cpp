// light_tree.cpp -- host build of the NEE light tree.
//
// The tree and its packed form follow packages/shared-samplers/src/
// lightTree.ts (buildSubtree :350-448, mergeCones :250-307, packing
// :611-737): median split on the longest centroid axis (a power-sorted median
// when every centroid coincides), preorder nodes with the left subtree first,
// node power = sum of its leaves, bounds = union of its leaves' bounds, cones
// merged bottom-up.  Packed powers are normalised by the largest node power
// and floored at the smallest normal f32; cone cosines are rounded outward by
// 64 f32 steps and 2^-18 so the device-side f32 cone test never culls a
// supported direction.  Builder decisions are binary64, as the reference's
// JavaScript numbers.
#include "light_tree.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace vtl {
namespace {

constexpr double kPi = 3.14159265358979323846;

struct Cone {
  double axis[3] = {0, 0, 0};
  double theta_o = kPi, theta_e = kPi;  // full sphere
};

struct Node {
  int emitter = -1;
  double power = 0;
  double lo[3], hi[3];
  int left = -1, right = -1;
  Cone cone;
};

// Math.hypot of three finite values: max-scaled, compensated sum of squares.
double hypot3(double a, double b, double c) {
  const double v[3] = {std::fabs(a), std::fabs(b), std::fabs(c)};
  const double m = std::max(v[0], std::max(v[1], v[2]));
  if (std::isinf(m)) return m;
  if (m == 0) return 0;
  double sum = 0, comp = 0;
  for (double x : v) {
    const double n = x / m, s = n * n - comp, next = sum + s;
    comp = (next - sum) - s;
    sum = next;
  }
  return std::sqrt(sum) * m;
}

double vec_len(const double* v) {
  const double s = std::max(std::fabs(v[0]), std::max(std::fabs(v[1]), std::fabs(v[2])));
  return s == 0 ? 0 : s * hypot3(v[0] / s, v[1] / s, v[2] / s);
}

Cone full_sphere(double theta_e) {
  Cone c;
  c.theta_e = theta_e;
  return c;
}

// Smallest cone bounding two cones: the wider one's axis rotated towards the
// other in their common plane, or the full sphere when no stable plane exists.
Cone merge(const Cone& a, const Cone& b) {
  const double la = vec_len(a.axis), lb = vec_len(b.axis);
  if (la == 0 || lb == 0 || a.theta_o >= kPi || b.theta_o >= kPi)
    return full_sphere(std::min(kPi, std::max(a.theta_e, b.theta_e)));
  const Cone& w = b.theta_o > a.theta_o ? b : a;  // wider
  const Cone& n = b.theta_o > a.theta_o ? a : b;
  const double lw = vec_len(w.axis), ln = vec_len(n.axis);
  double wa[3], na[3];
  for (int k = 0; k < 3; ++k) wa[k] = w.axis[k] / lw, na[k] = n.axis[k] / ln;
  const double cos_d = std::max(-1.0, std::min(1.0, wa[0] * na[0] + wa[1] * na[1] + wa[2] * na[2]));
  const double d_theta = std::acos(cos_d);
  const double theta_e = std::min(kPi, std::max(w.theta_e, n.theta_e));
  if (std::min(d_theta + n.theta_o, kPi) <= w.theta_o) {
    Cone c;
    std::memcpy(c.axis, wa, sizeof wa);
    c.theta_o = w.theta_o;
    c.theta_e = theta_e;
    return c;
  }
  const double theta_o = (w.theta_o + d_theta + n.theta_o) * 0.5;
  if (theta_o >= kPi) return full_sphere(theta_e);
  double t[3];
  for (int k = 0; k < 3; ++k) t[k] = na[k] - cos_d * wa[k];
  const double tl = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  if (tl < 1e-8) return full_sphere(theta_e);  // antiparallel: no stable plane
  for (double& x : t) x /= tl;
  const double rot = theta_o - w.theta_o, cr = std::cos(rot), sr = std::sin(rot);
  Cone c;
  for (int k = 0; k < 3; ++k) c.axis[k] = wa[k] * cr + t[k] * sr;
  c.theta_o = theta_o;
  c.theta_e = theta_e;
  return c;
}

struct Builder {
  const std::vector<Leaf>& leaves;
  std::vector<Cone> leaf_cone;
  std::vector<Node> nodes;

  int build(uint32_t* idx, uint32_t n) {
    const int at = (int)nodes.size();
    nodes.emplace_back();
    if (n == 1) {
      const Leaf& L = leaves[idx[0]];
      Node& d = nodes[at];
      d.emitter = (int)idx[0];
      d.power = L.power;
      for (int k = 0; k < 3; ++k) d.lo[k] = L.lo[k], d.hi[k] = L.hi[k];
      d.cone = leaf_cone[idx[0]];
      return at;
    }
    double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    Node d;
    d.power = 0;
    for (int k = 0; k < 3; ++k) d.lo[k] = INFINITY, d.hi[k] = -INFINITY;
    for (uint32_t i = 0; i < n; ++i) {  // in the order received (the sum's order)
      const Leaf& L = leaves[idx[i]];
      d.power += L.power;
      for (int k = 0; k < 3; ++k) {
        clo[k] = std::min(clo[k], L.centroid[k]);
        chi[k] = std::max(chi[k], L.centroid[k]);
        d.lo[k] = std::min(d.lo[k], L.lo[k]);
        d.hi[k] = std::max(d.hi[k], L.hi[k]);
      }
    }
    if (!std::isfinite(d.power)) throw std::range_error("light tree power overflowed");
    const double sx = chi[0] - clo[0], sy = chi[1] - clo[1], sz = chi[2] - clo[2];
    if (std::max(sx, std::max(sy, sz)) < 1e-4) {
      std::sort(idx, idx + n, [&](uint32_t a, uint32_t b) {
        if (leaves[a].power != leaves[b].power) return leaves[a].power > leaves[b].power;
        return a < b;
      });
    } else {
      const int axis = (sx >= sy && sx >= sz) ? 0 : (sy >= sz ? 1 : 2);
      std::sort(idx, idx + n, [&](uint32_t a, uint32_t b) {
        if (leaves[a].centroid[axis] != leaves[b].centroid[axis])
          return leaves[a].centroid[axis] < leaves[b].centroid[axis];
        return a < b;
      });
    }
    nodes[at] = d;
    const uint32_t mid = n / 2;
    const int l = build(idx, mid);
    const int r = build(idx + mid, n - mid);
    nodes[at].left = l;
    nodes[at].right = r;
    nodes[at].cone = merge(nodes[l].cone, nodes[r].cone);
    return at;
  }
};

float step_down(float v) {
  if (v == 0) return -0x1p-149f;
  uint32_t b;
  std::memcpy(&b, &v, 4);
  b = v > 0 ? b - 1 : b + 1;
  std::memcpy(&v, &b, 4);
  return v;
}

// Cosine threshold of a cone angle, rounded outward (lightTree.ts:193-212).
float outward_cos(double angle) {
  const float c = (float)std::cos(std::min(kPi, angle));
  float t = c;
  for (int i = 0; i < 64; ++i) t = step_down(t);
  const float pad = (float)((double)c - 0x1p-18);
  return std::max(-1.0f, std::min(t, pad));
}

}  // namespace

std::vector<float> build_packed(const std::vector<Leaf>& leaves) {
  const uint32_t n = (uint32_t)leaves.size();
  if (n == 0) return {};
  Builder B{leaves, {}, {}};
  B.leaf_cone.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    const Leaf& L = leaves[i];
    if (!(L.power >= 0)) throw std::range_error("light power must be finite and non-negative");
    const double l = L.oriented ? vec_len(L.axis) : 0;
    if (l > 0) {
      for (int k = 0; k < 3; ++k) B.leaf_cone[i].axis[k] = L.axis[k] / l;
      B.leaf_cone[i].theta_o = L.theta_o;
      B.leaf_cone[i].theta_e = std::min(kPi, L.theta_e);
    }
  }
  std::vector<uint32_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0u);
  B.nodes.reserve(2 * (size_t)n);
  B.build(idx.data(), n);
  const std::vector<Node>& N = B.nodes;
  double max_power = 0;
  for (const Node& d : N) max_power = std::max(max_power, d.power);
  std::vector<float> out(N.size() * kNodeFloats, 0.0f);
  for (size_t i = 0; i < N.size(); ++i) {
    const Node& d = N[i];
    float* o = &out[i * kNodeFloats];
    o[0] = (float)d.emitter;
    float power = 0;
    if (d.power > 0 && max_power > 0) {
      const float r = (float)(d.power / max_power);
      power = r > 1.175494351e-38f ? r : 1.175494351e-38f;
    }
    o[1] = power;
    o[2] = (float)d.left;
    o[3] = (float)d.right;
    const double al = vec_len(d.cone.axis);
    for (int k = 0; k < 3; ++k) {
      o[4 + k] = (float)d.lo[k];
      o[7 + k] = (float)d.hi[k];
      o[10 + k] = al > 0 ? (float)(d.cone.axis[k] / al) : 0.0f;
    }
    o[13] = outward_cos(d.cone.theta_o);
    o[14] = outward_cos(d.cone.theta_o + d.cone.theta_e);
  }
  for (size_t i = N.size(); i-- > 0;) {  // subtree leaf counts, children later
    const Node& d = N[i];
    out[i * kNodeFloats + 15] =
        d.left < 0 ? 1.0f
                   : out[(size_t)d.left * kNodeFloats + 15] + out[(size_t)d.right * kNodeFloats + 15];
  }
  return out;
}

double js_hypot3(double a, double b, double c) { return hypot3(a, b, c); }

float area_f32(const float u[3], const float v[3], float coefficient) {
  // classifyAreaVectorF32 (areaEmitterGeometry.ts:51-128): equilibrated edge
  // cross product, one rounding per operation.
  float area = 0;
  const float es = std::max({std::fabs(u[0]), std::fabs(u[1]), std::fabs(u[2]), std::fabs(v[0]),
                             std::fabs(v[1]), std::fabs(v[2])});
  if (es > 0 && std::isfinite(es)) {
    float su[3], sv[3];
    for (int k = 0; k < 3; ++k) su[k] = u[k] / es, sv[k] = v[k] / es;
    const float cx = su[1] * sv[2] - su[2] * sv[1];
    const float cy = su[2] * sv[0] - su[0] * sv[2];
    const float cz = su[0] * sv[1] - su[1] * sv[0];
    const float cs = std::max({std::fabs(cx), std::fabs(cy), std::fabs(cz)});
    if (cs > 0 && std::isfinite(cs)) {
      const float dl = (float)hypot3(cx / cs, cy / cs, cz / cs);
      if (dl > 0 && std::isfinite(dl)) {
        const float scaled = cs * dl;
        const float weighted = coefficient * scaled;
        const float once = weighted * es;
        const float ar = once * es;
        const float inv = 1.0f / ar;
        if (ar > 0 && std::isfinite(ar) && inv > 0 && std::isfinite(inv)) area = ar;
      }
    }
  }
  return area;
}

double power_of(const float le[3], double area) {
  // emitterPower (emitterPacking.ts:900-934): f32 luminance of the f32
  // radiance, times the f32 area for a finite emitter.
  const double lum_d = 0.2126 * (double)le[0] + 0.7152 * (double)le[1] + 0.0722 * (double)le[2];
  const float lum = (float)std::max(0.0, lum_d);
  if (area < 0) return lum;  // delta light
  return (float)((double)lum * area);
}

Leaf triangle_leaf(const float a[3], const float b[3], const float c[3], const float le[3], bool two_sided) {
  Leaf L{};
  // f32 triangle area (areaEmitterGeometry.ts:141-173)
  float u[3], v[3];
  for (int k = 0; k < 3; ++k) u[k] = b[k] - a[k], v[k] = c[k] - a[k];
  const float area = area_f32(u, v, 0.5f);
  L.power = power_of(le, area);
  for (int k = 0; k < 3; ++k) {
    L.centroid[k] = ((double)a[k] + (double)b[k] + (double)c[k]) / 3;
    L.lo[k] = std::min({(double)a[k], (double)b[k], (double)c[k]});
    L.hi[k] = std::max({(double)a[k], (double)b[k], (double)c[k]});
  }
  L.oriented = !two_sided;
  if (L.oriented) {
    // norm3(cross(b - a, c - a)) in binary64 (emitterPacking.ts:1062-1070)
    const double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
    const double e2[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
    const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                         e1[0] * e2[1] - e1[1] * e2[0]};
    const double l = hypot3(n[0], n[1], n[2]);
    for (int k = 0; k < 3; ++k) L.axis[k] = l > 1e-12 ? n[k] / l : 0.0;
    L.theta_o = 0;
    L.theta_e = kPi / 2;
  }
  return L;
}

}  // namespace vtl
