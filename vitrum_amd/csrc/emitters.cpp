// emitters.cpp -- packing of the analytic emitters and their light-tree
// leaves, following packages/pt-webgpu/src/scene/emitterPacking.ts (record
// contents, validation) and flatEmitterWalk.ts (leaf geometry).  Host
// arithmetic is binary64 (JavaScript numbers) with values published to the
// device rounded to f32.
#include "emitters.h"

#include <algorithm>
#include <cmath>

#include "render.h"

namespace vte {
namespace {

constexpr double kPi = 3.14159265358979323846;

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// packRadianceRgbScaleF32 (shared-bvh/src/radianceFloat32.ts:90-124): f32
// colour times f32 intensity, rounded once; non-negative and finite.
bool radiance(const vt_emitter& e, float out[3]) {
  if (!finite3(e.color) || !std::isfinite(e.intensity) || e.intensity < 0) return false;
  for (int k = 0; k < 3; ++k) {
    if (e.color[k] < 0) return false;
    out[k] = (float)((double)e.color[k] * (double)e.intensity);
    if (!std::isfinite(out[k])) return false;
  }
  return true;
}

// normalizeDirectionF32 (areaEmitterGeometry.ts:181-198)
bool normalize_f32(const double v[3], float out[3]) {
  const double s = std::max({std::fabs(v[0]), std::fabs(v[1]), std::fabs(v[2])});
  if (!(s > 0) || !std::isfinite(s)) return false;
  const double sx = v[0] / s, sy = v[1] / s, sz = v[2] / s;
  const double l = vtl::js_hypot3(sx, sy, sz);
  if (!(l > 0) || !std::isfinite(l)) return false;
  out[0] = (float)(sx / l), out[1] = (float)(sy / l), out[2] = (float)(sz / l);
  return finite3(out);
}

void push(std::vector<float>& r, std::initializer_list<float> v) { r.insert(r.end(), v); }

}  // namespace

std::string pack(const vt_emitter* em, uint32_t n, Packed& out) {
  out = Packed();
  if (n && !em) return "emitters is NULL";
  for (uint32_t i = 0; i < n; ++i)
    if (em[i].kind > VT_EMITTER_DISC_AREA) return "emitter " + std::to_string(i) + ": unknown kind";
  const auto label = [](uint32_t i) { return "emitter " + std::to_string(i) + ": "; };
  std::vector<float>& r = out.records;
  float rad[3];
  // directional (emitterPacking.ts:607-648): the packed direction points at
  // the light (the authored one points the way the light travels)
  for (uint32_t i = 0; i < n; ++i) {
    const vt_emitter& e = em[i];
    if (e.kind != VT_EMITTER_DIRECTIONAL) continue;
    if (!radiance(e, rad) || !finite3(e.direction)) return label(i) + "colour x intensity and direction must be finite";
    const double x = e.direction[0], y = e.direction[1], z = e.direction[2];
    const double len = vtl::js_hypot3(x, y, z);
    const double t[3] = {len < 1e-8 ? 0.0 : -x / len, len < 1e-8 ? 1.0 : -y / len, len < 1e-8 ? 0.0 : -z / len};
    const float ad = std::isfinite(e.angular_diameter) && e.angular_diameter > 0 ? e.angular_diameter : 0.0f;
    push(r, {(float)t[0], (float)t[1], (float)t[2], ad, rad[0], rad[1], rad[2], e.cast_shadow ? 0.0f : 1.0f,
             0, 0, 0, 0, 0, 0, 0, 0});
    ++out.n_dir;
  }
  // point (:650-668)
  for (uint32_t i = 0; i < n; ++i) {
    const vt_emitter& e = em[i];
    if (e.kind != VT_EMITTER_POINT) continue;
    if (!radiance(e, rad) || !finite3(e.position)) return label(i) + "colour x intensity and position must be finite";
    const float dist = std::isfinite(e.distance) && e.distance > 0 ? e.distance : 0.0f;
    if (!std::isfinite(e.decay)) return label(i) + "decay must be finite";
    push(r, {e.position[0], e.position[1], e.position[2], e.cast_shadow ? 0.0f : 1.0f, rad[0], rad[1], rad[2], 0,
             dist, e.decay, 0, 0, 0, 0, 0, 0});
    ++out.n_point;
  }
  // spot (:670-702): inner angle = outer x (1 - penumbra)
  for (uint32_t i = 0; i < n; ++i) {
    const vt_emitter& e = em[i];
    if (e.kind != VT_EMITTER_SPOT) continue;
    if (!radiance(e, rad) || !finite3(e.position) || !finite3(e.direction) || !std::isfinite(e.angle))
      return label(i) + "colour x intensity, position, direction and angle must be finite";
    const double x = e.direction[0], y = e.direction[1], z = e.direction[2];
    const double len = vtl::js_hypot3(x, y, z);
    const double d[3] = {len < 1e-8 ? 0.0 : x / len, len < 1e-8 ? -1.0 : y / len, len < 1e-8 ? 0.0 : z / len};
    const double pen = std::min(1.0, std::max(0.0, std::isfinite(e.penumbra) ? (double)e.penumbra : 0.0));
    const double outer = e.angle, inner = outer * (1 - pen);
    const float dist = std::isfinite(e.distance) && e.distance > 0 ? e.distance : 0.0f;
    if (!std::isfinite(e.decay)) return label(i) + "decay must be finite";
    push(r, {e.position[0], e.position[1], e.position[2], e.cast_shadow ? 0.0f : 1.0f, (float)d[0], (float)d[1],
             (float)d[2], (float)std::cos(outer), rad[0], rad[1], rad[2], (float)std::cos(inner), dist, e.decay,
             0, 0});
    ++out.n_spot;
  }
  // rect areas (:704-727), then discs (:728-735, packDiscAsRect :261-320)
  for (uint32_t i = 0; i < n; ++i) {
    const vt_emitter& e = em[i];
    if (e.kind != VT_EMITTER_RECT_AREA) continue;
    if (!radiance(e, rad) || !finite3(e.position) || !finite3(e.u_axis) || !finite3(e.v_axis))
      return label(i) + "colour x intensity, position and axes must be finite";
    if (!(vtl::area_f32(e.u_axis, e.v_axis, 4.0f) > 0))
      return label(i) + "rect-area emitter does not retain strictly positive finite area in f32";
    push(r, {e.position[0], e.position[1], e.position[2], e.cast_shadow ? 0.0f : 1.0f, e.u_axis[0], e.u_axis[1],
             e.u_axis[2], 0, e.v_axis[0], e.v_axis[1], e.v_axis[2], 0, rad[0], rad[1], rad[2], 0.0f});
    ++out.n_rect;
  }
  for (uint32_t i = 0; i < n; ++i) {
    const vt_emitter& e = em[i];
    if (e.kind != VT_EMITTER_DISC_AREA) continue;
    if (!radiance(e, rad) || !finite3(e.position) || !finite3(e.normal))
      return label(i) + "colour x intensity, position and normal must be finite";
    if (!std::isfinite(e.radius) || !(e.radius > 0)) return label(i) + "disc-area emitter has a non-positive radius";
    const double nd[3] = {e.normal[0], e.normal[1], e.normal[2]};
    float nn[3], tn[3];
    if (!normalize_f32(nd, nn)) return label(i) + "disc-area emitter has a degenerate normal";
    const double ux = nn[0], uy = nn[1], uz = nn[2];
    double ax = 0, ay = 1, az = 0;
    if (std::fabs(uy) > 0.999) ax = 1, ay = 0, az = 0;
    const double td[3] = {ay * uz - az * uy, az * ux - ax * uz, ax * uy - ay * ux};
    if (!normalize_f32(td, tn)) return label(i) + "disc-area emitter has a degenerate tangent basis";
    const double tx = tn[0], ty = tn[1], tz = tn[2];
    const double b[3] = {uy * tz - uz * ty, uz * tx - ux * tz, ux * ty - uy * tx};
    const double rr = e.radius;
    const float u[3] = {(float)(tx * rr), (float)(ty * rr), (float)(tz * rr)};
    const float v[3] = {(float)(b[0] * rr), (float)(b[1] * rr), (float)(b[2] * rr)};
    if (!(vtl::area_f32(u, v, (float)kPi) > 0))
      return label(i) + "disc-area emitter does not retain strictly positive finite area in f32";
    push(r, {e.position[0], e.position[1], e.position[2], e.cast_shadow ? 0.0f : 1.0f, u[0], u[1], u[2], 0, v[0],
             v[1], v[2], 0, rad[0], rad[1], rad[2], 1.0f});
    ++out.n_rect;
  }
  return "";
}

std::vector<vtl::Leaf> positional_leaves(const Packed& p) {
  std::vector<vtl::Leaf> out;
  const uint32_t first = p.n_dir;
  for (uint32_t k = first; k < p.count(); ++k) {
    const float* q = &p.records[(size_t)k * vtr::kLightFloats];
    vtl::Leaf L{};
    const bool spot = k >= p.n_dir + p.n_point && k < p.n_dir + p.n_point + p.n_spot;
    const bool area = k >= p.n_dir + p.n_point + p.n_spot;
    for (int c = 0; c < 3; ++c) L.centroid[c] = L.lo[c] = L.hi[c] = q[c];
    if (!area) {
      L.power = vtl::power_of(q + (spot ? 8 : 4), -1.0);
      if (spot) {
        const double n = vtl::js_hypot3(q[4], q[5], q[6]);
        for (int c = 0; c < 3; ++c) L.axis[c] = n > 1e-12 ? q[4 + c] / n : 0.0;
        L.oriented = true;
        L.theta_o = 0;
        L.theta_e = std::acos(std::max(-1.0, std::min(1.0, (double)q[7])));
      }
    } else {
      const bool disc = std::fabs(q[15] - 1.0f) < 0.5f;
      const float ar = vtl::area_f32(q + 4, q + 8, disc ? (float)kPi : 4.0f);
      L.power = vtl::power_of(q + 12, ar);
      for (int c = 0; c < 3; ++c) {
        const double ext = disc ? vtl::js_hypot3(q[4], q[5], q[6]) : std::fabs((double)q[4 + c]) + std::fabs((double)q[8 + c]);
        L.lo[c] = (double)q[c] - ext;
        L.hi[c] = (double)q[c] + ext;
      }
      const double u[3] = {q[4], q[5], q[6]}, v[3] = {q[8], q[9], q[10]};
      const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
      const double l = vtl::js_hypot3(n[0], n[1], n[2]);
      for (int c = 0; c < 3; ++c) L.axis[c] = l > 1e-12 ? n[c] / l : 0.0;
      L.oriented = true;
      L.theta_o = 0;
      L.theta_e = kPi / 2;
    }
    out.push_back(L);
  }
  return out;
}

vtl::Leaf directional_leaf(const Packed& p, uint32_t d, const double lo[3], const double hi[3]) {
  vtl::Leaf L{};
  L.power = vtl::power_of(&p.records[(size_t)d * vtr::kLightFloats + 4], -1.0);
  for (int c = 0; c < 3; ++c) {
    L.lo[c] = lo[c];
    L.hi[c] = hi[c];
    L.centroid[c] = (lo[c] + hi[c]) / 2;
  }
  L.oriented = false;
  return L;
}

}  // namespace vte
