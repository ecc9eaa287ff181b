// light_tree.h -- the spatial light tree that selects the NEE light
// (packages/shared-samplers/src/lightTree.ts; selection in
// packages/pt-webgpu/src/wgsl/pathTrace/mediumNee.wgsl.ts:14-38).
#pragma once

#include <cstdint>
#include <vector>

namespace vtl {

constexpr uint32_t kNodeFloats = 16;        // packed f32 per node
constexpr uint32_t kBuckets = 0x01000000u;  // 24-bit root draw
constexpr float kDist2Floor = 1e-3f;        // LT_DIST2_FLOOR (material.wgsl.ts:2131)

// One selectable light: its power, centroid, bounds and emission cone
// (oriented = false: emits in every direction).
struct Leaf {
  double power;
  double centroid[3], lo[3], hi[3];
  bool oriented;
  double axis[3], theta_o, theta_e;
};

// buildLightTree + packLightTreeForGPU (lightTree.ts:492-737): leaf i of the
// result is light slot i; 16 floats per node in preorder.
std::vector<float> build_packed(const std::vector<Leaf>& leaves);

// Leaf of a one-sided or two-sided emissive triangle with f32 world vertices
// and radiance le (emitterPacking.ts:1148-1172 with emitterPower :900-934 and
// the f32 triangle area of areaEmitterGeometry.ts:141-173).
Leaf triangle_leaf(const float a[3], const float b[3], const float c[3], const float le[3], bool two_sided);

// Math.hypot of three finite numbers as V8 computes it (max-scaled,
// compensated sum of squares).
double js_hypot3(double a, double b, double c);
// classifyAreaVectorF32 (areaEmitterGeometry.ts:51-128): coefficient * |u x v|
// in f32 arithmetic, 0 when degenerate or not finite.
float area_f32(const float u[3], const float v[3], float coefficient);
// emitterPower (emitterPacking.ts:900-934); area < 0 marks a delta light.
double power_of(const float le[3], double area);

}  // namespace vtl
