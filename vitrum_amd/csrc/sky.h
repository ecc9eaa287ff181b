// sky.h -- the procedural daylight environment (ProceduralSkyEnvironment,
// packages/core/src/scene/environment.ts:59-71): the Preetham, Shirley and
// Smits analytic daylight model (SIGGRAPH 1999) baked on the host into an
// equirectangular binary32 radiance map, which the HDRI importance path then
// samples (packages/pt-webgpu/src/scene/environmentPacking.ts:222-282).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace vts {

struct SkyParams {
  double sun[3];           // towards the sun; normalised here ([0, 1, 0] when ~0)
  double turbidity;        // clamped to [1.5, 30]
  double rayleigh;         // >= 0
  double mie_coefficient;  // >= 0
  double mie_g;            // Henyey-Greenstein g, clamped to +-0.999999
  double intensity;        // >= 0
};

// Solid-angle occupancy of the solar cap (angular radius 0.00436) per texel,
// as a multiplier of the texel's centre-sampled measure
// sin(theta_c) dtheta dphi (shared-samplers/src/solarDiscCoverage.ts:48-182).
std::string solar_coverage(uint32_t width, uint32_t height, const double sun[3], std::vector<double>& out);

// bakePreethamSkyEquirect (shared-samplers/src/preethamSky.ts:161-313): the
// width x height map's binary32 rgb (3 floats per texel, row 0 at +Y) and
// the luminance x solid angle integral of those texels.  Empty string on
// success, otherwise the RangeError's message.
std::string bake_preetham(const SkyParams& p, uint32_t width, uint32_t height, std::vector<float>& rgb,
                          double* luminance_integral);

}  // namespace vts
