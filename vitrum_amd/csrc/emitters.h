// emitters.h -- host packing of the analytic emitters (directional, point,
// spot, rect and disc areas) and their light-tree leaves.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vitrum_hip.h"
#include "light_tree.h"

namespace vte {

struct Packed {
  std::vector<float> records;  // render.h kLightFloats per light, selection order
  uint32_t n_dir = 0, n_point = 0, n_spot = 0, n_rect = 0;
  uint32_t count() const { return n_dir + n_point + n_spot + n_rect; }
};

// packEmitterArrays (packages/pt-webgpu/src/scene/emitterPacking.ts:600-760):
// directional, point, spot, rect then disc records in scene order within each
// kind.  Returns an error message (RangeError in the reference) or "".
std::string pack(const vt_emitter* emitters, uint32_t count, Packed& out);

// Light-tree leaves of the positional analytic lights (point, spot, rect,
// disc) in record order (buildLightTreeInputForScene, emitterPacking.ts:
// 1005-1230).
std::vector<vtl::Leaf> positional_leaves(const Packed& p);

// Leaf of directional light d: its delta power, placed on the union of the
// positional lights' bounds (lo/hi/centroid given).
vtl::Leaf directional_leaf(const Packed& p, uint32_t d, const double lo[3], const double hi[3]);

}  // namespace vte
