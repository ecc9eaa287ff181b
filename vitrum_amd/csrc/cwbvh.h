// cwbvh.h -- pt-webgpu's opt-in `cwbvh-closest` traversal mode
// (bvhTraversal: 'cwbvh-closest'): an 8-wide compressed BVH collapsed from the
// reference binary BVH (packages/shared-bvh/src/compressedWideBvh.ts:215-418),
// walked on the device with a Moller-Trumbore proposal that the watertight
// test re-verifies, the binary traversal being the fallback
// (packages/pt-webgpu/src/wgsl/pathTrace/intersection.wgsl.ts:576-838).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace vtw {

constexpr uint32_t kKids = 8;
constexpr uint32_t kInvalidCount = 0xFFFFFFFFu;

// Concatenated wide forest: one tree per binary root.
struct Forest {
  std::vector<float> node_bounds;      // 6 per node: parent min xyz, max xyz
  std::vector<uint32_t> child_bounds;  // 24 per node: 8 children x 3 words (lo16 | hi16 << 16)
  std::vector<uint32_t> child_meta;    // 24 per node: kind, node index or triangle offset, count
  std::vector<uint32_t> child_count;   // per node (kInvalidCount: corrupt tree, use the binary path)
  std::vector<uint32_t> roots;         // wide root per requested binary root
};

// Collapses the binary trees at `roots` of a reference-layout node array (8
// words per node).  tri_count == 0 gives each root the empty wide node.
Forest collapse(const uint32_t* nodes, uint32_t node_count, const std::vector<uint32_t>& roots,
                uint32_t tri_count);

// Device view of a packed scene in the reference layout plus its wide forest.
struct Scene {
  const uint32_t* tlas_nodes;
  uint32_t tlas_node_count;
  const uint32_t* inst_indices;
  uint32_t inst_index_count;
  const float* w2l;
  const float* l2w;
  uint32_t inst_count;
  const uint32_t* blas_roots;  // binary BLAS root per packed instance
  const uint32_t* wide_roots;  // wide root per packed instance
  const uint32_t* bvh_nodes;
  uint32_t bvh_node_count;
  const float4* positions;
  uint32_t vertex_count;
  const uint4* indices;
  uint32_t tri_count;
  const float* node_bounds;
  const uint32_t* child_bounds;
  const uint32_t* child_meta;
  const uint32_t* child_count;
  uint32_t wide_count;
  const uint8_t* single_sided;   // per packed triangle: back faces are not surfaces
  const uint8_t* inst_reversed;  // per packed instance: orientation-reversing L2W
};

struct Params {
  Scene s;
  const float4* rays;     // 2 per ray
  uint32_t* hits;         // 5 words per ray
  float* hit_pos;         // optional, 3 per ray
  uint32_t n;
  const uint32_t* n_ptr;  // when set, the ray count is read from device memory (n = capacity)
  uint32_t sided;         // 1: traceClosest's sided re-trace (intersection.wgsl.ts:462-490)
};

hipError_t launch_trace(const Params& p, uint32_t grid, hipStream_t s);

}  // namespace vtw
