// builder.h -- host-side acceleration-structure builders of the MI355X backend.
//
// The builders reproduce the reference's binned-SAH BLAS builder
// (packages/shared-bvh/src/buildArrayBvh.ts:342), its TLAS builder
// (packages/shared-bvh/src/tlas.ts:859) and the scene packer
// (packages/shared-bvh/src/scenePack.ts:1104) byte for byte, so traversal on
// the GPU returns the same hit records as the reference.  The implementation
// is iterative with in-place stable partitions over an index permutation and
// builds independent BLASes on a thread pool.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace vt {

constexpr uint32_t kLeafFlag = 0xFFFF0000u;
constexpr uint32_t kBvhStackDepth = 60;   // strides.ts:110
constexpr uint32_t kTlasStackDepth = 64;  // strides.ts:112
constexpr uint32_t kMaxBuildDepth = 58;   // strides.ts:119-125

struct BuildError {
  int code = 0;  // VT_ERR_* value, 0 = ok
  std::string message;
  explicit operator bool() const { return code != 0; }
};

struct BlasOut {
  std::vector<uint32_t> nodes;  // 8 words per node
  std::vector<uint32_t> reordered_indices;
  std::vector<uint32_t> reordered_mat;
  std::vector<uint32_t> reordered_to_source;
};

BuildError build_blas(const float* positions, uint32_t position_floats, uint32_t position_stride,
                      const uint32_t* indices, uint32_t index_words, uint32_t index_stride,
                      const uint32_t* tri_material_ids, uint32_t max_leaf, uint32_t num_bins,
                      uint32_t max_depth, BlasOut& out);

// The argument checks of build_blas (buildArrayBvh.ts:342-420), shared with
// the device builder so both reject the same inputs with the same messages.
BuildError check_blas_inputs(const float* positions, uint32_t position_floats, uint32_t position_stride,
                             const uint32_t* indices, uint32_t index_words, uint32_t index_stride,
                             uint32_t max_leaf, uint32_t num_bins, uint32_t max_depth);

struct TlasOut {
  std::vector<uint32_t> nodes;
  std::vector<uint32_t> instance_indices;
  std::vector<uint32_t> blas_roots;
  std::vector<float> transforms;
  uint32_t fallbacks[4] = {0, 0, 0, 0};
};

BuildError build_tlas(uint32_t n, const uint32_t* blas_ids, const double* aabb_min,
                      const double* aabb_max, const float* world_to_local, uint32_t max_leaf,
                      uint32_t num_bins, uint32_t max_depth, TlasOut& out);

// validateBvhEncoding (packages/shared-bvh/src/validateBvhEncoding.ts:55).
BuildError validate_bvh(const uint32_t* words, uint32_t node_count, const uint32_t* roots,
                        uint32_t root_count, int64_t triangle_count, uint32_t stack_capacity);

bool invert_mat4(const float* m, float* out);

struct PrimitiveIn {
  uint32_t kind;  // 0 mesh, 1 instanced mesh
  uint32_t material;
  const float* positions;
  uint32_t position_floats;
  const float* normals;
  uint32_t normal_floats;
  const uint32_t* indices;
  uint32_t index_count;
  const float* transforms;
  uint32_t transform_count;
};

struct Binding {
  uint32_t blas_root, instance_count, vertex_start, vertex_count, tri_start, tri_count,
      primitive, kind;
  float local_min[3], local_max[3];
};

struct Pack {
  std::vector<float> positions, normals;   // vec4 per vertex
  std::vector<uint32_t> indices;           // vec4u per triangle
  std::vector<uint32_t> tri_material, tri_source, tri_primitive;
  std::vector<uint32_t> bvh_nodes;
  std::vector<uint32_t> tlas_nodes, tlas_instance_indices, tlas_blas_roots;
  std::vector<float> tlas_w2l, tlas_l2w;
  std::vector<uint32_t> instance_primitive, instance_source;
  std::vector<Binding> bindings;
  std::vector<std::string> warnings;
  // World AABB of the packed instances (binary64), kept for scene sizing.
  double world_min[3], world_max[3];
};

// One mesh for a BLAS builder: stride-4 positions and indices.
struct BlasJobIn {
  const float* pos4;
  uint32_t position_floats;
  const uint32_t* idx4;
  uint32_t index_words;
  const uint32_t* mats;
};
// A builder of many BLASes at once (the device builder) with build_blas's
// contract per mesh, called with the 4-triangle leaves, 16 bins and
// maxDepth 58 that packing uses.
using BlasBuilder = std::function<BuildError(const std::vector<BlasJobIn>& meshes, std::vector<BlasOut>& outs)>;

// Meshes of at least large_min_triangles triangles are built by `large`
// (when given), all in one call on the calling thread after the host threads
// finish the rest; should it fail, they are built on the host.  The pack is
// the same either way.
BuildError pack_scene(const PrimitiveIn* prims, uint32_t prim_count, uint32_t threads, Pack& out,
                      const BlasBuilder* large = nullptr, uint32_t large_min_triangles = 0);

// computeLocalAabb (scenePack.ts:172-195) of `count` stride-4 vertices, in
// binary64 with Math.min / Math.max.
void local_aabb(const float* pos4, uint32_t count, double* mn, double* mx);
// transformAabb (scenePack.ts:197-223): the 8 corners of a local box through a
// column-major local-to-world matrix, binary64.
void world_aabb(const float* l2w, const double* lmn, const double* lmx, double* mn, double* mx);

// refitTlasInstances (tlas.ts:1141-1209) in place over reference TLAS nodes:
// the leaves holding a changed instance take its binary64 box (rounded to
// binary32), their ancestors the union of their children.  aabb_min/max hold
// 3 doubles per original instance; every instance's box must be ordered and
// finite (validateTlasRefitAabbs, tlas.ts:1012-1039).
BuildError refit_tlas_instances(std::vector<uint32_t>& nodes, const std::vector<uint32_t>& perm,
                                uint32_t inst_count, const double* aabb_min, const double* aabb_max,
                                const std::vector<uint32_t>& changed);

}  // namespace vt
