// refit.h -- device refit of BVH boxes (refit.hip).
//
// refitBvhBounds (packages/shared-bvh/src/refitBvhBounds.ts:80-179)
// recomputes every node box of a 32-byte preorder BVH from its triangles
// without changing the topology.  The device version runs bottom-up in one
// launch: one thread per leaf computes the leaf box, then climbs through the
// parent links; at each interior node an arrival counter lets the second
// child to finish compute the union (the first one stops), so every node is
// written exactly once, after both of its children, with no thread waiting on
// another.  A node's box is accumulated with the reference's strict
// comparisons in the reference's order (triangles of a leaf in index order,
// vertices 0,1,2; left child then right child), so the boxes match the
// reference's bit for bit, signed zeros included.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vtf {

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Rejections found by the device preflight (lowest offending node wins).
enum RefitError : uint32_t {
  kRefitOk = 0,
  kRefitCorruptInterior = 1,  // split word > 2 or a child outside the array
  kRefitCorruptLeaf = 2,      // triangle range outside the index array
  kRefitBadVertex = 3,        // vertex index outside the position array
  kRefitNonFinite = 4,        // non-finite vertex coordinate
  kRefitRepeated = 5,         // node reached from two parents
  kRefitUnreachable = 6,      // node not reached from node 0
};

struct RefitArgs {
  uint32_t* nodes;           // reference nodes, 8 words each
  uint32_t first, count;     // node range to refit (one tree)
  const uint32_t* indices;   // triangle vertex indices
  uint32_t index_stride;     // words per triangle (3 or 4)
  const float* positions;
  uint32_t position_stride;  // floats per vertex (3 or 4)
  const uint32_t* parent;    // per node: its parent, kNone at a root
  uint32_t* arrivals;        // per node: zero on entry
  // Optional copies of the boxes in the traversal trees: the Node64 child
  // slot (index * 2 + side) and Node128 slot (index * 4 + slot) of each node.
  const uint32_t* dst64;
  uint32_t* node64;
  const uint32_t* dst128;
  uint32_t* node128;
};

// Preflight of a caller's tree (preflightRefitInputs, refitBvhBounds.ts:
// 3-63): links parents (parent must hold kNone) and reports the first
// rejection into *err as (node << 8 | RefitError); *err must hold ~0.
hipError_t launch_refit_link(const uint32_t* nodes, uint32_t count, const uint32_t* indices, uint64_t tri_count,
                             uint32_t index_stride, const float* positions, uint64_t vertex_count,
                             uint32_t position_stride, uint32_t* parent, unsigned long long* err,
                             hipStream_t s);
hipError_t launch_refit(const RefitArgs& a, hipStream_t s);
// Writes the vertex words of triangle records first..first+count-1 (12 words
// each: a.xyz b.xyz c.xyz, the rest kept) from stride-4 indices and positions.
hipError_t launch_refit_tris(float* tris, const uint32_t* indices, const float* positions, uint32_t first,
                             uint32_t count, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t value, hipStream_t s);

}  // namespace vtf
