// gpubuild.h -- binned-SAH BLAS construction on the device.
//
// The same trees as the host builder (builder.cpp, buildArrayBvh.ts:342-780),
// byte for byte, built one tree level per round for every BLAS of a scene at
// once: per level one workgroup per node computes its box and centroid box
// and, for a node that is to be split, bins its triangles (16 bins per axis,
// binary64 bin index) with integer atomics on order keys in LDS and sweeps
// the bins on one thread in the host builder's order (first strict cost <,
// axis-major); a second kernel partitions the node's range of the
// triangle permutation stably (bin <= split first).  Box minima and maxima
// are JavaScript Math.min / Math.max over order keys, so they do not depend
// on the order the atomics land.  The host keeps the tree's shape (a few
// words per node per level), sorts the rare median-fallback ranges
// (buildArrayBvh.ts:677-694) as the host builder does, and writes the
// preorder node array at the end.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "builder.h"

struct vt_engine;

namespace vtg {

struct LevelArgs {
  const uint32_t* node_list; // level nodes this launch covers
  const uint32_t* begin;     // per level node: range of perm
  const uint32_t* end;
  const uint8_t* split_wanted;  // 1: more triangles than a leaf holds, depth allows a split
  const uint32_t* perm;      // triangle permutation (global ids)
  const float* bmin;         // 3 x n_tri, axis-major
  const float* bmax;
  const double* cen;         // 3 x n_tri centroids
  uint8_t* bin_of;           // 3 x n_tri bin of each triangle on each axis
  float* node_box;           // out: 6 per level node
  double* cen_box;           // out: 6 per level node
  uint32_t* decision;        // out: axis, split bin, SAH split taken
  // Wide nodes (ranges split over several workgroups): wide_slot[node] is
  // the node's slot in the wide arrays, ~0u for a node the level kernel
  // reduces and bins alone.  A wide node's box keys, centroid keys and bins
  // come from the wide kernels.
  const uint32_t* wide_slot;
  const uint32_t* w_box_key;             // 6 per wide node
  const unsigned long long* w_cen_key;   // 6 per wide node
  const uint32_t* w_bins;                // 3 * num_bins * 7 per wide node
  uint32_t n_tri, num_bins;
};

// Chunks of the wide nodes of a level: chunk c covers [c_begin[c], c_end[c])
// of wide node c_slot[c] (level node c_node[c]); a wide node's chunks are
// consecutive, from w_first[slot], w_count[slot] of them.
struct WideArgs {
  const uint32_t* c_begin;
  const uint32_t* c_end;
  const uint32_t* c_slot;
  const uint32_t* c_node;
  const uint32_t* w_first;
  const uint32_t* w_count;
  const uint32_t* level_begin;  // per level node
  const uint32_t* decision;     // per level node
  uint32_t* perm;
  uint32_t* tmp;
  const float* bmin;
  const float* bmax;
  const double* cen;
  uint8_t* bin_of;
  uint32_t* w_box_key;
  unsigned long long* w_cen_key;
  uint32_t* w_bins;
  uint32_t* chunk_left;         // per chunk
  uint32_t* n_left;             // per level node
  uint32_t n_tri, num_bins;
};

struct PartitionArgs {
  const uint32_t* nodes;     // level nodes to partition
  const uint32_t* begin;
  const uint32_t* end;
  const uint32_t* decision;
  const uint8_t* bin_of;
  uint32_t* perm;
  uint32_t* tmp;
  uint32_t* n_left;          // out per level node
  uint32_t n_tri;
};

// Dynamic LDS of a level launch with threads_per_node threads per node.
size_t level_lds_bytes(uint32_t num_bins, uint32_t threads_per_node);
// Vertex indices must be below the vertex count (checked on the host).
hipError_t launch_tri_records(const float* pos, uint32_t pstride, const uint32_t* idx, uint32_t istride, uint32_t n,
                              float* bmin, float* bmax, double* cen, hipStream_t s);
// threads_per_node: 256 (one node per workgroup; wide nodes only here) or
// 64 (four nodes per workgroup, for small nodes).
hipError_t launch_level(const LevelArgs& a, uint32_t nodes, uint32_t threads_per_node, hipStream_t s);
hipError_t launch_partition(const PartitionArgs& a, uint32_t nodes, hipStream_t s);
// Box keys, centroid keys and bins of `wide` wide nodes over `chunks` chunks
// (before launch_level).
hipError_t launch_wide_reduce(const WideArgs& a, uint32_t wide, uint32_t chunks, hipStream_t s);
// Stable partition of the wide nodes whose decision takes the SAH split
// (after launch_level).
hipError_t launch_wide_partition(const WideArgs& a, uint32_t chunks, hipStream_t s);

// vt::build_blas with the SAH work on `eng`'s device and stream: the same
// checks, errors and output, byte for byte.
vt::BuildError build_blas_device(vt_engine* eng, const float* positions, uint32_t position_floats,
                                 uint32_t position_stride, const uint32_t* indices, uint32_t index_words,
                                 uint32_t index_stride, const uint32_t* tri_material_ids, uint32_t max_leaf,
                                 uint32_t num_bins, uint32_t max_depth, vt::BlasOut& out);

// build_blas_device over several meshes at once (stride-4 positions and
// indices): one tree per mesh, all grown together level by level; outs[i]
// equals vt::build_blas of mesh i.
vt::BuildError build_blas_forest(vt_engine* eng, const std::vector<vt::BlasJobIn>& meshes, uint32_t max_leaf,
                                 uint32_t num_bins, uint32_t max_depth, std::vector<vt::BlasOut>& outs);

}  // namespace vtg
