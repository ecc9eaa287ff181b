// traverse.h -- launch interface of the traversal kernels (traverse.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vtk {

constexpr uint32_t kTraceBlock = 128;     // threads per block (2 waves)
constexpr uint32_t kLdsDepth = 20;        // stack entries per lane kept in LDS
constexpr uint32_t kStackCapacity = 128;  // TLAS + BLAS entries per lane (>= 2 * 58 + 2)
// Wide nodes: children per node (a multiple of 4) and node size in float4s:
// six coordinate rows and one child-ref row of kWideKids words, padded to
// 8 rows (device_layout.h).
constexpr uint32_t kWideKids = 4;
constexpr uint32_t kWideNodeFloat4s = 2 * kWideKids;

struct DeviceScene {
  const float4* blas_nodes;  // Node64, 4 x float4 each
  const float4* tlas_nodes;
  const float4* tris;        // 3 x float4 per triangle
  const float4* instances;   // 9 x float4 per instance slot
  uint32_t tlas_root;
  uint32_t tlas_node_count;
  uint32_t tri_count;
  uint32_t inst_count;
  const float4* blas4_nodes;  // Node128, 8 x float4 each (null: use the Node64 trees)
  const float4* tlas4_nodes;
  uint32_t tlas4_root;
  uint32_t reserved;
};

struct TraceParams {
  DeviceScene scene;
  const float4* rays;  // 2 x float4 per ray
  uint32_t* hits;      // 5 words per ray (closest)
  uint32_t* occluded;  // 1 word per ray (any)
  uint4* counts;       // per ray (counting mode)
  uint32_t n;
  const uint32_t* n_ptr;  // when set, the ray count is read from device memory
  float* hit_pos;         // optional: world hit position, 3 floats per ray
  uint32_t* counter;   // work counter, zeroed by the launcher
  uint32_t* error;     // set to non-zero on stack overflow
  uint2* spill;        // (kStackCapacity - kLdsDepth) x grid threads
  uint32_t refill_min;    // refill a wave once this many lanes are idle
  uint32_t steps;         // traversal steps between refill checks
  uint32_t near_first;    // closest hit: visit the nearer child first (ties as the reference)
  uint32_t defer_min;     // run a leaf / instance operation once this many lanes need it
  uint32_t wide;          // 1: walk the Node128 trees (every scene must have them)
  // 1: the path tracer's wrappers -- closest hit skips back faces of
  // single-sided materials (re-tracing beyond them), any hit ignores them
  // (triangle record word 9 = 1 marks a single-sided triangle, instance word
  // 35 = 1 an orientation-reversing instance); 0: raw traversal.
  uint32_t sided;
  // Batched launch: rays [offsets[k], offsets[k+1]) belong to scenes[k].
  const DeviceScene* scenes;
  const uint32_t* scene_offsets;
  uint32_t scene_count;
};

// Closest (any = false) or any hit; `count` selects the work-counting variant
// (Node64 trees, the reference's visiting order).  Otherwise the Node128
// trees are walked when p.wide is set.
hipError_t launch_trace(const TraceParams& p, bool any, bool count, uint32_t grid, hipStream_t s);
// Closest hit over a batch of scenes (p.scenes / p.scene_offsets / p.scene_count).
hipError_t launch_trace_batch(const TraceParams& p, uint32_t grid, hipStream_t s);
// Any hit over a batch of scenes (occluded words into p.occluded).
hipError_t launch_trace_batch_any(const TraceParams& p, uint32_t grid, hipStream_t s);
hipError_t trace_occupancy(bool any, int* blocks_per_cu);  // the smallest over the kernels launched
hipError_t trace_occupancy_batch(int* blocks_per_cu);
hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t s);
// Streaming copy of n16 16-byte words.
hipError_t launch_copy16(uint4* dst, const uint4* src, size_t n16, hipStream_t s);
hipError_t launch_fp32_selftest(const float* a, const float* b, float* q, float* r, uint32_t n,
                                hipStream_t s);

}  // namespace vtk
