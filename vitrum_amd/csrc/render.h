// render.h -- device data and launchers of the wavefront path tracer (render.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "traverse.h"

namespace vtr {

constexpr uint32_t kMatFloats = 12;   // base rgb, roughness, metallic, Le rgb, twoSided, pad x3
constexpr uint32_t kEmitFloats = 16;  // a, b, c, Le, twoSided, packed instance, packed tri, pad
constexpr uint32_t kInstFloats = 36;  // W2L (16), L2W (16), mirrored, emitter base, pad x2

struct ShadeScene {
  const float4* positions;      // packed vertices (vec4)
  const float4* normals;
  const uint4* indices;         // packed triangles (vec4u)
  const uint32_t* tri_material;
  const uint32_t* tri_source;   // source triangle of each packed triangle
  const float* materials;       // kMatFloats per material
  const float* emitters;        // kEmitFloats per emitter, light-selection order
  const float* instances;       // kInstFloats per packed instance
  const float* env_cdf;         // env_w * env_h + 1
  const float* env_pdf;         // per texel
  uint32_t emitter_count;
  uint32_t env_present, env_w, env_h;
  float env_radiance[3];
  // Light tree over the light slots (16 floats per node); 0 nodes = uniform.
  const float* light_tree;
  uint32_t light_tree_nodes;
};

constexpr uint32_t kTrainSlots = 4;   // recorded vertices per self-training path
constexpr uint32_t kTrainRec = 22;     // 16 feature floats + throughput (3) + radiance (3)

struct FrameParams {
  float inv_vp[16];
  uint32_t width, height;
  uint32_t frame_seed, frame_index;
  uint32_t bounce_limit;
  float ray_bias, tmin;
  // Neural radiance cache: paths [0, pixels) are image paths, the next
  // train_paths are self-training paths started at pseudo-random pixels.
  uint32_t nrc;
  uint32_t pixels, train_paths;
  // Self-training termination: a training path that reaches vertex
  // `train_vertices` ends there with a cache query, whose prediction stands
  // in for the rest of the path (0: training paths run to bounce_limit).
  uint32_t train_vertices;
  float scene_min[3], scene_inv_extent[3];
  // Sampler: PCG, or Owen-scrambled Sobol (frame key, direction table).
  uint32_t sobol;
  uint32_t sobol_key;
  const uint16_t* sobol_dirs;
};

// Structure-of-arrays path state, one path per pixel.
struct PathState {
  uint32_t* rng;       // sampler state word
  uint32_t* rng_dim;   // next Sobol dimension (0xFFFFFFFF: PCG)
  float* throughput;   // 3 per path: throughput entering the current vertex
  float* conn;         // 3 per path: throughput carried by the BSDF sample (pre-RR)
  float* radiance;     // 3 per path
  float* prev_pdf;     // > 0 when the previous sample takes part in area/env MIS
  uint32_t* flags;     // bit 0: survives RR; bits 8-15: next vertex index
  uint64_t* signature; // hash of the path's discrete decisions
  // Extension queue (rays traced with closest hit).
  uint32_t* queue;      // path index per queued ray
  float4* rays;         // 2 per queued ray
  uint32_t* hits;       // 5 per queued ray
  float* hit_pos;       // 3 per queued ray
  uint32_t* count;      // [0] extension rays, [1] next extension rays, [2] shadow rays
  uint32_t* next_queue;
  float4* next_rays;
  // Shadow queue (any-hit).
  uint32_t* shadow_path;
  float4* shadow_rays;
  float* shadow_contrib;  // 3 per shadow ray
  uint32_t* shadow_occ;
  // Accumulation.
  float* accum;           // 4 per pixel (rgb, unused)
  // Neural radiance cache.
  uint32_t* nrc_count;    // [0] cache queries, [1] training samples
  float* q_feat;          // 16 per query: features of the query vertex
  float* q_thr;           // 3 per query: throughput arriving at it
  uint32_t* q_path;       // path of each query
  float* q_rgb;           // 3 per query: cache prediction
  float* train_rec;       // kTrainSlots x kTrainRec per training path
  uint32_t* train_n;      // recorded vertices per training path
  float* train_feat;      // 16 per training sample (compacted)
  float* train_tgt;       // 3 per training sample
};

hipError_t launch_camera(const FrameParams& f, const PathState& s, hipStream_t st);
hipError_t launch_shade(const FrameParams& f, const ShadeScene& sc, const PathState& s,
                        uint32_t vertex, uint32_t capacity, hipStream_t st);
hipError_t launch_shadow_resolve(const PathState& s, uint32_t capacity, hipStream_t st);
hipError_t launch_accumulate(const FrameParams& f, const PathState& s, hipStream_t st);
// Adds throughput x cache prediction to every queried path.
hipError_t launch_nrc_apply(const PathState& s, uint32_t capacity, hipStream_t st);
// Turns the recorded vertices of the training paths into (features, target) samples.
hipError_t launch_nrc_targets(const FrameParams& f, const PathState& s, hipStream_t st);
// which & 1: extension counts rotate; which & 2: the shadow count is cleared.
hipError_t launch_reset_counts(const PathState& s, uint32_t which, hipStream_t st);
hipError_t launch_zero_f32(float* p, uint32_t n, hipStream_t st);

}  // namespace vtr
