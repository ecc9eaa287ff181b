// render.h -- device data and launchers of the wavefront path tracer (render.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "traverse.h"

namespace vtr {

// Material record: base rgb, roughness, metallic, Le rgb, twoSided,
// clearcoat, clearcoatRoughness, sheen, sheenRoughness, sheenColor rgb,
// anisotropy, cos and sin of anisotropyRotation, specularColor rgb,
// specularIntensity, pad.
constexpr uint32_t kMatFloats = 24;
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
  const float4* env_texels;     // per texel: scaled radiance rgb, pdf per steradian
  uint32_t emitter_count;
  uint32_t env_present, env_w, env_h;
  float env_rot_c, env_rot_s;   // the dome's rotation about +Y
  // Light tree over the light slots (16 floats per node); 0 nodes = uniform.
  const float* light_tree;
  uint32_t light_tree_nodes;
  // Analytic emitters (kLightFloats each) in light-selection order:
  // directional, point, spot, then rect and disc areas; the mesh-area
  // emitters follow them in the selection index, the environment comes last.
  const float* lights;
  uint32_t n_dir, n_point, n_spot, n_rect;
};

// Analytic emitter records (16 floats; no_shadow = castShadow false):
//   directional: toward-light xyz, angular diameter, irradiance rgb, no_shadow
//   point:       position xyz, no_shadow, radiance rgb, 0, cutoff distance, decay
//   spot:        position xyz, no_shadow, axis xyz, cos outer, radiance rgb,
//                cos inner, cutoff distance, decay
//   rect / disc: centre xyz, no_shadow, u axis xyz, 0, v axis xyz, 0,
//                radiance rgb, shape (0 rect with half-axes u, v; 1 disc, the
//                affine image of the unit disc)
constexpr uint32_t kLightFloats = 16;

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
  // Shade each workgroup's queue entries in material-sorted order.
  uint32_t shade_sort;
  // When set, the frame's seed, index and Sobol frame key are read from these
  // three device words instead (launches recorded once in a HIP graph and
  // replayed frame after frame).
  const uint32_t* frame_dev;
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
  float* moments;         // 2 per pixel: sums of sample luminance and its square
  // First-hit auxiliary images of the latest frame (accumulateFrame,
  // kernel.wgsl.ts:920-960): shading normal * 0.5 + 0.5 and the primary hit
  // distance; base colour, 1 (a miss writes (0.5, 1, 0.5, 0) and zeros).
  float* aux_nd;          // 4 per pixel
  float* aux_albedo;      // 4 per pixel
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
// Present pass (present.wgsl.ts:58-95): running mean of `accum` (divided by
// `samples`), exposure, tonemap operator `mode` (0 aces, 1 agx, 2 reinhard,
// 3 linear, 4 none), sRGB OETF unless `linear_out`, clamped to [0, 65504] and
// rounded to binary16 like the reference's rgba16float target; 4 floats per
// pixel out (alpha 1).
hipError_t launch_present(const float* accum, uint32_t pixels, uint32_t samples, uint32_t mode, float exposure,
                          uint32_t linear_out, float* out, hipStream_t st);
// Per-pixel luminance variance of the accumulated samples: max(0, E[L^2] -
// E[L]^2) in the first of 4 floats per pixel.
hipError_t launch_variance(const float* moments, uint32_t pixels, uint32_t samples, float* out, hipStream_t st);
// Adds throughput x cache prediction to every queried path.
hipError_t launch_nrc_apply(const PathState& s, uint32_t capacity, hipStream_t st);
// Turns the recorded vertices of the training paths into (features, target) samples.
hipError_t launch_nrc_targets(const FrameParams& f, const PathState& s, hipStream_t st);
// which & 1: extension counts rotate; which & 2: the shadow count is cleared.
hipError_t launch_reset_counts(const PathState& s, uint32_t which, hipStream_t st);

// Batched frames: the device view of one member engine's queues.
struct BatchMember {
  const float4* rays[2];      // extension rays of the two ping-pong queues
  uint32_t* hits;             // 5 words per queue entry
  float* hit_pos;             // 3 floats per queue entry
  const float4* shadow_rays;
  uint32_t* shadow_occ;
  const uint32_t* count;      // the member's queue counts (word 0: current, 2: shadow)
};
hipError_t launch_batch_prefix(const BatchMember* m, uint32_t count, uint32_t word, uint32_t* offsets,
                               hipStream_t st);
hipError_t launch_batch_gather(const BatchMember* m, uint32_t count, const uint32_t* offsets, uint32_t shadow,
                               uint32_t cur, float4* out, uint32_t capacity, hipStream_t st);
hipError_t launch_batch_scatter(const BatchMember* m, uint32_t count, const uint32_t* offsets, uint32_t shadow,
                                const uint32_t* words, const float* pos, uint32_t capacity, hipStream_t st);
hipError_t launch_zero_f32(float* p, uint32_t n, hipStream_t st);

}  // namespace vtr
