// nrc.h -- neural radiance cache (walkaround-hybrid src/neural/nrc,
// DEFAULT_NRC_CONFIG in nrcSubsystem.ts:115-131): a multiresolution hash-grid
// position encoding (8 levels x 4096 rows x 2 trainable features, level
// resolution floor(4 * 2^l)), one-blob encoded direction (8 Gaussian bins,
// sigma 1/8, for each octahedral coordinate) and raw surface features feed a
// 64-64-64-64-64-3 BF16 MLP (4 hidden ReLU layers, linear RGB output, no
// biases), with fused MFMA inference and training kernels and Adam over FP32
// master parameters (nrc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vtn {

constexpr uint32_t kFeat = 16;          // raw feature floats per sample
constexpr uint32_t kDims = 64;          // MLP input width (39 encoded dimensions, zero padded)
constexpr uint32_t kWidth = 64;         // hidden width
constexpr uint32_t kOut = 3;            // RGB
constexpr uint32_t kHidden = 4;         // hidden ReLU layers = 64x64 matrices
constexpr uint32_t kMlpParams = kHidden * kWidth * kWidth + kOut * kWidth;  // 16,576
// Hash grid (nrcEncoding.ts): kLevels levels, kRows rows of kLevelFeat
// features each, resolution floor(kMinRes * 2^l); the Instant-NGP spatial hash.
constexpr uint32_t kLevels = 8;
constexpr uint32_t kRows = 4096;
constexpr uint32_t kLevelFeat = 2;
constexpr uint32_t kMinRes = 4;
constexpr uint32_t kTableParams = kLevels * kRows * kLevelFeat;  // 65,536
constexpr uint32_t kParams = kMlpParams + kTableParams;          // master vector: MLP, then tables
constexpr uint32_t kBlobBins = 8;
// bf16 MFMA fragments (one 16-byte fragment per lane per (row block, k step)).
constexpr uint32_t kFragsHidden = 4 * 2 * 64;  // per 64x64 layer: 4 row blocks x 2 k steps x 64 lanes
constexpr uint32_t kFragsOut = 1 * 2 * 64;     // output layer padded to 16 rows
constexpr uint32_t kFragsFwd = kHidden * kFragsHidden + kFragsOut;          // forward (W)
constexpr uint32_t kFragsBwd = kHidden * kFragsHidden;                       // backward (W^T) of layers 1..4
constexpr uint32_t kFragsIn = 2 * 64;          // W0^T rows 0..15: gradient of the hash-grid inputs
constexpr uint32_t kFrags = kFragsFwd + kFragsBwd + kFragsIn;

// Raw feature record (16 floats):
//   0-2   position normalised to [0,1]^3 by the scene bounds (clamped)
//   3-5   direction (unit; towards the previous path vertex)
//   6-8   shading normal (unit)
//   9     roughness, 10-12 diffuse albedo, 13-15 zero
// Encoded input (nrcEncoding.ts:380-400 order): dims 0-15 hash grid (level
// l, feature f at 2l + f), 16-23 one-blob of the direction's octahedral u,
// 24-31 of v, 32-34 normal, 35 roughness, 36-38 albedo, 39-63 zero.
struct AdamParams {
  float lr, beta1, beta2, eps;
  float bc1, bc2;  // 1 - beta^t
  float lr_table;  // learning rate of the hash-grid tables
};

// Per-workgroup partial MLP weight gradients of the training kernel (each
// workgroup owns one slot of kMlpParams floats and accumulates its tiles there
// with plain stores; a reduction kernel then adds the slots into grad).
constexpr uint32_t kGradSlots = 1024;

hipError_t launch_pack(const float* master, uint4* frags, hipStream_t s);
// When n_ptr is set the sample count is read from device memory (at most n).
// `tables` = master + kMlpParams (the FP32 hash-grid features).
hipError_t launch_infer(const float* feats, uint32_t n, const uint32_t* n_ptr, const uint4* frags,
                        const float* tables, float* rgb, hipStream_t s);
// grad += d loss / d parameters over the batch (kParams: MLP weights, then the
// tables); loss[0] += relative-L2 loss of the batch.  Scratch: `part` holds
// kGradSlots * kMlpParams floats, `dx` 16 floats per sample (n).
hipError_t launch_train(const float* feats, const float* targets, uint32_t n, const uint32_t* n_ptr,
                        const uint4* frags, const float* tables, float* grad, float* part, float* dx, float* loss,
                        hipStream_t s);
hipError_t launch_adam(float* master, float* grad, float* m, float* v, AdamParams p, hipStream_t s);
hipError_t launch_fill(float* p, uint32_t n, float value, hipStream_t s);

}  // namespace vtn
