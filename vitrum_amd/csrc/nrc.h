// nrc.h -- neural radiance cache: a 64-64-64-64-64-3 BF16 MLP (4 hidden ReLU
// layers, linear RGB output, no biases: a constant input dimension plays that
// role) with fused MFMA inference and training kernels and an Adam optimizer
// over FP32 master weights (nrc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vtn {

constexpr uint32_t kFeat = 16;          // raw feature floats per sample
constexpr uint32_t kDims = 64;          // encoded input width
constexpr uint32_t kWidth = 64;         // hidden width
constexpr uint32_t kOut = 3;            // RGB
constexpr uint32_t kHidden = 4;         // hidden ReLU layers = 64x64 matrices
constexpr uint32_t kParams = kHidden * kWidth * kWidth + kOut * kWidth;  // 16,576
// bf16 MFMA fragments (one 16-byte fragment per lane per (row block, k step)).
constexpr uint32_t kFragsHidden = 4 * 2 * 64;  // per 64x64 layer: 4 row blocks x 2 k steps x 64 lanes
constexpr uint32_t kFragsOut = 1 * 2 * 64;     // output layer padded to 16 rows
constexpr uint32_t kFragsFwd = kHidden * kFragsHidden + kFragsOut;          // forward (W)
constexpr uint32_t kFragsBwd = kHidden * kFragsHidden;                       // backward (W^T) of layers 1..4
constexpr uint32_t kFrags = kFragsFwd + kFragsBwd;

// Raw feature record (16 floats):
//   0-2  position normalised to [0,1]^3 by the scene bounds
//   3-4  direction (theta/pi, phi/(2 pi)), 5-6 normal (theta/pi, phi/(2 pi))
//   7    roughness, 8-10 diffuse albedo, 11-13 specular albedo, 14-15 zero
struct AdamParams {
  float lr, beta1, beta2, eps;
  float bc1, bc2;  // 1 - beta^t
};

hipError_t launch_pack(const float* master, uint4* frags, hipStream_t s);
// When n_ptr is set the sample count is read from device memory (at most n).
hipError_t launch_infer(const float* feats, uint32_t n, const uint32_t* n_ptr, const uint4* frags, float* rgb,
                        hipStream_t s);
// grad += d loss / d W over the batch; loss[0] += relative-L2 loss of the batch.
hipError_t launch_train(const float* feats, const float* targets, uint32_t n, const uint32_t* n_ptr,
                        const uint4* frags, float* grad, float* loss, hipStream_t s);
hipError_t launch_adam(float* master, float* grad, float* m, float* v, AdamParams p, hipStream_t s);
hipError_t launch_fill(float* p, uint32_t n, float value, hipStream_t s);

}  // namespace vtn
