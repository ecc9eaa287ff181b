// sobol.h -- Owen-scrambled Sobol sampler of pt-webgpu's Sobol mode
// (packages/pt-webgpu/src/wgsl/common.wgsl.ts:22-197;
// packages/shared-samplers/src/sobol.ts:55-245): Joe-Kuo D(6) directions
// (dimensions 2..513 of the source data -> 0..511), 16 direction bits stored
// bit-reversed, Laine-Karras nested uniform scrambling, an additive ranked
// 8x8 rotation tile and a PCG continuation after dimension 511.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vts {

constexpr uint32_t kDims = 512;
constexpr uint32_t kBits = 16;

__host__ __device__ __forceinline__ uint32_t rev32(uint32_t x) {
  x = ((x & 0xAAAAAAAAu) >> 1) | ((x & 0x55555555u) << 1);
  x = ((x & 0xCCCCCCCCu) >> 2) | ((x & 0x33333333u) << 2);
  x = ((x & 0xF0F0F0F0u) >> 4) | ((x & 0x0F0F0F0Fu) << 4);
  x = ((x & 0xFF00FF00u) >> 8) | ((x & 0x00FF00FFu) << 8);
  return (x >> 16) | (x << 16);
}
__host__ __device__ __forceinline__ uint32_t hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}
__host__ __device__ __forceinline__ uint32_t combine(uint32_t s, uint32_t v) { return s ^ (v + ((s << 6) + (s >> 2))); }
__host__ __device__ __forceinline__ uint32_t scramble(uint32_t v, uint32_t seed) {
  uint32_t x = v + seed;
  x ^= x * 0x6C50B47Cu;
  x ^= x * 0xB82F1E52u;
  x ^= x * 0xC7AFE638u;
  x ^= x * 0x8D22F6E6u;
  return rev32(x);
}
__host__ __device__ __forceinline__ uint32_t frame_key(uint32_t seed, uint32_t frame) {
  const uint32_t block_key = ((hash(seed) & 0xFFFFu) + (frame >> 16) * 0x9E37u) & 0xFFFFu;
  return (block_key << 16) | (frame & 0xFFFFu);
}
__host__ __device__ __forceinline__ uint32_t rotation(uint32_t tile, uint32_t dim) {
  // ranked 8x8 tile (sobol.ts:30-39), 4 ranks per word (6 bits each)
  constexpr uint8_t kRank[64] = {0,  63, 12, 60, 3,  55, 15, 62, 53, 23, 57, 17, 44, 19, 32, 22,
                                 10, 40, 5,  41, 8,  35, 7,  47, 45, 28, 48, 25, 54, 29, 36, 24,
                                 2,  38, 13, 46, 1,  37, 14, 51, 58, 30, 49, 16, 59, 20, 43, 18,
                                 11, 56, 6,  34, 9,  39, 4,  50, 52, 31, 33, 27, 42, 26, 61, 21};
  const uint32_t rank = kRank[tile & 63u];
  return rank == 0 ? 0u : (hash(combine(rank, dim)) & 0x00FFFFFFu);
}

// Sampler word of (sample index, dimension) for a pixel stream; `dirs` holds
// kDims x kBits bit-reversed 16-bit direction components.
__host__ __device__ __forceinline__ uint32_t owen_u32(const uint16_t* dirs, uint32_t sample, uint32_t dim,
                                                      uint32_t tile, uint32_t stream_seed) {
  const uint32_t seed = hash(combine(stream_seed, dim));
  const uint32_t shuffled = scramble(rev32(sample & 0xFFFFu), combine(stream_seed, 0u)) & 0xFFFFu;
  uint32_t comp = 0;
  const uint16_t* d = dirs + dim * kBits;
#pragma unroll
  for (uint32_t b = 0; b < kBits; ++b)
    if ((shuffled >> b) & 1u) comp ^= d[b];
  const uint32_t s24 = (scramble(comp, combine(seed, 1u + (dim & 3u))) >> 8) & 0x00FFFFFFu;
  return ((s24 + rotation(tile, dim)) & 0x00FFFFFFu) << 8;
}

__host__ __device__ __forceinline__ uint32_t pixel_seed(uint32_t px, uint32_t py) {
  return hash(combine(hash(px), py));
}

// Initial PCG continuation state of a stream (sobol.ts:150-156).
__host__ __device__ __forceinline__ uint32_t fallback_init(uint32_t px, uint32_t py, uint32_t key) {
  return hash(combine(combine(combine(pixel_seed(px, py), key >> 16), key & 0xFFFFu), kDims));
}

// Fills kDims * kBits components from the embedded Joe-Kuo rows (host).
void build_directions(uint16_t* out);

}  // namespace vts
