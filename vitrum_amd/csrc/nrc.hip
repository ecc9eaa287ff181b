// nrc.hip -- fused BF16 MFMA kernels of the neural radiance cache.
//
// Network (Mueller et al. 2021, Real-time Neural Radiance Caching): encoded
// input x (64) -> H1..H4 = ReLU(W_l H_{l-1}) (64 wide) -> y = W4 H4 (RGB).
// Input encoding (walkaround-hybrid src/neural/nrc/nrcEncoding.ts): a
// multiresolution hash grid of the position with trainable feature tables
// (Instant-NGP), one-blob bins of the octahedral direction, raw surface
// features.  The tables receive d loss / d input through W0^T, scattered with
// the trilinear weights.
// Loss: relative L2 of walkaround-hybrid/src/neural/nrc/wgsl/fusedMlp.wgsl.ts
// :65-120 (luminance of the prediction as a stop-gradient denominator,
// epsilon 1e-2); Adam as in walkaround-hybrid/src/neural/nrc/
// fusedMlpTrainer.ts:30-84 (FP32 master weights, bf16 copies for the MFMAs).
//
// MFMA formulation.  Every layer is computed transposed, H_out^T = W H_in^T,
// with v_mfma_f32_16x16x32_bf16: A = 16 output neurons x 32 inputs of W,
// B = 32 inputs x 16 samples of H_in^T, C = 16 neurons x 16 samples.  A lane
// (g = lane / 16, r = lane % 16) of C holds neurons 4g..4g+3 of sample r.  The
// 32-wide K of each MFMA step kb is permuted so that lane (g, r) supplies
// inputs {32kb + 4g + j} and {32kb + 16 + 4g + j}, j < 4 -- exactly the values
// the same lane holds in the C tiles of neuron blocks 2kb and 2kb + 1 of the
// previous layer.  Activations therefore pass from layer to layer in
// registers (ReLU, bf16 rounding), with no LDS round trip; the weight
// fragments are packed in the same permuted order.  Backward uses W^T the same
// way.  Weight gradients dW = dH^T H need the sample dimension as K; those
// operands are staged transposed ([neuron][sample]) in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "nrc.h"

namespace vtn {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Round to nearest even (NaN -> quiet NaN).
__host__ __device__ __forceinline__ uint32_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC0u;
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

__device__ __forceinline__ bf16x8 frag_of(uint4 q) { return __builtin_bit_cast(bf16x8, q); }

__device__ __forceinline__ uint4 pack8(const float* v) {
  uint4 q;
  q.x = f2bf(v[0]) | (f2bf(v[1]) << 16);
  q.y = f2bf(v[2]) | (f2bf(v[3]) << 16);
  q.z = f2bf(v[4]) | (f2bf(v[5]) << 16);
  q.w = f2bf(v[6]) | (f2bf(v[7]) << 16);
  return q;
}

// Input neuron of element t (0..7) of lane group g in K step kb.
__host__ __device__ __forceinline__ uint32_t nmap(uint32_t kb, uint32_t g, uint32_t t) {
  return (2u * kb + (t >> 2)) * 16u + g * 4u + (t & 3u);
}

// ---- input encoding (nrcEncoding.ts) --------------------------------------

// Instant-NGP spatial hash (nrcEncoding.ts:spatialHash3D): XOR of the corner
// coordinates times 1, 2654435761 and 805459861 (32-bit wraparound), modulo
// the row count (a power of two).
__device__ __forceinline__ uint32_t grid_hash(uint32_t ix, uint32_t iy, uint32_t iz) {
  return (ix ^ (iy * 2654435761u) ^ (iz * 805459861u)) & (kRows - 1u);
}

// Trilinear corners of level l for a normalised position: row and weight.
struct Corners {
  uint32_t row[8];
  float w[8];
};

__device__ __forceinline__ void grid_corners(uint32_t l, float nx, float ny, float nz, Corners& c) {
  const float res = (float)(kMinRes << l);
  const float px = nx * res, py = ny * res, pz = nz * res;
  const float x0 = floorf(px), y0 = floorf(py), z0 = floorf(pz);
  const float fx = px - x0, fy = py - y0, fz = pz - z0;
  const uint32_t ix = (uint32_t)x0, iy = (uint32_t)y0, iz = (uint32_t)z0;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t cx = k & 1u, cy = (k >> 1) & 1u, cz = (k >> 2) & 1u;
    const float wx = cx ? fx : 1.0f - fx, wy = cy ? fy : 1.0f - fy, wz = cz ? fz : 1.0f - fz;
    c.w[k] = (wx * wy) * wz;
    c.row[k] = l * kRows + grid_hash(ix + cx, iy + cy, iz + cz);
  }
}

// octEncodeDir (nrcEncoding.ts): octahedral map of a direction to [0,1]^2.
__device__ __forceinline__ void oct_encode(float dx, float dy, float dz, float& u, float& v) {
  const float sc = fmaxf(fabsf(dx), fmaxf(fabsf(dy), fabsf(dz)));
  if (!(sc > 0.0f) || !(sc <= 3.402823e38f)) {
    u = v = 0.5f;
    return;
  }
  const float sx = dx / sc, sy = dy / sc, sz = dz / sc;
  const float inv = 1.0f / ((fabsf(sx) + fabsf(sy)) + fabsf(sz));
  float px = sx * inv, py = sy * inv;
  if (sz < 0.0f) {
    const float ox = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
    const float oy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    px = ox;
    py = oy;
  }
  u = px * 0.5f + 0.5f;
  v = py * 0.5f + 0.5f;
}

// One-blob bins b0..b0+3 of a scalar (oneBlobEncodeScalar: Gaussian kernels
// at (i + 0.5) / k, sigma 1/k, normalised to unit L1 over the k bins).
__device__ __forceinline__ void one_blob4(float x, uint32_t b0, float* out) {
  const float uc = fminf(fmaxf(x, 0.0f), 1.0f);
  float a[kBlobBins], sum = 0.0f;
#pragma unroll
  for (uint32_t i = 0; i < kBlobBins; ++i) {
    const float d = (uc - ((float)i + 0.5f) / (float)kBlobBins) * (float)kBlobBins;
    a[i] = expf(-0.5f * (d * d));
    sum += a[i];
  }
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) {
    float x4 = 0.0f;
#pragma unroll
    for (uint32_t i = 0; i < kBlobBins; ++i)
      if (i == b0 + j) x4 = a[i];
    out[j] = sum > 1e-20f ? x4 / sum : x4;
  }
}

// The 16 encoded inputs a lane of group g supplies for one sample, in the
// K order of nmap: kb 0 -> dims {4g..4g+3, 16+4g..16+4g+3}, kb 1 -> dims
// {32+4g..32+4g+3, 48+4g..48+4g+3}.  Lane group g owns hash-grid levels 2g
// and 2g + 1.
// Levels kLdsFirstLevel.. of the tables are read from a workgroup copy in LDS
// when `lds` is given (the inference kernel stages them: their rows are hit
// at random, while the coarse levels' few rows stay in the caches).
constexpr uint32_t kLdsFirstLevel = 3;
constexpr uint32_t kLdsRows = (kLevels - kLdsFirstLevel) * kRows;  // 20,480 float2 = 160 KiB

__device__ __forceinline__ void encode_lane(uint32_t g, const float* f, const float* __restrict__ tables,
                                            float (&v)[2][8], const float2* lds = nullptr) {
  // hash grid, levels 2g, 2g + 1
#pragma unroll
  for (uint32_t j = 0; j < 2; ++j) {
    Corners c;
    const uint32_t level = 2u * g + j;
    grid_corners(level, f[0], f[1], f[2], c);
    float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
      const float2 t = (lds && level >= kLdsFirstLevel)
                           ? lds[c.row[k] - kLdsFirstLevel * kRows]
                           : *reinterpret_cast<const float2*>(tables + 2u * (size_t)c.row[k]);
      f0 += c.w[k] * t.x;
      f1 += c.w[k] * t.y;
    }
    v[0][2 * j] = f0;
    v[0][2 * j + 1] = f1;
  }
  // one-blob of the direction: groups 0, 1 -> u bins 0-3, 4-7; groups 2, 3 -> v
  float ou, ov;
  oct_encode(f[3], f[4], f[5], ou, ov);
  one_blob4(g < 2u ? ou : ov, (g & 1u) * 4u, &v[0][4]);
  // raw features: dims 32-34 normal, 35 roughness, 36-38 albedo
#pragma unroll
  for (uint32_t t = 0; t < 8; ++t) v[1][t] = 0.0f;
  if (g == 0u) {
    v[1][0] = f[6];
    v[1][1] = f[7];
    v[1][2] = f[8];
    v[1][3] = f[9];
  } else if (g == 1u) {
    v[1][0] = f[10];
    v[1][1] = f[11];
    v[1][2] = f[12];
  }
}

// ---- packing: FP32 master weights -> permuted bf16 fragments --------------

__global__ void pack_kernel(const float* __restrict__ w, uint4* __restrict__ frags) {
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;  // one fragment (lane) each
  if (id >= kFrags) return;
  const uint32_t lane = id & 63u, g = lane >> 4, r = lane & 15u;
  float v[8];
  if (id < kHidden * kFragsHidden) {
    // forward, hidden layer l: A = W_l, row o = ob*16 + r, k = nmap(kb, g, t)
    const uint32_t l = id / kFragsHidden, rem = id % kFragsHidden;
    const uint32_t ob = rem / 128u, kb = (rem / 64u) & 1u;
    for (uint32_t t = 0; t < 8; ++t) v[t] = w[l * 4096u + (ob * 16u + r) * 64u + nmap(kb, g, t)];
  } else if (id < kFragsFwd) {
    // forward, output layer: rows 0..2 of W4, zero padding to 16 rows
    const uint32_t kb = ((id - kHidden * kFragsHidden) / 64u) & 1u;
    for (uint32_t t = 0; t < 8; ++t) v[t] = r < kOut ? w[16384u + r * 64u + nmap(kb, g, t)] : 0.0f;
  } else if (id >= kFragsFwd + kFragsBwd) {
    // input gradient of dims 0..15: A = W0^T, row i = r, k = output neuron
    const uint32_t kb = ((id - kFragsFwd - kFragsBwd) / 64u) & 1u;
    for (uint32_t t = 0; t < 8; ++t) v[t] = w[nmap(kb, g, t) * 64u + r];
  } else {
    // backward through W_m (m = 1..4): A = W_m^T, row i = ib*16 + r, k = output neuron
    const uint32_t b = id - kFragsFwd;
    const uint32_t m = 1u + b / kFragsHidden, rem = b % kFragsHidden;
    const uint32_t ib = rem / 128u, kb = (rem / 64u) & 1u;
    const uint32_t i = ib * 16u + r;
    for (uint32_t t = 0; t < 8; ++t) {
      const uint32_t o = nmap(kb, g, t);
      if (m < 4u)
        v[t] = w[m * 4096u + o * 64u + i];
      else
        v[t] = (kb == 0u && o < kOut) ? w[16384u + o * 64u + i] : 0.0f;
    }
  }
  frags[id] = pack8(v);
}

// ---- shared forward ------------------------------------------------------------

// Builds the B fragments of the encoded inputs of 4 sample blocks.
__device__ __forceinline__ void encode_tile(const float* __restrict__ feats, const float* __restrict__ tables,
                                            uint32_t n, uint32_t base, uint32_t g, uint32_t r, uint4 (&B)[4][2],
                                            const float2* lds = nullptr) {
#pragma unroll
  for (uint32_t sb = 0; sb < 4; ++sb) {
    const uint32_t s = base + sb * 16u + r;
    float v[2][8];
    if (s < n) {
      float f[kFeat];
      const float4* fp = reinterpret_cast<const float4*>(feats + (size_t)s * kFeat);
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) {
        const float4 x = fp[q];
        f[4 * q] = x.x, f[4 * q + 1] = x.y, f[4 * q + 2] = x.z, f[4 * q + 3] = x.w;
      }
      encode_lane(g, f, tables, v, lds);
    } else {
#pragma unroll
      for (uint32_t t = 0; t < 8; ++t) v[0][t] = v[1][t] = 0.0f;
    }
    B[sb][0] = pack8(v[0]);
    B[sb][1] = pack8(v[1]);
  }
}

// One hidden layer: B (inputs) -> B (ReLU outputs, bf16); returns the
// positive-output bit mask (bit sb*16 + kb*8 + t).
__device__ __forceinline__ uint64_t hidden_layer(const uint4* __restrict__ frags, uint32_t lane,
                                                 uint4 (&B)[4][2]) {
  f32x4 C[4][4];
#pragma unroll
  for (uint32_t ob = 0; ob < 4; ++ob)
#pragma unroll
    for (uint32_t sb = 0; sb < 4; ++sb) C[ob][sb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (uint32_t ob = 0; ob < 4; ++ob)
#pragma unroll
    for (uint32_t kb = 0; kb < 2; ++kb) {
      const bf16x8 a = frag_of(frags[(ob * 2u + kb) * 64u + lane]);
#pragma unroll
      for (uint32_t sb = 0; sb < 4; ++sb) C[ob][sb] = mfma(a, frag_of(B[sb][kb]), C[ob][sb]);
    }
  uint64_t mask = 0;
#pragma unroll
  for (uint32_t sb = 0; sb < 4; ++sb)
#pragma unroll
    for (uint32_t kb = 0; kb < 2; ++kb) {
      float v[8];
#pragma unroll
      for (uint32_t t = 0; t < 8; ++t) {
        const float x = C[2u * kb + (t >> 2)][sb][t & 3u];
        v[t] = x > 0.0f ? x : 0.0f;
        if (x > 0.0f) mask |= 1ull << (sb * 16u + kb * 8u + t);
      }
      B[sb][kb] = pack8(v);
    }
  return mask;
}

__device__ __forceinline__ void output_layer(const uint4* __restrict__ frags, uint32_t lane,
                                             const uint4 (&B)[4][2], f32x4 (&Y)[4]) {
#pragma unroll
  for (uint32_t sb = 0; sb < 4; ++sb) Y[sb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (uint32_t kb = 0; kb < 2; ++kb) {
    const bf16x8 a = frag_of(frags[kHidden * kFragsHidden + kb * 64u + lane]);
#pragma unroll
    for (uint32_t sb = 0; sb < 4; ++sb) Y[sb] = mfma(a, frag_of(B[sb][kb]), Y[sb]);
  }
}

// ---- inference: one wave = 64 samples -------------------------------------------
//
// Persistent: one 1024-thread workgroup per CU copies levels 3-7 of the
// tables (160 KiB) into LDS once, then each of its 16 waves takes tiles of 64
// samples, grid-strided.  The encoding reads the same f32 table values as from
// memory, so the outputs are those of a direct gather.

__global__ __launch_bounds__(1024) void infer_kernel(const float* __restrict__ feats, uint32_t cap,
                                                     const uint32_t* __restrict__ n_ptr,
                                                     const uint4* __restrict__ frags,
                                                     const float* __restrict__ tables, float* __restrict__ rgb) {
  __shared__ float2 tab[kLdsRows];
  const float2* src = reinterpret_cast<const float2*>(tables) + kLdsFirstLevel * kRows;
  for (uint32_t i = threadIdx.x; i < kLdsRows; i += 1024u) tab[i] = src[i];
  __syncthreads();
  const uint32_t n = n_ptr ? min(*n_ptr, cap) : cap;
  const uint32_t lane = threadIdx.x & 63u, g = lane >> 4, r = lane & 15u;
  const uint32_t tiles = (n + 63u) / 64u, waves = gridDim.x * 16u;
  for (uint32_t tile = blockIdx.x * 16u + (threadIdx.x >> 6); tile < tiles; tile += waves) {
    const uint32_t base = tile * 64u;
    // re-read the fragments every tile (an opaque pointer copy keeps them
    // from being hoisted out of the loop into spilled registers)
    const uint4* fw = frags;
    asm volatile("" : "+s"(fw));
    uint4 B[4][2];
    encode_tile(feats, tables, n, base, g, r, B, tab);
#pragma unroll
    for (uint32_t l = 0; l < kHidden; ++l) (void)hidden_layer(fw + l * kFragsHidden, lane, B);
    f32x4 Y[4];
    output_layer(fw, lane, B, Y);
    if (g == 0u) {
#pragma unroll
      for (uint32_t sb = 0; sb < 4; ++sb) {
        const uint32_t s = base + sb * 16u + r;
        if (s < n) {
          rgb[3 * (size_t)s] = Y[sb][0];
          rgb[3 * (size_t)s + 1] = Y[sb][1];
          rgb[3 * (size_t)s + 2] = Y[sb][2];
        }
      }
    }
  }
}

// ---- training: forward, relative-L2 gradient, backward, dW ------------------------
//
// A workgroup of four waves trains on tiles of 64 samples; wave w owns the
// samples w*16 .. w*16+15 of a tile (one 16-column block of every MFMA), so a
// wave's activations are 2 fragments per layer and its registers leave room
// for three workgroups per CU.  The layer inputs are staged transposed
// ([neuron][sample], bf16) in the workgroup's LDS for the weight gradients,
// dW = D^T H over the tile's 64 samples, which wave w computes for output
// rows w*16 .. w*16+15 into accumulators that stay in registers across all of
// the workgroup's tiles; they are written to the workgroup's slot once.

constexpr uint32_t kTile = 64u * 64u;  // one transposed [neuron][sample] bf16 tile

// Encoded inputs of sample s for lane group g (zero past n).
__device__ __forceinline__ void encode_block(const float* __restrict__ feats, const float* __restrict__ tables,
                                             uint32_t n, uint32_t s, uint32_t g, uint4 (&B)[2]) {
  float v[2][8];
  if (s < n) {
    float f[kFeat];
    const float4* fp = reinterpret_cast<const float4*>(feats + (size_t)s * kFeat);
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
      const float4 x = fp[q];
      f[4 * q] = x.x, f[4 * q + 1] = x.y, f[4 * q + 2] = x.z, f[4 * q + 3] = x.w;
    }
    encode_lane(g, f, tables, v);
  } else {
#pragma unroll
    for (uint32_t t = 0; t < 8; ++t) v[0][t] = v[1][t] = 0.0f;
  }
  B[0] = pack8(v[0]);
  B[1] = pack8(v[1]);
}

// One hidden layer of one 16-sample block: B -> ReLU(W B) in bf16; returns
// the positive-output mask (bit kb*8 + t).
__device__ __forceinline__ uint32_t hidden_block(const uint4* __restrict__ frags, uint32_t lane, uint4 (&B)[2]) {
  f32x4 C[4];
#pragma unroll
  for (uint32_t ob = 0; ob < 4; ++ob) C[ob] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (uint32_t ob = 0; ob < 4; ++ob)
#pragma unroll
    for (uint32_t kb = 0; kb < 2; ++kb) C[ob] = mfma(frag_of(frags[(ob * 2u + kb) * 64u + lane]), frag_of(B[kb]), C[ob]);
  uint32_t mask = 0;
#pragma unroll
  for (uint32_t kb = 0; kb < 2; ++kb) {
    float v[8];
#pragma unroll
    for (uint32_t t = 0; t < 8; ++t) {
      const float x = C[2u * kb + (t >> 2)][t & 3u];
      v[t] = x > 0.0f ? x : 0.0f;
      if (x > 0.0f) mask |= 1u << (kb * 8u + t);
    }
    B[kb] = pack8(v);
  }
  return mask;
}

// Backward through W_m^T for one block: dH = (W_m^T dNext) masked by ReLU'.
__device__ __forceinline__ void backward_block(const uint4* __restrict__ fr, uint32_t lane, uint32_t kbs,
                                               uint32_t mask, uint4 (&B)[2]) {
  f32x4 C[4];
#pragma unroll
  for (uint32_t ib = 0; ib < 4; ++ib) C[ib] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (uint32_t ib = 0; ib < 4; ++ib)
    for (uint32_t kb = 0; kb < kbs; ++kb) C[ib] = mfma(frag_of(fr[(ib * 2u + kb) * 64u + lane]), frag_of(B[kb]), C[ib]);
#pragma unroll
  for (uint32_t kb = 0; kb < 2; ++kb) {
    float v[8];
#pragma unroll
    for (uint32_t t = 0; t < 8; ++t) v[t] = ((mask >> (kb * 8u + t)) & 1u) ? C[2u * kb + (t >> 2)][t & 3u] : 0.0f;
    B[kb] = pack8(v);
  }
}

// Stores a block's fragments transposed into column block w: T[n * 64 + w*16 + r].
__device__ __forceinline__ void store_block(uint16_t* T, uint32_t w, uint32_t g, uint32_t r, const uint4 (&B)[2]) {
#pragma unroll
  for (uint32_t kb = 0; kb < 2; ++kb) {
    const uint32_t q[4] = {B[kb].x, B[kb].y, B[kb].z, B[kb].w};
#pragma unroll
    for (uint32_t t = 0; t < 8; ++t)
      T[nmap(kb, g, t) * 64u + w * 16u + r] = (uint16_t)((q[t >> 1] >> ((t & 1u) * 16u)) & 0xFFFFu);
  }
}

// acc[ib] += D^T H for output rows ob*16.. (A from DT) and input columns
// ib*16.. (B from HT), K = the tile's 64 samples.
__device__ __forceinline__ void weight_grad_rows(const uint16_t* DT, const uint16_t* HT, uint32_t ob, uint32_t g,
                                                 uint32_t r, f32x4 (&acc)[4]) {
#pragma unroll
  for (uint32_t ks = 0; ks < 2; ++ks) {
    const uint4 a = *reinterpret_cast<const uint4*>(DT + (ob * 16u + r) * 64u + ks * 32u + g * 8u);
#pragma unroll
    for (uint32_t ib = 0; ib < 4; ++ib) {
      const uint4 b = *reinterpret_cast<const uint4*>(HT + (ib * 16u + r) * 64u + ks * 32u + g * 8u);
      acc[ib] = mfma(frag_of(a), frag_of(b), acc[ib]);
    }
  }
}

__global__ __launch_bounds__(256, 3) void train_kernel(const float* __restrict__ feats, const float* __restrict__ targets,
                                                    uint32_t cap, const uint32_t* __restrict__ n_ptr,
                                                    const uint4* __restrict__ frags, const float* __restrict__ tables,
                                                    float* __restrict__ partials, float* __restrict__ dx,
                                                    float* __restrict__ loss) {
  const uint32_t n = n_ptr ? min(*n_ptr, cap) : cap;
  const uint32_t tiles = (n + 63u) / 64u;
  __shared__ __align__(16) uint16_t HT[5 * kTile];  // inputs of W0..W4, transposed
  __shared__ __align__(16) uint16_t DT[kTile];      // deltas, transposed
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, g = lane >> 4, r = lane & 15u;
  f32x4 acc[kHidden][4];  // dW_0..dW_3, rows w*16.., all 64 inputs
  f32x4 acc_out;          // dW_4 (rows 0..2 of a 16-row block), inputs w*16..
#pragma unroll
  for (uint32_t m = 0; m < kHidden; ++m)
#pragma unroll
    for (uint32_t ib = 0; ib < 4; ++ib) acc[m][ib] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  acc_out = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float part = 0.0f;
  // Tiles of 64 samples, strided over the workgroups (uniform loop bounds).
  for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const uint32_t s = tile * 64u + w * 16u + r;
    // The fragments are re-read from the cache every tile: an opaque copy of
    // the pointer keeps the compiler from hoisting all 68 KB of them out of
    // the loop into (spilled) registers.
    const uint4* fw = frags;
    asm volatile("" : "+s"(fw));
    const uint4* bw = fw + kFragsFwd;
    __syncthreads();  // the previous tile's last reads of HT/DT
    uint4 B[2];
    uint32_t mask[kHidden];
    encode_block(feats, tables, n, s, g, B);
    store_block(HT, w, g, r, B);
#pragma unroll
    for (uint32_t l = 0; l < kHidden; ++l) {
      // one layer's fragments at a time: loads are not hoisted across layers
      __builtin_amdgcn_sched_barrier(0);
      mask[l] = hidden_block(fw + l * kFragsHidden, lane, B);
      store_block(HT + (l + 1u) * kTile, w, g, r, B);
    }
    __builtin_amdgcn_sched_barrier(0);
    f32x4 Y = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (uint32_t kb = 0; kb < 2; ++kb)
      Y = mfma(frag_of(fw[kHidden * kFragsHidden + kb * 64u + lane]), frag_of(B[kb]), Y);

    // Relative L2 (fusedMlp.wgsl.ts:65-120): d/dy = 2 (y - t) / (3 n (L^2 + eps)),
    // L = luminance of the prediction, normalised by its largest component.
    // Lanes of group 0 hold outputs 0..2 of their sample.
    float d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (g == 0u && s < n) {
      const float p0 = Y[0], p1 = Y[1], p2 = Y[2];
      const float t0 = targets[3 * (size_t)s], t1 = targets[3 * (size_t)s + 1], t2 = targets[3 * (size_t)s + 2];
      const float sc = fmaxf(fabsf(p0), fmaxf(fabsf(p1), fabsf(p2)));
      float lum = 0.0f;
      if (sc > 0.0f) {
        const float nl = (0.299f * (p0 / sc) + 0.587f * (p1 / sc)) + 0.114f * (p2 / sc);
        lum = fminf(fmaxf(nl, -1.0f), 1.0f) * sc;
      }
      const float bl = fminf(fabsf(lum), 1e19f);
      const float den = bl * bl + 1e-2f;
      const float inv = 2.0f / (float)n / 3.0f / den;
      const float p[3] = {p0, p1, p2}, tg[3] = {t0, t1, t2};
#pragma unroll
      for (uint32_t c = 0; c < 3; ++c) {
        const float e = p[c] - tg[c];
        float x = e * inv;
        x = x != x ? 0.0f : fminf(fmaxf(x, -65504.0f), 65504.0f);
        d[c] = x;
        if (e == e && den == den) part += e * e / den;
      }
    }
    B[0] = pack8(d);
    B[1] = make_uint4(0u, 0u, 0u, 0u);
    // dY^T rows 0..15 of this wave's columns: lane (g, r) writes rows 4g..4g+3
    // (the outputs it holds; rows 3..15 are zero padding).
    {
      const uint32_t q0 = B[0].x, q1 = B[0].y;
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t bits = g == 0u ? ((j < 2u ? q0 : q1) >> ((j & 1u) * 16u)) & 0xFFFFu : 0u;
        DT[(4u * g + j) * 64u + w * 16u + r] = (uint16_t)(j < 3u ? bits : 0u);
      }
    }
    __syncthreads();
    // dW4 (rows 0..2 padded to 16) for inputs w*16..: one 16x16 block per wave
#pragma unroll
    for (uint32_t ks = 0; ks < 2; ++ks) {
      const uint4 a = *reinterpret_cast<const uint4*>(DT + r * 64u + ks * 32u + g * 8u);
      const uint4 b = *reinterpret_cast<const uint4*>(HT + 4u * kTile + (w * 16u + r) * 64u + ks * 32u + g * 8u);
      acc_out = mfma(frag_of(a), frag_of(b), acc_out);
    }
    // dH4 through W4^T
    backward_block(bw + 3u * kFragsHidden, lane, 1u, mask[3], B);
#pragma unroll
    for (uint32_t m = 3; m >= 1; --m) {
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();  // every wave's reads of DT are done
      store_block(DT, w, g, r, B);  // dH_{m+1} (pre-activation delta of layer m's output)
      __syncthreads();
      weight_grad_rows(DT, HT + m * kTile, w, g, r, acc[m]);
      backward_block(bw + (m - 1u) * kFragsHidden, lane, 2u, mask[m - 1u], B);
    }
    __syncthreads();
    store_block(DT, w, g, r, B);  // dH1
    __syncthreads();
    weight_grad_rows(DT, HT, w, g, r, acc[0]);
    // d loss / d x of the hash-grid inputs (dims 0..15) = W0^T dH1: lane
    // (g, r) receives dims 4g..4g+3 -- the features of levels 2g, 2g + 1 that
    // it encoded -- of its sample; table_grad_kernel scatters them into the
    // tables with the trilinear weights (hashGridBackward).
    f32x4 X = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (uint32_t kb = 0; kb < 2; ++kb) X = mfma(frag_of(fw[kFragsFwd + kFragsBwd + kb * 64u + lane]), frag_of(B[kb]), X);
    if (s < n) *reinterpret_cast<float4*>(dx + (size_t)s * 16u + 4u * g) = make_float4(X[0], X[1], X[2], X[3]);
  }
  // loss: wave sum, one atomic per wave
  for (uint32_t off = 32; off > 0; off >>= 1) part += __shfl_xor(part, (int)off);
  if (lane == 0 && part != 0.0f) unsafeAtomicAdd(loss, part / (3.0f * (float)n));
  // The workgroup's slot: every element is written by exactly one lane.
  float* __restrict__ slot = partials + (size_t)blockIdx.x * kMlpParams;
#pragma unroll
  for (uint32_t m = 0; m < kHidden; ++m)
#pragma unroll
    for (uint32_t ib = 0; ib < 4; ++ib)
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) slot[m * 4096u + (w * 16u + g * 4u + j) * 64u + ib * 16u + r] = acc[m][ib][j];
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j)
    if (g * 4u + j < kOut) slot[16384u + (g * 4u + j) * 64u + w * 16u + r] = acc_out[j];
}

// Hash-grid table gradient (hashGridBackward): each input gradient of a
// level's two features, times the trilinear weight of each of the 8 corners,
// accumulates into that corner's row.  One workgroup per (level, chunk of
// samples) sums its chunk into a private copy of the level's 4096 x 2
// gradient in LDS and stores the copy whole into its own partial slice
// (part[(chunk * kLevels + level) * 8192 ..]); table_reduce_kernel then sums
// the chunks' slices.  The LDS copy holds fixed-point integers (2^46 per
// unit): LDS integer atomics run about four times faster than float ones
// here, and integer sums do not depend on the order of the additions.  A
// contribution is clamped to +-4 (far beyond any gradient of a training
// step; 16,384 contributions per element per chunk then stay below 2^63); a
// non-finite one makes the chunk's whole slice NaN, as a float sum would
// carry it, so Adam skips that step for the level.
constexpr float kGradFix = 0x1p46f;
constexpr float kGradFixClamp = 4.0f;

__global__ __launch_bounds__(256) void table_grad_kernel(const float* __restrict__ feats, const float* __restrict__ dx,
                                                         uint32_t cap, const uint32_t* __restrict__ n_ptr,
                                                         uint32_t per_chunk, float* __restrict__ part) {
  __shared__ unsigned long long acc[kRows * kLevelFeat];  // two's complement fixed point
  __shared__ uint32_t nonfinite;
  const uint32_t l = blockIdx.x, tid = threadIdx.x;
  for (uint32_t i = tid; i < kRows * kLevelFeat; i += 256u) acc[i] = 0ull;
  if (tid == 0) nonfinite = 0u;
  __syncthreads();
  const uint32_t n = n_ptr ? min(*n_ptr, cap) : cap;
  const uint32_t lo = blockIdx.y * per_chunk, hi = min(n, lo + per_chunk);
  bool bad = false;
#pragma unroll 2
  for (uint32_t s = lo + tid; s < hi; s += 256u) {
    const float* fp = feats + (size_t)s * kFeat;
    Corners c;
    grid_corners(l, fp[0], fp[1], fp[2], c);
    const float d0 = dx[(size_t)s * 16u + 2u * l], d1 = dx[(size_t)s * 16u + 2u * l + 1u];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
      const uint32_t row = c.row[k] - l * kRows;
      const float v0 = c.w[k] * d0, v1 = c.w[k] * d1;
      bad = bad || !(fabsf(v0) <= 3.402823e38f) || !(fabsf(v1) <= 3.402823e38f);
      const float q0 = fminf(fmaxf(v0, -kGradFixClamp), kGradFixClamp) * kGradFix;
      const float q1 = fminf(fmaxf(v1, -kGradFixClamp), kGradFixClamp) * kGradFix;
      atomicAdd(&acc[2u * row], (unsigned long long)__float2ll_rn(q0));
      atomicAdd(&acc[2u * row + 1u], (unsigned long long)__float2ll_rn(q1));
    }
  }
  if (bad) nonfinite = 1u;
  __syncthreads();
  float* __restrict__ out = part + ((size_t)blockIdx.y * kLevels + l) * (kRows * kLevelFeat);
  const bool poisoned = nonfinite != 0u;
  for (uint32_t i = tid; i < kRows * kLevelFeat; i += 256u)
    out[i] = poisoned ? __builtin_nanf("") : (float)((double)(long long)acc[i] * (1.0 / 70368744177664.0));
}

// grad[kMlpParams + e] += the sum over the chunks of element e of the
// partial table gradients (one thread per element).
__global__ __launch_bounds__(256) void table_reduce_kernel(const float* __restrict__ part, uint32_t chunks,
                                                           float* __restrict__ tgrad) {
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= kTableParams) return;
  const uint32_t l = e / (kRows * kLevelFeat), i = e % (kRows * kLevelFeat);
  float sum = 0.0f;
  for (uint32_t c = 0; c < chunks; ++c) sum += part[((size_t)c * kLevels + l) * (kRows * kLevelFeat) + i];
  tgrad[e] += sum;
}

// grad[i] += sum of the partial MLP weight gradients of the workgroups that
// processed at least one tile; blockIdx.y picks every gridDim.y-th slot.
__global__ __launch_bounds__(256) void reduce_slots_kernel(const float* __restrict__ part, uint32_t slots,
                                                           uint32_t cap, const uint32_t* __restrict__ n_ptr,
                                                           float* __restrict__ grad) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const uint32_t n = n_ptr ? min(*n_ptr, cap) : cap;
  const uint32_t used = min(slots, (n + 63u) / 64u);
  if (i >= kMlpParams) return;
  float acc = 0.0f;
  for (uint32_t k = blockIdx.y; k < used; k += gridDim.y) acc += part[(size_t)k * kMlpParams + i];
  if (acc != 0.0f) unsafeAtomicAdd(grad + i, acc);
}

// ---- Adam (fusedMlpTrainer.ts:30-84) -------------------------------------------

__device__ __forceinline__ bool finite_(float x) { return x == x && fabsf(x) <= 3.402823e38f; }

__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ grad, float* __restrict__ m,
                            float* __restrict__ v, AdamParams p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float cur = w[i], om = m[i], ov = v[i], gr = grad[i];
  grad[i] = 0.0f;
  if (!finite_(cur) || !finite_(om) || !finite_(ov) || ov < 0.0f) {
    w[i] = finite_(cur) ? cur : 0.0f;
    m[i] = 0.0f;
    v[i] = 0.0f;
    return;
  }
  if (!finite_(gr)) return;
  const float mi = p.beta1 * om + (1.0f - p.beta1) * gr;
  const float vi = p.beta2 * ov + (1.0f - p.beta2) * gr * gr;
  const float mhat = mi / p.bc1, vhat = vi / p.bc2;
  if (!finite_(mi) || !finite_(vi) || !finite_(mhat) || !finite_(vhat) || vi < 0.0f || vhat < 0.0f) return;
  const float lr = i < kMlpParams ? p.lr : p.lr_table;
  const float upd = cur - lr * mhat / (sqrtf(vhat) + p.eps);
  if (!finite_(upd)) return;
  m[i] = mi;
  v[i] = vi;
  w[i] = fminf(fmaxf(upd, -65504.0f), 65504.0f);
}

__global__ void fill_kernel(float* p, uint32_t n, float value) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

// ---- launchers ----------------------------------------------------------------

hipError_t launch_pack(const float* master, uint4* frags, hipStream_t s) {
  hipLaunchKernelGGL(pack_kernel, dim3((kFrags + 255) / 256), dim3(256), 0, s, master, frags);
  return hipGetLastError();
}

hipError_t launch_infer(const float* feats, uint32_t n, const uint32_t* n_ptr, const uint4* frags,
                        const float* tables, float* rgb, hipStream_t s) {
  if (n == 0) return hipSuccess;
  // one workgroup per CU (each holds 160 KiB of LDS), fewer for small batches
  static uint32_t cus = 0;
  if (cus == 0) {
    int dev = 0, c = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    cus = (uint32_t)std::max(1, c);
  }
  const uint32_t tiles = (n + 63u) / 64u;
  const uint32_t blocks = std::min(cus, (tiles + 15u) / 16u);
  hipLaunchKernelGGL(infer_kernel, dim3(blocks), dim3(1024), 0, s, feats, n, n_ptr, frags, tables, rgb);
  return hipGetLastError();
}

hipError_t launch_train(const float* feats, const float* targets, uint32_t n, const uint32_t* n_ptr,
                        const uint4* frags, const float* tables, float* grad, float* part, float* dx, float* loss,
                        hipStream_t s) {
  if (n == 0) return hipSuccess;
  // As many workgroups as are resident at once (bounded by the slots), each
  // looping over tiles, so no workgroup waits for a second wave of launches.
  static uint32_t resident = 0;
  if (resident == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, train_kernel, 256, 0);
    if (e != hipSuccess) return e;
    resident = (uint32_t)std::max(1, cus * std::max(per_cu, 1));
  }
  const uint32_t tiles = (n + 63u) / 64u;
  const uint32_t blocks = std::min(std::min(tiles, resident), kGradSlots);
  hipLaunchKernelGGL(train_kernel, dim3(blocks), dim3(256), 0, s, feats, targets, n, n_ptr, frags, tables, part, dx,
                     loss);
  if (hipError_t e = hipGetLastError()) return e;
  hipLaunchKernelGGL(reduce_slots_kernel, dim3((kMlpParams + 255u) / 256u, 32u), dim3(256), 0, s, part, blocks, n,
                     n_ptr, grad);
  if (hipError_t e = hipGetLastError()) return e;
  // About 4096 samples per workgroup, at most 64 chunks per level; the
  // chunks' partial tables go to `part` (free again once reduce_slots_kernel
  // has read the MLP slots: 64 x 65,536 floats <= kGradSlots x kMlpParams).
  static_assert((size_t)64 * kTableParams <= (size_t)kGradSlots * kMlpParams, "table partials fit the slots");
  const uint32_t chunks = std::min(64u, (n + 4095u) / 4096u);
  const uint32_t per_chunk = (n + chunks - 1u) / chunks;
  hipLaunchKernelGGL(table_grad_kernel, dim3(kLevels, chunks), dim3(256), 0, s, feats, dx, n, n_ptr, per_chunk,
                     part);
  if (hipError_t e = hipGetLastError()) return e;
  hipLaunchKernelGGL(table_reduce_kernel, dim3((kTableParams + 255u) / 256u), dim3(256), 0, s, part, chunks,
                     grad + kMlpParams);
  return hipGetLastError();
}

hipError_t launch_adam(float* master, float* grad, float* m, float* v, AdamParams p, hipStream_t s) {
  hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, master, grad, m, v, p);
  return hipGetLastError();
}

hipError_t launch_fill(float* p, uint32_t n, float value, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, value);
  return hipGetLastError();
}

}  // namespace vtn
