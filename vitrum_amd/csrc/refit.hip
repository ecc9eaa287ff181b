// refit.hip -- bottom-up BVH refit on the device (see refit.h).
#include "refit.h"

namespace vtf {
namespace {

constexpr uint32_t kBlock = 256;

__device__ inline bool is_leaf(uint32_t split) { return (split >> 16) == 0xFFFFu; }

__device__ inline void report(unsigned long long* err, uint64_t node, uint32_t kind) {
  atomicMin(err, (unsigned long long)((node << 8) | kind));
}

__global__ __launch_bounds__(kBlock) void refit_link_kernel(const uint32_t* __restrict__ nodes, uint32_t count,
                                                            const uint32_t* __restrict__ indices, uint64_t tri_count,
                                                            uint32_t index_stride,
                                                            const float* __restrict__ positions,
                                                            uint64_t vertex_count, uint32_t position_stride,
                                                            uint32_t* parent, unsigned long long* err) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= count) return;
  const uint32_t* q = nodes + 8 * (size_t)i;
  const uint32_t split = q[7];
  if (!is_leaf(split)) {
    const uint64_t left = (uint64_t)i + 1, right = (uint64_t)i + q[6];
    if (split > 2 || left >= count || right <= i || right >= count) {
      report(err, i, kRefitCorruptInterior);
      return;
    }
    // A second parent (or left == right) finds the link already set.
    if (atomicCAS(&parent[left], kNone, i) != kNone) report(err, left, kRefitRepeated);
    if (atomicCAS(&parent[right], kNone, i) != kNone) report(err, right, kRefitRepeated);
    return;
  }
  const uint64_t n = split & 0xFFFFu, off = q[6];
  if (off > tri_count || n > tri_count - off) {
    report(err, i, kRefitCorruptLeaf);
    return;
  }
  for (uint64_t t = off; t < off + n; ++t)
    for (uint32_t lane = 0; lane < 3; ++lane) {
      const uint32_t v = indices[t * index_stride + lane];
      if (v >= vertex_count) {
        report(err, i, kRefitBadVertex);
        return;
      }
      const float* p = positions + (size_t)v * position_stride;
      if (!isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2])) {
        report(err, i, kRefitNonFinite);
        return;
      }
    }
}

__global__ __launch_bounds__(kBlock) void refit_reach_kernel(const uint32_t* __restrict__ parent, uint32_t count,
                                                             unsigned long long* err) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= 1 && i < count && parent[i] == kNone) report(err, i, kRefitUnreachable);
}

__device__ inline void store_box(const RefitArgs& a, uint32_t j, const float* mn, const float* mx) {
  uint32_t* q = a.nodes + 8 * (size_t)j;
  for (int k = 0; k < 3; ++k) q[k] = __float_as_uint(mn[k]), q[3 + k] = __float_as_uint(mx[k]);
  if (a.dst64) {
    const uint32_t d = a.dst64[j];
    uint32_t* w = a.node64 + 16 * (size_t)(d >> 1) + 6 * (d & 1);
    for (int k = 0; k < 3; ++k) w[k] = __float_as_uint(mn[k]), w[3 + k] = __float_as_uint(mx[k]);
  }
  if (a.dst128) {
    const uint32_t d = a.dst128[j];
    if (d != kNone) {
      uint32_t* w = a.node128 + 32 * (size_t)(d >> 2) + (d & 3);
      for (int k = 0; k < 3; ++k) w[4 * k] = __float_as_uint(mn[k]), w[4 * (3 + k)] = __float_as_uint(mx[k]);
    }
  }
}

__global__ __launch_bounds__(kBlock) void refit_kernel(RefitArgs a) {
  const uint32_t i = a.first + blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.first + a.count) return;
  const uint32_t split = a.nodes[8 * (size_t)i + 7];
  if (!is_leaf(split)) return;
  float mn[3], mx[3];
  const uint32_t n = split & 0xFFFFu, off = a.nodes[8 * (size_t)i + 6];
  // A leaf without triangles (the canonical empty BVH) keeps finite point
  // bounds (refitBvhBounds.ts:133-138).
  const float lo = n == 0 ? 0.0f : INFINITY, hi = n == 0 ? 0.0f : -INFINITY;
  for (int k = 0; k < 3; ++k) mn[k] = lo, mx[k] = hi;
  for (uint32_t t = off; t < off + n; ++t)
    for (uint32_t lane = 0; lane < 3; ++lane) {
      const float* p = a.positions + (size_t)a.indices[(size_t)t * a.index_stride + lane] * a.position_stride;
      for (int k = 0; k < 3; ++k) {
        const float x = p[k];
        if (x < mn[k]) mn[k] = x;
        if (x > mx[k]) mx[k] = x;
      }
    }
  store_box(a, i, mn, mx);
  for (uint32_t j = i;;) {
    const uint32_t p = a.parent[j];
    if (p == kNone) return;
    // Release this subtree's boxes, then count the arrival; the second
    // arrival acquires its sibling's boxes and computes the parent.
    __threadfence();
    if (atomicAdd(&a.arrivals[p], 1u) == 0u) return;
    __threadfence();
    const uint32_t* c[2] = {a.nodes + 8 * ((size_t)p + 1), a.nodes + 8 * ((size_t)p + a.nodes[8 * (size_t)p + 6])};
    for (int k = 0; k < 3; ++k) mn[k] = INFINITY, mx[k] = -INFINITY;
    for (int s = 0; s < 2; ++s)
      for (int k = 0; k < 3; ++k) {
        const float l = __uint_as_float(c[s][k]), h = __uint_as_float(c[s][3 + k]);
        if (l < mn[k]) mn[k] = l;
        if (h > mx[k]) mx[k] = h;
      }
    store_box(a, p, mn, mx);
    j = p;
  }
}

__global__ __launch_bounds__(kBlock) void refit_tris_kernel(float* __restrict__ tris,
                                                            const uint32_t* __restrict__ indices,
                                                            const float* __restrict__ positions, uint32_t first,
                                                            uint32_t count) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= count) return;
  const size_t t = (size_t)first + g;
  float* w = tris + 12 * t;
  for (int v = 0; v < 3; ++v) {
    const float* p = positions + 4 * (size_t)indices[4 * t + v];
    for (int k = 0; k < 3; ++k) w[3 * v + k] = p[k];
  }
}

__global__ __launch_bounds__(kBlock) void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t value) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
    p[i] = value;
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_refit_link(const uint32_t* nodes, uint32_t count, const uint32_t* indices, uint64_t tri_count,
                             uint32_t index_stride, const float* positions, uint64_t vertex_count,
                             uint32_t position_stride, uint32_t* parent, unsigned long long* err,
                             hipStream_t s) {
  if (count == 0) return hipSuccess;
  refit_link_kernel<<<blocks_for(count), kBlock, 0, s>>>(nodes, count, indices, tri_count, index_stride, positions,
                                                         vertex_count, position_stride, parent, err);
  hipError_t r = hipGetLastError();
  if (r != hipSuccess) return r;
  refit_reach_kernel<<<blocks_for(count), kBlock, 0, s>>>(parent, count, err);
  return hipGetLastError();
}

hipError_t launch_refit(const RefitArgs& a, hipStream_t s) {
  if (a.count == 0) return hipSuccess;
  refit_kernel<<<blocks_for(a.count), kBlock, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_refit_tris(float* tris, const uint32_t* indices, const float* positions, uint32_t first,
                             uint32_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  refit_tris_kernel<<<blocks_for(count), kBlock, 0, s>>>(tris, indices, positions, first, count);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t value, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t b = (n + kBlock - 1) / kBlock;
  fill_u32_kernel<<<(uint32_t)(b < 4096 ? b : 4096), kBlock, 0, s>>>(p, n, value);
  return hipGetLastError();
}

}  // namespace vtf
