// gpubuild.hip -- binned-SAH BLAS construction on the device, one tree level
// per round (see gpubuild.h).
#include "gpubuild.h"

#include <cfloat>

namespace vtg {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kKeyPosInf = 0xFF800000u;  // fkey(+inf)
constexpr uint32_t kKeyNegInf = 0x007FFFFFu;  // fkey(-inf)
constexpr unsigned long long kDKeyPosInf = 0xFFF0000000000000ull;  // dkey(+inf)
constexpr unsigned long long kDKeyNegInf = 0x000FFFFFFFFFFFFFull;  // dkey(-inf)

// Math.min / Math.max of binary32 values (-0 is below +0), as order keys: an
// unsigned key that sorts like the IEEE total order, so an integer minimum
// of keys is the JavaScript minimum of the values.
__device__ inline uint32_t fkey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float unkey(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k); }
__device__ inline uint64_t dkey(double d) {
  const uint64_t u = (uint64_t)__double_as_longlong(d);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double undkey(uint64_t k) {
  return __longlong_as_double((long long)((k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k));
}
__device__ inline float jminf(float a, float b) { return fkey(a) <= fkey(b) ? a : b; }
__device__ inline float jmaxf(float a, float b) { return fkey(a) >= fkey(b) ? a : b; }
__device__ inline double jmax(double a, double b) { return dkey(a) >= dkey(b) ? a : b; }

// surfaceArea with extents clamped at zero (buildArrayBvh.ts:223-231), binary64.
__device__ inline double half_area(const float* b) {
  const double dx = jmax(0.0, (double)b[3] - (double)b[0]);
  const double dy = jmax(0.0, (double)b[4] - (double)b[1]);
  const double dz = jmax(0.0, (double)b[5] - (double)b[2]);
  return dx * dy + dy * dz + dz * dx;
}

__device__ inline void init_bins(uint32_t* bins, uint32_t K, uint32_t first, uint32_t step) {
  for (uint32_t i = first; i < 3 * K * 7; i += step) {
    const uint32_t f = i % 7;
    bins[i] = f < 3 ? kKeyPosInf : (f < 6 ? kKeyNegInf : 0u);
  }
}

// Box and centroid-box order keys of perm[i], i = first, first + step, ... < end.
__device__ inline void range_keys(const uint32_t* perm, const float* bmin, const float* bmax, const double* cen,
                                  uint32_t n_tri, uint32_t first, uint32_t end, uint32_t step, uint32_t lk[6],
                                  unsigned long long ck[6]) {
  for (int k = 0; k < 3; ++k) {
    lk[k] = kKeyPosInf, lk[3 + k] = kKeyNegInf;
    ck[k] = kDKeyPosInf, ck[3 + k] = kDKeyNegInf;
  }
  for (uint32_t i = first; i < end; i += step) {
    const uint32_t t = perm[i];
    for (int k = 0; k < 3; ++k) {
      lk[k] = min(lk[k], fkey(bmin[(size_t)k * n_tri + t]));
      lk[3 + k] = max(lk[3 + k], fkey(bmax[(size_t)k * n_tri + t]));
      const unsigned long long c = dkey(cen[(size_t)k * n_tri + t]);
      ck[k] = min(ck[k], c);
      ck[3 + k] = max(ck[3 + k], c);
    }
  }
}

// Bins perm[i] on every axis whose centroid span exceeds 1e-9: bin index
// floor((c - cmin) / span * K) in binary64, clamped to K - 1
// (buildArrayBvh.ts:342-780); LDS bins hold order keys and a count.
__device__ inline void bin_range(const uint32_t* perm, const float* bmin, const float* bmax, const double* cen,
                                 uint32_t n_tri, uint32_t K, const double* cmin, const double* span, uint32_t first,
                                 uint32_t end, uint32_t step, uint32_t* bins, uint8_t* bin_of) {
  for (uint32_t i = first; i < end; i += step) {
    const uint32_t t = perm[i];
    for (int axis = 0; axis < 3; ++axis) {
      if (!(span[axis] > 1e-9)) continue;
      const double f = floor((cen[(size_t)axis * n_tri + t] - cmin[axis]) / span[axis] * (double)K);
      const uint32_t b = f < (double)(K - 1) ? (uint32_t)f : K - 1;
      bin_of[(size_t)axis * n_tri + t] = (uint8_t)b;
      uint32_t* bb = bins + (axis * K + b) * 7;
      for (int k = 0; k < 3; ++k) {
        atomicMin(&bb[k], fkey(bmin[(size_t)k * n_tri + t]));
        atomicMax(&bb[3 + k], fkey(bmax[(size_t)k * n_tri + t]));
      }
      atomicAdd(&bb[6], 1u);
    }
  }
}

// Exclusive prefix count of `flag` over a 256-thread workgroup (four 64-lane
// waves, ballot per wave); *total receives the workgroup's count.  Every
// thread must call it.
__device__ inline uint32_t block_excl_count(bool flag, uint32_t* wave_tot, uint32_t* total) {
  const uint64_t m = __ballot(flag);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t before = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wave_tot[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t off = 0, sum = 0;
  for (uint32_t w = 0; w < kBlock / 64; ++w) {
    off += w < wave ? wave_tot[w] : 0u;
    sum += wave_tot[w];
  }
  __syncthreads();
  *total = sum;
  return off + before;
}

// Per-triangle records: binary32 bounds by Math.min / Math.max of the three
// vertices, binary64 centroid 0.5 (lo + hi) (buildArrayBvh.ts:165-190).
__global__ __launch_bounds__(kBlock) void tri_records_kernel(const float* __restrict__ pos, uint32_t pstride,
                                                             const uint32_t* __restrict__ idx, uint32_t istride,
                                                             uint32_t n, float* __restrict__ bmin,
                                                             float* __restrict__ bmax, double* __restrict__ cen) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const float* v[3];
  for (int j = 0; j < 3; ++j) v[j] = pos + pstride * (size_t)idx[istride * (size_t)t + j];
  for (int k = 0; k < 3; ++k) {
    const float lo = jminf(jminf(v[0][k], v[1][k]), v[2][k]);
    const float hi = jmaxf(jmaxf(v[0][k], v[1][k]), v[2][k]);
    bmin[(size_t)k * n + t] = lo;
    bmax[(size_t)k * n + t] = hi;
    cen[(size_t)k * n + t] = 0.5 * ((double)lo + (double)hi);
  }
}

// LDS words of one node's slot: per axis K bins of (3 min keys, 3 max keys,
// count), then 3 (K - 1) binary64 candidate costs.
__host__ __device__ inline uint32_t slot_words(uint32_t K) { return ((3 * K * 7 + 1) & ~1u) + 6 * (K - 1); }

// TPN threads per node, kBlock / TPN nodes per workgroup (node_list[i] for
// i = blockIdx.x * kBlock / TPN + slot, i < count): the node's box and
// centroid box, and for a node that is to be split the binned SAH decision.
// Bins live in LDS as order keys updated with integer atomics; every
// (axis, split) candidate is costed on its own thread and the first strict
// minimum is taken in the host builder's order (axis-major, then split).
// Barriers are uniform over the workgroup: a slot without a node, or whose
// node is a leaf, runs through them idle.
template <uint32_t TPN>
__global__ __launch_bounds__(kBlock) void level_kernel(LevelArgs a, uint32_t count) {
  constexpr uint32_t kSlots = kBlock / TPN;
  extern __shared__ unsigned long long lds_raw[];  // 8-byte aligned: the costs are binary64
  uint32_t* lds = reinterpret_cast<uint32_t*>(lds_raw);
  __shared__ uint32_t box_key_s[kSlots][6];
  __shared__ unsigned long long cen_key_s[kSlots][6];
  __shared__ double cmin_s[kSlots][3], span_s[kSlots][3];
  const uint32_t slot = threadIdx.x / TPN, lane = threadIdx.x % TPN;
  const uint32_t li = blockIdx.x * kSlots + slot;
  const bool live = li < count;
  const uint32_t node = live ? a.node_list[li] : 0u;
  const uint32_t begin = live ? a.begin[node] : 0u, end = live ? a.end[node] : 0u;
  const uint32_t n = end - begin;
  const uint32_t K = a.num_bins;
  uint32_t* box_key = box_key_s[slot];
  unsigned long long* cen_key = cen_key_s[slot];
  double* s_cmin = cmin_s[slot];
  double* s_span = span_s[slot];
  uint32_t* bins = lds + slot * slot_words(K);
  const uint32_t ws = live ? a.wide_slot[node] : ~0u;
  const bool wide = ws != ~0u;  // only in the one-node-per-workgroup launch
  if (wide) {
    // reduced and binned by the wide kernels
    if (lane < 6) box_key[lane] = a.w_box_key[6 * (size_t)ws + lane], cen_key[lane] = a.w_cen_key[6 * (size_t)ws + lane];
    const uint32_t* gb = a.w_bins + (size_t)ws * 3 * K * 7;
    for (uint32_t i = lane; i < 3 * K * 7; i += TPN) bins[i] = gb[i];
  } else {
    if (lane < 3) box_key[lane] = kKeyPosInf, box_key[3 + lane] = kKeyNegInf;
    if (lane < 3) cen_key[lane] = kDKeyPosInf, cen_key[3 + lane] = kDKeyNegInf;
    init_bins(bins, K, lane, TPN);
  }
  __syncthreads();
  if (!wide) {
    // node box and centroid box
    uint32_t lk[6];
    unsigned long long ck[6];
    range_keys(a.perm, a.bmin, a.bmax, a.cen, a.n_tri, begin + lane, end, TPN, lk, ck);
    for (int k = 0; k < 3; ++k) {
      atomicMin(&box_key[k], lk[k]);
      atomicMax(&box_key[3 + k], lk[3 + k]);
      atomicMin(&cen_key[k], ck[k]);
      atomicMax(&cen_key[3 + k], ck[3 + k]);
    }
  }
  __syncthreads();
  if (live && lane < 6) a.node_box[6 * (size_t)node + lane] = unkey(box_key[lane]);
  if (lane < 3) {
    const double lo = undkey(cen_key[lane]), hi = undkey(cen_key[3 + lane]);
    if (live) {
      a.cen_box[6 * (size_t)node + lane] = lo;
      a.cen_box[6 * (size_t)node + 3 + lane] = hi;
    }
    s_cmin[lane] = lo;
    s_span[lane] = hi - lo;
  }
  __syncthreads();
  const bool want = live && a.split_wanted[node];
  if (want && !wide)
    bin_range(a.perm, a.bmin, a.bmax, a.cen, a.n_tri, K, s_cmin, s_span, begin + lane, end, TPN, bins, a.bin_of);
  __syncthreads();
  // SAH cost of every (axis, split) candidate: the left box is the
  // Math.min / Math.max union of bins 0..split (exact, so equal to the host
  // builder's running prefix), the right box that of the rest.
  const uint32_t nc = 3 * (K - 1);
  double* cost = reinterpret_cast<double*>(bins + ((3 * K * 7 + 1) & ~1u));
  if (want) {
    float pbox[6];
    for (int k = 0; k < 6; ++k) pbox[k] = unkey(box_key[k]);
    const double parent_sa = half_area(pbox);
    for (uint32_t c = lane; c < nc; c += TPN) {
      const uint32_t axis = c / (K - 1), sp = c % (K - 1);
      double v = INFINITY;
      if (s_span[axis] > 1e-9) {
        const uint32_t* B = bins + axis * K * 7;
        float lb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        float rb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        int32_t nl = 0, nr = 0;
        for (uint32_t b = 0; b < K; ++b) {
          float* box = b <= sp ? lb : rb;
          for (int k = 0; k < 3; ++k) {
            box[k] = jminf(box[k], unkey(B[b * 7 + k]));
            box[3 + k] = jmaxf(box[3 + k], unkey(B[b * 7 + 3 + k]));
          }
          (b <= sp ? nl : nr) += (int32_t)B[b * 7 + 6];
        }
        if (nl != 0 && nr != 0) {
          const double raw = half_area(lb) * (double)nl + half_area(rb) * (double)nr;
          v = parent_sa > 0 ? raw / parent_sa : raw;
        }
      }
      cost[c] = v;
    }
  }
  __syncthreads();
  if (!want || lane != 0) return;
  // first strictly smaller cost in axis-major, split order
  double best = INFINITY;
  uint32_t best_axis = 0, best_split = 0;
  for (uint32_t c = 0; c < nc; ++c)
    if (cost[c] < best) best = cost[c], best_axis = c / (K - 1), best_split = c % (K - 1);
  a.decision[3 * (size_t)node + 0] = best_axis;
  a.decision[3 * (size_t)node + 1] = best_split;
  a.decision[3 * (size_t)node + 2] = (best < (double)n && best != INFINITY) ? 1u : 0u;
}

// Stable scatter of perm[first, last) of a node with split (axis, split):
// triangles with bin <= split go to tmp[left_at...], the rest to
// tmp[right_at...], both in their order.  Every thread of the workgroup
// calls it.
__device__ inline void scatter_range(const uint32_t* perm, const uint8_t* bin, uint32_t split, uint32_t first,
                                     uint32_t last, uint32_t left_at, uint32_t right_at, uint32_t* tmp,
                                     uint32_t* wave_tot) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t c = first; c < last; c += kBlock) {
    const uint32_t i = c + tid;
    const bool in = i < last;
    const uint32_t t = in ? perm[i] : 0u;
    const bool left = in && bin[t] <= split;
    uint32_t chunk_left;
    const uint32_t l_before = block_excl_count(left, wave_tot, &chunk_left);
    if (in) tmp[left ? left_at + l_before : right_at + (tid - l_before)] = t;
    const uint32_t chunk_n = min(kBlock, last - c);
    left_at += chunk_left;
    right_at += chunk_n - chunk_left;
  }
}

// One workgroup per node with an SAH split: a stable partition of its
// segment of the permutation (bin <= split first), through tmp.
__global__ __launch_bounds__(kBlock) void partition_kernel(PartitionArgs a) {
  __shared__ uint32_t wave_tot[kBlock / 64];
  __shared__ uint32_t total_left;
  const uint32_t node = a.nodes[blockIdx.x];
  const uint32_t begin = a.begin[node], end = a.end[node];
  const uint32_t axis = a.decision[3 * (size_t)node], split = a.decision[3 * (size_t)node + 1];
  const uint32_t tid = threadIdx.x;
  const uint8_t* bin = a.bin_of + (size_t)axis * a.n_tri;
  if (tid == 0) total_left = 0;
  __syncthreads();
  uint32_t mine = 0;
  for (uint32_t i = begin + tid; i < end; i += kBlock) mine += bin[a.perm[i]] <= split ? 1u : 0u;
  atomicAdd(&total_left, mine);
  __syncthreads();
  const uint32_t n_left = total_left;
  scatter_range(a.perm, bin, split, begin, end, begin, begin + n_left, a.tmp, wave_tot);
  __syncthreads();
  for (uint32_t i = begin + tid; i < end; i += kBlock) a.perm[i] = a.tmp[i];
  if (tid == 0) a.n_left[node] = n_left;
}

// ---- wide nodes: one range over many workgroups ---------------------------

__global__ __launch_bounds__(kBlock) void wide_init_kernel(WideArgs a) {
  const uint32_t w = blockIdx.x, K = a.num_bins;
  if (threadIdx.x < 3) {
    a.w_box_key[6 * (size_t)w + threadIdx.x] = kKeyPosInf;
    a.w_box_key[6 * (size_t)w + 3 + threadIdx.x] = kKeyNegInf;
    a.w_cen_key[6 * (size_t)w + threadIdx.x] = kDKeyPosInf;
    a.w_cen_key[6 * (size_t)w + 3 + threadIdx.x] = kDKeyNegInf;
  }
  init_bins(a.w_bins + (size_t)w * 3 * K * 7, K, threadIdx.x, kBlock);
}

__global__ __launch_bounds__(kBlock) void wide_box_kernel(WideArgs a) {
  __shared__ uint32_t box_key[6];
  __shared__ unsigned long long cen_key[6];
  const uint32_t c = blockIdx.x, tid = threadIdx.x;
  if (tid < 3) box_key[tid] = kKeyPosInf, box_key[3 + tid] = kKeyNegInf;
  if (tid < 3) cen_key[tid] = kDKeyPosInf, cen_key[3 + tid] = kDKeyNegInf;
  __syncthreads();
  uint32_t lk[6];
  unsigned long long ck[6];
  range_keys(a.perm, a.bmin, a.bmax, a.cen, a.n_tri, a.c_begin[c] + tid, a.c_end[c], kBlock, lk, ck);
  for (int k = 0; k < 3; ++k) {
    atomicMin(&box_key[k], lk[k]);
    atomicMax(&box_key[3 + k], lk[3 + k]);
    atomicMin(&cen_key[k], ck[k]);
    atomicMax(&cen_key[3 + k], ck[3 + k]);
  }
  __syncthreads();
  const uint32_t w = a.c_slot[c];
  if (tid < 3) {
    atomicMin(&a.w_box_key[6 * (size_t)w + tid], box_key[tid]);
    atomicMax(&a.w_box_key[6 * (size_t)w + 3 + tid], box_key[3 + tid]);
    atomicMin(&a.w_cen_key[6 * (size_t)w + tid], cen_key[tid]);
    atomicMax(&a.w_cen_key[6 * (size_t)w + 3 + tid], cen_key[3 + tid]);
  }
}

__global__ __launch_bounds__(kBlock) void wide_bin_kernel(WideArgs a) {
  extern __shared__ uint32_t lds[];
  __shared__ double s_cmin[3], s_span[3];
  const uint32_t c = blockIdx.x, tid = threadIdx.x, K = a.num_bins;
  const uint32_t w = a.c_slot[c];
  if (tid < 3) {
    const double lo = undkey(a.w_cen_key[6 * (size_t)w + tid]), hi = undkey(a.w_cen_key[6 * (size_t)w + 3 + tid]);
    s_cmin[tid] = lo;
    s_span[tid] = hi - lo;
  }
  init_bins(lds, K, threadIdx.x, kBlock);
  __syncthreads();
  bin_range(a.perm, a.bmin, a.bmax, a.cen, a.n_tri, K, s_cmin, s_span, a.c_begin[c] + tid, a.c_end[c], kBlock, lds,
            a.bin_of);
  __syncthreads();
  uint32_t* gb = a.w_bins + (size_t)w * 3 * K * 7;
  for (uint32_t i = tid; i < 3 * K * 7; i += kBlock) {
    const uint32_t f = i % 7;
    if (lds[i - f + 6] == 0) continue;  // empty bin: nothing to merge
    if (f < 3)
      atomicMin(&gb[i], lds[i]);
    else if (f < 6)
      atomicMax(&gb[i], lds[i]);
    else
      atomicAdd(&gb[i], lds[i]);
  }
}

__device__ inline bool wide_split(const WideArgs& a, uint32_t node, uint32_t* axis, uint32_t* split) {
  *axis = a.decision[3 * (size_t)node];
  *split = a.decision[3 * (size_t)node + 1];
  return a.decision[3 * (size_t)node + 2] != 0;
}

__global__ __launch_bounds__(kBlock) void wide_count_kernel(WideArgs a) {
  __shared__ uint32_t total;
  const uint32_t c = blockIdx.x, tid = threadIdx.x;
  uint32_t axis, split;
  if (!wide_split(a, a.c_node[c], &axis, &split)) return;
  const uint8_t* bin = a.bin_of + (size_t)axis * a.n_tri;
  if (tid == 0) total = 0;
  __syncthreads();
  uint32_t mine = 0;
  for (uint32_t i = a.c_begin[c] + tid; i < a.c_end[c]; i += kBlock) mine += bin[a.perm[i]] <= split ? 1u : 0u;
  atomicAdd(&total, mine);
  __syncthreads();
  if (tid == 0) a.chunk_left[c] = total;
}

__global__ __launch_bounds__(kBlock) void wide_scatter_kernel(WideArgs a) {
  __shared__ uint32_t wave_tot[kBlock / 64];
  __shared__ uint32_t s_before, s_all;
  const uint32_t c = blockIdx.x, tid = threadIdx.x;
  const uint32_t node = a.c_node[c], w = a.c_slot[c];
  uint32_t axis, split;
  if (!wide_split(a, node, &axis, &split)) return;
  if (tid == 0) s_before = 0, s_all = 0;
  __syncthreads();
  const uint32_t first = a.w_first[w], count = a.w_count[w];
  uint32_t before = 0, all = 0;
  for (uint32_t j = tid; j < count; j += kBlock) {
    const uint32_t v = a.chunk_left[first + j];
    all += v;
    before += first + j < c ? v : 0u;
  }
  atomicAdd(&s_before, before);
  atomicAdd(&s_all, all);
  __syncthreads();
  const uint32_t nb = a.level_begin[node], cb = a.c_begin[c];
  const uint32_t n_left = s_all, left_at = nb + s_before;
  const uint32_t right_at = nb + n_left + (cb - nb - s_before);
  scatter_range(a.perm, a.bin_of + (size_t)axis * a.n_tri, split, cb, a.c_end[c], left_at, right_at, a.tmp,
                wave_tot);
  if (c == first && tid == 0) a.n_left[node] = n_left;
}

__global__ __launch_bounds__(kBlock) void wide_copy_kernel(WideArgs a) {
  const uint32_t c = blockIdx.x;
  uint32_t axis, split;
  if (!wide_split(a, a.c_node[c], &axis, &split)) return;
  for (uint32_t i = a.c_begin[c] + threadIdx.x; i < a.c_end[c]; i += kBlock) a.perm[i] = a.tmp[i];
}

}  // namespace

size_t level_lds_bytes(uint32_t num_bins, uint32_t threads_per_node) {
  return (size_t)(kBlock / threads_per_node) * slot_words(num_bins) * 4;
}

hipError_t launch_tri_records(const float* pos, uint32_t pstride, const uint32_t* idx, uint32_t istride, uint32_t n,
                              float* bmin, float* bmax, double* cen, hipStream_t s) {
  if (n == 0) return hipSuccess;
  tri_records_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(pos, pstride, idx, istride, n, bmin, bmax, cen);
  return hipGetLastError();
}

hipError_t launch_level(const LevelArgs& a, uint32_t nodes, uint32_t threads_per_node, hipStream_t s) {
  if (nodes == 0) return hipSuccess;
  if (threads_per_node == kBlock) {
    level_kernel<kBlock><<<nodes, kBlock, level_lds_bytes(a.num_bins, kBlock), s>>>(a, nodes);
  } else if (threads_per_node == 64) {
    constexpr uint32_t kSlots = kBlock / 64;
    level_kernel<64><<<(nodes + kSlots - 1) / kSlots, kBlock, level_lds_bytes(a.num_bins, 64), s>>>(a, nodes);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_partition(const PartitionArgs& a, uint32_t nodes, hipStream_t s) {
  if (nodes == 0) return hipSuccess;
  partition_kernel<<<nodes, kBlock, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_wide_reduce(const WideArgs& a, uint32_t wide, uint32_t chunks, hipStream_t s) {
  if (wide == 0) return hipSuccess;
  wide_init_kernel<<<wide, kBlock, 0, s>>>(a);
  wide_box_kernel<<<chunks, kBlock, 0, s>>>(a);
  wide_bin_kernel<<<chunks, kBlock, (size_t)3 * a.num_bins * 7 * 4, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_wide_partition(const WideArgs& a, uint32_t chunks, hipStream_t s) {
  if (chunks == 0) return hipSuccess;
  wide_count_kernel<<<chunks, kBlock, 0, s>>>(a);
  wide_scatter_kernel<<<chunks, kBlock, 0, s>>>(a);
  wide_copy_kernel<<<chunks, kBlock, 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace vtg
