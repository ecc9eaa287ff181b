// builder.cpp -- binned-SAH BLAS/TLAS builders and scene packer (host side).
//
// Output bytes equal the reference's (buildArrayBvh.ts, tlas.ts,
// scenePack.ts).  The decisions that fix those bytes are reproduced exactly:
//  * SAH cost, centroids and bin indices are binary64, as JS numbers are;
//  * bounds use Math.min/Math.max semantics, including signed zeros, because
//    a -0.0f bound survives into the packed node;
//  * the first strictly-cheaper split (axis-major, then plane order) wins;
//  * partitions keep the subset order (left, then right), the median fallback
//    sorts by (centroid, source index), and nodes are emitted in preorder.
// Because every partition here is an in-place *stable* partition of one index
// permutation, a node's triangles occupy one contiguous range and a leaf's
// offset is simply the start of that range.
#include "builder.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

#include "../../include/vitrum_hip.h"

namespace vt {
namespace {

inline double jmin(double a, double b) {
  if (a < b) return a;
  if (b < a) return b;
  return (a == 0.0 && (std::signbit(a) || std::signbit(b))) ? -0.0 : a;
}
inline double jmax(double a, double b) {
  if (a > b) return a;
  if (b > a) return b;
  return (a == 0.0 && !(std::signbit(a) && std::signbit(b))) ? 0.0 : a;
}
inline float jminf(float a, float b) { return (float)jmin(a, b); }
inline float jmaxf(float a, float b) { return (float)jmax(a, b); }

inline uint32_t fbits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float bitsf(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

BuildError fail(int code, std::string msg) { return BuildError{code, std::move(msg)}; }

// Half surface area with extents clamped at zero (buildArrayBvh.ts:223-231).
inline double half_area(double x0, double y0, double z0, double x1, double y1, double z1) {
  double dx = jmax(0.0, x1 - x0), dy = jmax(0.0, y1 - y0), dz = jmax(0.0, z1 - z0);
  return dx * dy + dy * dz + dz * dx;
}

struct Box3f {
  float mn[3], mx[3];
  void reset() {
    for (int k = 0; k < 3; ++k) mn[k] = INFINITY, mx[k] = -INFINITY;
  }
};

// Per-triangle records in structure-of-arrays form.
struct TriSoA {
  std::vector<float> mn[3], mx[3];
  std::vector<double> c[3];
};

struct BlasJob {
  uint32_t begin, end, depth, parent;  // parent != ~0u: this job is parent's right child
};

}  // namespace

BuildError check_blas_inputs(const float* positions, uint32_t position_floats, uint32_t position_stride,
                             const uint32_t* indices, uint32_t index_words, uint32_t index_stride,
                             uint32_t max_leaf, uint32_t num_bins, uint32_t max_depth) {
  if (position_stride != 3 && position_stride != 4)
    return fail(VT_ERR_RANGE, "positionStride must be exactly 3 or 4.");
  if (index_stride != 3 && index_stride != 4)
    return fail(VT_ERR_RANGE, "indexStride must be exactly 3 or 4.");
  if (max_leaf < 1 || max_leaf > 0xFFFFu)
    return fail(VT_ERR_RANGE, "maxLeafTriangles must be a safe integer in [1, 65535].");
  if (num_bins < 2 || num_bins > 256)
    return fail(VT_ERR_RANGE, "binCount must be a safe integer in [2, 256].");
  if (max_depth > kMaxBuildDepth)
    return fail(VT_ERR_RANGE, "maxDepth must be a safe integer in [0, 58].");
  if (position_floats % position_stride)
    return fail(VT_ERR_RANGE, "positions.length must be divisible by positionStride.");
  if (index_words % index_stride)
    return fail(VT_ERR_RANGE, "indices.length must be divisible by indexStride.");
  for (uint32_t i = 0; i < position_floats; ++i)
    if (!std::isfinite(positions[i]))
      return fail(VT_ERR_RANGE, "positions[" + std::to_string(i) + "] must be finite.");
  const uint32_t n_vert = position_floats / position_stride;
  for (uint32_t t = 0; t < index_words / index_stride; ++t) {
    const uint32_t* tri = indices + (size_t)t * index_stride;
    if (tri[0] >= n_vert || tri[1] >= n_vert || tri[2] >= n_vert)
      return fail(VT_ERR_RANGE,
                  "triangle " + std::to_string(t) + " references an out-of-range vertex.");
  }
  return {};
}

BuildError build_blas(const float* positions, uint32_t position_floats, uint32_t position_stride,
                      const uint32_t* indices, uint32_t index_words, uint32_t index_stride,
                      const uint32_t* tri_material_ids, uint32_t max_leaf, uint32_t num_bins,
                      uint32_t max_depth, BlasOut& out) {
  if (BuildError e = check_blas_inputs(positions, position_floats, position_stride, indices, index_words,
                                      index_stride, max_leaf, num_bins, max_depth))
    return e;
  const uint32_t n_tri = index_words / index_stride;
  out.nodes.clear();
  if (n_tri == 0) {
    out.nodes.assign(8, 0u);
    out.nodes[7] = kLeafFlag;
    out.reordered_indices.assign(indices, indices + index_words);
    out.reordered_mat.clear();
    out.reordered_to_source.clear();
    return {};
  }
  const uint32_t n_vert = position_floats / position_stride;
  TriSoA R;
  for (int k = 0; k < 3; ++k) {
    R.mn[k].resize(n_tri);
    R.mx[k].resize(n_tri);
    R.c[k].resize(n_tri);
  }
  for (uint32_t t = 0; t < n_tri; ++t) {
    const uint32_t* tri = indices + (size_t)t * index_stride;
    if (tri[0] >= n_vert || tri[1] >= n_vert || tri[2] >= n_vert)
      return fail(VT_ERR_RANGE,
                  "triangle " + std::to_string(t) + " references an out-of-range vertex.");
    const float* a = positions + (size_t)tri[0] * position_stride;
    const float* b = positions + (size_t)tri[1] * position_stride;
    const float* c = positions + (size_t)tri[2] * position_stride;
    for (int k = 0; k < 3; ++k) {
      float lo = jminf(jminf(a[k], b[k]), c[k]);
      float hi = jmaxf(jmaxf(a[k], b[k]), c[k]);
      R.mn[k][t] = lo;
      R.mx[k][t] = hi;
      R.c[k][t] = 0.5 * ((double)lo + (double)hi);
    }
  }

  std::vector<uint32_t> perm(n_tri);
  for (uint32_t t = 0; t < n_tri; ++t) perm[t] = t;
  std::vector<uint8_t> bin_of[3];
  for (int k = 0; k < 3; ++k) bin_of[k].resize(n_tri);
  std::vector<uint32_t> scratch(n_tri);
  const uint32_t K = num_bins;
  std::vector<Box3f> bins(K);
  std::vector<uint32_t> bin_count(K);
  std::vector<float> pre(6 * K), suf(6 * K);
  std::vector<int32_t> pre_n(K), suf_n(K);

  out.nodes.reserve(16 * (size_t)n_tri);
  std::vector<BlasJob> stack;
  stack.push_back({0, n_tri, 0, ~0u});
  while (!stack.empty()) {
    BlasJob job = stack.back();
    stack.pop_back();
    const uint32_t node = (uint32_t)(out.nodes.size() / 8);
    out.nodes.resize(out.nodes.size() + 8, 0u);
    if (job.parent != ~0u) out.nodes[8 * (size_t)job.parent + 6] = node - job.parent;
    const uint32_t n = job.end - job.begin;
    const uint32_t* sub = perm.data() + job.begin;
    Box3f nb;
    nb.reset();
    for (uint32_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        nb.mn[k] = jminf(nb.mn[k], R.mn[k][sub[i]]);
        nb.mx[k] = jmaxf(nb.mx[k], R.mx[k][sub[i]]);
      }
    uint32_t* w = &out.nodes[8 * (size_t)node];
    for (int k = 0; k < 3; ++k) w[k] = fbits(nb.mn[k]), w[3 + k] = fbits(nb.mx[k]);
    if (n <= max_leaf) {
      w[6] = job.begin;
      w[7] = kLeafFlag | n;
      continue;
    }
    if (job.depth >= max_depth) {
      std::string extra = n > 0xFFFFu ? " and the packed 16-bit limit (0xFFFF = 65535)" : "";
      return fail(VT_ERR_BUILD, "maxDepth=" + std::to_string(max_depth) +
                                    " prevents enforcing maxLeafTriangles=" +
                                    std::to_string(max_leaf) + ": depth " +
                                    std::to_string(job.depth) + " still contains " +
                                    std::to_string(n) + " triangles" + extra + ".");
    }
    double cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        cmin[k] = jmin(cmin[k], R.c[k][sub[i]]);
        cmax[k] = jmax(cmax[k], R.c[k][sub[i]]);
      }
    const double parent_sa = half_area(nb.mn[0], nb.mn[1], nb.mn[2], nb.mx[0], nb.mx[1], nb.mx[2]);
    double best_cost = INFINITY;
    int best_axis = 0;
    uint32_t best_split = 0;
    for (int axis = 0; axis < 3; ++axis) {
      const double span = cmax[axis] - cmin[axis];
      if (span <= 1e-9) continue;
      for (uint32_t b = 0; b < K; ++b) bins[b].reset(), bin_count[b] = 0;
      const std::vector<double>& cax = R.c[axis];
      uint8_t* cache = bin_of[axis].data();
      for (uint32_t i = 0; i < n; ++i) {
        const uint32_t t = sub[i];
        const double f = std::floor((cax[t] - cmin[axis]) / span * (double)K);
        const uint32_t b = f < (double)(K - 1) ? (uint32_t)f : K - 1;
        cache[t] = (uint8_t)b;
        Box3f& bb = bins[b];
        for (int k = 0; k < 3; ++k) {
          bb.mn[k] = jminf(bb.mn[k], R.mn[k][t]);
          bb.mx[k] = jmaxf(bb.mx[k], R.mx[k][t]);
        }
        ++bin_count[b];
      }
      Box3f acc;
      acc.reset();
      int32_t cnt = 0;
      for (uint32_t b = 0; b < K; ++b) {
        for (int k = 0; k < 3; ++k) {
          acc.mn[k] = jminf(acc.mn[k], bins[b].mn[k]);
          acc.mx[k] = jmaxf(acc.mx[k], bins[b].mx[k]);
          pre[6 * b + k] = acc.mn[k];
          pre[6 * b + 3 + k] = acc.mx[k];
        }
        cnt += (int32_t)bin_count[b];
        pre_n[b] = cnt;
      }
      acc.reset();
      cnt = 0;
      for (int32_t b = (int32_t)K - 1; b >= 0; --b) {
        for (int k = 0; k < 3; ++k) {
          acc.mn[k] = jminf(acc.mn[k], bins[b].mn[k]);
          acc.mx[k] = jmaxf(acc.mx[k], bins[b].mx[k]);
          suf[6 * b + k] = acc.mn[k];
          suf[6 * b + 3 + k] = acc.mx[k];
        }
        cnt += (int32_t)bin_count[b];
        suf_n[b] = cnt;
      }
      for (uint32_t s = 0; s + 1 < K; ++s) {
        const int32_t nl = pre_n[s], nr = suf_n[s + 1];
        if (nl == 0 || nr == 0) continue;
        const float* L = &pre[6 * s];
        const float* Rb = &suf[6 * (s + 1)];
        const double lsa = half_area(L[0], L[1], L[2], L[3], L[4], L[5]);
        const double rsa = half_area(Rb[0], Rb[1], Rb[2], Rb[3], Rb[4], Rb[5]);
        const double raw = lsa * (double)nl + rsa * (double)nr;
        const double cost = parent_sa > 0 ? raw / parent_sa : raw;
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_split = s;
        }
      }
    }
    uint32_t split_axis = (uint32_t)best_axis;
    uint32_t n_left = 0;
    bool partitioned = false;
    if (best_cost < (double)n && best_cost != INFINITY) {
      // Stable in-place partition on the cached bin index of the chosen axis.
      const uint8_t* cache = bin_of[best_axis].data();
      uint32_t* s = perm.data() + job.begin;
      uint32_t nr = 0;
      for (uint32_t i = 0; i < n; ++i) {
        if (cache[s[i]] <= best_split)
          s[n_left++] = s[i];
        else
          scratch[nr++] = s[i];
      }
      std::memcpy(s + n_left, scratch.data(), 4 * (size_t)nr);
      partitioned = n_left != 0 && nr != 0;
    }
    if (!partitioned) {
      // Median fallback (buildArrayBvh.ts:677-694).
      split_axis = 0;
      double widest = cmax[0] - cmin[0];
      for (int axis = 1; axis < 3; ++axis)
        if (cmax[axis] - cmin[axis] > widest) split_axis = axis, widest = cmax[axis] - cmin[axis];
      const std::vector<double>& cax = R.c[split_axis];
      std::sort(perm.begin() + job.begin, perm.begin() + job.end, [&](uint32_t a, uint32_t b) {
        return cax[a] != cax[b] ? cax[a] < cax[b] : a < b;
      });
      n_left = n / 2;
    }
    out.nodes[8 * (size_t)node + 7] = split_axis;
    stack.push_back({job.begin + n_left, job.end, job.depth + 1, node});
    stack.push_back({job.begin, job.begin + n_left, job.depth + 1, ~0u});
  }
  const uint32_t node_count = (uint32_t)(out.nodes.size() / 8);
  out.reordered_indices.assign((size_t)n_tri * index_stride, 0u);
  out.reordered_mat.resize(n_tri);
  out.reordered_to_source = perm;
  for (uint32_t nt = 0; nt < n_tri; ++nt) {
    const uint32_t ot = perm[nt];
    for (int k = 0; k < 3; ++k)
      out.reordered_indices[(size_t)nt * index_stride + k] = indices[(size_t)ot * index_stride + k];
    out.reordered_mat[nt] = tri_material_ids ? tri_material_ids[ot] : 0u;
  }
  return validate_bvh(out.nodes.data(), node_count, nullptr, 0, n_tri, kBvhStackDepth);
}

BuildError validate_bvh(const uint32_t* w, uint32_t total, const uint32_t* roots,
                        uint32_t root_count, int64_t triangle_count, uint32_t cap) {
  const uint32_t root0 = 0;
  if (!roots) {
    roots = &root0;
    root_count = total ? 1 : 0;
  }
  if (total == 0) {
    if (root_count) return fail(VT_ERR_BUILD, "an empty node buffer cannot declare roots.");
    if (triangle_count > 0) return fail(VT_ERR_BUILD, "empty node buffer cannot cover triangles.");
    return {};
  }
  if (root_count == 0)
    return fail(VT_ERR_BUILD, "non-empty node buffer must declare at least one root.");
  std::vector<uint8_t> seen(total, 0), is_root(total, 0);
  std::vector<std::pair<uint32_t, uint32_t>> ranges;
  std::vector<std::pair<uint32_t, uint32_t>> st;
  for (uint32_t r = 0; r < root_count; ++r) {
    const uint32_t root = roots[r];
    if (root >= total) return fail(VT_ERR_BUILD, "root is outside the node buffer.");
    if (is_root[root]) return fail(VT_ERR_BUILD, "root declared more than once.");
    is_root[root] = 1;
    st.assign(1, {root, 0u});
    while (!st.empty()) {
      auto [node, depth] = st.back();
      st.pop_back();
      if (seen[node]) return fail(VT_ERR_BUILD, "node reachable more than once.");
      seen[node] = 1;
      if (depth + 1 > cap)
        return fail(VT_ERR_BUILD, "traversal requires " + std::to_string(depth + 1) +
                                      " live stack entries, exceeding shader capacity " +
                                      std::to_string(cap) + ".");
      const uint32_t* q = w + 8 * (size_t)node;
      float mn[3], mx[3];
      for (int k = 0; k < 3; ++k) mn[k] = bitsf(q[k]), mx[k] = bitsf(q[3 + k]);
      for (int k = 0; k < 3; ++k)
        if (!std::isfinite(mn[k]) || !std::isfinite(mx[k]))
          return fail(VT_ERR_BUILD, "node has non-finite bounds.");
      for (int k = 0; k < 3; ++k)
        if (mn[k] > mx[k]) return fail(VT_ERR_BUILD, "node has inverted bounds.");
      if ((q[7] >> 16) == 0xFFFFu) {
        const uint32_t count = q[7] & 0xFFFFu, off = q[6];
        if (triangle_count >= 0) {
          if ((int64_t)off > triangle_count || (int64_t)count > triangle_count - off)
            return fail(VT_ERR_BUILD, "leaf triangle range exceeds triangle capacity.");
          ranges.push_back({off, count});
        }
        continue;
      }
      if (q[7] > 2) return fail(VT_ERR_BUILD, "interior node has invalid split axis.");
      const uint64_t lc = (uint64_t)node + 1, rc = (uint64_t)node + q[6];
      if (q[6] <= 1 || lc >= total || rc >= total)
        return fail(VT_ERR_BUILD, "interior node has invalid relative right-child offset.");
      for (uint64_t ch : {lc, rc}) {
        const uint32_t* c = w + 8 * ch;
        if (bitsf(c[0]) < mn[0] || bitsf(c[1]) < mn[1] || bitsf(c[2]) < mn[2] ||
            bitsf(c[3]) > mx[0] || bitsf(c[4]) > mx[1] || bitsf(c[5]) > mx[2])
          return fail(VT_ERR_BUILD, "interior node bounds do not enclose a child.");
      }
      st.push_back({(uint32_t)rc, depth + 1});
      st.push_back({(uint32_t)lc, depth + 1});
    }
  }
  for (uint32_t i = 0; i < total; ++i)
    if (!seen[i]) return fail(VT_ERR_BUILD, "node(s) unreachable from the declared roots.");
  if (triangle_count >= 0) {
    std::sort(ranges.begin(), ranges.end(),
              [](const auto& a, const auto& b) { return a.first < b.first; });
    uint64_t cursor = 0;
    for (auto& r : ranges) {
      if (r.first != cursor) return fail(VT_ERR_BUILD, "leaf triangle ranges do not tile.");
      cursor += r.second;
    }
    if ((int64_t)cursor != triangle_count)
      return fail(VT_ERR_BUILD, "leaf ranges do not cover every triangle.");
  }
  return {};
}

// ---------------------------------------------------------------------------
// TLAS (tlas.ts:859)
// ---------------------------------------------------------------------------

namespace {

struct InstRec {
  uint32_t orig;
  double c[3], mn[3], mx[3];
};
struct TNode {
  double mn[3], mx[3];
  uint32_t right, split;
};

uint32_t balanced_depth(uint32_t count, uint32_t max_leaf) {
  uint64_t leaves = ((uint64_t)count + max_leaf - 1) / max_leaf;
  uint32_t d = 0;
  while ((1ull << d) < leaves) ++d;
  return leaves <= 1 ? 0 : d;
}

struct TlasBuilder {
  std::vector<InstRec> R;
  std::vector<TNode> nodes;
  std::vector<uint32_t> perm;
  uint32_t max_leaf, K, max_depth;
  uint32_t fallbacks[4] = {0, 0, 0, 0};
  BuildError err;

  bool fits(uint32_t l, uint32_t r, uint32_t child_depth) const {
    int64_t rem = (int64_t)max_depth - (int64_t)child_depth;
    return (int64_t)balanced_depth(l, max_leaf) <= rem && (int64_t)balanced_depth(r, max_leaf) <= rem;
  }

  uint32_t bin_index(double c, double cmin, double ext) const {
    double f = std::floor((c - cmin) / ext * (double)K);
    if (f > (double)(K - 1)) f = (double)(K - 1);
    if (f < 0) f = 0;
    return (uint32_t)f;
  }

  uint32_t build(std::vector<uint32_t>& ids, uint32_t depth) {
    const uint32_t self = (uint32_t)nodes.size();
    const uint32_t n = (uint32_t)ids.size();
    if (n <= max_leaf) {
      TNode leaf;
      for (int k = 0; k < 3; ++k) leaf.mn[k] = INFINITY, leaf.mx[k] = -INFINITY;
      for (uint32_t id : ids)
        for (int k = 0; k < 3; ++k) {
          if (R[id].mn[k] < leaf.mn[k]) leaf.mn[k] = R[id].mn[k];
          if (R[id].mx[k] > leaf.mx[k]) leaf.mx[k] = R[id].mx[k];
        }
      leaf.split = kLeafFlag | n;
      leaf.right = (uint32_t)perm.size();
      for (uint32_t id : ids) perm.push_back(R[id].orig);
      nodes.push_back(leaf);
      return self;
    }
    if (depth >= max_depth || balanced_depth(n, max_leaf) > max_depth - depth) {
      err = fail(VT_ERR_BUILD, "Cannot encode " + std::to_string(n) + " instances at depth " +
                                   std::to_string(depth) +
                                   " within the traversal-safe maximum depth.");
      return self;
    }
    double cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t id : ids)
      for (int k = 0; k < 3; ++k) {
        if (R[id].c[k] < cmin[k]) cmin[k] = R[id].c[k];
        if (R[id].c[k] > cmax[k]) cmax[k] = R[id].c[k];
      }
    const double ext[3] = {cmax[0] - cmin[0], cmax[1] - cmin[1], cmax[2] - cmin[2]};
    bool have = false;
    int s_axis = 0;
    uint32_t s_bin = 0;
    double s_cost = 0;
    if (!(ext[0] <= 0 && ext[1] <= 0 && ext[2] <= 0)) {
      struct Bin {
        double mn[3], mx[3];
        uint32_t n;
      };
      std::vector<Bin> bins(K);
      std::vector<double> lsa(K), rsa(K);
      std::vector<uint32_t> ln(K), rn(K);
      for (int axis = 0; axis < 3; ++axis) {
        if (ext[axis] <= 0) continue;
        for (auto& b : bins) {
          for (int k = 0; k < 3; ++k) b.mn[k] = INFINITY, b.mx[k] = -INFINITY;
          b.n = 0;
        }
        for (uint32_t id : ids) {
          Bin& b = bins[bin_index(R[id].c[axis], cmin[axis], ext[axis])];
          for (int k = 0; k < 3; ++k) {
            if (R[id].mn[k] < b.mn[k]) b.mn[k] = R[id].mn[k];
            if (R[id].mx[k] > b.mx[k]) b.mx[k] = R[id].mx[k];
          }
          ++b.n;
        }
        auto sweep = [&](int32_t from, int32_t to, int32_t step, std::vector<double>& sa,
                         std::vector<uint32_t>& cnt) {
          Bin acc;
          for (int k = 0; k < 3; ++k) acc.mn[k] = INFINITY, acc.mx[k] = -INFINITY;
          acc.n = 0;
          for (int32_t b = from; b != to; b += step) {
            if (bins[b].n) {
              for (int k = 0; k < 3; ++k) {
                if (bins[b].mn[k] < acc.mn[k]) acc.mn[k] = bins[b].mn[k];
                if (bins[b].mx[k] > acc.mx[k]) acc.mx[k] = bins[b].mx[k];
              }
              acc.n += bins[b].n;
            }
            double dx = jmax(0.0, acc.mx[0] - acc.mn[0]), dy = jmax(0.0, acc.mx[1] - acc.mn[1]),
                   dz = jmax(0.0, acc.mx[2] - acc.mn[2]);
            sa[b] = 2.0 * (dx * dy + dy * dz + dz * dx);
            cnt[b] = acc.n;
          }
        };
        sweep(0, (int32_t)K, 1, lsa, ln);
        sweep((int32_t)K - 1, -1, -1, rsa, rn);
        const double total = lsa[K - 1];
        if (total <= 0) continue;
        for (uint32_t b = 0; b + 1 < K; ++b) {
          if (ln[b] == 0 || rn[b + 1] == 0) continue;
          const double sah = ((double)ln[b] * lsa[b] + (double)rn[b + 1] * rsa[b + 1]) / total;
          if (!have || sah < s_cost) have = true, s_axis = axis, s_bin = b, s_cost = sah;
        }
      }
    }
    uint32_t axis = have ? (uint32_t)s_axis : 0;
    std::vector<uint32_t> left, right;
    int reason = -1;
    if (!have) {
      reason = 0;
    } else if (s_cost >= (double)n) {
      reason = 1;
    } else {
      for (uint32_t id : ids)
        (bin_index(R[id].c[s_axis], cmin[s_axis], ext[s_axis]) <= s_bin ? left : right).push_back(id);
      if (left.empty() || right.empty())
        reason = 2;
      else if (!fits((uint32_t)left.size(), (uint32_t)right.size(), depth + 1))
        reason = 3;
    }
    if (reason >= 0) {
      axis = 0;
      if (ext[1] > ext[axis]) axis = 1;
      if (ext[2] > ext[axis]) axis = 2;
      std::vector<uint32_t> ord = ids;
      std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
        const double ca = R[a].c[axis], cb = R[b].c[axis];
        return ca != cb ? ca < cb : R[a].orig < R[b].orig;
      });
      left.assign(ord.begin(), ord.begin() + n / 2);
      right.assign(ord.begin() + n / 2, ord.end());
      ++fallbacks[reason];
    }
    if (left.empty() || right.empty() ||
        !fits((uint32_t)left.size(), (uint32_t)right.size(), depth + 1)) {
      err = fail(VT_ERR_BUILD, "Internal balanced fallback failed to make a partition.");
      return self;
    }
    nodes.push_back(TNode{{0, 0, 0}, {0, 0, 0}, 0, axis});
    ids.clear();
    ids.shrink_to_fit();
    build(left, depth + 1);
    if (err) return self;
    const uint32_t r = build(right, depth + 1);
    if (err) return self;
    TNode& me = nodes[self];
    me.right = r - self;
    for (int k = 0; k < 3; ++k) {
      me.mn[k] = jmin(nodes[self + 1].mn[k], nodes[r].mn[k]);
      me.mx[k] = jmax(nodes[self + 1].mx[k], nodes[r].mx[k]);
    }
    return self;
  }
};

// validateTlasBuild (tlas.ts:630-850).
BuildError validate_tlas(const std::vector<uint32_t>& nodes, const std::vector<uint32_t>& perm,
                         uint32_t n_inst, const float* xf, uint32_t max_leaf, uint32_t max_depth) {
  const uint32_t count = (uint32_t)(nodes.size() / 8);
  if (count == 0) return fail(VT_ERR_BUILD, "validateTlasBuild: nodeCount must be positive.");
  std::vector<uint8_t> vn(count), vp(perm.size()), vi(n_inst);
  std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 0u}};
  uint32_t visited = 0;
  while (!st.empty()) {
    auto [ni, depth] = st.back();
    st.pop_back();
    if (ni >= count || vn[ni]) return fail(VT_ERR_BUILD, "validateTlasBuild: bad node reference.");
    if (depth > max_depth) return fail(VT_ERR_BUILD, "validateTlasBuild: node too deep.");
    if (depth + 1 > kTlasStackDepth)
      return fail(VT_ERR_BUILD, "validateTlasBuild: traversal stack exceeds capacity.");
    vn[ni] = 1;
    ++visited;
    const uint32_t* q = &nodes[8 * (size_t)ni];
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(bitsf(q[k])) || !std::isfinite(bitsf(q[3 + k])) ||
          bitsf(q[k]) > bitsf(q[3 + k]))
        return fail(VT_ERR_BUILD, "validateTlasBuild: node has invalid bounds.");
    if ((q[7] >> 16) == 0xFFFFu) {
      const uint32_t c = q[7] & 0xFFFFu, off = q[6];
      if (c == 0 || c > max_leaf || off > perm.size() || c > perm.size() - off)
        return fail(VT_ERR_BUILD, "validateTlasBuild: leaf has invalid range.");
      for (uint32_t i = 0; i < c; ++i) {
        if (vp[off + i]) return fail(VT_ERR_BUILD, "validateTlasBuild: slot referenced twice.");
        vp[off + i] = 1;
        const uint32_t ii = perm[off + i];
        if (ii >= n_inst || vi[ii]) return fail(VT_ERR_BUILD, "validateTlasBuild: bad instance.");
        vi[ii] = 1;
        for (int l = 0; l < 16; ++l)
          if (!std::isfinite(xf[16 * (size_t)ii + l]))
            return fail(VT_ERR_BUILD, "validateTlasBuild: non-finite transform.");
      }
      continue;
    }
    if (depth >= max_depth || q[7] > 2)
      return fail(VT_ERR_BUILD, "validateTlasBuild: invalid interior node.");
    const uint64_t lc = (uint64_t)ni + 1, rc = (uint64_t)ni + q[6];
    if (q[6] <= 1 || lc >= count || rc >= count)
      return fail(VT_ERR_BUILD, "validateTlasBuild: invalid child references.");
    for (uint64_t ch : {lc, rc}) {
      const uint32_t* c = &nodes[8 * ch];
      for (int k = 0; k < 3; ++k)
        if (bitsf(c[k]) < bitsf(q[k]) || bitsf(c[3 + k]) > bitsf(q[3 + k]))
          return fail(VT_ERR_BUILD, "validateTlasBuild: bounds do not enclose child.");
    }
    st.push_back({(uint32_t)rc, depth + 1});
    st.push_back({(uint32_t)lc, depth + 1});
  }
  if (visited != count) return fail(VT_ERR_BUILD, "validateTlasBuild: unreachable nodes.");
  for (uint8_t v : vp)
    if (!v) return fail(VT_ERR_BUILD, "validateTlasBuild: unreferenced permutation slot.");
  return {};
}

}  // namespace

BuildError build_tlas(uint32_t n, const uint32_t* blas_ids, const double* aabb_min,
                      const double* aabb_max, const float* w2l, uint32_t max_leaf,
                      uint32_t num_bins, uint32_t max_depth, TlasOut& out) {
  if (n == 0) return fail(VT_ERR_BUILD, "buildTlas: instances list is empty.");
  if (max_leaf < 1 || max_leaf > 0xFFFFu)
    return fail(VT_ERR_RANGE, "buildTlas: maxLeafInstances must be in [1, 65535].");
  if (num_bins < 2 || num_bins > 256)
    return fail(VT_ERR_RANGE, "buildTlas: numBins must be in [2, 256].");
  if (max_depth > kMaxBuildDepth) return fail(VT_ERR_RANGE, "buildTlas: maxDepth must be in [0, 58].");
  TlasBuilder b;
  b.max_leaf = max_leaf;
  b.K = num_bins;
  b.max_depth = max_depth;
  b.R.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    const double* mn = aabb_min + 3 * (size_t)i;
    const double* mx = aabb_max + 3 * (size_t)i;
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(mn[k]) || !std::isfinite(mx[k]) || !std::isfinite((float)mn[k]) ||
          !std::isfinite((float)mx[k]))
        return fail(VT_ERR_RANGE, "buildTlas: instance aabb must be finite and representable as float32.");
    if (mx[0] < mn[0] || mx[1] < mn[1] || mx[2] < mn[2])
      return fail(VT_ERR_BUILD, "buildTlas: instance has inverted AABB.");
    for (int l = 0; l < 16; ++l)
      if (!std::isfinite(w2l[16 * (size_t)i + l]))
        return fail(VT_ERR_RANGE, "buildTlas: instance has non-finite worldToLocal data.");
    InstRec& r = b.R[i];
    r.orig = i;
    for (int k = 0; k < 3; ++k) r.mn[k] = mn[k], r.mx[k] = mx[k], r.c[k] = (mn[k] + mx[k]) * 0.5;
  }
  if (balanced_depth(n, max_leaf) > max_depth)
    return fail(VT_ERR_RANGE, "buildTlas: instances cannot fit within the requested depth.");
  std::vector<uint32_t> all(n);
  for (uint32_t i = 0; i < n; ++i) all[i] = i;
  b.nodes.reserve(2 * (size_t)n);
  b.build(all, 0);
  if (b.err) return b.err;
  out.nodes.assign(b.nodes.size() * 8, 0u);
  for (size_t i = 0; i < b.nodes.size(); ++i) {
    uint32_t* w = &out.nodes[8 * i];
    for (int k = 0; k < 3; ++k)
      w[k] = fbits((float)b.nodes[i].mn[k]), w[3 + k] = fbits((float)b.nodes[i].mx[k]);
    w[6] = b.nodes[i].right;
    w[7] = b.nodes[i].split;
  }
  out.instance_indices = b.perm;
  out.blas_roots.assign(blas_ids, blas_ids + n);
  out.transforms.assign(w2l, w2l + 16 * (size_t)n);
  std::memcpy(out.fallbacks, b.fallbacks, sizeof b.fallbacks);
  return validate_tlas(out.nodes, out.instance_indices, n, w2l, max_leaf, max_depth);
}

// invertMat4 (packages/shared-bvh/src/mathUtils.ts:21-105).
bool invert_mat4(const float* mf, float* out) {
  double a[16];
  for (int i = 0; i < 16; ++i) {
    a[i] = mf[i];
    if (!std::isfinite(a[i])) return false;
  }
  const double b00 = a[0] * a[5] - a[1] * a[4], b01 = a[0] * a[6] - a[2] * a[4];
  const double b02 = a[0] * a[7] - a[3] * a[4], b03 = a[1] * a[6] - a[2] * a[5];
  const double b04 = a[1] * a[7] - a[3] * a[5], b05 = a[2] * a[7] - a[3] * a[6];
  const double b06 = a[8] * a[13] - a[9] * a[12], b07 = a[8] * a[14] - a[10] * a[12];
  const double b08 = a[8] * a[15] - a[11] * a[12], b09 = a[9] * a[14] - a[10] * a[13];
  const double b10 = a[9] * a[15] - a[11] * a[13], b11 = a[10] * a[15] - a[11] * a[14];
  const double det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
  if (!std::isfinite(det) || det == 0.0) return false;
  const double inv = 1.0 / det;
  float o[16];
  o[0] = (float)((a[5] * b11 - a[6] * b10 + a[7] * b09) * inv);
  o[1] = (float)((-a[1] * b11 + a[2] * b10 - a[3] * b09) * inv);
  o[2] = (float)((a[13] * b05 - a[14] * b04 + a[15] * b03) * inv);
  o[3] = (float)((-a[9] * b05 + a[10] * b04 - a[11] * b03) * inv);
  o[4] = (float)((-a[4] * b11 + a[6] * b08 - a[7] * b07) * inv);
  o[5] = (float)((a[0] * b11 - a[2] * b08 + a[3] * b07) * inv);
  o[6] = (float)((-a[12] * b05 + a[14] * b02 - a[15] * b01) * inv);
  o[7] = (float)((a[8] * b05 - a[10] * b02 + a[11] * b01) * inv);
  o[8] = (float)((a[4] * b10 - a[5] * b08 + a[7] * b06) * inv);
  o[9] = (float)((-a[0] * b10 + a[1] * b08 - a[3] * b06) * inv);
  o[10] = (float)((a[12] * b04 - a[13] * b02 + a[15] * b00) * inv);
  o[11] = (float)((-a[8] * b04 + a[9] * b02 - a[11] * b00) * inv);
  o[12] = (float)((-a[4] * b09 + a[5] * b07 - a[6] * b06) * inv);
  o[13] = (float)((a[0] * b09 - a[1] * b07 + a[2] * b06) * inv);
  o[14] = (float)((-a[12] * b03 + a[13] * b01 - a[14] * b00) * inv);
  o[15] = (float)((a[8] * b03 - a[9] * b01 + a[10] * b00) * inv);
  for (float v : o)
    if (!std::isfinite(v)) return false;
  for (int pass = 0; pass < 2; ++pass)
    for (int col = 0; col < 4; ++col)
      for (int row = 0; row < 4; ++row) {
        double prod = 0, sum_abs = 0;
        for (int k = 0; k < 4; ++k) {
          const double l = pass == 0 ? a[k * 4 + row] : (double)o[k * 4 + row];
          const double r = pass == 0 ? (double)o[col * 4 + k] : a[col * 4 + k];
          prod += l * r;
          sum_abs += std::fabs(l * r);
        }
        const double tol = 1e-5 * std::max(1.0, sum_abs);
        if (!std::isfinite(prod) || std::fabs(prod - (row == col ? 1.0 : 0.0)) > tol) return false;
      }
  std::memcpy(out, o, sizeof o);
  return true;
}

// ---------------------------------------------------------------------------
// Scene packing (scenePack.ts:1104-1307)
// ---------------------------------------------------------------------------

namespace {

const float kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

struct Slice {
  bool present = false;
  uint32_t vcount = 0, tcount = 0;
  std::vector<float> lpos, lnrm;
  BlasOut blas;
  double lmin[3], lmax[3];
  BuildError err;
  bool deferred = false;  // BLAS left to the large-mesh builder
  std::vector<uint32_t> lidx, lmat;
};

void build_slice(const PrimitiveIn& p, Slice& s, uint32_t defer_min) {
  const uint32_t vcount = p.position_floats / 3;
  if (vcount < 3) return;
  const uint32_t idx_len = p.indices ? p.index_count : vcount;
  const uint32_t tcount = idx_len / 3;
  if (tcount == 0) return;
  s.vcount = vcount;
  s.tcount = tcount;
  s.lpos.assign((size_t)vcount * 4, 0.f);
  s.lnrm.assign((size_t)vcount * 4, 0.f);
  for (uint32_t i = 0; i < vcount; ++i)
    for (int k = 0; k < 3; ++k) {
      s.lpos[4 * (size_t)i + k] = p.positions[3 * (size_t)i + k];
      const size_t o = 3 * (size_t)i + k;
      s.lnrm[4 * (size_t)i + k] = o < p.normal_floats ? p.normals[o] : (k == 1 ? 1.f : 0.f);
    }
  std::vector<uint32_t> lidx((size_t)tcount * 4, 0u), lmat(tcount, p.material);
  for (uint32_t t = 0; t < tcount; ++t)
    for (int k = 0; k < 3; ++k)
      lidx[4 * (size_t)t + k] = p.indices ? p.indices[3 * (size_t)t + k] : 3 * t + k;
  if (defer_min && tcount >= defer_min) {
    s.deferred = true;
    s.lidx.swap(lidx);
    s.lmat.swap(lmat);
  } else {
    s.err = build_blas(s.lpos.data(), vcount * 4, 4, lidx.data(), tcount * 4, 4, lmat.data(), 4, 16,
                       kMaxBuildDepth, s.blas);
  }
  local_aabb(s.lpos.data(), vcount, s.lmin, s.lmax);
  s.present = true;
}

// transformPoint (scenePack.ts:161-170), binary64.
void xform_point(const float* m, const double* p, double* q) {
  double t[4];
  for (int r = 0; r < 4; ++r)
    t[r] = (double)m[r] * p[0] + (double)m[4 + r] * p[1] + (double)m[8 + r] * p[2] + (double)m[12 + r];
  if (std::fabs(t[3]) > 1e-8)
    q[0] = t[0] / t[3], q[1] = t[1] / t[3], q[2] = t[2] / t[3];
  else
    q[0] = t[0], q[1] = t[1], q[2] = t[2];
}

}  // namespace

void local_aabb(const float* pos4, uint32_t count, double* mn, double* mx) {
  for (int k = 0; k < 3; ++k) mn[k] = INFINITY, mx[k] = -INFINITY;
  for (uint32_t i = 0; i < count; ++i)
    for (int k = 0; k < 3; ++k) {
      mn[k] = jmin(mn[k], pos4[4 * (size_t)i + k]);
      mx[k] = jmax(mx[k], pos4[4 * (size_t)i + k]);
    }
}

void world_aabb(const float* l2w, const double* lmn, const double* lmx, double* mn, double* mx) {
  for (int k = 0; k < 3; ++k) mn[k] = INFINITY, mx[k] = -INFINITY;
  for (int c = 0; c < 8; ++c) {
    const double corner[3] = {(c & 1) ? lmx[0] : lmn[0], (c & 2) ? lmx[1] : lmn[1], (c & 4) ? lmx[2] : lmn[2]};
    double q[3];
    xform_point(l2w, corner, q);
    for (int k = 0; k < 3; ++k) mn[k] = jmin(mn[k], q[k]), mx[k] = jmax(mx[k], q[k]);
  }
}

BuildError refit_tlas_instances(std::vector<uint32_t>& nodes, const std::vector<uint32_t>& perm,
                                uint32_t inst_count, const double* aabb_min, const double* aabb_max,
                                const std::vector<uint32_t>& changed) {
  auto fail = [](int code, const std::string& m) { return BuildError{code, m}; };
  for (uint32_t i : perm) {
    if (i >= inst_count) return fail(VT_ERR_BUILD, "refitTlas: instance reference out of range");
    const double* a = aabb_min + 3 * (size_t)i;
    const double* b = aabb_max + 3 * (size_t)i;
    if (b[0] < a[0] || b[1] < a[1] || b[2] < a[2])
      return fail(VT_ERR_BUILD, "refitTlas: instance " + std::to_string(i) + " has inverted AABB.");
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(a[k]) || !std::isfinite(b[k]))
        return fail(VT_ERR_RANGE, "refitTlas: instance " + std::to_string(i) + " has non-finite AABB.");
  }
  if (changed.empty()) return {};
  const uint32_t n = (uint32_t)(nodes.size() / 8);
  std::vector<uint8_t> hit(inst_count, 0), affected(n, 0);
  for (uint32_t i : changed) {
    if (i >= inst_count) return fail(VT_ERR_RANGE, "refitTlas: changed instance out of range");
    hit[i] = 1;
  }
  // Preorder: children follow their parent, so a parent array built in one
  // pass lets every marked leaf flag its ancestors.
  std::vector<uint32_t> parent(n, UINT32_MAX);
  for (uint32_t node = 0; node < n; ++node) {
    const uint32_t* q = &nodes[8 * (size_t)node];
    if ((q[7] >> 16) == 0xFFFFu) {
      const uint32_t off = q[6], cnt = q[7] & 0xFFFFu;
      bool touched = false;
      for (uint32_t k = 0; k < cnt && off + k < perm.size(); ++k) touched = touched || hit[perm[off + k]];
      for (uint32_t c = node; touched && c != UINT32_MAX && !affected[c]; c = parent[c]) affected[c] = 1;
      continue;
    }
    parent[node + 1] = node;
    parent[node + q[6]] = node;
  }
  // Children carry higher indices than their parent: walking down from the
  // last node refits every child before its parent.
  for (uint32_t node = n; node-- > 0;) {
    if (!affected[node]) continue;
    uint32_t* q = &nodes[8 * (size_t)node];
    if ((q[7] >> 16) == 0xFFFFu) {
      double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (uint32_t k = 0; k < (q[7] & 0xFFFFu); ++k) {
        const uint32_t orig = perm[q[6] + k];
        for (int a = 0; a < 3; ++a) {
          if (aabb_min[3 * (size_t)orig + a] < mn[a]) mn[a] = aabb_min[3 * (size_t)orig + a];
          if (aabb_max[3 * (size_t)orig + a] > mx[a]) mx[a] = aabb_max[3 * (size_t)orig + a];
        }
      }
      for (int a = 0; a < 3; ++a) q[a] = fbits((float)mn[a]), q[3 + a] = fbits((float)mx[a]);
    } else {
      const uint32_t* l = &nodes[8 * ((size_t)node + 1)];
      const uint32_t* r = &nodes[8 * ((size_t)node + q[6])];
      for (int a = 0; a < 3; ++a) {
        q[a] = fbits(jminf(bitsf(l[a]), bitsf(r[a])));
        q[3 + a] = fbits(jmaxf(bitsf(l[3 + a]), bitsf(r[3 + a])));
      }
    }
  }
  return {};
}

BuildError pack_scene(const PrimitiveIn* prims, uint32_t prim_count, uint32_t threads, Pack& out,
                      const BlasBuilder* large, uint32_t large_min_triangles) {
  out = Pack();
  std::vector<Slice> slices(prim_count);
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  threads = std::min<uint32_t>(threads, std::max<uint32_t>(1, prim_count));
  std::atomic<uint32_t> next{0};
  auto worker = [&]() {
    for (uint32_t i; (i = next.fetch_add(1)) < prim_count;)
      build_slice(prims[i], slices[i], large ? std::max(1u, large_min_triangles) : 0u);
  };
  if (threads <= 1) {
    worker();
  } else {
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
  }
  std::vector<Slice*> deferred;
  for (Slice& s : slices)
    if (s.deferred) deferred.push_back(&s);
  if (!deferred.empty()) {
    std::vector<BlasJobIn> jobs;
    for (Slice* s : deferred)
      jobs.push_back({s->lpos.data(), s->vcount * 4, s->lidx.data(), s->tcount * 4, s->lmat.data()});
    std::vector<BlasOut> outs;
    const bool ok = !(*large)(jobs, outs) && outs.size() == deferred.size();
    for (size_t i = 0; i < deferred.size(); ++i) {
      Slice& s = *deferred[i];
      if (ok)
        s.blas.nodes.swap(outs[i].nodes), s.blas.reordered_indices.swap(outs[i].reordered_indices),
            s.blas.reordered_mat.swap(outs[i].reordered_mat),
            s.blas.reordered_to_source.swap(outs[i].reordered_to_source);
      else
        s.err = build_blas(s.lpos.data(), s.vcount * 4, 4, s.lidx.data(), s.tcount * 4, 4, s.lmat.data(), 4, 16,
                           kMaxBuildDepth, s.blas);
      std::vector<uint32_t>().swap(s.lidx);
      std::vector<uint32_t>().swap(s.lmat);
    }
  }
  std::vector<double> inst_mn, inst_mx;
  std::vector<uint32_t> inst_root, roots;
  for (int k = 0; k < 3; ++k) out.world_min[k] = INFINITY, out.world_max[k] = -INFINITY;
  for (uint32_t pi = 0; pi < prim_count; ++pi) {
    const PrimitiveIn& p = prims[pi];
    Slice& s = slices[pi];
    if (!s.present) {
      out.warnings.push_back("Primitive " + std::to_string(pi) + " has no triangles; skipping.");
      continue;
    }
    if (s.err) return s.err;
    if (p.kind == 1 && p.transform_count == 0) {
      out.warnings.push_back("Primitive " + std::to_string(pi) +
                             " is an instanced-mesh with zero instances; skipping.");
      continue;
    }
    const uint32_t vbase = (uint32_t)(out.positions.size() / 4);
    const uint32_t tbase = (uint32_t)out.tri_material.size();
    const uint32_t nbase = (uint32_t)(out.bvh_nodes.size() / 8);
    const uint32_t ntx = p.kind == 1 ? p.transform_count : 1;
    uint32_t survivors = 0;
    for (uint32_t si = 0; si < ntx; ++si) {
      const float* l2w = (p.kind == 1 || (p.transforms && p.transform_count))
                             ? p.transforms + 16 * (size_t)si
                             : kIdentity;
      float w2l[16];
      if (!invert_mat4(l2w, w2l)) {
        out.warnings.push_back("Primitive " + std::to_string(pi) +
                               " has non-invertible instance transform; skipping this instance.");
        continue;
      }
      double mn[3], mx[3];
      world_aabb(l2w, s.lmin, s.lmax, mn, mx);
      for (int k = 0; k < 3; ++k) {
        inst_mn.push_back(mn[k]);
        inst_mx.push_back(mx[k]);
        out.world_min[k] = std::min(out.world_min[k], mn[k]);
        out.world_max[k] = std::max(out.world_max[k], mx[k]);
      }
      out.tlas_w2l.insert(out.tlas_w2l.end(), w2l, w2l + 16);
      out.tlas_l2w.insert(out.tlas_l2w.end(), l2w, l2w + 16);
      inst_root.push_back(nbase);
      out.instance_primitive.push_back(pi);
      out.instance_source.push_back(si);
      ++survivors;
    }
    if (survivors == 0) {
      out.warnings.push_back("Primitive " + std::to_string(pi) +
                             " has no invertible placements; skipping its geometry.");
      continue;
    }
    out.positions.insert(out.positions.end(), s.lpos.begin(), s.lpos.end());
    out.normals.insert(out.normals.end(), s.lnrm.begin(), s.lnrm.end());
    for (uint32_t t = 0; t < s.tcount; ++t) {
      const uint32_t* q = &s.blas.reordered_indices[4 * (size_t)t];
      out.indices.insert(out.indices.end(), {q[0] + vbase, q[1] + vbase, q[2] + vbase, q[3]});
      out.tri_material.push_back(s.blas.reordered_mat[t]);
      out.tri_source.push_back(s.blas.reordered_to_source[t]);
      out.tri_primitive.push_back(pi);
    }
    const size_t n0 = out.bvh_nodes.size();
    out.bvh_nodes.insert(out.bvh_nodes.end(), s.blas.nodes.begin(), s.blas.nodes.end());
    for (size_t w = n0; w < out.bvh_nodes.size(); w += 8)
      if ((out.bvh_nodes[w + 7] >> 16) == 0xFFFFu) out.bvh_nodes[w + 6] += tbase;
    Binding b{nbase, survivors, vbase, s.vcount, tbase, s.tcount, pi, p.kind, {}, {}};
    for (int k = 0; k < 3; ++k) b.local_min[k] = (float)s.lmin[k], b.local_max[k] = (float)s.lmax[k];
    out.bindings.push_back(b);
    roots.push_back(nbase);
  }
  const uint32_t n_inst = (uint32_t)inst_root.size();
  if (n_inst > 0) {
    TlasOut t;
    BuildError e = build_tlas(n_inst, inst_root.data(), inst_mn.data(), inst_mx.data(),
                              out.tlas_w2l.data(), 1, 16, kMaxBuildDepth, t);
    if (e) return e;
    out.tlas_nodes = std::move(t.nodes);
    out.tlas_instance_indices = std::move(t.instance_indices);
    out.tlas_blas_roots = std::move(t.blas_roots);
  }
  if (!roots.empty()) {
    BuildError e = validate_bvh(out.bvh_nodes.data(), (uint32_t)(out.bvh_nodes.size() / 8),
                                roots.data(), (uint32_t)roots.size(),
                                (int64_t)out.tri_material.size(), kBvhStackDepth);
    if (e) return e;
  }
  return {};
}

}  // namespace vt
