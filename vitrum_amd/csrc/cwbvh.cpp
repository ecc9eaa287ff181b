// cwbvh.cpp -- host collapse of the reference binary BVH into the 8-wide
// compressed forest of pt-webgpu's `cwbvh-closest` mode, and the engine's
// switch between the binary and the wide traversal.
//
// Collapse rules (packages/shared-bvh/src/compressedWideBvh.ts:180-418): a
// wide node takes the bounds of the binary node it starts from; its child
// slots start as the binary node's two children and the interior slot of
// largest surface area (first on ties) is replaced in place by its two
// children until 8 slots are used or only leaves remain; wide nodes are
// allocated in preorder; child bounds are quantized to 16 bits of the parent
// box, rounded outward by one step (floor - 1, ceil + 1).  A tree whose child
// boxes leave their parent, or whose depth-first schedule needs more than 64
// stack entries, is published with an invalid child count so the device
// falls back to the binary traversal.  Binary64 arithmetic, as the
// reference's JavaScript numbers.
#include "cwbvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "engine_internal.h"

namespace vtw {
namespace {

struct BNode {
  double lo[3], hi[3];
  uint32_t w6, w7;
};

BNode read(const uint32_t* w, uint32_t i) {
  BNode n;
  for (int k = 0; k < 3; ++k) {
    float a, b;
    std::memcpy(&a, &w[8 * (size_t)i + k], 4);
    std::memcpy(&b, &w[8 * (size_t)i + 3 + k], 4);
    n.lo[k] = a;
    n.hi[k] = b;
  }
  n.w6 = w[8 * (size_t)i + 6];
  n.w7 = w[8 * (size_t)i + 7];
  return n;
}

bool leaf(uint32_t w7) { return (w7 >> 16) == 0xFFFFu; }

// Unit offset of value inside [lo, hi], overflow-safe; false when undefined.
bool unit(double v, double lo, double hi, double* out) {
  if (!std::isfinite(v) || !std::isfinite(lo) || !std::isfinite(hi) || lo > hi) return false;
  if (hi == lo) {
    *out = 0;
    return true;
  }
  const double ext = hi - lo;
  if (std::isfinite(ext)) {
    *out = (v - lo) / ext;
    return true;
  }
  const double half = hi * 0.5 - lo * 0.5;
  if (!(half > 0) || !std::isfinite(half)) return false;
  *out = (v * 0.5 - lo * 0.5) / half;
  return true;
}

uint32_t quant(double v, double lo, double hi, bool up) {
  double u;
  if (!unit(v, lo, hi, &u)) return 0;
  if (hi == lo) return up ? 65535u : 0u;
  const double q = up ? std::ceil(u * 65535) + 1 : std::floor(u * 65535) - 1;
  if (!std::isfinite(q)) return 0;
  return (uint32_t)std::min(65535.0, std::max(0.0, q));
}

double area(const BNode& n) {
  const double dx = std::max(0.0, n.hi[0] - n.lo[0]), dy = std::max(0.0, n.hi[1] - n.lo[1]),
               dz = std::max(0.0, n.hi[2] - n.lo[2]);
  return dx * dy + dy * dz + dz * dx;
}

struct Builder {
  const uint32_t* w;
  Forest& f;
  bool valid = true;

  uint32_t alloc(uint32_t bin) {
    const uint32_t i = (uint32_t)f.child_count.size();
    const BNode n = read(w, bin);
    for (int k = 0; k < 3; ++k) f.node_bounds.push_back((float)n.lo[k]);
    for (int k = 0; k < 3; ++k) f.node_bounds.push_back((float)n.hi[k]);
    f.child_count.push_back(0);
    f.child_bounds.resize(f.child_bounds.size() + kKids * 3, 0);
    f.child_meta.resize(f.child_meta.size() + kKids * 3, 0);
    return i;
  }

  uint32_t build(uint32_t bin) {
    const BNode parent = read(w, bin);
    const uint32_t wide = alloc(bin);
    uint32_t kids[kKids], n = 0;
    if (leaf(parent.w7)) {
      kids[n++] = bin;
    } else {
      kids[n++] = bin + 1;
      kids[n++] = bin + parent.w6;
      while (n < kKids) {
        int at = -1;
        double best = -1;
        for (uint32_t i = 0; i < n; ++i) {
          const BNode c = read(w, kids[i]);
          if (leaf(c.w7)) continue;
          const double a = area(c);
          if (a > best) best = a, at = (int)i;
        }
        if (at < 0) break;
        const uint32_t open = kids[at];
        const BNode o = read(w, open);
        std::memmove(kids + at + 2, kids + at + 1, (n - (uint32_t)at - 1) * sizeof(uint32_t));
        kids[at] = open + 1;
        kids[at + 1] = open + o.w6;
        ++n;
      }
    }
    f.child_count[wide] = n;
    for (uint32_t s = 0; s < n; ++s) {
      const BNode c = read(w, kids[s]);
      uint32_t q[6];
      for (int a = 0; a < 3; ++a) {
        double t;
        if (c.lo[a] < parent.lo[a] || c.hi[a] > parent.hi[a] || c.lo[a] > c.hi[a] ||
            !unit(c.lo[a], parent.lo[a], parent.hi[a], &t) || !unit(c.hi[a], parent.lo[a], parent.hi[a], &t))
          valid = false;
        q[a] = quant(c.lo[a], parent.lo[a], parent.hi[a], false);
        q[3 + a] = quant(c.hi[a], parent.lo[a], parent.hi[a], true);
      }
      uint32_t* b = &f.child_bounds[((size_t)wide * kKids + s) * 3];
      b[0] = q[0] | (q[1] << 16);
      b[1] = q[2] | (q[3] << 16);
      b[2] = q[4] | (q[5] << 16);
      uint32_t m0, m1, m2;
      if (leaf(c.w7)) {
        m0 = 2;
        m1 = c.w6;
        m2 = c.w7 & 0xFFFFu;
      } else {
        m0 = 1;
        m1 = build(kids[s]);
        m2 = 0;
      }
      uint32_t* m = &f.child_meta[((size_t)wide * kKids + s) * 3];  // after the recursion grew the array
      m[0] = m0;
      m[1] = m1;
      m[2] = m2;
    }
    return wide;
  }
};

// Largest depth-first stack of a wide tree (compressedWideBvh.ts:60-92).
uint32_t stack_need(const Forest& f, uint32_t root, uint32_t end) {
  std::vector<uint32_t> stack{root};
  uint32_t most = 1;
  while (!stack.empty()) {
    const uint32_t node = stack.back();
    stack.pop_back();
    for (uint32_t s = 0; s < f.child_count[node]; ++s) {
      const uint32_t* m = &f.child_meta[((size_t)node * kKids + s) * 3];
      if (m[0] != 1) continue;
      if (m[1] >= end || m[1] <= node) return UINT32_MAX;  // not a preorder tree
      stack.push_back(m[1]);
    }
    most = std::max<uint32_t>(most, (uint32_t)stack.size());
  }
  return most;
}

}  // namespace

Forest collapse(const uint32_t* nodes, uint32_t node_count, const std::vector<uint32_t>& roots,
                uint32_t tri_count) {
  Forest f;
  bool layout_ok = true;
  for (uint32_t i = 0; i < node_count && layout_ok; ++i) {
    const BNode n = read(nodes, i);
    for (int a = 0; a < 3; ++a) {
      double t;
      if (!unit(n.lo[a], n.lo[a], n.hi[a], &t)) layout_ok = false;
    }
  }
  for (uint32_t r : roots) {
    const uint32_t first = (uint32_t)f.child_count.size();
    f.roots.push_back(first);
    const BNode rn = read(nodes, r);
    if (tri_count == 0 || !layout_ok || (leaf(rn.w7) && (rn.w7 & 0xFFFFu) == 0)) {
      Builder b{nodes, f};
      const uint32_t i = b.alloc(r);
      for (int k = 0; k < 6; ++k) f.node_bounds[6 * (size_t)i + k] = 0.0f;
      f.child_count[i] = layout_ok ? 0u : kInvalidCount;
      continue;
    }
    Builder b{nodes, f};
    b.build(r);
    const uint32_t end = (uint32_t)f.child_count.size();
    if (!b.valid || stack_need(f, first, end) > 64)
      for (uint32_t i = first; i < end; ++i) f.child_count[i] = kInvalidCount;
  }
  return f;
}

}  // namespace vtw

// ---- engine integration -----------------------------------------------------

namespace {
vt_status cw_err(vt_engine* e, vt_status code, const std::string& msg) {
  if (e) e->last_error = msg;
  return code;
}
}  // namespace

struct CwbvhState {
  DeviceBuffer tlas_nodes, inst_indices, w2l, l2w, blas_roots, wide_roots, bvh_nodes;
  DeviceBuffer node_bounds, child_bounds, child_meta, child_count, single_sided, inst_reversed;
  vtw::Scene view{};
};

void engine_release_cwbvh(vt_engine* e) {
  if (!e->cwbvh) return;
  for (DeviceBuffer* b : {&e->cwbvh->tlas_nodes, &e->cwbvh->inst_indices, &e->cwbvh->w2l, &e->cwbvh->l2w,
                          &e->cwbvh->blas_roots, &e->cwbvh->wide_roots, &e->cwbvh->bvh_nodes,
                          &e->cwbvh->node_bounds, &e->cwbvh->child_bounds, &e->cwbvh->child_meta,
                          &e->cwbvh->child_count, &e->cwbvh->single_sided, &e->cwbvh->inst_reversed})
    b->release();
  delete e->cwbvh;
  e->cwbvh = nullptr;
}

// Builds (once per scene) the wide forest and the reference-layout buffers the
// wide traversal reads.
vt_status engine_cwbvh_ensure(vt_engine* e) {
  if (e->cwbvh) return VT_OK;
  if (!e->has_scene || !e->pack) return cw_err(e, VT_ERR_STATE, "no scene is set");
  const vt::Pack& P = *e->pack;
  const uint32_t n_nodes = (uint32_t)(P.bvh_nodes.size() / 8);
  const uint32_t n_inst = (uint32_t)P.tlas_blas_roots.size();
  std::vector<uint32_t> uniq(P.tlas_blas_roots.begin(), P.tlas_blas_roots.end());
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  for (uint32_t r : uniq)
    if (r >= n_nodes) return cw_err(e, VT_ERR_BUILD, "instance BLAS root outside the node array");
  vtw::Forest f = vtw::collapse(P.bvh_nodes.data(), n_nodes, uniq, (uint32_t)(P.indices.size() / 4));
  std::vector<uint32_t> wide_root(std::max<uint32_t>(1, n_inst), 0);
  for (uint32_t i = 0; i < n_inst; ++i) {
    const size_t k = (size_t)(std::lower_bound(uniq.begin(), uniq.end(), P.tlas_blas_roots[i]) - uniq.begin());
    wide_root[i] = f.roots[k];
  }
  std::unique_ptr<CwbvhState> st(new CwbvhState());
  auto up = [&](DeviceBuffer& b, const void* src, size_t bytes) {
    const uint32_t zero[4] = {0, 0, 0, 0};
    return engine_upload(e, b, bytes ? src : zero, bytes ? bytes : sizeof zero);
  };
  vt_status s = VT_OK;
  s = s ? s : up(st->tlas_nodes, P.tlas_nodes.data(), P.tlas_nodes.size() * 4);
  s = s ? s : up(st->inst_indices, P.tlas_instance_indices.data(), P.tlas_instance_indices.size() * 4);
  s = s ? s : up(st->w2l, P.tlas_w2l.data(), P.tlas_w2l.size() * 4);
  s = s ? s : up(st->l2w, P.tlas_l2w.data(), P.tlas_l2w.size() * 4);
  s = s ? s : up(st->blas_roots, P.tlas_blas_roots.data(), P.tlas_blas_roots.size() * 4);
  s = s ? s : up(st->wide_roots, wide_root.data(), wide_root.size() * 4);
  s = s ? s : up(st->bvh_nodes, P.bvh_nodes.data(), P.bvh_nodes.size() * 4);
  s = s ? s : up(st->node_bounds, f.node_bounds.data(), f.node_bounds.size() * 4);
  s = s ? s : up(st->child_bounds, f.child_bounds.data(), f.child_bounds.size() * 4);
  s = s ? s : up(st->child_meta, f.child_meta.data(), f.child_meta.size() * 4);
  s = s ? s : up(st->child_count, f.child_count.data(), f.child_count.size() * 4);
  std::vector<uint8_t> single(P.indices.size() / 4, 1u), reversed(std::max<uint32_t>(1, n_inst), 0u);
  if (e->layout)
    for (size_t t = 0; t < single.size() && t < e->layout->tri_count; ++t)
      single[t] = e->layout->tris[t * vt::kTriWords + 9] > 0.5f ? 1u : 0u;
  for (uint32_t i = 0; i < n_inst; ++i) reversed[i] = vt::orientation_reversed(&P.tlas_l2w[16 * (size_t)i]) ? 1u : 0u;
  s = s ? s : up(st->single_sided, single.data(), single.size());
  s = s ? s : up(st->inst_reversed, reversed.data(), reversed.size());
  if (s) {
    e->cwbvh = st.release();
    engine_release_cwbvh(e);
    return s;
  }
  vtw::Scene& v = st->view;
  v.tlas_nodes = (const uint32_t*)st->tlas_nodes.ptr;
  v.tlas_node_count = (uint32_t)(P.tlas_nodes.size() / 8);
  v.inst_indices = (const uint32_t*)st->inst_indices.ptr;
  v.inst_index_count = (uint32_t)P.tlas_instance_indices.size();
  v.w2l = (const float*)st->w2l.ptr;
  v.l2w = (const float*)st->l2w.ptr;
  v.inst_count = n_inst;
  v.blas_roots = (const uint32_t*)st->blas_roots.ptr;
  v.wide_roots = (const uint32_t*)st->wide_roots.ptr;
  v.bvh_nodes = (const uint32_t*)st->bvh_nodes.ptr;
  v.bvh_node_count = n_nodes;
  v.positions = (const float4*)e->scene.positions.ptr;
  v.vertex_count = (uint32_t)(P.positions.size() / 4);
  v.indices = (const uint4*)e->scene.indices.ptr;
  v.tri_count = (uint32_t)(P.indices.size() / 4);
  v.node_bounds = (const float*)st->node_bounds.ptr;
  v.child_bounds = (const uint32_t*)st->child_bounds.ptr;
  v.child_meta = (const uint32_t*)st->child_meta.ptr;
  v.child_count = (const uint32_t*)st->child_count.ptr;
  v.wide_count = (uint32_t)f.child_count.size();
  v.single_sided = (const uint8_t*)st->single_sided.ptr;
  v.inst_reversed = (const uint8_t*)st->inst_reversed.ptr;
  const hipError_t r = hipStreamSynchronize(e->stream);
  e->cwbvh = st.release();
  if (r != hipSuccess) return cw_err(e, VT_ERR_HIP, hipGetErrorString(r));
  return VT_OK;
}

const vtw::Scene* engine_cwbvh_view(vt_engine* e) { return e->cwbvh ? &e->cwbvh->view : nullptr; }

vt_status engine_trace_cwbvh(vt_engine* e, const float4* d_rays, uint32_t n, const uint32_t* d_n,
                             uint32_t* d_hits, float* d_hit_pos, uint32_t sided) {
  if (n == 0) return VT_OK;
  if (vt_status s = engine_cwbvh_ensure(e)) return s;
  vtw::Params p{};
  p.s = e->cwbvh->view;
  p.rays = d_rays;
  p.hits = d_hits;
  p.hit_pos = d_hit_pos;
  p.n = n;
  p.n_ptr = d_n;
  p.sided = sided;
  // One ray per lane, grid-stride: enough blocks to cover the rays, capped at
  // 16 blocks of 128 per CU.
  const uint32_t cap = std::max<uint32_t>(1, e->cu_count) * 16u;
  const uint32_t grid = std::min<uint32_t>((n + 127u) / 128u, cap);
  const hipError_t r = vtw::launch_trace(p, grid, e->stream);
  if (r != hipSuccess) return cw_err(e, VT_ERR_HIP, hipGetErrorString(r));
  return VT_OK;
}

extern "C" vt_status vt_set_bvh_traversal(vt_engine* e, uint32_t mode) {
  if (!e) return VT_ERR_INVALID_ARGUMENT;
  if (mode > 1) return cw_err(e, VT_ERR_RANGE, "bvh traversal must be 0 (binary) or 1 (cwbvh-closest)");
  e->opts.bvh_traversal = mode;
  return VT_OK;
}
