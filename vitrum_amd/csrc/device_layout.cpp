// device_layout.cpp -- converts a packed scene into the Node64 HBM layout.
#include "device_layout.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "builder.h"

namespace vt {
namespace {

inline uint32_t fbits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// Converts one preorder 32-byte node forest (roots given) into Node64s.
// src64, when given, receives the reference node behind each Node64 child
// slot (2 per Node64, kNoNode for an empty slot).
std::string convert_forest(const std::vector<uint32_t>& ref_nodes, const std::vector<uint32_t>& roots,
                           std::vector<uint32_t>& out, std::vector<uint32_t>& wrapper_of_root,
                           std::vector<uint32_t>* src64 = nullptr) {
  const uint32_t count = (uint32_t)(ref_nodes.size() / 8);
  std::vector<uint32_t> map(count, ~0u);
  uint32_t next = 0;
  for (uint32_t i = 0; i < count; ++i)
    if ((ref_nodes[8 * (size_t)i + 7] >> 16) != 0xFFFFu) map[i] = next++;
  const uint32_t interior = next;
  out.assign((size_t)(interior + roots.size()) * kNode64Words, 0u);
  if (src64) src64->assign((size_t)(interior + roots.size()) * 2, kNoNode);
  std::string err;
  auto child_ref = [&](uint32_t j) -> uint32_t {
    const uint32_t* q = &ref_nodes[8 * (size_t)j];
    if ((q[7] >> 16) != 0xFFFFu) return map[j];
    const uint32_t c = q[7] & 0xFFFFu, off = q[6];
    if (c == 0) return kEmptyRef;
    if (c > kMaxLeafCountInRef || off > kMaxLeafOffsetInRef) {
      err = "leaf with " + std::to_string(c) + " primitives at offset " + std::to_string(off) +
            " exceeds the device leaf encoding (count <= 15, offset < 2^27)";
      return kEmptyRef;
    }
    return kRefLeaf | (c << 27) | off;
  };
  auto put_box = [&](uint32_t* w, uint32_t j) {
    std::memcpy(w, &ref_nodes[8 * (size_t)j], 24);
  };
  auto put_empty_box = [&](uint32_t* w) {
    for (int k = 0; k < 3; ++k) w[k] = fbits(INFINITY), w[3 + k] = fbits(-INFINITY);
  };
  for (uint32_t i = 0; i < count; ++i) {
    if (map[i] == ~0u) continue;
    const uint32_t* q = &ref_nodes[8 * (size_t)i];
    const uint32_t l = i + 1, r = i + q[6];
    uint32_t* w = &out[(size_t)map[i] * kNode64Words];
    put_box(w, l);
    put_box(w + 6, r);
    w[12] = child_ref(l);
    w[13] = child_ref(r);
    if (src64) (*src64)[2 * (size_t)map[i]] = l, (*src64)[2 * (size_t)map[i] + 1] = r;
  }
  wrapper_of_root.resize(roots.size());
  for (size_t k = 0; k < roots.size(); ++k) {
    const uint32_t idx = interior + (uint32_t)k;
    uint32_t* w = &out[(size_t)idx * kNode64Words];
    put_box(w, roots[k]);
    put_empty_box(w + 6);
    w[12] = child_ref(roots[k]);
    w[13] = kEmptyRef;
    if (src64) (*src64)[2 * (size_t)idx] = roots[k];
    wrapper_of_root[k] = idx;
  }
  return err;
}

}  // namespace

bool orientation_reversed(const float* m) {
  auto amax = [](float a, float b, float c) {
    const float x = std::fabs(a), y = std::fabs(b), z = std::fabs(c);
    const float yz = y > z ? y : z;
    return x > yz ? x : yz;
  };
  const float s0 = amax(m[0], m[1], m[2]), s1 = amax(m[4], m[5], m[6]), s2 = amax(m[8], m[9], m[10]);
  const float big = 3.402823e38f;
  if (!(s0 > 0.0f) || s0 > big || !(s1 > 0.0f) || s1 > big || !(s2 > 0.0f) || s2 > big) return false;
  const float ax = m[0] / s0, ay = m[1] / s0, az = m[2] / s0;
  const float bx = m[4] / s1, by = m[5] / s1, bz = m[6] / s1;
  const float cx = m[8] / s2, cy = m[9] / s2, cz = m[10] / s2;
  const float xx = ay * bz - az * by, xy = az * bx - ax * bz, xz = ax * by - ay * bx;
  const float det = (xx * cx + xy * cy) + xz * cz;
  return !(det >= 0.0f);
}

namespace {

struct Slot {
  uint32_t ref;
  float box[6];
  uint32_t src;  // reference node of the box (kNoNode when not tracked)
};

inline float area(const float* b) {
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

inline bool inside(const float* c, const float* p) {
  return c[0] >= p[0] && c[1] >= p[1] && c[2] >= p[2] && c[3] <= p[3] && c[4] <= p[4] && c[5] <= p[5];
}

// Collapses a Node64 forest into Node128s (see device_layout.h).  With
// src64 (the reference node of each Node64 child slot) it records in dst128
// the Node128 slot that holds each reference node's box.
class Collapser {
 public:
  Collapser(const std::vector<uint32_t>& n64, std::vector<uint32_t>& out,
            const std::vector<uint32_t>* src64 = nullptr, std::vector<uint32_t>* dst128 = nullptr)
      : n64_(n64), out_(out), src64_(src64), dst128_(dst128) {}

  // Node128 index of the tree under Node64 `node`; `need` receives the
  // worst-case stack entries of a depth-first walk of that tree.
  uint32_t build(uint32_t node, uint32_t* need) {
    std::vector<Slot> slots;
    children(node, slots);
    while (slots.size() < kWideKids) {
      int pick = -1;
      float best = -1.0f;
      std::vector<Slot> kids;
      for (size_t i = 0; i < slots.size(); ++i) {
        if (slots[i].ref & kRefLeaf) continue;
        std::vector<Slot> k;
        children(slots[i].ref, k);
        if (slots.size() - 1 + k.size() > kWideKids) continue;
        bool nested = true;
        for (const Slot& c : k) nested = nested && inside(c.box, slots[i].box);
        const float a = area(slots[i].box);
        if (!nested || !(a > best)) continue;
        best = a;
        pick = (int)i;
        kids = k;
      }
      if (pick < 0) break;
      slots.erase(slots.begin() + pick);
      slots.insert(slots.begin() + pick, kids.begin(), kids.end());
    }
    const uint32_t idx = (uint32_t)(out_.size() / kNode128Words);
    out_.resize(out_.size() + kNode128Words, 0u);
    uint32_t deepest = 0;
    for (size_t k = 0; k < kWideKids; ++k) {
      uint32_t ref = kEmptyRef;
      float box[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      if (k < slots.size()) {
        ref = slots[k].ref;
        std::memcpy(box, slots[k].box, 24);
        if (!(ref & kRefLeaf)) {
          uint32_t sub = 0;
          ref = build(ref, &sub);
          deepest = std::max(deepest, sub);
        }
      }
      uint32_t* w = &out_[(size_t)idx * kNode128Words];
      for (int a = 0; a < 6; ++a) w[kWideKids * a + k] = fbits(box[a]);
      w[6 * kWideKids + k] = ref;
      if (dst128_ && k < slots.size() && slots[k].src != kNoNode)
        (*dst128_)[slots[k].src] = idx * kWideKids + (uint32_t)k;
    }
    // Expanding this node pushes at most (slots - 1) entries.
    *need = (slots.empty() ? 0u : (uint32_t)slots.size() - 1u) + deepest;
    return idx;
  }

 private:
  void children(uint32_t node, std::vector<Slot>& out) const {
    const uint32_t* w = &n64_[(size_t)node * kNode64Words];
    for (int c = 0; c < 2; ++c) {
      const uint32_t ref = w[12 + c];
      if (ref == kEmptyRef) continue;
      Slot s;
      s.ref = ref;
      std::memcpy(s.box, w + 6 * c, 24);
      s.src = src64_ ? (*src64_)[2 * (size_t)node + c] : kNoNode;
      out.push_back(s);
    }
  }

  const std::vector<uint32_t>& n64_;
  std::vector<uint32_t>& out_;
  const std::vector<uint32_t>* src64_;
  std::vector<uint32_t>* dst128_;
};

// Worst-case stack entries of a depth-first walk of a Node64 tree: one push
// per expansion with two passing children.
uint32_t stack_need2(const std::vector<uint32_t>& n64, uint32_t node) {
  const uint32_t* w = &n64[(size_t)node * kNode64Words];
  uint32_t deepest = 0, kids = 0;
  for (int c = 0; c < 2; ++c) {
    const uint32_t ref = w[12 + c];
    if (ref == kEmptyRef) continue;
    ++kids;
    if (!(ref & kRefLeaf)) deepest = std::max(deepest, stack_need2(n64, ref));
  }
  return (kids > 0 ? kids - 1 : 0) + deepest;
}

}  // namespace

std::string build_tlas_layout(const std::vector<uint32_t>& tlas_ref_nodes, uint32_t inst_count, TlasLayout& out) {
  out = TlasLayout();
  if (inst_count == 0) return {};
  std::vector<uint32_t> wrapper;
  std::string err = convert_forest(tlas_ref_nodes, {0u}, out.nodes, wrapper);
  if (!err.empty()) return err;
  out.root = wrapper[0];
  out.node_count = (uint32_t)(out.nodes.size() / kNode64Words);
  Collapser ct(out.nodes, out.nodes4);
  out.root4 = ct.build(out.root, &out.need4);
  out.need2 = stack_need2(out.nodes, out.root);
  return {};
}

void set_tlas_layout(TlasLayout&& t, DeviceLayout& out) {
  out.tlas_nodes = std::move(t.nodes);
  out.tlas4_nodes = std::move(t.nodes4);
  out.tlas_root = t.root;
  out.tlas4_root = t.root4;
  out.tlas_node_count = t.node_count;
  out.stack4_need = t.need4 + out.blas_need4;
  out.stack2_need = t.need2 + out.blas_need2;
}

std::string build_device_layout(const Pack& p, DeviceLayout& out) {
  out = DeviceLayout();
  // BLAS forest: one wrapper per distinct root.
  std::vector<uint32_t> roots;
  std::unordered_map<uint32_t, uint32_t> root_slot;
  for (const Binding& b : p.bindings) {
    root_slot[b.blas_root] = (uint32_t)roots.size();
    roots.push_back(b.blas_root);
  }
  std::vector<uint32_t> blas_wrapper, src64;
  std::string err = convert_forest(p.bvh_nodes, roots, out.blas_nodes, blas_wrapper, &src64);
  if (!err.empty()) return err;
  out.blas_node_count = (uint32_t)(out.blas_nodes.size() / kNode64Words);
  // Refit maps: parents and the Node64 slot of every reference node.
  const uint32_t n_ref = (uint32_t)(p.bvh_nodes.size() / 8);
  out.blas_parent.assign(n_ref, kNoNode);
  out.blas_dst64.assign(n_ref, kNoNode);
  out.blas_dst128.assign(n_ref, kNoNode);
  for (uint32_t i = 0; i < n_ref; ++i) {
    const uint32_t* q = &p.bvh_nodes[8 * (size_t)i];
    if ((q[7] >> 16) == 0xFFFFu) continue;
    out.blas_parent[i + 1] = i;
    out.blas_parent[i + q[6]] = i;
  }
  for (size_t slot = 0; slot < src64.size(); ++slot)
    if (src64[slot] != kNoNode) out.blas_dst64[src64[slot]] = (uint32_t)slot;

  const uint32_t n_inst = (uint32_t)p.tlas_blas_roots.size();
  out.inst_count = n_inst;
  out.instances.assign((size_t)n_inst * kInstWords, 0u);
  for (uint32_t s = 0; s < (uint32_t)p.tlas_instance_indices.size(); ++s) {
    const uint32_t inst = p.tlas_instance_indices[s];
    uint32_t* w = &out.instances[(size_t)s * kInstWords];
    std::memcpy(w, &p.tlas_w2l[16 * (size_t)inst], 64);
    std::memcpy(w + 16, &p.tlas_l2w[16 * (size_t)inst], 64);
    auto it = root_slot.find(p.tlas_blas_roots[inst]);
    if (it == root_slot.end()) return "instance references an unknown BLAS root";
    w[32] = blas_wrapper[it->second];
    w[33] = inst;
    w[35] = orientation_reversed(&p.tlas_l2w[16 * (size_t)inst]) ? 1u : 0u;
  }
  // Node128 trees: one per BLAS wrapper, one for the TLAS.
  std::vector<uint32_t> blas4_root(blas_wrapper.size(), 0);
  {
    Collapser cb(out.blas_nodes, out.blas4_nodes, &src64, &out.blas_dst128);
    for (size_t k = 0; k < blas_wrapper.size(); ++k) {
      uint32_t need = 0;
      blas4_root[k] = cb.build(blas_wrapper[k], &need);
      out.blas_need4 = std::max(out.blas_need4, need);
      out.blas_need2 = std::max(out.blas_need2, stack_need2(out.blas_nodes, blas_wrapper[k]));
    }
  }
  TlasLayout tl;
  err = build_tlas_layout(p.tlas_nodes, n_inst, tl);
  if (!err.empty()) return err;
  set_tlas_layout(std::move(tl), out);
  for (uint32_t s = 0; s < (uint32_t)p.tlas_instance_indices.size(); ++s) {
    const uint32_t inst = p.tlas_instance_indices[s];
    out.instances[(size_t)s * kInstWords + 34] = blas4_root[root_slot[p.tlas_blas_roots[inst]]];
  }
  const uint32_t n_tri = (uint32_t)(p.indices.size() / 4);
  out.tri_count = n_tri;
  out.tris.assign((size_t)n_tri * kTriWords, 0.f);
  for (uint32_t t = 0; t < n_tri; ++t) {
    float* w = &out.tris[(size_t)t * kTriWords];
    for (int v = 0; v < 3; ++v) {
      const uint32_t vi = p.indices[4 * (size_t)t + v];
      for (int k = 0; k < 3; ++k) w[3 * v + k] = p.positions[4 * (size_t)vi + k];
    }
  }
  return {};
}

}  // namespace vt
