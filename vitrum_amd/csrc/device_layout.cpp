// device_layout.cpp -- converts a packed scene into the Node64 HBM layout.
#include "device_layout.h"

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
// leaf_offset_limit bounds the offsets a leaf ref can carry.
std::string convert_forest(const std::vector<uint32_t>& ref_nodes, const std::vector<uint32_t>& roots,
                           std::vector<uint32_t>& out, std::vector<uint32_t>& wrapper_of_root) {
  const uint32_t count = (uint32_t)(ref_nodes.size() / 8);
  std::vector<uint32_t> map(count, ~0u);
  uint32_t next = 0;
  for (uint32_t i = 0; i < count; ++i)
    if ((ref_nodes[8 * (size_t)i + 7] >> 16) != 0xFFFFu) map[i] = next++;
  const uint32_t interior = next;
  out.assign((size_t)(interior + roots.size()) * kNode64Words, 0u);
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
  }
  wrapper_of_root.resize(roots.size());
  for (size_t k = 0; k < roots.size(); ++k) {
    const uint32_t idx = interior + (uint32_t)k;
    uint32_t* w = &out[(size_t)idx * kNode64Words];
    put_box(w, roots[k]);
    put_empty_box(w + 6);
    w[12] = child_ref(roots[k]);
    w[13] = kEmptyRef;
    wrapper_of_root[k] = idx;
  }
  return err;
}

}  // namespace

std::string build_device_layout(const Pack& p, DeviceLayout& out) {
  out = DeviceLayout();
  // BLAS forest: one wrapper per distinct root.
  std::vector<uint32_t> roots;
  std::unordered_map<uint32_t, uint32_t> root_slot;
  for (const Binding& b : p.bindings) {
    root_slot[b.blas_root] = (uint32_t)roots.size();
    roots.push_back(b.blas_root);
  }
  std::vector<uint32_t> blas_wrapper;
  std::string err = convert_forest(p.bvh_nodes, roots, out.blas_nodes, blas_wrapper);
  if (!err.empty()) return err;
  out.blas_node_count = (uint32_t)(out.blas_nodes.size() / kNode64Words);

  const uint32_t n_inst = (uint32_t)p.tlas_blas_roots.size();
  if (n_inst > 0) {
    std::vector<uint32_t> tlas_wrapper;
    err = convert_forest(p.tlas_nodes, {0u}, out.tlas_nodes, tlas_wrapper);
    if (!err.empty()) return err;
    out.tlas_root = tlas_wrapper[0];
    out.tlas_node_count = (uint32_t)(out.tlas_nodes.size() / kNode64Words);
  }
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
