// refit.cpp -- refits behind the C ABI: a caller's BVH on the device
// (refitBvhBounds), and a bound scene after one primitive's vertices move.
//
// The scene refit follows the reference's refit path for moved geometry
// (walkaround-hybrid HybridEnginePrimitiveUpdates.ts:1482-1560): the vertex
// slice is replaced, refitBvhBounds recomputes the primitive's BLAS boxes
// with the topology kept, and refitTlasTransforms with localBoundsChanged
// (shared-bvh scenePack.ts:1310-1420, refitTlasInstances tlas.ts:1141-1209)
// refits the TLAS over the primitive's instances from their new world boxes;
// accumulation restarts.  On the device the BLAS refit writes each node's box
// into the reference nodes and straight into the Node64 and Node128 slots
// that carry it, so the traversal trees are current when the launch ends;
// the (small) TLAS is refitted on the host and its trees rebuilt and
// uploaded.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "engine_internal.h"
#include "refit.h"

namespace {

vt_status rerr(vt_engine* e, vt_status code, const std::string& msg) { return engine_error(e, code, msg); }

#define RF_HIP(eng, expr)                                                                  \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) return rerr((eng), VT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

const char* refit_error_text(uint32_t kind) {
  switch (kind) {
    case vtf::kRefitCorruptInterior: return "interior node is corrupt";
    case vtf::kRefitCorruptLeaf: return "leaf triangle range is corrupt";
    case vtf::kRefitBadVertex: return "leaf references a vertex outside the positions";
    case vtf::kRefitNonFinite: return "leaf references a non-finite vertex";
    case vtf::kRefitRepeated: return "node is reached from two parents";
    case vtf::kRefitUnreachable: return "node is unreachable from the root";
    default: return "rejected";
  }
}

}  // namespace

// Device copies the scene refit works on, made on the first refit of a scene
// and dropped with it: the reference BLAS nodes, their parent links and
// derived-tree slots, and arrival counters; plus scratch for refits of a
// caller's tree.
struct RefitState {
  DeviceBuffer nodes, parent, dst64, dst128, arrivals;
  bool ready = false;
  DeviceBuffer user_parent, user_arrivals, user_err;
  void release_scene() {
    for (DeviceBuffer* b : {&nodes, &parent, &dst64, &dst128, &arrivals}) b->release();
    ready = false;
  }
  void release() {
    release_scene();
    for (DeviceBuffer* b : {&user_parent, &user_arrivals, &user_err}) b->release();
  }
};

void engine_release_refit(vt_engine* e) {
  if (!e->refit) return;
  e->refit->release();
  delete e->refit;
  e->refit = nullptr;
}

void engine_refit_scene_changed(vt_engine* e) {
  if (e->refit) e->refit->release_scene();
}

static vt_status ensure_refit_state(vt_engine* e) {
  if (!e->refit) e->refit = new RefitState();
  return VT_OK;
}

static vt_status ensure_scene_refit(vt_engine* e) {
  if (vt_status s = ensure_refit_state(e)) return s;
  RefitState& R = *e->refit;
  if (R.ready) return VT_OK;
  const vt::Pack& P = *e->pack;
  const vt::DeviceLayout& L = *e->layout;
  const size_t n = P.bvh_nodes.size() / 8;
  vt_status s = VT_OK;
  s = s ? s : engine_upload(e, R.nodes, P.bvh_nodes.data(), P.bvh_nodes.size() * 4);
  s = s ? s : engine_upload(e, R.parent, L.blas_parent.data(), n * 4);
  s = s ? s : engine_upload(e, R.dst64, L.blas_dst64.data(), n * 4);
  s = s ? s : engine_upload(e, R.dst128, L.blas_dst128.data(), n * 4);
  s = s ? s : R.arrivals.ensure(e, std::max<size_t>(n * 4, 16));
  if (s) return s;
  RF_HIP(e, hipStreamSynchronize(e->stream));
  R.ready = true;
  return VT_OK;
}

extern "C" {

vt_status vt_refit_bvh_bounds_device(vt_engine* e, uint32_t* d_nodes, uint32_t node_count, const uint32_t* d_indices,
                                     uint32_t index_words, uint32_t index_stride, const float* d_positions,
                                     uint32_t position_floats, uint32_t position_stride) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    if (position_stride != 3 && position_stride != 4)
      return rerr(e, VT_ERR_RANGE, "refitBvhBounds: position stride must be 3 or 4.");
    if (index_stride != 3 && index_stride != 4)
      return rerr(e, VT_ERR_RANGE, "refitBvhBounds: index stride must be 3 or 4.");
    if (position_floats % position_stride || index_words % index_stride)
      return rerr(e, VT_ERR_BUILD, "refitBvhBounds: geometry buffer is not stride-aligned.");
    if (node_count == 0) return VT_OK;
    if (!d_nodes || (index_words && !d_indices) || (position_floats && !d_positions))
      return rerr(e, VT_ERR_INVALID_ARGUMENT, "refitBvhBounds: NULL buffer");
    RF_HIP(e, hipSetDevice(e->device));
    if (vt_status s = ensure_refit_state(e)) return s;
    RefitState& R = *e->refit;
    const size_t bytes = std::max<size_t>((size_t)node_count * 4, 16);
    vt_status s = R.user_parent.ensure(e, bytes);
    s = s ? s : R.user_arrivals.ensure(e, bytes);
    s = s ? s : R.user_err.ensure(e, 16);
    if (s) return s;
    uint32_t* parent = (uint32_t*)R.user_parent.ptr;
    uint32_t* arrivals = (uint32_t*)R.user_arrivals.ptr;
    auto* err = (unsigned long long*)R.user_err.ptr;
    // Preflight first; nothing of the caller's is written unless it passes.
    RF_HIP(e, vtf::launch_fill_u32(parent, node_count, vtf::kNone, e->stream));
    RF_HIP(e, vtf::launch_fill_u32((uint32_t*)err, 2, 0xFFFFFFFFu, e->stream));
    RF_HIP(e, vtf::launch_refit_link(d_nodes, node_count, d_indices, index_words / index_stride, index_stride,
                                     d_positions, position_floats / position_stride, position_stride, parent, err,
                                     e->stream));
    unsigned long long found = 0;
    RF_HIP(e, hipMemcpyAsync(&found, err, 8, hipMemcpyDeviceToHost, e->stream));
    RF_HIP(e, hipStreamSynchronize(e->stream));
    if (found != ~0ull)
      return rerr(e, VT_ERR_BUILD, std::string("refitBvhBounds: ") + refit_error_text((uint32_t)(found & 0xFF)) +
                                       " (node " + std::to_string(found >> 8) + ")");
    RF_HIP(e, vtf::launch_fill_u32(arrivals, node_count, 0u, e->stream));
    vtf::RefitArgs a{};
    a.nodes = d_nodes;
    a.first = 0;
    a.count = node_count;
    a.indices = d_indices;
    a.index_stride = index_stride;
    a.positions = d_positions;
    a.position_stride = position_stride;
    a.parent = parent;
    a.arrivals = arrivals;
    RF_HIP(e, vtf::launch_refit(a, e->stream));
    return VT_OK;
  });
}

vt_status vt_refit_primitive(vt_engine* e, uint32_t primitive, const float* positions, uint32_t position_floats,
                             const float* normals, uint32_t normal_floats) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    if (!e->has_scene || !e->pack || !e->layout) return rerr(e, VT_ERR_STATE, "no scene is set");
    if (!positions) return rerr(e, VT_ERR_INVALID_ARGUMENT, "positions is NULL");
    vt::Pack& P = *e->pack;
    size_t bi = P.bindings.size();
    for (size_t k = 0; k < P.bindings.size(); ++k)
      if (P.bindings[k].primitive == primitive) bi = k;
    if (bi == P.bindings.size())
      return rerr(e, VT_ERR_RANGE, "primitive " + std::to_string(primitive) + " has no packed geometry");
    const vt::Binding& b = P.bindings[bi];
    if ((uint64_t)position_floats != 3ull * b.vertex_count)
      return rerr(e, VT_ERR_RANGE, "expected " + std::to_string(3ull * b.vertex_count) +
                                       " position floats (the primitive's vertex count is kept)");
    if (normals && normal_floats != position_floats)
      return rerr(e, VT_ERR_RANGE, "normals must have as many floats as positions");
    for (uint32_t i = 0; i < position_floats; ++i)
      if (!std::isfinite(positions[i]) || (normals && !std::isfinite(normals[i])))
        return rerr(e, VT_ERR_RANGE, "vertex " + std::to_string(i / 3) + " is non-finite");
    // Mesh-area emitters are built from the emissive triangles' world
    // vertices; moving them needs the full scene path.
    for (uint32_t t = b.tri_start; t < b.tri_start + b.tri_count; ++t) {
      const uint32_t mid = P.tri_material[t];
      if (mid >= e->materials.size()) continue;
      const vt_material& m = e->materials[mid];
      for (int k = 0; k < 3; ++k)
        if ((float)((double)m.emissive[k] * (double)m.emissive_intensity) > 0.0f)
          return rerr(e, VT_ERR_STATE, "primitive " + std::to_string(primitive) +
                                           " is emissive: its mesh-area emitters move with it; set the scene again");
    }
    // ---- host side: everything that can fail happens before any write ----
    std::vector<float> pos4((size_t)b.vertex_count * 4), nrm4;
    for (uint32_t v = 0; v < b.vertex_count; ++v) {
      for (int k = 0; k < 3; ++k) pos4[4 * (size_t)v + k] = positions[3 * (size_t)v + k];
      pos4[4 * (size_t)v + 3] = P.positions[4 * ((size_t)b.vertex_start + v) + 3];
    }
    if (normals) {
      nrm4.resize(pos4.size());
      for (uint32_t v = 0; v < b.vertex_count; ++v) {
        for (int k = 0; k < 3; ++k) nrm4[4 * (size_t)v + k] = normals[3 * (size_t)v + k];
        nrm4[4 * (size_t)v + 3] = P.normals[4 * ((size_t)b.vertex_start + v) + 3];
      }
    }
    double lmn[3], lmx[3];
    vt::local_aabb(pos4.data(), b.vertex_count, lmn, lmx);
    // Every instance's world box (the others' from their bindings' boxes,
    // which hold the binary32 values the binary64 boxes were taken over).
    const uint32_t n_inst = (uint32_t)P.instance_primitive.size();
    std::vector<const vt::Binding*> of_prim;
    for (const vt::Binding& x : P.bindings) {
      if (x.primitive >= of_prim.size()) of_prim.resize(x.primitive + 1, nullptr);
      of_prim[x.primitive] = &x;
    }
    std::vector<double> amn(3 * (size_t)n_inst), amx(3 * (size_t)n_inst);
    std::vector<uint32_t> changed;
    double wmn[3] = {INFINITY, INFINITY, INFINITY}, wmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n_inst; ++i) {
      const uint32_t pi = P.instance_primitive[i];
      double bmn[3], bmx[3];
      if (pi == primitive) {
        std::memcpy(bmn, lmn, sizeof bmn);
        std::memcpy(bmx, lmx, sizeof bmx);
        changed.push_back(i);
      } else {
        const vt::Binding* x = of_prim[pi];
        for (int k = 0; k < 3; ++k) bmn[k] = x->local_min[k], bmx[k] = x->local_max[k];
      }
      vt::world_aabb(&P.tlas_l2w[16 * (size_t)i], bmn, bmx, &amn[3 * (size_t)i], &amx[3 * (size_t)i]);
      for (int k = 0; k < 3; ++k) {
        wmn[k] = std::min(wmn[k], amn[3 * (size_t)i + k]);
        wmx[k] = std::max(wmx[k], amx[3 * (size_t)i + k]);
      }
    }
    std::vector<uint32_t> tlas = P.tlas_nodes;
    vt::BuildError be = vt::refit_tlas_instances(tlas, P.tlas_instance_indices, n_inst, amn.data(), amx.data(), changed);
    if (be) return rerr(e, be.code, be.message);
    vt::TlasLayout tl;
    std::string le = vt::build_tlas_layout(tlas, n_inst, tl);
    if (!le.empty()) return rerr(e, VT_ERR_RANGE, le);
    // The BLAS's node range: its root up to the next BLAS's root.
    const uint32_t first = b.blas_root;
    uint32_t end = (uint32_t)(P.bvh_nodes.size() / 8);
    for (const vt::Binding& x : P.bindings)
      if (x.blas_root > first && x.blas_root < end) end = x.blas_root;
    std::vector<uint32_t> refitted((size_t)(end - first) * 8);
    RF_HIP(e, hipSetDevice(e->device));
    // Frames queued on the stream finish with the old geometry first.
    RF_HIP(e, hipStreamSynchronize(e->stream));
    if (vt_status s = ensure_scene_refit(e)) return s;
    SceneBuffers& S = e->scene;
    if (tl.nodes4.size() * 4 > S.tlas4.bytes || tl.nodes.size() * 4 > S.tlas.bytes)
      return rerr(e, VT_ERR_BUILD, "refitted TLAS does not fit its buffers");
    // ---- device side ----
    RefitState& R = *e->refit;
    char* dpos = (char*)S.positions.ptr + 16 * (size_t)b.vertex_start;
    RF_HIP(e, hipMemcpyAsync(dpos, pos4.data(), pos4.size() * 4, hipMemcpyHostToDevice, e->stream));
    if (normals) {
      char* dn = (char*)S.normals.ptr + 16 * (size_t)b.vertex_start;
      RF_HIP(e, hipMemcpyAsync(dn, nrm4.data(), nrm4.size() * 4, hipMemcpyHostToDevice, e->stream));
    }
    RF_HIP(e, vtf::launch_refit_tris((float*)S.tris.ptr, (const uint32_t*)S.indices.ptr, (const float*)S.positions.ptr,
                                     b.tri_start, b.tri_count, e->stream));
    RF_HIP(e, vtf::launch_fill_u32((uint32_t*)R.arrivals.ptr + first, end - first, 0u, e->stream));
    vtf::RefitArgs a{};
    a.nodes = (uint32_t*)R.nodes.ptr;
    a.first = first;
    a.count = end - first;
    a.indices = (const uint32_t*)S.indices.ptr;
    a.index_stride = 4;
    a.positions = (const float*)S.positions.ptr;
    a.position_stride = 4;
    a.parent = (const uint32_t*)R.parent.ptr;
    a.arrivals = (uint32_t*)R.arrivals.ptr;
    a.dst64 = (const uint32_t*)R.dst64.ptr;
    a.node64 = (uint32_t*)S.blas.ptr;
    const bool wide = !e->layout->blas4_nodes.empty();
    a.dst128 = wide ? (const uint32_t*)R.dst128.ptr : nullptr;
    a.node128 = wide ? (uint32_t*)S.blas4.ptr : nullptr;
    RF_HIP(e, vtf::launch_refit(a, e->stream));
    if (!tl.nodes.empty())
      RF_HIP(e, hipMemcpyAsync(S.tlas.ptr, tl.nodes.data(), tl.nodes.size() * 4, hipMemcpyHostToDevice, e->stream));
    if (!tl.nodes4.empty())
      RF_HIP(e, hipMemcpyAsync(S.tlas4.ptr, tl.nodes4.data(), tl.nodes4.size() * 4, hipMemcpyHostToDevice,
                               e->stream));
    RF_HIP(e, hipMemcpyAsync(refitted.data(), (const uint32_t*)R.nodes.ptr + 8 * (size_t)first,
                             refitted.size() * 4, hipMemcpyDeviceToHost, e->stream));
    RF_HIP(e, hipStreamSynchronize(e->stream));
    // ---- host copies follow (nothing below fails) ----
    std::memcpy(&P.positions[4 * (size_t)b.vertex_start], pos4.data(), pos4.size() * 4);
    if (normals) std::memcpy(&P.normals[4 * (size_t)b.vertex_start], nrm4.data(), nrm4.size() * 4);
    std::memcpy(&P.bvh_nodes[8 * (size_t)first], refitted.data(), refitted.size() * 4);
    vt::Binding& bw = P.bindings[bi];
    for (int k = 0; k < 3; ++k) bw.local_min[k] = (float)lmn[k], bw.local_max[k] = (float)lmx[k];
    P.tlas_nodes = std::move(tlas);
    for (int k = 0; k < 3; ++k) P.world_min[k] = wmn[k], P.world_max[k] = wmx[k];
    vt::set_tlas_layout(std::move(tl), *e->layout);
    engine_render_rebound(e);
    engine_release_cwbvh(e);
    ++e->scene_generation;
    return vt_reset(e);
  });
}

}  // extern "C"
