// gpubuild.cpp -- the device BLAS builder behind the C ABI.
//
// Drives the level kernels of gpubuild.hip: the tree grows one depth per
// round, every node of a depth in one launch.  A node's result depends only
// on the multiset of its triangles and their order inside its range, and the
// range a node owns is the one the host builder's depth-first loop gives it
// (partitions are stable and touch only the node's own range), so the nodes,
// the decisions and the final permutation equal the host builder's; the host
// then lays the nodes out in the same preorder (left child next, right child
// at a relative offset; buildArrayBvh.ts:342-780).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "engine_internal.h"
#include "gpubuild.h"

// Device buffers of the builder, kept by the engine between builds (they
// only grow) and freed with it.
struct GpuBuildScratch {
  DeviceBuffer pos, idx, bmin, bmax, cen, perm, tmp, bin_of;
  DeviceBuffer lbegin, lend, lwant, nbox, cbox, decision, nleft, pnodes;
  DeviceBuffer wslot, wbox, wcen, wbins, wfirst, wcount, chunks, cleft, lnodes;
  void release() {
    for (DeviceBuffer* b : {&pos, &idx, &bmin, &bmax, &cen, &perm, &tmp, &bin_of, &lbegin, &lend, &lwant, &nbox,
                            &cbox, &decision, &nleft, &pnodes, &wslot, &wbox, &wcen, &wbins, &wfirst, &wcount,
                            &chunks, &cleft, &lnodes})
      b->release();
  }
};

void engine_release_gpubuild(vt_engine* e) {
  if (!e->gpubuild) return;
  e->gpubuild->release();
  delete e->gpubuild;
  e->gpubuild = nullptr;
}

namespace {

#define GB_HIP(eng, expr)                                                                  \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) return vt::BuildError{VT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
  } while (0)

struct GNode {
  uint32_t begin, end, depth;
  uint32_t left = ~0u, right = ~0u, axis = 0;
  float box[6];
};

// A node with more triangles than kWideMin is reduced, binned and
// partitioned by several workgroups, kChunk triangles each; smaller nodes
// take one workgroup.
constexpr uint32_t kWideMin = 4096;
constexpr uint32_t kChunk = 1024;
// Nodes of at most kSmallMax triangles share a workgroup, 64 threads each,
// when four nodes' bins fit the LDS budget.
constexpr uint32_t kSmallMax = 256;
constexpr size_t kSmallLdsBudget = 64 * 1024;

// The level loop over a forest: triangles [roots[r].first, roots[r].second)
// of the (already checked) inputs form tree r, whose root is nodes[r].  On
// return `nodes` holds every node and `perm` the final permutation (global
// triangle ids).
vt::BuildError build_core(vt_engine* eng, const float* positions, uint32_t position_floats, uint32_t position_stride,
                          const uint32_t* indices, uint32_t index_words, uint32_t index_stride,
                          const std::vector<std::pair<uint32_t, uint32_t>>& roots, uint32_t max_leaf,
                          uint32_t num_bins, uint32_t max_depth, std::vector<GNode>& nodes,
                          std::vector<uint32_t>& perm) {
    const uint32_t n_tri = index_words / index_stride;
    hipStream_t s = eng->stream;
    GB_HIP(eng, hipSetDevice(eng->device));
    if (!eng->gpubuild) eng->gpubuild = new GpuBuildScratch();
    GpuBuildScratch& b = *eng->gpubuild;
    const size_t n = n_tri;
    // A level holds at most n_tri nodes: every node owns at least one
    // triangle; wide nodes own more than kWideMin each.
    const size_t max_wide = n / kWideMin + 1, max_chunks = n / kChunk + max_wide;
    struct Sz {
      DeviceBuffer* buf;
      size_t bytes;
    };
    for (const Sz& z : {Sz{&b.pos, (size_t)position_floats * 4}, Sz{&b.idx, (size_t)index_words * 4},
                        Sz{&b.bmin, 12 * n}, Sz{&b.bmax, 12 * n}, Sz{&b.cen, 24 * n}, Sz{&b.perm, 4 * n},
                        Sz{&b.tmp, 4 * n}, Sz{&b.bin_of, 3 * n}, Sz{&b.lbegin, 4 * n}, Sz{&b.lend, 4 * n},
                        Sz{&b.lwant, n}, Sz{&b.nbox, 24 * n}, Sz{&b.cbox, 48 * n}, Sz{&b.decision, 12 * n},
                        Sz{&b.nleft, 4 * n}, Sz{&b.pnodes, 4 * n}, Sz{&b.wslot, 4 * n},
                        Sz{&b.wbox, 24 * max_wide}, Sz{&b.wcen, 48 * max_wide},
                        Sz{&b.wbins, (size_t)84 * num_bins * max_wide}, Sz{&b.wfirst, 4 * max_wide},
                        Sz{&b.wcount, 4 * max_wide}, Sz{&b.chunks, 16 * max_chunks},
                        Sz{&b.cleft, 4 * max_chunks}, Sz{&b.lnodes, 4 * n}})
      if (vt_status st = z.buf->ensure(eng, std::max<size_t>(z.bytes, 16)))
        return vt::BuildError{st, "device BLAS build: hipMalloc(" + std::to_string(z.bytes) + ") failed"};
    GB_HIP(eng, hipMemcpyAsync(b.pos.ptr, positions, (size_t)position_floats * 4, hipMemcpyHostToDevice, s));
    GB_HIP(eng, hipMemcpyAsync(b.idx.ptr, indices, (size_t)index_words * 4, hipMemcpyHostToDevice, s));
    perm.resize(n_tri);
    for (uint32_t t = 0; t < n_tri; ++t) perm[t] = t;
    GB_HIP(eng, hipMemcpyAsync(b.perm.ptr, perm.data(), 4 * n, hipMemcpyHostToDevice, s));
    float* d_bmin = (float*)b.bmin.ptr;
    float* d_bmax = (float*)b.bmax.ptr;
    double* d_cen = (double*)b.cen.ptr;
    uint32_t* d_perm = (uint32_t*)b.perm.ptr;
    GB_HIP(eng, vtg::launch_tri_records((const float*)b.pos.ptr, position_stride, (const uint32_t*)b.idx.ptr,
                                        index_stride, n_tri, d_bmin, d_bmax, d_cen, s));

    vtg::LevelArgs la{};
    la.begin = (const uint32_t*)b.lbegin.ptr;
    la.end = (const uint32_t*)b.lend.ptr;
    la.split_wanted = (const uint8_t*)b.lwant.ptr;
    la.perm = d_perm;
    la.bmin = d_bmin;
    la.bmax = d_bmax;
    la.cen = d_cen;
    la.bin_of = (uint8_t*)b.bin_of.ptr;
    la.node_box = (float*)b.nbox.ptr;
    la.cen_box = (double*)b.cbox.ptr;
    la.decision = (uint32_t*)b.decision.ptr;
    la.n_tri = n_tri;
    la.num_bins = num_bins;
    vtg::PartitionArgs pa{};
    pa.nodes = (const uint32_t*)b.pnodes.ptr;
    pa.begin = la.begin;
    pa.end = la.end;
    pa.decision = la.decision;
    pa.bin_of = la.bin_of;
    pa.perm = d_perm;
    pa.tmp = (uint32_t*)b.tmp.ptr;
    pa.n_left = (uint32_t*)b.nleft.ptr;
    pa.n_tri = n_tri;
    la.wide_slot = (const uint32_t*)b.wslot.ptr;
    const bool small_ok = vtg::level_lds_bytes(num_bins, 64) <= kSmallLdsBudget;
    std::vector<uint32_t> hlist;
    la.w_box_key = (const uint32_t*)b.wbox.ptr;
    la.w_cen_key = (const unsigned long long*)b.wcen.ptr;
    la.w_bins = (const uint32_t*)b.wbins.ptr;
    vtg::WideArgs wa{};
    uint32_t* d_chunks = (uint32_t*)b.chunks.ptr;  // begin, end, slot, node: max_chunks each
    wa.c_begin = d_chunks;
    wa.c_end = d_chunks + max_chunks;
    wa.c_slot = d_chunks + 2 * max_chunks;
    wa.c_node = d_chunks + 3 * max_chunks;
    wa.w_first = (const uint32_t*)b.wfirst.ptr;
    wa.w_count = (const uint32_t*)b.wcount.ptr;
    wa.level_begin = la.begin;
    wa.decision = la.decision;
    wa.perm = d_perm;
    wa.tmp = pa.tmp;
    wa.bmin = d_bmin;
    wa.bmax = d_bmax;
    wa.cen = d_cen;
    wa.bin_of = la.bin_of;
    wa.w_box_key = (uint32_t*)b.wbox.ptr;
    wa.w_cen_key = (unsigned long long*)b.wcen.ptr;
    wa.w_bins = (uint32_t*)b.wbins.ptr;
    wa.chunk_left = (uint32_t*)b.cleft.ptr;
    wa.n_left = pa.n_left;
    wa.n_tri = n_tri;
    wa.num_bins = num_bins;
    std::vector<uint32_t> hslot, hchunk, hwfirst, hwcount;

    nodes.clear();
    nodes.reserve(2 * n);
    std::vector<uint32_t> level, next, hb, he, hdec, hnl, psel;
    for (const auto& r : roots) {
      level.push_back((uint32_t)nodes.size());
      nodes.push_back({r.first, r.second, 0});
    }
    std::vector<uint8_t> hw;
    std::vector<float> hbox;
    std::vector<double> hcbox, hcen;  // hcen: centroids, fetched on the first median fallback
    while (!level.empty()) {
      const uint32_t L = (uint32_t)level.size();
      hb.resize(L), he.resize(L), hw.resize(L);
      bool any_split = false;
      for (uint32_t i = 0; i < L; ++i) {
        const GNode& g = nodes[level[i]];
        const uint32_t cnt = g.end - g.begin;
        hb[i] = g.begin, he[i] = g.end;
        hw[i] = cnt > max_leaf ? 1 : 0;
        any_split |= hw[i] != 0;
        if (cnt > max_leaf && g.depth >= max_depth) {
          // Every node of a level has the same depth, so the first such node
          // in level order is the first the depth-first builder meets.
          std::string extra = cnt > 0xFFFFu ? " and the packed 16-bit limit (0xFFFF = 65535)" : "";
          return vt::BuildError{VT_ERR_BUILD,
                              "maxDepth=" + std::to_string(max_depth) + " prevents enforcing maxLeafTriangles=" +
                                  std::to_string(max_leaf) + ": depth " + std::to_string(g.depth) +
                                  " still contains " + std::to_string(cnt) + " triangles" + extra + "."};
        }
      }
      // wide nodes and their chunks
      hslot.assign(L, ~0u);
      hwfirst.clear(), hwcount.clear();
      std::vector<uint32_t> cb, ce, cs, cn;
      for (uint32_t i = 0; i < L; ++i) {
        if (!hw[i] || he[i] - hb[i] <= kWideMin) continue;
        const uint32_t w = (uint32_t)hwfirst.size();
        hslot[i] = w;
        hwfirst.push_back((uint32_t)cb.size());
        for (uint32_t x = hb[i]; x < he[i]; x += kChunk) {
          cb.push_back(x);
          ce.push_back(std::min(he[i], x + kChunk));
          cs.push_back(w);
          cn.push_back(i);
        }
        hwcount.push_back((uint32_t)cb.size() - hwfirst.back());
      }
      const uint32_t W = (uint32_t)hwfirst.size(), NC = (uint32_t)cb.size();
      GB_HIP(eng, hipMemcpyAsync(b.lbegin.ptr, hb.data(), 4 * (size_t)L, hipMemcpyHostToDevice, s));
      GB_HIP(eng, hipMemcpyAsync(b.lend.ptr, he.data(), 4 * (size_t)L, hipMemcpyHostToDevice, s));
      GB_HIP(eng, hipMemcpyAsync(b.lwant.ptr, hw.data(), (size_t)L, hipMemcpyHostToDevice, s));
      GB_HIP(eng, hipMemcpyAsync(b.wslot.ptr, hslot.data(), 4 * (size_t)L, hipMemcpyHostToDevice, s));
      if (W) {
        hchunk.resize(4 * max_chunks);
        std::copy(cb.begin(), cb.end(), hchunk.begin());
        std::copy(ce.begin(), ce.end(), hchunk.begin() + max_chunks);
        std::copy(cs.begin(), cs.end(), hchunk.begin() + 2 * max_chunks);
        std::copy(cn.begin(), cn.end(), hchunk.begin() + 3 * max_chunks);
        GB_HIP(eng, hipMemcpyAsync(d_chunks, hchunk.data(), 16 * max_chunks, hipMemcpyHostToDevice, s));
        GB_HIP(eng, hipMemcpyAsync(b.wfirst.ptr, hwfirst.data(), 4 * (size_t)W, hipMemcpyHostToDevice, s));
        GB_HIP(eng, hipMemcpyAsync(b.wcount.ptr, hwcount.data(), 4 * (size_t)W, hipMemcpyHostToDevice, s));
        GB_HIP(eng, vtg::launch_wide_reduce(wa, W, NC, s));
      }
      // one workgroup per large or wide node, four small nodes per workgroup
      hlist.clear();
      for (uint32_t i = 0; i < L; ++i)
        if (!small_ok || hslot[i] != ~0u || he[i] - hb[i] > kSmallMax) hlist.push_back(i);
      const uint32_t n_big = (uint32_t)hlist.size();
      for (uint32_t i = 0; i < L; ++i)
        if (small_ok && hslot[i] == ~0u && he[i] - hb[i] <= kSmallMax) hlist.push_back(i);
      GB_HIP(eng, hipMemcpyAsync(b.lnodes.ptr, hlist.data(), 4 * (size_t)L, hipMemcpyHostToDevice, s));
      vtg::LevelArgs lb = la;
      lb.node_list = (const uint32_t*)b.lnodes.ptr;
      GB_HIP(eng, vtg::launch_level(lb, n_big, 256, s));
      lb.node_list += n_big;
      GB_HIP(eng, vtg::launch_level(lb, L - n_big, 64, s));
      hbox.resize(6 * (size_t)L);
      hdec.resize(3 * (size_t)L);
      hcbox.resize(6 * (size_t)L);
      GB_HIP(eng, hipMemcpyAsync(hbox.data(), b.nbox.ptr, 24 * (size_t)L, hipMemcpyDeviceToHost, s));
      if (any_split)
        GB_HIP(eng, hipMemcpyAsync(hdec.data(), b.decision.ptr, 12 * (size_t)L, hipMemcpyDeviceToHost, s));
      bool have_cbox = false;  // centroid boxes: fetched for a median fallback only
      GB_HIP(eng, hipStreamSynchronize(s));
      psel.clear();
      for (uint32_t i = 0; i < L; ++i) {
        std::memcpy(nodes[level[i]].box, &hbox[6 * (size_t)i], 24);
        if (hw[i] && hdec[3 * (size_t)i + 2] && hslot[i] == ~0u) psel.push_back(i);
      }
      hnl.assign(L, 0u);
      if (!psel.empty() || W) {
        GB_HIP(eng, hipMemcpyAsync(b.pnodes.ptr, psel.data(), 4 * psel.size(), hipMemcpyHostToDevice, s));
        GB_HIP(eng, vtg::launch_partition(pa, (uint32_t)psel.size(), s));
        GB_HIP(eng, vtg::launch_wide_partition(wa, NC, s));
        GB_HIP(eng, hipMemcpyAsync(hnl.data(), b.nleft.ptr, 4 * (size_t)L, hipMemcpyDeviceToHost, s));
        GB_HIP(eng, hipStreamSynchronize(s));
      }
      next.clear();
      for (uint32_t i = 0; i < L; ++i) {
        if (!hw[i]) continue;
        const uint32_t id = level[i];
        const uint32_t nb = nodes[id].begin, ne = nodes[id].end, cnt = ne - nb;
        uint32_t axis = hdec[3 * (size_t)i], n_left = hnl[i];
        const bool partitioned = hdec[3 * (size_t)i + 2] && n_left != 0 && n_left != cnt;
        if (!partitioned) {
          // Median fallback (buildArrayBvh.ts:677-694): widest centroid axis,
          // sort by centroid then triangle id, split at n/2.
          if (!have_cbox) {
            GB_HIP(eng, hipMemcpyAsync(hcbox.data(), b.cbox.ptr, 48 * (size_t)L, hipMemcpyDeviceToHost, s));
            GB_HIP(eng, hipStreamSynchronize(s));
            have_cbox = true;
          }
          const double* cb = &hcbox[6 * (size_t)i];
          axis = 0;
          double widest = cb[3] - cb[0];
          for (uint32_t k = 1; k < 3; ++k)
            if (cb[3 + k] - cb[k] > widest) axis = k, widest = cb[3 + k] - cb[k];
          if (hcen.empty()) {
            hcen.resize(3 * n);
            GB_HIP(eng, hipMemcpyAsync(hcen.data(), d_cen, 24 * n, hipMemcpyDeviceToHost, s));
          }
          std::vector<uint32_t> seg(cnt);
          GB_HIP(eng, hipMemcpyAsync(seg.data(), d_perm + nb, 4 * (size_t)cnt, hipMemcpyDeviceToHost, s));
          GB_HIP(eng, hipStreamSynchronize(s));
          const double* cax = hcen.data() + (size_t)axis * n;
          std::sort(seg.begin(), seg.end(),
                    [&](uint32_t x, uint32_t y) { return cax[x] != cax[y] ? cax[x] < cax[y] : x < y; });
          GB_HIP(eng, hipMemcpyAsync(d_perm + nb, seg.data(), 4 * (size_t)cnt, hipMemcpyHostToDevice, s));
          GB_HIP(eng, hipStreamSynchronize(s));
          n_left = cnt / 2;
        }
        const uint32_t d = nodes[id].depth + 1;
        const uint32_t lc = (uint32_t)nodes.size();
        nodes.push_back({nb, nb + n_left, d});
        nodes.push_back({nb + n_left, ne, d});
        nodes[id].left = lc;
        nodes[id].right = lc + 1;
        nodes[id].axis = axis;
        next.push_back(lc);
        next.push_back(lc + 1);
      }
      level.swap(next);
    }
    GB_HIP(eng, hipMemcpyAsync(perm.data(), d_perm, 4 * n, hipMemcpyDeviceToHost, s));
    GB_HIP(eng, hipStreamSynchronize(s));

    return {};
}

// Tree `root` of a forest built by build_core as the host builder's output
// for its mesh: preorder nodes (left child next, right child at a relative
// offset), leaf offsets and permutation relative to the mesh's first
// triangle tri_base, reordered indices and materials from the mesh's own
// arrays.
vt::BuildError emit_tree(const std::vector<GNode>& nodes, uint32_t root, const std::vector<uint32_t>& perm,
                         uint32_t tri_base, uint32_t n_tri, const uint32_t* indices, uint32_t index_stride,
                         const uint32_t* tri_material_ids, vt::BlasOut& out) {
    std::vector<uint32_t> words;
    words.reserve(16 * (size_t)n_tri);
    std::vector<std::pair<uint32_t, uint32_t>> st{{root, ~0u}};  // (node, preorder parent whose right child it is)
    uint32_t slot = 0;
    while (!st.empty()) {
      auto [id, parent] = st.back();
      st.pop_back();
      const uint32_t at = slot++;
      words.resize(words.size() + 8, 0u);
      if (parent != ~0u) words[8 * (size_t)parent + 6] = at - parent;
      const GNode& g = nodes[id];
      uint32_t* w = &words[8 * (size_t)at];
      std::memcpy(w, g.box, 24);
      if (g.left == ~0u) {
        w[6] = g.begin - tri_base;
        w[7] = vt::kLeafFlag | (g.end - g.begin);
        continue;
      }
      w[7] = g.axis;
      st.push_back({g.right, at});
      st.push_back({g.left, ~0u});
    }
    const uint32_t total = slot;
    out.nodes.swap(words);
    out.reordered_indices.assign((size_t)n_tri * index_stride, 0u);
    out.reordered_mat.resize(n_tri);
    out.reordered_to_source.resize(n_tri);
    for (uint32_t nt = 0; nt < n_tri; ++nt) {
      const uint32_t ot = perm[tri_base + nt] - tri_base;
      for (int k = 0; k < 3; ++k)
        out.reordered_indices[(size_t)nt * index_stride + k] = indices[(size_t)ot * index_stride + k];
      out.reordered_mat[nt] = tri_material_ids ? tri_material_ids[ot] : 0u;
      out.reordered_to_source[nt] = ot;
    }
    return vt::validate_bvh(out.nodes.data(), total, nullptr, 0, n_tri, vt::kBvhStackDepth);
}

}  // namespace

namespace vtg {

vt::BuildError build_blas_device(vt_engine* eng, const float* positions, uint32_t position_floats,
                                 uint32_t position_stride, const uint32_t* indices, uint32_t index_words,
                                 uint32_t index_stride, const uint32_t* tri_material_ids, uint32_t max_leaf,
                                 uint32_t num_bins, uint32_t max_depth, vt::BlasOut& out) {
  if (vt::BuildError e = vt::check_blas_inputs(positions, position_floats, position_stride, indices, index_words,
                                               index_stride, max_leaf, num_bins, max_depth))
    return e;
  const uint32_t n_tri = index_words / index_stride;
  out.nodes.clear();
  if (n_tri == 0) {
    out.nodes.assign(8, 0u);
    out.nodes[7] = vt::kLeafFlag;
    out.reordered_indices.assign(indices, indices + index_words);
    out.reordered_mat.clear();
    out.reordered_to_source.clear();
    return {};
  }
  std::vector<GNode> nodes;
  std::vector<uint32_t> perm;
  if (vt::BuildError e = build_core(eng, positions, position_floats, position_stride, indices, index_words,
                                    index_stride, {{0u, n_tri}}, max_leaf, num_bins, max_depth, nodes, perm))
    return e;
  return emit_tree(nodes, 0, perm, 0, n_tri, indices, index_stride, tri_material_ids, out);
}

vt::BuildError build_blas_forest(vt_engine* eng, const std::vector<vt::BlasJobIn>& meshes, uint32_t max_leaf,
                                 uint32_t num_bins, uint32_t max_depth, std::vector<vt::BlasOut>& outs) {
  outs.assign(meshes.size(), vt::BlasOut());
  std::vector<float> pos;
  std::vector<uint32_t> idx;
  std::vector<std::pair<uint32_t, uint32_t>> roots;
  std::vector<uint32_t> root_of(meshes.size(), ~0u), tri_base(meshes.size(), 0u);
  size_t vsum = 0, tsum = 0;
  for (const vt::BlasJobIn& m : meshes) vsum += m.position_floats, tsum += m.index_words;
  if (vsum / 4 > 0xFFFFFFFFull || tsum / 4 > 0xFFFFFFFFull)
    return vt::BuildError{VT_ERR_RANGE, "device BLAS build: the meshes exceed 2^32 vertices or triangles"};
  pos.reserve(vsum);
  idx.reserve(tsum);
  for (size_t i = 0; i < meshes.size(); ++i) {
    const vt::BlasJobIn& m = meshes[i];
    if (vt::BuildError e = vt::check_blas_inputs(m.pos4, m.position_floats, 4, m.idx4, m.index_words, 4, max_leaf,
                                                 num_bins, max_depth))
      return e;
    const uint32_t nt = m.index_words / 4, vbase = (uint32_t)(pos.size() / 4);
    if (nt == 0) {
      outs[i].nodes.assign(8, 0u);
      outs[i].nodes[7] = vt::kLeafFlag;
      continue;
    }
    tri_base[i] = (uint32_t)(idx.size() / 4);
    root_of[i] = (uint32_t)roots.size();
    roots.push_back({tri_base[i], tri_base[i] + nt});
    pos.insert(pos.end(), m.pos4, m.pos4 + m.position_floats);
    for (uint32_t t = 0; t < nt; ++t) {
      for (int k = 0; k < 3; ++k) idx.push_back(m.idx4[4 * (size_t)t + k] + vbase);
      idx.push_back(0u);
    }
  }
  if (roots.empty()) return {};
  std::vector<GNode> nodes;
  std::vector<uint32_t> perm;
  if (vt::BuildError e = build_core(eng, pos.data(), (uint32_t)pos.size(), 4, idx.data(), (uint32_t)idx.size(), 4,
                                    roots, max_leaf, num_bins, max_depth, nodes, perm))
    return e;
  for (size_t i = 0; i < meshes.size(); ++i) {
    if (root_of[i] == ~0u) continue;
    const vt::BlasJobIn& m = meshes[i];
    if (vt::BuildError e = emit_tree(nodes, root_of[i], perm, tri_base[i], m.index_words / 4, m.idx4, 4, m.mats,
                                     outs[i]))
      return e;
  }
  return {};
}

}  // namespace vtg
// Same arguments, checks, outputs and errors as vt_build_blas; the SAH work
// runs on the engine's device.
vt_status vt_build_blas_device(vt_engine* eng, const float* positions, uint32_t position_floats,
                               uint32_t position_stride, const uint32_t* indices, uint32_t index_words,
                               uint32_t index_stride, const uint32_t* tri_material_ids, uint32_t max_leaf,
                               uint32_t num_bins, uint32_t max_depth, uint32_t* out_node_words,
                               uint32_t* out_node_count, uint32_t* out_reordered_indices,
                               uint32_t* out_reordered_mat, uint32_t* out_reordered_to_source) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return engine_error(nullptr, VT_ERR_INVALID_ARGUMENT, "NULL engine");
    if ((position_floats && !positions) || (index_words && !indices) || !out_node_words || !out_node_count ||
        (index_words && !out_reordered_indices))
      return engine_error(eng, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    vt::BlasOut o;
    vt::BuildError e = vtg::build_blas_device(eng, positions, position_floats, position_stride, indices, index_words,
                                              index_stride, tri_material_ids, max_leaf, num_bins, max_depth, o);
    if (e) return engine_error(eng, e.code, e.message);
    std::memcpy(out_node_words, o.nodes.data(), o.nodes.size() * 4);
    *out_node_count = (uint32_t)(o.nodes.size() / 8);
    if (!o.reordered_indices.empty())
      std::memcpy(out_reordered_indices, o.reordered_indices.data(), o.reordered_indices.size() * 4);
    if (out_reordered_mat && !o.reordered_mat.empty())
      std::memcpy(out_reordered_mat, o.reordered_mat.data(), o.reordered_mat.size() * 4);
    if (out_reordered_to_source && !o.reordered_to_source.empty())
      std::memcpy(out_reordered_to_source, o.reordered_to_source.data(), o.reordered_to_source.size() * 4);
    return VT_OK;
  });
}
