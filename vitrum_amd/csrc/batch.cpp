// batch.cpp -- many scenes in one persistent closest-hit launch, its HIP graph
// and the launch-parameter autotuner.
//
// A batch owns the traversal buffers of each scene and a device table of the
// scenes' buffer addresses.  The rays of all scenes sit in one array, grouped
// by scene (ray_offsets[k] .. ray_offsets[k+1] belong to scene k); every lane
// of the persistent kernel looks the scene of a new ray up in that table, so
// scenes of very different sizes share one launch and one load-balanced work
// queue.  Hit records are the per-scene ones (primitive / instance indices of
// that scene).
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "engine_internal.h"
#include "traverse.h"

struct vt_batch {
  vt_engine* eng = nullptr;
  std::vector<std::unique_ptr<SceneBuffers>> scenes;
  std::vector<vt::DeviceLayout> layouts;
  DeviceBuffer table, offsets;
  std::vector<uint32_t> host_offsets;
  bool table_dirty = true;
  bool wide = false;  // every scene's Node128 trees fit the kernel's stack
  // Launch parameters (the autotuner's choice).
  uint32_t refill_min = 8, steps = 16, near_first = 1, defer_min = 12, blocks_per_cu = 1;
  // Captured graph of (counter reset, batched trace).
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string error;
};

static vt_status berr(vt_batch* b, vt_status s, const std::string& m) {
  b->eng->last_error = m;
  return s;
}

#define B_HIP(b, call)                                                         \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess) return berr((b), VT_ERR_HIP, std::string(#call ": ") + hipGetErrorString(_e)); \
  } while (0)

static void drop_graph(vt_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static vt_status upload_table(vt_batch* b) {
  if (!b->table_dirty) return VT_OK;
  std::vector<vtk::DeviceScene> t(b->scenes.size());
  bool all_wide = true;
  for (size_t k = 0; k < t.size(); ++k) {
    bool wide = false;
    t[k] = engine_device_scene(*b->scenes[k], b->layouts[k], &wide);
    all_wide = all_wide && wide;
  }
  b->wide = all_wide;
  if (vt_status s = engine_upload(b->eng, b->table, t.data(), t.size() * sizeof(vtk::DeviceScene))) return s;
  b->table_dirty = false;
  return VT_OK;
}

static vt_status set_offsets(vt_batch* b, const uint32_t* ray_offsets, uint32_t* n_out) {
  const size_t k = b->scenes.size();
  if (k == 0) return berr(b, VT_ERR_STATE, "the batch has no scenes");
  if (!ray_offsets) return berr(b, VT_ERR_INVALID_ARGUMENT, "ray_offsets is NULL");
  if (ray_offsets[0] != 0) return berr(b, VT_ERR_RANGE, "ray_offsets[0] must be 0");
  for (size_t i = 0; i < k; ++i)
    if (ray_offsets[i + 1] < ray_offsets[i]) return berr(b, VT_ERR_RANGE, "ray_offsets must be non-decreasing");
  if (ray_offsets[k] > 0x7FFFFFFFu) return berr(b, VT_ERR_RANGE, "too many rays");
  std::vector<uint32_t> v(ray_offsets, ray_offsets + k + 1);
  if (v != b->host_offsets) {
    if (vt_status s = engine_upload(b->eng, b->offsets, v.data(), v.size() * 4)) return s;
    b->host_offsets = v;
  }
  *n_out = ray_offsets[k];
  return upload_table(b);
}

static vtk::TraceParams batch_params(vt_batch* b, const vt_ray* d_rays, vt_hit* d_hits, uint32_t n) {
  vt_engine* e = b->eng;
  vtk::TraceParams p{};
  p.rays = (const float4*)d_rays;
  p.hits = (uint32_t*)d_hits;
  p.n = n;
  p.counter = (uint32_t*)e->counter.ptr;
  p.error = (uint32_t*)e->error_word.ptr;
  p.spill = (uint2*)e->spill.ptr;
  p.refill_min = b->refill_min;
  p.steps = b->steps;
  p.near_first = b->near_first;
  p.defer_min = b->defer_min;
  p.wide = (b->wide && e->node_width == 4) ? 1u : 0u;
  p.scenes = (const vtk::DeviceScene*)b->table.ptr;
  p.scene_offsets = (const uint32_t*)b->offsets.ptr;
  p.scene_count = (uint32_t)b->scenes.size();
  return p;
}

static uint32_t batch_grid(vt_batch* b) { return b->eng->cu_count * b->blocks_per_cu; }

static vt_status check_overflow(vt_batch* b) {
  uint32_t err = 0;
  B_HIP(b, hipMemcpyAsync(&err, b->eng->error_word.ptr, 4, hipMemcpyDeviceToHost, b->eng->stream));
  B_HIP(b, hipStreamSynchronize(b->eng->stream));
  if (err) return berr(b, VT_ERR_HIP, "traversal stack capacity exceeded");
  return VT_OK;
}

extern "C" {

vt_status vt_batch_create(vt_engine* eng, vt_batch** out) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng || !out) return VT_ERR_INVALID_ARGUMENT;
    auto b = std::make_unique<vt_batch>();
    b->eng = eng;
    int per_cu = 0;
    if (hipSetDevice(eng->device) != hipSuccess || vtk::trace_occupancy_batch(&per_cu) != hipSuccess)
      return VT_ERR_HIP;
    // The spill area is sized for the engine's own grid: never exceed it.
    b->blocks_per_cu = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)per_cu, eng->trace_blocks_per_cu));
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) return VT_ERR_HIP;
    *out = b.release();
    return VT_OK;
  });
}

vt_status vt_batch_destroy(vt_batch* b) {
  return vt_guarded(b ? b->eng : nullptr, [&]() -> vt_status {
    if (!b) return VT_OK;
    (void)hipSetDevice(b->eng->device);
    (void)hipStreamSynchronize(b->eng->stream);
    drop_graph(b);
    for (auto& s : b->scenes) s->release();
    b->table.release();
    b->offsets.release();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
    return VT_OK;
  });
}

vt_status vt_batch_add_scene(vt_batch* b, const vt_scene_desc* scene, uint32_t* index) {
  return vt_guarded(b ? b->eng : nullptr, [&]() -> vt_status {
    if (!b || !scene) return VT_ERR_INVALID_ARGUMENT;
    auto buf = std::make_unique<SceneBuffers>();
    vt::DeviceLayout layout;
    if (vt_status s = engine_build_trace_scene(b->eng, scene, *buf, layout)) return s;
    if (index) *index = (uint32_t)b->scenes.size();
    b->scenes.push_back(std::move(buf));
    b->layouts.push_back(std::move(layout));
    b->table_dirty = true;
    drop_graph(b);
    return VT_OK;
  });
}

vt_status vt_batch_scene_count(const vt_batch* b, uint32_t* count) {
  return vt_guarded(b ? b->eng : nullptr, [&]() -> vt_status {
    if (!b || !count) return VT_ERR_INVALID_ARGUMENT;
    *count = (uint32_t)b->scenes.size();
    return VT_OK;
  });
}

vt_status vt_batch_trace_closest_device(vt_batch* b, const vt_ray* d_rays, const uint32_t* ray_offsets,
                                        vt_hit* d_hits, float* ms) {
  return vt_guarded(b ? b->eng : nullptr, [&]() -> vt_status {
    if (!b) return VT_ERR_INVALID_ARGUMENT;
    uint32_t n = 0;
    if (vt_status s = set_offsets(b, ray_offsets, &n)) return s;
    if (n && (!d_rays || !d_hits)) return berr(b, VT_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (ms) *ms = 0.0f;
    if (n == 0) return VT_OK;
    vt_engine* e = b->eng;
    B_HIP(b, hipSetDevice(e->device));
    B_HIP(b, hipEventRecord(b->ev0, e->stream));
    B_HIP(b, vtk::launch_trace_batch(batch_params(b, d_rays, d_hits, n), batch_grid(b), e->stream));
    B_HIP(b, hipEventRecord(b->ev1, e->stream));
    if (ms) {
      B_HIP(b, hipEventSynchronize(b->ev1));
      B_HIP(b, hipEventElapsedTime(ms, b->ev0, b->ev1));
    }
    return check_overflow(b);
  });
}

// Records (counter reset, batched trace) into a HIP graph on the engine
// stream.  The graph holds kernel nodes only (no memset nodes); replaying it
// re-traces whatever the ray buffer holds at that time.
vt_status vt_batch_capture(vt_batch* b, const vt_ray* d_rays, const uint32_t* ray_offsets, vt_hit* d_hits) {
  return vt_guarded(b ? b->eng : nullptr, [&]() -> vt_status {
    if (!b) return VT_ERR_INVALID_ARGUMENT;
    uint32_t n = 0;
    if (vt_status s = set_offsets(b, ray_offsets, &n)) return s;
    if (n == 0 || !d_rays || !d_hits) return berr(b, VT_ERR_INVALID_ARGUMENT, "empty batch or NULL buffer");
    vt_engine* e = b->eng;
    B_HIP(b, hipSetDevice(e->device));
    B_HIP(b, hipStreamSynchronize(e->stream));
    drop_graph(b);
    const vtk::TraceParams p = batch_params(b, d_rays, d_hits, n);
    B_HIP(b, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = vtk::launch_trace_batch(p, batch_grid(b), e->stream);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return berr(b, VT_ERR_HIP, std::string("graph capture failed: ") +
                                     hipGetErrorString(le != hipSuccess ? le : ce));
    }
    b->graph = g;
    B_HIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return VT_OK;
  });
}

vt_status vt_batch_replay(vt_batch* b, float* ms) {
  return vt_guarded(b ? b->eng : nullptr, [&]() -> vt_status {
    if (!b) return VT_ERR_INVALID_ARGUMENT;
    if (!b->exec) return berr(b, VT_ERR_STATE, "no graph has been captured");
    vt_engine* e = b->eng;
    B_HIP(b, hipSetDevice(e->device));
    B_HIP(b, hipEventRecord(b->ev0, e->stream));
    B_HIP(b, hipGraphLaunch(b->exec, e->stream));
    B_HIP(b, hipEventRecord(b->ev1, e->stream));
    if (ms) {
      B_HIP(b, hipEventSynchronize(b->ev1));
      B_HIP(b, hipEventElapsedTime(ms, b->ev0, b->ev1));
    }
    return check_overflow(b);
  });
}

// Times every candidate (median of `repeats` launches on the caller's rays)
// and keeps the fastest: wave refill threshold, steps per refill, child
// order (reference left-first or nearest-first inside BLASes, both exact) and
// resident blocks per CU.  Any captured graph is dropped (re-capture after).
vt_status vt_batch_autotune(vt_batch* b, const vt_ray* d_rays, const uint32_t* ray_offsets, vt_hit* d_hits,
                            uint32_t repeats, vt_tuning* chosen) {
  return vt_guarded(b ? b->eng : nullptr, [&]() -> vt_status {
    if (!b || !chosen) return VT_ERR_INVALID_ARGUMENT;
    if (repeats == 0 || repeats > 15) return berr(b, VT_ERR_RANGE, "repeats must be in [1, 15]");
    uint32_t n = 0;
    if (vt_status s = set_offsets(b, ray_offsets, &n)) return s;
    if (n == 0 || !d_rays || !d_hits) return berr(b, VT_ERR_INVALID_ARGUMENT, "empty batch or NULL buffer");
    drop_graph(b);
    const uint32_t max_bpc = b->blocks_per_cu;
    const uint32_t refills[] = {8, 16, 32}, steps[] = {16, 48}, orders[] = {1, 0}, defers[] = {1, 12};
    vt_tuning best{};
    best.ms = -1.0f;
    for (uint32_t bpc = max_bpc; bpc >= 1 && bpc + 2 > max_bpc; --bpc)
      for (uint32_t r : refills)
        for (uint32_t st : steps)
          for (uint32_t o : orders)
            for (uint32_t d : defers) {
              b->refill_min = r;
              b->steps = st;
              b->near_first = o;
              b->defer_min = d;
              b->blocks_per_cu = bpc;
              std::vector<float> t;
              for (uint32_t k = 0; k < repeats; ++k) {
                float ms = 0;
                if (vt_status s = vt_batch_trace_closest_device(b, d_rays, ray_offsets, d_hits, &ms)) return s;
                t.push_back(ms);
              }
              std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
              const float med = t[t.size() / 2];
              if (best.ms < 0 || med < best.ms) best = vt_tuning{r, st, o, d, bpc, med};
            }
    b->refill_min = best.refill_min;
    b->steps = best.steps;
    b->near_first = best.near_first;
    b->defer_min = best.defer_min;
    b->blocks_per_cu = best.blocks_per_cu;
    *chosen = best;
    return VT_OK;
  });
}

vt_status vt_batch_set_tuning(vt_batch* b, const vt_tuning* t) {
  return vt_guarded(b ? b->eng : nullptr, [&]() -> vt_status {
    if (!b || !t) return VT_ERR_INVALID_ARGUMENT;
    if (t->refill_min < 1 || t->refill_min > 64 || t->steps < 1 || t->near_first > 1 || t->defer_min < 1 ||
        t->defer_min > 64 || t->blocks_per_cu < 1 ||
        t->blocks_per_cu > b->eng->trace_blocks_per_cu)
      return berr(b, VT_ERR_RANGE, "tuning out of range");
    b->refill_min = t->refill_min;
    b->steps = t->steps;
    b->near_first = t->near_first;
    b->defer_min = t->defer_min;
    b->blocks_per_cu = t->blocks_per_cu;
    drop_graph(b);
    return VT_OK;
  });
}

}  // extern "C"
