// noalloc.hip -- closest-hit traversal that allocates nothing: the launch's
// work counter, error word and stack-spill area live in `scratch`, a fixed
// device array of this file (one copy per GPU, held until the process ends),
// instead of in the engine's allocations.  For callers that cannot allocate
// device memory (code under stream capture, tight serving loops).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine_internal.h"
#include "traverse.h"

namespace {

// Layout of `scratch`: words [0, 16) the launch's ray counters, word 16 the
// error word, then the stack-spill area from byte 256 on.
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr size_t kSpillOffset = 256;
constexpr size_t kBlockSpillBytes =
    (size_t)(vtk::kStackCapacity - vtk::kLdsDepth) * vtk::kTraceBlock * sizeof(uint2);  // 110,592
constexpr uint32_t kMaxBlocks = (uint32_t)((kScratchBytes - kSpillOffset) / kBlockSpillBytes);  // 606
constexpr int kMaxDevices = 64;

std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};  // device address of each GPU's copy

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Looks up this GPU's copy of `scratch` (once per device, with the device
// current).  The array's host-side name is not a device address; only the
// pointer returned here is passed to kernels.
vt_status engine_noalloc_setup(vt_engine* e) {
  if (e->device < 0 || e->device >= kMaxDevices)
    return engine_error(e, VT_ERR_RANGE, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[e->device]) {
    hipError_t r = hipSetDevice(e->device);
    void* p = nullptr;
    if (r == hipSuccess) r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !p)
      return engine_error(e, VT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r));
    g_scratch[e->device] = p;
  }
  e->noalloc_scratch = g_scratch[e->device];
  return VT_OK;
}

extern "C" vt_status vt_trace_closest_device_noalloc(vt_engine* eng, const vt_ray* d_rays, uint32_t n,
                                                     vt_hit* d_hits) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return engine_error(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
    if (!eng->has_scene) return engine_error(eng, VT_ERR_STATE, "no scene is set");
    if (eng->opts.bvh_traversal == 1)
      return engine_error(eng, VT_ERR_STATE, "the allocation-free trace walks the binary BVH only");
    if (n && (!d_rays || !d_hits)) return engine_error(eng, VT_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (n > 0x7FFFFFFFu) return engine_error(eng, VT_ERR_RANGE, "at most 2^31 - 1 rays per call");
    if (!eng->noalloc_scratch) return engine_error(eng, VT_ERR_STATE, "scratch was not set up");
    hipError_t r = hipSetDevice(eng->device);
    if (r != hipSuccess) return engine_error(eng, VT_ERR_HIP, hipGetErrorString(r));
    if (n == 0) return VT_OK;
    // Persistent grid: as many blocks as the spill area holds, a whole
    // number of blocks per CU, at most the engine's own grid.
    uint32_t grid = eng->trace_grid < kMaxBlocks ? eng->trace_grid : kMaxBlocks;
    if (grid > eng->cu_count) grid -= grid % eng->cu_count;
    unsigned char* base = (unsigned char*)eng->noalloc_scratch;
    vtk::TraceParams p = engine_trace_params(eng);
    p.counter = (uint32_t*)base;
    p.error = (uint32_t*)base + 16;
    p.spill = (uint2*)(base + kSpillOffset);
    p.rays = (const float4*)d_rays;
    p.hits = (uint32_t*)d_hits;
    p.occluded = nullptr;
    p.counts = nullptr;
    p.n = n;
    // The error word starts each call at zero (a fill kernel, so the call
    // can be captured); launch_trace zeroes the counter the same way.
    r = vtk::launch_zero(p.error, 1, eng->stream);
    if (r == hipSuccess) r = vtk::launch_trace(p, false, false, grid, eng->stream);
    if (r != hipSuccess) return engine_error(eng, VT_ERR_HIP, hipGetErrorString(r));
    eng->last_ray_count = n;
    return VT_OK;
  });
}
