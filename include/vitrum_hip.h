/*
 * vitrum_hip.h -- C ABI of the MI355X path-tracing backend (libvthip.so).
 *
 * The reference's plug-in boundary is the TypeScript `Engine` contract
 * (packages/core/src/engine/index.ts:109-486) created by an `EngineFactory`
 * (packages/core/src/engine/factory.ts) and registered in BACKEND_CONSTRUCTORS
 * (packages/engine/src/createEngine.ts:99-103).  These entry points are what a
 * Node N-API binding of that contract calls; each cites the reference
 * interface it replaces.
 *
 * Conventions
 *  - Every function returns vt_status (0 = success, negative = error).  The
 *    message of the last failure on an engine is available through
 *    vt_last_error(); engine-less functions report through vt_last_error(NULL).
 *  - Host pointers are borrowed for the duration of the call only.
 *  - "_device" entry points take device pointers and enqueue work on the
 *    engine's HIP stream without synchronising; everything else is blocking.
 *  - Arguments are validated before any state changes (the reference throws
 *    TypeError/RangeError before mutation, buildArrayBvh.ts:348-408).
 *  - One engine per thread at a time; calls on one engine must not overlap.
 */
#ifndef VITRUM_HIP_H
#define VITRUM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t vt_status;
enum {
  VT_OK = 0,
  VT_ERR_INVALID_ARGUMENT = -1, /* TypeError in the reference */
  VT_ERR_RANGE = -2,            /* RangeError in the reference */
  VT_ERR_BUILD = -3,            /* Error thrown by a builder/validator */
  VT_ERR_HIP = -4,              /* EngineError{kind:'gpu-internal'} */
  VT_ERR_NO_DEVICE = -5,        /* EngineError{kind:'device-lost'} at creation */
  VT_ERR_STATE = -6,            /* call made in the wrong engine state */
  VT_ERR_OUT_OF_MEMORY = -7
};

typedef struct vt_engine vt_engine;

/* ---- rays and hits ------------------------------------------------------- */

/* 32-byte ray: origin, tmin, unit direction, tmax. */
typedef struct vt_ray {
  float o[3];
  float tmin;
  float d[3];
  float tmax;
} vt_ray;

/* 20-byte closest-hit record.  prim is the global packed (BVH-reordered)
 * triangle index and inst the packed instance index, the pair
 * SceneHit{triIndex, instanceIndex} of
 * packages/pt-webgpu/src/wgsl/pathTrace/intersectionCore.wgsl.ts:33-60;
 * u, v are its baryVW.  A miss has inst == 0xFFFFFFFF and t == +inf. */
typedef struct vt_hit {
  uint32_t prim;
  uint32_t inst;
  float t;
  float u;
  float v;
} vt_hit;

/* ---- scene description (packages/core/src/scene/index.ts:42) ----------- */

enum { VT_PRIM_MESH = 0, VT_PRIM_INSTANCED_MESH = 1 };

/* Mesh / instanced mesh primitive (packages/core/src/scene/primitives.ts:
 * 167-227).  positions/normals are xyz float triples; indices are optional
 * (NULL = generated 0..vertexCount-1); transforms are column-major mat4s (a
 * mesh takes 0 or 1, an instanced mesh one per instance). */
typedef struct vt_primitive {
  uint32_t kind;
  uint32_t material;
  const float *positions;
  uint32_t position_floats;
  const float *normals;
  uint32_t normal_floats;
  const uint32_t *indices;
  uint32_t index_count;
  const float *transforms;
  uint32_t transform_count;
} vt_primitive;

/* MaterialSpec subset rendered by the opaque BSDF: the Lambert + GGX base
 * and the clearcoat, sheen and anisotropy extension lobes
 * (packages/core/src/scene/material.ts:167-458; bsdf.wgsl.ts:392-916).
 * The extension scalars are clamped as materialPacking.ts:249-316 packs
 * them; 0 turns a lobe off, so a zero-initialised extension block renders
 * the base BSDF.  Non-finite values are VT_ERR_RANGE at vt_set_scene. */
typedef struct vt_material {
  float base_color[3];
  float roughness;
  float metallic;
  float emissive[3];
  float emissive_intensity;
  uint32_t double_sided;
  float clearcoat, clearcoat_roughness;
  float sheen, sheen_roughness, sheen_color[3];
  float anisotropy, anisotropy_rotation; /* strength in [0,1]; radians */
  /* KHR_materials_specular: used when has_specular is non-zero, else the
   * defaults specularColor (1,1,1) and specularIntensity 1. */
  uint32_t has_specular;
  float specular_color[3], specular_intensity;
} vt_material;

/* Environment (packages/core/src/scene/environment.ts:16-37).
 *   VT_ENV_NONE:     no environment (misses are black).
 *   VT_ENV_CONSTANT: constant radiance[] in every direction (a constant
 *                    64 x 32 HDRI; the importance table is built for that).
 *   VT_ENV_HDRI:     equirectangular map HdriEnvironment.hdri as linear
 *                    binary32 texels, width x height x channels, row-major,
 *                    row 0 at +Y (theta 0), column 0 at phi = -pi, phi
 *                    measured from +X towards +Z (connect.wgsl.ts:126-153).
 *                    channels 1 (grey), 2 (r, g; b = r), 3 (rgb) or 4
 *                    (rgba, alpha ignored), as environmentMapPixels.ts:
 *                    103-160 expands them.  intensity scales the texels
 *                    (HdriEnvironment.intensity), rotation_y is the dome's
 *                    counter-clockwise rotation about +Y in radians.  The map
 *                    is copied; the pointer need not outlive vt_set_scene.  A
 *                    map whose luminance is zero everywhere renders as none
 *                    (environmentPacking.ts:420-440).
 *   VT_ENV_PROCEDURAL_SKY: ProceduralSkyEnvironment (environment.ts:59-71),
 *                    the Preetham daylight model baked on the host into a
 *                    256 x 128 map that is then sampled as an HDRI
 *                    (environmentPacking.ts:222-282, preethamSky.ts:161-313):
 *                    sun_direction (normalised; [0, 1, 0] when ~0),
 *                    turbidity (clamped to [1.5, 30]), rayleigh, mie_coefficient,
 *                    mie_directional_g (clamped to +-0.999999) and intensity. */
typedef enum vt_environment_kind {
  VT_ENV_NONE = 0,
  VT_ENV_CONSTANT = 1,
  VT_ENV_HDRI = 2,
  VT_ENV_PROCEDURAL_SKY = 3
} vt_environment_kind;

typedef struct vt_environment {
  uint32_t kind;
  float radiance[3];     /* VT_ENV_CONSTANT */
  const float *texels;   /* VT_ENV_HDRI: width * height * channels floats */
  uint32_t width, height, channels;
  float intensity;       /* VT_ENV_HDRI, VT_ENV_PROCEDURAL_SKY; >= 0 */
  float rotation_y;      /* VT_ENV_HDRI; radians */
  double sun_direction[3];                 /* VT_ENV_PROCEDURAL_SKY */
  double turbidity, rayleigh, mie_coefficient, mie_directional_g;
} vt_environment;

/* Scene.emitters (packages/core/src/scene/emitters.ts:7-114): the analytic
 * emitter kinds, packed as emitterPacking.ts:600-760 packs them.  Fields a
 * kind does not use are ignored.  radiance = f32(color) x f32(intensity). */
typedef enum vt_emitter_kind {
  VT_EMITTER_DIRECTIONAL = 0, /* direction (the way the light travels), angular_diameter */
  VT_EMITTER_POINT = 1,       /* position, distance (0: no cutoff), decay (2: inverse square) */
  VT_EMITTER_SPOT = 2,        /* position, direction, angle (half cone), penumbra, distance, decay */
  VT_EMITTER_RECT_AREA = 3,   /* position, u_axis, v_axis (half extents; u x v is the emitting side) */
  VT_EMITTER_DISC_AREA = 4    /* position, normal, radius */
} vt_emitter_kind;

typedef struct vt_emitter {
  uint32_t kind;
  float color[3];
  float intensity;
  uint32_t cast_shadow; /* 0: no shadow rays (castShadow: false) */
  float position[3];
  float direction[3];
  float normal[3];
  float u_axis[3];
  float v_axis[3];
  float radius;
  float angle;
  float penumbra;
  float distance;
  float decay;
  float angular_diameter;
} vt_emitter;

typedef struct vt_scene_desc {
  const vt_primitive *primitives;
  uint32_t primitive_count;
  const vt_material *materials;
  uint32_t material_count;
  vt_environment environment;
  float scene_radius; /* ray-origin bias scale (sceneScalePolicy.ts:83-123);
                         0 = derive from the scene bounds */
  const vt_emitter *emitters; /* analytic emitters; mesh-area light comes from emissive materials */
  uint32_t emitter_count;
} vt_scene_desc;

/* bakePreethamSkyEquirect (shared-samplers/src/preethamSky.ts:161-313) on
 * the host, no device needed: the procedural sky of `env` (kind
 * VT_ENV_PROCEDURAL_SKY) as width x height binary32 rgb texels (3 floats per
 * texel into rgb_out, row 0 at +Y) and the texels' luminance x solid-angle
 * integral.  VT_ERR_RANGE with the reason for invalid parameters. */
vt_status vt_bake_procedural_sky(const vt_environment *env, uint32_t width, uint32_t height, float *rgb_out,
                                 double *luminance_integral);

/* ---- engine lifecycle ---------------------------------------------------- */

/* EngineOptions (packages/core/src/engine/factory.ts) + the pt-webgpu
 * options this backend honours (packages/pt-webgpu/src/index.ts:219-364). */
typedef struct vt_options {
  uint32_t max_bounces;           /* 0 = default 8 */
  uint32_t max_samples_per_pixel; /* 0 = unlimited */
  uint32_t sampling;              /* 0 = pcg, 1 = sobol */
  /* NEE light selection: 0 = the spatial light tree when the scene has two
   * or more lights (lightTreeImportanceSampling, the reference's default on a
   * full-tier device), 1 = uniform (lightTreeImportanceSampling: false). */
  uint32_t light_selection;
  /* Closest-hit traversal: 0 = the binary BVH (pt-webgpu's default), 1 =
   * bvhTraversal 'cwbvh-closest' (8-wide compressed BVH with a
   * Moller-Trumbore proposal re-verified by the watertight test). */
  uint32_t bvh_traversal;
  uint32_t reserved[3];
} vt_options;

/* createPTEngine_WebGPU (packages/pt-webgpu/src/index.ts:2363).  The device
 * ordinal is chosen by the host; the engine never resets the device. */
vt_status vt_create(int32_t device_ordinal, const vt_options *opts, vt_engine **out);
/* Engine.dispose (packages/core/src/engine/index.ts:348).  NULL is a no-op. */
vt_status vt_destroy(vt_engine *eng);
/* Copies the last error message (NUL-terminated, truncated to len). */
vt_status vt_last_error(const vt_engine *eng, char *buf, size_t len);

/* Engine.setScene (packages/core/src/engine/index.ts:120;
 * packages/pt-webgpu/src/index.ts:1120; atomic rebuild at :1229-1260).
 * Validates (material indices, Le = emissive x emissiveIntensity finite and
 * >= 0 in f32), packs and builds on the host (BLAS per primitive, TLAS over
 * instances), converts to the device layout, and builds and uploads the
 * traversal buffers and the shading data (materials, emitters, environment
 * table, light tree) into fresh allocations.  Only when every step has
 * succeeded is the new scene bound and accumulation reset.  Transactional:
 * on any failure -- a validation or build error, an allocation failure on
 * the host (VT_ERR_OUT_OF_MEMORY) or the device -- the previous scene stays
 * bound and renders as before. */
vt_status vt_set_scene(vt_engine *eng, const vt_scene_desc *scene);

/* Host threads of the scene builder (no counterpart in the reference, whose
 * builder is single-threaded JavaScript): 0 (default) = every hardware
 * thread; a host that shares its CPUs passes its own share. */
vt_status vt_set_build_threads(vt_engine *eng, uint32_t threads);

/* Meshes of at least `min_triangles` triangles get their BLAS from the device
 * builder (vt_build_blas_device) when a scene is bound; 0 (default) builds
 * every BLAS on the host.  The bound scene is the same either way (no
 * counterpart in the reference, whose BLAS build is buildArrayBvh.ts:342 in
 * JavaScript). */
vt_status vt_set_device_build(vt_engine *eng, uint32_t min_triangles);

/* Test hook (no counterpart in the reference): each bit of `flags` makes the
 * next matching step fail once -- 1: scene packing throws std::bad_alloc
 * (vt_set_scene on `eng`; vt_pack_scene when `eng` is NULL), 2: the
 * shading-data upload of vt_set_scene reports VT_ERR_OUT_OF_MEMORY, 4: its
 * preparation throws std::runtime_error.  Every entry point catches
 * exceptions and returns VT_ERR_OUT_OF_MEMORY (std::bad_alloc) or
 * VT_ERR_BUILD (any other), so none unwinds into the caller. */
vt_status vt_debug_inject_failure(vt_engine *eng, uint32_t flags);

/* ---- frames ---------------------------------------------------------------- */

/* FrameInput (packages/core/src/frame.ts:91-105) reduced to what the path
 * tracer consumes: the inverse of projMatrix * viewMatrix (column-major), the
 * viewport, frameIndex / frameSeed and quality.bounces / samplesTarget. */
typedef struct vt_frame {
  float inv_view_proj[16];
  uint32_t width;
  uint32_t height;
  uint32_t frame_index;
  uint32_t frame_seed;
  uint32_t max_bounces;    /* 0 = the engine's max_bounces */
  uint32_t samples_target; /* 0 = unlimited */
} vt_frame;

/* Engine.renderFrame (packages/core/src/engine/index.ts:251;
 * packages/pt-webgpu/src/index.ts:1817): one sample per pixel accumulated
 * into the running mean.  Enqueued on the engine stream; returns without
 * waiting.  A changed viewport restarts accumulation. */
vt_status vt_render_frame(vt_engine *eng, const vt_frame *frame);

/* Batched frames (no counterpart in the reference, whose engines render one
 * scene each): several engines, each with its own
 * scene, camera and viewport, render one frame each with every bounce's
 * closest-hit rays of all of them traced in ONE persistent launch over the
 * scenes' table (the shadow rays in another), all on the host engine's
 * stream.  A member's image is bit-identical to rendering the same frame on
 * the member alone.  vt_render_batch_capture records a frame into a HIP
 * graph; vt_render_batch_replay replays it for another frame index (each
 * member keeps its captured camera, viewport and seed; the frame's seed,
 * index and Sobol key are read from device memory), so a frame costs one
 * graph launch.  Members: engines on the host's device with a scene set,
 * the binary traversal and no radiance cache.  A member's own stream is
 * ordered after each batched frame. */
typedef struct vt_render_batch vt_render_batch;
vt_status vt_render_batch_create(vt_engine *host, vt_render_batch **out);
vt_status vt_render_batch_destroy(vt_render_batch *batch);
vt_status vt_render_batch_add(vt_render_batch *batch, vt_engine *member);
/* frames[k] is member k's frame (vt_render_frame's arguments). */
vt_status vt_render_batch_frame(vt_render_batch *batch, const vt_frame *frames);
vt_status vt_render_batch_capture(vt_render_batch *batch, const vt_frame *frames);
/* VT_ERR_STATE when nothing was captured or a member's scene or viewport
 * changed since. */
vt_status vt_render_batch_replay(vt_render_batch *batch, uint32_t frame_index);
/* captureFrame (packages/core/src/engine/index.ts:466; pt-webgpu
 * index.ts:2019): Float32 RGBA running mean, row-major, top-left origin,
 * alpha 1.  Blocking.  If a traversal ran out of stack since the last check,
 * returns VT_ERR_HIP once, clears the condition and leaves `rgba` unwritten. */
vt_status vt_capture(vt_engine *eng, float *rgba, uint32_t width, uint32_t height);

/* FrameQualitySettings.tonemap / exposure / outputColorSpace
 * (packages/core/src/frame.ts:32-105) as the present pass applies them
 * (packages/pt-webgpu/src/wgsl/present.wgsl.ts:58-95). */
typedef struct vt_present {
  uint32_t tonemap;     /* 0 aces (default), 1 agx, 2 reinhard, 3 linear, 4 none */
  float exposure;       /* linear multiplier applied first, >= 0 (default 1) */
  uint32_t color_space; /* 0 srgb (IEC 61966-2-1 OETF, default), 1 linear */
} vt_present;
/* captureFrame({colorSpace: 'output'}) (packages/core/src/engine/index.ts:
 * 466; pt-webgpu index.ts:2019-2035, which reads the present pass's
 * rgba16float texture): the running mean tonemapped, OETF-encoded, clamped to
 * [0, 65504] and rounded to binary16, returned as float32 RGBA (alpha 1).
 * Blocking. */
vt_status vt_capture_output(vt_engine *eng, float *rgba, uint32_t width, uint32_t height,
                            const vt_present *present);
/* FrameRendered.normalDepth / albedo / variance (packages/core/src/frame.ts:
 * 350-421), written by accumulateFrame (packages/pt-webgpu/src/wgsl/
 * pathTrace/kernel.wgsl.ts:920-960).  which 0: first-hit shading normal *
 * 0.5 + 0.5 and hit distance (a miss gives (0.5, 1, 0.5, 0)); 1: first-hit
 * base colour, 1 (a miss gives zeros); both from the latest frame.  2: the
 * luminance variance of the accumulated samples, max(0, E[L^2] - E[L]^2), in
 * the first of the four floats.  4 floats per pixel.  Blocking. */
vt_status vt_capture_aux(vt_engine *eng, uint32_t which, float *out, uint32_t width, uint32_t height);
/* Engine.reset (packages/core/src/engine/index.ts:256): clears accumulation. */
vt_status vt_reset(vt_engine *eng);
/* GPU milliseconds of the last vt_render_frame (HIP events on the engine
 * stream around the whole frame).  Blocking. */
vt_status vt_frame_time(vt_engine *eng, float *ms);
/* Test hook (no counterpart in the reference): a 64-bit hash per pixel of the
 * last frame's discrete path decisions (hit triangles, light picks, lobes,
 * roulette), for path-by-path comparison with a CPU integrator. */
vt_status vt_capture_signatures(vt_engine *eng, uint64_t *sig, uint32_t width, uint32_t height);
/* Mesh-area emitter count and the ray-origin bias of the bound scene. */
vt_status vt_render_info(vt_engine *eng, uint32_t *emitter_count, float *ray_bias);
/* The bound scene's NEE light tree as packLightTreeForGPU lays it out
 * (packages/shared-samplers/src/lightTree.ts:611-737; 16 floats per node,
 * leaves = mesh emitters in order, then the environment); *count = 0 when
 * selection is uniform.  nodes may be NULL to query the count. */
vt_status vt_light_tree(vt_engine *eng, float *nodes, uint32_t capacity, uint32_t *count);

/* ---- traversal (the hot path) -------------------------------------------- */

/* traceTlasClosest (packages/shared-bvh/src/wgsl/tlasSceneHitTraversal.wgsl.ts:
 * 185) with the pt-webgpu instance augmentation (packages/pt-webgpu/src/wgsl/
 * pathTrace/intersection.wgsl.ts:31-152).  Host buffers; blocking. */
vt_status vt_trace_closest(vt_engine *eng, const vt_ray *rays, uint32_t n, vt_hit *hits);
/* traceTlasAny (tlasSceneHitTraversal.wgsl.ts:227): occluded[i] = 0 or 1. */
vt_status vt_trace_any(vt_engine *eng, const vt_ray *rays, uint32_t n, uint32_t *occluded);
/* Device-pointer variants, asynchronous on the engine stream. */
vt_status vt_trace_closest_device(vt_engine *eng, const vt_ray *d_rays, uint32_t n,
                                  vt_hit *d_hits);
vt_status vt_trace_any_device(vt_engine *eng, const vt_ray *d_rays, uint32_t n,
                              uint32_t *d_occluded);
/* No counterpart in the reference.  vt_trace_closest_device for callers that
 * cannot allocate device memory (code under stream capture, serving loops,
 * harnesses that forbid hipMalloc): it allocates nothing.  The launch's ray
 * counter, error word and stack-spill area live in a fixed 64 MiB device
 * array of the library (one copy per GPU, looked up at vt_create and held
 * until the process ends), which bounds the persistent grid at 606 blocks
 * (512 on a 256-CU MI355X); rays and hits are the caller's.  Same hit
 * records as vt_trace_closest_device (binary / 4-wide walk; VT_ERR_STATE in
 * the cwbvh-closest mode); at most 2^31 - 1 rays per call (VT_ERR_RANGE).
 * Enqueued on the engine stream.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other engines' streams, other host
 * threads) race and must be serialised by the caller. */
vt_status vt_trace_closest_device_noalloc(vt_engine *eng, const vt_ray *d_rays, uint32_t n,
                                          vt_hit *d_hits);
/* Work counters of the last closest-hit launch made with counting enabled
 * (vt_set_counting).  counts: 4 u32 per ray (BLAS boxes tested, triangles
 * tested, TLAS boxes tested, instances entered), host buffer. */
vt_status vt_set_counting(vt_engine *eng, int32_t enabled);
vt_status vt_trace_closest_counts(vt_engine *eng, const vt_ray *rays, uint32_t n, vt_hit *hits,
                                  uint32_t *counts);

/* Shading order (no counterpart in the reference; a material sort ahead of
 * shading): 1 = each workgroup of the shading kernel counting-sorts its
 * 256 queue entries by material in LDS (misses first) before shading them;
 * 0 (default) = queue order.  Images are identical either way. */
vt_status vt_set_shade_sort(vt_engine *eng, uint32_t enabled);

/* Traversal scheduling knobs (no counterpart in the reference): a wave is
 * refilled once `refill_min` (1..64) of its lanes are idle, lanes run up to
 * `steps` traversal steps between refill checks, and `near_first` orders the
 * closest-hit traversal: 0 = the reference's left-first order; 1 = nearer
 * child first inside each BLAS, equal triangle distances resolved as the
 * reference resolves them (lowest reordered index), instances in the
 * reference's TLAS order -- hit records identical to the reference's except
 * where two distinct triangles of one BLAS are hit at exactly the same
 * distance and the lower-indexed one's box test rounds above it (the
 * order-dependent pop-time cull can then pick differently); 2 (default) =
 * also nearest-first across instances, with every world bound widened by a
 * near-tie margin and a ray whose hits in two instances come within that
 * margin walked again in the reference's TLAS order, so instances add no
 * further difference.  A lane whose next
 * operation is a BLAS leaf's triangle tests, an instance entry or an instance
 * exit waits until `defer_min` (1..64) lanes of its wave need the same
 * operation (or no lane has node work left), so the wave runs it once for
 * many lanes; every lane's own sequence of operations is unchanged. */
vt_status vt_set_trace_tuning(vt_engine *eng, uint32_t refill_min, uint32_t steps, uint32_t near_first,
                              uint32_t defer_min);

/* Node width of the binary-BVH traversal (no counterpart in the reference):
 * 4 (default) walks 4-wide trees collapsed from the reference's binary BVH
 * and TLAS (a binary node is replaced by its children only when their boxes
 * lie inside its box, so every box test the walk makes is one the reference
 * makes, with the same arithmetic, and the hit records are unchanged); 2
 * walks the binary trees.  Scenes whose 4-wide trees could need more stack
 * than the kernel holds, and the work-counting launches, always use 2. */
vt_status vt_set_node_width(vt_engine *eng, uint32_t width);

/* Hit filter of vt_trace_closest* / vt_trace_any* (the path tracer always
 * uses 1): 0 (default) = the raw traversal (traceClosestRaw and
 * traceTlasAny); 1 = the path tracer's wrappers (traceClosest,
 * packages/pt-webgpu/src/wgsl/pathTrace/intersection.wgsl.ts:462-490, and
 * traceSurfaceVisibility, connect.wgsl.ts:37-103): a hit on the back face of
 * a single-sided opaque material (materialAcceptsSidedHit,
 * material.wgsl.ts:3259-3263; doubleSided defaults to false) is not a
 * surface -- closest hit re-traces beyond it, any hit is not blocked by it.
 * The front face follows the watertight test's side and the instance's
 * orientation (intersection.wgsl.ts:45). */
vt_status vt_set_hit_filter(vt_engine *eng, uint32_t filter);

/* bvhTraversal (packages/pt-webgpu/src/index.ts option; opt-in
 * 'cwbvh-closest', packages/pt-webgpu/src/wgsl/pathTrace/intersection.wgsl.ts:
 * 576-838): 0 = binary, 1 = wide.  Applies to vt_trace_closest* and to the
 * path tracer's extension rays; any-hit rays stay binary, as in the
 * reference. */
vt_status vt_set_bvh_traversal(vt_engine *eng, uint32_t mode);

/* Test hook (no counterpart in the reference): q = a / b and r = sqrt(a)
 * computed on the device, to check that the device's binary32 divide and
 * square root round exactly like the host's. Host buffers. */
vt_status vt_selftest_fp32(vt_engine *eng, const float *a, const float *b, float *q, float *r,
                           uint32_t n);

/* ---- device memory and stream plumbing ----------------------------------- */

vt_status vt_device_alloc(vt_engine *eng, size_t bytes, void **d_ptr);
vt_status vt_device_free(vt_engine *eng, void *d_ptr);
vt_status vt_memcpy_h2d(vt_engine *eng, void *d_dst, const void *src, size_t bytes);
vt_status vt_memcpy_d2h(vt_engine *eng, void *dst, const void *d_src, size_t bytes);
/* Device-to-device copy on the engine stream by the library's own streaming
 * kernel (16-byte vector loads and stores; 16-byte aligned pointers and
 * sizes), so its HBM counters are those of a plain copy of known size. */
vt_status vt_memcpy_d2d(vt_engine *eng, void *d_dst, const void *d_src, size_t bytes);
vt_status vt_synchronize(vt_engine *eng);
/* The engine's hipStream_t, for callers that order their own work with it.
 * The renderer also forks work onto streams of its own and joins it back
 * before the frame's image is complete; a cache training step may still run
 * after that until the next cache use or vt_synchronize. */
vt_status vt_get_stream(vt_engine *eng, void **stream);

/* Kernel timing: milliseconds of the last traversal launch measured with HIP
 * events recorded on the engine stream around that launch. */
typedef struct vt_frame_stats {
  float last_trace_ms;
  float last_frame_ms;
  uint32_t samples_accumulated;
  uint32_t last_ray_count;
  uint64_t device_bytes;
} vt_frame_stats;
/* FrameStats telemetry (packages/core/src/engine/telemetry.ts). */
vt_status vt_stats(vt_engine *eng, vt_frame_stats *out);

/* ---- host-side builders (no device needed) ------------------------------- */

/* buildArrayBvh (packages/shared-bvh/src/buildArrayBvh.ts:342).  Same
 * arguments and outputs; out_node_words holds max(1, 2*triCount-1)*8 words. */
vt_status vt_build_blas(const float *positions, uint32_t position_floats,
                        uint32_t position_stride, const uint32_t *indices,
                        uint32_t index_words, uint32_t index_stride,
                        const uint32_t *tri_material_ids, uint32_t max_leaf,
                        uint32_t num_bins, uint32_t max_depth, uint32_t *out_node_words,
                        uint32_t *out_node_count, uint32_t *out_reordered_indices,
                        uint32_t *out_reordered_mat, uint32_t *out_reordered_to_source);

/* buildArrayBvh (packages/shared-bvh/src/buildArrayBvh.ts:342-780) with the
 * binned-SAH work on the engine's device: the same arguments, checks, error
 * messages and outputs as vt_build_blas, byte for byte.  The tree grows one
 * depth per round (every node of a depth in one launch); the host keeps the
 * node list, sorts the median-fallback ranges and writes the preorder
 * layout.  Its device scratch stays with the engine between calls (it only
 * grows) and is freed by vt_destroy. */
vt_status vt_build_blas_device(vt_engine *eng, const float *positions, uint32_t position_floats,
                               uint32_t position_stride, const uint32_t *indices,
                               uint32_t index_words, uint32_t index_stride,
                               const uint32_t *tri_material_ids, uint32_t max_leaf,
                               uint32_t num_bins, uint32_t max_depth, uint32_t *out_node_words,
                               uint32_t *out_node_count, uint32_t *out_reordered_indices,
                               uint32_t *out_reordered_mat, uint32_t *out_reordered_to_source);

/* buildTlas (packages/shared-bvh/src/tlas.ts:859).  AABBs are binary64, as
 * the reference's JS numbers; out_nodes holds (2n-1)*8 words. */
vt_status vt_build_tlas(uint32_t n, const uint32_t *blas_ids, const double *aabb_min,
                        const double *aabb_max, const float *world_to_local,
                        uint32_t max_leaf, uint32_t num_bins, uint32_t max_depth,
                        uint32_t *out_nodes, uint32_t *out_node_count,
                        uint32_t *out_instance_indices, uint32_t *out_blas_roots,
                        float *out_transforms);

/* packSceneFromCore (packages/shared-bvh/src/scenePack.ts:1104) into a
 * host-side pack whose fields can be read back for inspection. */
typedef struct vt_pack vt_pack;
enum {
  VT_PACK_POSITIONS = 0, VT_PACK_NORMALS, VT_PACK_INDICES, VT_PACK_TRI_MATERIAL,
  VT_PACK_TRI_SOURCE, VT_PACK_TRI_PRIMITIVE, VT_PACK_BVH_NODES, VT_PACK_TLAS_NODES,
  VT_PACK_TLAS_INSTANCE_INDICES, VT_PACK_TLAS_BLAS_ROOTS, VT_PACK_TLAS_W2L,
  VT_PACK_TLAS_L2W, VT_PACK_INSTANCE_PRIMITIVE, VT_PACK_INSTANCE_SOURCE,
  VT_PACK_FIELD_COUNT
};
vt_status vt_pack_scene(const vt_primitive *prims, uint32_t prim_count, uint32_t threads,
                        vt_pack **out);
/* Field size in 32-bit words (negative on a bad field). */
int64_t vt_pack_field_words(const vt_pack *p, int32_t field);
vt_status vt_pack_copy_field(const vt_pack *p, int32_t field, void *dst, uint64_t dst_words);
void vt_pack_free(vt_pack *p);

/* Inspection of the bound scene: the VT_PACK_* fields of its pack (host
 * copies, kept current by refits) and its device traversal trees
 * (VT_SCENE_BLAS_NODE64 / _NODE128: the BLAS forests in the Node64 and
 * Node128 layouts; VT_SCENE_TRIANGLES: the 12-word triangle records).
 * *words receives the field's size in 32-bit words; dst may be NULL to ask
 * for it.  VT_ERR_RANGE when dst_words is too small. */
enum { VT_SCENE_BLAS_NODE64 = 64, VT_SCENE_BLAS_NODE128 = 65, VT_SCENE_TRIANGLES = 66 };
vt_status vt_scene_copy_field(vt_engine *eng, int32_t field, void *dst, uint64_t dst_words, uint64_t *words);

/* refitBvhBounds (packages/shared-bvh/src/refitBvhBounds.ts:80-179) on
 * device buffers: recomputes, in place, every box of the 32-byte preorder
 * BVH in d_nodes (node_count nodes, root at node 0) from its triangles
 * (index_stride words per triangle, 3 or 4; position_stride floats per
 * vertex, 3 or 4; the fourth lane is ignored), keeping the topology.  Boxes
 * equal the reference's bit for bit.  The preflight of refitBvhBounds.ts:3-63
 * runs first on the device: a corrupt or unreachable node, a bad triangle
 * range or vertex index, or a non-finite vertex returns VT_ERR_BUILD without
 * writing the nodes; a stride other than 3 or 4 is VT_ERR_RANGE.  Ordered on
 * the engine's stream (the refit itself completes asynchronously). */
vt_status vt_refit_bvh_bounds_device(vt_engine *eng, uint32_t *d_nodes, uint32_t node_count,
                                     const uint32_t *d_indices, uint32_t index_words, uint32_t index_stride,
                                     const float *d_positions, uint32_t position_floats,
                                     uint32_t position_stride);

/* Moves the vertices of a bound scene's primitive (scene order index) with
 * the topology kept, as the reference's refit path for moved geometry
 * (walkaround-hybrid HybridEnginePrimitiveUpdates.ts:1482-1560): positions
 * (3 floats per vertex, the primitive's vertex count) and optionally normals
 * replace its vertices, its BLAS is refitted on the device
 * (refitBvhBounds), the TLAS over its instances' new world boxes
 * (refitTlasTransforms, scenePack.ts:1310-1420; refitTlasInstances,
 * tlas.ts:1141-1209), the ray bias follows the new scene bounds, and
 * accumulation restarts.  Errors leave the scene unchanged: VT_ERR_RANGE for
 * a primitive without packed geometry, another vertex count or non-finite
 * data; VT_ERR_STATE without a scene or for an emissive primitive (its
 * mesh-area emitters would move: set the scene again). */
vt_status vt_refit_primitive(vt_engine *eng, uint32_t primitive, const float *positions,
                             uint32_t position_floats, const float *normals, uint32_t normal_floats);


/* ---- neural radiance cache ------------------------------------------------
 * The online-trained cache of walkaround-hybrid's NRC subsystem
 * (packages/walkaround-hybrid/src/neural/nrc/nrcSubsystem.ts,
 * fusedMlpTrainer.ts, wgsl/fusedMlp.wgsl.ts): a 64-wide MLP with 4 hidden
 * ReLU layers and RGB output, bf16 weights and activations, FP32 accumulation,
 * FP32 master weights trained by Adam on the relative-L2 loss.
 * Samples are 16-float feature records: position normalised to [0,1]^3,
 * direction and normal as (theta/pi, phi/(2 pi)), roughness, diffuse albedo
 * (3), specular albedo (3), 2 zero floats; targets are RGB. */
typedef struct vt_nrc_options {
  float learning_rate;       /* 1e-2 (MLP weights) */
  float beta1;               /* 0.9 */
  float beta2;               /* 0.999 */
  float epsilon;             /* 1e-8 */
  uint32_t seed;             /* initial weights */
  float table_learning_rate; /* hash-grid tables; 0 = 1e-1 (nrcSubsystem.ts tableLearningRate) */
  uint32_t reserved[2];
} vt_nrc_options;
/* Number of FP32 master parameters: W0..W3 (64x64) then W4 (3x64), row-major
 * [out][in], then the hash-grid tables (8 levels x 4096 rows x 2 features,
 * row-major [level][row][feature]) -- 82,112 in all. */
vt_status vt_nrc_param_count(uint32_t *count);
/* Re-initialises weights (from opts->seed) and Adam state; NULL keeps options. */
vt_status vt_nrc_reset(vt_engine *eng, const vt_nrc_options *opts);
vt_status vt_nrc_set_weights(vt_engine *eng, const float *weights, uint32_t count);
vt_status vt_nrc_get_weights(vt_engine *eng, float *weights, uint32_t count);
/* RGB prediction (3 floats per sample). Host buffers / device pointers. */
vt_status vt_nrc_infer(vt_engine *eng, const float *features, uint32_t n, float *rgb);
vt_status vt_nrc_infer_device(vt_engine *eng, const float *d_features, uint32_t n, float *d_rgb);
/* One Adam step on the batch (fusedMlpTrainer.ts train step); loss may be NULL. */
vt_status vt_nrc_train_step(vt_engine *eng, const float *features, const float *targets, uint32_t n,
                            float *loss);
vt_status vt_nrc_train_step_device(vt_engine *eng, const float *d_features, const float *d_targets,
                                   uint32_t n);
vt_status vt_nrc_last_loss(vt_engine *eng, float *loss);
/* Milliseconds of the last inference kernel and of the last training
 * (forward + backward + weight-gradient) kernel, from HIP events recorded on
 * the engine stream around them; 0 when not run yet. */
vt_status vt_nrc_timing(vt_engine *eng, float *infer_ms, float *train_ms);
/* d loss / d weights of the batch (no update), for gradient checks. */
vt_status vt_nrc_gradient(vt_engine *eng, const float *features, const float *targets, uint32_t n,
                          float *grad, float *loss);

/* Path termination by the cache in vt_render_frame (nrcSubsystem.ts): with
 * `enabled`, each image path ends at the vertex after its first
 * diffuse-dominant vertex (specular lobe probability < 1/2) and adds
 * throughput x cache prediction there; `train_paths` extra self-training
 * paths per frame (pseudo-random pixels) record up to 4 vertices each and
 * drive one Adam step per frame.  Resets accumulation. */
vt_status vt_nrc_render_config(vt_engine *eng, int32_t enabled, uint32_t train_paths);

/* Length of the self-training paths' teacher suffix (NRC_TEACHER_MAX_VERTICES
 * = 4, walkaround-hybrid shaders/nrcIndependentSuffix.wgsl.ts:23): a training
 * path shades `vertices` vertices (all of them recorded for training) and
 * ends at the next one with a cache query whose prediction x throughput is
 * added to its radiance before the targets are formed (self-training
 * termination, Mueller et al. 2021, sec. 5); the reference truncates there
 * without an estimate of the remainder.  0 = training paths run to the
 * bounce limit with no terminal query.  Range [0, 4], default 4.  Resets
 * accumulation; VT_ERR_STATE without a scene. */
vt_status vt_nrc_teacher_vertices(vt_engine *eng, uint32_t vertices);

/* ---- batches: many scenes in one persistent launch --------------------------
 * No single counterpart in the reference (it renders one scene per engine);
 * a batch carries several scenes' traversal buffers so that the closest-hit
 * rays of all of them run in one load-balanced persistent launch, which can
 * be captured once into a HIP graph and replayed.  ray_offsets (host, scene
 * count + 1 entries, ray_offsets[0] = 0) gives each scene's ray range in the
 * ray array; hit records are per scene (that scene's primitive/instance
 * indices), identical to vt_trace_closest on that scene alone. */
typedef struct vt_batch vt_batch;
typedef struct vt_tuning {
  uint32_t refill_min;    /* wave refill threshold */
  uint32_t steps;         /* traversal steps between refills */
  uint32_t near_first;    /* 0 left-first, 1 nearest-first inside BLASes */
  uint32_t defer_min;     /* lanes that must need a leaf/instance operation */
  uint32_t blocks_per_cu; /* resident 128-thread blocks per CU */
  float ms;               /* measured median launch time (autotuner output) */
} vt_tuning;
vt_status vt_batch_create(vt_engine *eng, vt_batch **out);
vt_status vt_batch_destroy(vt_batch *b);
vt_status vt_batch_add_scene(vt_batch *b, const vt_scene_desc *scene, uint32_t *index);
vt_status vt_batch_scene_count(const vt_batch *b, uint32_t *count);
/* One batched closest-hit launch (device pointers); *ms, if given, receives
 * the event-timed launch duration (and the call then waits for it). */
vt_status vt_batch_trace_closest_device(vt_batch *b, const vt_ray *d_rays, const uint32_t *ray_offsets,
                                        vt_hit *d_hits, float *ms);
/* Captures the launch into a HIP graph; vt_batch_replay re-runs it. */
vt_status vt_batch_capture(vt_batch *b, const vt_ray *d_rays, const uint32_t *ray_offsets, vt_hit *d_hits);
vt_status vt_batch_replay(vt_batch *b, float *ms);
/* Measures candidate launch parameters on these rays, keeps the fastest. */
vt_status vt_batch_autotune(vt_batch *b, const vt_ray *d_rays, const uint32_t *ray_offsets, vt_hit *d_hits,
                            uint32_t repeats, vt_tuning *chosen);
vt_status vt_batch_set_tuning(vt_batch *b, const vt_tuning *t);

/* invertMat4 (packages/shared-bvh/src/mathUtils.ts:21-110): binary64
 * cofactor inverse of a column-major mat4 rounded to binary32, with the
 * reciprocal-product check.  VT_ERR_RANGE when the matrix is not invertible.
 * The frame driver uses it for inverse(projMatrix * viewMatrix)
 * (packages/pt-webgpu/src/frameParamsPacker.ts:117-118). */
vt_status vt_invert_mat4(const float m[16], float out[16]);

/* Library version string, e.g. "vitrum-hip 0.1.0 gfx950". */
const char *vt_version(void);

#ifdef __cplusplus
}
#endif
#endif
