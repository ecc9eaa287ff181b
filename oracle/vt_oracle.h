/*
 * vt_oracle.h -- CPU restatement of vitrum's path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links,
 * loads or calls it.
 *
 * Every function restates the reference algorithm it cites (paths relative to
 * the vitrum repository root).  Arithmetic follows the reference's own CPU
 * semantics: builder decisions in IEEE binary64 (JavaScript numbers), the
 * traversal in IEEE binary32 with one rounding per operation and no fused
 * multiply-add (the `Math.fround` discipline of
 * packages/shared-bvh/src/opticalWatertightTriangle.ts).
 *
 * All entry points return 0 on success and a negative status on failure; the
 * failure message is available from orc_last_error().
 */
#ifndef VT_ORACLE_H
#define VT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *orc_last_error(void);

/* ---- BLAS build: packages/shared-bvh/src/buildArrayBvh.ts:342 ----------- */
/* out_node_words capacity: max(1, 2*tri_count-1) * 8 words.                  */
int orc_build_blas(const float *positions, uint32_t position_floats,
                   uint32_t position_stride, const uint32_t *indices,
                   uint32_t index_words, uint32_t index_stride,
                   const uint32_t *tri_material_ids, uint32_t max_leaf,
                   uint32_t num_bins, uint32_t max_depth,
                   uint32_t *out_node_words, uint32_t *out_node_count,
                   uint32_t *out_reordered_indices, uint32_t *out_reordered_mat,
                   uint32_t *out_reordered_to_source);

/* ---- validateBvhEncoding: packages/shared-bvh/src/validateBvhEncoding.ts:55
 * triangle_count < 0 means "not supplied".  out_proof (7 words): nodeCount,
 * rootCount, interiorNodeCount, leafNodeCount, maxDepth,
 * maxTraversalStackEntries, traversalStackCapacity.                          */
int orc_validate_bvh(const uint32_t *node_words, uint32_t node_count,
                     const uint32_t *roots, int32_t root_count,
                     int64_t triangle_count, uint32_t stack_capacity,
                     uint32_t *out_proof);

/* ---- TLAS build: packages/shared-bvh/src/tlas.ts:859 --------------------
 * aabb_min/aabb_max are binary64 (the reference passes JS numbers).
 * out_nodes capacity: (2n-1)*8 words.  out_fallbacks (4): degenerate-centroids,
 * non-improving-sah, degenerate-partition, depth-safety.                     */
int orc_build_tlas(uint32_t n, const uint32_t *blas_ids, const double *aabb_min,
                   const double *aabb_max, const float *world_to_local,
                   uint32_t max_leaf, uint32_t num_bins, uint32_t max_depth,
                   uint32_t *out_nodes, uint32_t *out_node_count,
                   uint32_t *out_instance_indices, uint32_t *out_blas_roots,
                   float *out_transforms, uint32_t *out_fallbacks);

/* ---- invertMat4: packages/shared-bvh/src/mathUtils.ts:21 (1 = invertible) */
int orc_invert_mat4(const float *m, float *out);

/* ---- refitBvhBounds: packages/shared-bvh/src/refitBvhBounds.ts:3-179 -----
 * In place over the tree rooted at node 0 of `node_words` (pass a view at a
 * BLAS root to refit one tree of a forest; leaf triangle offsets index
 * `indices`).  -2: a stride other than 3 or 4; -1: rejected topology or
 * geometry (nothing is written).                                            */
int orc_refit_bvh_bounds(uint32_t *node_words, uint32_t node_count, const uint32_t *indices,
                         uint32_t index_words, uint32_t index_stride, const float *positions,
                         uint32_t position_floats, uint32_t position_stride);

/* ---- refitTlasInstances: packages/shared-bvh/src/tlas.ts:1012-1209 -------
 * aabb_min / aabb_max: binary64, 3 per original instance.  out_affected
 * (capacity node_count) receives the refitted nodes, children first.        */
int orc_refit_tlas_instances(uint32_t *node_words, uint32_t node_count, const uint32_t *perm,
                             uint32_t perm_count, uint32_t instance_count, const double *aabb_min,
                             const double *aabb_max, const uint32_t *changed, uint32_t changed_count,
                             uint32_t *out_affected, uint32_t *out_affected_count);

/* ---- packSceneFromCore: packages/shared-bvh/src/scenePack.ts:1104 -------
 * A primitive is a mesh (kind 0, transform_count 0 or 1; NULL transforms =
 * identity) or an instanced mesh (kind 1, transform_count instances).       */
typedef struct orc_primitive {
  uint32_t kind;
  uint32_t material_id;
  const float *positions;   /* xyz triples */
  uint32_t position_floats;
  const float *normals;     /* xyz triples */
  uint32_t normal_floats;
  const uint32_t *indices;  /* NULL = generated 0..vertexCount-1 */
  uint32_t index_count;
  const float *transforms;  /* column-major mat4, 16 floats each */
  uint32_t transform_count;
} orc_primitive;

typedef struct orc_pack orc_pack;

/* Field ids for orc_pack_field_size / orc_pack_copy_field. */
enum {
  ORC_POSITIONS = 0,          /* f32 vec4 per vertex */
  ORC_NORMALS = 1,            /* f32 vec4 per vertex */
  ORC_INDICES = 2,            /* u32 vec4 per triangle (rebased) */
  ORC_TRI_MATERIAL = 3,       /* u32 */
  ORC_TRI_SOURCE = 4,         /* u32 */
  ORC_TRI_PRIMITIVE = 5,      /* u32 */
  ORC_BVH_NODES = 6,          /* u32 x8 per node */
  ORC_TLAS_NODES = 7,         /* u32 x8 per node */
  ORC_TLAS_INSTANCE_INDICES = 8,
  ORC_TLAS_BLAS_ROOTS = 9,
  ORC_TLAS_W2L = 10,          /* f32 x16 per instance */
  ORC_TLAS_L2W = 11,          /* f32 x16 per instance */
  ORC_INSTANCE_PRIMITIVE = 12,
  ORC_INSTANCE_SOURCE = 13,
  ORC_BINDINGS = 14,          /* u32 x8 per binding: blasRoot, instanceCount,
                                 vertexStart, vertexCount, triStart, triCount,
                                 primitiveIndex, kind */
  ORC_BINDING_AABB = 15,      /* f32 x6 per binding: local min xyz, max xyz */
  ORC_FIELD_COUNT = 16
};

int orc_pack_scene(const orc_primitive *prims, uint32_t prim_count,
                   orc_pack **out);
/* Size of a field in 32-bit words. */
int64_t orc_pack_field_size(const orc_pack *p, int field);
int orc_pack_copy_field(const orc_pack *p, int field, void *dst,
                        uint64_t dst_words);
uint32_t orc_pack_warning_count(const orc_pack *p);
void orc_pack_free(orc_pack *p);

/* ---- Traversal ---------------------------------------------------------- */
typedef struct orc_ray {
  float o[3];
  float tmin;
  float d[3];
  float tmax;
} orc_ray;

typedef struct orc_hit {
  uint32_t prim; /* global packed (BVH-reordered) triangle index */
  uint32_t inst; /* packed instance index; 0xFFFFFFFF = miss */
  float t;
  float u; /* barycentric weight of vertex b */
  float v; /* barycentric weight of vertex c */
} orc_hit;

/* Views of packed scene buffers (the layout packSceneFromCore returns). */
typedef struct orc_scene {
  const float *positions;      /* vec4 */
  uint32_t vertex_count;
  const uint32_t *indices;     /* vec4u */
  uint32_t triangle_count;
  const uint32_t *bvh_nodes;   /* 8 words per node */
  uint32_t bvh_node_count;
  const uint32_t *tlas_nodes;
  uint32_t tlas_node_count;
  const uint32_t *instance_indices;
  uint32_t instance_index_count;
  const uint32_t *blas_roots;
  const float *w2l;
  const float *l2w;
  uint32_t instance_count;
} orc_scene;

/* Watertight triangle test, f32 mirror of
 * packages/shared-bvh/src/opticalWatertightTriangle.ts:98.
 * out (10 floats): hit, t, point xyz, bary xyz, side, zeroEdgeMask.          */
void orc_watertight(const float *origin, const float *direction,
                    const float *a, const float *b, const float *c,
                    float exclusive_min_t, float *out);

/* Closest hit: traceTlasClosest (packages/shared-bvh/src/wgsl/
 * tlasSceneHitTraversal.wgsl.ts:185) as patched by pt-webgpu
 * (packages/pt-webgpu/src/wgsl/pathTrace/intersection.wgsl.ts:31-152).
 * counts (optional, 4 u32 per ray): BLAS nodes slab-tested, triangles tested,
 * TLAS nodes slab-tested, instances entered.                                */
int orc_trace_closest(const orc_scene *s, const orc_ray *rays, uint32_t n,
                      orc_hit *out, uint32_t *counts, uint32_t threads);
/* Closest hit plus the world hit position (3 f32 per ray, the traversal's
 * L2W * local hit point, tlasSceneHitTraversal.wgsl.ts:145-152). */
int orc_trace_closest_pos(const orc_scene *s, const orc_ray *rays, uint32_t n,
                          orc_hit *out, float *pos, uint32_t threads);
/* Any hit: traceTlasAny (tlasSceneHitTraversal.wgsl.ts:227). */
int orc_trace_any(const orc_scene *s, const orc_ray *rays, uint32_t n,
                  uint32_t *out, uint32_t threads);
/* The path tracer's wrappers over the closest hit (a8): traceClosest
 * (intersection.wgsl.ts:462-490) skips hits on the back face of a
 * single-sided opaque material by re-tracing from the hit distance;
 * traceSurfaceVisibility (connect.wgsl.ts:37-103) reports a ray occluded by
 * the first accepted surface, after at most hit_limit surfaces.
 * single_sided: one byte per packed triangle (nonzero: back faces are not
 * surfaces), or NULL (every hit accepted).  orc_orientation_sign: sign of an
 * instance's local-to-world 3x3 determinant (transformLinearOrientationSign,
 * intersectionCore.wgsl.ts:96-112); negative flips the local front face. */
int orc_trace_closest_sided(const orc_scene *s, const uint8_t *single_sided, const orc_ray *rays, uint32_t n,
                            orc_hit *out, float *pos, uint32_t threads);
int orc_trace_occluded_sided(const orc_scene *s, const uint8_t *single_sided, uint32_t hit_limit,
                             const orc_ray *rays, uint32_t n, uint32_t *occluded, uint32_t threads);
int orc_orientation_sign(const float *l2w);

/* FP64 ray/triangle intersection of the hit triangle in world space: the
 * instance's local vertices are carried to world space through its
 * local-to-world matrix in binary64, then a closed-form binary64
 * Moller-Trumbore solve gives t and the (u, v) weights of vertices b, c.
 * out: 3 doubles per ray (NaN when the ray misses the supporting plane).    */
int orc_fp64_hit_params(const orc_scene *s, const orc_ray *rays,
                        const orc_hit *hits, uint32_t n, double *out);

/* ---- The reference's CPU closest hit: 8-wide compressed BVH -------------
 * buildCompressedWideBvhFromArrayBvh (packages/shared-bvh/src/
 * compressedWideBvh.ts:248-418) over the binary trees at `roots` of a node
 * array, concatenated into one forest (roots[] -> wide roots), and
 * intersectCompressedWideBvhFirstHit (:754-860, binary64 Moller-Trumbore)
 * composed with tlasIntersect (packages/shared-bvh/src/tlas.ts:1223).
 * status per root: 0 wide, 1 empty, 2 binary-fallback.                     */
typedef struct orc_cwbvh orc_cwbvh;
typedef struct orc_hit64 {
  uint32_t hit;
  uint32_t prim;  /* packed (BVH-reordered) triangle index */
  uint32_t inst;  /* packed instance index */
  uint32_t pad;
  double t;
  double bary[3]; /* weights of vertices a, b, c */
} orc_hit64;
int orc_cwbvh_build(const uint32_t *bvh_words, uint32_t node_count,
                    const uint32_t *roots, uint32_t root_count,
                    uint32_t tri_count, orc_cwbvh **out);
void orc_cwbvh_free(orc_cwbvh *c);
uint32_t orc_cwbvh_node_count(const orc_cwbvh *c);
/* Any output pointer may be NULL.  Sizes: node_bounds 6/node, child_bounds
 * 48/node (u16), child_meta 24/node, child_count 1/node; roots, status and
 * stack_entries one per binary root. */
int orc_cwbvh_copy(const orc_cwbvh *c, float *node_bounds,
                   uint16_t *child_bounds, uint32_t *child_meta,
                   uint32_t *child_count, uint32_t *roots, uint32_t *status,
                   double *stack_entries);
/* requiredCwbvhTraversalStackEntries (:60-92) of a given topology. */
double orc_cwbvh_stack_entries(const uint32_t *child_meta,
                               const uint32_t *child_count,
                               uint32_t node_count, uint32_t root);
/* pt-webgpu's opt-in `cwbvh-closest` traversal in binary32
 * (pathTrace/intersection.wgsl.ts:576-838 over cwbvhIntersect.wgsl.ts:197-336):
 * Moller-Trumbore proposal in the wide forest, watertight re-verification,
 * binary fallback per BLAS.  The view points at an orc_cwbvh's arrays. */
typedef struct orc_cwbvh_view {
  const float *node_bounds;
  const uint16_t *child_bounds;
  const uint32_t *child_meta;
  const uint32_t *child_count;
  uint32_t node_count;
  const uint32_t *inst_root; /* wide root per packed instance */
} orc_cwbvh_view;
int orc_trace_closest_cwbvh(const orc_scene *s, const orc_cwbvh_view *cw,
                            const orc_ray *rays, uint32_t n, orc_hit *out,
                            uint32_t threads);
/* The same with traceClosest's sided loop (single_sided as for
 * orc_trace_closest_sided; NULL accepts every hit). */
int orc_trace_closest_cwbvh_sided(const orc_scene *s, const orc_cwbvh_view *cw, const uint8_t *single_sided,
                                  const orc_ray *rays, uint32_t n, orc_hit *out, uint32_t threads);
/* blas_wide_root: wide root of each packed instance's BLAS. */
int orc_cwbvh_trace_closest(const orc_scene *s, const orc_cwbvh *c,
                            const uint32_t *blas_wide_root,
                            const orc_ray *rays, uint32_t n, orc_hit64 *out,
                            uint32_t threads);

/* ---- Light tree: packages/shared-samplers/src/lightTree.ts --------------
 * buildLightTree + packLightTreeForGPU (16 f32 per node; out capacity
 * (2n-1)*16), packed importance, samplePackedLightTreeCPU (bucket = the
 * 24-bit root draw) and packedLightTreePdfCPU.  has_cone may be NULL.      */
int orc_light_tree_build(uint32_t n, const double *power, const double *centroid,
                         const double *aabb_min, const double *aabb_max,
                         const uint8_t *has_cone, const double *axis,
                         const double *theta_o, const double *theta_e,
                         float *out, uint32_t *out_nodes);
int orc_light_tree_pack_nodes(uint32_t count, const int32_t *emitter,
                              const double *power, const int32_t *left,
                              const int32_t *right, const double *aabb_min,
                              const double *aabb_max, const double *axis,
                              const double *theta_o, const double *theta_e,
                              float *out);
float orc_light_tree_importance(const float *packed, uint32_t node,
                                const float *p, float floor_d2);
int32_t orc_light_tree_sample(const float *packed, uint32_t nodes,
                              const float *p, float floor_d2, uint32_t bucket,
                              float *pdf);
float orc_light_tree_pdf(const float *packed, uint32_t nodes, const float *p,
                         float floor_d2, int32_t emitter);

/* ---- PCG32: packages/shared-samplers/src/wgsl/pcg.wgsl.ts:5-22 ---------- */
uint32_t orc_pcg_init(uint32_t px, uint32_t py, uint32_t frame_seed);
uint32_t orc_pcg_next(uint32_t *state);
float orc_rand_f32(uint32_t *state);

#ifdef __cplusplus
}
#endif
#endif
