// traverse.hip -- closest-hit and any-hit TLAS->BLAS traversal for gfx950.
//
// Semantics are those of the reference's traversal, so hit records are
// bit-identical to it:
//   traceTlasClosest / traceTlasAny  packages/shared-bvh/src/wgsl/tlasSceneHitTraversal.wgsl.ts:92-269
//   pt-webgpu instance augmentation  packages/pt-webgpu/src/wgsl/pathTrace/intersection.wgsl.ts:31-152
//   traceMeshBvh                     packages/pt-webgpu/src/wgsl/pathTrace/intersectionCore.wgsl.ts:461-596
//   watertight triangle test         packages/shared-bvh/src/opticalWatertightTriangle.ts:98-220
//   safeInvDir / slab test           packages/shared-bvh/src/wgsl/bvhIntersect.wgsl.ts:86-111,
//                                    intersectionCore.wgsl.ts:24-31
// Every float operation rounds once (built with -ffp-contract=off and
// correctly rounded division and square root).
//
// Design (MI355X):
//  * Node64 BVH2 nodes: the two child boxes sit in the parent, so a node
//    expansion is one 64-byte load and two slab tests.  The traversal order
//    is the reference's left-first depth-first order; a child's entry
//    distance is kept on the stack and re-checked against the current best
//    when it is popped, which is exactly the reference's pop-time cull.
//  * Persistent waves: each wave pulls rays from a global counter.  Lanes that
//    finish are refilled together once enough of the wave is idle: a ballot
//    of idle lanes, one atomicAdd per refill, and a prefix count (mbcnt) to
//    hand out consecutive ray indices.  A launch with fewer rays than 64 per
//    resident wave hands each wave at most its even share, so small batches
//    (late bounces, training paths) occupy every wave slot of every SIMD and
//    the dependent node loads of one wave hide behind the others'.
//  * One combined TLAS+BLAS stack per lane in LDS (entry-major so a wave's
//    accesses at equal depth hit 64 distinct banks), spilling the rare deep
//    entries to a per-thread global area.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>

#include "traverse.h"
#include "cwbvh.h"

namespace vtk {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0x80000000u;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMiss = 0xFFFFFFFFu;
// Deferred per-lane operations of the traversal loop.
constexpr uint32_t kPendNone = 0, kPendLeaf = 1, kPendEnter = 2, kPendExit = 3;

__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float safe_max() { return __uint_as_float(0x7F7FFFFDu); }  // 3.402823e38
__device__ __forceinline__ float inf_const() { return __uint_as_float(0x7F7FFFFFu); }  // 3.402823466e38

__device__ __forceinline__ float fabs_(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ bool finite_(float x) { return __builtin_isfinite(x); }
__device__ __forceinline__ bool nan_(float x) { return __builtin_isnan(x); }
// Math.max: NaN-propagating (the TS mirror's semantics).
__device__ __forceinline__ float jmax(float a, float b) {
  return (nan_(a) || nan_(b)) ? __builtin_nanf("") : (a > b ? a : b);
}
__device__ __forceinline__ float max3abs(float x, float y, float z) {
  return jmax(fabs_(x), jmax(fabs_(y), fabs_(z)));
}
__device__ __forceinline__ float sel3(int k, float x, float y, float z) {
  return k == 0 ? x : (k == 1 ? y : z);
}

// Clang vector types travel in registers through the (rare) out-of-line
// calls below; plain structs would be passed through scratch memory.
typedef float V3 __attribute__((ext_vector_type(3)));
typedef float F4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

// ---------------------------------------------------------------------------
// Watertight triangle test (opticalWatertightTriangle.ts:15-220)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float edge_fn(float ax, float ay, float bx, float by) {
  if (ax == bx && ay == by) return 0.0f;
  const bool ordered = ax < bx || (ax == bx && ay < by);
  const float fx = ordered ? ax : bx, fy = ordered ? ay : by;
  const float sx = ordered ? bx : ax, sy = ordered ? by : ay;
  const float canonical = fx * sy - fy * sx;
  return ordered ? canonical : -canonical;
}

__device__ __noinline__ float ray_param(V3 o, V3 dir, V3 p) {
  const V3 delta = sub(p, o);
  const float ds = max3abs(delta.x, delta.y, delta.z);
  const float rs = max3abs(dir.x, dir.y, dir.z);
  if (!(rs > 0.0f) || !finite_(rs) || !finite_(ds)) return -1.0f;
  if (!(ds > 0.0f)) return 0.0f;
  const V3 d = v3(dir.x / rs, dir.y / rs, dir.z / rs);
  const V3 q = v3(delta.x / ds, delta.y / ds, delta.z / ds);
  const float num = (q.x * d.x + q.y * d.y) + q.z * d.z;
  const float den = (d.x * d.x + d.y * d.y) + d.z * d.z;
  return (num / den) * (ds / rs);
}

__device__ __forceinline__ bool lex_before(V3 a, V3 b) {
  return a.x < b.x || (a.x == b.x && (a.y < b.y || (a.y == b.y && a.z < b.z)));
}

// Returns (t, point) of an edge hit in registers.
__device__ __noinline__ F4 edge_ray_param(V3 o, V3 dir, V3 fi, V3 si, float fpx, float fpy, float spx,
                                          float spy) {
  const bool ordered = lex_before(fi, si);
  const V3 f = ordered ? fi : si, s = ordered ? si : fi;
  const float ax = ordered ? fpx : spx, ay = ordered ? fpy : spy;
  const float bx = ordered ? spx : fpx, by = ordered ? spy : fpy;
  const float ddx = bx - ax, ddy = by - ay;
  float interp;
  if (fabs_(ddx) >= fabs_(ddy)) {
    if (ddx == 0.0f) return F4{-1.0f, 0.0f, 0.0f, 0.0f};
    interp = -ax / ddx;
  } else {
    if (ddy == 0.0f) return F4{-1.0f, 0.0f, 0.0f, 0.0f};
    interp = -ay / ddy;
  }
  const V3 point = v3(f.x + interp * (s.x - f.x), f.y + interp * (s.y - f.y), f.z + interp * (s.z - f.z));
  return F4{ray_param(o, dir, point), point.x, point.y, point.z};
}

struct TriHit {
  float t;
  float u, v;  // bary.y, bary.z
  V3 point;
  bool front;  // side > 0: the ray meets the authored front face
};

// Ray-only part of the watertight test (opticalWatertightTriangle.ts:106-124):
// the dominant axis permutation and shear, computed once per ray per instance.
struct WtRay {
  float ds;          // max |dir| component
  float sx, sy, sz;  // shear constants
  uint32_t perm;     // kx | ky << 2 | kz << 4; 0xFFFFFFFF = invalid direction
};

__device__ __forceinline__ WtRay wt_setup(V3 dir) {
  WtRay w;
  w.ds = max3abs(dir.x, dir.y, dir.z);
  if (!(w.ds > 0.0f) || !finite_(w.ds)) {
    w.perm = 0xFFFFFFFFu;
    w.sx = w.sy = w.sz = 0.0f;
    return w;
  }
  const V3 d = v3(dir.x / w.ds, dir.y / w.ds, dir.z / w.ds);
  int kz = fabs_(d.y) > fabs_(d.x) ? 1 : 0;
  if (fabs_(d.z) > fabs_(sel3(kz, d.x, d.y, d.z))) kz = 2;
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  const float dkz = sel3(kz, d.x, d.y, d.z);
  if (dkz < 0.0f) {
    const int t = kx;
    kx = ky;
    ky = t;
  }
  w.sx = sel3(kx, d.x, d.y, d.z) / dkz;
  w.sy = sel3(ky, d.x, d.y, d.z) / dkz;
  w.sz = 1.0f / dkz;
  w.perm = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
  return w;
}

// Rare tail of an in-window hit (opticalWatertightTriangle.ts:164-220): the
// authored-normal side test.  Returns 0 for a rejected hit, 1 for the front
// face (side > 0), -1 for the back face.
__device__ __noinline__ int wt_side(V3 dir, float ds, V3 a, V3 b, V3 c) {
  const V3 e1r = sub(b, a), e2r = sub(c, a);
  const float es = jmax(max3abs(e1r.x, e1r.y, e1r.z), max3abs(e2r.x, e2r.y, e2r.z));
  if (!(es > 0.0f) || !finite_(es)) return 0;
  const V3 s1 = v3(e1r.x / es, e1r.y / es, e1r.z / es);
  const V3 s2 = v3(e2r.x / es, e2r.y / es, e2r.z / es);
  const V3 nr = v3(s1.y * s2.z - s1.z * s2.y, s1.z * s2.x - s1.x * s2.z, s1.x * s2.y - s1.y * s2.x);
  const float ns = max3abs(nr.x, nr.y, nr.z);
  if (!(ns > 0.0f) || !finite_(ns)) return 0;
  const V3 nd = v3(nr.x / ns, nr.y / ns, nr.z / ns);
  const float nl = __builtin_sqrtf((nd.x * nd.x + nd.y * nd.y) + nd.z * nd.z);
  if (!(nl > 0.0f) || !finite_(nl)) return 0;
  const V3 n = v3(nd.x / nl, nd.y / nl, nd.z / nl);
  const V3 d = v3(dir.x / ds, dir.y / ds, dir.z / ds);
  const float side = -((d.x * n.x + d.y * n.y) + d.z * n.z);
  if (side == 0.0f || !finite_(side)) return 0;
  return side > 0.0f ? 1 : -1;
}

// Returns true and fills h when the triangle is hit with lo < t < hi (or
// t <= hi when `incl_hi`), where lo = tmin (also the predicate's exclusive
// minimum) and hi the current bound.  Work that can only turn an
// out-of-window hit into a miss is skipped for out-of-window hits, which
// leaves the accepted set unchanged.
__device__ __forceinline__ bool tri_test(V3 o, V3 dir, const WtRay& w, V3 a, V3 b, V3 c, float tmin,
                                         float hi, bool incl_hi, TriHit& h) {
  if (w.perm == 0xFFFFFFFFu || nan_(tmin)) return false;
  const int kx = w.perm & 3, ky = (w.perm >> 2) & 3, kz = w.perm >> 4;
  const V3 pa = sub(a, o), pb = sub(b, o), pc = sub(c, o);
  const float paz = sel3(kz, pa.x, pa.y, pa.z), pbz = sel3(kz, pb.x, pb.y, pb.z),
              pcz = sel3(kz, pc.x, pc.y, pc.z);
  const float apx = sel3(kx, pa.x, pa.y, pa.z) - w.sx * paz, apy = sel3(ky, pa.x, pa.y, pa.z) - w.sy * paz;
  const float bpx = sel3(kx, pb.x, pb.y, pb.z) - w.sx * pbz, bpy = sel3(ky, pb.x, pb.y, pb.z) - w.sy * pbz;
  const float cpx = sel3(kx, pc.x, pc.y, pc.z) - w.sx * pcz, cpy = sel3(ky, pc.x, pc.y, pc.z) - w.sy * pcz;
  const float e0 = edge_fn(bpx, bpy, cpx, cpy);
  const float e1 = edge_fn(cpx, cpy, apx, apy);
  const float e2 = edge_fn(apx, apy, bpx, bpy);
  const float det = (e0 + e1) + e2;
  if (det == 0.0f || !finite_(det)) return false;
  const bool positive = det > 0.0f;
  const float n0 = positive ? e0 : -e0, n1 = positive ? e1 : -e1, n2 = positive ? e2 : -e2;
  if (n0 < 0.0f || n1 < 0.0f || n2 < 0.0f) return false;
  const float scaled_t = (e0 * (paz * w.sz) + e1 * (pbz * w.sz)) + e2 * (pcz * w.sz);
  float t = (scaled_t / det) / w.ds;
  V3 point = v3(0, 0, 0);
  const int zc = (n0 == 0.0f) + (n1 == 0.0f) + (n2 == 0.0f);
  if (zc != 0) {
    if (zc >= 2) {
      if (n0 != 0.0f) point = a, t = ray_param(o, dir, point);
      if (n1 != 0.0f) point = b, t = ray_param(o, dir, point);
      if (n2 != 0.0f) point = c, t = ray_param(o, dir, point);
    } else {
      F4 ep;
      if (n0 == 0.0f)
        ep = edge_ray_param(o, dir, b, c, bpx, bpy, cpx, cpy);
      else if (n1 == 0.0f)
        ep = edge_ray_param(o, dir, c, a, cpx, cpy, apx, apy);
      else
        ep = edge_ray_param(o, dir, a, b, apx, apy, bpx, bpy);
      t = ep.x;
      point = v3(ep.y, ep.z, ep.w);
    }
  }
  const float lo = tmin > 0.0f ? tmin : 0.0f;
  if (!(t > lo) || !finite_(t)) return false;
  if (!(t > tmin && (t < hi || (incl_hi && t == hi)))) return false;  // traversal acceptance window
  // opticalWatertightTriangle.ts:164-219: authored-normal side test, then
  // barycentrics and the canonical hit point.
  const int side = wt_side(dir, w.ds, a, b, c);
  if (side == 0) return false;
  h.front = side > 0;
  h.t = t;
  h.u = e1 / det;
  h.v = e2 / det;
  if (zc == 0) {
    point = v3(a.x + (h.u * (b.x - a.x) + h.v * (c.x - a.x)), a.y + (h.u * (b.y - a.y) + h.v * (c.y - a.y)),
               a.z + (h.u * (b.z - a.z) + h.v * (c.z - a.z)));
  }
  h.point = point;
  return true;
}

// ---------------------------------------------------------------------------
// Slab test, transforms (bvhIntersect.wgsl.ts:86-111; intersectionCore.wgsl.ts:24-83)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float safe_inv(float v) {
  const float m = safe_max();
  if (v == 0.0f) return v >= 0.0f ? m : -m;
  const float r = 1.0f / v;
  return r < -m ? -m : (r > m ? m : r);
}

// Computes tNear and returns the bound-independent part of the test:
// !(tNear > tFar || tFar < tMin).
__device__ __forceinline__ bool slab(V3 o, V3 inv, float mnx, float mny, float mnz, float mxx,
                                     float mxy, float mxz, float tmin, float& tnear) {
  const float t1x = (mnx - o.x) * inv.x, t1y = (mny - o.y) * inv.y, t1z = (mnz - o.z) * inv.z;
  const float t2x = (mxx - o.x) * inv.x, t2y = (mxy - o.y) * inv.y, t2z = (mxz - o.z) * inv.z;
  const float tn = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
  const float tf = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
  tnear = tn;
  return !(tn > tf || tf < tmin);
}

__device__ __forceinline__ V3 xform_point(float4 c0, float4 c1, float4 c2, float4 c3, V3 p) {
  const float rx = ((c0.x * p.x + c1.x * p.y) + c2.x * p.z) + c3.x;
  const float ry = ((c0.y * p.x + c1.y * p.y) + c2.y * p.z) + c3.y;
  const float rz = ((c0.z * p.x + c1.z * p.y) + c2.z * p.z) + c3.z;
  const float rw = ((c0.w * p.x + c1.w * p.y) + c2.w * p.z) + c3.w;
  const float s = fmaxf(fmaxf(fabs_(rx), fabs_(ry)), fmaxf(fabs_(rz), fabs_(rw)));
  if (!(s > 0.0f) || s > safe_max()) return v3(0, 0, 0);
  const float qx = rx / s, qy = ry / s, qz = rz / s, qw = rw / s;
  if (qw == 0.0f) return v3(0, 0, 0);
  const V3 q = v3(qx / qw, qy / qw, qz / qw);
  if (nan_(q.x) || nan_(q.y) || nan_(q.z) || fabs_(q.x) > safe_max() || fabs_(q.y) > safe_max() ||
      fabs_(q.z) > safe_max())
    return v3(0, 0, 0);
  return q;
}

__device__ __forceinline__ V3 safe_normalize(V3 v) {
  const float s = fmaxf(fabs_(v.x), fmaxf(fabs_(v.y), fabs_(v.z)));
  if (!(s > 0.0f) || s > safe_max()) return v3(0, 1, 0);
  const V3 q = v3(v.x / s, v.y / s, v.z / s);
  const float len = __builtin_sqrtf((q.x * q.x + q.y * q.y) + q.z * q.z);
  return v3(q.x / len, q.y / len, q.z / len);
}

__device__ __forceinline__ V3 xform_dir(float4 c0, float4 c1, float4 c2, V3 d) {
  return safe_normalize(v3((c0.x * d.x + c1.x * d.y) + c2.x * d.z, (c0.y * d.x + c1.y * d.y) + c2.y * d.z,
                           (c0.z * d.x + c1.z * d.y) + c2.z * d.z));
}

// ---------------------------------------------------------------------------
// Persistent traversal kernel
// ---------------------------------------------------------------------------

// Single-step persistent loop: each iteration a lane pops (with the
// reference's pop-time cull), expands one Node64 (two child boxes), tests
// one leaf's triangles or enters / leaves one instance.
//
// Child order.  The reference visits the left child first.  With
// near_first = 1 (closest hit only) the child with the smaller entry distance
// is visited first inside each BLAS, and equal triangle distances are
// resolved as the reference's order resolves them (the lower reordered index
// wins); instances are still visited in the reference's TLAS order, because
// with coplanar geometry in different instances the reference's choice
// depends on that order through the rounding of the instance transforms.
// near_first = 2 also orders the TLAS by distance.  The reference's choice
// between instances depends on its order only through near-ties: an
// instance's local search is bounded by the best world distance found so far,
// mapped into the instance's space, and where two instances' hits lie within
// the rounding of that mapping the earlier-visited one can exclude the other.
// So in this mode every world bound is widened by a margin (2^-13 of the
// distance plus the origin's magnitude, far above the transforms' rounding),
// which keeps every hit near the current best visible, and a lane that sees
// two instances' hits within that margin of each other walks its ray again in
// the reference's TLAS order.  Without such a near-tie the minimum world
// distance is what any order finds, so the hit records equal the reference's.
// Node culling stays `tNear > bound` (equal distances are still visited).  The
// counting variant always uses the reference order, so its work counts are
// the reference's.
template <bool ANY, bool COUNT, bool MULTI, bool W4, uint32_t BLOCK, uint32_t LDS_DEPTH>
// Register budget: four waves per SIMD (128 VGPRs; with the 20-entry LDS
// stack, 8 blocks of 128 fit a CU), which beats three waves without spills.
__global__ __launch_bounds__(BLOCK, 4) void trace_kernel(TraceParams p) {
  // Per-lane stack: LDS_DEPTH entries in LDS (entry-major, so a wave's
  // accesses at equal depth touch 64 distinct banks), deeper entries in a
  // per-thread global spill area.  The shared array is indexed directly so the
  // compiler emits ds_* instructions, and the spill area is accessed as 64-bit
  // words, a different type, so the two branches cannot be merged into one
  // generic (flat) access with the latency of global memory.
  __shared__ uint2 lds_stack[LDS_DEPTH * BLOCK];
  const uint32_t tid = threadIdx.x;
  const size_t spill_stride = (size_t)gridDim.x * BLOCK, gtid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  unsigned long long* __restrict__ spill = reinterpret_cast<unsigned long long*>(p.spill);
  auto stk_store = [&](uint32_t i, uint2 v) {
    if (i < LDS_DEPTH)
      lds_stack[i * BLOCK + tid] = v;
    else
      spill[(i - LDS_DEPTH) * spill_stride + gtid] = (unsigned long long)v.x | ((unsigned long long)v.y << 32);
  };
  auto stk_load = [&](uint32_t i) -> uint2 {
    if (i < LDS_DEPTH) return lds_stack[i * BLOCK + tid];
    const unsigned long long w = spill[(i - LDS_DEPTH) * spill_stride + gtid];
    return make_uint2((uint32_t)w, (uint32_t)(w >> 32));
  };
  // Scene of the current ray: the launch's scene, or (batched) the one whose
  // ray range holds the ray, looked up when the lane takes a new ray.
  const float4* __restrict__ blas = W4 ? p.scene.blas4_nodes : p.scene.blas_nodes;
  const float4* __restrict__ tlas = W4 ? p.scene.tlas4_nodes : p.scene.tlas_nodes;
  const float4* __restrict__ tris = p.scene.tris;
  const float4* __restrict__ insts = p.scene.instances;
  uint32_t tlas_root = W4 ? p.scene.tlas4_root : p.scene.tlas_root, tlas_count = p.scene.tlas_node_count;
  const uint32_t lane = __lane_id();
  const uint64_t lane_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const bool near = !ANY && !COUNT && p.near_first != 0u;        // inside BLASes
  const bool near_tlas = !ANY && !COUNT && p.near_first == 2u;  // also across instances

  const uint32_t n_rays = p.n_ptr ? *p.n_ptr : p.n;
  const uint32_t n_waves = gridDim.x * (BLOCK / 64u);
  const uint32_t take_cap = max(1u, min(64u, n_rays / n_waves + (n_rays % n_waves != 0u)));
  bool active = false;
  uint32_t ray_idx = 0;
  // World ray and best hit.
  V3 ro{}, rd{}, rinv{};
  float tmin = 0, tmax = 0, best = 0, bu = 0, bv = 0;
  V3 bpos{};
  uint32_t bprim = 0, binst = kMiss, bslot = 0;
  bool bfront = false;  // the best hit is on the front face (world orientation)
  // near_first = 2: the lane walks in the reference's TLAS order (after a
  // near-tie), a near-tie was seen, and the magnitude of the ray origin.
  bool ref_order = false, recheck = false;
  float rscale = 0.0f;
  // Instance-local state.
  V3 lo{}, ld{}, linv{}, lpos{};
  WtRay lw{};
  bool lw_ready = false;
  float ltmin = 0, lbest = 0, lu = 0, lv = 0;
  uint32_t lprim = 0, inst_id = 0;
  bool lhit = false, lfront = false;
  bool lmirror = false;  // the instance's local-to-world map reverses orientation
  uint32_t slot = 0, slot_end = 0, enter_slot = 0;
  uint32_t level = 0, cur = kNone, sp = 0, blas_base = 0;
  uint32_t pend = 0;  // pending deferred operation (kPend*)
  uint32_t c_blas = 0, c_tri = 0, c_tlas = 0, c_inst = 0;

  // The world-space bound of the TLAS walk: the best distance, widened by the
  // near-tie margin while the lane orders the TLAS by distance.
  auto world_bound = [&]() -> float {
    if (ANY) return tmax;
    if (near_tlas && !ref_order && binst != kMiss) return best + (best + rscale) * 0x1p-13f;
    return best;
  };

  auto enter_instance = [&](uint32_t s) {
    const float4* ir = insts + 9 * (size_t)s;
    const float4 c0 = ir[0], c1 = ir[1], c2 = ir[2], c3 = ir[3];
    const uint4 meta = *reinterpret_cast<const uint4*>(ir + 8);
    lo = xform_point(c0, c1, c2, c3, ro);
    ld = xform_dir(c0, c1, c2, rd);
    linv = v3(safe_inv(ld.x), safe_inv(ld.y), safe_inv(ld.z));
    lw_ready = false;  // the watertight set-up waits for the first leaf
    const V3 start = v3(ro.x + rd.x * tmin, ro.y + rd.y * tmin, ro.z + rd.z * tmin);
    ltmin = fmaxf(dot3(sub(xform_point(c0, c1, c2, c3, start), lo), ld), 0.0f);
    // Closest: localTMax starts at INFINITY; any-hit: at the ray's tMax
    // (tlasSceneHitTraversal.wgsl.ts:103-108 and :168-173).
    const float bound = world_bound();
    float ltmax = ANY ? tmax : inf_const();
    if (bound < inf_const() * 0.5f) {
      const V3 end = v3(ro.x + rd.x * bound, ro.y + rd.y * bound, ro.z + rd.z * bound);
      ltmax = fmaxf(dot3(sub(xform_point(c0, c1, c2, c3, end), lo), ld), ltmin);
    }
    lbest = ltmax;
    lhit = false;
    lfront = false;
    lmirror = meta.w != 0u;
    inst_id = meta.y;
    level = 1;
    blas_base = sp;
    cur = W4 ? meta.z : meta.x;
    if (COUNT) ++c_inst;
  };

  auto finalize_instance = [&]() {
    if (!ANY && lhit && lbest > ltmin) {
      const float4* ir = insts + 9 * (size_t)slot + 4;
      const V3 w = xform_point(ir[0], ir[1], ir[2], ir[3], lpos);
      const float wd = dot3(sub(w, ro), rd);
      if (near_tlas && !ref_order && binst != kMiss && fabsf(wd - best) <= (best + rscale) * 0x1p-13f)
        recheck = true;
      if (wd > tmin && wd < best) {
        best = wd;
        bpos = w;
        bprim = lprim;
        binst = inst_id;
        bslot = slot;
        bu = lu;
        bv = lv;
        bfront = lfront != lmirror;
      }
    }
  };

  auto finish = [&](bool occluded) {
    if (ANY) {
      p.occluded[ray_idx] = occluded ? 1u : 0u;
    } else {
      const bool hit = binst != kMiss;
      uint32_t* h = p.hits + 5 * (size_t)ray_idx;
      h[0] = hit ? bprim : kMiss;
      h[1] = hit ? binst : kMiss;
      h[2] = hit ? f2u(best) : 0x7F800000u;
      h[3] = hit ? f2u(bu) : 0u;
      h[4] = hit ? f2u(bv) : 0u;
      if (p.hit_pos) {
        float* hp = p.hit_pos + 3 * (size_t)ray_idx;
        hp[0] = bpos.x;
        hp[1] = bpos.y;
        hp[2] = bpos.z;
      }
    }
    if (COUNT) p.counts[ray_idx] = make_uint4(c_blas, c_tri, c_tlas, c_inst);
    active = false;
  };

  auto push = [&](uint32_t ref, float tn) {
    if (sp >= kStackCapacity) {
      atomicOr(p.error, 1u);
      return false;
    }
    stk_store(sp++, make_uint2(ref, f2u(tn)));
    return true;
  };

  // Once a refill hands out indices past the last ray the supply is gone for
  // good; the wave then only finishes its active lanes (no further atomics on
  // the shared counter, which every finishing wave would otherwise hit).
  // The first fill is static -- wave w takes rays [w * take_cap, (w + 1) *
  // take_cap) -- so a launch starts without every wave hitting the shared
  // counter at once (and an empty launch touches it not at all); the counter
  // hands out the rays after that range.
  bool exhausted = false, first_fill = true;
  const uint32_t wave_id = blockIdx.x * (BLOCK / 64u) + threadIdx.x / 64u;
  const uint32_t static_end = n_waves * take_cap;
  for (;;) {
    // ---- wave-level refill: ballot idle lanes, one atomic, prefix ranks ----
    const uint64_t idle = __ballot(!active);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (!exhausted && (n_idle >= p.refill_min || idle == __ballot(1))) {
      const uint32_t take = min(n_idle, take_cap);
      uint32_t base = 0;
      if (first_fill) {
        base = wave_id * take_cap;  // every lane is idle: take == take_cap
      } else {
        if (lane == 0) base = static_end + atomicAdd(p.counter, take);
        base = __shfl(base, 0);
      }
      first_fill = false;
      exhausted = base + take >= n_rays;
      const uint32_t rank = (uint32_t)__popcll(idle & lane_below);
      if (!active && rank < take) {
        const uint32_t idx = base + rank;
        if (idx < n_rays) {
          ray_idx = idx;
          if (MULTI) {
            uint32_t lo_s = 0, hi_s = p.scene_count;
            while (lo_s + 1u < hi_s) {
              const uint32_t mid = (lo_s + hi_s) >> 1;
              if (p.scene_offsets[mid] <= idx) lo_s = mid; else hi_s = mid;
            }
            const DeviceScene& S = p.scenes[lo_s];
            blas = W4 ? S.blas4_nodes : S.blas_nodes;
            tlas = W4 ? S.tlas4_nodes : S.tlas_nodes;
            tris = S.tris;
            insts = S.instances;
            tlas_root = W4 ? S.tlas4_root : S.tlas_root;
            tlas_count = S.tlas_node_count;
          }
          const float4 a = p.rays[2 * (size_t)idx], b = p.rays[2 * (size_t)idx + 1];
          ro = v3(a.x, a.y, a.z);
          tmin = a.w;
          rd = v3(b.x, b.y, b.z);
          tmax = b.w;
          rinv = v3(safe_inv(rd.x), safe_inv(rd.y), safe_inv(rd.z));
          rscale = fmaxf(fabs_(ro.x), fmaxf(fabs_(ro.y), fabs_(ro.z)));
          ref_order = false;
          recheck = false;
          best = tmax;
          binst = kMiss;
          bprim = 0;
          bslot = 0;
          bu = bv = 0;
          level = 0;
          sp = 0;
          pend = kPendNone;
          cur = tlas_count ? tlas_root : kNone;
          if (COUNT) c_blas = c_tri = c_tlas = c_inst = 0;
          active = true;
          if (cur == kNone) finish(false);  // no TLAS: fail closed
        }
      }
    }
    if (!__any(active)) break;

    for (uint32_t step = 0; step < p.steps; ++step) {
      // ---- deferred expensive work, shared by the lanes that need it ----
      // A lane whose next operation is an instance exit (merge of the local
      // hit), an instance entry (ray transform) or a BLAS leaf (triangle
      // tests) waits until `defer_min` lanes of the wave -- or half of its
      // active lanes, when fewer are left -- need the same operation, or
      // until no lane has cheap node work left; each lane still performs
      // exactly its own sequence of operations.
      const uint64_t m_exit = __ballot(active && pend == kPendExit);
      const uint64_t m_enter = __ballot(active && pend == kPendEnter);
      const uint64_t m_leaf = __ballot(active && pend == kPendLeaf);
      const uint64_t m_cheap = __ballot(active && pend == kPendNone);
      if ((m_exit | m_enter | m_leaf | m_cheap) == 0ull) break;
      const bool all = m_cheap == 0ull;
      const uint32_t n_act = (uint32_t)__popcll(m_exit | m_enter | m_leaf | m_cheap);
      const uint32_t dm = min(p.defer_min, (n_act + 1u) >> 1);
      if (m_exit && (all || (uint32_t)__popcll(m_exit) >= dm) && active && pend == kPendExit) {
        finalize_instance();
        level = 0;
        if (++slot < slot_end) {
          enter_slot = slot;
          pend = kPendEnter;
        } else {
          pend = kPendNone;
        }
      }
      if (m_enter && (all || (uint32_t)__popcll(m_enter) >= dm) && active && pend == kPendEnter) {
        slot = enter_slot;
        enter_instance(slot);
        pend = kPendNone;
      }
      if (m_leaf && (all || (uint32_t)__popcll(m_leaf) >= dm)) {
        const bool mine = active && pend == kPendLeaf;
        uint32_t cnt = 0, tfirst = 0;
        if (mine) {
          cnt = (cur >> 27) & 15u;
          tfirst = cur & 0x07FFFFFFu;
          cur = kNone;
          pend = kPendNone;
          if (!lw_ready) {
            lw = wt_setup(ld);
            lw_ready = true;
          }
        }
        if (COUNT) {
          // Work counting: each lane tests its own leaf, one triangle at a time.
          for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t t = tfirst + i;
            const float4* tr = tris + 3 * (size_t)t;
            const float4 w0 = tr[0], w1 = tr[1], w2 = tr[2];
            ++c_tri;
            TriHit h;
            if (tri_test(lo, ld, lw, v3(w0.x, w0.y, w0.z), v3(w0.w, w1.x, w1.y), v3(w1.z, w1.w, w2.x), ltmin,
                         lbest, near && lhit, h)) {
              if (h.t == lbest && lhit && !(t < lprim)) continue;  // equal t: lower index wins
              lhit = true;
              lbest = h.t;
              lprim = t;
              lu = h.u;
              lv = h.v;
              lpos = h.point;
              lfront = h.front;
            }
          }
        } else {
          // The leaves' (ray, triangle) pairs are spread over all 64 lanes,
          // 64 pairs per round: pair q belongs to the lane whose exclusive
          // prefix of triangle counts is the largest one <= q.  Every pair is
          // tested against its ray's state at the start of the round, then
          // each ray keeps the lexicographic minimum (t, triangle index) of
          // its pairs -- exactly what testing them one after another gives
          // (a later triangle replaces the best only with a smaller t or, at
          // equal t, a lower index), so the hit records are unchanged.
          uint32_t excl = 0, total = 0;
#pragma unroll
          for (uint32_t b = 0; b < 4; ++b) {  // counts are below 16: one ballot per bit
            const uint64_t m = __ballot((cnt >> b) & 1u);
            excl += (uint32_t)__popcll(m & lane_below) << b;
            total += (uint32_t)__popcll(m) << b;
          }
          const uint32_t incl = excl + cnt;
          bool occl = false;
          for (uint32_t base = 0; base < total; base += 64u) {
            const uint32_t q = base + lane;
            const bool work = q < total;
            // owner: the lowest lane whose inclusive prefix exceeds q
            uint32_t o_lo = 0, o_hi = 63;
#pragma unroll
            for (uint32_t it = 0; it < 6; ++it) {
              const uint32_t mid = (o_lo + o_hi) >> 1;
              if ((uint32_t)__shfl((int)incl, (int)mid) > q) o_hi = mid; else o_lo = mid + 1u;
            }
            const int o = (int)min(o_lo, 63u);
            const uint32_t o_first = (uint32_t)__shfl((int)tfirst, o), o_excl = (uint32_t)__shfl((int)excl, o);
            const V3 wo = v3(__shfl(lo.x, o), __shfl(lo.y, o), __shfl(lo.z, o));
            const V3 wd = v3(__shfl(ld.x, o), __shfl(ld.y, o), __shfl(ld.z, o));
            WtRay ww;
            ww.ds = __shfl(lw.ds, o);
            ww.sx = __shfl(lw.sx, o);
            ww.sy = __shfl(lw.sy, o);
            ww.sz = __shfl(lw.sz, o);
            ww.perm = (uint32_t)__shfl((int)lw.perm, o);
            const float w_tmin = __shfl(ltmin, o), w_hi = __shfl(lbest, o);
            const bool w_incl = __shfl((int)(near && lhit), o) != 0;
            const bool w_mirror = __shfl((int)lmirror, o) != 0;
            uint32_t kt = 0xFFFFFFFFu, ki = 0xFFFFFFFFu, kl = lane;
            TriHit h;
            h.t = h.u = h.v = 0.0f;
            h.point = v3(0, 0, 0);
            h.front = false;
            // A batched launch can hold rays of different scenes in one wave:
            // the pair's triangle is in its owner's scene.
            const float4* otris = tris;
            if (MULTI) {
              const uint64_t tp = (uint64_t)(uintptr_t)tris;
              const uint32_t p_lo = (uint32_t)__shfl((int)(uint32_t)tp, o);
              const uint32_t p_hi = (uint32_t)__shfl((int)(uint32_t)(tp >> 32), o);
              otris = reinterpret_cast<const float4*>((uintptr_t)(((uint64_t)p_hi << 32) | p_lo));
            }
            if (work) {
              const uint32_t t = o_first + (q - o_excl);
              const float4* tr = otris + 3 * (size_t)t;
              const float4 w0 = tr[0], w1 = tr[1], w2 = tr[2];
              // Any hit through the path tracer's wrappers (traceAny /
              // traceSurfaceVisibility): a back face of a single-sided
              // material does not block.
              if (tri_test(wo, wd, ww, v3(w0.x, w0.y, w0.z), v3(w0.w, w1.x, w1.y), v3(w1.z, w1.w, w2.x), w_tmin,
                           w_hi, w_incl, h) &&
                  (!ANY || !p.sided || h.front != w_mirror || !(w2.y > 0.5f))) {
                kt = f2u(h.t);  // t > 0: the bit patterns order as the values
                ki = t;
              }
            }
            // Segmented minimum over each owner's run of lanes (at most 15).
#pragma unroll
            for (uint32_t d = 1; d < 16; d <<= 1) {
              const uint32_t ot = (uint32_t)__shfl_down((int)kt, d), oi = (uint32_t)__shfl_down((int)ki, d);
              const uint32_t ol = (uint32_t)__shfl_down((int)kl, d);
              const int oo = __shfl_down(o, d);
              if (lane + d < 64u && oo == o && (ot < kt || (ot == kt && oi < ki))) {
                kt = ot;
                ki = oi;
                kl = ol;
              }
            }
            // Each ray reads its run's minimum (its run in this round starts at
            // lane max(excl, base) - base) and the winner's hit data.
            const bool here = mine && cnt > 0u && excl < base + 64u && incl > base;
            const int src = (int)((max(excl, base) - base) & 63u);
            const uint32_t rt = (uint32_t)__shfl((int)kt, src), ri = (uint32_t)__shfl((int)ki, src);
            const int rl = __shfl((int)kl, src);
            const float ru = __shfl(h.u, rl), rv = __shfl(h.v, rl);
            const V3 rp = v3(__shfl(h.point.x, rl), __shfl(h.point.y, rl), __shfl(h.point.z, rl));
            const bool rf = __shfl((int)h.front, rl) != 0;
            if (here && rt != 0xFFFFFFFFu) {
              if (ANY) {
                occl = true;
              } else {
                const float t = u2f(rt);
                if (!lhit || t < lbest || (t == lbest && ri < lprim)) {
                  lhit = true;
                  lbest = t;
                  lprim = ri;
                  lu = ru;
                  lv = rv;
                  lpos = rp;
                  lfront = rf;
                }
              }
            }
          }
          if (ANY && occl) finish(true);
        }
      }
      if (!active || pend != kPendNone) continue;
      // ---- cheap work: pop with the reference's pop-time cull, expand ----
      if (cur == kNone) {
        for (;;) {
          if (level == 1) {
            if (sp > blas_base) {
              const uint2 e = stk_load(--sp);
              if (!(u2f(e.y) > lbest)) {
                cur = e.x;
                break;
              }
              continue;
            }
            pend = kPendExit;  // BLAS exhausted
            break;
          }
          if (sp > 0) {
            const uint2 e = stk_load(--sp);
            if (!(u2f(e.y) > world_bound())) {
              cur = e.x;
              break;
            }
            continue;
          }
          if (recheck) {
            // A near-tie between instances: walk the ray again, from its
            // original interval, in the reference's TLAS order.
            const float4 a = p.rays[2 * (size_t)ray_idx], b = p.rays[2 * (size_t)ray_idx + 1];
            tmin = a.w;
            tmax = b.w;
            best = tmax;
            binst = kMiss;
            bprim = 0;
            bslot = 0;
            bu = bv = 0.0f;
            level = 0;
            sp = 0;
            ref_order = true;
            recheck = false;
            cur = tlas_root;
            break;
          }
          // traceClosest (intersection.wgsl.ts:462-490): a back-face hit on a
          // single-sided material is not a surface; the trace starts again
          // with its distance as the exclusive lower bound, and ends as a
          // miss once that bound stops advancing or reaches tMax.
          if (!ANY && p.sided && binst != kMiss && !bfront && tris[3 * (size_t)bprim + 2].y > 0.5f) {
            if (best > tmin && best < tmax) {
              tmin = best;
              best = tmax;
              binst = kMiss;
              bprim = 0;
              bslot = 0;
              bu = bv = 0.0f;
              level = 0;
              sp = 0;
              cur = tlas_root;
              break;
            }
            binst = kMiss;
          }
          finish(false);
          break;
        }
        if (!active || pend != kPendNone) continue;
      }
      if (W4 && (cur & kLeafBit) == 0) {
        // ---- interior: one Node128, four child boxes tested axis by axis ----
        static_assert(kWideKids == 4, "the expansion below is written for 4-wide nodes");
        const float4* nd = (level ? blas : tlas) + (size_t)kWideNodeFloat4s * cur;
        const V3 o = level ? lo : ro;
        const V3 inv = level ? linv : rinv;
        const float tm = level ? ltmin : tmin;
        const float bound = level ? lbest : world_bound();
        float4 a = nd[0], b = nd[3];
        float n0, n1, n2, n3, f0, f1, f2, f3;
        {
          const float p0 = (a.x - o.x) * inv.x, q0 = (b.x - o.x) * inv.x;
          const float p1 = (a.y - o.x) * inv.x, q1 = (b.y - o.x) * inv.x;
          const float p2 = (a.z - o.x) * inv.x, q2 = (b.z - o.x) * inv.x;
          const float p3 = (a.w - o.x) * inv.x, q3 = (b.w - o.x) * inv.x;
          n0 = fminf(p0, q0), f0 = fmaxf(p0, q0);
          n1 = fminf(p1, q1), f1 = fmaxf(p1, q1);
          n2 = fminf(p2, q2), f2 = fmaxf(p2, q2);
          n3 = fminf(p3, q3), f3 = fmaxf(p3, q3);
        }
        a = nd[1], b = nd[4];
        {
          const float p0 = (a.x - o.y) * inv.y, q0 = (b.x - o.y) * inv.y;
          const float p1 = (a.y - o.y) * inv.y, q1 = (b.y - o.y) * inv.y;
          const float p2 = (a.z - o.y) * inv.y, q2 = (b.z - o.y) * inv.y;
          const float p3 = (a.w - o.y) * inv.y, q3 = (b.w - o.y) * inv.y;
          n0 = fmaxf(n0, fminf(p0, q0)), f0 = fminf(f0, fmaxf(p0, q0));
          n1 = fmaxf(n1, fminf(p1, q1)), f1 = fminf(f1, fmaxf(p1, q1));
          n2 = fmaxf(n2, fminf(p2, q2)), f2 = fminf(f2, fmaxf(p2, q2));
          n3 = fmaxf(n3, fminf(p3, q3)), f3 = fminf(f3, fmaxf(p3, q3));
        }
        a = nd[2], b = nd[5];
        {
          const float p0 = (a.x - o.z) * inv.z, q0 = (b.x - o.z) * inv.z;
          const float p1 = (a.y - o.z) * inv.z, q1 = (b.y - o.z) * inv.z;
          const float p2 = (a.z - o.z) * inv.z, q2 = (b.z - o.z) * inv.z;
          const float p3 = (a.w - o.z) * inv.z, q3 = (b.w - o.z) * inv.z;
          n0 = fmaxf(n0, fminf(p0, q0)), f0 = fminf(f0, fmaxf(p0, q0));
          n1 = fmaxf(n1, fminf(p1, q1)), f1 = fminf(f1, fmaxf(p1, q1));
          n2 = fmaxf(n2, fminf(p2, q2)), f2 = fminf(f2, fmaxf(p2, q2));
          n3 = fmaxf(n3, fminf(p3, q3)), f3 = fminf(f3, fmaxf(p3, q3));
        }
        const uint4 r = *reinterpret_cast<const uint4*>(nd + 6);
        // The slab test of each child: !(tNear > tFar || tFar < tMin), then
        // the bound (the pop-time cull re-checks it against the best hit).
        const bool k0 = r.x != kEmpty && !(n0 > f0 || f0 < tm) && !(n0 > bound);
        const bool k1 = r.y != kEmpty && !(n1 > f1 || f1 < tm) && !(n1 > bound);
        const bool k2 = r.z != kEmpty && !(n2 > f2 || f2 < tm) && !(n2 > bound);
        const bool k3 = r.w != kEmpty && !(n3 > f3 || f3 < tm) && !(n3 > bound);
        if (COUNT) {
          const uint32_t tested = (r.x != kEmpty) + (r.y != kEmpty) + (r.z != kEmpty) + (r.w != kEmpty);
          if (level) c_blas += tested; else c_tlas += tested;
        }
        // Visit first: the nearest passing child (inside BLASes with
        // near_first, lowest slot on equal distances), else the first passing
        // slot, which is the reference's depth-first order.  The others are
        // pushed so that they pop in slot order.
        uint32_t first = k0 ? 0u : (k1 ? 1u : (k2 ? 2u : (k3 ? 3u : 4u)));
        if ((level ? near : (near_tlas && !ref_order)) && first < 4u) {
          float bt = n0;
          bt = first == 1u ? n1 : (first == 2u ? n2 : (first == 3u ? n3 : bt));
          if (k1 && n1 < bt) first = 1u, bt = n1;
          if (k2 && n2 < bt) first = 2u, bt = n2;
          if (k3 && n3 < bt) first = 3u, bt = n3;
        }
        bool ok = true;
        if (k3 && first != 3u) ok = push(r.w, n3);
        if (ok && k2 && first != 2u) ok = push(r.z, n2);
        if (ok && k1 && first != 1u) ok = push(r.y, n1);
        if (ok && k0 && first != 0u) ok = push(r.x, n0);
        if (!ok) {
          finish(false);
          continue;
        }
        cur = first == 0u ? r.x : (first == 1u ? r.y : (first == 2u ? r.z : (first == 3u ? r.w : kNone)));
      } else if (!W4 && (cur & kLeafBit) == 0) {
        // ---- interior: one Node64 load, two child boxes ----
        const float4* nd = (level ? blas : tlas) + 4 * (size_t)cur;
        const float4 q0 = nd[0], q1 = nd[1], q2 = nd[2], q3 = nd[3];
        const uint32_t ref_l = f2u(q3.x), ref_r = f2u(q3.y);
        const V3 o = level ? lo : ro;
        const V3 inv = level ? linv : rinv;
        const float tm = level ? ltmin : tmin;
        const float bound = level ? lbest : world_bound();
        float tn_l, tn_r;
        const bool pass_l = ref_l != kEmpty && slab(o, inv, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, tm, tn_l);
        const bool pass_r = ref_r != kEmpty && slab(o, inv, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, tm, tn_r);
        if (COUNT) {
          const uint32_t tested = (ref_l != kEmpty) + (ref_r != kEmpty);
          if (level) c_blas += tested; else c_tlas += tested;
        }
        const bool take_l = pass_l && !(tn_l > bound), take_r = pass_r && !(tn_r > bound);
        if (take_l && take_r) {
          const bool right_first = (level ? near : (near_tlas && !ref_order)) && tn_r < tn_l;
          if (!push(right_first ? ref_l : ref_r, right_first ? tn_l : tn_r)) {
            finish(false);
            continue;
          }
          cur = right_first ? ref_r : ref_l;
        } else {
          cur = take_l ? ref_l : (take_r ? ref_r : kNone);
        }
      }
      // A leaf reached by the pop or the expansion becomes pending work.
      if (cur != kNone && (cur & kLeafBit)) {
        if (level == 0) {
          enter_slot = cur & 0x07FFFFFFu;
          slot_end = enter_slot + ((cur >> 27) & 15u);
          cur = kNone;
          pend = enter_slot < slot_end ? kPendEnter : kPendNone;
        } else {
          pend = kPendLeaf;
        }
      }
    }
  }
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void copy16_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ void fp32_selftest_kernel(const float* a, const float* b, float* q, float* r, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    q[i] = a[i] / b[i];
    r[i] = __builtin_sqrtf(a[i]);
  }
}

}  // namespace vtk

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

namespace vtk {

template <bool ANY, bool COUNT, bool MULTI, bool W4>
static hipError_t launch_one(const TraceParams& p, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((trace_kernel<ANY, COUNT, MULTI, W4, kTraceBlock, kLdsDepth>), dim3(grid),
                     dim3(kTraceBlock), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_trace_batch(const TraceParams& p, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, p.counter, 1u);
  return p.wide ? launch_one<false, false, true, true>(p, grid, s)
                             : launch_one<false, false, true, false>(p, grid, s);
}

hipError_t launch_trace_batch_any(const TraceParams& p, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, p.counter, 1u);
  return p.wide ? launch_one<true, false, true, true>(p, grid, s) : launch_one<true, false, true, false>(p, grid, s);
}

hipError_t launch_trace(const TraceParams& p, bool any, bool count, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, p.counter, 1u);
  if (count) return launch_one<false, true, false, false>(p, grid, s);
  const bool w4 = p.wide != 0;
  if (any) return w4 ? launch_one<true, false, false, true>(p, grid, s) : launch_one<true, false, false, false>(p, grid, s);
  return w4 ? launch_one<false, false, false, true>(p, grid, s) : launch_one<false, false, false, false>(p, grid, s);
}

template <bool ANY, bool COUNT, bool MULTI, bool W4>
static int occupancy_of() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_kernel<ANY, COUNT, MULTI, W4, kTraceBlock, kLdsDepth>,
                                                   kTraceBlock, 0) != hipSuccess)
    return -1;
  return n;
}

static hipError_t min_occupancy(std::initializer_list<int> v, int* out) {
  int m = 1 << 30;
  for (int x : v) {
    if (x < 0) return hipErrorInvalidValue;
    m = x < m ? x : m;
  }
  *out = m;
  return hipSuccess;
}

hipError_t trace_occupancy_batch(int* blocks_per_cu) {
  return min_occupancy({occupancy_of<false, false, true, false>(), occupancy_of<false, false, true, true>()},
                       blocks_per_cu);
}

hipError_t trace_occupancy(bool any, int* blocks_per_cu) {
  if (any)
    return min_occupancy({occupancy_of<true, false, false, false>(), occupancy_of<true, false, false, true>()},
                         blocks_per_cu);
  // The counting variant (measurement only) is left out: its extra blocks
  // simply start as others retire.
  return min_occupancy({occupancy_of<false, false, false, false>(), occupancy_of<false, false, false, true>()},
                       blocks_per_cu);
}

hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_copy16(uint4* dst, const uint4* src, size_t n16, hipStream_t s) {
  if (n16 == 0) return hipSuccess;
  const size_t want = (n16 + 255) / 256;
  const uint32_t blocks = (uint32_t)(want < 8192 ? want : 8192);
  hipLaunchKernelGGL(copy16_kernel, dim3(blocks), dim3(256), 0, s, dst, src, n16);
  return hipGetLastError();
}

hipError_t launch_fp32_selftest(const float* a, const float* b, float* q, float* r, uint32_t n,
                                hipStream_t s) {
  hipLaunchKernelGGL(fp32_selftest_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a, b, q, r, n);
  return hipGetLastError();
}

}  // namespace vtk

// ---------------------------------------------------------------------------
// cwbvh-closest mode: pt-webgpu's opt-in wide traversal
// (pathTrace/intersection.wgsl.ts:576-838).  One ray per lane over the
// reference-layout TLAS; each BLAS is walked in its 8-wide compressed tree
// (cwbvhIntersect.wgsl.ts:197-336) with the Moller-Trumbore core
// (bvhIntersect.wgsl.ts:160-238); the proposed hit is accepted only when the
// watertight test reproduces its distance bit for bit, and a disagreement, an
// invalid wide layout or a stack overflow answers that BLAS with the binary
// traversal (intersectionCore.wgsl.ts:461-596).  Any-hit rays never use it.
// ---------------------------------------------------------------------------

namespace vtw {

using vtk::F4;
using vtk::TriHit;
using vtk::V3;
using vtk::WtRay;

constexpr uint32_t kWideStack = 64;  // cwbvhIntersect.wgsl.ts stack depth
constexpr uint32_t kBinStack = 60;   // strides.ts:110
constexpr uint32_t kTlasStack = 64;  // strides.ts:112
constexpr uint32_t kLeafMask = 0xFFFF0000u;

// The WGSL select-based min / max (NaN compares false).
__device__ __forceinline__ float max2(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float min2(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float cw_inf() { return 3.402823e38f; }

__device__ __forceinline__ float decode(uint32_t q, float mn, float mx) {
  if (!(mx > mn)) return mn;
  if (q == 0) return mn;
  if (q >= 65535u) return mx;
  const float t = (float)q / 65535.0f;
  return mn * (1.0f - t) + mx * t;
}

__device__ __forceinline__ bool bounds_valid(V3 mn, V3 mx) {
  const float m = cw_inf();
  return mn.x <= mx.x && mn.y <= mx.y && mn.z <= mx.z && vtk::fabs_(mn.x) <= m && vtk::fabs_(mn.y) <= m &&
         vtk::fabs_(mn.z) <= m && vtk::fabs_(mx.x) <= m && vtk::fabs_(mx.y) <= m && vtk::fabs_(mx.z) <= m;
}

__device__ __forceinline__ bool aabb_entry(V3 o, V3 inv, V3 bmn, V3 bmx, float tmax) {
  const float t0x = (bmn.x - o.x) * inv.x, t0y = (bmn.y - o.y) * inv.y, t0z = (bmn.z - o.z) * inv.z;
  const float t1x = (bmx.x - o.x) * inv.x, t1y = (bmx.y - o.y) * inv.y, t1z = (bmx.z - o.z) * inv.z;
  const float tn = max2(max2(min2(t0x, t1x), min2(t0y, t1y)), min2(t0z, t1z));
  const float tf = min2(min2(max2(t0x, t1x), max2(t0y, t1y)), max2(t0z, t1z));
  return !(tn > tf || tf < 0.0f || tn > tmax);
}

__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
  return vtk::v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 scale_div(V3 a, float s) { return vtk::v3(a.x / s, a.y / s, a.z / s); }

// mollerTrumboreCore: equilibrated edges and direction, angular determinant
// rejection at 1e-7, barycentric tolerance 1e-6.  Returns t, or -1 for a miss.
__device__ __noinline__ float mt_core(V3 o, V3 dir, V3 a, V3 b, V3 c, float tmin) {
  using vtk::dot3;
  using vtk::fabs_;
  using vtk::nan_;
  const V3 e1r = vtk::sub(b, a), e2r = vtk::sub(c, a);
  const float es = max2(max2(max2(fabs_(e1r.x), fabs_(e1r.y)), fabs_(e1r.z)),
                        max2(max2(fabs_(e2r.x), fabs_(e2r.y)), fabs_(e2r.z)));
  const float ds = max2(fabs_(dir.x), max2(fabs_(dir.y), fabs_(dir.z)));
  if (!(es > 0.0f) || es > cw_inf() || !(ds > 0.0f) || ds > cw_inf()) return -1.0f;
  const V3 e1 = scale_div(e1r, es), e2 = scale_div(e2r, es), dn = scale_div(dir, ds);
  const V3 ao = scale_div(vtk::sub(o, a), es);
  const V3 n = cross3(e1, e2);
  const float ns = max2(fabs_(n.x), max2(fabs_(n.y), fabs_(n.z)));
  if (!(ns > 0.0f) || ns > cw_inf()) return -1.0f;
  const V3 nd = scale_div(n, ns);
  const float nl = __builtin_sqrtf(dot3(nd, nd));
  const float dl = __builtin_sqrtf(dot3(dn, dn));
  const V3 rd = scale_div(dn, dl);
  const float det = -dot3(dn, n);
  const V3 nn = scale_div(nd, nl);
  const float ang = fabs_(dot3(rd, nn));
  if (!(ang > 1e-7f) || det == 0.0f) return -1.0f;
  const float inv_det = 1.0f / det;
  const V3 dao = cross3(ao, dn);
  const float u = dot3(e2, dao) * inv_det;
  const float v = -dot3(e1, dao) * inv_det;
  const float st = dot3(ao, n) * inv_det;
  const float t = (st * es) / ds;
  const float w = (1.0f - u) - v;
  if (nan_(u) || nan_(v) || nan_(w) || nan_(t) || fabs_(u) > cw_inf() || fabs_(v) > cw_inf() ||
      fabs_(w) > cw_inf() || fabs_(t) > cw_inf() || u < -1e-6f || v < -1e-6f || w < -1e-6f ||
      t < max2(tmin, 0.0f))
    return -1.0f;
  return t;
}

__device__ __forceinline__ V3 vload(const Scene& s, uint32_t i) {
  const float4 p = s.positions[i];
  return vtk::v3(p.x, p.y, p.z);
}

struct LocalHit {
  bool hit;
  float t;
  V3 pos;
  uint32_t tri;
  float u, v;
  bool front;  // local front face (the watertight test's side)
};

// traceMeshBvh over the reference node layout (8 words per node), left child
// first, closest hit.
__device__ __noinline__ LocalHit mesh_binary(const Scene& s, V3 o, V3 d, V3 inv, const WtRay& w, float tmin,
                                             float tmax, uint32_t root) {
  LocalHit h{false, tmax, vtk::v3(0, 0, 0), 0u, 0.0f, 0.0f, false};
  if (s.bvh_node_count == 0 || root >= s.bvh_node_count) return h;
  uint32_t stack[kBinStack];
  uint32_t sp = 0;
  stack[sp++] = root;
  while (sp > 0) {
    const uint32_t ni = stack[--sp];
    if (ni >= s.bvh_node_count) continue;
    const uint4* nd = reinterpret_cast<const uint4*>(s.bvh_nodes + 8 * (size_t)ni);
    const uint4 lo = nd[0], hi = nd[1];
    float tn;
    if (!vtk::slab(o, inv, vtk::u2f(lo.x), vtk::u2f(lo.y), vtk::u2f(lo.z), vtk::u2f(lo.w), vtk::u2f(hi.x),
                   vtk::u2f(hi.y), tmin, tn) ||
        tn > h.t)
      continue;
    if ((hi.w & kLeafMask) == kLeafMask) {
      const uint32_t count = hi.w & 0xFFFFu, start = hi.z;
      for (uint32_t i = 0; i < count; ++i) {
        const uint32_t t = start + i;
        if (t >= s.tri_count) continue;
        const uint4 ix = s.indices[t];
        if (ix.x >= s.vertex_count || ix.y >= s.vertex_count || ix.z >= s.vertex_count) continue;
        TriHit th;
        if (vtk::tri_test(o, d, w, vload(s, ix.x), vload(s, ix.y), vload(s, ix.z), tmin, h.t, false, th)) {
          h.hit = true;
          h.t = th.t;
          h.pos = th.point;
          h.tri = t;
          h.u = th.u;
          h.v = th.v;
          h.front = th.front;
        }
      }
    } else {
      if (sp + 2 > kBinStack) return h;
      stack[sp++] = ni + hi.z;
      stack[sp++] = ni + 1;
    }
  }
  return h;
}

// Wide walk.  status: 0 complete, 1 stack overflow, 2 invalid layout.
struct WideHit {
  uint32_t status;
  bool hit;
  float t;
  uint32_t tri;
};

__device__ __noinline__ WideHit wide_walk(const Scene& s, V3 o, V3 d, V3 inv, float tmin, float tmax,
                                          uint32_t root) {
  WideHit best{0u, false, tmax, 0u};
  const uint32_t nc = s.wide_count;
  if (nc == 0) return best;
  if (root >= nc) {
    best.status = 2;
    return best;
  }
  uint32_t stack[kWideStack];
  uint32_t sp = 0;
  stack[sp++] = root;
  while (sp > 0) {
    const uint32_t node = stack[--sp];
    if (node >= nc) {
      best.status = 2;
      return best;
    }
    const uint32_t count = s.child_count[node];
    if (count > kKids) {
      best.status = 2;
      return best;
    }
    const float* pb = s.node_bounds + 6 * (size_t)node;
    const V3 pmn = vtk::v3(pb[0], pb[1], pb[2]), pmx = vtk::v3(pb[3], pb[4], pb[5]);
    if (!bounds_valid(pmn, pmx)) {
      best.status = 2;
      return best;
    }
    for (uint32_t slot = 0; slot < count; ++slot) {
      const size_t k = ((size_t)node * kKids + slot) * 3;
      const uint32_t kind = s.child_meta[k], ref = s.child_meta[k + 1], cnt = s.child_meta[k + 2];
      if (kind == 0) {
        best.status = 2;
        return best;
      }
      const uint32_t q0 = s.child_bounds[k], q1 = s.child_bounds[k + 1], q2 = s.child_bounds[k + 2];
      const V3 bmn = vtk::v3(decode(q0 & 0xFFFFu, pmn.x, pmx.x), decode(q0 >> 16, pmn.y, pmx.y),
                             decode(q1 & 0xFFFFu, pmn.z, pmx.z));
      const V3 bmx = vtk::v3(decode(q1 >> 16, pmn.x, pmx.x), decode(q2 & 0xFFFFu, pmn.y, pmx.y),
                             decode(q2 >> 16, pmn.z, pmx.z));
      if (!bounds_valid(bmn, bmx)) {
        best.status = 2;
        return best;
      }
      if (!aabb_entry(o, inv, bmn, bmx, best.t)) continue;
      if (kind == 1) {
        if (ref >= nc) {
          best.status = 2;
          return best;
        }
        if (sp >= kWideStack) {
          best.status = 1;
          return best;
        }
        stack[sp++] = ref;
      } else if (kind == 2) {
        if (cnt == 0 || ref > s.tri_count || cnt > s.tri_count - ref) {
          best.status = 2;
          return best;
        }
        for (uint32_t i = 0; i < cnt; ++i) {
          const uint32_t tri = ref + i;
          const uint4 ix = s.indices[tri];
          if (ix.x >= s.vertex_count || ix.y >= s.vertex_count || ix.z >= s.vertex_count) {
            best.status = 2;
            return best;
          }
          const float t = mt_core(o, d, vload(s, ix.x), vload(s, ix.y), vload(s, ix.z), tmin);
          if (t >= 0.0f && t > tmin && t < best.t) {
            best.hit = true;
            best.t = t;
            best.tri = tri;
          }
        }
      } else {
        best.status = 2;
        return best;
      }
    }
  }
  return best;
}

// traceMeshCwbvhClosest.
__device__ __forceinline__ LocalHit mesh_cwbvh(const Scene& s, V3 o, V3 d, V3 inv, const WtRay& w, float tmin,
                                               float tmax, uint32_t wide_root, uint32_t binary_root) {
  if (s.wide_count == 0 || wide_root >= s.wide_count) return mesh_binary(s, o, d, inv, w, tmin, tmax, binary_root);
  const WideHit c = wide_walk(s, o, d, inv, tmin, tmax, wide_root);
  if (c.status != 0) return mesh_binary(s, o, d, inv, w, tmin, tmax, binary_root);
  LocalHit h{false, tmax, vtk::v3(0, 0, 0), 0u, 0.0f, 0.0f, false};
  if (!c.hit || c.t <= tmin || c.t >= tmax) return h;
  if (c.tri >= s.tri_count) return mesh_binary(s, o, d, inv, w, tmin, tmax, binary_root);
  const uint4 ix = s.indices[c.tri];
  TriHit e;
  if (!vtk::tri_test(o, d, w, vload(s, ix.x), vload(s, ix.y), vload(s, ix.z), tmin, vtk::inf_const(), true, e) ||
      vtk::f2u(e.t) != vtk::f2u(c.t))
    return mesh_binary(s, o, d, inv, w, tmin, tmax, binary_root);
  h.hit = true;
  h.t = c.t;
  h.pos = e.point;
  h.tri = c.tri;
  h.u = e.u;
  h.v = e.v;
  h.front = e.front;
  return h;
}

__device__ __forceinline__ void load_cols(const float* m, float4& c0, float4& c1, float4& c2, float4& c3) {
  const float4* q = reinterpret_cast<const float4*>(m);
  c0 = q[0];
  c1 = q[1];
  c2 = q[2];
  c3 = q[3];
}

struct WorldHit {
  float t;
  uint32_t prim, inst;
  float u, v;
  V3 pos;
  bool front;  // world front face (flipped by an orientation-reversing instance)
};

// One instance of traceTlasClosestCwbvh / tlasInstanceClosest: the local ray,
// its [ltmin, ltmax] window, the BLAS walk and the world-distance acceptance.
__device__ __forceinline__ void instance_closest(const Scene& s, uint32_t inst, bool wide, V3 ro, V3 rd,
                                                 float tmin, WorldHit& b) {
  if (inst >= s.inst_count) return;
  float4 c0, c1, c2, c3;
  load_cols(s.w2l + 16 * (size_t)inst, c0, c1, c2, c3);
  const V3 lo = vtk::xform_point(c0, c1, c2, c3, ro);
  const V3 ld = vtk::xform_dir(c0, c1, c2, rd);
  const V3 linv = vtk::v3(vtk::safe_inv(ld.x), vtk::safe_inv(ld.y), vtk::safe_inv(ld.z));
  const V3 sp = vtk::v3(ro.x + rd.x * tmin, ro.y + rd.y * tmin, ro.z + rd.z * tmin);
  const float ltmin = max2(vtk::dot3(vtk::sub(vtk::xform_point(c0, c1, c2, c3, sp), lo), ld), 0.0f);
  float ltmax = vtk::inf_const();
  if (b.t < vtk::inf_const() * 0.5f) {
    const V3 ep = vtk::v3(ro.x + rd.x * b.t, ro.y + rd.y * b.t, ro.z + rd.z * b.t);
    ltmax = max2(vtk::dot3(vtk::sub(vtk::xform_point(c0, c1, c2, c3, ep), lo), ld), ltmin);
  }
  const WtRay w = vtk::wt_setup(ld);
  const LocalHit lh = wide ? mesh_cwbvh(s, lo, ld, linv, w, ltmin, ltmax, s.wide_roots[inst], s.blas_roots[inst])
                           : mesh_binary(s, lo, ld, linv, w, ltmin, ltmax, s.blas_roots[inst]);
  if (!lh.hit || !(lh.t > ltmin)) return;
  load_cols(s.l2w + 16 * (size_t)inst, c0, c1, c2, c3);
  const V3 wpos = vtk::xform_point(c0, c1, c2, c3, lh.pos);
  const float wd = vtk::dot3(vtk::sub(wpos, ro), rd);
  // intersection.wgsl.ts:811 rejects with `<=` / `>=` (a NaN distance is kept),
  // unlike tlasInstanceClosest's `>` / `<` acceptance.
  if (wide && (wd <= tmin || wd >= b.t)) return;
  if (!wide && !(wd > tmin && wd < b.t)) return;
  b.t = wd;
  b.prim = lh.tri;
  b.inst = inst;
  b.u = lh.u;
  b.v = lh.v;
  b.pos = wpos;
  b.front = lh.front != (s.inst_reversed[inst] != 0);
}

__global__ __launch_bounds__(128) void trace_cwbvh_kernel(Params p) {
  const Scene& s = p.s;
  // With a device count, p.n is the capacity of the ray and hit buffers.
  const uint32_t n = p.n_ptr ? min(*p.n_ptr, p.n) : p.n;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float4 ra = p.rays[2 * (size_t)i], rb = p.rays[2 * (size_t)i + 1];
    const V3 ro = vtk::v3(ra.x, ra.y, ra.z), rd = vtk::v3(rb.x, rb.y, rb.z);
    const float tmax = rb.w;
    const V3 rinv = vtk::v3(vtk::safe_inv(rd.x), vtk::safe_inv(rd.y), vtk::safe_inv(rd.z));
    WorldHit b;
    // traceClosest (intersection.wgsl.ts:462-490) with p.sided: a back-face
    // hit on a single-sided material is skipped by tracing again from its
    // distance (exclusive); a bound that stops advancing or reaches tMax
    // ends the ray as a miss.
    float tmin = ra.w;
    for (;;) {
      // Mode 0: wide BLAS walks; 1: binary BLAS walks (a TLAS stack overflow in
      // mode 0 restarts here, as traceTlasClosestCwbvh falls back to
      // traceTlasClosest); 2: every instance in permutation order
      // (tlasClosestFallback, on a TLAS stack overflow in mode 1).
      uint32_t mode = (s.tlas_node_count == 0 || s.inst_index_count == 0) ? 3u : 0u;
      for (;;) {
        b.t = tmax;
        b.prim = 0;
        b.inst = vtk::kMiss;
        b.u = b.v = 0.0f;
        b.pos = vtk::v3(0, 0, 0);
        b.front = false;
        if (mode == 3) break;  // no TLAS: fail closed
        if (mode == 2) {
          for (uint32_t pi = 0; pi < s.inst_index_count; ++pi)
            instance_closest(s, s.inst_indices[pi], false, ro, rd, tmin, b);
          break;
        }
        uint32_t stack[kTlasStack];
        uint32_t sp = 0;
        stack[sp++] = 0;
        bool overflow = false;
        while (sp > 0 && !overflow) {
          const uint32_t ni = stack[--sp];
          if (ni >= s.tlas_node_count) continue;
          const uint4* nd = reinterpret_cast<const uint4*>(s.tlas_nodes + 8 * (size_t)ni);
          const uint4 lo = nd[0], hi = nd[1];
          float tn;
          if (!vtk::slab(ro, rinv, vtk::u2f(lo.x), vtk::u2f(lo.y), vtk::u2f(lo.z), vtk::u2f(lo.w),
                         vtk::u2f(hi.x), vtk::u2f(hi.y), tmin, tn) ||
              tn > b.t)
            continue;
          if ((hi.w & kLeafMask) == kLeafMask) {
            const uint32_t count = hi.w & 0xFFFFu, start = hi.z;
            for (uint32_t k = 0; k < count; ++k) {
              const uint32_t pi = start + k;
              if (pi >= s.inst_index_count) continue;
              instance_closest(s, s.inst_indices[pi], mode == 0, ro, rd, tmin, b);
            }
          } else if (sp + 2 <= kTlasStack) {
            stack[sp++] = ni + hi.z;
            stack[sp++] = ni + 1;
          } else {
            overflow = true;
          }
        }
        if (!overflow) break;
        ++mode;
      }
      if (!p.sided || b.inst == vtk::kMiss || b.front || b.prim >= s.tri_count || !s.single_sided[b.prim]) break;
      if (!(b.t > tmin) || !(b.t < tmax)) {
        b.inst = vtk::kMiss;
        break;
      }
      tmin = b.t;
    }
    const bool hit = b.inst != vtk::kMiss;
    uint32_t* h = p.hits + 5 * (size_t)i;
    h[0] = hit ? b.prim : vtk::kMiss;
    h[1] = hit ? b.inst : vtk::kMiss;
    h[2] = hit ? vtk::f2u(b.t) : 0x7F800000u;
    h[3] = hit ? vtk::f2u(b.u) : 0u;
    h[4] = hit ? vtk::f2u(b.v) : 0u;
    if (p.hit_pos) {
      float* hp = p.hit_pos + 3 * (size_t)i;
      hp[0] = b.pos.x;
      hp[1] = b.pos.y;
      hp[2] = b.pos.z;
    }
  }
}

hipError_t launch_trace(const Params& p, uint32_t grid, hipStream_t s) {
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(trace_cwbvh_kernel, dim3(grid), dim3(128), 0, s, p);
  return hipGetLastError();
}

}  // namespace vtw
