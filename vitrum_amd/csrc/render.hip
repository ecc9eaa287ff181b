// render.hip -- wavefront unidirectional path tracer for gfx950.
//
// The estimator is the one of pt-webgpu's `main` megakernel
// (packages/pt-webgpu/src/wgsl/pathTrace/kernel.wgsl.ts:962-2032) for opaque
// Lambert + GGX materials, mesh-area emitters and an environment map, with
// the per-path RNG consumed in the same order.  The megakernel is split into
// stages that run over compacted queues:
//
//   camera  -> for each vertex: closest-hit (traverse.hip) -> shade -> any-hit
//              shadow rays (traverse.hip) -> shadow resolve -> ... -> accumulate
//
// The BSDF->light MIS connections (connect.wgsl.ts:478-700) are evaluated
// where the wavefront naturally has the data: the BSDF-sampled ray *is* the
// next extension ray (same origin, direction and tMin), so its closest hit
// decides which emitter the connection sees and whether the environment is
// visible.  The connection is weighted with the throughput before Russian
// roulette, and a path killed by roulette still has its last ray traced for
// the connection, exactly as the megakernel adds the connection before the
// roulette draw.
//
// Queues are appended with wave-aggregated atomics: a ballot of the lanes
// that emit, one atomicAdd by the first of them, and mbcnt ranks.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "render.h"
#include "sobol.h"

namespace vtr {

constexpr float kPi = 3.14159265358979f;
constexpr float kInvPi = 0.31830988618f;
constexpr float kInv2Pi = 0.15915494309189535f;
constexpr uint32_t kMiss = 0xFFFFFFFFu;

__device__ __forceinline__ float safe_max() { return __uint_as_float(0x7F7FFFFDu); }
__device__ __forceinline__ float f32_max() { return __uint_as_float(0x7F7FFFFFu); }

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ F3 load3(const float* p) { return f3(p[0], p[1], p[2]); }
__device__ __forceinline__ void store3(float* p, F3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

// safe_normalize (pt-webgpu/src/wgsl/common.wgsl.ts)
__device__ F3 safe_normalize(F3 v) {
  const float s = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
  if (!(s > 0.0f) || s > safe_max()) return f3(0, 1, 0);
  const F3 q = f3(v.x / s, v.y / s, v.z / s);
  const float l = __builtin_sqrtf(dot(q, q));
  return f3(q.x / l, q.y / l, q.z / l);
}

// safe_length (common.wgsl.ts)
__device__ float safe_length(F3 v) {
  const float s = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
  if (!(s > 0.0f) || s > f32_max()) return 0.0f;
  const F3 q = f3(v.x / s, v.y / s, v.z / s);
  const float r = s * __builtin_sqrtf(dot(q, q));
  return (!(r > 0.0f) || r > f32_max()) ? 0.0f : r;
}

// ---- RNG (shared-samplers/src/wgsl/pcg.wgsl.ts; common.wgsl.ts) ---------

__device__ __forceinline__ uint32_t pcg_init(uint32_t px, uint32_t py, uint32_t key) {
  uint32_t s = px * 1664525u + py * 1013904223u + key * 22695477u;
  s ^= s >> 17;
  s ^= s << 31;
  s ^= s >> 11;
  return s;
}
__device__ __forceinline__ uint32_t pcg_next(uint32_t& s) {
  s = s * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
  return (w >> 22) ^ w;
}
// Per-path sampler: PCG (dim == kPcgOnly) or pt-webgpu's Owen-scrambled
// Sobol stream, which continues with PCG after its last dimension.
constexpr uint32_t kPcgOnly = 0xFFFFFFFFu;
struct Rng {
  uint32_t s;     // PCG state / Sobol continuation state
  uint32_t dim;   // next Sobol dimension, kPcgOnly for PCG
  uint32_t seed;  // Sobol stream seed (pixel and sequence key)
  uint32_t tile;  // Sobol rotation tile
  uint32_t sample;
  const uint16_t* dirs;
};
__device__ __forceinline__ uint32_t rng_next(Rng& r) {
  if (r.dim >= vts::kDims) {
    if (r.dim != kPcgOnly) ++r.dim;
    return pcg_next(r.s);
  }
  const uint32_t v = vts::owen_u32(r.dirs, r.sample, r.dim, r.tile, r.seed);
  ++r.dim;
  return v;
}
__device__ __forceinline__ float rand01(Rng& r) { return (float)(rng_next(r) >> 8) / 16777216.0f; }
// Polar disc coordinates sqrt(u1)·cos(2π·u2), sqrt(u1)·sin(2π·u2), evaluated
// in binary64 and rounded once to binary32.  A phase rounded to binary32
// first (2π·u2 near 6 carries up to 2.4e-7 rad) and a binary32 cosine would
// tilt low-roughness GGX samples enough to move the next hit measurably; the
// binary64 restatement of the integrator rounds the same two values once.
__device__ __forceinline__ void disc_polar(float u1, float u2, float& a, float& b) {
  double s, c;
  sincos(6.283185307179586 * (double)u2, &s, &c);
  const double r = __builtin_sqrt((double)u1);
  a = (float)(r * c);
  b = (float)(r * s);
}
__device__ float represented(float p) {
  if (!(p > 0.0f)) return 0.0f;
  if (p >= 1.0f) return 1.0f;
  const float b = clampf(floorf(p * 16777216.0f + 0.5f), 1.0f, 16777215.0f);
  return b / 16777216.0f;
}
__device__ uint32_t rand_bounded(Rng& s, uint32_t bound) {
  if (bound <= 1u) return 0u;
  const uint32_t threshold = ((0xFFFFFFFFu % bound) + 1u) % bound;
  for (;;) {
    const uint32_t hi = rng_next(s) & 0xFFFFFF00u;
    const uint32_t lo = rng_next(s) >> 24;
    const uint32_t w = hi | lo;
    if (w >= threshold) return w % bound;
  }
}

// ---- light tree (packages/shared-samplers/src/wgsl/lightTree.wgsl.ts) -----
// Represented importance of a packed node for shading point p in binary32,
// one rounding per operation; log2/exp2 are evaluated in binary64 and
// rounded once.
constexpr float kF32MinNormal = 1.175494351e-38f;
constexpr float kF32Max = 3.402823466e38f;
__device__ __forceinline__ float lt_log2(float x) { return (float)log2((double)x); }
__device__ __forceinline__ float lt_exp2(float x) { return (float)exp2((double)x); }
__device__ __forceinline__ float lt_dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float max3abs(F3 v) { return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fabsf(v.z)); }

__device__ float lt_dist2(F3 p, F3 lo, F3 hi) {
  const F3 h = f3(fmaxf(fmaxf(0.5f * lo.x - 0.5f * p.x, 0.0f), 0.5f * p.x - 0.5f * hi.x),
                  fmaxf(fmaxf(0.5f * lo.y - 0.5f * p.y, 0.0f), 0.5f * p.y - 0.5f * hi.y),
                  fmaxf(fmaxf(0.5f * lo.z - 0.5f * p.z, 0.0f), 0.5f * p.z - 0.5f * hi.z));
  const float hs = max3abs(h);
  if (!(hs >= kF32MinNormal)) return 0.0f;
  if (hs > 0.5f * kF32Max) return kF32Max;
  const F3 sc = f3(h.x / hs, h.y / hs, h.z / hs);
  const float sq = lt_dot(sc, sc);
  const float scale = 2.0f * hs;
  if (scale >= __builtin_sqrtf(kF32Max / sq)) return kF32Max;
  return fminf((scale * scale) * sq, kF32Max);
}

__device__ float lt_cone(F3 axis, float cos_o, float cos_oe, F3 p, F3 lo, F3 hi) {
  const float as = max3abs(axis);
  if (!(as > 0.0f) || as > kF32Max) return 1.0f;
  const F3 sa = f3(axis.x / as, axis.y / as, axis.z / as);
  const float al = __builtin_sqrtf(lt_dot(sa, sa));
  const F3 dv = f3((0.5f * p.x - 0.25f * lo.x) - 0.25f * hi.x, (0.5f * p.y - 0.25f * lo.y) - 0.25f * hi.y,
                   (0.5f * p.z - 0.25f * lo.z) - 0.25f * hi.z);
  const F3 rv = f3(0.25f * hi.x - 0.25f * lo.x, 0.25f * hi.y - 0.25f * lo.y, 0.25f * hi.z - 0.25f * lo.z);
  const float ds = max3abs(dv);
  if (!(ds >= kF32MinNormal) || ds > kF32Max) return 1.0f;
  const F3 sd = f3(dv.x / ds, dv.y / ds, dv.z / ds);
  const float sdl = __builtin_sqrtf(lt_dot(sd, sd));
  const float rs = max3abs(rv);
  float r_over_d = 0.0f;
  if (rs >= kF32MinNormal) {
    const F3 sr = f3(rv.x / rs, rv.y / rs, rv.z / rs);
    const float srl = __builtin_sqrtf(lt_dot(sr, sr));
    const float lr = ((lt_log2(rs) + lt_log2(srl)) - lt_log2(ds)) - lt_log2(sdl);
    if (lr >= 0.0f) return 1.0f;
    r_over_d = lt_exp2(fmaxf(lr, -126.0f));
  }
  const F3 d = f3(sd.x / sdl, sd.y / sdl, sd.z / sdl);
  const F3 a = f3(sa.x / al, sa.y / al, sa.z / al);
  const float cos_t = clampf(lt_dot(a, d), -1.0f, 1.0f);
  const float sin_u = clampf(r_over_d, 0.0f, 1.0f);
  const float cos_u = __builtin_sqrtf(fmaxf(0.0f, 1.0f - sin_u * sin_u));
  float cos_adj = 1.0f;
  if (cos_t < cos_u) {
    const float sin_t = __builtin_sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    cos_adj = clampf(cos_t * cos_u + sin_t * sin_u, -1.0f, 1.0f);
  }
  if (cos_adj < cos_oe) return 0.0f;
  if (cos_adj >= cos_o) return 1.0f;
  const float span = cos_o - cos_oe;
  if (!(span > 0.0f)) return 1.0f;
  return clampf((cos_adj - cos_oe) / span, kF32MinNormal, 1.0f);
}

__device__ float lt_importance(const float* node, F3 p) {
  const float power = node[1];
  if (!(power > 0.0f)) return 0.0f;
  const F3 lo = f3(node[4], node[5], node[6]), hi = f3(node[7], node[8], node[9]);
  const float d2 = fmaxf(fmaxf(lt_dist2(p, lo, hi), 1e-3f), kF32MinNormal);
  const float c = lt_cone(f3(node[10], node[11], node[12]), node[13], node[14], p, lo, hi);
  if (!(c > 0.0f)) return 0.0f;
  const float li = (lt_log2(power) - lt_log2(d2)) + lt_log2(fmaxf(c, kF32MinNormal));
  if (li >= 127.99999f) return kF32Max;
  if (li <= -126.0f) return kF32MinNormal;
  return clampf(lt_exp2(li), kF32MinNormal, kF32Max);
}

// sampleLightTree: one 24-bit root draw apportioned down the tree (every leaf
// keeps one bucket, the spare buckets follow the represented importances).
// Returns the light slot, or -1; *pdf = leaf buckets / 2^24.
__device__ int lt_sample(const float* tree, uint32_t nodes, F3 p, Rng& rng, float* pdf) {
  *pdf = 0.0f;
  uint32_t node = 0, cur = 0x01000000u, local = 0;
  if ((int)tree[2] >= 0 && (int)tree[3] >= 0) local = min(0x00FFFFFFu, rng_next(rng) >> 8);
  for (uint32_t guard = 0; guard < nodes + 1u; ++guard) {
    const float* b = tree + 16 * (size_t)node;
    const int l = (int)b[2], r = (int)b[3];
    if (l < 0 || r < 0) {
      *pdf = (float)cur / 16777216.0f;
      return (int)b[0];
    }
    const float* lb = tree + 16 * (size_t)l;
    const float* rb = tree + 16 * (size_t)r;
    const uint32_t ll = (uint32_t)lb[15], rl = (uint32_t)rb[15];
    const float li = lt_importance(lb, p), ri = lt_importance(rb, p);
    const uint32_t remaining = cur - (ll + rl);
    float p_left = 0.5f;
    const float s = fmaxf(li, ri);
    if (s > 0.0f) {
      const float lf = li / s, rf = ri / s;
      p_left = lf / (lf + rf);
    }
    const float rounded = floorf((float)remaining * p_left + 0.5f);
    const uint32_t left = ll + min(remaining, (uint32_t)fmaxf(0.0f, rounded));
    if (local < left) {
      cur = left;
      node = (uint32_t)l;
    } else {
      local -= left;
      cur -= left;
      node = (uint32_t)r;
    }
  }
  return -1;
}

__device__ __forceinline__ uint64_t mix64(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h * 0xBF58476D1CE4E5B9ull;
}

__device__ __forceinline__ float luminance(F3 c) { return (0.2126f * c.x + 0.7152f * c.y) + 0.0722f * c.z; }

// ---- BSDF (pt-webgpu/src/wgsl/pathTrace/bsdf.wgsl.ts, material.wgsl.ts) --

struct Mat {
  F3 base;
  float rough, metal;
  F3 emit;
  bool two_sided;
  // extension lobes (bsdf.wgsl.ts:392-808): clearcoat, sheen, anisotropy
  float cc, cc_rough, sheen, sheen_rough;
  F3 sheen_color;
  float aniso, aniso_c, aniso_s;  // strength, cos / sin of the rotation
  F3 spec_color;                  // KHR_materials_specular
  float spec_int;
};

__device__ Mat load_mat(const float* m) {
  Mat r;
  r.base = load3(m);
  r.rough = m[3];
  r.metal = m[4];
  r.emit = load3(m + 5);
  r.two_sided = m[8] > 0.5f;
  r.cc = m[9];
  r.cc_rough = m[10];
  r.sheen = m[11];
  r.sheen_rough = m[12];
  r.sheen_color = load3(m + 13);
  r.aniso = m[16];
  r.aniso_c = m[17];
  r.aniso_s = m[18];
  r.spec_color = load3(m + 19);
  r.spec_int = m[22];
  return r;
}

// materialSpecularF0 with ior 1.5 (bsdf.wgsl.ts:192-209): the IOR F0 tinted by
// specularColor (clamped to 1) and scaled by specularIntensity, mixed towards
// the base colour by metallic.
__device__ F3 spec_f0(const Mat& m) {
  const float ratio = (1.5f - 1.0f) / (1.5f + 1.0f);
  const float f = ratio * ratio;
  const float si = clampf(m.spec_int, 0.0f, 1.0f);
  const F3 d = f3(fminf(f * fmaxf(m.spec_color.x, 0.0f), 1.0f) * si, fminf(f * fmaxf(m.spec_color.y, 0.0f), 1.0f) * si,
                  fminf(f * fmaxf(m.spec_color.z, 0.0f), 1.0f) * si);
  const float mt = clampf(m.metal, 0.0f, 1.0f);
  return f3(d.x + (m.base.x - d.x) * mt, d.y + (m.base.y - d.y) * mt, d.z + (m.base.z - d.z) * mt);
}
// The Schlick F90 of materialSpecularFresnelSchlick: specularIntensity for
// dielectrics, 1 for metals (bsdf.wgsl.ts:211-222).
__device__ float spec_f90(const Mat& m) {
  const float d = clampf(m.spec_int, 0.0f, 1.0f);
  return d + (1.0f - d) * clampf(m.metal, 0.0f, 1.0f);
}
// materialSpecularFresnelSchlick, clamped to [0, 1] as the layered interface
// response clamps it (bsdf.wgsl.ts:283-292).
__device__ F3 schlick01(float c, F3 f0, float f90) {
  const float x = clampf(1.0f - c, 0.0f, 1.0f);
  const float x2 = x * x;
  const float x5 = x2 * x2 * x;
  return f3(clampf(f0.x + (f90 - f0.x) * x5, 0.0f, 1.0f), clampf(f0.y + (f90 - f0.y) * x5, 0.0f, 1.0f),
            clampf(f0.z + (f90 - f0.z) * x5, 0.0f, 1.0f));
}
// ggxD (material.wgsl.ts:2797-2804) at the half vector h:
// a^2 / (pi ((1 - cos^2) + cos^2 a^2)^2), with 1 - cos^2 taken as |n x h|^2.
// Near the lobe's peak 1 - cos^2 is smaller than a^2; from a binary32 cosine
// it would carry an absolute error of ~1e-7 (relative 1e-7 / a^2, 4e-4 for a
// coat of roughness 0.15), from the cross product its error shrinks with the
// angle.
__device__ float ggx_d(F3 n, F3 h, float alpha) {
  const float a2 = alpha * alpha;
  const float c = fmaxf(dot(n, h), 0.0f);
  const F3 x = cross(n, h);
  const float s2 = c > 0.0f ? fminf(dot(x, x), 1.0f) : 1.0f;
  const float d = s2 + (c * c) * a2;
  return a2 / (kPi * d * d);
}
__device__ float smith_g1(float ndv, float rough) {  // roughDielectric.ts:52-60
  const float c = fmaxf(ndv, 0.0f);
  if (c <= 1e-6f) return 0.0f;
  const float alpha = fmaxf(rough * rough, 1e-3f);
  const float tan2 = fmaxf(0.0f, 1.0f - c * c) / fmaxf(c * c, 1e-10f);
  return 2.0f / (1.0f + __builtin_sqrtf(1.0f + alpha * alpha * tan2));
}
__device__ float power_heuristic(float a, float b) {  // material.wgsl.ts:2808
  if (!(a >= 0.0f) || !(b >= 0.0f) || a > f32_max() || b > f32_max()) return 0.0f;
  const float s = fmaxf(a, b);
  if (!(s > 0.0f)) return 0.0f;
  const float x = a / s, y = b / s;
  const float x2 = x * x, y2 = y * y;
  return x2 / (x2 + y2);
}

__constant__ float kGgxE[64] = {
    0.1375f, 0.5617f, 0.7546f, 0.8522f, 0.9111f, 0.9505f, 0.9788f, 1.0f,
    0.2955f, 0.515f,  0.7091f, 0.8192f, 0.889f,  0.937f,  0.9721f, 0.9988f,
    0.5794f, 0.5541f, 0.6677f, 0.7691f, 0.8451f, 0.9021f, 0.9457f, 0.98f,
    0.7011f, 0.6486f, 0.6669f, 0.7199f, 0.7776f, 0.8305f, 0.8764f, 0.9155f,
    0.7335f, 0.6901f, 0.6696f, 0.6756f, 0.6972f, 0.7262f, 0.7578f, 0.7893f,
    0.7153f, 0.6712f, 0.6355f, 0.6145f, 0.6052f, 0.6045f, 0.6101f, 0.6199f,
    0.6669f, 0.6137f, 0.5657f, 0.5286f, 0.5f,    0.478f,  0.4611f, 0.4483f,
    0.6017f, 0.537f,  0.4773f, 0.4296f, 0.3905f, 0.358f,  0.3305f, 0.3069f};
__constant__ float kGgxEavg[8] = {0.9106f, 0.8931f, 0.8629f, 0.8094f, 0.725f, 0.6147f, 0.4931f, 0.3766f};

__device__ float lut_e(float c, float rough) {  // ggxDirectionalAlbedo (material.wgsl.ts:2869)
  const float mu = clampf(c, 0.0f, 1.0f), r = clampf(rough, 0.0f, 1.0f);
  const float fr = r * 7.0f, fm = mu * 7.0f;
  const uint32_t r0 = (uint32_t)floorf(fr), m0 = (uint32_t)floorf(fm);
  const uint32_t r1 = min(r0 + 1u, 7u), m1 = min(m0 + 1u, 7u);
  const float tr = fr - (float)r0, tm = fm - (float)m0;
  const float e00 = kGgxE[r0 * 8 + m0], e01 = kGgxE[r0 * 8 + m1];
  const float e10 = kGgxE[r1 * 8 + m0], e11 = kGgxE[r1 * 8 + m1];
  const float e0 = e00 + (e01 - e00) * tm, e1 = e10 + (e11 - e10) * tm;
  return clampf(e0 + (e1 - e0) * tr, 0.02f, 1.0f);
}
__device__ float lut_eavg(float rough) {
  const float r = clampf(rough, 0.0f, 1.0f), fr = r * 7.0f;
  const uint32_t r0 = (uint32_t)floorf(fr), r1 = min(r0 + 1u, 7u);
  const float tr = fr - (float)r0;
  return clampf(kGgxEavg[r0] + (kGgxEavg[r1] - kGgxEavg[r0]) * tr, 0.3f, 1.0f);
}
// ggxMultiscatterLobeRoughness (material.wgsl.ts:2906-2926): view and light
// lookups at their own roughness, the average albedo at a third.
__device__ F3 ggx_ms(F3 f0, float rough_v, float rough_l, float rough_avg, float ndv, float ndl) {
  const float eavg = lut_eavg(rough_avg), om = 1.0f - eavg;
  if (om < 1e-4f) return f3(0, 0, 0);
  const float eo = lut_e(ndv, rough_v), ei = lut_e(ndl, rough_l);
  const float shape = (1.0f - eo) * (1.0f - ei) / fmaxf(kPi * om, 1e-6f);
  const float c[3] = {f0.x, f0.y, f0.z};
  float o[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float favg = c[k] + (1.0f - c[k]) * (1.0f / 21.0f);
    const float fms = favg * favg * eavg / fmaxf(1.0f - favg * om, 1e-4f);
    o[k] = fms * shape;
  }
  return f3(o[0], o[1], o[2]);
}

// buildOnb (bsdf.wgsl.ts:1632-1640)
__device__ void build_onb(F3 n, F3& t, F3& b) {
  const F3 up = fabsf(n.y) > 0.999f ? f3(1, 0, 0) : f3(0, 1, 0);
  const F3 c = cross(up, n);
  const float l = __builtin_sqrtf(dot(c, c));
  t = f3(c.x / l, c.y / l, c.z / l);
  b = cross(n, t);
}
// The tangent frame rotated by anisotropyRotation (bsdf.wgsl.ts:2582-2595);
// the rotation's cos and sin come from the material table.
__device__ void aniso_frame(const Mat& m, F3 n, F3& t, F3& b) {
  F3 t0, b0;
  build_onb(n, t0, b0);
  t = m.aniso_c * t0 + m.aniso_s * b0;
  b = (-m.aniso_s) * t0 + m.aniso_c * b0;
}

__device__ float smoothstep_unit(float e0, float e1, float x) {
  const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

// ---- anisotropic GGX (bsdf.wgsl.ts:597-808) ----
// computeAnisotropicAxes (bsdf.wgsl.ts:1928-1933)
__device__ void aniso_axes(float alpha, float aniso, float& ax, float& ay) {
  const float aspect = __builtin_sqrtf(fmaxf(1.0f - 0.9f * aniso, 1e-4f));
  ax = fmaxf(alpha / aspect, 1e-4f);
  ay = fmaxf(alpha * aspect, 1e-4f);
}
__device__ float ggx_d_aniso(float ht, float hb, float hn, float ax, float ay) {
  const float d = (ht / ax) * (ht / ax) + (hb / ay) * (hb / ay) + hn * hn;
  return 1.0f / fmaxf(kPi * ax * ay * d * d, 1e-10f);
}
__device__ float smith_g1_aniso(float vt, float vb, float vn, float ax, float ay) {
  const float vn2 = fmaxf(vn * vn, 1e-10f);
  const float den = vn + __builtin_sqrtf(vn2 + (vt * ax) * (vt * ax) + (vb * ay) * (vb * ay));
  return (2.0f * vn) / fmaxf(den, 1e-6f);
}
// anisotropicProjectedRoughness / anisotropicAverageRoughness (bsdf.wgsl.ts:755-783)
__device__ float aniso_projected_rough(F3 dir, F3 t, F3 b, float rough, float aniso) {
  float ax, ay;
  aniso_axes(fmaxf(rough * rough, 1e-3f), aniso, ax, ay);
  const float dt = dot(dir, t), db = dot(dir, b);
  const float tl2 = dt * dt + db * db;
  const float blend = clampf(0.15f * aniso, 0.0f, 0.15f);
  float projected;
  if (tl2 <= 1e-6f) {
    projected = __builtin_sqrtf(clampf(0.5f * (ax + ay), 1e-4f, 1.0f));
  } else {
    const float eff = __builtin_sqrtf(((dt * ax) * (dt * ax) + (db * ay) * (db * ay)) / tl2);
    projected = __builtin_sqrtf(clampf(eff, 1e-4f, 1.0f));
  }
  return rough + (projected - rough) * blend;
}
__device__ float aniso_average_rough(float rough, float aniso) {
  float ax, ay;
  aniso_axes(fmaxf(rough * rough, 1e-3f), aniso, ax, ay);
  const float rms = __builtin_sqrtf(0.5f * (ax * ax + ay * ay));
  const float projected = __builtin_sqrtf(clampf(rms, 1e-4f, 1.0f));
  return rough + (projected - rough) * clampf(0.15f * aniso, 0.0f, 0.15f);
}
// anisotropicMultiscatterScale (bsdf.wgsl.ts:785-796)
__device__ float aniso_ms_scale(float aniso, float rough) {
  const float reduction = smoothstep_unit(0.0f, 0.35f, clampf(aniso, 0.0f, 1.0f));
  const float scale = 0.6f * smoothstep_unit(0.35f, 0.9f, rough);
  return 1.0f + (scale - 1.0f) * reduction;
}
// brdfAnisotropicSpecPdf (bsdf.wgsl.ts:727-753)
__device__ float aniso_spec_pdf(const Mat& m, F3 n, F3 t, F3 b, F3 wo, F3 wi) {
  const float ndv = fmaxf(dot(n, wo), 1e-6f), ndl = fmaxf(dot(n, wi), 0.0f);
  if (ndl <= 1e-5f) return 0.0f;
  const F3 h = safe_normalize(wo + wi);
  float ax, ay;
  aniso_axes(fmaxf(m.rough * m.rough, 1e-3f), m.aniso, ax, ay);
  const float d = ggx_d_aniso(dot(h, t), dot(h, b), fmaxf(dot(h, n), 0.0f), ax, ay);
  const float g1 = smith_g1_aniso(dot(wo, t), dot(wo, b), fmaxf(dot(wo, n), 1e-6f), ax, ay);
  return (d * g1) / fmaxf(4.0f * ndv, 1e-6f);
}

// ---- clearcoat and sheen layers (bsdf.wgsl.ts:398-596) ----
// clearcoatLayerWeight: the coat's fixed-IOR (F0 0.04) Fresnel at |V.N|.
__device__ float cc_weight(const Mat& m, F3 n, F3 wo) {
  const float x = clampf(1.0f - clampf(fabsf(dot(n, wo)), 0.0f, 1.0f), 0.0f, 1.0f);
  const float x2 = x * x;
  const float fcc = 0.04f + (1.0f - 0.04f) * (x2 * x2 * x);
  return clampf(m.cc, 0.0f, 1.0f) * fcc;
}
// evalClearcoatLobe (the kernel without N.L)
__device__ float cc_eval(const Mat& m, F3 n, F3 wo, F3 wi) {
  if (m.cc <= 0.0f) return 0.0f;
  const float ndl = fmaxf(dot(n, wi), 0.0f), ndv = fmaxf(dot(n, wo), 0.0f);
  if (ndl <= 1e-5f || ndv <= 1e-5f) return 0.0f;
  const F3 h = safe_normalize(wi + wo);
  const float alpha = fmaxf(m.cc_rough * m.cc_rough, 1e-3f);
  const float d = ggx_d(n, h, alpha);
  const float g = smith_g1(ndv, m.cc_rough) * smith_g1(ndl, m.cc_rough);
  return cc_weight(m, n, wo) * ((d * g) / fmaxf(4.0f * ndv * ndl, 1e-6f));
}
__device__ float cc_pdf(const Mat& m, F3 n, F3 wo, F3 wi) {  // clearcoatPdf
  if (m.cc <= 0.0f) return 0.0f;
  const float ndv = fmaxf(dot(n, wo), 0.0f);
  if (ndv <= 1e-5f) return 0.0f;
  const float ndl = fmaxf(dot(n, wi), 0.0f);
  if (ndl <= 1e-5f) return 0.0f;
  const F3 h = safe_normalize(wi + wo);
  const float alpha = fmaxf(m.cc_rough * m.cc_rough, 1e-3f);
  return (ggx_d(n, h, alpha) * smith_g1(ndv, m.cc_rough)) / fmaxf(4.0f * ndv, 1e-6f);
}
// charlieD (Estevez and Kulla 2017) at the half vector h; sin(theta_h) is
// |n x h| (see ggx_d), the power evaluated in binary64 and rounded once.
__device__ float charlie_d(F3 n, F3 h, float alpha) {
  const float inv = 1.0f / fmaxf(alpha, 1e-4f);
  const F3 x = cross(n, h);
  const float sin_h = fmaxf(dot(n, h), 0.0f) > 0.0f ? __builtin_sqrtf(fminf(dot(x, x), 1.0f)) : 1.0f;
  return (2.0f + inv) * (float)pow((double)sin_h, (double)inv) / (2.0f * kPi);
}
__device__ float sheen_vis(float ndl, float ndv) { return 1.0f / fmaxf(4.0f * (ndl + ndv - ndl * ndv), 1e-6f); }
// evalSheenLobe (the kernel without N.L)
__device__ F3 sheen_eval(const Mat& m, F3 n, F3 wo, F3 wi) {
  if (m.sheen <= 0.0f) return f3(0, 0, 0);
  const float ndl = fmaxf(dot(n, wi), 0.0f), ndv = fmaxf(dot(n, wo), 0.0f);
  if (ndl <= 1e-5f || ndv <= 1e-5f) return f3(0, 0, 0);
  const F3 h = safe_normalize(wi + wo);
  const float alpha = fmaxf(m.sheen_rough * m.sheen_rough, 1e-3f);
  const float k = charlie_d(n, h, alpha) * sheen_vis(ndl, ndv);
  return (m.sheen * m.sheen_color) * k;
}
// sheenDirectionalAlbedo (Estevez and Kulla 2017, section 5)
__device__ float sheen_albedo(float c_raw, float alpha) {
  const float c = 1.0f - clampf(c_raw, 0.0f, 1.0f);
  return 0.65584461f * (c * c * c) + 1.0f / (4.16526551f + expf(-7.97291361f * __builtin_sqrtf(alpha) + 6.33516894f));
}
__device__ float sheen_attenuation(const Mat& m, F3 n, F3 wo, F3 wi) {  // sheenBaseAttenuation
  if (m.sheen <= 0.0f) return 1.0f;
  const float ndl = clampf(fabsf(dot(n, wi)), 0.0f, 1.0f), ndv = clampf(fabsf(dot(n, wo)), 0.0f, 1.0f);
  const float a = fmaxf(m.sheen_rough, 0.07f), a2 = a * a;
  const float mc = fmaxf(fmaxf(m.sheen_color.x, m.sheen_color.y), m.sheen_color.z);
  const float full = fminf(1.0f - mc * sheen_albedo(ndv, a2), 1.0f - mc * sheen_albedo(ndl, a2));
  return clampf(1.0f + (full - 1.0f) * clampf(m.sheen, 0.0f, 1.0f), 0.0f, 1.0f);
}
__device__ float sheen_pdf(const Mat& m, F3 n, F3 wo, F3 wi) {  // charlieSheenPdf
  if (m.sheen <= 0.0f) return 0.0f;
  const float ndl = fmaxf(dot(n, wi), 0.0f), ndv = fmaxf(dot(n, wo), 0.0f);
  if (ndl <= 1e-5f || ndv <= 1e-5f) return 0.0f;
  const F3 h = safe_normalize(wi + wo);
  const float ndh = fmaxf(dot(n, h), 0.0f), vdh = fmaxf(dot(wo, h), 1e-6f);
  const float alpha = fmaxf(m.sheen_rough * m.sheen_rough, 1e-3f);
  return (charlie_d(n, h, alpha) * ndh) / fmaxf(4.0f * vdh, 1e-6f);
}

// evaluateBrdfFullWithClearcoatNormal (bsdf.wgsl.ts:809-916): the base lobes
// (Lambert, isotropic or anisotropic GGX, Kulla-Conty) under the sheen layer
// under the clearcoat layer.
__device__ F3 bsdf_eval(const Mat& m, F3 n, F3 wo, F3 wi) {
  const float ndl = fmaxf(dot(n, wi), 0.0f), ndv = fmaxf(dot(n, wo), 0.0f);
  if (ndl <= 1e-5f || ndv <= 1e-5f) return f3(0, 0, 0);
  const F3 h = safe_normalize(wi + wo);
  const float vdh = fmaxf(dot(wo, h), 0.0f);
  const F3 f0 = spec_f0(m);
  const F3 F = schlick01(vdh, f0, spec_f90(m));
  const float kdm = 1.0f - m.metal;
  const F3 diff = f3((1.0f - F.x) * kdm * m.base.x * kInvPi, (1.0f - F.y) * kdm * m.base.y * kInvPi,
                     (1.0f - F.z) * kdm * m.base.z * kInvPi);
  const F3 f0c = f3(clampf(f0.x, 0, 1), clampf(f0.y, 0, 1), clampf(f0.z, 0, 1));
  F3 spec, ms;
  if (m.aniso > 0.0f) {
    F3 t, b;
    aniso_frame(m, n, t, b);
    // evalBrdfSpecAnisotropic (bsdf.wgsl.ts:693-725)
    float ax, ay;
    aniso_axes(fmaxf(m.rough * m.rough, 1e-3f), m.aniso, ax, ay);
    const float d = ggx_d_aniso(dot(h, t), dot(h, b), fmaxf(dot(h, n), 0.0f), ax, ay);
    const float g = smith_g1_aniso(dot(wo, t), dot(wo, b), fmaxf(dot(wo, n), 1e-6f), ax, ay) *
                    smith_g1_aniso(dot(wi, t), dot(wi, b), fmaxf(dot(wi, n), 1e-6f), ax, ay);
    spec = (d * g) * F * (1.0f / fmaxf(4.0f * fmaxf(ndv, 1e-6f) * ndl, 1e-6f));
    const float ravg = aniso_average_rough(m.rough, m.aniso);
    ms = ggx_ms(f0c, aniso_projected_rough(wo, t, b, m.rough, m.aniso),
                aniso_projected_rough(wi, t, b, m.rough, m.aniso), ravg, ndv, ndl) *
         aniso_ms_scale(m.aniso, ravg);
  } else {
    const float alpha = fmaxf(m.rough * m.rough, 1e-3f);
    const float d = ggx_d(n, h, alpha);
    const float g = smith_g1(ndv, m.rough) * smith_g1(ndl, m.rough);
    spec = (d * g) * F * (1.0f / fmaxf(4.0f * ndv * ndl, 1e-6f));
    ms = ggx_ms(f0c, m.rough, m.rough, m.rough, ndv, ndl);
  }
  const F3 base = diff + spec + ms;
  if (m.cc <= 0.0f && m.sheen <= 0.0f) return base;
  const float cc_att = m.cc > 0.0f ? clampf(1.0f - cc_weight(m, n, wo), 0.0f, 1.0f) : 1.0f;
  const F3 layered = (base * sheen_attenuation(m, n, wo, wi) + sheen_eval(m, n, wo, wi)) * cc_att;
  const float cc = cc_eval(m, n, wo, wi);
  return f3(layered.x + cc, layered.y + cc, layered.z + cc);
}

// brdfFiniteBaseLobeWeights specular probability (bsdf.wgsl.ts:1172-1210).
__device__ float spec_prob(const Mat& m, float ndv) {
  const F3 F = schlick01(ndv, spec_f0(m), spec_f90(m));
  const float x = fmaxf(luminance(F), m.metal);
  return represented(clampf(0.04f + (0.96f - 0.04f) * x, 0.04f, 0.96f));
}

// brdfRepresentedExtensionLobeProbabilities (bsdf.wgsl.ts:1219-1262): joint
// P(base), P(clearcoat), P(sheen), and P(clearcoat | extension).
struct ExtProbs {
  float base, cc, sheen, cc_given_ext;
};
__device__ ExtProbs ext_probs(const Mat& m) {
  const float cc = fmaxf(m.cc, 0.0f), sh = fmaxf(m.sheen, 0.0f), ext = cc + sh;
  if (!(ext > 0.0f)) return ExtProbs{1.0f, 0.0f, 0.0f, 0.0f};
  const float total = 1.0f + ext;
  float pb, pe;
  if (ext <= 1.0f) {
    pe = represented(ext / total);
    pb = 1.0f - pe;
  } else {
    pb = represented(1.0f / total);
    pe = 1.0f - pb;
  }
  float pc;
  if (!(cc > 0.0f))
    pc = 0.0f;
  else if (!(sh > 0.0f))
    pc = 1.0f;
  else if (cc <= sh)
    pc = represented(cc / ext);
  else
    pc = 1.0f - represented(sh / ext);
  const float jc = pe * pc;
  return ExtProbs{pb, jc, pe - jc, pc};
}

// brdfDirectionalPdfFullWithClearcoatNormal (bsdf.wgsl.ts:1370-1489): the
// base lobes' density (brdfDirectionalPdfWithIridescence, :1555-1631, or its
// anisotropic VNDF form) and the extension lobes', by their joint
// probabilities.
__device__ float bsdf_pdf(const Mat& m, F3 n, F3 wo, F3 wi) {
  const float wdn = dot(n, wi), odn = dot(n, wo);
  const float ndv = fmaxf(odn, 0.0f);
  if (ndv <= 1e-5f) return 0.0f;
  const float sp = spec_prob(m, ndv), dp = 1.0f - sp;
  if (!(wdn * odn > 0.0f)) return 0.0f;
  const float ndl = fmaxf(wdn, 0.0f);
  if (ndl <= 1e-5f) return 0.0f;
  float pspec;
  if (m.aniso > 0.0f) {
    F3 t, b;
    aniso_frame(m, n, t, b);
    pspec = aniso_spec_pdf(m, n, t, b, wo, wi);
  } else {
    const F3 h = safe_normalize(wi + wo);
    const float alpha = fmaxf(m.rough * m.rough, 1e-3f);
    pspec = (ggx_d(n, h, alpha) * smith_g1(ndv, m.rough)) / fmaxf(4.0f * ndv, 1e-6f);
  }
  const float base = dp * (ndl * kInvPi) + sp * pspec;
  if (m.cc <= 0.0f && m.sheen <= 0.0f) return base;
  const ExtProbs e = ext_probs(m);
  return (e.base * base + e.cc * cc_pdf(m, n, wo, wi)) + e.sheen * sheen_pdf(m, n, wo, wi);
}

// sampleGgxVndfTangent (bsdf.wgsl.ts:1710-1735) and, with ax != ay,
// sampleGgxVndfAnisTangent (bsdf.wgsl.ts:612-637).
__device__ F3 vndf_sample(F3 wo, float ax, float ay, Rng& rng) {
  const F3 vh = safe_normalize(f3(ax * wo.x, ay * wo.y, wo.z));
  const float lensq = vh.x * vh.x + vh.y * vh.y;
  const float inv = lensq > 1e-10f ? 1.0f / __builtin_sqrtf(lensq) : 0.0f;
  const F3 t1 = lensq > 1e-10f ? f3(-vh.y * inv, vh.x * inv, 0.0f) : f3(1, 0, 0);
  const F3 t2 = cross(vh, t1);
  const float u1 = rand01(rng), u2 = rand01(rng);
  float a, b;
  disc_polar(u1, u2, a, b);
  const float s = 0.5f * (1.0f + vh.z);
  b = (1.0f - s) * __builtin_sqrtf(fmaxf(0.0f, 1.0f - a * a)) + s * b;
  const F3 nh = a * t1 + b * t2 + __builtin_sqrtf(fmaxf(0.0f, 1.0f - a * a - b * b)) * vh;
  return safe_normalize(f3(ax * nh.x, ay * nh.y, fmaxf(1e-6f, nh.z)));
}

// measureAreaVector (common.wgsl.ts:310-362): unit normal and area
// coefficient * |u x v| (0.5 for a triangle's edges, 4 for a rect's half
// axes, pi for a disc's).
__device__ bool area_measure(F3 u, F3 v, F3& normal, float& area, float coefficient = 0.5f) {
  const float es = fmaxf(fmaxf(fabsf(u.x), fmaxf(fabsf(u.y), fabsf(u.z))),
                         fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z))));
  if (!(es > 0.0f) || es > safe_max()) return false;
  const F3 av = cross(f3(u.x / es, u.y / es, u.z / es), f3(v.x / es, v.y / es, v.z / es));
  const float cs = fmaxf(fabsf(av.x), fmaxf(fabsf(av.y), fabsf(av.z)));
  if (!(cs > 0.0f) || cs > safe_max()) return false;
  const F3 ad = f3(av.x / cs, av.y / cs, av.z / cs);
  const float dl = __builtin_sqrtf(dot(ad, ad));
  normal = f3(ad.x / dl, ad.y / dl, ad.z / dl);
  float a = coefficient * (cs * dl);
  a = (a * es) * es;
  area = a;
  return a > 0.0f && a <= safe_max();
}
// ptAreaToSolidAnglePdf (material.wgsl.ts:242-275), area-measure form.
__device__ float area_to_solid_angle(float dist, float cosl, float area, float es) {
  if (!(dist > 0.0f) || !(cosl > 0.0f) || !(area > 0.0f)) return 0.0f;
  const float na = (area / es) / es;
  const float de = dist / es;
  const float pdf = (de * de) / (cosl * na);
  return (pdf > 0.0f && pdf <= f32_max()) ? pdf : 0.0f;
}
__device__ float edge_scale(F3 u, F3 v) {
  return fmaxf(fmaxf(fabsf(u.x), fmaxf(fabsf(u.y), fabsf(u.z))), fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z))));
}

// ---- environment (connect.wgsl.ts:126-219) -------------------------------

// The dome is rotated by rot about +Y: a world direction d looks up the
// unrotated map at RY(-rot) d, and a direction sampled in the map leaves as
// RY(+rot) d (connectCore.wgsl.ts:20-44).
__device__ __forceinline__ F3 env_to_map(const ShadeScene& sc, F3 d) {
  const float c = sc.env_rot_c, s = sc.env_rot_s;
  return f3(c * d.x - s * d.z, d.y, s * d.x + c * d.z);
}
__device__ __forceinline__ F3 env_to_world(const ShadeScene& sc, F3 d) {
  const float c = sc.env_rot_c, s = sc.env_rot_s;
  return f3(c * d.x + s * d.z, d.y, -s * d.x + c * d.z);
}

// The texel a world direction looks up: (scaled radiance, pdf per steradian).
__device__ float4 env_lookup(const ShadeScene& sc, F3 wd) {
  const F3 d = env_to_map(sc, wd);
  const float phi = atan2f(d.z, d.x);
  const float theta = acosf(clampf(d.y, -1.0f, 1.0f));
  float u = phi * kInv2Pi + 0.5f;
  u = u - floorf(u);
  const float v = clampf(theta * kInvPi, 0.0f, 0.999999f);
  const uint32_t x = min((uint32_t)floorf(u * (float)sc.env_w), sc.env_w - 1u);
  const uint32_t y = min((uint32_t)floorf(v * (float)sc.env_h), sc.env_h - 1u);
  return sc.env_texels[y * sc.env_w + x];
}

// A direction inside environment texel (x, y), uniform in phi and cos(theta)
// (environmentPacking.ts texel layout), evaluated in binary64 and rounded once
// per component, as the binary64 restatement rounds it.
__device__ F3 env_cell_dir(uint32_t x, uint32_t y, uint32_t w, uint32_t h, float xa, float xb) {
  const double kTwoPi = 6.283185307179586, kPiD = 3.141592653589793;
  const double phi0 = ((double)x / w - 0.5) * kTwoPi, phi1 = ((double)(x + 1u) / w - 0.5) * kTwoPi;
  const double th0 = (double)y / h * kPiD, th1 = (double)(y + 1u) / h * kPiD;
  const double phi = phi0 + (phi1 - phi0) * (double)xa;
  const double c0 = cos(th0), c1 = cos(th1);
  const double ct = c0 + (c1 - c0) * (double)xb;
  const double st = __builtin_sqrt(fmax(0.0, 1.0 - ct * ct));
  double sp, cp;
  sincos(phi, &sp, &cp);
  return f3((float)(cp * st), (float)ct, (float)(sp * st));
}

// ---- analytic emitters (kernel.wgsl.ts:903-918, 1198-1497) --------------------

// cos and sin of a binary32 angle, evaluated in binary64 and rounded once (the
// binary64 restatement rounds the same values).
__device__ __forceinline__ void cos_sin_f(float a, float& c, float& s) {
  double sd, cd;
  sincos((double)a, &sd, &cd);
  c = (float)cd;
  s = (float)sd;
}
// x^y evaluated in binary64 and rounded once.
__device__ __forceinline__ float pow_f(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __forceinline__ F3 normalize_f(F3 v) {
  const float l = __builtin_sqrtf(dot(v, v));
  return f3(v.x / l, v.y / l, v.z / l);
}
// ptDirectionalConeIsDelta (material.wgsl.ts:208-216)
__device__ bool cone_is_delta(float ad) {
  if (!(ad > 0.0f)) return true;
  float c, s;
  cos_sin_f(ad * 0.25f, c, s);
  return s < 7.666467e-20f;
}
// sampleDirectionalCone (kernel.wgsl.ts:903-918) with ptDirectionalConeSinCos
// (material.wgsl.ts:224-235): uniform in solid angle inside the cone.
__device__ F3 sample_directional_cone(F3 axis, float ad, Rng& rng) {
  F3 d = safe_normalize(axis);
  if (!cone_is_delta(ad)) {
    const float xi1 = rand01(rng), xi2 = rand01(rng);
    float cq, sq;
    cos_sin_f(ad * 0.25f, cq, sq);
    const float one_minus_cos_half = (2.0f * sq) * sq;
    const float q = (1.0f - clampf(xi1, 0.0f, 1.0f)) * one_minus_cos_half;
    const float sin_t = __builtin_sqrtf(fmaxf(q * (2.0f - q), 0.0f));
    const float cos_t = __builtin_sqrtf(fmaxf(1.0f - sin_t * sin_t, 0.0f));
    float cp, sp;
    cos_sin_f(6.28318530718f * xi2, cp, sp);
    const F3 tx = fabsf(d.x) > 0.9f ? f3(0, 1, 0) : f3(1, 0, 0);
    const F3 by = normalize_f(cross(d, tx));
    const F3 bx = cross(by, d);
    d = normalize_f(((sin_t * cp) * bx + (sin_t * sp) * by) + cos_t * d);
  }
  return d;
}
// pointSpotDistanceAttenuation (material.wgsl.ts:2176-2204)
__device__ float point_attenuation(float dist, float cutoff, float decay) {
  if (!(dist > 0.0f) || dist > f32_max()) return 0.0f;
  float att = 1.0f;
  if (decay > 0.01f) {
    const float den = pow_f(dist, decay);
    if (!(den > 0.0f)) att = f32_max();
    else if (den > f32_max()) return 0.0f;
    else att = fminf(1.0f / den, f32_max());
  }
  if (cutoff > 0.0f) att = att * clampf(1.0f - pow_f(dist / cutoff, 4.0f), 0.0f, 1.0f);
  return att;
}
__device__ float smoothstep_f(float e0, float e1, float x) {
  const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
// concentricDiscSample (kernelCore.wgsl.ts:73-85), Shirley & Chiu 1997
__device__ void concentric_disc(float a, float b, float& x, float& y) {
  if (a == 0.0f && b == 0.0f) {
    x = y = 0.0f;
    return;
  }
  float cr, cphi;
  if (fabsf(a) >= fabsf(b)) {
    cr = a;
    cphi = (kPi / 4.0f) * (b / a);
  } else {
    cr = b;
    cphi = kPi / 2.0f - (kPi / 4.0f) * (a / b);
  }
  float c, s;
  cos_sin_f(cphi, c, s);
  x = cr * c;
  y = cr * s;
}
// intersectRectAreaLightRay (connect.wgsl.ts:395-446) with
// solveAreaVectorCoordinates (common.wgsl.ts:370-412): distance and solid-
// angle pdf of the analytic rect or disc light `q` along a ray, or false.
__device__ bool intersect_rect_light(const float* q, F3 o, F3 d, float tmin, float& t_out, float& pdf_out) {
  const F3 c = load3(q), u = load3(q + 4), v = load3(q + 8);
  const bool disc = fabsf(q[15] - 1.0f) < 0.5f;
  F3 ln;
  float area;
  if (!area_measure(u, v, ln, area, disc ? kPi : 4.0f)) return false;
  const float denom = dot(ln, d);
  if (denom == 0.0f) return false;
  const float t = dot(ln, c - o) / denom;
  if (t <= tmin) return false;
  const F3 rel = (o + d * t) - c;
  const float es = edge_scale(u, v);
  const F3 su = f3(u.x / es, u.y / es, u.z / es), sv = f3(v.x / es, v.y / es, v.z / es);
  const F3 sr = f3(rel.x / es, rel.y / es, rel.z / es);
  const F3 an = f3(fabsf(ln.x), fabsf(ln.y), fabsf(ln.z));
  float det, sn, tn;
  if (an.x >= an.y && an.x >= an.z) {
    det = su.y * sv.z - su.z * sv.y;
    sn = sr.y * sv.z - sr.z * sv.y;
    tn = su.y * sr.z - su.z * sr.y;
  } else if (an.y >= an.z) {
    det = su.z * sv.x - su.x * sv.z;
    sn = sr.z * sv.x - sr.x * sv.z;
    tn = su.z * sr.x - su.x * sr.z;
  } else {
    det = su.x * sv.y - su.y * sv.x;
    sn = sr.x * sv.y - sr.y * sv.x;
    tn = su.x * sr.y - su.y * sr.x;
  }
  if (det == 0.0f) return false;
  const float uc = sn / det, vc = tn / det;
  if (!(fabsf(uc) <= f32_max()) || !(fabsf(vc) <= f32_max())) return false;
  const bool inside = disc ? (uc * uc + vc * vc <= 1.0f) : (fabsf(uc) <= 1.0f && fabsf(vc) <= 1.0f);
  if (!inside) return false;
  const float cl = fmaxf(dot(ln, f3(-d.x, -d.y, -d.z)), 0.0f);
  if (cl <= 0.0f) return false;
  const float pdf = area_to_solid_angle(t, cl, area, es);
  if (!(pdf > 0.0f)) return false;
  t_out = t;
  pdf_out = pdf;
  return true;
}

// ---- wave queue ------------------------------------------------------------

__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
  const uint64_t m = __ballot(want);
  if (m == 0ull) return kMiss;
  const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
  const uint32_t lane = __lane_id();
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, (int)leader);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64u - lane));
  return want ? base + (uint32_t)__popcll(m & below) : kMiss;
}

// Same, callable under divergent control flow: the lanes that reach the call
// take part, the first wanting lane does the atomic.
__device__ __forceinline__ uint32_t wave_append_active(uint32_t* counter, bool want) {
  const uint64_t m = __ballot(want);
  if (m == 0ull) return kMiss;
  const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
  const uint32_t lane = __lane_id();
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, (int)leader);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64u - lane));
  return want ? base + (uint32_t)__popcll(m & below) : kMiss;
}


// Cache input record of a surface vertex (layout in nrc.h).
// Raw cache-query record (vtn feature layout): the position normalised by the
// scene bounds, the direction towards the previous vertex, the shading normal,
// roughness and diffuse albedo (nrcQueryRadiance(xs, ns, -wi, rough, albedo),
// risGiNrc.wgsl.ts:477-480).
__device__ void nrc_features(const FrameParams& f, F3 hp, F3 wo, F3 nrm, const Mat& m, float* out) {
  out[0] = clampf((hp.x - f.scene_min[0]) * f.scene_inv_extent[0], 0.0f, 1.0f);
  out[1] = clampf((hp.y - f.scene_min[1]) * f.scene_inv_extent[1], 0.0f, 1.0f);
  out[2] = clampf((hp.z - f.scene_min[2]) * f.scene_inv_extent[2], 0.0f, 1.0f);
  out[3] = wo.x;
  out[4] = wo.y;
  out[5] = wo.z;
  out[6] = nrm.x;
  out[7] = nrm.y;
  out[8] = nrm.z;
  out[9] = m.rough;
  const float kd = 1.0f - m.metal;
  out[10] = m.base.x * kd;
  out[11] = m.base.y * kd;
  out[12] = m.base.z * kd;
  out[13] = 0.0f;
  out[14] = 0.0f;
  out[15] = 0.0f;
}

// ---- kernels -------------------------------------------------------------

// Pixel of self-training path p (a pseudo-random pixel per frame).
__device__ __forceinline__ uint32_t train_pixel(const FrameParams& f, uint32_t p) {
  uint32_t h = pcg_init(p - f.pixels, f.frame_index, f.frame_seed ^ 0x5BD1E995u);
  return pcg_next(h) % f.pixels;
}

// Sampler of a path (the state word `s` is set by the caller): Sobol for
// image paths in Sobol mode (pcgInit of common.wgsl.ts:130-150), else PCG.
__device__ __forceinline__ Rng path_rng(const FrameParams& f, uint32_t px, uint32_t py, bool train) {
  Rng r;
  r.s = 0;
  r.dirs = f.sobol_dirs;
  r.sample = f.sobol_key & 0xFFFFu;
  if (f.sobol && !train) {
    r.dim = 0;
    r.seed = vts::hash(vts::combine(vts::pixel_seed(px, py), f.sobol_key >> 16));
    r.tile = (px & 7u) | ((py & 7u) << 3);
  } else {
    r.dim = kPcgOnly;
    r.seed = r.tile = 0;
  }
  return r;
}

// generatePrimaryRay / unproject_ray_common (kernelCore.wgsl.ts:22-26; common.wgsl.ts:412-469)
// The frame's parameters with the per-frame words taken from device memory
// when the launch carries frame_dev.
__device__ __forceinline__ FrameParams frame_of(const FrameParams& f0) {
  FrameParams f = f0;
  if (f0.frame_dev) {
    f.frame_seed = f0.frame_dev[0];
    f.frame_index = f0.frame_dev[1];
    f.sobol_key = f0.frame_dev[2];
  }
  return f;
}

__global__ void camera_kernel(FrameParams f0, PathState s) {
  const FrameParams f = frame_of(f0);
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t total = f.pixels + (f.nrc ? f.train_paths : 0u);
  if (p >= total) return;
  uint32_t pix = p, key = f.frame_seed ^ f.frame_index;
  const bool train = p >= f.pixels;
  if (train) {
    // Self-training path: a pseudo-random pixel and its own PCG stream.
    pix = train_pixel(f, p);
    key ^= 0xA511E9B3u;
    s.train_n[p - f.pixels] = 0u;
  }
  const uint32_t px = pix % f.width, py = pix / f.width;
  Rng rng = path_rng(f, px, py, train);
  rng.s = train || !f.sobol ? pcg_init(px, py, key) : vts::fallback_init(px, py, f.sobol_key);
  const float jx = rand01(rng), jy = rand01(rng);
  const float u = ((float)px + jx) / (float)f.width, v = ((float)py + jy) / (float)f.height;
  const float nx = u * 2.0f - 1.0f, ny = 1.0f - v * 2.0f;
  const float* m = f.inv_vp;
  float fh[4], nh[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    fh[r] = ((m[r] * nx + m[4 + r] * ny) + m[8 + r] * 1.0f) + m[12 + r];
    nh[r] = ((m[r] * nx + m[4 + r] * ny) + m[8 + r] * -1.0f) + m[12 + r];
  }
  const float fs = fmaxf(fmaxf(fabsf(fh[0]), fabsf(fh[1])), fmaxf(fabsf(fh[2]), fabsf(fh[3])));
  const float ns = fmaxf(fmaxf(fabsf(nh[0]), fabsf(nh[1])), fmaxf(fabsf(nh[2]), fabsf(nh[3])));
  F3 o = f3(0, 0, 0), d = f3(0, 0, 0);
  bool ok = fs > 0.0f && fs <= safe_max() && ns > 0.0f && ns <= safe_max();
  if (ok) {
#pragma unroll
    for (int r = 0; r < 4; ++r) fh[r] = fh[r] / fs, nh[r] = nh[r] / ns;
    ok = nh[3] != 0.0f;
  }
  if (ok) {
    o = f3(nh[0] / nh[3], nh[1] / nh[3], nh[2] / nh[3]);
    const float orient = fh[3] != 0.0f ? (fh[3] * nh[3] > 0.0f ? 1.0f : (fh[3] * nh[3] < 0.0f ? -1.0f : 0.0f)) : 1.0f;
    const F3 num = f3((fh[0] * nh[3] - nh[0] * fh[3]) * orient, (fh[1] * nh[3] - nh[1] * fh[3]) * orient,
                      (fh[2] * nh[3] - nh[2] * fh[3]) * orient);
    d = safe_normalize(num);
  }
  s.rng[p] = rng.s;
  s.rng_dim[p] = rng.dim;
  store3(s.throughput + 3 * (size_t)p, f3(1, 1, 1));
  store3(s.conn + 3 * (size_t)p, f3(0, 0, 0));
  store3(s.radiance + 3 * (size_t)p, f3(0, 0, 0));
  s.prev_pdf[p] = 0.0f;
  s.flags[p] = ok ? (train ? 5u : 1u) : 0u;  // bit 0 alive, bit 2 training path
  s.signature[p] = 0x1234ull;
  s.queue[p] = p;
  s.rays[2 * (size_t)p] = make_float4(o.x, o.y, o.z, f.tmin);
  s.rays[2 * (size_t)p + 1] = make_float4(d.x, d.y, d.z, f32_max());
  if (p == 0) s.count[0] = total;
}

// One vertex of every queued path: MIS connection of the previous BSDF sample,
// emission, NEE (enqueues a shadow ray), BSDF sample (enqueues the next
// extension ray) and Russian roulette.
// Material sort (f.shade_sort).  Each workgroup shades its 256 queue entries
// in the order of a counting sort by a small key -- misses first, then hits
// grouped by material (ids past 13 share the last group), entries past the
// queue's end last -- so a wave's lanes mostly take the same branches and
// read the same material records.  The ranks come from per-wave ballots and
// a fixed wave order, so the order is deterministic; each path's result does
// not depend on it.  Off by default: with one uber-BSDF the branches barely
// diverge, and on the 1080p Cornell frames the sort cost 2-3% of the frame.
constexpr uint32_t kSortBins = 16;

__global__ __launch_bounds__(256) void shade_kernel(FrameParams f0, ShadeScene sc, PathState s,
                                                     uint32_t vertex) {
  const FrameParams f = frame_of(f0);
  __shared__ uint32_t s_count[4][kSortBins];
  __shared__ uint16_t s_order[256];
  const uint32_t n = s.count[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t li = threadIdx.x, wave = li >> 6, lane = li & 63u;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64u - lane));
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    if (f.shade_sort) {
      uint32_t bin = kSortBins - 1u;
      if (base + li < n) {
        const uint32_t* h = s.hits + 5 * (size_t)(base + li);
        bin = h[1] == kMiss ? 0u : 1u + min(sc.tri_material[h[0]], kSortBins - 3u);
      }
      uint32_t rank = 0;
#pragma unroll
      for (uint32_t b = 0; b < kSortBins; ++b) {
        const uint64_t m = __ballot(bin == b);
        if (bin == b) rank = (uint32_t)__popcll(m & below);
        if (lane == 0) s_count[wave][b] = (uint32_t)__popcll(m);
      }
      __syncthreads();
      uint32_t off = rank;
      for (uint32_t b = 0; b < bin; ++b) off += s_count[0][b] + s_count[1][b] + s_count[2][b] + s_count[3][b];
      for (uint32_t w = 0; w < wave; ++w) off += s_count[w][bin];
      s_order[off] = (uint16_t)li;
      __syncthreads();
    }
    const uint32_t i = base + (f.shade_sort ? (uint32_t)s_order[li] : li);
    bool want_shadow = false, want_ext = false;
    float4 shadow_a, shadow_b, ext_a, ext_b;
    F3 shadow_c = f3(0, 0, 0);
    uint32_t p = 0;
    if (i < n) {
      p = s.queue[i];
      const uint32_t* h = s.hits + 5 * (size_t)i;
      const uint32_t prim = h[0], inst = h[1];
      const float t_hit = __uint_as_float(h[2]), bu = __uint_as_float(h[3]), bv = __uint_as_float(h[4]);
      const bool miss = inst == kMiss;
      const float4 rb = s.rays[2 * (size_t)i + 1];
      const F3 dir = f3(rb.x, rb.y, rb.z);
      Rng rng;
      {
        const bool train = p >= f.pixels;
        const uint32_t pix = train ? train_pixel(f, p) : p;
        rng = path_rng(f, pix % f.width, pix / f.width, train);
        rng.s = s.rng[p];
        rng.dim = s.rng_dim[p];
      }
      F3 T = load3(s.throughput + 3 * (size_t)p);
      const F3 Cn = load3(s.conn + 3 * (size_t)p);
      F3 L = load3(s.radiance + 3 * (size_t)p);
      const float prev_pdf = s.prev_pdf[p];
      uint32_t flags = s.flags[p];
      uint64_t sig = s.signature[p];
      const bool prev_mis = prev_pdf > 0.0f;
      // 1. BSDF -> light / environment connection of the previous sample.
      Mat m{};
      const float* inst_rec = miss ? nullptr : sc.instances + (size_t)inst * kInstFloats;
      if (!miss) m = load_mat(sc.materials + (size_t)sc.tri_material[prim] * kMatFloats);
      if (prev_mis) {
        // The closest visible area light along the previous BSDF sample
        // (bsdfAreaLightConnectionContribution, connect.wgsl.ts:478-700):
        // analytic rect and disc lights are intersected here (they are not in
        // the BVH; one is visible when no surface lies before it, which the
        // sample's own closest hit tells), a mesh emitter is the hit itself.
        float best_t = f32_max(), best_pdf = 0.0f;
        F3 best_le = f3(0, 0, 0);
        if (sc.n_rect) {
          const float4 ra = s.rays[2 * (size_t)i];
          const F3 ro = f3(ra.x, ra.y, ra.z);
          const uint32_t r0 = sc.n_dir + sc.n_point + sc.n_spot;
          for (uint32_t k = 0; k < sc.n_rect; ++k) {
            const float* q = sc.lights + (size_t)(r0 + k) * kLightFloats;
            float tl, pl;
            if (intersect_rect_light(q, ro, dir, f.tmin, tl, pl)) {
              const bool visible = q[3] > 0.5f || miss || !(t_hit < fmaxf(tl - f.ray_bias * 2.0f, f.tmin));
              if (visible && tl < best_t) {
                best_t = tl;
                best_pdf = pl;
                best_le = load3(q + 12);
              }
            }
          }
        }
        if (!miss && (m.emit.x > 0.0f || m.emit.y > 0.0f || m.emit.z > 0.0f) && sc.emitter_count) {
          const uint32_t ebase = __float_as_uint(inst_rec[33]);
          if (ebase != kMiss) {
            const float* e = sc.emitters + (size_t)(ebase + sc.tri_source[prim]) * kEmitFloats;
            const F3 a = load3(e), b = load3(e + 3), c = load3(e + 6);
            F3 ln;
            float area;
            if (area_measure(b - a, c - a, ln, area)) {
              float cl = dot(ln, f3(-dir.x, -dir.y, -dir.z));
              cl = e[12] > 0.5f ? fabsf(cl) : fmaxf(cl, 0.0f);
              const float lpdf = area_to_solid_angle(t_hit, cl, area, edge_scale(b - a, c - a));
              if (lpdf > 0.0f && t_hit < best_t) {
                best_t = t_hit;
                best_pdf = lpdf;
                best_le = load3(e + 9);
              }
            }
          }
        }
        if (best_pdf > 0.0f) L = L + Cn * best_le * power_heuristic(prev_pdf, best_pdf);
        if (miss && sc.env_present) {
          const float4 et = env_lookup(sc, dir);
          L = L + Cn * f3(et.x, et.y, et.z) * power_heuristic(prev_pdf, et.w);
        }
      }
      const bool alive = (flags & 1u) != 0u;
      bool done = !alive || vertex >= f.bounce_limit;
      if (!done) sig = mix64(sig, miss ? 0xFFFFFFFFFFFFull : (((uint64_t)inst << 32) | prim));
      if (!done && miss) {
        if (!prev_mis && sc.env_present) {
          const float4 et = env_lookup(sc, dir);
          L = L + T * f3(et.x, et.y, et.z);
        }
        done = true;
        if (vertex == 0 && p < f.pixels) {
          *(float4*)(s.aux_nd + 4 * (size_t)p) = make_float4(0.5f, 1.0f, 0.5f, 0.0f);
          *(float4*)(s.aux_albedo + 4 * (size_t)p) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
      }
      if (!done) {
        // 2. Surface at the hit (shadePrologue.wgsl.ts:55-130).
        const uint4 tri = sc.indices[prim];
        const float4 n0 = sc.normals[tri.x], n1 = sc.normals[tri.y], n2 = sc.normals[tri.z];
        const float ba = (1.0f - bu) - bv;
        const F3 nl = safe_normalize(f3(n0.x, n0.y, n0.z) * ba + f3(n1.x, n1.y, n1.z) * bu +
                                     f3(n2.x, n2.y, n2.z) * bv);
        const float* w2l = inst_rec;
        const F3 nw = safe_normalize(f3(dot(load3(w2l), nl), dot(load3(w2l + 4), nl), dot(load3(w2l + 8), nl)));
        const float4 p0 = sc.positions[tri.x], p1 = sc.positions[tri.y], p2 = sc.positions[tri.z];
        const F3 ng = cross(f3(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z), f3(p2.x - p0.x, p2.y - p0.y, p2.z - p0.z));
        const F3 ld = f3(dot(f3(w2l[0], w2l[4], w2l[8]), dir), dot(f3(w2l[1], w2l[5], w2l[9]), dir),
                         dot(f3(w2l[2], w2l[6], w2l[10]), dir));
        const bool mirrored = inst_rec[32] > 0.5f;
        const bool front = (dot(ld, ng) < 0.0f) != mirrored;
        const F3 nrm = front ? nw : f3(-nw.x, -nw.y, -nw.z);
        const F3 hp = load3(s.hit_pos + 3 * (size_t)i);
        const F3 wo = f3(-dir.x, -dir.y, -dir.z);
        if (vertex == 0 && p < f.pixels) {
          // firstHit* (shadePrologue.wgsl.ts:173-179): normal, base colour and
          // the distance from the primary ray's origin.
          const float4 ra = s.rays[2 * (size_t)i];
          const F3 dv = hp - f3(ra.x, ra.y, ra.z);
          float4* nd = (float4*)(s.aux_nd + 4 * (size_t)p);
          *nd = make_float4(nrm.x * 0.5f + 0.5f, nrm.y * 0.5f + 0.5f, nrm.z * 0.5f + 0.5f,
                            __builtin_sqrtf(dot(dv, dv)));
          *(float4*)(s.aux_albedo + 4 * (size_t)p) = make_float4(m.base.x, m.base.y, m.base.z, 1.0f);
        }
        if (!prev_mis && (front || m.two_sided)) L = L + T * m.emit;
        // Neural radiance cache: an image path that has passed a
        // diffuse-dominant vertex ends here with a cache query (its
        // prediction times T is added after the bounce loop); a training path
        // records the vertex (features, T, radiance so far) for its target,
        // and at vertex f.train_vertices ends with a query of its own, so its
        // targets carry the cache's estimate of the truncated remainder
        // (self-training, Mueller et al. 2021, "Real-time Neural Radiance
        // Caching for Path Tracing", sec. 5).
        bool query = false;
        if (f.nrc) {
          const bool train = (flags & 4u) != 0u;
          query = train ? (f.train_vertices != 0u && vertex >= f.train_vertices) : (flags & 2u) != 0u;
          if (train && !query && vertex < kTrainSlots) {
            float* rec = s.train_rec + ((size_t)(p - f.pixels) * kTrainSlots + vertex) * kTrainRec;
            nrc_features(f, hp, wo, nrm, m, rec);
            store3(rec + 16, T);
            store3(rec + 19, L);
            s.train_n[p - f.pixels] = vertex + 1u;
          }
          const uint32_t qi = wave_append_active(s.nrc_count, query);
          if (query) {
            nrc_features(f, hp, wo, nrm, m, s.q_feat + 16 * (size_t)qi);
            store3(s.q_thr + 3 * (size_t)qi, T);
            s.q_path[qi] = p;
            flags = 0u;
          }
        }
        if (!query) {
        // 3. NEE over one light picked by the light tree, or uniformly when
        //    the tree is off or declines (mediumNee.wgsl.ts:14-38;
        //    kernel.wgsl.ts:1496-1690).
        const uint32_t n_ana = sc.n_dir + sc.n_point + sc.n_spot + sc.n_rect;
        const uint32_t n_light = n_ana + sc.emitter_count + (sc.env_present ? 1u : 0u);
        if (n_light > 0u) {
          uint32_t picked = 0xFFFFFFFFu;
          float inv_sel = 0.0f;
          if (sc.light_tree_nodes) {
            float pdf;
            const int e = lt_sample(sc.light_tree, sc.light_tree_nodes, hp, rng, &pdf);
            if (e >= 0 && pdf > 0.0f && (uint32_t)e < n_light) {
              picked = (uint32_t)e;
              inv_sel = 1.0f / pdf;
            }
          }
          if (picked == 0xFFFFFFFFu) {
            picked = rand_bounded(rng, n_light);
            inv_sel = (float)n_light;
          }
          sig = mix64(sig, 0xA000000000ull | picked);
          // A shadow ray carries the contribution; a light with castShadow
          // false adds it unoccluded.
          auto queue_shadow = [&](F3 contrib, F3 origin, F3 wi, float tmax, bool no_shadow) {
            if (no_shadow) {
              L = L + contrib;
              return;
            }
            shadow_c = contrib;
            shadow_a = make_float4(origin.x, origin.y, origin.z, f.tmin);
            shadow_b = make_float4(wi.x, wi.y, wi.z, tmax);
            want_shadow = true;
          };
          if (picked < n_ana) {
            const float* q = sc.lights + (size_t)picked * kLightFloats;
            const F3 neg_n = f3(-nrm.x, -nrm.y, -nrm.z);
            if (picked < sc.n_dir) {
              // directional: irradiance along a (cone-sampled) direction
              const F3 wi = sample_directional_cone(load3(q), q[3], rng);
              const F3 off = dot(nrm, wi) > 0.0f ? nrm : neg_n;
              const float ndl = fabsf(dot(nrm, wi));
              const F3 fr = bsdf_eval(m, nrm, wo, wi);
              queue_shadow(T * fr * load3(q + 4) * (ndl * inv_sel), hp + off * f.ray_bias, wi, f32_max(),
                           q[7] > 0.5f);
            } else if (picked < sc.n_dir + sc.n_point + sc.n_spot) {
              // point / spot: delta position, authored decay and cutoff
              const bool spot = picked >= sc.n_dir + sc.n_point;
              const float cutoff = q[spot ? 12 : 8], decay = q[spot ? 13 : 9];
              const F3 tl = load3(q) - hp;
              const float dist = safe_length(tl);
              if (dist > 0.0f && !(cutoff > 0.0f && dist > cutoff)) {
                const F3 wi = f3(tl.x / dist, tl.y / dist, tl.z / dist);
                float soft = 1.0f;
                bool lit = true;
                if (spot) {
                  const F3 axis = safe_normalize(load3(q + 4));
                  const float cos_outer = q[7], cos_inner = q[11];
                  const float cone_cos = dot(f3(-wi.x, -wi.y, -wi.z), axis);
                  lit = cone_cos >= cos_outer;
                  soft = smoothstep_f(cos_outer, fmaxf(cos_inner, cos_outer + 1e-6f), cone_cos);
                }
                if (lit) {
                  const F3 off = dot(nrm, wi) > 0.0f ? nrm : neg_n;
                  const float ndl = fabsf(dot(nrm, wi));
                  const F3 fr = bsdf_eval(m, nrm, wo, wi);
                  const float att = point_attenuation(dist, cutoff, decay);
                  const F3 rad = load3(q + (spot ? 8 : 4));
                  queue_shadow(T * fr * rad * (ndl * soft * att * inv_sel), hp + off * f.ray_bias, wi,
                               fmaxf(dist - f.ray_bias * 2.0f, f.tmin), q[3] > 0.5f);
                }
              }
            } else {
              // rect / disc area: a point on the emitter, area measure to
              // solid angle, MIS against the BSDF
              const F3 c = load3(q), u = load3(q + 4), v = load3(q + 8);
              const bool disc = fabsf(q[15] - 1.0f) < 0.5f;
              F3 ln;
              float area;
              const bool ok = area_measure(u, v, ln, area, disc ? kPi : 4.0f);
              const float xi1 = rand01(rng), xi2 = rand01(rng);
              F3 lp;
              if (disc) {
                float dx, dy;
                concentric_disc(xi1 * 2.0f - 1.0f, xi2 * 2.0f - 1.0f, dx, dy);
                lp = (c + u * dx) + v * dy;
              } else {
                lp = (c + u * (xi1 * 2.0f - 1.0f)) + v * (xi2 * 2.0f - 1.0f);
              }
              const F3 tl = lp - hp;
              const float dist = safe_length(tl);
              const F3 wi = safe_normalize(tl);
              const float ndl = fabsf(dot(nrm, wi));
              if (dist > 0.0f && ndl > 0.0f && ok) {
                const float cl = fmaxf(dot(ln, f3(-wi.x, -wi.y, -wi.z)), 0.0f);
                if (cl > 0.0f) {
                  const float lpdf = area_to_solid_angle(dist, cl, area, edge_scale(u, v));
                  if (lpdf > 0.0f) {
                    const F3 fr = bsdf_eval(m, nrm, wo, wi);
                    const float bpdf = bsdf_pdf(m, nrm, wo, wi);
                    const float w = power_heuristic(lpdf, bpdf);
                    const F3 off = dot(nrm, wi) > 0.0f ? nrm : neg_n;
                    queue_shadow(T * fr * load3(q + 12) * (ndl * w / lpdf * inv_sel), hp + off * f.ray_bias, wi,
                                 fmaxf(dist - f.ray_bias * 2.0f, f.tmin), q[3] > 0.5f);
                  }
                }
              }
            }
          } else if (picked - n_ana < sc.emitter_count) {
            const float* e = sc.emitters + (size_t)(picked - n_ana) * kEmitFloats;
            const F3 a = load3(e), b = load3(e + 3), c = load3(e + 6);
            const float r1 = rand01(rng), r2 = rand01(rng);
            const float su = __builtin_sqrtf(r1), uu = 1.0f - su, vv = r2 * su, ww = (1.0f - uu) - vv;
            const F3 lp = a * uu + b * vv + c * ww;
            const F3 tl = lp - hp;
            const float dist = safe_length(tl);
            F3 ln;
            float area;
            const bool ok = area_measure(b - a, c - a, ln, area);
            const F3 wi = safe_normalize(tl);
            const float ndl = fabsf(dot(nrm, wi));
            if (dist > 0.0f && ndl > 0.0f && ok) {
              const F3 fr = bsdf_eval(m, nrm, wo, wi);
              float cl = dot(ln, f3(-wi.x, -wi.y, -wi.z));
              cl = e[12] > 0.5f ? fabsf(cl) : fmaxf(cl, 0.0f);
              if (cl > 0.0f) {
                const float lpdf = area_to_solid_angle(dist, cl, area, edge_scale(b - a, c - a));
                if (lpdf > 0.0f) {
                  const float bpdf = bsdf_pdf(m, nrm, wo, wi);
                  const float w = power_heuristic(lpdf, bpdf);
                  const F3 off = dot(nrm, wi) > 0.0f ? nrm : f3(-nrm.x, -nrm.y, -nrm.z);
                  const F3 so = hp + off * f.ray_bias;
                  const float tmax = fmaxf(dist - f.ray_bias * 2.0f, f.tmin);
                  shadow_c = T * fr * load3(e + 9) * (ndl * w / lpdf * inv_sel);
                  shadow_a = make_float4(so.x, so.y, so.z, f.tmin);
                  shadow_b = make_float4(wi.x, wi.y, wi.z, tmax);
                  want_shadow = true;
                }
              }
            }
          } else {
            const float xi = rand01(rng);
            const uint32_t count = sc.env_w * sc.env_h;
            uint32_t lo = 0, hi = count;
            while (lo + 1u < hi) {
              const uint32_t mid = (lo + hi) >> 1;
              if (sc.env_cdf[mid] <= xi) lo = mid; else hi = mid;
            }
            const uint32_t idx = min(lo, count - 1u);
            const float xa = rand01(rng), xb = rand01(rng);
            const uint32_t ex = idx % sc.env_w, ey = idx / sc.env_w;
            const F3 wi = safe_normalize(env_to_world(sc, env_cell_dir(ex, ey, sc.env_w, sc.env_h, xa, xb)));
            const float4 et = sc.env_texels[idx];
            const float epdf = et.w;
            const F3 env = f3(et.x, et.y, et.z);
            const float ndl = fabsf(dot(nrm, wi));
            if (ndl > 0.0f) {
              const F3 off = dot(nrm, wi) > 0.0f ? nrm : f3(-nrm.x, -nrm.y, -nrm.z);
              const F3 so = hp + off * f.ray_bias;
              const F3 fr = bsdf_eval(m, nrm, wo, wi);
              const float bpdf = bsdf_pdf(m, nrm, wo, wi);
              const float w = power_heuristic(epdf, bpdf);
              shadow_c = T * fr * env * (ndl * w / fmaxf(epdf, 1e-8f) * inv_sel);
              shadow_a = make_float4(so.x, so.y, so.z, f.tmin);
              shadow_b = make_float4(wi.x, wi.y, wi.z, f32_max());
              want_shadow = true;
            }
          }
        }
        // 4. BSDF sample, opaque branch (bsdf.wgsl.ts:2920-2997) and the finite
        //    finalizer (bsdf.wgsl.ts:1863-1926).
        const float ndv = fmaxf(dot(nrm, wo), 0.0f);
        const float sp = spec_prob(m, ndv);
        const ExtProbs ep = ext_probs(m);
        F3 tb, bb;
        if (m.aniso > 0.0f)
          aniso_frame(m, nrm, tb, bb);
        else
          build_onb(nrm, tb, bb);
        const F3 inc = f3(-wo.x, -wo.y, -wo.z);
        F3 wi;
        uint32_t lobe;  // 0 diffuse, 1 specular, 2 clearcoat, 3 sheen
        if (rand01(rng) < ep.base) {
          lobe = rand01(rng) < sp ? 1u : 0u;
          if (lobe == 1u) {
            // glossyReflectionSample / its anisotropic form (bsdf.wgsl.ts:640-691, 1752-1778)
            const float alpha = fmaxf(m.rough * m.rough, 1e-3f);
            float ax = alpha, ay = alpha;
            if (m.aniso > 0.0f) aniso_axes(alpha, m.aniso, ax, ay);
            const F3 wl = f3(dot(wo, tb), dot(wo, bb), dot(wo, nrm));
            const F3 hl = vndf_sample(wl, ax, ay, rng);
            const F3 hw = safe_normalize(hl.x * tb + hl.y * bb + hl.z * nrm);
            wi = safe_normalize(inc - hw * (2.0f * dot(hw, inc)));
          } else {
            // cosineHemisphereSample (bsdf.wgsl.ts:1648-1664), its own frame
            F3 tc, bc;
            build_onb(nrm, tc, bc);
            const float u1 = rand01(rng), u2 = rand01(rng);
            float a, b;
            disc_polar(u1, u2, a, b);
            const float ct = __builtin_sqrtf(fmaxf(0.0f, 1.0f - u1));
            wi = safe_normalize(a * tc + b * bc + ct * nrm);
          }
        } else if (rand01(rng) < ep.cc_given_ext) {
          // the clearcoat lobe: an isotropic VNDF sample at the coat roughness
          lobe = 2u;
          F3 tc, bc;
          build_onb(nrm, tc, bc);
          const float alpha = fmaxf(m.cc_rough * m.cc_rough, 1e-3f);
          const F3 wl = f3(dot(wo, tc), dot(wo, bc), dot(wo, nrm));
          const F3 hl = vndf_sample(wl, alpha, alpha, rng);
          const F3 hw = safe_normalize(hl.x * tc + hl.y * bc + hl.z * nrm);
          wi = safe_normalize(inc - hw * (2.0f * dot(hw, inc)));
        } else {
          // charlieSheenSample (bsdf.wgsl.ts:1665-1700): sin(theta_h) = u1^(1/(1/alpha + 2))
          lobe = 3u;
          const float u1 = rand01(rng), u2 = rand01(rng);
          const float alpha = fmaxf(m.sheen_rough * m.sheen_rough, 1e-3f);
          const float inv = 1.0f / fmaxf(alpha, 1e-4f);
          const float sin_h = (float)pow((double)u1, 1.0 / (double)(inv + 2.0f));
          const float cos_h = __builtin_sqrtf(fmaxf(0.0f, 1.0f - sin_h * sin_h));
          float c, s;
          disc_polar(1.0f, u2, c, s);
          const F3 hw = safe_normalize((sin_h * c) * tb + (sin_h * s) * bb + cos_h * nrm);
          wi = safe_normalize(inc - hw * (2.0f * dot(hw, inc)));
        }
        sig = mix64(sig, 0xB0ull + lobe);
        const float mpdf = bsdf_pdf(m, nrm, wo, wi);
        const F3 fr = bsdf_eval(m, nrm, wo, wi);
        const float cosine = fabsf(dot(nrm, wi));
        const bool allows = mpdf > 0.0f && cosine > 1e-8f;
        const F3 C2 = allows ? T * (fr * (cosine / mpdf)) : f3(0, 0, 0);
        // 5. Russian roulette after the connection (kernel.wgsl.ts:2009-2013).
        //    A rejected sample leaves zero throughput: the reference keeps
        //    tracing that path, but all it adds later is multiplied by zero,
        //    so it ends here.
        bool lives = allows;
        float rr_mul = 1.0f;
        if (allows && vertex > 2u) {
          const float surv = represented(clampf(fmaxf(C2.x, fmaxf(C2.y, C2.z)), 0.1f, 0.95f));
          lives = rand01(rng) < surv;
          sig = mix64(sig, lives ? 0xC1ull : 0xC0ull);
          rr_mul = 1.0f / surv;
        }
        T = C2 * rr_mul;
        store3(s.conn + 3 * (size_t)p, C2);
        s.prev_pdf[p] = allows ? mpdf : 0.0f;
        // bit 0 alive, bit 1 a diffuse-dominant vertex has been passed, bit 2 training path
        flags = (flags & 4u) | (lives ? 1u : 0u) | (((flags & 2u) != 0u || sp < 0.5f) ? 2u : 0u);
        if (allows) {
          const F3 no = hp + nrm * f.ray_bias;
          ext_a = make_float4(no.x, no.y, no.z, f.tmin);
          ext_b = make_float4(wi.x, wi.y, wi.z, f32_max());
          want_ext = true;
        }
        }  // !query
      }
      s.rng[p] = rng.s;
      s.rng_dim[p] = rng.dim;
      store3(s.throughput + 3 * (size_t)p, T);
      store3(s.radiance + 3 * (size_t)p, L);
      s.flags[p] = flags;
      s.signature[p] = sig;
    }
    // Wave-aggregated queue appends (all lanes of the wave reach these).
    const uint32_t si = wave_append(s.count + 2, want_shadow);
    if (want_shadow) {
      s.shadow_path[si] = p;
      s.shadow_rays[2 * (size_t)si] = shadow_a;
      s.shadow_rays[2 * (size_t)si + 1] = shadow_b;
      store3(s.shadow_contrib + 3 * (size_t)si, shadow_c);
    }
    const uint32_t ei = wave_append(s.count + 1, want_ext);
    if (want_ext) {
      s.next_queue[ei] = p;
      s.next_rays[2 * (size_t)ei] = ext_a;
      s.next_rays[2 * (size_t)ei + 1] = ext_b;
    }
  }
}

__global__ void shadow_resolve_kernel(PathState s) {
  const uint32_t n = s.count[2];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (s.shadow_occ[i]) continue;
    const uint32_t p = s.shadow_path[i];
    float* L = s.radiance + 3 * (size_t)p;
    const float* c = s.shadow_contrib + 3 * (size_t)i;
    L[0] = L[0] + c[0];
    L[1] = L[1] + c[1];
    L[2] = L[2] + c[2];
  }
}

__global__ void nrc_apply_kernel(PathState s) {
  const uint32_t n = s.nrc_count[0];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t p = s.q_path[i];
    const F3 t = load3(s.q_thr + 3 * (size_t)i), y = load3(s.q_rgb + 3 * (size_t)i);
    const F3 c = f3(fmaxf(y.x, 0.0f), fmaxf(y.y, 0.0f), fmaxf(y.z, 0.0f));
    store3(s.radiance + 3 * (size_t)p, load3(s.radiance + 3 * (size_t)p) + t * c);
  }
}

// Target of a recorded vertex: (L_final - L_at_vertex) / T_at_vertex, the
// radiance scattered at that vertex toward the previous one.
__global__ void nrc_targets_kernel(FrameParams f, PathState s) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  bool any = false;
  uint32_t nrec = 0;
  F3 L = f3(0, 0, 0);
  if (t < f.train_paths) {
    nrec = s.train_n[t];
    L = load3(s.radiance + 3 * (size_t)(f.pixels + t));
    any = nrec > 0u;
  }
  for (uint32_t k = 0; k < kTrainSlots; ++k) {
    bool want = false;
    const float* rec = nullptr;
    F3 tg = f3(0, 0, 0);
    if (any && k < nrec) {
      rec = s.train_rec + ((size_t)t * kTrainSlots + k) * kTrainRec;
      const F3 T = load3(rec + 16), Lr = load3(rec + 19);
      if (T.x > 0.0f && T.y > 0.0f && T.z > 0.0f) {
        tg = f3(fmaxf(L.x - Lr.x, 0.0f) / T.x, fmaxf(L.y - Lr.y, 0.0f) / T.y, fmaxf(L.z - Lr.z, 0.0f) / T.z);
        want = tg.x <= 3.0e38f && tg.y <= 3.0e38f && tg.z <= 3.0e38f;
      }
    }
    const uint32_t j = wave_append(s.nrc_count + 1, want);
    if (want) {
      float* out = s.train_feat + 16 * (size_t)j;
      for (uint32_t q = 0; q < 16; ++q) out[q] = rec[q];
      store3(s.train_tgt + 3 * (size_t)j, tg);
    }
  }
}

__global__ void accumulate_kernel(FrameParams f, PathState s) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= f.width * f.height) return;
  const float* L = s.radiance + 3 * (size_t)p;
  float* a = s.accum + 4 * (size_t)p;
  const F3 c = f3(fmaxf(L[0], 0.0f), fmaxf(L[1], 0.0f), fmaxf(L[2], 0.0f));
  a[0] = a[0] + c.x;
  a[1] = a[1] + c.y;
  a[2] = a[2] + c.z;
  // variance moments (accumulateFrame, kernel.wgsl.ts:936-940)
  const float l = luminance(c);
  float* mo = s.moments + 2 * (size_t)p;
  mo[0] = mo[0] + l;
  mo[1] = mo[1] + l * l;
}

// ---- present pass (present.wgsl.ts; tonemap.wgsl.ts) -----------------------

__device__ float tm_exposure(float c, float exposure) {  // vt_safeExposureScalar
  if (c != c || !(exposure > 0.0f)) return 0.0f;
  const float fmax = 3.402823466e38f;
  const float mag = fminf(fabsf(c) * fminf(exposure, fmax), fmax);
  return c < 0.0f ? -mag : mag;
}
__device__ float tm_aces(float x) {  // Narkowicz 2015, numerator and denominator over v^2
  const float v = fmaxf(x, 0.0f);
  const float inv = 1.0f / fmaxf(v, 1e-20f);
  const float curved = (2.51f + 0.03f * inv) / ((2.43f + 0.59f * inv) + (0.14f * inv) * inv);
  return v == 0.0f ? 0.0f : fminf(fmaxf(curved, 0.0f), 1.0f);
}
__device__ float tm_agx(float c) {
  const float v = fmaxf(c, 1e-6f);
  const float lx = fminf(fmaxf((log2f(v) + 12.47393f) / (12.47393f + 4.026069f), 0.0f), 1.0f);
  const float x2 = lx * lx, x4 = x2 * x2;
  const float y = (((((15.5f * x4 * x2 - 40.14f * x4 * lx) + 31.96f * x4) - 6.868f * x2 * lx) + 0.4298f * x2) +
                   0.1191f * lx) - 0.00232f;
  return fminf(fmaxf(y, 0.0f), 1.0f);
}
__device__ float tm_apply(float c, uint32_t mode, float exposure) {  // vitrumTonemap
  const float x = tm_exposure(c, exposure);
  if (mode == 1u) return tm_agx(x);
  if (mode == 2u) {
    const float v = fmaxf(x, 0.0f);
    return v / (1.0f + v);
  }
  if (mode == 3u) return fminf(fmaxf(x, 0.0f), 1.0f);
  if (mode == 4u) return x;
  return tm_aces(x);
}
__device__ float srgb_oetf(float c) {  // vt_linearToSrgb (IEC 61966-2-1)
  const float v = fmaxf(c, 0.0f);
  return v <= 0.0031308f ? v * 12.92f : 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
}

__global__ void present_kernel(const float* accum, uint32_t pixels, float inv_samples, uint32_t mode,
                               float exposure, uint32_t linear_out, float* out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pixels) return;
  const float4 a = *(const float4*)(accum + 4 * (size_t)p);
  const float hdr[3] = {a.x * inv_samples, a.y * inv_samples, a.z * inv_samples};
  float o[3];
  for (int k = 0; k < 3; ++k) {
    float t = tm_apply(fmaxf(hdr[k], 0.0f), mode, exposure);
    if (!linear_out) t = srgb_oetf(t);
    t = fminf(fmaxf(t, 0.0f), 65504.0f);
    o[k] = __half2float(__float2half_rn(t));  // the rgba16float store
  }
  *(float4*)(out + 4 * (size_t)p) = make_float4(o[0], o[1], o[2], 1.0f);
}

__global__ void variance_kernel(const float* moments, uint32_t pixels, float samples, float* out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pixels) return;
  const float count = fmaxf(samples, 1.0f);
  const float mean = moments[2 * (size_t)p] / count;
  const float var = fmaxf(0.0f, moments[2 * (size_t)p + 1] / count - mean * mean);
  *(float4*)(out + 4 * (size_t)p) = make_float4(var, 0.0f, 0.0f, 0.0f);
}

// Promotes the next extension queue to the current one and clears the
// shadow count, on the device so a frame needs no host round trip.
// which & 1: extension counts (next queue becomes current); which & 2:
// shadow count.
__global__ void rotate_counts_kernel(uint32_t* count, uint32_t which) {
  if (threadIdx.x == 0) {
    if (which & 1u) {
      count[0] = count[1];
      count[1] = 0u;
    }
    if (which & 2u) count[2] = 0u;
  }
}

// ---- batched frames: one traversal launch over the queues of several engines ----

// offsets[k] = the rays of members 0..k-1 in queue word `word` (exclusive
// prefix; offsets[count] = the total).  One thread: a batch has few members.
__global__ void batch_prefix_kernel(const BatchMember* __restrict__ m, uint32_t count, uint32_t word,
                                    uint32_t* __restrict__ offsets) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    uint32_t total = 0;
    for (uint32_t k = 0; k < count; ++k) {
      offsets[k] = total;
      total += m[k].count[word];
    }
    offsets[count] = total;
  }
}

__device__ __forceinline__ uint32_t batch_member_of(const uint32_t* offsets, uint32_t count, uint32_t i) {
  uint32_t lo = 0, hi = count;
  while (lo + 1u < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// Copies each member's extension (shadow = 0, from rays[cur]) or shadow rays
// into the batch's ray array, member k's at offsets[k].
__global__ void batch_gather_kernel(const BatchMember* __restrict__ m, uint32_t count,
                                    const uint32_t* __restrict__ offsets, uint32_t shadow, uint32_t cur,
                                    float4* __restrict__ out) {
  const uint32_t total = offsets[count];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const uint32_t k = batch_member_of(offsets, count, i), li = i - offsets[k];
    const float4* src = shadow ? m[k].shadow_rays : m[k].rays[cur];
    out[2 * (size_t)i] = src[2 * (size_t)li];
    out[2 * (size_t)i + 1] = src[2 * (size_t)li + 1];
  }
}

// Copies the batch's closest-hit records and hit positions (shadow = 0) or
// occlusion words (shadow = 1) back to each member.
__global__ void batch_scatter_kernel(const BatchMember* __restrict__ m, uint32_t count,
                                     const uint32_t* __restrict__ offsets, uint32_t shadow,
                                     const uint32_t* __restrict__ words, const float* __restrict__ pos) {
  const uint32_t total = offsets[count];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const uint32_t k = batch_member_of(offsets, count, i), li = i - offsets[k];
    if (shadow) {
      m[k].shadow_occ[li] = words[i];
    } else {
#pragma unroll
      for (uint32_t w = 0; w < 5; ++w) m[k].hits[5 * (size_t)li + w] = words[5 * (size_t)i + w];
#pragma unroll
      for (uint32_t c = 0; c < 3; ++c) m[k].hit_pos[3 * (size_t)li + c] = pos[3 * (size_t)i + c];
    }
  }
}

__global__ void zero_f32_kernel(float* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

hipError_t launch_camera(const FrameParams& f, const PathState& s, hipStream_t st) {
  const uint32_t n = f.pixels + (f.nrc ? f.train_paths : 0u);
  hipLaunchKernelGGL(camera_kernel, dim3((n + 255) / 256), dim3(256), 0, st, f, s);
  return hipGetLastError();
}

hipError_t launch_shade(const FrameParams& f, const ShadeScene& sc, const PathState& s,
                        uint32_t vertex, uint32_t capacity, hipStream_t st) {
  const uint32_t blocks = (capacity + 255) / 256;
  hipLaunchKernelGGL(shade_kernel, dim3(blocks < 8192 ? blocks : 8192), dim3(256), 0, st, f, sc, s, vertex);
  return hipGetLastError();
}

hipError_t launch_shadow_resolve(const PathState& s, uint32_t capacity, hipStream_t st) {
  const uint32_t blocks = (capacity + 255) / 256;
  hipLaunchKernelGGL(shadow_resolve_kernel, dim3(blocks < 4096 ? blocks : 4096), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_nrc_apply(const PathState& s, uint32_t capacity, hipStream_t st) {
  const uint32_t blocks = (capacity + 255) / 256;
  hipLaunchKernelGGL(nrc_apply_kernel, dim3(blocks < 4096 ? (blocks ? blocks : 1) : 4096), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_nrc_targets(const FrameParams& f, const PathState& s, hipStream_t st) {
  if (f.train_paths == 0) return hipSuccess;
  hipLaunchKernelGGL(nrc_targets_kernel, dim3((f.train_paths + 255) / 256), dim3(256), 0, st, f, s);
  return hipGetLastError();
}

hipError_t launch_accumulate(const FrameParams& f, const PathState& s, hipStream_t st) {
  const uint32_t n = f.width * f.height;
  hipLaunchKernelGGL(accumulate_kernel, dim3((n + 255) / 256), dim3(256), 0, st, f, s);
  return hipGetLastError();
}

hipError_t launch_present(const float* accum, uint32_t pixels, uint32_t samples, uint32_t mode, float exposure,
                          uint32_t linear_out, float* out, hipStream_t st) {
  if (pixels == 0) return hipSuccess;
  // accum.xyz / max(accum.w, 1) of the reference's running sum
  const float inv = 1.0f / (float)(samples > 0 ? samples : 1u);
  hipLaunchKernelGGL(present_kernel, dim3((pixels + 255) / 256), dim3(256), 0, st, accum, pixels, inv, mode,
                     exposure, linear_out, out);
  return hipGetLastError();
}

hipError_t launch_variance(const float* moments, uint32_t pixels, uint32_t samples, float* out, hipStream_t st) {
  if (pixels == 0) return hipSuccess;
  hipLaunchKernelGGL(variance_kernel, dim3((pixels + 255) / 256), dim3(256), 0, st, moments, pixels,
                     (float)samples, out);
  return hipGetLastError();
}

hipError_t launch_batch_prefix(const BatchMember* m, uint32_t count, uint32_t word, uint32_t* offsets,
                               hipStream_t st) {
  hipLaunchKernelGGL(batch_prefix_kernel, dim3(1), dim3(64), 0, st, m, count, word, offsets);
  return hipGetLastError();
}

hipError_t launch_batch_gather(const BatchMember* m, uint32_t count, const uint32_t* offsets, uint32_t shadow,
                               uint32_t cur, float4* out, uint32_t capacity, hipStream_t st) {
  const uint32_t want = (capacity + 255u) / 256u, blocks = want < 4096u ? want : 4096u;
  hipLaunchKernelGGL(batch_gather_kernel, dim3(blocks ? blocks : 1u), dim3(256), 0, st, m, count, offsets, shadow,
                     cur, out);
  return hipGetLastError();
}

hipError_t launch_batch_scatter(const BatchMember* m, uint32_t count, const uint32_t* offsets, uint32_t shadow,
                                const uint32_t* words, const float* pos, uint32_t capacity, hipStream_t st) {
  const uint32_t want = (capacity + 255u) / 256u, blocks = want < 4096u ? want : 4096u;
  hipLaunchKernelGGL(batch_scatter_kernel, dim3(blocks ? blocks : 1u), dim3(256), 0, st, m, count, offsets, shadow,
                     words, pos);
  return hipGetLastError();
}

hipError_t launch_reset_counts(const PathState& s, uint32_t which, hipStream_t st) {
  hipLaunchKernelGGL(rotate_counts_kernel, dim3(1), dim3(64), 0, st, s.count, which);
  return hipGetLastError();
}

hipError_t launch_zero_f32(float* p, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(zero_f32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace vtr
