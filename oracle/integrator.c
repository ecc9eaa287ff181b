/*
 * integrator.c -- binary64 mirror of pt-webgpu's unidirectional path tracer
 * for opaque Lambert + GGX materials, mesh-area emitters and a constant
 * environment.  TEST INFRASTRUCTURE ONLY (see vt_oracle.h).
 *
 * Follows packages/pt-webgpu/src/wgsl/pathTrace/kernel.wgsl.ts:962-2032 (the
 * `main` megakernel) with the RNG consumed in the same order:
 *   pcgInit; 2 jitter draws; per bounce: [light pick: ptRandBoundedU32]
 *   [mesh light: 2 draws | environment: 1 CDF draw + 2 cell draws]
 *   [BSDF: extension-lobe draw, specular/diffuse draw, 2 direction draws]
 *   [RR draw when bounce > 2].
 * Shading arithmetic is binary64; ray/scene intersection uses the f32
 * traversal of traverse.c on the f32-rounded rays (the device traverses f32
 * rays), and hit geometry is re-solved in binary64.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "vt_oracle.h"

int orc_fail(const char *fmt, ...);
int orc_trace_closest(const orc_scene *s, const orc_ray *rays, uint32_t n, orc_hit *out,
                      uint32_t *counts, uint32_t threads);
int orc_trace_any(const orc_scene *s, const orc_ray *rays, uint32_t n, uint32_t *out,
                  uint32_t threads);

#define PI 3.14159265358979323846
#define INV_PI (1.0 / PI)
#define F32_MAX 3.4028234663852886e38
#define SAFE_MAX 3.4028229765e38 /* 0x7F7FFFFD */
#define ALPHA_FLOOR 1e-3         /* roughDielectric.ts:38 */

typedef struct {
  double x, y, z;
} d3;

static d3 D3(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static d3 add(d3 a, d3 b) { return D3(a.x + b.x, a.y + b.y, a.z + b.z); }
static d3 sub(d3 a, d3 b) { return D3(a.x - b.x, a.y - b.y, a.z - b.z); }
static d3 mul(d3 a, d3 b) { return D3(a.x * b.x, a.y * b.y, a.z * b.z); }
static d3 scl(d3 a, double s) { return D3(a.x * s, a.y * s, a.z * s); }
static double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 cross(d3 a, d3 b) {
  return D3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static double maxd(double a, double b) { return a > b ? a : b; }
static double mind(double a, double b) { return a < b ? a : b; }
static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double max3abs(d3 v) { return maxd(fabs(v.x), maxd(fabs(v.y), fabs(v.z))); }

/* safe_normalize (pt-webgpu/src/wgsl/common.wgsl.ts) */
static d3 safe_normalize(d3 v) {
  double s = max3abs(v);
  if (!(s > 0) || s > SAFE_MAX) return D3(0, 1, 0);
  d3 q = scl(v, 1.0 / s);
  return scl(q, 1.0 / sqrt(dot(q, q)));
}

/* ---- RNG (shared-samplers/src/wgsl/pcg.wgsl.ts; common.wgsl.ts) ------- */

/* Sampler: PCG, or the Owen-scrambled Sobol stream of sobol.c (pt-webgpu's
 * Sobol mode, common.wgsl.ts:22-197) continuing with PCG after dimension
 * 511; dim == ORC_PCG_ONLY selects PCG. */
#define ORC_PCG_ONLY 0xFFFFFFFFu
typedef struct {
  uint32_t s, dim, sample, tile, seed;
} orng;
uint32_t orc_sobol_owen_u32(uint32_t path, uint32_t dim, uint32_t tile, uint32_t stream_seed);
uint32_t orc_sobol_hash(uint32_t x);
uint32_t orc_sobol_combine(uint32_t s, uint32_t v);
uint32_t orc_sobol_frame_key(uint32_t seed, uint32_t frame);

static uint32_t pcg_next(orng *r) {
  if (r->dim < 512u) {
    const uint32_t v = orc_sobol_owen_u32(r->sample, r->dim, r->tile, r->seed);
    r->dim += 1u;
    return v;
  }
  if (r->dim != ORC_PCG_ONLY) r->dim += 1u;
  return orc_pcg_next(&r->s);
}
static double rand01(orng *s) { return (double)(pcg_next(s) >> 8) / 16777216.0; }
static double rand01_(void *s) { return rand01((orng *)s); }
/* Polar disc coordinates sqrt(u1)cos(2 pi u2), sqrt(u1)sin(2 pi u2) rounded
 * once to binary32: the reference draws these in f32, and the product
 * evaluates them in binary64 and rounds once, so both sides start a sampled
 * direction from the same two values (bsdf.wgsl.ts:1648-1664, 1710-1735). */
static void disc_polar(double u1, double u2, double *a, double *b) {
  double phi = 2.0 * PI * u2, r = sqrt(u1);
  *a = (double)(float)(r * cos(phi));
  *b = (double)(float)(r * sin(phi));
}

/* represented_bernoulli_probability_f32 (pcg.wgsl.ts:31-40) */
static double represented(double p) {
  if (!(p > 0)) return 0;
  if (p >= 1) return 1;
  double b = clampd(floor(p * 16777216.0 + 0.5), 1.0, 16777215.0);
  return b / 16777216.0;
}

/* ptRandomU32 / ptRandBoundedU32 (pt-webgpu/src/wgsl/common.wgsl.ts) */
static uint32_t rand_bounded(orng *s, uint32_t bound) {
  if (bound <= 1) return 0;
  uint32_t threshold = ((0xFFFFFFFFu % bound) + 1u) % bound;
  for (;;) {
    uint32_t hi = pcg_next(s) & 0xFFFFFF00u;
    uint32_t lo = pcg_next(s) >> 24;
    uint32_t w = hi | lo;
    if (w >= threshold) return w % bound;
  }
}

static double luminance(d3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }

/* ---- BSDF (bsdf.wgsl.ts, material.wgsl.ts) ------------------------------ */

typedef struct {
  d3 base;
  double rough, metal;
  d3 emit;
  int two_sided;
  /* extension lobes (bsdf.wgsl.ts:392-808) */
  double cc, cc_rough, sheen, sheen_rough;
  d3 sheen_color;
  double aniso, aniso_c, aniso_s; /* strength; cos, sin of the rotation (binary32 table values) */
  d3 spec_color;
  double spec_int;
} mat_t;

/* materialSpecularF0 (bsdf.wgsl.ts:192-209), ior 1.5: the IOR F0 tinted by
 * specularColor (clamped to 1), scaled by specularIntensity, mixed towards the
 * base colour by metallic. */
static d3 spec_f0(const mat_t *m) {
  double ratio = (1.5 - 1.0) / (1.5 + 1.0);
  double f = ratio * ratio;
  double si = clampd(m->spec_int, 0, 1);
  d3 d = D3(mind(f * maxd(m->spec_color.x, 0), 1) * si, mind(f * maxd(m->spec_color.y, 0), 1) * si,
            mind(f * maxd(m->spec_color.z, 0), 1) * si);
  double mt = clampd(m->metal, 0, 1);
  return D3(d.x + (m->base.x - d.x) * mt, d.y + (m->base.y - d.y) * mt, d.z + (m->base.z - d.z) * mt);
}
/* F90 of materialSpecularFresnelSchlick: specularIntensity for dielectrics,
 * 1 for metals. */
static double spec_f90(const mat_t *m) {
  double d = clampd(m->spec_int, 0, 1);
  return d + (1 - d) * clampd(m->metal, 0, 1);
}

/* materialSpecularFresnelSchlick (bsdf.wgsl.ts:211-222) */
static d3 schlick(double cos_t, d3 f0, double f90) {
  double x = clampd(1.0 - cos_t, 0, 1);
  double x5 = x * x * x * x * x;
  return D3(f0.x + (f90 - f0.x) * x5, f0.y + (f90 - f0.y) * x5, f0.z + (f90 - f0.z) * x5);
}
static d3 clamp01v(d3 v) { return D3(clampd(v.x, 0, 1), clampd(v.y, 0, 1), clampd(v.z, 0, 1)); }

/* ggxD (material.wgsl.ts:2797-2804) */
static double ggx_d(double ndh, double alpha) {
  double a2 = alpha * alpha;
  double n2 = clampd(ndh * ndh, 0, 1);
  double d = (1.0 - n2) + n2 * a2;
  return a2 / (PI * d * d);
}

/* smithG1 (pt-webgpu/src/math/roughDielectric.ts:52-60) */
static double smith_g1(double ndv, double rough) {
  double c = maxd(ndv, 0);
  if (c <= 1e-6) return 0;
  double alpha = maxd(rough * rough, ALPHA_FLOOR);
  double tan2 = maxd(0, 1 - c * c) / maxd(c * c, 1e-10);
  return 2.0 / (1.0 + sqrt(1.0 + alpha * alpha * tan2));
}

/* powerHeuristic (material.wgsl.ts:2808-2822) */
static double power_heuristic(double a, double b) {
  if (!(a >= 0) || !(b >= 0) || a > F32_MAX || b > F32_MAX) return 0;
  double s = maxd(a, b);
  if (!(s > 0)) return 0;
  a /= s;
  b /= s;
  return (a * a) / (a * a + b * b);
}

/* Kulla-Conty tables and lookups (material.wgsl.ts:2852-2931) */
static const double GGX_E_LUT[64] = {
    0.1375, 0.5617, 0.7546, 0.8522, 0.9111, 0.9505, 0.9788, 1.0,
    0.2955, 0.515,  0.7091, 0.8192, 0.889,  0.937,  0.9721, 0.9988,
    0.5794, 0.5541, 0.6677, 0.7691, 0.8451, 0.9021, 0.9457, 0.98,
    0.7011, 0.6486, 0.6669, 0.7199, 0.7776, 0.8305, 0.8764, 0.9155,
    0.7335, 0.6901, 0.6696, 0.6756, 0.6972, 0.7262, 0.7578, 0.7893,
    0.7153, 0.6712, 0.6355, 0.6145, 0.6052, 0.6045, 0.6101, 0.6199,
    0.6669, 0.6137, 0.5657, 0.5286, 0.5,    0.478,  0.4611, 0.4483,
    0.6017, 0.537,  0.4773, 0.4296, 0.3905, 0.358,  0.3305, 0.3069};
static const double GGX_EAVG_LUT[8] = {0.9106, 0.8931, 0.8629, 0.8094,
                                       0.725,  0.6147, 0.4931, 0.3766};

static double lut_e(double cos_t, double rough) {
  double mu = clampd(cos_t, 0, 1), r = clampd(rough, 0, 1);
  double fr = r * 7.0, fm = mu * 7.0;
  unsigned r0 = (unsigned)floor(fr), m0 = (unsigned)floor(fm);
  unsigned r1 = r0 + 1 < 7 ? r0 + 1 : 7, m1 = m0 + 1 < 7 ? m0 + 1 : 7;
  double tr = fr - r0, tm = fm - m0;
  double e0 = GGX_E_LUT[r0 * 8 + m0] + (GGX_E_LUT[r0 * 8 + m1] - GGX_E_LUT[r0 * 8 + m0]) * tm;
  double e1 = GGX_E_LUT[r1 * 8 + m0] + (GGX_E_LUT[r1 * 8 + m1] - GGX_E_LUT[r1 * 8 + m0]) * tm;
  return clampd(e0 + (e1 - e0) * tr, 0.02, 1.0);
}
static double lut_eavg(double rough) {
  double r = clampd(rough, 0, 1), fr = r * 7.0;
  unsigned r0 = (unsigned)floor(fr), r1 = r0 + 1 < 7 ? r0 + 1 : 7;
  double tr = fr - r0;
  return clampd(GGX_EAVG_LUT[r0] + (GGX_EAVG_LUT[r1] - GGX_EAVG_LUT[r0]) * tr, 0.3, 1.0);
}
/* ggxMultiscatterLobeRoughness (material.wgsl.ts:2906-2926) */
static d3 ggx_ms(d3 f0, double rough_v, double rough_l, double rough_avg, double ndv, double ndl) {
  double eavg = lut_eavg(rough_avg), om = 1.0 - eavg;
  if (om < 1e-4) return D3(0, 0, 0);
  double eo = lut_e(ndv, rough_v), ei = lut_e(ndl, rough_l);
  double shape = (1.0 - eo) * (1.0 - ei) / maxd(PI * om, 1e-6);
  double c[3] = {f0.x, f0.y, f0.z}, o[3];
  for (int k = 0; k < 3; ++k) {
    double favg = c[k] + (1.0 - c[k]) * (1.0 / 21.0);
    double fms = favg * favg * eavg / maxd(1.0 - favg * om, 1e-4);
    o[k] = fms * shape;
  }
  return D3(o[0], o[1], o[2]);
}

/* buildOnb (bsdf.wgsl.ts:1632-1640) */
static void build_onb(d3 n, d3 *t, d3 *b) {
  d3 up = fabs(n.y) > 0.999 ? D3(1, 0, 0) : D3(0, 1, 0);
  d3 c = cross(up, n);
  double l = sqrt(dot(c, c));
  *t = D3(c.x / l, c.y / l, c.z / l);
  *b = cross(n, *t);
}
/* the frame rotated by anisotropyRotation (bsdf.wgsl.ts:2582-2595) */
static void aniso_frame(const mat_t *m, d3 n, d3 *t, d3 *b) {
  d3 t0, b0;
  build_onb(n, &t0, &b0);
  *t = add(scl(t0, m->aniso_c), scl(b0, m->aniso_s));
  *b = add(scl(t0, -m->aniso_s), scl(b0, m->aniso_c));
}
static double smoothstep_d(double e0, double e1, double x) {
  double t = clampd((x - e0) / (e1 - e0), 0, 1);
  return t * t * (3 - 2 * t);
}
/* computeAnisotropicAxes (bsdf.wgsl.ts:1928-1933) */
static void aniso_axes(double alpha, double aniso, double *ax, double *ay) {
  double aspect = sqrt(maxd(1.0 - 0.9 * aniso, 1e-4));
  *ax = maxd(alpha / aspect, 1e-4);
  *ay = maxd(alpha * aspect, 1e-4);
}
/* ggxDAnis, smithG1Anis (bsdf.wgsl.ts:597-608) */
static double ggx_d_aniso(double ht, double hb, double hn, double ax, double ay) {
  double d = (ht / ax) * (ht / ax) + (hb / ay) * (hb / ay) + hn * hn;
  return 1.0 / maxd(PI * ax * ay * d * d, 1e-10);
}
static double smith_g1_aniso(double vt, double vb, double vn, double ax, double ay) {
  double vn2 = maxd(vn * vn, 1e-10);
  return (2.0 * vn) / maxd(vn + sqrt(vn2 + (vt * ax) * (vt * ax) + (vb * ay) * (vb * ay)), 1e-6);
}
/* anisotropicProjectedRoughness / anisotropicAverageRoughness /
 * anisotropicMultiscatterScale (bsdf.wgsl.ts:755-796) */
static double aniso_projected_rough(d3 dir, d3 t, d3 b, double rough, double aniso) {
  double ax, ay;
  aniso_axes(maxd(rough * rough, ALPHA_FLOOR), aniso, &ax, &ay);
  double dt = dot(dir, t), db = dot(dir, b), tl2 = dt * dt + db * db;
  double blend = clampd(0.15 * aniso, 0, 0.15), projected;
  if (tl2 <= 1e-6)
    projected = sqrt(clampd(0.5 * (ax + ay), 1e-4, 1));
  else
    projected = sqrt(clampd(sqrt(((dt * ax) * (dt * ax) + (db * ay) * (db * ay)) / tl2), 1e-4, 1));
  return rough + (projected - rough) * blend;
}
static double aniso_average_rough(double rough, double aniso) {
  double ax, ay;
  aniso_axes(maxd(rough * rough, ALPHA_FLOOR), aniso, &ax, &ay);
  double projected = sqrt(clampd(sqrt(0.5 * (ax * ax + ay * ay)), 1e-4, 1));
  return rough + (projected - rough) * clampd(0.15 * aniso, 0, 0.15);
}
static double aniso_ms_scale(double aniso, double rough) {
  double reduction = smoothstep_d(0, 0.35, clampd(aniso, 0, 1));
  double scale = 0.6 * smoothstep_d(0.35, 0.9, rough);
  return 1.0 + (scale - 1.0) * reduction;
}
/* brdfAnisotropicSpecPdf (bsdf.wgsl.ts:727-753) */
static double aniso_spec_pdf(const mat_t *m, d3 n, d3 t, d3 b, d3 wo, d3 wi) {
  double ndv = maxd(dot(n, wo), 1e-6), ndl = maxd(dot(n, wi), 0);
  if (ndl <= 1e-5) return 0;
  d3 h = safe_normalize(add(wo, wi));
  double ax, ay;
  aniso_axes(maxd(m->rough * m->rough, ALPHA_FLOOR), m->aniso, &ax, &ay);
  double d = ggx_d_aniso(dot(h, t), dot(h, b), maxd(dot(h, n), 0), ax, ay);
  double g1 = smith_g1_aniso(dot(wo, t), dot(wo, b), maxd(dot(wo, n), 1e-6), ax, ay);
  return (d * g1) / maxd(4.0 * ndv, 1e-6);
}
/* clearcoatLayerWeight / evalClearcoatLobe / clearcoatPdf (bsdf.wgsl.ts:398-472) */
static double cc_weight(const mat_t *m, d3 n, d3 wo) {
  double x = clampd(1.0 - clampd(fabs(dot(n, wo)), 0, 1), 0, 1);
  double fcc = 0.04 + (1.0 - 0.04) * (x * x * x * x * x);
  return clampd(m->cc, 0, 1) * fcc;
}
static double cc_eval(const mat_t *m, d3 n, d3 wo, d3 wi) {
  if (m->cc <= 0) return 0;
  double ndl = maxd(dot(n, wi), 0), ndv = maxd(dot(n, wo), 0);
  if (ndl <= 1e-5 || ndv <= 1e-5) return 0;
  d3 h = safe_normalize(add(wi, wo));
  double d = ggx_d(maxd(dot(n, h), 0), maxd(m->cc_rough * m->cc_rough, ALPHA_FLOOR));
  double g = smith_g1(ndv, m->cc_rough) * smith_g1(ndl, m->cc_rough);
  return cc_weight(m, n, wo) * ((d * g) / maxd(4.0 * ndv * ndl, 1e-6));
}
static double cc_pdf(const mat_t *m, d3 n, d3 wo, d3 wi) {
  if (m->cc <= 0) return 0;
  double ndv = maxd(dot(n, wo), 0);
  if (ndv <= 1e-5) return 0;
  double ndl = maxd(dot(n, wi), 0);
  if (ndl <= 1e-5) return 0;
  d3 h = safe_normalize(add(wi, wo));
  double d = ggx_d(maxd(dot(n, h), 0), maxd(m->cc_rough * m->cc_rough, ALPHA_FLOOR));
  return (d * smith_g1(ndv, m->cc_rough)) / maxd(4.0 * ndv, 1e-6);
}
/* charlieD, sheenVisibility, evalSheenLobe, sheenDirectionalAlbedo,
 * sheenBaseAttenuation, charlieSheenPdf (bsdf.wgsl.ts:474-596) */
static double charlie_d(double ndh, double alpha) {
  double inv = 1.0 / maxd(alpha, 1e-4);
  double sin_h = sqrt(maxd(0, 1 - ndh * ndh));
  return (2.0 + inv) * pow(sin_h, inv) / (2.0 * PI);
}
static double sheen_vis(double ndl, double ndv) { return 1.0 / maxd(4.0 * (ndl + ndv - ndl * ndv), 1e-6); }
static d3 sheen_eval(const mat_t *m, d3 n, d3 wo, d3 wi) {
  if (m->sheen <= 0) return D3(0, 0, 0);
  double ndl = maxd(dot(n, wi), 0), ndv = maxd(dot(n, wo), 0);
  if (ndl <= 1e-5 || ndv <= 1e-5) return D3(0, 0, 0);
  d3 h = safe_normalize(add(wi, wo));
  double alpha = maxd(m->sheen_rough * m->sheen_rough, ALPHA_FLOOR);
  return scl(m->sheen_color, m->sheen * charlie_d(maxd(dot(n, h), 0), alpha) * sheen_vis(ndl, ndv));
}
static double sheen_albedo(double c_raw, double alpha) {
  double c = 1.0 - clampd(c_raw, 0, 1);
  return 0.65584461 * (c * c * c) + 1.0 / (4.16526551 + exp(-7.97291361 * sqrt(alpha) + 6.33516894));
}
static double sheen_attenuation(const mat_t *m, d3 n, d3 wo, d3 wi) {
  if (m->sheen <= 0) return 1;
  double ndl = clampd(fabs(dot(n, wi)), 0, 1), ndv = clampd(fabs(dot(n, wo)), 0, 1);
  double a = maxd(m->sheen_rough, 0.07), a2 = a * a;
  double mc = maxd(maxd(m->sheen_color.x, m->sheen_color.y), m->sheen_color.z);
  double full = mind(1.0 - mc * sheen_albedo(ndv, a2), 1.0 - mc * sheen_albedo(ndl, a2));
  return clampd(1.0 + (full - 1.0) * clampd(m->sheen, 0, 1), 0, 1);
}
static double sheen_pdf(const mat_t *m, d3 n, d3 wo, d3 wi) {
  if (m->sheen <= 0) return 0;
  double ndl = maxd(dot(n, wi), 0), ndv = maxd(dot(n, wo), 0);
  if (ndl <= 1e-5 || ndv <= 1e-5) return 0;
  d3 h = safe_normalize(add(wi, wo));
  double ndh = maxd(dot(n, h), 0), vdh = maxd(dot(wo, h), 1e-6);
  double alpha = maxd(m->sheen_rough * m->sheen_rough, ALPHA_FLOOR);
  return (charlie_d(ndh, alpha) * ndh) / maxd(4.0 * vdh, 1e-6);
}

/* evaluateBrdfFullWithClearcoatNormal (bsdf.wgsl.ts:809-916): base lobes
 * (Lambert, isotropic or anisotropic GGX with evalBrdfSpecAnisotropic,
 * Kulla-Conty), under sheen, under clearcoat:
 *   (base * sheenAttenuation + sheen) * clearcoatAttenuation + clearcoat. */
static d3 bsdf_eval(const mat_t *m, d3 n, d3 wo, d3 wi) {
  double ndl = maxd(dot(n, wi), 0), ndv = maxd(dot(n, wo), 0);
  if (ndl <= 1e-5 || ndv <= 1e-5) return D3(0, 0, 0);
  d3 h = safe_normalize(add(wi, wo));
  double ndh = maxd(dot(n, h), 0), vdh = maxd(dot(wo, h), 0);
  d3 f0 = spec_f0(m);
  d3 F = clamp01v(schlick(vdh, f0, spec_f90(m)));
  double kdm = 1.0 - m->metal;
  d3 diff = D3((1 - F.x) * kdm * m->base.x * INV_PI, (1 - F.y) * kdm * m->base.y * INV_PI,
               (1 - F.z) * kdm * m->base.z * INV_PI);
  d3 spec, ms;
  if (m->aniso > 0) {
    d3 t, b;
    aniso_frame(m, n, &t, &b);
    double ax, ay;
    aniso_axes(maxd(m->rough * m->rough, ALPHA_FLOOR), m->aniso, &ax, &ay);
    double d = ggx_d_aniso(dot(h, t), dot(h, b), maxd(dot(h, n), 0), ax, ay);
    double g = smith_g1_aniso(dot(wo, t), dot(wo, b), maxd(dot(wo, n), 1e-6), ax, ay) *
               smith_g1_aniso(dot(wi, t), dot(wi, b), maxd(dot(wi, n), 1e-6), ax, ay);
    spec = scl(F, (d * g) / maxd(4.0 * maxd(ndv, 1e-6) * ndl, 1e-6));
    double ravg = aniso_average_rough(m->rough, m->aniso);
    ms = scl(ggx_ms(clamp01v(f0), aniso_projected_rough(wo, t, b, m->rough, m->aniso),
                    aniso_projected_rough(wi, t, b, m->rough, m->aniso), ravg, ndv, ndl),
             aniso_ms_scale(m->aniso, ravg));
  } else {
    double alpha = maxd(m->rough * m->rough, ALPHA_FLOOR);
    double d = ggx_d(ndh, alpha);
    double g = smith_g1(ndv, m->rough) * smith_g1(ndl, m->rough);
    spec = scl(F, (d * g) / maxd(4.0 * ndv * ndl, 1e-6));
    ms = ggx_ms(clamp01v(f0), m->rough, m->rough, m->rough, ndv, ndl);
  }
  d3 base = add(add(diff, spec), ms);
  if (m->cc <= 0 && m->sheen <= 0) return base;
  double cc_att = m->cc > 0 ? clampd(1.0 - cc_weight(m, n, wo), 0, 1) : 1.0;
  d3 layered = scl(add(scl(base, sheen_attenuation(m, n, wo, wi)), sheen_eval(m, n, wo, wi)), cc_att);
  double cc = cc_eval(m, n, wo, wi);
  return D3(layered.x + cc, layered.y + cc, layered.z + cc);
}

/* brdfFiniteBaseLobeWeights (bsdf.wgsl.ts:1172-1210): specular probability. */
static double spec_prob(const mat_t *m, double ndv) {
  d3 F = clamp01v(schlick(ndv, spec_f0(m), spec_f90(m)));
  double x = maxd(luminance(F), m->metal);
  return represented(clampd(0.04 + (0.96 - 0.04) * x, 0.04, 0.96));
}

/* brdfRepresentedExtensionLobeProbabilities (bsdf.wgsl.ts:1219-1262):
 * p[0..3] = joint base, joint clearcoat, joint sheen, P(clearcoat | ext). */
static void ext_probs(const mat_t *m, double p[4]) {
  double cc = maxd(m->cc, 0), sh = maxd(m->sheen, 0), ext = cc + sh;
  if (!(ext > 0)) {
    p[0] = 1, p[1] = p[2] = p[3] = 0;
    return;
  }
  double total = 1.0 + ext, pb, pe, pc;
  if (ext <= 1.0) {
    pe = represented(ext / total);
    pb = 1.0 - pe;
  } else {
    pb = represented(1.0 / total);
    pe = 1.0 - pb;
  }
  if (!(cc > 0))
    pc = 0;
  else if (!(sh > 0))
    pc = 1;
  else if (cc <= sh)
    pc = represented(cc / ext);
  else
    pc = 1.0 - represented(sh / ext);
  p[0] = pb, p[1] = pe * pc, p[2] = pe - pe * pc, p[3] = pc;
}

/* brdfDirectionalPdfFullWithClearcoatNormal (bsdf.wgsl.ts:1370-1489) over
 * brdfDirectionalPdfWithIridescence (:1555-1631) or the anisotropic VNDF. */
static double bsdf_pdf(const mat_t *m, d3 n, d3 wo, d3 wi) {
  double wdn = dot(n, wi), odn = dot(n, wo);
  double ndv = maxd(odn, 0);
  if (ndv <= 1e-5) return 0;
  double sp = spec_prob(m, ndv), dp = 1.0 - sp;
  if (!(wdn * odn > 0)) return 0;
  double ndl = maxd(wdn, 0);
  if (ndl <= 1e-5) return 0;
  double pspec;
  if (m->aniso > 0) {
    d3 t, b;
    aniso_frame(m, n, &t, &b);
    pspec = aniso_spec_pdf(m, n, t, b, wo, wi);
  } else {
    d3 h = safe_normalize(add(wi, wo));
    double d = ggx_d(maxd(dot(n, h), 0), maxd(m->rough * m->rough, ALPHA_FLOOR));
    pspec = (d * smith_g1(ndv, m->rough)) / maxd(4.0 * ndv, 1e-6);
  }
  double base = dp * ndl * INV_PI + sp * pspec;
  if (m->cc <= 0 && m->sheen <= 0) return base;
  double p[4];
  ext_probs(m, p);
  return (p[0] * base + p[1] * cc_pdf(m, n, wo, wi)) + p[2] * sheen_pdf(m, n, wo, wi);
}

/* ---- camera (kernelCore.wgsl.ts:22-26; common.wgsl.ts:412-469) --------- */


/* generatePrimaryRay / unproject_ray_common (kernelCore.wgsl.ts:22-26;
 * common.wgsl.ts:412-469).  Camera rays are generated in f32, as the reference
 * shader does: the near/far unprojection cancels heavily in f32 when
 * far/near is large, so a binary64 camera would trace different rays.  The
 * matrix-vector products sum the columns in order 0..3; divisions and square
 * roots round once (IEEE binary32). */
static float f32max3abs(float a, float b, float c) {
  float x = fabsf(a), y = fabsf(b), z = fabsf(c);
  return x > y ? (x > z ? x : z) : (y > z ? y : z);
}
static int primary_ray(const float *m, uint32_t px, uint32_t py, uint32_t w, uint32_t h, float jx,
                       float jy, d3 *o, d3 *d) {
  const float u = ((float)px + jx) / (float)w, v = ((float)py + jy) / (float)h;
  const float nx = u * 2.0f - 1.0f, ny = 1.0f - v * 2.0f;
  float fh[4], nh[4];
  for (int r = 0; r < 4; ++r) {
    fh[r] = ((m[r] * nx + m[4 + r] * ny) + m[8 + r] * 1.0f) + m[12 + r];
    nh[r] = ((m[r] * nx + m[4 + r] * ny) + m[8 + r] * -1.0f) + m[12 + r];
  }
  const float fs = fmaxf(fmaxf(fabsf(fh[0]), fabsf(fh[1])), fmaxf(fabsf(fh[2]), fabsf(fh[3])));
  const float ns = fmaxf(fmaxf(fabsf(nh[0]), fabsf(nh[1])), fmaxf(fabsf(nh[2]), fabsf(nh[3])));
  const float safe = 3.4028229765e38f; /* 0x7F7FFFFD */
  if (!(fs > 0.0f && fs <= safe && ns > 0.0f && ns <= safe)) return 0;
  for (int r = 0; r < 4; ++r) fh[r] = fh[r] / fs, nh[r] = nh[r] / ns;
  if (nh[3] == 0.0f) return 0;
  *o = D3(nh[0] / nh[3], nh[1] / nh[3], nh[2] / nh[3]);
  const float orient = fh[3] != 0.0f ? (fh[3] * nh[3] > 0.0f ? 1.0f : (fh[3] * nh[3] < 0.0f ? -1.0f : 0.0f)) : 1.0f;
  float q[3] = {(fh[0] * nh[3] - nh[0] * fh[3]) * orient, (fh[1] * nh[3] - nh[1] * fh[3]) * orient,
                (fh[2] * nh[3] - nh[2] * fh[3]) * orient};
  /* safe_normalize in f32 (common.wgsl.ts) */
  const float sc = f32max3abs(q[0], q[1], q[2]);
  if (!(sc > 0.0f) || sc > safe) {
    *d = D3(0, 1, 0);
    return 1;
  }
  for (int k = 0; k < 3; ++k) q[k] = q[k] / sc;
  const float l = sqrtf((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]);
  *d = D3(q[0] / l, q[1] / l, q[2] / l);
  return 1;
}

/* ---- emitters (scene/emitterPacking.ts:748-820) -------------------------- */

typedef struct {
  d3 a, b, c; /* world vertices (f32 values) */
  d3 le;
  int two_sided;
  uint32_t inst, prim;
} emitter_t;

/* measureAreaVector (common.wgsl.ts:310-362) in binary64: unit normal and area. */
static int area_measure(d3 u, d3 v, d3 *normal, double *area) {
  d3 c = cross(u, v);
  double l = sqrt(dot(c, c));
  if (!(l > 0)) return 0;
  *normal = scl(c, 1.0 / l);
  *area = 0.5 * l;
  return 1;
}

/* ptAreaToSolidAnglePdf (material.wgsl.ts:242-275) */
static double area_to_solid_angle(double dist, double cos_l, double area) {
  if (!(dist > 0) || !(cos_l > 0) || !(area > 0)) return 0;
  return dist * dist / (cos_l * area);
}

/* ---- analytic emitters (kernel.wgsl.ts:903-918, 1198-1497) --------------- */

/* the area element of a rect (half axes u, v: 4|u x v|) or disc (the affine
 * image of the unit disc: pi |u x v|) and its unit normal u x v */
static int analytic_area(d3 u, d3 v, int disc, d3 *normal, double *area) {
  d3 c = cross(u, v);
  double l = sqrt(dot(c, c));
  if (!(l > 0)) return 0;
  *normal = scl(c, 1.0 / l);
  *area = (disc ? PI : 4.0) * l;
  return 1;
}
/* ptDirectionalConeIsDelta / ptDirectionalConeSinCos (material.wgsl.ts:
 * 208-235) and sampleDirectionalCone (kernel.wgsl.ts:903-918) */
static int cone_is_delta(double ad) { return !(ad > 0) || sin(ad * 0.25) < 7.666467e-20; }
static d3 sample_directional_cone(d3 axis, double ad, orng *rng) {
  d3 d = safe_normalize(axis);
  if (cone_is_delta(ad)) return d;
  double xi1 = rand01(rng), xi2 = rand01(rng);
  double sq = sin(ad * 0.25), omc = 2.0 * sq * sq;
  double q = (1.0 - clampd(xi1, 0, 1)) * omc;
  double st = sqrt(maxd(q * (2.0 - q), 0)), ct = sqrt(maxd(1.0 - st * st, 0));
  double phi = 6.28318530718 * xi2;
  d3 tx = fabs(d.x) > 0.9 ? D3(0, 1, 0) : D3(1, 0, 0);
  d3 by = safe_normalize(cross(d, tx));
  d3 bx = cross(by, d);
  return safe_normalize(add(add(scl(bx, st * cos(phi)), scl(by, st * sin(phi))), scl(d, ct)));
}
/* pointSpotDistanceAttenuation (material.wgsl.ts:2176-2204) */
static double point_attenuation(double dist, double cutoff, double decay) {
  if (!(dist > 0)) return 0;
  double att = 1.0;
  if (decay > 0.01) att = 1.0 / pow(dist, decay);
  if (cutoff > 0) att *= clampd(1.0 - pow(dist / cutoff, 4.0), 0, 1);
  return att;
}
/* concentricDiscSample (kernelCore.wgsl.ts:73-85) */
static void concentric_disc(double a, double b, double *x, double *y) {
  if (a == 0 && b == 0) {
    *x = *y = 0;
    return;
  }
  double r, phi;
  if (fabs(a) >= fabs(b)) {
    r = a;
    phi = (PI / 4.0) * (b / a);
  } else {
    r = b;
    phi = PI / 2.0 - (PI / 4.0) * (a / b);
  }
  *x = r * cos(phi);
  *y = r * sin(phi);
}
/* intersectRectAreaLightRay (connect.wgsl.ts:395-446), binary64 */
static int intersect_rect_light(const float *q, d3 o, d3 d, double tmin, double *t_out, double *pdf_out) {
  d3 c = D3(q[0], q[1], q[2]), u = D3(q[4], q[5], q[6]), v = D3(q[8], q[9], q[10]);
  int disc = fabs(q[15] - 1.0f) < 0.5f;
  d3 n;
  double area;
  if (!analytic_area(u, v, disc, &n, &area)) return 0;
  double den = dot(n, d);
  if (den == 0) return 0;
  double t = dot(n, sub(c, o)) / den;
  if (t <= tmin) return 0;
  d3 rel = sub(add(o, scl(d, t)), c);
  /* coordinates of rel in (u, v): least squares through the normal equations */
  double uu = dot(u, u), uv = dot(u, v), vv = dot(v, v), ru = dot(rel, u), rv = dot(rel, v);
  double det = uu * vv - uv * uv;
  if (det == 0) return 0;
  double a = (ru * vv - rv * uv) / det, b = (rv * uu - ru * uv) / det;
  int inside = disc ? (a * a + b * b <= 1.0) : (fabs(a) <= 1.0 && fabs(b) <= 1.0);
  if (!inside) return 0;
  double cl = maxd(dot(n, scl(d, -1)), 0);
  if (cl <= 0) return 0;
  double pdf = area_to_solid_angle(t, cl, area);
  if (!(pdf > 0)) return 0;
  *t_out = t;
  *pdf_out = pdf;
  return 1;
}

/* ---- environment (connect.wgsl.ts:126-219; connectCore.wgsl.ts:46-61) --- */

typedef struct {
  int present;
  uint32_t w, h;
  const float *cdf;    /* w*h + 1 entries, f32 */
  const float *texels; /* per texel: intensity-scaled rgb, solid-angle pdf of the sampler */
  float rot_c, rot_s;  /* the dome's rotation about +Y */
} env_t;

/* RY(-rot) d: the unrotated map's direction for world direction d, and
 * RY(+rot) d back (connectCore.wgsl.ts:20-44), in binary32. */
static d3 env_to_map(const env_t *e, d3 d) {
  float x = (float)d.x, y = (float)d.y, z = (float)d.z, c = e->rot_c, s = e->rot_s;
  float a = c * x, b = s * z, p = s * x, q = c * z;
  return D3((double)(a - b), (double)y, (double)(p + q));
}
static d3 env_to_world(const env_t *e, d3 d) {
  float x = (float)d.x, y = (float)d.y, z = (float)d.z, c = e->rot_c, s = e->rot_s;
  float a = c * x, b = s * z, p = -s * x, q = c * z;
  return D3((double)(a + b), (double)y, (double)(p + q));
}

/* environmentLookup's texel (connect.wgsl.ts:126-153): angles in binary64. */
static const float *env_lookup(const env_t *e, d3 wd) {
  d3 d = env_to_map(e, wd);
  double phi = atan2(d.z, d.x);
  double theta = acos(d.y < -1 ? -1 : (d.y > 1 ? 1 : d.y));
  double u = phi * (0.5 / PI) + 0.5;
  u = u - floor(u);
  double v = theta / PI;
  v = v < 0 ? 0 : (v > 0.999999 ? 0.999999 : v);
  uint32_t x = (uint32_t)floor(u * e->w), y = (uint32_t)floor(v * e->h);
  if (x > e->w - 1) x = e->w - 1;
  if (y > e->h - 1) y = e->h - 1;
  return e->texels + 4 * ((size_t)y * e->w + x);
}
static d3 env_rgb(const float *t) { return D3(t[0], t[1], t[2]); }

static d3 env_cell_dir(uint32_t x, uint32_t y, uint32_t w, uint32_t h, double xa, double xb) {
  double phi0 = ((double)x / w - 0.5) * (2.0 * PI), phi1 = ((double)(x + 1) / w - 0.5) * (2.0 * PI);
  double th0 = (double)y / h * PI, th1 = (double)(y + 1) / h * PI;
  double phi = phi0 + (phi1 - phi0) * xa;
  double c0 = cos(th0), c1 = cos(th1);
  double ct = c0 + (c1 - c0) * xb;
  double st = sqrt(maxd(0, 1 - ct * ct));
  /* each component rounded once to binary32, as the product rounds it */
  return D3((double)(float)(cos(phi) * st), (double)(float)ct, (double)(float)(sin(phi) * st));
}

/* ---- render ----------------------------------------------------------- */

typedef struct orc_render_desc {
  const orc_scene *scene;
  const float *normals;           /* vec4 per packed vertex */
  const uint32_t *tri_material;   /* packed triangle -> material */
  const float *materials;         /* 12 floats: base rgb, rough, metal, emit rgb, two_sided, pad */
  uint32_t material_count;
  const float *emitters;          /* 16 floats: a xyz, b xyz, c xyz, le rgb, two_sided, inst, prim, pad */
  uint32_t emitter_count;
  uint32_t env_present;
  uint32_t env_w, env_h;
  const float *env_cdf;
  const float *env_texels;        /* 4 floats per texel: scaled rgb, pdf */
  float env_rot[2];               /* cos, sin of the rotation about +Y */
  float inv_view_proj[16];        /* column-major, f32 */
  uint32_t width, height;
  uint32_t frame_seed;
  uint32_t max_bounces;
  float ray_bias;
  uint32_t sampling; /* 0 = PCG, 1 = Owen-scrambled Sobol */
  const float *light_tree;        /* packed light tree (light_tree.c), or NULL */
  uint32_t light_tree_nodes;      /* 0 = uniform light selection */
  /* NULL, or 8 doubles per rendered pixel: the first hit's f32 shading
   * normal * 0.5 + 0.5, its distance from the camera ray's origin, base
   * colour, 1 -- (0.5, 1, 0.5, 0, 0, 0, 0, 0) for a miss (accumulateFrame,
   * kernel.wgsl.ts:946-957). */
  double *aux;
  /* Analytic emitters, 16 floats each in light-selection order (directional,
   * point, spot, rect/disc; the records of the product's render.h): they come
   * before the mesh-area emitters in the selection index. */
  const float *lights;
  uint32_t n_dir, n_point, n_spot, n_rect;
} orc_render_desc;

typedef struct {
  const orc_render_desc *d;
  emitter_t *em;
  mat_t *mats;
  env_t env;
  double bias, tmin;
  uint8_t *single_sided; /* per packed triangle: back faces are not surfaces */
  uint32_t hit_limit;    /* sceneSurfaceHitLimit (material.wgsl.ts:423-445) */
} rctx;

static float f32(double x) { return (float)x; }

static orc_ray mk_ray(d3 o, d3 d, double tmin, double tmax) {
  orc_ray r;
  r.o[0] = f32(o.x), r.o[1] = f32(o.y), r.o[2] = f32(o.z);
  r.d[0] = f32(d.x), r.d[1] = f32(d.y), r.d[2] = f32(d.z);
  r.tmin = f32(tmin);
  r.tmax = tmax >= F32_MAX ? (float)F32_MAX : f32(tmax);
  return r;
}

/* traceSurfaceVisibility (connect.wgsl.ts:37-103), opaque materials. */
static int visible(const rctx *c, d3 o, d3 d, double tmax) {
  orc_ray r = mk_ray(o, d, c->tmin, tmax);
  uint32_t occ = 0;
  orc_trace_occluded_sided(c->d->scene, c->single_sided, c->hit_limit, &r, 1, &occ, 1);
  return occ == 0;
}

/* ---- binary32 ray generation ----------------------------------------------
 * The reference computes the shading normal, the sampled direction and the
 * next ray's origin in f32 (shadePrologue.wgsl.ts, bsdf.wgsl.ts:1632-1778,
 * kernel.wgsl.ts:1993-2013), and every later decision depends on those bits:
 * the next hit point feeds the light tree, whose 2^24-bucket split moves by
 * one bucket for a relative change of 1e-7 in a child's importance, so a
 * binary64 ray would reach a different light-selection pdf (by 1/buckets,
 * ~1e-4 for a leaf holding a few thousand buckets).  These helpers restate
 * that arithmetic operation for operation: one IEEE binary32 rounding per
 * operation, sums left to right, sqrt and divisions correctly rounded.
 * Shading values (BSDF, pdfs, MIS, radiance) stay binary64. */
typedef struct {
  float x, y, z;
} f3_t;
static f3_t F3f(float x, float y, float z) {
  f3_t r = {x, y, z};
  return r;
}
static float dotf3(f3_t a, f3_t b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static f3_t crossf3(f3_t a, f3_t b) {
  return F3f(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static f3_t addf3(f3_t a, f3_t b) { return F3f(a.x + b.x, a.y + b.y, a.z + b.z); }
static f3_t subf3(f3_t a, f3_t b) { return F3f(a.x - b.x, a.y - b.y, a.z - b.z); }
static f3_t sclf3(f3_t a, float s) { return F3f(a.x * s, a.y * s, a.z * s); }
static float maxf_(float a, float b) { return a > b ? a : b; }
/* safe_normalize (common.wgsl.ts): max-abs scaling, then unit length */
static f3_t safe_normalize_f(f3_t v) {
  const float s = maxf_(fabsf(v.x), maxf_(fabsf(v.y), fabsf(v.z)));
  if (!(s > 0.0f) || s > 3.4028229765e38f) return F3f(0, 1, 0);
  const f3_t q = F3f(v.x / s, v.y / s, v.z / s);
  const float l = sqrtf(dotf3(q, q));
  return F3f(q.x / l, q.y / l, q.z / l);
}
/* buildOnb (bsdf.wgsl.ts:1632-1640) */
static void build_onb_f(f3_t n, f3_t *t, f3_t *b) {
  const f3_t up = fabsf(n.y) > 0.999f ? F3f(1, 0, 0) : F3f(0, 1, 0);
  const f3_t c = crossf3(up, n);
  const float l = sqrtf(dotf3(c, c));
  *t = F3f(c.x / l, c.y / l, c.z / l);
  *b = crossf3(n, *t);
}
static d3 D3f(f3_t v) { return D3(v.x, v.y, v.z); }
/* a shadow or connection ray's origin: the hit point pushed off along the
 * normal on the side the ray leaves (hitPos + offsetNormal * bias, f32) */
static d3 shadow_origin(f3_t hp, f3_t n, int upper, float bias) {
  const f3_t off = upper ? n : F3f(-n.x, -n.y, -n.z);
  return D3f(addf3(hp, sclf3(off, bias)));
}

static void disc_polar(double u1, double u2, double *a, double *b);
static double rand01_(void *rng);
/* sampleGgxVndfTangent (bsdf.wgsl.ts:1710-1735) and, with ax != ay,
 * sampleGgxVndfAnisTangent (:612-637), in f32 */
static f3_t vndf_sample_f(f3_t wo, float ax, float ay, void *rng) {
  const f3_t vh = safe_normalize_f(F3f(ax * wo.x, ay * wo.y, wo.z));
  const float lensq = vh.x * vh.x + vh.y * vh.y;
  const float inv = lensq > 1e-10f ? 1.0f / sqrtf(lensq) : 0.0f;
  const f3_t t1 = lensq > 1e-10f ? F3f(-vh.y * inv, vh.x * inv, 0.0f) : F3f(1, 0, 0);
  const f3_t t2 = crossf3(vh, t1);
  const double u1 = rand01_(rng), u2 = rand01_(rng);
  double ad, bd;
  disc_polar(u1, u2, &ad, &bd);
  const float a = (float)ad;
  float b = (float)bd;
  const float sw = 0.5f * (1.0f + vh.z);
  b = (1.0f - sw) * sqrtf(maxf_(0.0f, 1.0f - a * a)) + sw * b;
  const f3_t nh = addf3(addf3(sclf3(t1, a), sclf3(t2, b)), sclf3(vh, sqrtf(maxf_(0.0f, (1.0f - a * a) - b * b))));
  return safe_normalize_f(F3f(ax * nh.x, ay * nh.y, maxf_(1e-6f, nh.z)));
}

/* A material from its 24-float table record (see oracle.material_table). */
static void mat_from(const float *q, mat_t *m) {
  m->base = D3(q[0], q[1], q[2]);
  m->rough = q[3];
  m->metal = q[4];
  m->emit = D3(q[5], q[6], q[7]);
  m->two_sided = q[8] > 0.5f;
  m->cc = q[9];
  m->cc_rough = q[10];
  m->sheen = q[11];
  m->sheen_rough = q[12];
  m->sheen_color = D3(q[13], q[14], q[15]);
  m->aniso = q[16];
  m->aniso_c = q[17];
  m->aniso_s = q[18];
  m->spec_color = D3(q[19], q[20], q[21]);
  m->spec_int = q[22];
}

/* The opaque branch of sampleNextBounceDirectionWithClearcoatNormal
 * (bsdf.wgsl.ts:2920-2997): the base lobes or an extension lobe (clearcoat,
 * sheen) by their represented probabilities, then the direction in f32
 * (glossyReflectionSample and its anisotropic form, the cosine lobe,
 * charlieSheenSample; bsdf.wgsl.ts:640-691, 1648-1778): it is the next ray's
 * direction.  sp: the specular probability at the vertex. */
static f3_t bsdf_sample_f32(const mat_t *m, f3_t n32, f3_t wo32, double sp, orng *rng, unsigned *lobe_out) {
  double ep[4];
  ext_probs(m, ep);
  f3_t tb, bb, wi32;
  build_onb_f(n32, &tb, &bb);
  const float ac = (float)m->aniso_c, as = (float)m->aniso_s;
  if (m->aniso > 0) {
    const f3_t t0 = tb, b0 = bb;
    tb = addf3(sclf3(t0, ac), sclf3(b0, as));
    bb = addf3(sclf3(t0, -as), sclf3(b0, ac));
  }
  const f3_t inc = F3f(-wo32.x, -wo32.y, -wo32.z);
  unsigned lobe; /* 0 diffuse, 1 specular, 2 clearcoat, 3 sheen */
  if (rand01(rng) < ep[0]) {
    lobe = rand01(rng) < sp ? 1u : 0u;
    if (lobe == 1u) {
      const float rough = (float)m->rough, alpha = maxf_(rough * rough, 1e-3f);
      float ax = alpha, ay = alpha;
      if (m->aniso > 0) {
        const float aspect = sqrtf(maxf_(1.0f - 0.9f * (float)m->aniso, 1e-4f));
        ax = maxf_(alpha / aspect, 1e-4f);
        ay = maxf_(alpha * aspect, 1e-4f);
      }
      const f3_t wl = F3f(dotf3(wo32, tb), dotf3(wo32, bb), dotf3(wo32, n32));
      const f3_t hl = vndf_sample_f(wl, ax, ay, rng);
      const f3_t hw = safe_normalize_f(addf3(addf3(sclf3(tb, hl.x), sclf3(bb, hl.y)), sclf3(n32, hl.z)));
      wi32 = safe_normalize_f(subf3(inc, sclf3(hw, 2.0f * dotf3(hw, inc))));
    } else {
      f3_t tc, bc;
      build_onb_f(n32, &tc, &bc);
      const double u1 = rand01(rng), u2 = rand01(rng);
      double a, b;
      disc_polar(u1, u2, &a, &b);
      const float ct = sqrtf(maxf_(0.0f, 1.0f - (float)u1));
      wi32 = safe_normalize_f(addf3(addf3(sclf3(tc, (float)a), sclf3(bc, (float)b)), sclf3(n32, ct)));
    }
  } else if (rand01(rng) < ep[3]) {
    lobe = 2u;
    f3_t tc, bc;
    build_onb_f(n32, &tc, &bc);
    const float cr = (float)m->cc_rough, alpha = maxf_(cr * cr, 1e-3f);
    const f3_t wl = F3f(dotf3(wo32, tc), dotf3(wo32, bc), dotf3(wo32, n32));
    const f3_t hl = vndf_sample_f(wl, alpha, alpha, rng);
    const f3_t hw = safe_normalize_f(addf3(addf3(sclf3(tc, hl.x), sclf3(bc, hl.y)), sclf3(n32, hl.z)));
    wi32 = safe_normalize_f(subf3(inc, sclf3(hw, 2.0f * dotf3(hw, inc))));
  } else {
    lobe = 3u;
    const double u1 = rand01(rng), u2 = rand01(rng);
    const float sr = (float)m->sheen_rough, alpha = maxf_(sr * sr, 1e-3f);
    const float inv = 1.0f / maxf_(alpha, 1e-4f);
    const float sin_h = (float)pow(u1, 1.0 / (double)(inv + 2.0f));
    const float cos_h = sqrtf(maxf_(0.0f, 1.0f - sin_h * sin_h));
    double cd, sd;
    disc_polar(1.0, u2, &cd, &sd);
    const f3_t hw = safe_normalize_f(addf3(addf3(sclf3(tb, sin_h * (float)cd), sclf3(bb, sin_h * (float)sd)),
                                           sclf3(n32, cos_h)));
    wi32 = safe_normalize_f(subf3(inc, sclf3(hw, 2.0f * dotf3(hw, inc))));
  }
  *lobe_out = lobe;
  return wi32;
}

typedef struct {
  int hit;
  uint32_t inst, prim;
  d3 pos, normal;  /* binary64 hit and the shading normal (nrm32 widened) */
  int front;
  float tpos[3];   /* the traversal's f32 world hit position */
  f3_t nrm32;      /* f32 shading normal oriented to the front face */
} surf_t;

/* Closest hit (traceClosest, intersection.wgsl.ts:462-490: back faces of
 * single-sided materials are skipped) and binary64 surface at the hit. */
static int closest(const rctx *c, d3 o, d3 d, surf_t *s) {
  orc_ray r = mk_ray(o, d, c->tmin, F32_MAX);
  orc_hit h;
  float wpos[3];
  orc_trace_closest_sided(c->d->scene, c->single_sided, &r, 1, &h, wpos, 1);
  if (h.inst == 0xFFFFFFFFu) {
    s->hit = 0;
    return 0;
  }
  const orc_scene *sc = c->d->scene;
  const float *m = sc->l2w + 16 * (size_t)h.inst;
  const float *w2l = sc->w2l + 16 * (size_t)h.inst;
  const uint32_t *tri = sc->indices + 4 * (size_t)h.prim;
  d3 P[3], N[3];
  for (int v = 0; v < 3; ++v) {
    const float *p = sc->positions + 4 * (size_t)tri[v];
    double x = p[0], y = p[1], z = p[2];
    double q[4];
    for (int k = 0; k < 4; ++k)
      q[k] = (double)m[k] * x + (double)m[4 + k] * y + (double)m[8 + k] * z + (double)m[12 + k];
    P[v] = D3(q[0] / q[3], q[1] / q[3], q[2] / q[3]);
    const float *nn = c->d->normals + 4 * (size_t)tri[v];
    N[v] = D3(nn[0], nn[1], nn[2]);
  }
  /* binary64 Moller-Trumbore on the world triangle, ray as traced (f32). */
  d3 O = D3(r.o[0], r.o[1], r.o[2]), Dd = D3(r.d[0], r.d[1], r.d[2]);
  d3 e1 = sub(P[1], P[0]), e2 = sub(P[2], P[0]);
  d3 pv = cross(Dd, e2);
  double det = dot(e1, pv);
  double u = h.u, v = h.v;
  if (det != 0) {
    d3 tv = sub(O, P[0]);
    u = dot(tv, pv) / det;
    d3 qv = cross(tv, e1);
    v = dot(Dd, qv) / det;
  }
  s->hit = 1;
  s->inst = h.inst;
  s->prim = h.prim;
  s->tpos[0] = wpos[0];
  s->tpos[1] = wpos[1];
  s->tpos[2] = wpos[2];
  s->pos = add(P[0], add(scl(e1, u), scl(e2, v)));
  d3 nl = safe_normalize(add(add(scl(N[0], 1 - u - v), scl(N[1], u)), scl(N[2], v)));
  /* transformNormalFromWorldToLocalCols: dot with each W2L column */
  d3 nw = safe_normalize(D3(w2l[0] * nl.x + w2l[1] * nl.y + w2l[2] * nl.z,
                            w2l[4] * nl.x + w2l[5] * nl.y + w2l[6] * nl.z,
                            w2l[8] * nl.x + w2l[9] * nl.y + w2l[10] * nl.z));
  (void)nw;
  /* f32 shading normal and front face, as the reference computes them
   * (shadePrologue.wgsl.ts:55-130): barycentric blend of the local vertex
   * normals, the inverse-transpose (W2L columns) to world, and the front face
   * from the local geometric normal against the local ray direction, flipped
   * by a mirroring instance (intersectionCore.wgsl.ts:96-112). */
  {
    const float bu = h.u, bv = h.v, ba = (1.0f - bu) - bv;
    f3_t nn[3], pp[3];
    for (int v = 0; v < 3; ++v) {
      const float *q = c->d->normals + 4 * (size_t)tri[v];
      const float *pq = sc->positions + 4 * (size_t)tri[v];
      nn[v] = F3f(q[0], q[1], q[2]);
      pp[v] = F3f(pq[0], pq[1], pq[2]);
    }
    const f3_t nl = safe_normalize_f(addf3(addf3(sclf3(nn[0], ba), sclf3(nn[1], bu)), sclf3(nn[2], bv)));
    const f3_t nwf = safe_normalize_f(F3f(dotf3(F3f(w2l[0], w2l[1], w2l[2]), nl),
                                          dotf3(F3f(w2l[4], w2l[5], w2l[6]), nl),
                                          dotf3(F3f(w2l[8], w2l[9], w2l[10]), nl)));
    const f3_t ng = crossf3(subf3(pp[1], pp[0]), subf3(pp[2], pp[0]));
    const f3_t df = F3f(r.d[0], r.d[1], r.d[2]);
    const f3_t ld = F3f(dotf3(F3f(w2l[0], w2l[4], w2l[8]), df), dotf3(F3f(w2l[1], w2l[5], w2l[9]), df),
                        dotf3(F3f(w2l[2], w2l[6], w2l[10]), df));
    const int mirrored = orc_orientation_sign(m) < 0;
    s->front = (dotf3(ld, ng) < 0.0f) != mirrored;
    s->nrm32 = s->front ? nwf : F3f(-nwf.x, -nwf.y, -nwf.z);
    s->normal = D3f(s->nrm32);
  }
  return 1;
}

static int find_emitter(const rctx *c, uint32_t inst, uint32_t prim) {
  /* emitters are sorted by (inst, prim) in an index built at setup */
  int lo = 0, hi = (int)c->d->emitter_count - 1;
  const emitter_t *e = c->em;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    uint64_t km = ((uint64_t)e[mid].inst << 32) | e[mid].prim, k = ((uint64_t)inst << 32) | prim;
    if (km == k) return mid;
    if (km < k) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* Records, per path, the discrete decisions that drive the estimator. */
static uint64_t mix64(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h * 0xBF58476D1CE4E5B9ull;
}

/* Mixes one decision into the signature; with `ev` non-NULL also logs it
   (ev[0] = count, ev[1..ORC_EVENT_SLOTS-1] = decisions). */
#define ORC_EVENT_SLOTS 40
static uint64_t mixlog(uint64_t h, uint64_t v, uint64_t *ev) {
  if (ev && ev[0] + 1 < ORC_EVENT_SLOTS) ev[1 + ev[0]++] = v;
  return mix64(h, v);
}

static void render_path(const rctx *c, uint32_t px, uint32_t py, uint32_t frame_index, double *out,
                        uint64_t *sig_out, const int *sorted_to_orig, uint64_t *ev, double *aux) {
  const orc_render_desc *d = c->d;
  orng rng;
  memset(&rng, 0, sizeof rng);
  if (d->sampling == 1) {
    /* pcgInit of the Sobol mode (common.wgsl.ts:130-150) */
    const uint32_t key = orc_sobol_frame_key(d->frame_seed, frame_index);
    const uint32_t pixel_seed = orc_sobol_hash(orc_sobol_combine(orc_sobol_hash(px), py));
    rng.sample = key & 0xFFFFu;
    rng.dim = 0;
    rng.tile = (px & 7u) | ((py & 7u) << 3);
    rng.seed = orc_sobol_hash(orc_sobol_combine(pixel_seed, key >> 16));
    rng.s = orc_sobol_hash(orc_sobol_combine(
        orc_sobol_combine(orc_sobol_combine(pixel_seed, key >> 16), key & 0xFFFFu), 512u));
  } else {
    rng.dim = ORC_PCG_ONLY;
    rng.s = orc_pcg_init(px, py, d->frame_seed ^ frame_index);
  }
  const float jx = (float)rand01(&rng), jy = (float)rand01(&rng);
  d3 o, dir;
  d3 L = D3(0, 0, 0), T = D3(1, 1, 1);
  uint64_t sig = 0x1234;
  if (!primary_ray(d->inv_view_proj, px, py, d->width, d->height, jx, jy, &o, &dir)) {
    out[0] = out[1] = out[2] = 0;
    *sig_out = sig;
    return;
  }
  int prev_mis = 0;
  uint32_t limit = d->max_bounces < 1 ? 1 : (d->max_bounces > 8 ? 8 : d->max_bounces);
  uint32_t n_light = d->n_dir + d->n_point + d->n_spot + d->n_rect + d->emitter_count + (c->env.present ? 1u : 0u);
  for (uint32_t bounce = 0; bounce < limit; ++bounce) {
    surf_t s;
    closest(c, o, dir, &s);
    sig = mixlog(sig, s.hit ? (((uint64_t)s.inst << 32) | s.prim) : 0xFFFFFFFFFFFFull, ev);
    if (!s.hit) {
      if (!prev_mis && c->env.present) L = add(L, mul(T, env_rgb(env_lookup(&c->env, dir))));
      if (aux && bounce == 0) {
        const double miss_aux[8] = {0.5, 1.0, 0.5, 0.0, 0.0, 0.0, 0.0, 0.0};
        memcpy(aux, miss_aux, sizeof miss_aux);
      }
      break;
    }
    const mat_t *m = &c->mats[d->tri_material[s.prim]];
    /* shading runs at the traversal's f32 hit point with the f32 normal */
    d3 n = s.normal, wo = scl(dir, -1), hp = D3(s.tpos[0], s.tpos[1], s.tpos[2]);
    const f3_t hp32 = F3f(s.tpos[0], s.tpos[1], s.tpos[2]), n32 = s.nrm32;
    const float bias32 = (float)c->bias;
    if (aux && bounce == 0) {
      /* firstHit* (shadePrologue.wgsl.ts:173-179), f32 as the reference */
      const f3_t dv = subf3(hp32, F3f((float)o.x, (float)o.y, (float)o.z));
      aux[0] = n32.x * 0.5f + 0.5f;
      aux[1] = n32.y * 0.5f + 0.5f;
      aux[2] = n32.z * 0.5f + 0.5f;
      aux[3] = sqrtf(dotf3(dv, dv));
      aux[4] = m->base.x;
      aux[5] = m->base.y;
      aux[6] = m->base.z;
      aux[7] = 1.0;
    }
    if (!prev_mis && (s.front || m->two_sided)) L = add(L, mul(T, m->emit));
    /* NEE: one light picked by the light tree (one 24-bit draw when the
     * root is interior), or uniformly when the tree is off or declines
     * (sampleCanonicalDirectLight, mediumNee.wgsl.ts:14-38). */
    const uint32_t n_ana = d->n_dir + d->n_point + d->n_spot + d->n_rect;
    if (n_light > 0) {
      uint32_t picked = 0xFFFFFFFFu;
      double inv_sel = 0;
      if (d->light_tree_nodes) {
        const float *t = d->light_tree;
        uint32_t bucket = 0;
        if ((int)t[2] >= 0 && (int)t[3] >= 0) bucket = pcg_next(&rng) >> 8;
        /* the reference evaluates the tree at the traversal's f32 hit
         * position (hit.position of traceClosest) */
        const float *p = s.tpos;
        float pdf;
        const int32_t e = orc_light_tree_sample(t, d->light_tree_nodes, p, 1e-3f, bucket, &pdf);
        if (e >= 0 && pdf > 0 && (uint32_t)e < n_light) {
          picked = (uint32_t)e;
          inv_sel = 1.0 / (double)pdf;
        }
      }
      if (picked == 0xFFFFFFFFu) {
        picked = rand_bounded(&rng, n_light);
        inv_sel = (double)n_light;
      }
      sig = mixlog(sig, 0xA000000000ull | picked, ev);
      if (picked < n_ana) {
        const float *q = d->lights + 16 * (size_t)picked;
        if (picked < d->n_dir) {
          /* directional: irradiance along a (cone-sampled) direction, no MIS */
          d3 wi = sample_directional_cone(D3(q[0], q[1], q[2]), q[3], &rng);
          double ndl = fabs(dot(n, wi));
          if (q[7] > 0.5f || visible(c, shadow_origin(hp32, n32, dot(n, wi) > 0, bias32), wi, F32_MAX))
            L = add(L, scl(mul(mul(T, bsdf_eval(m, n, wo, wi)), D3(q[4], q[5], q[6])), ndl * inv_sel));
        } else if (picked < d->n_dir + d->n_point + d->n_spot) {
          /* point / spot: delta position with the authored decay and cutoff */
          int spot = picked >= d->n_dir + d->n_point;
          double cutoff = q[spot ? 12 : 8], decay = q[spot ? 13 : 9];
          d3 tl = sub(D3(q[0], q[1], q[2]), hp);
          double dist = sqrt(dot(tl, tl));
          if (dist > 0 && !(cutoff > 0 && dist > cutoff)) {
            d3 wi = scl(tl, 1.0 / dist);
            double soft = 1.0;
            int lit = 1;
            if (spot) {
              d3 axis = safe_normalize(D3(q[4], q[5], q[6]));
              double co = q[7], ci = q[11], cc = dot(scl(wi, -1), axis);
              lit = cc >= co;
              soft = smoothstep_d(co, maxd(ci, co + 1e-6), cc);
            }
            double tmax = maxd(dist - 2.0 * c->bias, c->tmin);
            if (lit && (q[3] > 0.5f || visible(c, shadow_origin(hp32, n32, dot(n, wi) > 0, bias32), wi, tmax))) {
              double ndl = fabs(dot(n, wi));
              d3 rad = spot ? D3(q[8], q[9], q[10]) : D3(q[4], q[5], q[6]);
              double att = point_attenuation(dist, cutoff, decay);
              L = add(L, scl(mul(mul(T, bsdf_eval(m, n, wo, wi)), rad), ndl * soft * att * inv_sel));
            }
          }
        } else {
          /* rect / disc area: a point on the emitter, MIS against the BSDF */
          d3 cc = D3(q[0], q[1], q[2]), u = D3(q[4], q[5], q[6]), v = D3(q[8], q[9], q[10]);
          int disc = fabs(q[15] - 1.0f) < 0.5f;
          d3 ln;
          double area;
          int ok = analytic_area(u, v, disc, &ln, &area);
          double xi1 = rand01(&rng), xi2 = rand01(&rng), a, b;
          if (disc) {
            concentric_disc(xi1 * 2.0 - 1.0, xi2 * 2.0 - 1.0, &a, &b);
          } else {
            a = xi1 * 2.0 - 1.0;
            b = xi2 * 2.0 - 1.0;
          }
          d3 lp = add(add(cc, scl(u, a)), scl(v, b));
          d3 tl = sub(lp, hp);
          double dist = sqrt(dot(tl, tl));
          d3 wi = safe_normalize(tl);
          double ndl = fabs(dot(n, wi));
          if (dist > 0 && ndl > 0 && ok) {
            double cl = maxd(dot(ln, scl(wi, -1)), 0);
            double lpdf = cl > 0 ? area_to_solid_angle(dist, cl, area) : 0;
            if (lpdf > 0) {
              double w = power_heuristic(lpdf, bsdf_pdf(m, n, wo, wi));
              double tmax = maxd(dist - 2.0 * c->bias, c->tmin);
              if (q[3] > 0.5f || visible(c, shadow_origin(hp32, n32, dot(n, wi) > 0, bias32), wi, tmax))
                L = add(L, scl(mul(mul(T, bsdf_eval(m, n, wo, wi)), D3(q[12], q[13], q[14])), ndl * w / lpdf * inv_sel));
            }
          }
        }
      } else if (picked - n_ana < d->emitter_count) {
        const emitter_t *e = &c->em[sorted_to_orig[picked - n_ana]];
        double r1 = rand01(&rng), r2 = rand01(&rng);
        double su = sqrt(r1), uu = 1 - su, vv = r2 * su, ww = 1 - uu - vv;
        d3 lp = add(add(scl(e->a, uu), scl(e->b, vv)), scl(e->c, ww));
        d3 tl = sub(lp, hp);
        double dist = sqrt(dot(tl, tl));
        d3 ln;
        double area;
        int ok = area_measure(sub(e->b, e->a), sub(e->c, e->a), &ln, &area);
        d3 wi = safe_normalize(tl);
        double ndl = fabs(dot(n, wi));
        if (dist > 0 && ndl > 0 && ok) {
          d3 f = bsdf_eval(m, n, wo, wi);
          double cl = dot(ln, scl(wi, -1));
          cl = e->two_sided ? fabs(cl) : maxd(cl, 0);
          if (cl > 0) {
            double lpdf = area_to_solid_angle(dist, cl, area);
            if (lpdf > 0) {
              double bpdf = bsdf_pdf(m, n, wo, wi);
              double w = power_heuristic(lpdf, bpdf);
              double tmax = maxd(dist - 2.0 * c->bias, c->tmin);
              if (visible(c, shadow_origin(hp32, n32, dot(n, wi) > 0, bias32), wi, tmax)) {
                double k = ndl * w / lpdf * inv_sel;
                L = add(L, scl(mul(mul(T, f), e->le), k));
              }
            }
          }
        }
      } else {
        double xi = rand01(&rng);
        uint32_t count = c->env.w * c->env.h, lo = 0, hi = count;
        while (lo + 1 < hi) {
          uint32_t mid = (lo + hi) >> 1;
          if ((double)c->env.cdf[mid] <= xi) lo = mid; else hi = mid;
        }
        uint32_t idx = lo < count - 1 ? lo : count - 1;
        double xa = rand01(&rng), xb = rand01(&rng);
        d3 wi = safe_normalize(env_to_world(&c->env, env_cell_dir(idx % c->env.w, idx / c->env.w, c->env.w,
                                                                  c->env.h, xa, xb)));
        const float *et = c->env.texels + 4 * (size_t)idx;
        double epdf = (double)et[3];
        double ndl = fabs(dot(n, wi));
        if (ndl > 0) {
          if (visible(c, shadow_origin(hp32, n32, dot(n, wi) > 0, bias32), wi, F32_MAX)) {
            d3 f = bsdf_eval(m, n, wo, wi);
            double bpdf = bsdf_pdf(m, n, wo, wi);
            double w = power_heuristic(epdf, bpdf);
            double k = ndl * w / maxd(epdf, 1e-8) * inv_sel;
            L = add(L, scl(mul(mul(T, f), env_rgb(et)), k));
          }
        }
      }
    }
    /* BSDF sample, opaque branch (bsdf.wgsl.ts:2920-2997): base lobes or
     * an extension lobe (clearcoat, sheen) by their represented
     * probabilities; the sampled direction in f32 (glossyReflectionSample
     * and its anisotropic form, the cosine lobe, charlieSheenSample,
     * bsdf.wgsl.ts:640-691, 1648-1778): it is the next ray's direction. */
    double ndv = maxd(dot(n, wo), 0);
    double sp = spec_prob(m, ndv);
    unsigned lobe;
    const f3_t wi32 = bsdf_sample_f32(m, n32, F3f(-(float)dir.x, -(float)dir.y, -(float)dir.z), sp, &rng, &lobe);
    d3 wi = D3f(wi32);
    sig = mixlog(sig, 0xB0 + lobe, ev);
    /* finalizeFiniteBounceSample (bsdf.wgsl.ts:1863-1926) */
    double mpdf = bsdf_pdf(m, n, wo, wi);
    d3 f = bsdf_eval(m, n, wo, wi);
    double cosine = fabs(dot(n, wi));
    int allows = 1;
    d3 tmul;
    if (mpdf <= 0 || cosine <= 1e-8) {
      tmul = D3(0, 0, 0);
      allows = 0;
      mpdf = 0;
    } else {
      tmul = scl(f, cosine / mpdf);
    }
    d3 T_at = T;
    T = mul(T, tmul);
    d3 new_o = D3f(addf3(hp32, sclf3(n32, bias32)));
    /* BSDF -> area light and environment connections (connect.wgsl.ts:478-700). */
    if (allows) {
      double ndl = fabs(dot(n, wi));
      double bpdf = mpdf;
      if (ndl > 0 && bpdf > 0) {
        d3 oo = shadow_origin(hp32, n32, dot(n, wi) > 0, bias32);
        double best_t = F32_MAX, best_pdf = 0;
        d3 best_le = D3(0, 0, 0);
        /* analytic rect and disc lights first (connect.wgsl.ts:544-572) */
        for (uint32_t k = 0; k < d->n_rect; ++k) {
          const float *q = d->lights + 16 * (size_t)(d->n_dir + d->n_point + d->n_spot + k);
          double t, lpdf;
          if (!intersect_rect_light(q, oo, wi, c->tmin, &t, &lpdf)) continue;
          if (!(q[3] > 0.5f) && !visible(c, oo, wi, maxd(t - 2.0 * c->bias, c->tmin))) continue;
          if (t < best_t) {
            best_t = t;
            best_pdf = lpdf;
            best_le = D3(q[12], q[13], q[14]);
          }
        }
        for (uint32_t li = 0; li < d->emitter_count; ++li) {
          const emitter_t *e = &c->em[sorted_to_orig[li]];
          /* intersectMeshAreaLightRay (connect.wgsl.ts:453-476), binary64 */
          d3 e1 = sub(e->b, e->a), e2 = sub(e->c, e->a);
          d3 pv = cross(wi, e2);
          double det = dot(e1, pv);
          if (det == 0) continue;
          d3 tv = sub(oo, e->a);
          double u = dot(tv, pv) / det;
          if (u < 0 || u > 1) continue;
          d3 qv = cross(tv, e1);
          double v = dot(wi, qv) / det;
          if (v < 0 || u + v > 1) continue;
          double t = dot(e2, qv) / det;
          if (!(t > c->tmin) || t >= F32_MAX) continue;
          d3 ln;
          double area;
          if (!area_measure(e1, e2, &ln, &area)) continue;
          double cl = dot(ln, scl(wi, -1));
          cl = e->two_sided ? fabs(cl) : maxd(cl, 0);
          if (cl <= 0) continue;
          double lpdf = area_to_solid_angle(t, cl, area);
          if (!(lpdf > 0)) continue;
          if (!(t < best_t)) continue;
          if (!visible(c, oo, wi, maxd(t - 2.0 * c->bias, c->tmin))) continue;
          best_t = t;
          best_pdf = lpdf;
          best_le = e->le;
        }
        if (best_t < F32_MAX && best_pdf > 0) {
          double w = power_heuristic(bpdf, best_pdf);
          L = add(L, scl(mul(mul(T_at, f), best_le), ndl * w / bpdf));
        }
        if (c->env.present && visible(c, oo, wi, F32_MAX)) {
          const float *et = env_lookup(&c->env, wi);
          double w = power_heuristic(bpdf, (double)et[3]);
          L = add(L, scl(mul(mul(T_at, f), env_rgb(et)), ndl * w / bpdf));
        }
      }
    }
    prev_mis = allows && mpdf > 0;
    /* A rejected sample leaves zero throughput.  The reference keeps tracing
       such a path (kernel.wgsl.ts:1993-2013), but every later term it adds is
       multiplied by that zero, so ending it here changes no radiance value. */
    if (!allows) break;
    o = new_o;
    dir = wi;
    if (bounce > 2) {
      double surv = represented(clampd(maxd(T.x, maxd(T.y, T.z)), 0.1, 0.95));
      int lives = rand01(&rng) < surv;
      sig = mixlog(sig, lives ? 0xC1 : 0xC0, ev);
      if (!lives) break;
      T = scl(T, 1.0 / surv);
    }
  }
  out[0] = maxd(L.x, 0);
  out[1] = maxd(L.y, 0);
  out[2] = maxd(L.z, 0);
  *sig_out = sig;
}

typedef struct {
  const rctx *c;
  const int *sorted_to_orig;
  uint32_t frame_index, x0, y0, w, h, row_begin, row_end;
  double *out;
  uint64_t *sig;
  uint64_t *events; /* NULL, or ORC_EVENT_SLOTS per pixel */
} rjob;

static void *render_rows(void *arg) {
  rjob *j = (rjob *)arg;
  for (uint32_t yy = j->row_begin; yy < j->row_end; ++yy)
    for (uint32_t xx = 0; xx < j->w; ++xx) {
      size_t i = (size_t)yy * j->w + xx;
      render_path(j->c, j->x0 + xx, j->y0 + yy, j->frame_index, j->out + 3 * i, j->sig + i,
                  j->sorted_to_orig, j->events ? j->events + ORC_EVENT_SLOTS * i : NULL,
                  j->c->d->aux ? j->c->d->aux + 8 * i : NULL);
    }
  return NULL;
}

static const emitter_t *g_em_sort;
static int cmp_em(const void *a, const void *b) {
  const emitter_t *x = &g_em_sort[*(const int *)a], *y = &g_em_sort[*(const int *)b];
  uint64_t kx = ((uint64_t)x->inst << 32) | x->prim, ky = ((uint64_t)y->inst << 32) | y->prim;
  return kx < ky ? -1 : (kx > ky ? 1 : 0);
}

/* Renders one frame of the tile [x0, x0+w) x [y0, y0+h).
 * out: 3 doubles per pixel (radiance of this frame's path), sig: path
 * signature per pixel, events: NULL or ORC_EVENT_SLOTS words per pixel (count,
 * then each decision mixed into the signature, in order). */
int orc_render(const orc_render_desc *d, uint32_t frame_index, uint32_t x0, uint32_t y0, uint32_t w,
               uint32_t h, double *out, uint64_t *sig, uint64_t *events, uint32_t threads) {
  rctx c;
  memset(&c, 0, sizeof c);
  c.d = d;
  c.mats = (mat_t *)calloc(d->material_count + 1, sizeof(mat_t));
  for (uint32_t i = 0; i < d->material_count; ++i) {
    mat_from(d->materials + 24 * (size_t)i, &c.mats[i]);
  }
  /* Emitters: the record list is in light-selection order; a sorted copy
   * (by instance, triangle) serves hit lookups. */
  emitter_t *em = (emitter_t *)calloc(d->emitter_count + 1, sizeof(emitter_t));
  for (uint32_t i = 0; i < d->emitter_count; ++i) {
    const float *q = d->emitters + 16 * (size_t)i;
    em[i].a = D3(q[0], q[1], q[2]);
    em[i].b = D3(q[3], q[4], q[5]);
    em[i].c = D3(q[6], q[7], q[8]);
    em[i].le = D3(q[9], q[10], q[11]);
    em[i].two_sided = q[12] > 0.5f;
    em[i].inst = (uint32_t)q[13];
    em[i].prim = (uint32_t)q[14];
  }
  int *order = (int *)malloc(sizeof(int) * (d->emitter_count + 1));
  for (uint32_t i = 0; i < d->emitter_count; ++i) order[i] = (int)i;
  c.em = em;
  c.env.present = d->env_present != 0;
  c.env.w = d->env_w;
  c.env.h = d->env_h;
  c.env.cdf = d->env_cdf;
  c.env.texels = d->env_texels;
  c.env.rot_c = d->env_rot[0];
  c.env.rot_s = d->env_rot[1];
  c.bias = (double)d->ray_bias;
  c.tmin = (double)f32(c.bias * 0.1);
  /* materialAcceptsSidedHit (material.wgsl.ts:3259-3263): opaque materials
   * that are not double-sided have no back-face surfaces. */
  const uint32_t ntri = d->scene->triangle_count;
  c.single_sided = (uint8_t *)calloc(ntri + 1, 1);
  for (uint32_t t = 0; t < ntri; ++t) {
    const uint32_t mid = d->tri_material[t];
    c.single_sided[t] = mid < d->material_count ? (uint8_t)!c.mats[mid].two_sided : 1u;
  }
  {
    const uint64_t lim = (uint64_t)ntri * (d->scene->tlas_node_count ? d->scene->instance_count : 1u);
    c.hit_limit = lim > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)lim;
  }
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t tid[64];
  rjob jobs[64];
  uint32_t per = (h + threads - 1) / threads;
  for (uint32_t t = 0; t < threads; ++t) {
    jobs[t] = (rjob){&c, order, frame_index, x0, y0, w, h, t * per < h ? t * per : h,
                     (t + 1) * per < h ? (t + 1) * per : h, out, sig, events};
    if (pthread_create(&tid[t], NULL, render_rows, &jobs[t]) != 0) {
      render_rows(&jobs[t]);
      tid[t] = 0;
    }
  }
  for (uint32_t t = 0; t < threads; ++t)
    if (tid[t]) pthread_join(tid[t], NULL);
  free(c.mats);
  free(c.single_sided);
  free(em);
  free(order);
  (void)cmp_em;
  (void)find_emitter;
  (void)mind;
  return 0;
}

/* ---- BSDF test hooks ------------------------------------------------------
 * Over a 24-float material record: the evaluation (the kernel without N.L),
 * the directional density, and one sampled direction from a PCG stream
 * (the state advances). */
void orc_bsdf_eval(const float *mat, const double *n, const double *wo, const double *wi, double *out) {
  mat_t m;
  mat_from(mat, &m);
  const d3 f = bsdf_eval(&m, D3(n[0], n[1], n[2]), D3(wo[0], wo[1], wo[2]), D3(wi[0], wi[1], wi[2]));
  out[0] = f.x, out[1] = f.y, out[2] = f.z;
}
double orc_bsdf_pdf(const float *mat, const double *n, const double *wo, const double *wi) {
  mat_t m;
  mat_from(mat, &m);
  return bsdf_pdf(&m, D3(n[0], n[1], n[2]), D3(wo[0], wo[1], wo[2]), D3(wi[0], wi[1], wi[2]));
}
void orc_bsdf_sample(const float *mat, const float *n, const float *wo, uint32_t *pcg_state, float *wi,
                     uint32_t *lobe) {
  mat_t m;
  mat_from(mat, &m);
  orng r = {*pcg_state, ORC_PCG_ONLY, 0, 0, 0};
  const d3 nd = D3(n[0], n[1], n[2]);
  const double sp = spec_prob(&m, maxd(dot(nd, D3(wo[0], wo[1], wo[2])), 0));
  unsigned lb = 0;
  const f3_t w = bsdf_sample_f32(&m, F3f(n[0], n[1], n[2]), F3f(wo[0], wo[1], wo[2]), sp, &r, &lb);
  wi[0] = w.x, wi[1] = w.y, wi[2] = w.z;
  *lobe = lb;
  *pcg_state = r.s;
}
