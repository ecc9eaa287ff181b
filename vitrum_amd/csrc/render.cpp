// render.cpp -- host side of the wavefront path tracer: per-scene shading
// data (materials, mesh-area emitters, environment importance table),
// per-resolution path state, and the frame schedule behind vt_render_frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "emitters.h"
#include "engine_internal.h"
#include "light_tree.h"
#include "render.h"
#include "sky.h"
#include "sobol.h"

namespace {

vt_status rerr(vt_engine* e, vt_status code, const std::string& msg) {
  e->last_error = msg;
  return code;
}

#define HIP_CHECK(e, expr)                                                                 \
  do {                                                                                     \
    hipError_t _r = (expr);                                                                \
    if (_r != hipSuccess) return rerr((e), VT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_r)); \
  } while (0)

// buildRepresentedDistributionF32 (packages/shared-samplers/src/
// representedDistribution.ts:44-175): Hamilton apportionment of 2^24
// buckets over the positive weights, then an f32 cumulative table.
std::vector<float> represented_cdf(const std::vector<double>& w) {
  const int64_t total = int64_t(1) << 24;
  std::vector<int64_t> buckets(w.size(), 0);
  std::vector<size_t> pos;
  for (size_t i = 0; i < w.size(); ++i)
    if (w[i] > 0) pos.push_back(i);
  if ((int64_t)pos.size() > total)
    throw std::range_error("the importance table has more positive entries than its 2^24 random buckets");
  if (!pos.empty()) {
    double mx = 0;
    for (size_t i : pos) mx = std::max(mx, w[i]);
    double sum = 0, comp = 0;
    for (size_t i : pos) {
      const double corr = w[i] / mx - comp;
      const double next = sum + corr;
      comp = (next - sum) - corr;
      sum = next;
    }
    const int64_t remaining = total - (int64_t)pos.size();
    int64_t allocated = (int64_t)pos.size();
    std::vector<std::pair<size_t, double>> rem;
    for (size_t i : pos) {
      const double quota = remaining == 0 ? 0.0 : ((w[i] / mx) / sum) * (double)remaining;
      const int64_t extra = (int64_t)std::floor(quota);
      buckets[i] = 1 + extra;
      allocated += extra;
      rem.push_back({i, quota - (double)extra});
    }
    std::stable_sort(rem.begin(), rem.end(), [](const auto& a, const auto& b) {
      return a.second != b.second ? a.second > b.second : a.first < b.first;
    });
    int64_t deficit = total - allocated;
    for (size_t c = 0; deficit > 0; ++c, --deficit) buckets[rem[c % rem.size()].first] += 1;
    if (deficit < 0) {
      auto asc = rem;
      std::stable_sort(asc.begin(), asc.end(), [](const auto& a, const auto& b) {
        return a.second != b.second ? a.second < b.second : a.first > b.first;
      });
      int64_t excess = -deficit;
      for (size_t c = 0; excess > 0; ++c) {
        const size_t i = asc[c % asc.size()].first;
        if (buckets[i] <= 1) continue;
        buckets[i] -= 1;
        --excess;
      }
    }
  }
  std::vector<float> cdf(w.size() + 1, 0.0f);
  int64_t cum = 0;
  for (size_t i = 0; i < w.size(); ++i) {
    cum += buckets[i];
    cdf[i + 1] = (float)((double)cum / (double)total);
  }
  return cdf;
}

// The environment's importance table and texels (environmentPacking.ts:
// 283-440): texel weights luminance x solid angle of the unscaled map, an
// f32 CDF over 2^24 represented buckets, and per texel the intensity-scaled
// radiance with the represented pmf over the texel's solid angle.
struct EnvTable {
  bool present = false;
  uint32_t w = 0, h = 0;
  float rot_c = 1.0f, rot_s = 0.0f;
  std::vector<float> cdf;     // w * h + 1
  std::vector<float> texels;  // w * h x (r, g, b, pdf per steradian)
  double power = 0;           // luminance x solid angle of the scaled map
};

double luminance(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }

}  // namespace

// The procedural sky's parameters as the reference defaults and packs them
// (environmentPacking.ts:237-262): intensity through the binary32 lane check
// (environmentRadianceScale.ts:81-98), then the Preetham bake.
std::string engine_bake_sky(const vt_environment& env, uint32_t W, uint32_t H, std::vector<float>& rgb,
                            double* integral) {
  const double in = env.intensity;
  if (!std::isfinite(in) || in < 0) return "procedural-sky intensity must be finite and non-negative";
  vts::SkyParams p{};
  for (int k = 0; k < 3; ++k) p.sun[k] = env.sun_direction[k];
  // the sun direction is normalised here and again by the bake
  const double len = std::hypot(p.sun[0], p.sun[1], p.sun[2]);
  if (std::isfinite(len)) {
    if (len < 1e-8)
      p.sun[0] = 0, p.sun[1] = 1, p.sun[2] = 0;
    else
      for (int k = 0; k < 3; ++k) p.sun[k] /= len;
  }
  p.turbidity = env.turbidity;
  p.rayleigh = env.rayleigh;
  p.mie_coefficient = env.mie_coefficient;
  p.mie_g = env.mie_directional_g;
  p.intensity = in;
  return vts::bake_preetham(p, W, H, rgb, integral);
}

namespace {

// Empty string on success (t.present = false for none or an all-black map).
std::string build_environment(const vt_environment& env, EnvTable& t) {
  if (env.kind == VT_ENV_NONE) return {};
  uint32_t W = 64, H = 32, ch = 0;
  float intensity = 1.0f, rot = 0.0f;
  const float* texels = env.texels;
  std::vector<float> sky;
  if (env.kind == VT_ENV_PROCEDURAL_SKY) {
    // the baked daylight map, sampled as an HDRI with intensity 1 and no
    // rotation (the sun direction is in world space; environmentPacking.ts:
    // 237-282)
    double integral = 0;
    const std::string se = engine_bake_sky(env, kSkyWidth, kSkyHeight, sky, &integral);
    if (!se.empty()) return se;
    W = kSkyWidth, H = kSkyHeight, ch = 3;
    texels = sky.data();
  } else if (env.kind == VT_ENV_CONSTANT) {
    for (int k = 0; k < 3; ++k)
      if (!(std::isfinite(env.radiance[k]) && env.radiance[k] >= 0.0f))
        return "environment radiance must be finite and non-negative";
  } else if (env.kind == VT_ENV_HDRI) {
    if (!env.texels || env.width == 0 || env.height == 0) return "an HDRI environment needs texels, width and height";
    if (env.channels < 1 || env.channels > 4) return "HDRI channels must be 1, 2, 3 or 4";
    if ((uint64_t)env.width * env.height * 4u > 0x7fffffffull) return "the HDRI has too many texels";
    if (!(std::isfinite(env.intensity) && env.intensity >= 0.0f)) return "HDRI intensity must be finite and non-negative";
    if (!std::isfinite(env.rotation_y)) return "HDRI rotationY must be finite";
    W = env.width, H = env.height, ch = env.channels;
    intensity = env.intensity;
    // canonicalizeEnvironmentRotationF32 (environmentRotation.ts:10-26)
    rot = (float)std::fmod((double)env.rotation_y, 2.0 * M_PI);
    if (rot == 0.0f) rot = 0.0f;
  } else {
    return "environment kind must be 0 (none), 1 (constant) or 2 (hdri)";
  }
  const size_t n = (size_t)W * H;
  std::vector<double> weights(n), omega(H);
  const double dphi = 2.0 * M_PI / W;
  for (uint32_t y = 0; y < H; ++y)
    omega[y] = dphi * (std::cos((double)y / H * M_PI) - std::cos((double)(y + 1) / H * M_PI));
  t.texels.assign(n * 4, 0.0f);
  double total = 0, scaled_total = 0;
  for (size_t i = 0; i < n; ++i) {
    float c[3];
    if (ch == 0) {
      c[0] = env.radiance[0], c[1] = env.radiance[1], c[2] = env.radiance[2];
    } else {
      // environmentMapPixels.ts:138-142: grey, (r, g) with b = r, rgb, rgba
      const float* q = texels + i * ch;
      c[0] = q[0];
      c[1] = ch > 1 ? q[1] : q[0];
      c[2] = ch > 2 ? q[2] : q[0];
    }
    for (int k = 0; k < 3; ++k)
      if (!(std::isfinite(c[k]) && c[k] >= 0.0f))
        return "HDRI texel " + std::to_string(i) + " must contain finite, non-negative radiance";
    const double om = omega[i / W];
    float sc[3];
    for (int k = 0; k < 3; ++k) {
      sc[k] = c[k] * intensity;  // one binary32 product (ptScaleEnvironmentRadiance)
      if (!std::isfinite(sc[k]))
        return "HDRI texel " + std::to_string(i) + " multiplied by its intensity must remain finite in binary32";
      t.texels[i * 4 + k] = sc[k];
    }
    weights[i] = std::max(0.0, luminance(c[0], c[1], c[2]) * om);
    total += weights[i];
    scaled_total += std::max(0.0, luminance(sc[0], sc[1], sc[2]) * om);
  }
  if (!std::isfinite(total) || !std::isfinite(scaled_total)) return "the HDRI importance integral overflowed";
  if (!(total > 0)) {
    t.texels.clear();
    return {};
  }
  if (intensity > 0.0f && scaled_total == 0)
    return "positive HDRI radiance multiplied by its positive intensity underflows to zero in binary32";
  try {
    t.cdf = represented_cdf(weights);
  } catch (const std::exception& ex) {
    return ex.what();
  }
  for (size_t i = 0; i < n; ++i)
    t.texels[i * 4 + 3] = (float)(((double)t.cdf[i + 1] - (double)t.cdf[i]) / omega[i / W]);
  t.present = true;
  t.w = W, t.h = H;
  t.rot_c = (float)std::cos((double)rot);
  t.rot_s = (float)std::sin((double)rot);
  t.power = scaled_total;
  return {};
}

}  // namespace

// Per-scene shading data.  vt_set_scene builds and uploads a complete set
// before it binds anything, so a failure leaves the previous scene's set.
struct RenderSceneData {
  DeviceBuffer materials, emitters, instances, env_cdf, env_texels, tri_source, light_tree, lights;
  uint32_t n_dir = 0, n_point = 0, n_spot = 0, n_rect = 0;  // analytic emitters in `lights`
  uint32_t light_tree_nodes = 0;  // 0: uniform light selection
  std::vector<float> light_tree_host;
  uint32_t emitter_count = 0;
  uint32_t env_present = 0, env_w = 0, env_h = 0;
  float env_rot[2] = {1, 0};  // cos, sin of the dome's rotation about +Y
  float ray_bias = 0, tmin = 0;
  float scene_min[3] = {0, 0, 0}, scene_inv_extent[3] = {1, 1, 1};
  void release_scene() {
    for (DeviceBuffer* b : {&materials, &emitters, &instances, &env_cdf, &env_texels, &tri_source, &light_tree, &lights})
      b->release();
  }
};

struct RenderState : RenderSceneData {
  DeviceBuffer sobol_dirs;  // 512 x 16 Sobol direction components (built once)
  // Per-resolution path state.
  uint32_t width = 0, height = 0, paths = 0;
  DeviceBuffer rng, rng_dim, throughput, conn, radiance, prev_pdf, flags, signature, count;
  DeviceBuffer queue[2], rays[2], hits, hit_pos;
  DeviceBuffer shadow_path, shadow_rays, shadow_contrib, shadow_occ, accum;
  DeviceBuffer moments, aux_nd, aux_albedo, present;  // variance moments, first-hit images, capture staging
  // Neural radiance cache (used when nrc_enabled).
  bool nrc_enabled = false;
  uint32_t nrc_train_paths = 65536;
  uint32_t nrc_train_vertices = 4;  // shaded teacher vertices before the terminal query (0: full length)
  DeviceBuffer nrc_count, q_feat, q_thr, q_path, q_rgb, train_rec, train_n, train_feat, train_tgt;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Shadow rays of a bounce are traced on `side` while the main stream traces
  // the next bounce's extension rays: fork after shading (ev_fork), join
  // before the next shading (ev_join).  Default priority, like the main
  // stream.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void release_paths() {
    for (DeviceBuffer* b : {&rng, &rng_dim, &throughput, &conn, &radiance, &prev_pdf, &flags, &signature, &count,
                            &queue[0], &queue[1], &rays[0], &rays[1], &hits, &hit_pos, &shadow_path,
                            &shadow_rays, &shadow_contrib, &shadow_occ, &accum, &moments, &aux_nd, &aux_albedo,
                            &present, &nrc_count, &q_feat, &q_thr,
                            &q_path, &q_rgb, &train_rec, &train_n, &train_feat, &train_tgt})
      b->release();
    width = height = paths = 0;
  }
};

void engine_release_render(vt_engine* e) {
  if (!e->render) return;
  (void)engine_nrc_quiesce(e);
  e->render->release_paths();
  e->render->release_scene();
  e->render->sobol_dirs.release();
  if (e->render->side) (void)hipStreamSynchronize(e->render->side);
  for (hipEvent_t ev : {e->render->ev0, e->render->ev1, e->render->ev_fork, e->render->ev_join})
    if (ev) (void)hipEventDestroy(ev);
  if (e->render->side) (void)hipStreamDestroy(e->render->side);
  delete e->render;
  e->render = nullptr;
}

// ptWebgpuRayOriginBias (packages/pt-webgpu/src/scene/sceneScalePolicy.ts:83-118).
static float ray_origin_bias(double radius, const double* center) {
  double mag = std::max(std::fabs((double)(float)center[0]),
                        std::max(std::fabs((double)(float)center[1]), std::fabs((double)(float)center[2])));
  double ulp = (mag == 0 || mag < 1.1754943508222875e-38) ? std::ldexp(1.0, -149)
                                                          : std::ldexp(1.0, (int)std::floor(std::log2(mag)) - 23);
  return (float)std::max(radius * (1e-3 / std::sqrt(3.0)), 2.0 * ulp);
}

// Ray-origin bias, tMin and the normalisation box of the scene, from its
// world bounds (or the caller's radius).
static void scene_bounds(RenderSceneData& R, const vt::Pack& P, float scene_radius) {
  double center[3], radius = scene_radius;
  for (int k = 0; k < 3; ++k) center[k] = 0.5 * (P.world_min[k] + P.world_max[k]);
  if (!(radius > 0)) {
    double d2 = 0;
    for (int k = 0; k < 3; ++k) d2 += (P.world_max[k] - P.world_min[k]) * (P.world_max[k] - P.world_min[k]);
    radius = 0.5 * std::sqrt(d2);
    if (!(radius > 0)) radius = std::sqrt(3.0);
  }
  if (!std::isfinite(center[0])) center[0] = center[1] = center[2] = 0;
  for (int k = 0; k < 3; ++k) {
    const double lo = P.world_min[k], hi = P.world_max[k];
    const bool ok = std::isfinite(lo) && std::isfinite(hi) && hi > lo;
    R.scene_min[k] = ok ? (float)lo : 0.0f;
    R.scene_inv_extent[k] = ok ? (float)(1.0 / (hi - lo)) : 1.0f;
  }
  R.ray_bias = ray_origin_bias(radius, center);
  R.tmin = std::max(R.ray_bias * 0.1f, 1.175494351e-38f);
}

void engine_render_rebound(vt_engine* e) {
  if (e->render && e->pack) scene_bounds(*e->render, *e->pack, e->scene_radius);
}

static vt_status ensure_render_state(vt_engine* e) {
  if (e->render) return VT_OK;
  std::unique_ptr<RenderState> r(new RenderState());
  HIP_CHECK(e, hipEventCreate(&r->ev0));
  HIP_CHECK(e, hipEventCreate(&r->ev1));
  HIP_CHECK(e, hipEventCreateWithFlags(&r->ev_fork, hipEventDisableTiming));
  HIP_CHECK(e, hipEventCreateWithFlags(&r->ev_join, hipEventDisableTiming));
  HIP_CHECK(e, hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking));
  std::vector<uint16_t> dirs(vts::kDims * vts::kBits);
  vts::build_directions(dirs.data());
  if (vt_status s = engine_upload(e, r->sobol_dirs, dirs.data(), dirs.size() * 2)) return s;
  HIP_CHECK(e, hipStreamSynchronize(e->stream));
  e->render = r.release();
  return VT_OK;
}

void engine_discard_render(RenderSceneData* d) {
  if (!d) return;
  d->release_scene();
  delete d;
}

void engine_commit_render(vt_engine* e, RenderSceneData* d) {
  RenderSceneData& cur = *e->render;
  cur.release_scene();
  cur = *d;
  delete d;  // its device buffers now belong to the engine
  e->samples_accumulated = 0;
}

vt_status engine_prepare_render(vt_engine* e, const vt::Pack& P, const std::vector<vt_material>& materials,
                                const vt_environment& env, float scene_radius, const vt_emitter* emitters,
                                uint32_t emitter_count, RenderSceneData** out) {
  *out = nullptr;
  if (vt_status s = ensure_render_state(e)) return s;
  if (e->fail_inject & kInjectRenderThrow) {
    e->fail_inject &= ~kInjectRenderThrow;
    throw std::runtime_error("injected failure while preparing the scene's shading data");
  }
  std::unique_ptr<RenderSceneData, void (*)(RenderSceneData*)> hold(new RenderSceneData(), engine_discard_render);
  RenderSceneData& R = *hold;
  // Materials: base rgb, roughness, metallic, Le = emissive * intensity, twoSided.
  std::vector<float> mats(std::max<size_t>(1, materials.size()) * vtr::kMatFloats, 0.0f);
  for (size_t i = 0; i < materials.size(); ++i) {
    const vt_material& m = materials[i];
    float* q = &mats[i * vtr::kMatFloats];
    // Le = emissive x emissiveIntensity in binary64, rounded once to f32 (the
    // reference multiplies JavaScript numbers and packs a Float32Array)
    for (int k = 0; k < 3; ++k)
      q[k] = m.base_color[k], q[5 + k] = (float)((double)m.emissive[k] * (double)m.emissive_intensity);
    q[3] = m.roughness;
    q[4] = m.metallic;
    q[8] = m.double_sided ? 1.0f : 0.0f;
    // Extension lobes, clamped as materialPacking.ts:259-316 packs them; the
    // anisotropy rotation is stored as its cosine and sine (binary64, rounded
    // once).
    auto c01 = [](float x) { return std::min(1.0f, std::max(0.0f, x)); };
    q[9] = c01(m.clearcoat);
    q[10] = c01(m.clearcoat_roughness);
    q[11] = c01(m.sheen);
    q[12] = c01(m.sheen_roughness);
    for (int k = 0; k < 3; ++k) q[13 + k] = c01(m.sheen_color[k]);
    q[16] = c01(m.anisotropy);
    q[17] = (float)std::cos((double)m.anisotropy_rotation);
    q[18] = (float)std::sin((double)m.anisotropy_rotation);
    for (int k = 0; k < 3; ++k) q[19 + k] = m.has_specular ? std::max(0.0f, m.specular_color[k]) : 1.0f;
    q[22] = m.has_specular ? c01(m.specular_intensity) : 1.0f;
  }
  if (materials.empty()) {
    mats[0] = mats[1] = mats[2] = 0.8f, mats[3] = 1.0f;
    mats[17] = 1.0f;
    mats[19] = mats[20] = mats[21] = mats[22] = 1.0f;
  }
  // Mesh-area emitters, one per emissive triangle (scene/emitterPacking.ts:
  // 748-820): primitives in scene order, each primitive's instances in packed
  // order, triangles in source order; world vertices are the binary64
  // local-to-world image of the local vertices, rounded to f32.
  const uint32_t n_inst = (uint32_t)P.tlas_blas_roots.size();
  std::vector<float> inst(std::max<uint32_t>(1, n_inst) * vtr::kInstFloats, 0.0f);
  for (uint32_t i = 0; i < n_inst; ++i) {
    float* q = &inst[(size_t)i * vtr::kInstFloats];
    std::memcpy(q, &P.tlas_w2l[16 * (size_t)i], 64);
    std::memcpy(q + 16, &P.tlas_l2w[16 * (size_t)i], 64);
    const float* m = &P.tlas_l2w[16 * (size_t)i];
    const double det = (double)m[0] * ((double)m[5] * m[10] - (double)m[9] * m[6]) -
                       (double)m[4] * ((double)m[1] * m[10] - (double)m[9] * m[2]) +
                       (double)m[8] * ((double)m[1] * m[6] - (double)m[5] * m[2]);
    q[32] = det < 0 ? 1.0f : 0.0f;
    uint32_t none = 0xFFFFFFFFu;
    std::memcpy(q + 33, &none, 4);
  }
  std::vector<float> em;
  for (const vt::Binding& b : P.bindings) {
    const uint32_t mid = P.tri_material[b.tri_start];
    if (mid * vtr::kMatFloats >= mats.size()) continue;
    const float* mq = &mats[(size_t)mid * vtr::kMatFloats];
    if (!(mq[5] > 0 || mq[6] > 0 || mq[7] > 0)) continue;
    std::vector<uint32_t> packed_of_source(b.tri_count);
    for (uint32_t t = 0; t < b.tri_count; ++t) packed_of_source[P.tri_source[b.tri_start + t]] = b.tri_start + t;
    for (uint32_t ii = 0; ii < n_inst; ++ii) {
      if (P.instance_primitive[ii] != b.primitive) continue;
      const uint32_t base = (uint32_t)(em.size() / vtr::kEmitFloats);
      std::memcpy(&inst[(size_t)ii * vtr::kInstFloats + 33], &base, 4);
      const float* m = &P.tlas_l2w[16 * (size_t)ii];
      for (uint32_t s = 0; s < b.tri_count; ++s) {
        const uint32_t t = packed_of_source[s];
        float rec[vtr::kEmitFloats] = {0};
        for (int v = 0; v < 3; ++v) {
          const float* p = &P.positions[4 * (size_t)P.indices[4 * (size_t)t + v]];
          // transformPoint (pt-webgpu/src/math/mat4.ts:20-33): binary64,
          // divided by w only when |w| > 1e-8
          double q[4];
          for (int k = 0; k < 4; ++k)
            q[k] = (double)m[k] * p[0] + (double)m[4 + k] * p[1] + (double)m[8 + k] * p[2] + (double)m[12 + k];
          const bool proj = std::fabs(q[3]) > 1e-8;
          for (int k = 0; k < 3; ++k) rec[3 * v + k] = (float)(proj ? q[k] / q[3] : q[k]);
        }
        rec[9] = mq[5], rec[10] = mq[6], rec[11] = mq[7];
        rec[12] = mq[8];
        rec[13] = (float)ii;
        rec[14] = (float)t;
        em.insert(em.end(), rec, rec + vtr::kEmitFloats);
      }
    }
  }
  R.emitter_count = (uint32_t)(em.size() / vtr::kEmitFloats);
  if (em.empty()) em.assign(vtr::kEmitFloats, 0.0f);
  // Environment importance table (environmentPacking.ts:283-440).
  EnvTable et;
  {
    const std::string ee = build_environment(env, et);
    if (!ee.empty()) return rerr(e, VT_ERR_RANGE, ee);
  }
  R.env_present = et.present;
  R.env_w = et.w, R.env_h = et.h;
  R.env_rot[0] = et.rot_c, R.env_rot[1] = et.rot_s;
  // Light tree over the light slots -- mesh emitters in order, then the
  // environment -- built when there are two or more lights and the engine
  // uses tree selection (the reference's default; uploadSceneBuffers.ts:
  // 1850-1862, emitterPacking.ts:1005-1230).
  // Analytic emitters: directional, point, spot, rect and disc records.
  vte::Packed ana;
  const std::string pe = vte::pack(emitters, emitter_count, ana);
  if (!pe.empty()) return rerr(e, VT_ERR_RANGE, pe);
  R.n_dir = ana.n_dir, R.n_point = ana.n_point, R.n_spot = ana.n_spot, R.n_rect = ana.n_rect;
  std::vector<float> tree;
  const uint32_t n_lights = ana.count() + R.emitter_count + (R.env_present ? 1u : 0u);
  if (e->opts.light_selection == 0 && n_lights >= 2) {
    // Leaves in the NEE walk's order (buildLightTreeInputForScene,
    // emitterPacking.ts:1005-1230): directional, point, spot, rect/disc,
    // mesh triangles, environment.  Directional lights and the environment
    // sit on the union of the positional lights' bounds.
    std::vector<vtl::Leaf> leaves = vte::positional_leaves(ana);
    leaves.reserve(n_lights);
    for (uint32_t i = 0; i < R.emitter_count; ++i) {
      const float* q = &em[(size_t)i * vtr::kEmitFloats];
      leaves.push_back(vtl::triangle_leaf(q, q + 3, q + 6, q + 9, q[12] > 0.5f));
    }
    double ulo[3] = {INFINITY, INFINITY, INFINITY}, uhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const vtl::Leaf& L : leaves)
      for (int k = 0; k < 3; ++k) {
        ulo[k] = std::min(ulo[k], L.lo[k]);
        uhi[k] = std::max(uhi[k], L.hi[k]);
      }
    if (ana.n_dir) {
      double lo[3], hi[3];
      for (int k = 0; k < 3; ++k) {
        const bool ok = std::isfinite(ulo[k]);
        lo[k] = ok ? ulo[k] : 0.0;
        hi[k] = ok ? uhi[k] : 0.0;
      }
      std::vector<vtl::Leaf> dirs;
      for (uint32_t d = 0; d < ana.n_dir; ++d) dirs.push_back(vte::directional_leaf(ana, d, lo, hi));
      leaves.insert(leaves.begin(), dirs.begin(), dirs.end());
    }
    if (R.env_present) {
      // environment leaf: luminance x solid angle over the intensity-scaled
      // map, rounded to f32 (environmentPacking.ts:183-202, 366-413), bounds =
      // the union of the positional lights' bounds (the origin when none)
      const double total = et.power;
      if (!std::isfinite((float)total)) return rerr(e, VT_ERR_RANGE, "environment light-tree power exceeds binary32 range");
      vtl::Leaf L{};
      const float pw = (float)total;
      L.power = total > 0 ? (pw > 0 ? pw : 0x1p-149f) : 0.0;
      for (int k = 0; k < 3; ++k) {
        const bool ok = std::isfinite(ulo[k]);
        L.lo[k] = ok ? ulo[k] : 0.0;
        L.hi[k] = ok ? uhi[k] : 0.0;
        L.centroid[k] = (L.lo[k] + L.hi[k]) / 2;
      }
      L.oriented = false;
      leaves.push_back(L);
    }
    try {
      tree = vtl::build_packed(leaves);
    } catch (const std::exception& ex) {
      return rerr(e, VT_ERR_RANGE, ex.what());
    }
  }
  R.light_tree_nodes = (uint32_t)(tree.size() / vtl::kNodeFloats);
  R.light_tree_host = tree;
  if (tree.empty()) tree.assign(vtl::kNodeFloats, 0.0f);
  scene_bounds(R, P, scene_radius);
  vt_status s = VT_OK;
  s = s ? s : engine_upload(e, R.materials, mats.data(), mats.size() * 4);
  s = s ? s : engine_upload(e, R.emitters, em.data(), em.size() * 4);
  s = s ? s : engine_upload(e, R.instances, inst.data(), inst.size() * 4);
  if (et.cdf.empty()) et.cdf.assign(2, 0.0f);
  if (et.texels.empty()) et.texels.assign(4, 0.0f);
  s = s ? s : engine_upload(e, R.env_cdf, et.cdf.data(), et.cdf.size() * 4);
  s = s ? s : engine_upload(e, R.env_texels, et.texels.data(), et.texels.size() * 4);
  s = s ? s : engine_upload(e, R.tri_source, P.tri_source.data(), P.tri_source.size() * 4);
  s = s ? s : engine_upload(e, R.light_tree, tree.data(), tree.size() * 4);
  if (ana.records.empty()) ana.records.assign(vtr::kLightFloats, 0.0f);
  s = s ? s : engine_upload(e, R.lights, ana.records.data(), ana.records.size() * 4);
  if (!s && (e->fail_inject & kInjectRenderAlloc)) {
    e->fail_inject &= ~kInjectRenderAlloc;
    s = rerr(e, VT_ERR_OUT_OF_MEMORY, "injected allocation failure while uploading the scene's shading data");
  }
  if (s) return s;
  *out = hold.release();
  return VT_OK;
}

static vt_status ensure_paths(vt_engine* e, uint32_t w, uint32_t h) {
  RenderState& R = *e->render;
  const uint32_t train = R.nrc_enabled ? R.nrc_train_paths : 0u;
  const uint32_t paths = w * h + train;
  if (R.width == w && R.height == h && R.paths == paths) return VT_OK;
  if (vt_status s = engine_nrc_quiesce(e)) return s;  // a training step may still read the records
  R.release_paths();
  const size_t n = (size_t)paths;
  const size_t pix = (size_t)w * h;
  vt_status s = VT_OK;
  auto need = [&](DeviceBuffer& b, size_t bytes) {
    if (!s) s = b.ensure(e, std::max<size_t>(bytes, 64));
  };
  need(R.rng, n * 4);
  need(R.rng_dim, n * 4);
  need(R.throughput, n * 12);
  need(R.conn, n * 12);
  need(R.radiance, n * 12);
  need(R.prev_pdf, n * 4);
  need(R.flags, n * 4);
  need(R.signature, n * 8);
  need(R.count, 64);
  need(R.queue[0], n * 4);
  need(R.queue[1], n * 4);
  need(R.rays[0], n * 32);
  need(R.rays[1], n * 32);
  need(R.hits, n * 20);
  need(R.hit_pos, n * 12);
  need(R.shadow_path, n * 4);
  need(R.shadow_rays, n * 32);
  need(R.shadow_contrib, n * 12);
  need(R.shadow_occ, n * 4);
  need(R.accum, pix * 16);
  need(R.moments, pix * 8);
  need(R.aux_nd, pix * 16);
  need(R.aux_albedo, pix * 16);
  need(R.present, pix * 16);
  if (train) {
    need(R.nrc_count, 64);
    const size_t nq = pix + train;  // image-path queries + training-path terminal queries
    need(R.q_feat, nq * 64);
    need(R.q_thr, nq * 12);
    need(R.q_path, nq * 4);
    need(R.q_rgb, nq * 12);
    need(R.train_rec, (size_t)train * vtr::kTrainSlots * vtr::kTrainRec * 4);
    need(R.train_n, (size_t)train * 4);
    need(R.train_feat, (size_t)train * vtr::kTrainSlots * 64);
    need(R.train_tgt, (size_t)train * vtr::kTrainSlots * 12);
  }
  if (s) {
    R.release_paths();
    return s;
  }
  R.width = w;
  R.height = h;
  R.paths = paths;
  e->samples_accumulated = 0;  // the next frame clears the accumulation
  return VT_OK;
}

static vtk::TraceParams trace_params(vt_engine* e) { return engine_trace_params(e); }

// A frame's launch parameters (validated), shared by vt_render_frame and the
// batched frames of vt_render_batch_frame.
struct FrameLaunch {
  vtr::FrameParams f{};
  vtr::ShadeScene sc{};
  vtr::PathState ps{};
  bool skip = false;   // the sample target is reached: render nothing
};

static vt_status frame_setup(vt_engine* e, const vt_frame* fr, FrameLaunch& L) {
  if (!e->has_scene || !e->render) return rerr(e, VT_ERR_STATE, "no scene is set");
  if (!fr) return rerr(e, VT_ERR_INVALID_ARGUMENT, "frame is NULL");
  if (fr->width == 0 || fr->height == 0 || (uint64_t)fr->width * fr->height > (1u << 28))
    return rerr(e, VT_ERR_RANGE, "viewport must be non-empty and at most 2^28 pixels");
  for (int i = 0; i < 16; ++i)
    if (!std::isfinite(fr->inv_view_proj[i])) return rerr(e, VT_ERR_RANGE, "inverse view-projection must be finite");
  HIP_CHECK(e, hipSetDevice(e->device));
  if (vt_status s = ensure_paths(e, fr->width, fr->height)) return s;
  L.skip = (e->opts.max_samples_per_pixel && e->samples_accumulated >= e->opts.max_samples_per_pixel) ||
           (fr->samples_target && e->samples_accumulated >= fr->samples_target);
  if (L.skip) return VT_OK;
  RenderState& R = *e->render;
  vtr::FrameParams& f = L.f;
  f = vtr::FrameParams{};
  std::memcpy(f.inv_vp, fr->inv_view_proj, 64);
  f.width = fr->width;
  f.height = fr->height;
  f.frame_seed = fr->frame_seed;
  f.frame_index = fr->frame_index;
  const uint32_t mb = fr->max_bounces ? fr->max_bounces : e->opts.max_bounces;
  f.bounce_limit = std::max(1u, std::min(mb, 8u));
  f.ray_bias = R.ray_bias;
  f.tmin = R.tmin;
  f.pixels = fr->width * fr->height;
  f.sobol = e->opts.sampling == 1 ? 1u : 0u;
  f.sobol_key = vts::frame_key(fr->frame_seed, fr->frame_index);
  f.sobol_dirs = (const uint16_t*)R.sobol_dirs.ptr;
  f.nrc = R.nrc_enabled ? 1u : 0u;
  f.train_paths = R.nrc_enabled ? R.nrc_train_paths : 0u;
  f.train_vertices = R.nrc_train_vertices;
  f.shade_sort = e->shade_sort;
  for (int k = 0; k < 3; ++k) f.scene_min[k] = R.scene_min[k], f.scene_inv_extent[k] = R.scene_inv_extent[k];
  vtr::ShadeScene& sc = L.sc;
  sc = vtr::ShadeScene{};
  sc.positions = (const float4*)e->scene.positions.ptr;
  sc.normals = (const float4*)e->scene.normals.ptr;
  sc.indices = (const uint4*)e->scene.indices.ptr;
  sc.tri_material = (const uint32_t*)e->scene.tri_material.ptr;
  sc.tri_source = (const uint32_t*)R.tri_source.ptr;
  sc.materials = (const float*)R.materials.ptr;
  sc.emitters = (const float*)R.emitters.ptr;
  sc.instances = (const float*)R.instances.ptr;
  sc.env_cdf = (const float*)R.env_cdf.ptr;
  sc.env_texels = (const float4*)R.env_texels.ptr;
  sc.emitter_count = R.emitter_count;
  sc.env_present = R.env_present;
  sc.env_w = R.env_w;
  sc.env_h = R.env_h;
  sc.env_rot_c = R.env_rot[0];
  sc.env_rot_s = R.env_rot[1];
  sc.light_tree = (const float*)R.light_tree.ptr;
  sc.light_tree_nodes = R.light_tree_nodes;
  sc.lights = (const float*)R.lights.ptr;
  sc.n_dir = R.n_dir;
  sc.n_point = R.n_point;
  sc.n_spot = R.n_spot;
  sc.n_rect = R.n_rect;
  vtr::PathState& ps = L.ps;
  ps = vtr::PathState{};
  ps.rng = (uint32_t*)R.rng.ptr;
  ps.rng_dim = (uint32_t*)R.rng_dim.ptr;
  ps.throughput = (float*)R.throughput.ptr;
  ps.conn = (float*)R.conn.ptr;
  ps.radiance = (float*)R.radiance.ptr;
  ps.prev_pdf = (float*)R.prev_pdf.ptr;
  ps.flags = (uint32_t*)R.flags.ptr;
  ps.signature = (uint64_t*)R.signature.ptr;
  ps.count = (uint32_t*)R.count.ptr;
  ps.hits = (uint32_t*)R.hits.ptr;
  ps.hit_pos = (float*)R.hit_pos.ptr;
  ps.shadow_path = (uint32_t*)R.shadow_path.ptr;
  ps.shadow_rays = (float4*)R.shadow_rays.ptr;
  ps.shadow_contrib = (float*)R.shadow_contrib.ptr;
  ps.shadow_occ = (uint32_t*)R.shadow_occ.ptr;
  ps.accum = (float*)R.accum.ptr;
  ps.moments = (float*)R.moments.ptr;
  ps.aux_nd = (float*)R.aux_nd.ptr;
  ps.aux_albedo = (float*)R.aux_albedo.ptr;
  return VT_OK;
}


extern "C" {

vt_status vt_render_frame(vt_engine* e, const vt_frame* fr) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    if (!fr) return rerr(e, VT_ERR_INVALID_ARGUMENT, "frame is NULL");
    FrameLaunch L;
    if (vt_status s = frame_setup(e, fr, L)) return s;
    if (L.skip) return VT_OK;
    RenderState& R = *e->render;
    vtr::FrameParams& f = L.f;
    vtr::ShadeScene& sc = L.sc;
    vtr::PathState& ps = L.ps;
    // Accumulation restarts (reset, new scene, new viewport, cache settings)
    // set the sample count to zero; the first frame after one clears the sums.
    if (e->samples_accumulated == 0) {
      HIP_CHECK(e, vtr::launch_zero_f32(ps.accum, f.pixels * 4, e->stream));
      HIP_CHECK(e, vtr::launch_zero_f32(ps.moments, f.pixels * 2, e->stream));
    }
    ps.nrc_count = (uint32_t*)R.nrc_count.ptr;
    ps.q_feat = (float*)R.q_feat.ptr;
    ps.q_thr = (float*)R.q_thr.ptr;
    ps.q_path = (uint32_t*)R.q_path.ptr;
    ps.q_rgb = (float*)R.q_rgb.ptr;
    ps.train_rec = (float*)R.train_rec.ptr;
    ps.train_n = (uint32_t*)R.train_n.ptr;
    ps.train_feat = (float*)R.train_feat.ptr;
    ps.train_tgt = (float*)R.train_tgt.ptr;
    const uint32_t cap = R.paths;
    // The previous frame's cache training step keeps running on its own
    // stream beside this frame's traces and shading: it reads only the
    // training features, targets and record count, which this frame rewrites
    // only after its cache inference has joined the step.  The frame-time
    // events (ev0 -> ev1) therefore exclude the training step.
    if (f.nrc) {
      if (vt_status s = engine_nrc_create(e)) return s;
    }
    HIP_CHECK(e, hipEventRecord(R.ev0, e->stream));
    HIP_CHECK(e, vtk::launch_zero((uint32_t*)R.count.ptr, 16, e->stream));
    // Query count only: the previous frame's training step may still read the
    // record count (word 1), which is cleared after the cache is joined below.
    if (f.nrc) HIP_CHECK(e, vtk::launch_zero((uint32_t*)R.nrc_count.ptr, 1, e->stream));
    int cur = 0;
    ps.queue = (uint32_t*)R.queue[cur].ptr;
    ps.rays = (float4*)R.rays[cur].ptr;
    HIP_CHECK(e, vtr::launch_camera(f, ps, e->stream));
    vtk::TraceParams tp = trace_params(e);
    // From the second bounce on, a bounce's extension trace runs beside the
    // previous bounce's shadow trace: each gets half of the persistent grid,
    // its own ray counter and its own half of the stack-spill area.
    const uint32_t half_grid = std::max(1u, e->trace_grid / 2u);
    const size_t spill_half = (size_t)(vtk::kStackCapacity - vtk::kLdsDepth) * half_grid * vtk::kTraceBlock;
    for (uint32_t v = 0; v <= f.bounce_limit; ++v) {
      ps.queue = (uint32_t*)R.queue[cur].ptr;
      ps.rays = (float4*)R.rays[cur].ptr;
      ps.next_queue = (uint32_t*)R.queue[cur ^ 1].ptr;
      ps.next_rays = (float4*)R.rays[cur ^ 1].ptr;
      vtk::TraceParams cp = tp;
      cp.rays = ps.rays;
      cp.hits = ps.hits;
      cp.hit_pos = ps.hit_pos;
      cp.n_ptr = ps.count;
      cp.sided = 1;  // traceClosest: back faces of single-sided materials are skipped
      if (e->opts.bvh_traversal == 1) {
        if (vt_status s = engine_trace_cwbvh(e, ps.rays, cap, ps.count, ps.hits, ps.hit_pos, 1u)) return s;
      } else {
        HIP_CHECK(e, vtk::launch_trace(cp, false, false, v == 0 ? e->trace_grid : half_grid, e->stream));
      }
      // The previous bounce's shadow rays are resolved (radiance, shadow
      // queue and count) before this bounce's shading touches them.
      if (v > 0) HIP_CHECK(e, hipStreamWaitEvent(e->stream, R.ev_join, 0));
      HIP_CHECK(e, vtr::launch_shade(f, sc, ps, v, cap, e->stream));
      if (v < f.bounce_limit) {
        HIP_CHECK(e, hipEventRecord(R.ev_fork, e->stream));
        HIP_CHECK(e, hipStreamWaitEvent(R.side, R.ev_fork, 0));
        vtk::TraceParams sp = tp;
        sp.rays = ps.shadow_rays;
        sp.occluded = ps.shadow_occ;
        sp.n_ptr = ps.count + 2;
        sp.sided = 1;  // traceSurfaceVisibility: back faces of single-sided materials do not block
        sp.counter = tp.counter + 8;
        sp.spill = tp.spill + spill_half;
        HIP_CHECK(e, vtk::launch_trace(sp, true, false, half_grid, R.side));
        HIP_CHECK(e, vtr::launch_shadow_resolve(ps, cap, R.side));
        HIP_CHECK(e, vtr::launch_reset_counts(ps, 2u, R.side));  // shadow count
        HIP_CHECK(e, hipEventRecord(R.ev_join, R.side));
      }
      HIP_CHECK(e, vtr::launch_reset_counts(ps, 1u, e->stream));  // extension counts
      cur ^= 1;
    }
    if (f.nrc) {
      // Cache queries of the image paths, then the online training step on the
      // self-training paths' recorded vertices.
      if (vt_status s = engine_nrc_infer_counted(e, ps.q_feat, ps.nrc_count, ps.q_rgb, f.pixels + f.train_paths))
        return s;
      HIP_CHECK(e, vtr::launch_nrc_apply(ps, f.pixels + f.train_paths, e->stream));
    }
    HIP_CHECK(e, vtr::launch_accumulate(f, ps, e->stream));
    if (f.nrc) {
      HIP_CHECK(e, vtk::launch_zero((uint32_t*)R.nrc_count.ptr + 1, 1, e->stream));
      HIP_CHECK(e, vtr::launch_nrc_targets(f, ps, e->stream));
      if (vt_status s = engine_nrc_train_counted(e, ps.train_feat, ps.train_tgt, ps.nrc_count + 1,
                                                 f.train_paths * vtr::kTrainSlots))
        return s;
    }
    HIP_CHECK(e, hipEventRecord(R.ev1, e->stream));
    e->samples_accumulated += 1;
    e->frame_timed = true;
    return VT_OK;
  });
}

vt_status vt_capture(vt_engine* e, float* rgba, uint32_t width, uint32_t height) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    if (!rgba) return rerr(e, VT_ERR_INVALID_ARGUMENT, "rgba is NULL");
    if (!e->render || e->render->width != width || e->render->height != height)
      return rerr(e, VT_ERR_STATE, "no frame of this size has been rendered");
    // The traversal's error word is read (and cleared) first: after a stack
    // overflow the capture reports it once and leaves `rgba` untouched, and
    // later captures succeed again.
    uint32_t err = 0;
    HIP_CHECK(e, hipMemcpyAsync(&err, e->error_word.ptr, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_CHECK(e, hipStreamSynchronize(e->stream));
    if (err) {
      HIP_CHECK(e, vtk::launch_zero((uint32_t*)e->error_word.ptr, 1, e->stream));
      HIP_CHECK(e, hipStreamSynchronize(e->stream));
      return rerr(e, VT_ERR_HIP, "traversal stack capacity exceeded during rendering");
    }
    HIP_CHECK(e, hipMemcpyAsync(rgba, e->render->accum.ptr, (size_t)width * height * 16,
                                hipMemcpyDeviceToHost, e->stream));
    HIP_CHECK(e, hipStreamSynchronize(e->stream));
    const float inv = e->samples_accumulated ? 1.0f / (float)e->samples_accumulated : 0.0f;
    for (size_t i = 0; i < (size_t)width * height; ++i) {
      rgba[4 * i] *= inv;
      rgba[4 * i + 1] *= inv;
      rgba[4 * i + 2] *= inv;
      rgba[4 * i + 3] = 1.0f;
    }
    return VT_OK;
  });
}

vt_status vt_capture_signatures(vt_engine* e, uint64_t* sig, uint32_t width, uint32_t height) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || !sig) return VT_ERR_INVALID_ARGUMENT;
    if (!e->render || e->render->width != width || e->render->height != height)
      return rerr(e, VT_ERR_STATE, "no frame of this size has been rendered");
    HIP_CHECK(e, hipMemcpyAsync(sig, e->render->signature.ptr, (size_t)width * height * 8,
                                hipMemcpyDeviceToHost, e->stream));
    HIP_CHECK(e, hipStreamSynchronize(e->stream));
    return VT_OK;
  });
}

vt_status vt_reset(vt_engine* e) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    e->samples_accumulated = 0;
    if (e->render && e->render->width) {
      const uint32_t pix = e->render->width * e->render->height;
      HIP_CHECK(e, vtr::launch_zero_f32((float*)e->render->accum.ptr, pix * 4, e->stream));
      HIP_CHECK(e, vtr::launch_zero_f32((float*)e->render->moments.ptr, pix * 2, e->stream));
      HIP_CHECK(e, hipStreamSynchronize(e->stream));
    }
    return VT_OK;
  });
}

vt_status vt_capture_output(vt_engine* e, float* rgba, uint32_t width, uint32_t height, const vt_present* pr) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    if (!rgba || !pr) return rerr(e, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (pr->tonemap > 4) return rerr(e, VT_ERR_RANGE, "tonemap must be 0 (aces) .. 4 (none)");
    if (!(pr->exposure >= 0.0f) || !std::isfinite(pr->exposure))
      return rerr(e, VT_ERR_RANGE, "exposure must be finite and non-negative");
    if (pr->color_space > 1) return rerr(e, VT_ERR_RANGE, "color_space must be 0 (srgb) or 1 (linear)");
    if (!e->render || e->render->width != width || e->render->height != height || !e->samples_accumulated)
      return rerr(e, VT_ERR_STATE, "no frame of this size has been rendered");
    HIP_CHECK(e, hipSetDevice(e->device));
    RenderState& R = *e->render;
    const uint32_t pix = width * height;
    HIP_CHECK(e, vtr::launch_present((const float*)R.accum.ptr, pix, e->samples_accumulated, pr->tonemap,
                                     pr->exposure, pr->color_space, (float*)R.present.ptr, e->stream));
    HIP_CHECK(e, hipMemcpyAsync(rgba, R.present.ptr, (size_t)pix * 16, hipMemcpyDeviceToHost, e->stream));
    HIP_CHECK(e, hipStreamSynchronize(e->stream));
    return VT_OK;
  });
}

vt_status vt_capture_aux(vt_engine* e, uint32_t which, float* out, uint32_t width, uint32_t height) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    if (!out) return rerr(e, VT_ERR_INVALID_ARGUMENT, "out is NULL");
    if (which > 2) return rerr(e, VT_ERR_RANGE, "which must be 0 (normal-depth), 1 (albedo) or 2 (variance)");
    if (!e->render || e->render->width != width || e->render->height != height || !e->samples_accumulated)
      return rerr(e, VT_ERR_STATE, "no frame of this size has been rendered");
    HIP_CHECK(e, hipSetDevice(e->device));
    RenderState& R = *e->render;
    const uint32_t pix = width * height;
    const void* src = which == 0 ? R.aux_nd.ptr : R.aux_albedo.ptr;
    if (which == 2) {
      HIP_CHECK(e, vtr::launch_variance((const float*)R.moments.ptr, pix, e->samples_accumulated,
                                        (float*)R.present.ptr, e->stream));
      src = R.present.ptr;
    }
    HIP_CHECK(e, hipMemcpyAsync(out, src, (size_t)pix * 16, hipMemcpyDeviceToHost, e->stream));
    HIP_CHECK(e, hipStreamSynchronize(e->stream));
    return VT_OK;
  });
}

vt_status vt_frame_time(vt_engine* e, float* ms) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || !ms) return VT_ERR_INVALID_ARGUMENT;
    *ms = 0;
    if (!e->render || !e->frame_timed) return VT_OK;
    HIP_CHECK(e, hipEventSynchronize(e->render->ev1));
    HIP_CHECK(e, hipEventElapsedTime(ms, e->render->ev0, e->render->ev1));
    e->last_frame_ms = *ms;
    return VT_OK;
  });
}

vt_status vt_nrc_render_config(vt_engine* e, int32_t enabled, uint32_t train_paths) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    if (enabled && (train_paths == 0 || train_paths > (1u << 22)))
      return rerr(e, VT_ERR_RANGE, "train_paths must be in [1, 2^22]");
    if (!e->render) return rerr(e, VT_ERR_STATE, "no scene is set");
    RenderState& R = *e->render;
    R.nrc_enabled = enabled != 0;
    if (enabled) R.nrc_train_paths = train_paths;
    R.width = R.height = 0;  // re-size path state on the next frame
    e->samples_accumulated = 0;
    return VT_OK;
  });
}

vt_status vt_nrc_teacher_vertices(vt_engine* e, uint32_t vertices) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    if (vertices > vtr::kTrainSlots)
      return rerr(e, VT_ERR_RANGE, "teacher vertices must be in [0, 4] (0: full-length training paths)");
    if (!e->render) return rerr(e, VT_ERR_STATE, "no scene is set");
    e->render->nrc_train_vertices = vertices;
    e->samples_accumulated = 0;
    return VT_OK;
  });
}

vt_status vt_light_tree(vt_engine* e, float* nodes, uint32_t capacity, uint32_t* count) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || !count) return VT_ERR_INVALID_ARGUMENT;
    if (!e->render) return rerr(e, VT_ERR_STATE, "no scene is set");
    const RenderState& R = *e->render;
    *count = R.light_tree_nodes;
    if (nodes && capacity) {
      if (capacity < R.light_tree_nodes) return rerr(e, VT_ERR_RANGE, "capacity below the light tree's node count");
      std::memcpy(nodes, R.light_tree_host.data(), (size_t)R.light_tree_nodes * vtl::kNodeFloats * sizeof(float));
    }
    return VT_OK;
  });
}

vt_status vt_render_info(vt_engine* e, uint32_t* emitter_count, float* ray_bias) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || !e->render) return VT_ERR_STATE;
    if (emitter_count) *emitter_count = e->render->emitter_count;
    if (ray_bias) *ray_bias = e->render->ray_bias;
    return VT_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched frames: several engines (each with its own scene, camera and
// viewport) render one frame each, with every bounce's closest-hit rays of
// all of them traced in one persistent launch over the scenes' table (and
// the shadow rays in another), on the batch's stream; a frame can be recorded
// once into a HIP graph and replayed.
// ---------------------------------------------------------------------------

struct vt_render_batch {
  vt_engine* host = nullptr;  // its stream, traversal scratch and tuning run the batch
  std::vector<vt_engine*> members;
  DeviceBuffer scenes, table, offsets, rays, hits, hit_pos, occ, words;
  // Graph of one frame and what it was recorded against.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<vt_frame> frames;
  std::vector<uintptr_t> state;  // each member's viewport, path-state and scene buffers at recording
  uint32_t* staged = nullptr;      // pinned frame words (seed, index, Sobol key per member)
  hipEvent_t staged_done = nullptr;
  hipEvent_t frame_done = nullptr;  // members' own streams wait for it
};

namespace {

vt_status batch_err(vt_render_batch* b, vt_status s, const std::string& m) { return rerr(b->host, s, m); }

void batch_drop_graph(vt_render_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->frames.clear();
  b->state.clear();
}

// Everything a batched frame needs before its launches: each member's frame
// parameters, the device tables, buffers sized for the members' queues.
struct BatchLaunch {
  std::vector<FrameLaunch> L;
  std::vector<vtr::BatchMember> bm;
  uint32_t capacity = 0, max_limit = 0;
  bool wide = true;
};

vt_status batch_prepare(vt_render_batch* b, const vt_frame* frames, bool device_words, BatchLaunch& B) {
  vt_engine* e = b->host;
  const uint32_t K = (uint32_t)b->members.size();
  if (K == 0) return batch_err(b, VT_ERR_STATE, "the batch has no members");
  if (!frames) return batch_err(b, VT_ERR_INVALID_ARGUMENT, "frames is NULL");
  B.L.assign(K, FrameLaunch{});
  B.bm.assign(K, vtr::BatchMember{});
  std::vector<vtk::DeviceScene> ds(K);
  uint64_t cap = 0;
  for (uint32_t k = 0; k < K; ++k) {
    vt_engine* m = b->members[k];
    if (vt_status s = frame_setup(m, &frames[k], B.L[k])) return batch_err(b, s, "member " + std::to_string(k) + ": " + m->last_error);
    if (B.L[k].skip) return batch_err(b, VT_ERR_STATE, "member " + std::to_string(k) + " has reached its sample target");
    RenderState& R = *m->render;
    if (R.nrc_enabled) return batch_err(b, VT_ERR_STATE, "batched frames do not run the radiance cache");
    vtr::BatchMember& q = B.bm[k];
    q.rays[0] = (const float4*)R.rays[0].ptr;
    q.rays[1] = (const float4*)R.rays[1].ptr;
    q.hits = (uint32_t*)R.hits.ptr;
    q.hit_pos = (float*)R.hit_pos.ptr;
    q.shadow_rays = (const float4*)R.shadow_rays.ptr;
    q.shadow_occ = (uint32_t*)R.shadow_occ.ptr;
    q.count = (const uint32_t*)R.count.ptr;
    bool w = false;
    ds[k] = engine_device_scene(m->scene, *m->layout, &w);
    B.wide = B.wide && w;
    cap += R.paths;
    B.max_limit = std::max(B.max_limit, B.L[k].f.bounce_limit);
    if (device_words) B.L[k].f.frame_dev = (const uint32_t*)b->words.ptr + 3u * k;
  }
  if (cap > 0x7FFFFFFFull) return batch_err(b, VT_ERR_RANGE, "the members' queues exceed 2^31 - 1 rays");
  B.capacity = (uint32_t)cap;
  vt_status s = VT_OK;
  s = s ? s : b->rays.ensure(e, cap * 32);
  s = s ? s : b->hits.ensure(e, cap * 20);
  s = s ? s : b->hit_pos.ensure(e, cap * 12);
  s = s ? s : b->occ.ensure(e, cap * 4);
  s = s ? s : b->offsets.ensure(e, (K + 1) * 4);
  s = s ? s : b->words.ensure(e, (size_t)K * 12);
  s = s ? s : engine_upload(e, b->scenes, ds.data(), ds.size() * sizeof(vtk::DeviceScene));
  s = s ? s : engine_upload(e, b->table, B.bm.data(), B.bm.size() * sizeof(vtr::BatchMember));
  if (s) return s;
  HIP_CHECK(e, hipStreamSynchronize(e->stream));  // the uploads read host vectors of this call
  if (device_words)  // the addresses above may have moved
    for (uint32_t k = 0; k < K; ++k) B.L[k].f.frame_dev = (const uint32_t*)b->words.ptr + 3u * k;
  return VT_OK;
}

// The launches of one batched frame on the host engine's stream (recordable:
// kernels only, no allocation, no memset, no host synchronisation).
vt_status batch_enqueue(vt_render_batch* b, BatchLaunch& B) {
  vt_engine* e = b->host;
  hipStream_t st = e->stream;
  const uint32_t K = (uint32_t)b->members.size();
  const vtr::BatchMember* table = (const vtr::BatchMember*)b->table.ptr;
  uint32_t* offsets = (uint32_t*)b->offsets.ptr;
  for (uint32_t k = 0; k < K; ++k) {
    RenderState& R = *b->members[k]->render;
    vtr::PathState& ps = B.L[k].ps;
    HIP_CHECK(e, vtk::launch_zero((uint32_t*)R.count.ptr, 16, st));
    ps.queue = (uint32_t*)R.queue[0].ptr;
    ps.rays = (float4*)R.rays[0].ptr;
    HIP_CHECK(e, vtr::launch_camera(B.L[k].f, ps, st));
  }
  // The host engine's traversal scratch and tuning (its own scene, if any,
  // plays no part: every ray takes its scene from the table).
  vtk::TraceParams tp{};
  tp.counter = (uint32_t*)e->counter.ptr;
  tp.error = (uint32_t*)e->error_word.ptr;
  tp.spill = (uint2*)e->spill.ptr;
  tp.refill_min = e->refill_min;
  tp.steps = e->trace_steps;
  tp.near_first = e->near_first;
  tp.defer_min = e->defer_min;
  tp.scenes = (const vtk::DeviceScene*)b->scenes.ptr;
  tp.scene_offsets = offsets;
  tp.scene_count = K;
  tp.wide = (B.wide && e->node_width == 4) ? 1u : 0u;
  tp.n_ptr = offsets + K;
  tp.sided = 1;  // traceClosest / traceSurfaceVisibility
  int cur = 0;
  for (uint32_t v = 0; v <= B.max_limit; ++v) {
    // extension rays of every member: gather, one traversal launch, scatter
    HIP_CHECK(e, vtr::launch_batch_prefix(table, K, 0u, offsets, st));
    HIP_CHECK(e, vtr::launch_batch_gather(table, K, offsets, 0u, (uint32_t)cur, (float4*)b->rays.ptr, B.capacity, st));
    vtk::TraceParams cp = tp;
    cp.rays = (const float4*)b->rays.ptr;
    cp.hits = (uint32_t*)b->hits.ptr;
    cp.hit_pos = (float*)b->hit_pos.ptr;
    HIP_CHECK(e, vtk::launch_trace_batch(cp, e->trace_grid, st));
    HIP_CHECK(e, vtr::launch_batch_scatter(table, K, offsets, 0u, (const uint32_t*)b->hits.ptr,
                                           (const float*)b->hit_pos.ptr, B.capacity, st));
    for (uint32_t k = 0; k < K; ++k) {
      if (v > B.L[k].f.bounce_limit) continue;
      RenderState& R = *b->members[k]->render;
      vtr::PathState& ps = B.L[k].ps;
      ps.queue = (uint32_t*)R.queue[cur].ptr;
      ps.rays = (float4*)R.rays[cur].ptr;
      ps.next_queue = (uint32_t*)R.queue[cur ^ 1].ptr;
      ps.next_rays = (float4*)R.rays[cur ^ 1].ptr;
      HIP_CHECK(e, vtr::launch_shade(B.L[k].f, B.L[k].sc, ps, v, R.paths, st));
    }
    if (v < B.max_limit) {
      // shadow rays of every member
      HIP_CHECK(e, vtr::launch_batch_prefix(table, K, 2u, offsets, st));
      HIP_CHECK(e, vtr::launch_batch_gather(table, K, offsets, 1u, 0u, (float4*)b->rays.ptr, B.capacity, st));
      vtk::TraceParams sp = tp;
      sp.rays = (const float4*)b->rays.ptr;
      sp.occluded = (uint32_t*)b->occ.ptr;
      HIP_CHECK(e, vtk::launch_trace_batch_any(sp, e->trace_grid, st));
      HIP_CHECK(e, vtr::launch_batch_scatter(table, K, offsets, 1u, (const uint32_t*)b->occ.ptr, nullptr,
                                             B.capacity, st));
      for (uint32_t k = 0; k < K; ++k) {
        if (v >= B.L[k].f.bounce_limit) continue;
        const RenderState& R = *b->members[k]->render;
        HIP_CHECK(e, vtr::launch_shadow_resolve(B.L[k].ps, R.paths, st));
        HIP_CHECK(e, vtr::launch_reset_counts(B.L[k].ps, 2u, st));
      }
    }
    for (uint32_t k = 0; k < K; ++k)
      if (v <= B.L[k].f.bounce_limit) HIP_CHECK(e, vtr::launch_reset_counts(B.L[k].ps, 1u, st));
    cur ^= 1;
  }
  for (uint32_t k = 0; k < K; ++k) HIP_CHECK(e, vtr::launch_accumulate(B.L[k].f, B.L[k].ps, st));
  return VT_OK;
}

// Clears the accumulation of members whose sample count restarted, after
// their own streams' work.
vt_status batch_begin(vt_render_batch* b, const BatchLaunch& B) {
  vt_engine* e = b->host;
  for (size_t k = 0; k < b->members.size(); ++k) {
    vt_engine* m = b->members[k];
    if (m != e) HIP_CHECK(e, hipStreamSynchronize(m->stream));
    if (m->render->side) HIP_CHECK(e, hipStreamSynchronize(m->render->side));
    if (m->samples_accumulated == 0) {
      HIP_CHECK(e, vtr::launch_zero_f32(B.L[k].ps.accum, B.L[k].f.pixels * 4, e->stream));
      HIP_CHECK(e, vtr::launch_zero_f32(B.L[k].ps.moments, B.L[k].f.pixels * 2, e->stream));
    }
  }
  return VT_OK;
}

// Orders each member's own stream after the batch's frame (its captures and
// later frames then see the result) and counts the sample.
vt_status batch_end(vt_render_batch* b) {
  vt_engine* e = b->host;
  HIP_CHECK(e, hipEventRecord(b->frame_done, e->stream));
  for (vt_engine* m : b->members) {
    if (m != e) HIP_CHECK(e, hipStreamWaitEvent(m->stream, b->frame_done, 0));
    m->samples_accumulated += 1;
    m->frame_timed = false;
  }
  return VT_OK;
}

// What a recorded graph depends on: each member's viewport, path-state and
// scene buffers.
std::vector<uintptr_t> batch_state(vt_render_batch* b) {
  std::vector<uintptr_t> v;
  for (vt_engine* m : b->members) {
    const RenderState& R = *m->render;
    v.push_back(((uintptr_t)R.width << 32) | R.height);
    v.push_back((uintptr_t)m->scene_generation);
    for (const void* p : {R.rays[0].ptr, R.rays[1].ptr, R.hits.ptr, R.accum.ptr, R.materials.ptr, R.lights.ptr,
                          R.env_texels.ptr, m->scene.tris.ptr, m->scene.blas4.ptr})
      v.push_back((uintptr_t)p);
  }
  return v;
}

}  // namespace

extern "C" {

vt_status vt_render_batch_create(vt_engine* host, vt_render_batch** out) {
  return vt_guarded(host, [&]() -> vt_status {
    if (!host || !out) return VT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    std::unique_ptr<vt_render_batch> b(new vt_render_batch());
    b->host = host;
    HIP_CHECK(host, hipSetDevice(host->device));
    HIP_CHECK(host, hipEventCreateWithFlags(&b->staged_done, hipEventDisableTiming));
    HIP_CHECK(host, hipEventCreateWithFlags(&b->frame_done, hipEventDisableTiming));
    *out = b.release();
    return VT_OK;
  });
}

vt_status vt_render_batch_destroy(vt_render_batch* b) {
  if (!b) return VT_OK;
  return vt_guarded(b->host, [&]() -> vt_status {
    (void)hipStreamSynchronize(b->host->stream);
    batch_drop_graph(b);
    for (DeviceBuffer* d : {&b->scenes, &b->table, &b->offsets, &b->rays, &b->hits, &b->hit_pos, &b->occ, &b->words})
      d->release();
    if (b->staged) (void)hipHostFree(b->staged);
    if (b->staged_done) (void)hipEventDestroy(b->staged_done);
    if (b->frame_done) (void)hipEventDestroy(b->frame_done);
    delete b;
    return VT_OK;
  });
}

vt_status vt_render_batch_add(vt_render_batch* b, vt_engine* member) {
  if (!b) return VT_ERR_INVALID_ARGUMENT;
  return vt_guarded(b->host, [&]() -> vt_status {
    if (!member) return batch_err(b, VT_ERR_INVALID_ARGUMENT, "member is NULL");
    if (member->device != b->host->device) return batch_err(b, VT_ERR_INVALID_ARGUMENT, "the member runs on another device");
    if (!member->has_scene || !member->render) return batch_err(b, VT_ERR_STATE, "the member has no scene");
    if (member->opts.bvh_traversal != 0) return batch_err(b, VT_ERR_STATE, "batched frames walk the binary BVH");
    if (b->members.size() >= 1024) return batch_err(b, VT_ERR_RANGE, "at most 1024 members");
    for (vt_engine* m : b->members)
      if (m == member) return batch_err(b, VT_ERR_INVALID_ARGUMENT, "the engine is already a member");
    batch_drop_graph(b);
    b->members.push_back(member);
    return VT_OK;
  });
}

vt_status vt_render_batch_frame(vt_render_batch* b, const vt_frame* frames) {
  if (!b) return VT_ERR_INVALID_ARGUMENT;
  return vt_guarded(b->host, [&]() -> vt_status {
    HIP_CHECK(b->host, hipSetDevice(b->host->device));
    BatchLaunch B;
    if (vt_status s = batch_prepare(b, frames, false, B)) return s;
    if (vt_status s = batch_begin(b, B)) return s;
    if (vt_status s = batch_enqueue(b, B)) return s;
    return batch_end(b);
  });
}

vt_status vt_render_batch_capture(vt_render_batch* b, const vt_frame* frames) {
  if (!b) return VT_ERR_INVALID_ARGUMENT;
  return vt_guarded(b->host, [&]() -> vt_status {
    vt_engine* e = b->host;
    HIP_CHECK(e, hipSetDevice(e->device));
    batch_drop_graph(b);
    BatchLaunch B;
    if (vt_status s = batch_prepare(b, frames, true, B)) return s;
    HIP_CHECK(e, hipStreamSynchronize(e->stream));  // the table uploads
    HIP_CHECK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    vt_status s = batch_enqueue(b, B);
    hipGraph_t g = nullptr;
    const hipError_t r = hipStreamEndCapture(e->stream, &g);
    if (s) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    if (r != hipSuccess || !g) return batch_err(b, VT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(r));
    b->graph = g;
    HIP_CHECK(e, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->frames.assign(frames, frames + b->members.size());
    b->state = batch_state(b);
    if (!b->staged) HIP_CHECK(e, hipHostMalloc((void**)&b->staged, 3 * 1024 * sizeof(uint32_t), hipHostMallocDefault));
    return VT_OK;
  });
}

vt_status vt_render_batch_replay(vt_render_batch* b, uint32_t frame_index) {
  if (!b) return VT_ERR_INVALID_ARGUMENT;
  return vt_guarded(b->host, [&]() -> vt_status {
    vt_engine* e = b->host;
    if (!b->exec) return batch_err(b, VT_ERR_STATE, "no batched frame has been captured");
    if (batch_state(b) != b->state) return batch_err(b, VT_ERR_STATE, "a member's scene or viewport changed; capture again");
    HIP_CHECK(e, hipSetDevice(e->device));
    const uint32_t K = (uint32_t)b->members.size();
    BatchLaunch B;
    B.L.resize(K);
    for (uint32_t k = 0; k < K; ++k) {
      vt_engine* m = b->members[k];
      B.L[k].ps.accum = (float*)m->render->accum.ptr;
      B.L[k].ps.moments = (float*)m->render->moments.ptr;
      B.L[k].f.pixels = b->frames[k].width * b->frames[k].height;
    }
    if (vt_status s = batch_begin(b, B)) return s;
    // this frame's words, staged in pinned memory once the previous copy is done
    HIP_CHECK(e, hipEventSynchronize(b->staged_done));
    for (uint32_t k = 0; k < K; ++k) {
      b->staged[3 * k] = b->frames[k].frame_seed;
      b->staged[3 * k + 1] = frame_index;
      b->staged[3 * k + 2] = vts::frame_key(b->frames[k].frame_seed, frame_index);
    }
    HIP_CHECK(e, hipMemcpyAsync(b->words.ptr, b->staged, (size_t)K * 12, hipMemcpyHostToDevice, e->stream));
    HIP_CHECK(e, hipEventRecord(b->staged_done, e->stream));
    HIP_CHECK(e, hipGraphLaunch(b->exec, e->stream));
    return batch_end(b);
  });
}

}  // extern "C"
