// nrc.cpp -- C ABI of the neural radiance cache (kernels in nrc.hip).
#include <cmath>
#include <cstring>
#include <vector>

#include "engine_internal.h"
#include "nrc.h"

struct NrcState {
  DeviceBuffer master, m, v, grad, frags, loss;
  DeviceBuffer part;  // per-workgroup partial MLP weight gradients of the training kernel
  DeviceBuffer dx;    // per-sample gradients of the hash-grid inputs (16 floats per sample)
  DeviceBuffer feats, targets, out;  // staging for the host-buffer entry points
  vt_nrc_options opts{};
  uint32_t step = 0;
  // HIP events around the last inference and the last training launch.
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool timed_infer = false, timed_train = false;
  // The path tracer's training step runs on `train` (default priority) so it
  // overlaps the start of the next frame; `trained` marks its end, and every
  // later use of the cache on the engine stream waits for it first.
  hipStream_t train = nullptr;
  hipEvent_t ready = nullptr, trained = nullptr;
  bool train_pending = false;
  void release() {
    if (train) (void)hipStreamSynchronize(train);
    for (DeviceBuffer* b : {&master, &m, &v, &grad, &frags, &loss, &part, &dx, &feats, &targets, &out}) b->release();
    for (hipEvent_t& x : ev)
      if (x) (void)hipEventDestroy(x), x = nullptr;
    for (hipEvent_t* x : {&ready, &trained})
      if (*x) (void)hipEventDestroy(*x), *x = nullptr;
    if (train) (void)hipStreamDestroy(train), train = nullptr;
  }
};

void engine_release_nrc(vt_engine* e) {
  if (!e->nrc) return;
  e->nrc->release();
  delete e->nrc;
  e->nrc = nullptr;
}

static vt_status nrc_err(vt_engine* e, vt_status s, const char* msg) {
  e->last_error = msg;
  return s;
}

#define NRC_HIP(e, call)                                        \
  do {                                                          \
    hipError_t _err = (call);                                   \
    if (_err != hipSuccess) {                                   \
      (e)->last_error = std::string(#call ": ") + hipGetErrorString(_err); \
      return VT_ERR_HIP;                                        \
    }                                                           \
  } while (0)

// Xavier-uniform initial weights from a PCG stream (deterministic per seed).
static void init_weights(uint32_t seed, std::vector<float>& w) {
  w.assign(vtn::kParams, 0.0f);
  uint64_t st = 0x853c49e6748fea9bull ^ ((uint64_t)seed << 1);
  auto next = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    const uint32_t x = (uint32_t)(((st >> 18) ^ st) >> 27), rot = (uint32_t)(st >> 59);
    return (x >> rot) | (x << ((32u - rot) & 31u));
  };
  auto fill = [&](size_t off, uint32_t rows, uint32_t cols) {
    const double a = std::sqrt(6.0 / (double)(rows + cols));
    for (size_t i = 0; i < (size_t)rows * cols; ++i)
      w[off + i] = (float)((((double)next() / 4294967296.0) * 2.0 - 1.0) * a);
  };
  for (uint32_t l = 0; l < vtn::kHidden; ++l) fill((size_t)l * 4096, 64, 64);
  fill(16384, vtn::kOut, 64);
  // Hash-grid tables: U(-1e-4, 1e-4) (nrcSubsystem.ts:541, Instant-NGP section 3).
  for (size_t i = 0; i < vtn::kTableParams; ++i)
    w[vtn::kMlpParams + i] = (float)((((double)next() / 4294967296.0) * 2.0 - 1.0) * 1e-4);
}

// Makes the engine stream wait for a training step still running on the
// training stream.
static vt_status nrc_join(vt_engine* e) {
  NrcState* s = e->nrc;
  if (!s || !s->train_pending) return VT_OK;
  s->train_pending = false;
  NRC_HIP(e, hipStreamWaitEvent(e->stream, s->trained, 0));
  return VT_OK;
}

vt_status engine_nrc_join(vt_engine* e) { return nrc_join(e); }

vt_status engine_nrc_quiesce(vt_engine* e) {
  if (e->nrc && e->nrc->train) NRC_HIP(e, hipStreamSynchronize(e->nrc->train));
  return VT_OK;
}

static vt_status nrc_ensure(vt_engine* e) {
  if (e->nrc) return nrc_join(e);
  NrcState* s = new NrcState();
  s->opts.learning_rate = 1e-2f;
  s->opts.beta1 = 0.9f;
  s->opts.beta2 = 0.999f;
  s->opts.epsilon = 1e-8f;
  s->opts.seed = 1;
  s->opts.table_learning_rate = 1e-1f;
  e->nrc = s;
  vt_status st = VT_OK;
  const size_t pb = vtn::kParams * sizeof(float);
  for (DeviceBuffer* b : {&s->master, &s->m, &s->v, &s->grad}) st = st ? st : b->ensure(e, pb);
  st = st ? st : s->frags.ensure(e, vtn::kFrags * 16);
  st = st ? st : s->loss.ensure(e, 64);
  st = st ? st : s->part.ensure(e, (size_t)vtn::kGradSlots * vtn::kMlpParams * sizeof(float));
  for (hipEvent_t& x : s->ev)
    if (!st && hipEventCreate(&x) != hipSuccess) st = VT_ERR_HIP;
  for (hipEvent_t* x : {&s->ready, &s->trained})
    if (!st && hipEventCreateWithFlags(x, hipEventDisableTiming) != hipSuccess) st = VT_ERR_HIP;
  if (!st && hipStreamCreateWithFlags(&s->train, hipStreamNonBlocking) != hipSuccess) st = VT_ERR_HIP;
  if (st) {
    engine_release_nrc(e);
    return st;
  }
  return vt_nrc_reset(e, nullptr);
}

static vt_status upload_master(vt_engine* e, const float* w) {
  NrcState& s = *e->nrc;
  NRC_HIP(e, hipSetDevice(e->device));
  NRC_HIP(e, hipMemcpyAsync(s.master.ptr, w, vtn::kParams * sizeof(float), hipMemcpyHostToDevice, e->stream));
  NRC_HIP(e, vtn::launch_fill((float*)s.m.ptr, vtn::kParams, 0.0f, e->stream));
  NRC_HIP(e, vtn::launch_fill((float*)s.v.ptr, vtn::kParams, 0.0f, e->stream));
  NRC_HIP(e, vtn::launch_fill((float*)s.grad.ptr, vtn::kParams, 0.0f, e->stream));
  NRC_HIP(e, vtn::launch_pack((const float*)s.master.ptr, (uint4*)s.frags.ptr, e->stream));
  NRC_HIP(e, hipStreamSynchronize(e->stream));
  s.step = 0;
  return VT_OK;
}

static bool finite_features(const float* f, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(f[i])) return false;
  return true;
}

static vt_status nrc_apply_adam(vt_engine* e, hipStream_t st);
static vt_status nrc_accumulate(vt_engine* e, const float* d_features, const float* d_targets, uint32_t n,
                                const uint32_t* d_count, hipStream_t st);

vt_status engine_nrc_ensure(vt_engine* e) { return nrc_ensure(e); }

// A new cache has no training step in flight, so creating one joins nothing.
vt_status engine_nrc_create(vt_engine* e) { return e->nrc ? VT_OK : nrc_ensure(e); }

vt_status engine_nrc_infer_counted(vt_engine* e, const float* d_feat, const uint32_t* d_count, float* d_rgb,
                                   uint32_t capacity) {
  if (vt_status s = nrc_ensure(e)) return s;
  NRC_HIP(e, vtn::launch_infer(d_feat, capacity, d_count, (const uint4*)e->nrc->frags.ptr,
                               (const float*)e->nrc->master.ptr + vtn::kMlpParams, d_rgb, e->stream));
  return VT_OK;
}

vt_status engine_nrc_train_counted(vt_engine* e, const float* d_feat, const float* d_tgt, const uint32_t* d_count,
                                   uint32_t capacity) {
  if (vt_status s = nrc_ensure(e)) return s;
  if (capacity == 0) return VT_OK;
  NrcState& n = *e->nrc;
  // Fork: the step reads the records written so far on the engine stream.
  NRC_HIP(e, hipEventRecord(n.ready, e->stream));
  NRC_HIP(e, hipStreamWaitEvent(n.train, n.ready, 0));
  if (vt_status s = nrc_accumulate(e, d_feat, d_tgt, capacity, d_count, n.train)) return s;
  if (vt_status s = nrc_apply_adam(e, n.train)) return s;
  NRC_HIP(e, hipEventRecord(n.trained, n.train));
  n.train_pending = true;
  return VT_OK;
}

extern "C" {

vt_status vt_nrc_param_count(uint32_t* count) {
  return vt_guarded(nullptr, [&]() -> vt_status {
    if (!count) return VT_ERR_INVALID_ARGUMENT;
    *count = vtn::kParams;
    return VT_OK;
  });
}

vt_status vt_nrc_reset(vt_engine* e, const vt_nrc_options* opts) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e) return VT_ERR_INVALID_ARGUMENT;
    if (!e->nrc) {
      vt_status s = nrc_ensure(e);  // allocates, then re-enters with defaults
      if (s || !opts) return s;
    }
    if (vt_status st = nrc_join(e)) return st;
    NrcState& s = *e->nrc;
    if (opts) {
      if (!(opts->learning_rate > 0) || !(opts->beta1 >= 0 && opts->beta1 < 1) ||
          !(opts->beta2 >= 0 && opts->beta2 < 1) || !(opts->epsilon > 0) || !(opts->table_learning_rate >= 0))
        return nrc_err(e, VT_ERR_RANGE,
                       "learning_rate > 0, table_learning_rate >= 0, 0 <= beta1, beta2 < 1 and epsilon > 0 are required");
      s.opts = *opts;
      if (s.opts.table_learning_rate == 0.0f) s.opts.table_learning_rate = 1e-1f;
    }
    std::vector<float> w;
    init_weights(s.opts.seed, w);
    return upload_master(e, w.data());
  });
}

vt_status vt_nrc_set_weights(vt_engine* e, const float* w, uint32_t count) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || !w) return VT_ERR_INVALID_ARGUMENT;
    if (count != vtn::kParams) return nrc_err(e, VT_ERR_RANGE, "weight count must equal vt_nrc_param_count()");
    if (!finite_features(w, count)) return nrc_err(e, VT_ERR_RANGE, "weights must be finite");
    if (vt_status s = nrc_ensure(e)) return s;
    return upload_master(e, w);
  });
}

vt_status vt_nrc_get_weights(vt_engine* e, float* w, uint32_t count) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || !w) return VT_ERR_INVALID_ARGUMENT;
    if (count != vtn::kParams) return nrc_err(e, VT_ERR_RANGE, "weight count must equal vt_nrc_param_count()");
    if (vt_status s = nrc_ensure(e)) return s;
    NRC_HIP(e, hipMemcpyAsync(w, e->nrc->master.ptr, count * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    NRC_HIP(e, hipStreamSynchronize(e->stream));
    return VT_OK;
  });
}

vt_status vt_nrc_infer_device(vt_engine* e, const float* d_features, uint32_t n, float* d_rgb) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || (n && (!d_features || !d_rgb))) return VT_ERR_INVALID_ARGUMENT;
    if (vt_status s = nrc_ensure(e)) return s;
    NRC_HIP(e, hipSetDevice(e->device));
    NRC_HIP(e, hipEventRecord(e->nrc->ev[0], e->stream));
    NRC_HIP(e, vtn::launch_infer(d_features, n, nullptr, (const uint4*)e->nrc->frags.ptr,
                                 (const float*)e->nrc->master.ptr + vtn::kMlpParams, d_rgb, e->stream));
    NRC_HIP(e, hipEventRecord(e->nrc->ev[1], e->stream));
    e->nrc->timed_infer = true;
    return VT_OK;
  });
}

vt_status vt_nrc_infer(vt_engine* e, const float* features, uint32_t n, float* rgb) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || (n && (!features || !rgb))) return VT_ERR_INVALID_ARGUMENT;
    if (!finite_features(features, (size_t)n * vtn::kFeat)) return nrc_err(e, VT_ERR_RANGE, "features must be finite");
    if (vt_status s = nrc_ensure(e)) return s;
    NrcState& s = *e->nrc;
    if (vt_status st = s.feats.ensure(e, (size_t)n * vtn::kFeat * 4 + 64)) return st;
    if (vt_status st = s.out.ensure(e, (size_t)n * 12 + 64)) return st;
    NRC_HIP(e, hipMemcpyAsync(s.feats.ptr, features, (size_t)n * vtn::kFeat * 4, hipMemcpyHostToDevice, e->stream));
    if (vt_status st = vt_nrc_infer_device(e, (const float*)s.feats.ptr, n, (float*)s.out.ptr)) return st;
    NRC_HIP(e, hipMemcpyAsync(rgb, s.out.ptr, (size_t)n * 12, hipMemcpyDeviceToHost, e->stream));
    NRC_HIP(e, hipStreamSynchronize(e->stream));
    return VT_OK;
  });
}

// Accumulates the batch gradient into the engine's gradient buffer and the
// batch loss into loss[0] (both device memory of the engine).
static vt_status nrc_accumulate(vt_engine* e, const float* d_features, const float* d_targets, uint32_t n,
                                const uint32_t* d_count, hipStream_t stream) {
  NrcState& s = *e->nrc;
  NRC_HIP(e, hipSetDevice(e->device));
  if (vt_status st = s.dx.ensure(e, (size_t)n * 16u * sizeof(float))) return st;
  NRC_HIP(e, vtn::launch_fill((float*)s.loss.ptr, 1, 0.0f, stream));
  NRC_HIP(e, hipEventRecord(s.ev[2], stream));
  NRC_HIP(e, vtn::launch_train(d_features, d_targets, n, d_count, (const uint4*)s.frags.ptr,
                               (const float*)s.master.ptr + vtn::kMlpParams, (float*)s.grad.ptr, (float*)s.part.ptr,
                               (float*)s.dx.ptr, (float*)s.loss.ptr, stream));
  NRC_HIP(e, hipEventRecord(s.ev[3], stream));
  s.timed_train = true;
  return VT_OK;
}

static vt_status nrc_apply_adam(vt_engine* e, hipStream_t stream) {
  NrcState& s = *e->nrc;
  s.step += 1;
  vtn::AdamParams p{};
  p.lr = s.opts.learning_rate;
  p.lr_table = s.opts.table_learning_rate;
  p.beta1 = s.opts.beta1;
  p.beta2 = s.opts.beta2;
  p.eps = s.opts.epsilon;
  p.bc1 = (float)(1.0 - std::pow((double)s.opts.beta1, (double)s.step));
  p.bc2 = (float)(1.0 - std::pow((double)s.opts.beta2, (double)s.step));
  NRC_HIP(e, vtn::launch_adam((float*)s.master.ptr, (float*)s.grad.ptr, (float*)s.m.ptr, (float*)s.v.ptr, p,
                              stream));
  NRC_HIP(e, vtn::launch_pack((const float*)s.master.ptr, (uint4*)s.frags.ptr, stream));
  return VT_OK;
}

vt_status vt_nrc_train_step_device(vt_engine* e, const float* d_features, const float* d_targets, uint32_t n) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || (n && (!d_features || !d_targets))) return VT_ERR_INVALID_ARGUMENT;
    if (n == 0) return VT_OK;
    if (vt_status s = nrc_ensure(e)) return s;
    if (vt_status s = nrc_accumulate(e, d_features, d_targets, n, nullptr, e->stream)) return s;
    return nrc_apply_adam(e, e->stream);
  });
}

vt_status vt_nrc_timing(vt_engine* e, float* infer_ms, float* train_ms) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || !infer_ms || !train_ms) return VT_ERR_INVALID_ARGUMENT;
    *infer_ms = *train_ms = 0.0f;
    if (!e->nrc) return VT_OK;
    NrcState& s = *e->nrc;
    if (s.timed_infer) {
      NRC_HIP(e, hipEventSynchronize(s.ev[1]));
      NRC_HIP(e, hipEventElapsedTime(infer_ms, s.ev[0], s.ev[1]));
    }
    if (s.timed_train) {
      NRC_HIP(e, hipEventSynchronize(s.ev[3]));
      NRC_HIP(e, hipEventElapsedTime(train_ms, s.ev[2], s.ev[3]));
    }
    return VT_OK;
  });
}

vt_status vt_nrc_last_loss(vt_engine* e, float* loss) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || !loss) return VT_ERR_INVALID_ARGUMENT;
    if (vt_status s = nrc_ensure(e)) return s;
    NRC_HIP(e, hipMemcpyAsync(loss, e->nrc->loss.ptr, 4, hipMemcpyDeviceToHost, e->stream));
    NRC_HIP(e, hipStreamSynchronize(e->stream));
    return VT_OK;
  });
}

static vt_status stage_batch(vt_engine* e, const float* features, const float* targets, uint32_t n) {
  if (!finite_features(features, (size_t)n * vtn::kFeat) || !finite_features(targets, (size_t)n * 3))
    return nrc_err(e, VT_ERR_RANGE, "features and targets must be finite");
  NrcState& s = *e->nrc;
  if (vt_status st = s.feats.ensure(e, (size_t)n * vtn::kFeat * 4 + 64)) return st;
  if (vt_status st = s.targets.ensure(e, (size_t)n * 12 + 64)) return st;
  NRC_HIP(e, hipMemcpyAsync(s.feats.ptr, features, (size_t)n * vtn::kFeat * 4, hipMemcpyHostToDevice, e->stream));
  NRC_HIP(e, hipMemcpyAsync(s.targets.ptr, targets, (size_t)n * 12, hipMemcpyHostToDevice, e->stream));
  return VT_OK;
}

vt_status vt_nrc_train_step(vt_engine* e, const float* features, const float* targets, uint32_t n, float* loss) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || (n && (!features || !targets))) return VT_ERR_INVALID_ARGUMENT;
    if (vt_status s = nrc_ensure(e)) return s;
    if (n == 0) return VT_OK;
    if (vt_status s = stage_batch(e, features, targets, n)) return s;
    NrcState& s = *e->nrc;
    if (vt_status st = vt_nrc_train_step_device(e, (const float*)s.feats.ptr, (const float*)s.targets.ptr, n))
      return st;
    if (loss) return vt_nrc_last_loss(e, loss);
    NRC_HIP(e, hipStreamSynchronize(e->stream));
    return VT_OK;
  });
}

vt_status vt_nrc_gradient(vt_engine* e, const float* features, const float* targets, uint32_t n, float* grad,
                          float* loss) {
  return vt_guarded(e, [&]() -> vt_status {
    if (!e || !grad || (n && (!features || !targets))) return VT_ERR_INVALID_ARGUMENT;
    if (vt_status s = nrc_ensure(e)) return s;
    NrcState& s = *e->nrc;
    NRC_HIP(e, hipSetDevice(e->device));
    NRC_HIP(e, vtn::launch_fill((float*)s.grad.ptr, vtn::kParams, 0.0f, e->stream));
    if (n) {
      if (vt_status st = stage_batch(e, features, targets, n)) return st;
      if (vt_status st = nrc_accumulate(e, (const float*)s.feats.ptr, (const float*)s.targets.ptr, n, nullptr, e->stream))
        return st;
    } else {
      NRC_HIP(e, vtn::launch_fill((float*)s.loss.ptr, 1, 0.0f, e->stream));
    }
    NRC_HIP(e, hipMemcpyAsync(grad, s.grad.ptr, vtn::kParams * 4, hipMemcpyDeviceToHost, e->stream));
    if (loss) NRC_HIP(e, hipMemcpyAsync(loss, s.loss.ptr, 4, hipMemcpyDeviceToHost, e->stream));
    NRC_HIP(e, vtn::launch_fill((float*)s.grad.ptr, vtn::kParams, 0.0f, e->stream));
    NRC_HIP(e, hipStreamSynchronize(e->stream));
    return VT_OK;
  });
}

}  // extern "C"
