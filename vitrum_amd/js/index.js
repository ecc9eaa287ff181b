'use strict';
// pt-hip: the MI355X path-tracing backend behind vitrum's Engine contract.
//
// createPTEngine_HIP(opts) is the EngineFactory of this backend (the pt-webgpu
// one is createPTEngine_WebGPU, packages/pt-webgpu/src/index.ts:2363), and
// constructPathTracerHipForDispatch has the BackendConstructor signature of
// packages/engine/src/createEngineInternals.ts:452 so it can be registered in
// BACKEND_CONSTRUCTORS (packages/engine/src/createEngine.ts:99-103).
//
// The returned object implements packages/core/src/engine/index.ts:109-486:
// state, capabilities, setScene, getScene, renderFrame, reset, pause, resume,
// dispose, onFrame/onProgress/onError/onWarning and captureFrame.  All device
// work goes through the N-API addon (../napi/vitrum_hip.node) into
// libvthip.so; there is no JavaScript fallback renderer.

const path = require('path');

let addon = null;
function loadAddon() {
  if (addon) return addon;
  const file = path.join(__dirname, '..', 'napi', 'vitrum_hip.node');
  try {
    addon = require(file);
  } catch (err) {
    const e = new Error(`pt-hip: native addon could not be loaded (${file}): ${err.message}`);
    e.code = 'VT_ERR_NO_ADDON';
    throw e;
  }
  return addon;
}

// multiplyMat4 (packages/pt-webgpu/src/math/mat4.ts:4-18): column-major
// a * b, each element summed in binary64 and stored as binary32.
function multiplyMat4(a, b) {
  const out = new Float32Array(16);
  for (let col = 0; col < 4; col += 1) {
    for (let row = 0; row < 4; row += 1) {
      out[col * 4 + row] =
        (a[row] || 0) * (b[col * 4] || 0) +
        (a[4 + row] || 0) * (b[col * 4 + 1] || 0) +
        (a[8 + row] || 0) * (b[col * 4 + 2] || 0) +
        (a[12 + row] || 0) * (b[col * 4 + 3] || 0);
    }
  }
  return out;
}

function isMat4(m) {
  return m != null && typeof m.length === 'number' && m.length === 16;
}

function toF32(v, label) {
  if (v instanceof Float32Array) return v;
  if (v == null || typeof v.length !== 'number') throw new TypeError(`${label} must be a Float32Array or array`);
  return Float32Array.from(v);
}

function toU32(v, label) {
  if (v == null) return null;
  if (v instanceof Uint32Array) return v;
  if (v instanceof Uint16Array || Array.isArray(v)) return Uint32Array.from(v);
  throw new TypeError(`${label} must be a Uint32Array, Uint16Array or array`);
}

// Environment maps (packages/shared-samplers/src/environmentMapPixels.ts:
// 103-160): raw {width, height, data} or DataTexture-shaped {image: {...}}
// payloads, with optional channels / dataType / colorSpace hints (direct or
// under __vitrum_hint__), decoded to linear RGB binary32 texels.
function halfToFloat(h) {
  const s = (h & 0x8000) ? -1 : 1;
  const e = (h & 0x7c00) >> 10;
  const f = h & 0x03ff;
  if (e === 0) return s * 2 ** -14 * (f / 1024);
  if (e === 0x1f) return f ? NaN : s * Infinity;
  return s * 2 ** (e - 15) * (1 + f / 1024);
}

function srgbToLinear(c) {
  return c <= 0.04045 ? c / 12.92 : Math.pow((c + 0.055) / 1.055, 2.4);
}

function mapHint(map) {
  if (map.__vitrum_hint__ != null) return map.__vitrum_hint__;
  const h = {};
  if ([1, 2, 3, 4].includes(map.channels)) h.channels = map.channels;
  if (['uint8', 'uint16', 'float16', 'half-float', 'float32'].includes(map.dataType)) h.dataType = map.dataType;
  if (map.colorSpace === 'linear' || map.colorSpace === 'srgb') h.colorSpace = map.colorSpace;
  return h;
}

function decodeScalar(src, i, hint) {
  const raw = Number(src[i] == null ? 0 : src[i]);
  switch (hint.dataType) {
    case 'uint16': return raw / 65535;
    case 'float16': case 'half-float': return halfToFloat(raw);
    case 'float32': return raw;
    case 'uint8': return raw / 255;
    default: break;
  }
  if (src instanceof Uint16Array) return raw / 65535;
  if (src instanceof Float32Array || src instanceof Float64Array || Array.isArray(src)) return raw;
  const bpe = src.BYTES_PER_ELEMENT || 1;
  return raw / (2 ** (8 * bpe) - 1);
}

function environmentMapTexels(map) {
  const image = map.image || {};
  const pick = (k) => (map[k] != null ? map[k] : image[k]);
  const src = pick('data');
  const width = Number(pick('width') || 0);
  const height = Number(pick('height') || 0);
  if (src == null || !Number.isSafeInteger(width) || !Number.isSafeInteger(height) || width <= 0 || height <= 0) {
    throw new TypeError('environment.hdri must carry {constant:[r,g,b]} or {width,height,data}');
  }
  const pixels = width * height;
  if (pixels * 4 > 0x7fffffff || src.length < pixels) throw new RangeError('environment.hdri data is too short or too large');
  const hint = mapHint(map);
  const per = src.length / pixels;
  const ch = hint.channels != null ? hint.channels : (per >= 4 ? 4 : per >= 3 ? 3 : per >= 2 ? 2 : 1);
  if (src.length < pixels * ch) throw new RangeError('environment.hdri data is shorter than width x height x channels');
  const srgb = hint.colorSpace === 'srgb';
  const texels = new Float32Array(pixels * 3);
  for (let p = 0; p < pixels; p += 1) {
    const b0 = p * ch;
    const r = decodeScalar(src, b0, hint);
    const rgb = [r, ch > 1 ? decodeScalar(src, b0 + 1, hint) : r, ch > 2 ? decodeScalar(src, b0 + 2, hint) : r];
    for (let k = 0; k < 3; k += 1) {
      const v = srgb ? srgbToLinear(rgb[k]) : rgb[k];
      if (!Number.isFinite(v) || !Number.isFinite(Math.fround(v)) || v < 0) {
        throw new RangeError(`environment.hdri texel ${p} must contain finite, non-negative radiance`);
      }
      texels[p * 3 + k] = v;
    }
  }
  return { width, height, texels };
}

// HdriEnvironment (packages/core/src/scene/environment.ts:21-56): a constant
// radiance {constant: [r, g, b]} (kind 1) or an equirectangular map (kind 2)
// with intensity and rotationY.
// An environment scale as the binary32 lane takes it
// (environmentRadianceScale.ts:81-98).
function environmentScale(value, label) {
  const v = value == null ? 1 : value;
  if (!Number.isFinite(v) || v < 0) throw new RangeError(`${label} must be finite and >= 0`);
  const packed = Math.fround(v);
  if (!Number.isFinite(packed) || (v > 0 && packed === 0)) {
    throw new RangeError(`${label} must remain finite and positive after binary32 packing`);
  }
  return packed;
}

function environmentOf(env) {
  if (env == null || env.kind === 'none') return { kind: 0, radiance: [0, 0, 0] };
  if (env.kind === 'procedural-sky') {
    // ProceduralSkyEnvironment (environment.ts:59-71); the addon fills the
    // defaults of environmentPacking.ts:245-262 for absent fields
    const out = { kind: 3, radiance: [0, 0, 0], intensity: environmentScale(env.intensity, 'procedural-sky intensity') };
    for (const k of ['turbidity', 'rayleigh', 'mieCoefficient', 'mieDirectionalG']) {
      if (env[k] != null) {
        if (!Number.isFinite(env[k])) throw new RangeError(`procedural-sky ${k} must be finite`);
        out[k] = env[k];
      }
    }
    if (env.sunDirection != null) {
      const d = Array.from(env.sunDirection);
      if (d.length !== 3 || !d.every(Number.isFinite)) throw new RangeError('procedural-sky sunDirection must be 3 finite numbers');
      out.sunDirection = d;
    }
    return out;
  }
  if (env.kind !== 'hdri') {
    throw new RangeError(`pt-hip renders 'none', 'hdri' and 'procedural-sky' environments, not '${env.kind}'`);
  }
  const map = env.hdri || {};
  const intensity = environmentScale(env.intensity, 'environment.intensity');
  let rotationY = env.rotationY == null ? 0 : env.rotationY;
  if (!Number.isFinite(rotationY)) throw new RangeError('environment.rotationY must be finite');
  // canonicalizeEnvironmentRotationF32 (environmentRotation.ts:10-26)
  rotationY = Math.fround(rotationY % (2 * Math.PI));
  if (Object.is(rotationY, -0)) rotationY = 0;
  if (Array.isArray(map.constant) || map.constant instanceof Float32Array) {
    const radiance = [map.constant[0], map.constant[1], map.constant[2]].map((c) => c * intensity);
    if (!radiance.every((c) => Number.isFinite(c) && c >= 0)) throw new RangeError('environment radiance must be finite and >= 0');
    return { kind: radiance.some((c) => c > 0) ? 1 : 0, radiance };
  }
  const { width, height, texels } = environmentMapTexels(map);
  return { kind: 2, radiance: [0, 0, 0], texels, width, height, channels: 3, intensity, rotationY };
}

// Scene (packages/core/src/scene/index.ts:42) -> the addon's scene blob.
function sceneBlob(scene, opts) {
  if (scene == null || !Array.isArray(scene.primitives)) throw new TypeError('scene.primitives must be an array');
  const primitives = [];
  const materials = [];
  scene.primitives.forEach((p, i) => {
    if (p.kind !== 'mesh' && p.kind !== 'instanced-mesh') {
      throw new RangeError(`pt-hip supports 'mesh' and 'instanced-mesh' primitives, not '${p.kind}' (primitive ${i})`);
    }
    const m = p.material || {};
    materials.push({
      baseColor: m.baseColor || [0.8, 0.8, 0.8],
      roughness: m.roughness == null ? 0.5 : m.roughness,
      metallic: m.metallic == null ? 0 : m.metallic,
      emissive: m.emissive || [0, 0, 0],
      emissiveIntensity: m.emissiveIntensity == null ? 1 : m.emissiveIntensity,
      doubleSided: !!m.doubleSided,
    });
    // extension lobes (MaterialSpec clearcoat / sheen / anisotropy / specular)
    for (const k of ['clearcoat', 'clearcoatRoughness', 'sheen', 'sheenRoughness', 'sheenColor', 'anisotropy',
      'anisotropyRotation', 'specularColor', 'specularIntensity']) {
      if (m[k] != null) materials[materials.length - 1][k] = m[k];
    }
    for (const k of ['transmission', 'iridescence']) {
      if (m[k] != null && m[k] > 0) throw new RangeError(`pt-hip renders opaque materials: ${k} is not supported (primitive ${i})`);
    }
    let transforms = null;
    if (p.kind === 'mesh') {
      if (p.transform != null) {
        if (!isMat4(p.transform)) throw new TypeError(`primitive ${i}: transform must be a mat4`);
        transforms = toF32(p.transform, 'transform');
      }
    } else {
      if (!Array.isArray(p.instances)) throw new TypeError(`primitive ${i}: instances must be an array of mat4`);
      transforms = new Float32Array(16 * p.instances.length);
      p.instances.forEach((t, k) => {
        if (!isMat4(t)) throw new TypeError(`primitive ${i}: instance ${k} must be a mat4`);
        transforms.set(t, 16 * k);
      });
    }
    primitives.push({
      kind: p.kind === 'mesh' ? 0 : 1,
      material: i,
      positions: toF32(p.positions, 'positions'),
      normals: p.normals == null ? null : toF32(p.normals, 'normals'),
      indices: toU32(p.indices, 'indices'),
      transforms,
    });
  });
  return {
    primitives,
    materials,
    environment: environmentOf(scene.environment),
    sceneRadius: opts && opts.sceneRadius ? opts.sceneRadius : 0,
    emitters: emittersOf(scene.emitters),
  };
}

// Scene.emitters (core/src/scene/emitters.ts:7-114): the analytic kinds, in
// scene order (the library packs them by kind as pt-webgpu does).  Mesh-area
// emitters are the emissive materials of mesh primitives.
const EMITTER_KINDS = { directional: 0, point: 1, spot: 2, 'rect-area': 3, 'disc-area': 4 };

function emittersOf(list) {
  if (list == null) return [];
  if (!Array.isArray(list)) throw new TypeError('scene.emitters must be an array');
  const out = [];
  list.forEach((e, i) => {
    if (e == null || typeof e !== 'object') throw new TypeError(`scene.emitters[${i}] must be an object`);
    if (e.kind === 'mesh-area') return;  // lit through the referenced mesh's emissive material
    if (!(e.kind in EMITTER_KINDS)) throw new RangeError(`scene.emitters[${i}].kind '${e.kind}' is not supported`);
    const vec = (v, name, def) => {
      if (v == null) return def;
      if (v.length !== 3 || !Array.prototype.every.call(v, Number.isFinite)) {
        throw new TypeError(`scene.emitters[${i}].${name} must be three finite numbers`);
      }
      return [v[0], v[1], v[2]];
    };
    const num = (v, name, def) => {
      if (v == null) return def;
      if (!Number.isFinite(v)) throw new TypeError(`scene.emitters[${i}].${name} must be a finite number`);
      return v;
    };
    out.push({
      kind: EMITTER_KINDS[e.kind],
      color: vec(e.color, 'color', [1, 1, 1]),
      intensity: num(e.intensity, 'intensity', 1),
      castShadow: e.castShadow !== false,
      position: vec(e.position, 'position', [0, 0, 0]),
      direction: vec(e.direction, 'direction', [0, -1, 0]),
      normal: vec(e.normal, 'normal', [0, -1, 0]),
      uAxis: vec(e.uAxis, 'uAxis', [0, 0, 0]),
      vAxis: vec(e.vAxis, 'vAxis', [0, 0, 0]),
      radius: num(e.radius, 'radius', 0),
      angle: num(e.angle, 'angle', 0),
      penumbra: num(e.penumbra, 'penumbra', 0),
      distance: num(e.distance, 'distance', 0),
      decay: num(e.decay, 'decay', 2),
      angularDiameter: num(e.angularDiameter, 'angularDiameter', 0),
    });
  });
  return out;
}

const SKIPPED = Object.freeze({ kind: 'skipped', samplesAccumulated: 0, isConverged: false });

// Fields of a ScenePrimitivePatch this backend renders, per primitive kind
// (core/src/scene/primitives.ts:167-227).
const PRIMITIVE_PATCH_KEYS = {
  mesh: new Set(['positions', 'normals', 'indices', 'transform', 'material']),
  'instanced-mesh': new Set(['positions', 'normals', 'indices', 'instances', 'material']),
};

function primitiveIndexOf(scene, id, method) {
  const i = scene.primitives.findIndex((p) => String(p.id) === String(id));
  if (i < 0) throw new Error(`${method}: primitive "${id}" not found in current scene`);
  return i;
}

// FrameQualitySettings.tonemap / exposure / outputColorSpace (core/src/
// frame.ts:32-105): defaults aces, 1, srgb (present.wgsl.ts:10-11).
const TONEMAP_MODES = { aces: 0, agx: 1, reinhard: 2, linear: 3, none: 4 };
function presentOf(q) {
  const tm = q.tonemap == null ? 'aces' : q.tonemap;
  if (!(tm in TONEMAP_MODES)) throw new RangeError(`quality.tonemap '${tm}' is not supported`);
  const exposure = q.exposure == null ? 1 : q.exposure;
  if (!Number.isFinite(exposure) || exposure < 0) throw new RangeError('quality.exposure must be finite and >= 0');
  const cs = q.outputColorSpace == null ? 'srgb' : q.outputColorSpace;
  if (cs !== 'srgb' && cs !== 'linear') throw new RangeError(`quality.outputColorSpace '${cs}' is not supported`);
  return { tonemap: TONEMAP_MODES[tm], exposure, colorSpace: cs === 'srgb' ? 0 : 1 };
}

class HipPathTracerEngine {
  constructor(native, handle, opts) {
    this._n = native;
    this._h = handle;
    this._opts = opts;
    this._scene = null;
    this._nrc = null;
    this._size = null;
    this._samples = 0;
    this._listeners = { frame: new Set(), progress: new Set(), error: new Set(), warning: new Set() };
    this.state = 'ready';
    this.backendId = 'pt-hip';
    this.capabilities = Object.freeze({
      supportsIncrementalScene: true,
      supportsAddRemovePrimitive: true,
      supportsAuxBuffers: false,
      accumulates: true,
      maxSamplesPerPixel: opts.maxSamplesPerPixel || Infinity,
    });
  }

  _live() {
    if (this.state === 'disposed') throw new Error('pt-hip: engine has been disposed');
    if (this.state === 'error') throw new Error('pt-hip: engine is in the error state; dispose and recreate it');
  }

  _emit(kind, value) {
    for (const cb of this._listeners[kind]) {
      try {
        cb(value);
      } catch (_) {
        // listener failures never reach the renderer
      }
    }
  }

  _fail(err) {
    const fatal = err && (err.code === 'VT_ERR_HIP' || err.code === 'VT_ERR_NO_DEVICE');
    if (fatal) this.state = 'error';
    this._emit('error', { kind: fatal ? 'gpu-internal' : 'render', message: String(err && err.message), fatal, raw: err });
  }

  // Engine.setScene (core/src/engine/index.ts:120): validates and converts
  // before any state changes; the previous scene stays bound on failure.
  setScene(scene) {
    this._live();
    const blob = sceneBlob(scene, this._opts);
    this._n.setScene(this._h, blob);
    this._scene = scene;
    this._samples = 0;
    if (this._nrc) {
      this._n.nrcConfig(this._h, this._nrc);
      this._nrc = keptNrc(this._nrc);  // seed / rate apply once
    }
  }

  // Radiance cache of the path tracer (the reference's nrcEnabled/nrcConfig,
  // WalkaroundGPUPipeline.ts:1860-1862).  nrcConfig.recordCap is the number
  // of self-training records gathered per frame (4 vertices per training
  // path); learningRate and seed re-initialise the cache.  Kept across
  // setScene; resets accumulation.
  setNeuralCache(enabled, config) {
    this._live();
    const nrc = nrcSettings(enabled, config);
    if (this._scene != null) {
      this._n.nrcConfig(this._h, nrc);
      // seed / learning rate apply once; later scenes keep the trained cache
      this._nrc = enabled ? keptNrc(nrc) : null;
    } else {
      this._nrc = enabled ? nrc : null;  // applied by setScene
    }
    this._samples = 0;
  }

  getScene() {
    return this._scene;
  }

  // Scene mutations (core/src/engine/index.ts:155-214).  Each validates its
  // arguments, derives the next scene copy-on-write (patchPrimitiveInScene,
  // core/src/scene/patchScene.ts:381: a material patch merges into the
  // material, other fields replace) and binds it through setScene, whose
  // result is what pt-webgpu's incremental paths produce (they equal a full
  // repack of the patched scene); accumulation restarts.  With
  // primitiveUpdates: 'refit', a patch of only positions (and normals) that
  // keeps the vertex count moves the vertices with the BVH topology kept
  // (the walkaround-hybrid refit path) instead; emissive primitives and
  // other patches take the setScene path.
  updatePrimitive(id, patch) {
    this._live();
    const scene = this._requireScene('updatePrimitive');
    const i = primitiveIndexOf(scene, id, 'updatePrimitive');
    if (patch == null || typeof patch !== 'object') throw new TypeError('updatePrimitive: patch must be an object');
    const cur = scene.primitives[i];
    const keys = [];
    for (const k of Object.keys(patch)) {
      if (k === 'id' || k === 'kind') {
        if (patch[k] !== undefined && patch[k] !== cur[k]) throw new TypeError(`updatePrimitive: '${k}' cannot change`);
        continue;
      }
      if (!PRIMITIVE_PATCH_KEYS[cur.kind] || !PRIMITIVE_PATCH_KEYS[cur.kind].has(k)) {
        throw new TypeError(`updatePrimitive: '${k}' is not a patchable field of a ${cur.kind} primitive`);
      }
      keys.push(k);
    }
    if (keys.length === 0) return;  // a validated empty patch is a no-op
    const next = Object.assign({}, cur, patch);
    if (patch.material != null) next.material = Object.assign({}, cur.material || {}, patch.material);
    const nextScene = Object.assign({}, scene, {
      primitives: scene.primitives.map((p, k) => (k === i ? next : p)),
    });
    if (this._opts.primitiveUpdates === 'refit' && keys.every((k) => k === 'positions' || k === 'normals')) {
      const positions = toF32(next.positions, 'positions');
      const normals = patch.normals == null ? null : toF32(next.normals, 'normals');
      try {
        this._n.refitPrimitive(this._h, i, { positions, normals });
        this._scene = nextScene;
        this._samples = 0;
        return;
      } catch (err) {
        if (err.code !== 'VT_ERR_RANGE' && err.code !== 'VT_ERR_STATE') throw err;
        this._emit('warning', { code: 'pt-hip.update-primitive-rebuild', backend: 'pt-hip', phase: 'mutation',
          method: 'updatePrimitive', message: `pt-hip: updatePrimitive("${id}") rebuilds the scene: ${err.message}` });
      }
    }
    this.setScene(nextScene);
  }

  addPrimitive(primitive) {
    this._live();
    const scene = this._requireScene('addPrimitive');
    if (primitive == null || typeof primitive !== 'object') throw new TypeError('addPrimitive: primitive must be an object');
    if (scene.primitives.some((p) => String(p.id) === String(primitive.id))) {
      throw new Error(`addPrimitive: primitive "${primitive.id}" already exists`);
    }
    this.setScene(Object.assign({}, scene, { primitives: scene.primitives.concat([primitive]) }));
  }

  removePrimitive(id) {
    this._live();
    const scene = this._requireScene('removePrimitive');
    const i = primitiveIndexOf(scene, id, 'removePrimitive');
    this.setScene(Object.assign({}, scene, { primitives: scene.primitives.filter((_, k) => k !== i) }));
  }

  updateEmitter(id, patch) {
    this._live();
    const scene = this._requireScene('updateEmitter');
    const list = scene.emitters || [];
    const i = list.findIndex((e) => String(e.id) === String(id));
    if (i < 0) throw new Error(`updateEmitter: emitter "${id}" not found in current scene`);
    if (patch == null || typeof patch !== 'object') throw new TypeError('updateEmitter: patch must be an object');
    const cur = list[i];
    if (patch.kind !== undefined && patch.kind !== cur.kind) {
      throw new Error(`updateEmitter: emitter "${id}" kind cannot change from "${cur.kind}" to "${patch.kind}"`);
    }
    if (patch.id !== undefined && patch.id !== cur.id) throw new Error(`updateEmitter: emitter "${id}" id cannot be changed`);
    if (Object.keys(patch).every((k) => k === 'id' || k === 'kind')) return;
    const next = Object.assign({}, cur, patch);
    this.setScene(Object.assign({}, scene, { emitters: list.map((e, k) => (k === i ? next : e)) }));
  }

  updateEnvironment(env) {
    this._live();
    const scene = this._requireScene('updateEnvironment');
    this.setScene(Object.assign({}, scene, { environment: env == null ? { kind: 'none' } : env }));
  }

  _requireScene(method) {
    if (this._scene == null) throw new Error(`${method}: no scene is set`);
    return this._scene;
  }

  // Engine.renderFrame (core/src/engine/index.ts:251): one sample per pixel,
  // accumulated into the running mean until quality.samplesTarget.
  renderFrame(input) {
    const a = this._frameArgs(input);
    if (a.done) return a.done;
    try {
      this._n.renderFrame(this._h, a.frame);
    } catch (err) {
      this._fail(err);
      throw err;
    }
    return this._frameDone(a.target);
  }

  // Validates a FrameInput; {done} when nothing is to be rendered, else the
  // addon's frame description and the sample target.
  _frameArgs(input) {
    this._live();
    if (this.state === 'paused' || this._scene == null) return { done: SKIPPED };
    if (input == null || !isMat4(input.viewMatrix) || !isMat4(input.projMatrix)) {
      throw new TypeError('FrameInput.viewMatrix and projMatrix must be mat4s');
    }
    const vp = input.viewport || {};
    const width = vp.width | 0, height = vp.height | 0;
    if (!(width > 0 && height > 0)) throw new RangeError('FrameInput.viewport must have positive width and height');
    const inv = this._n.invertMat4(multiplyMat4(input.projMatrix, input.viewMatrix));
    if (inv == null) throw new RangeError('projMatrix * viewMatrix is not invertible');
    const q = input.quality || {};
    const target = q.samplesTarget == null ? 0 : q.samplesTarget;
    this._present = presentOf(q);
    if (this._size && (this._size[0] !== width || this._size[1] !== height)) this._samples = 0;
    this._size = [width, height];
    if (target && this._samples >= target) {
      return { done: { kind: 'rendered', primaryRadiance: this._texture(), samplesAccumulated: this._samples, isConverged: true } };
    }
    return {
      target,
      frame: {
        invViewProj: inv,
        width,
        height,
        frameIndex: input.frameIndex >>> 0,
        frameSeed: input.frameSeed >>> 0,
        maxBounces: q.bounces == null ? 0 : q.bounces,
        samplesTarget: target,
      },
    };
  }

  _frameDone(target) {
    this._samples += 1;
    const out = {
      kind: 'rendered',
      primaryRadiance: this._texture(),
      normalDepth: this._texture('normalDepth'),
      albedo: this._texture('albedo'),
      variance: this._texture('variance'),
      samplesAccumulated: this._samples,
      isConverged: target > 0 && this._samples >= target,
    };
    this._emit('frame', out);
    this._emit('progress', { samplesAccumulated: this._samples, samplesTarget: target });
    return out;
  }

  // BackendTexture branded 'hip': one of the engine's images (the
  // accumulation by default; read with captureFrame / captureAux).
  _texture(image) {
    return Object.freeze({ backend: 'hip', engine: this, image: image || 'primaryRadiance', width: this._size[0],
      height: this._size[1], format: 'rgba32float' });
  }

  reset() {
    this._live();
    this._n.reset(this._h);
    this._samples = 0;
  }

  pause() {
    if (this.state === 'ready') this.state = 'paused';
  }

  resume() {
    if (this.state === 'paused') this.state = 'ready';
  }

  dispose() {
    if (this.state === 'disposed') return;
    this._n.destroy(this._h);
    this._h = null;
    this.state = 'disposed';
    for (const k of Object.keys(this._listeners)) this._listeners[k].clear();
  }

  _on(kind, cb) {
    if (typeof cb !== 'function') throw new TypeError('listener must be a function');
    this._listeners[kind].add(cb);
    return () => this._listeners[kind].delete(cb);
  }

  onFrame(cb) { return this._on('frame', cb); }
  onProgress(cb) { return this._on('progress', cb); }
  onError(cb) { return this._on('error', cb); }
  onWarning(cb) { return this._on('warning', cb); }

  // captureFrame (core/src/engine/index.ts:466): row-major, top-left origin,
  // float32 RGBA.  colorSpace 'linear' (default): the running mean;
  // 'output': the present pass of the last frame's quality settings
  // (tonemap, exposure, outputColorSpace; present.wgsl.ts), binary16 values.
  captureFrame(opts) {
    return new Promise((resolve, reject) => {
      try {
        this._live();
        const cs = opts && opts.colorSpace != null ? opts.colorSpace : 'linear';
        if (cs !== 'linear' && cs !== 'output') throw new RangeError(`captureFrame colorSpace '${cs}' is not supported`);
        if (!this._size || this._samples === 0) {
          resolve(null);
          return;
        }
        const [w, h] = this._size;
        if (cs === 'output') {
          const p = this._present;
          resolve({ width: w, height: h, rgba: this._n.captureOutput(this._h, w, h, p.tonemap, p.exposure, p.colorSpace) });
          return;
        }
        resolve({ width: w, height: h, rgba: this._n.capture(this._h, w, h) });
      } catch (err) {
        reject(err);
      }
    });
  }

  // The auxiliary images behind FrameRendered.normalDepth / albedo /
  // variance (accumulateFrame, kernel.wgsl.ts:920-960): float32 RGBA.
  captureAux(name) {
    this._live();
    const which = { normalDepth: 0, albedo: 1, variance: 2 }[name];
    if (which == null) throw new RangeError(`captureAux: unknown image '${name}'`);
    if (!this._size || this._samples === 0) return null;
    const [w, h] = this._size;
    return { width: w, height: h, rgba: this._n.captureAux(this._h, which, w, h) };
  }

  stats() {
    this._live();
    return this._n.stats(this._h);
  }
}

// Validated addon settings of the radiance cache from nrcEnabled/nrcConfig.
function nrcSettings(enabled, config) {
  const c = Object.assign({}, config || {});
  for (const k of Object.keys(c)) {
    if (k !== 'recordCap' && k !== 'learningRate' && k !== 'tableLearningRate' && k !== 'seed' &&
        k !== 'teacherVertices') {
      throw new TypeError(
        `pt-hip: nrcConfig.${k} is not supported (recordCap, learningRate, tableLearningRate, seed, teacherVertices)`);
    }
  }
  const cap = c.recordCap == null ? 4 * 65536 : c.recordCap;
  if (!(Number.isInteger(cap) && cap >= 4 && cap <= 4 * 4194304)) {
    throw new RangeError('pt-hip: nrcConfig.recordCap must be an integer in [4, 2^24]');
  }
  if (c.learningRate != null && !(typeof c.learningRate === 'number' && c.learningRate > 0)) {
    throw new RangeError('pt-hip: nrcConfig.learningRate must be a positive number');
  }
  const nrc = { enabled: enabled ? 1 : 0, trainPaths: Math.ceil(cap / 4) };
  if (c.tableLearningRate != null && !(typeof c.tableLearningRate === 'number' && c.tableLearningRate >= 0)) {
    throw new RangeError('pt-hip: nrcConfig.tableLearningRate must be a non-negative number');
  }
  if (c.learningRate != null) nrc.learningRate = c.learningRate;
  if (c.tableLearningRate != null) nrc.tableLearningRate = c.tableLearningRate;
  if (c.seed != null) nrc.seed = c.seed;
  if (c.teacherVertices != null) {
    if (!(Number.isInteger(c.teacherVertices) && c.teacherVertices >= 0 && c.teacherVertices <= 4)) {
      throw new RangeError('pt-hip: nrcConfig.teacherVertices must be an integer in [0, 4] (0: full-length training paths)');
    }
    nrc.teacherVertices = c.teacherVertices;
  }
  return nrc;
}

// The part of a cache configuration that is re-applied to every new scene
// (the seed and learning rates initialise the cache once).
function keptNrc(nrc) {
  const kept = { enabled: 1, trainPaths: nrc.trainPaths };
  if (nrc.teacherVertices != null) kept.teacherVertices = nrc.teacherVertices;
  return kept;
}

// EngineFactory (packages/core/src/engine/factory.ts): opts.device is the HIP
// device ordinal chosen by the host (default 0).
async function createPTEngine_HIP(opts) {
  const o = Object.assign({}, opts || {});
  const known = new Set(['device', 'maxBounces', 'maxSamplesPerPixel', 'sampling', 'sceneRadius', 'debug',
    'onWarning', 'denoiser', 'extensions', 'nrcEnabled', 'nrcConfig', 'lightTreeImportanceSampling',
    'bvhTraversal', 'primitiveUpdates']);
  for (const k of Object.keys(o)) {
    if (!known.has(k)) throw new TypeError(`pt-hip: unknown option '${k}'`);
  }
  if (o.denoiser && o.denoiser !== 'none') throw new RangeError("pt-hip has no denoiser (use denoiser: 'none')");
  // bvhTraversal (ptWebgpuValidation.ts:75,507): 'binary' (default) or
  // 'cwbvh-closest'.
  if (o.bvhTraversal !== undefined && o.bvhTraversal !== 'binary' && o.bvhTraversal !== 'cwbvh-closest') {
    throw new RangeError(`pt-hip: bvhTraversal is unsupported (got ${JSON.stringify(o.bvhTraversal)}); ` +
      'expected one of "binary", "cwbvh-closest"');
  }
  if (o.primitiveUpdates !== undefined && o.primitiveUpdates !== 'rebuild' && o.primitiveUpdates !== 'refit') {
    throw new RangeError(`pt-hip: primitiveUpdates must be "rebuild" or "refit" (got ${JSON.stringify(o.primitiveUpdates)})`);
  }
  if (o.nrcEnabled) nrcSettings(true, o.nrcConfig);  // validated before any device work
  const native = loadAddon();
  const handle = native.create(o.device == null ? 0 : o.device, {
    maxBounces: o.maxBounces || 0,
    maxSamplesPerPixel: Number.isFinite(o.maxSamplesPerPixel) ? o.maxSamplesPerPixel : 0,
    sampling: o.sampling === 'sobol' ? 1 : 0,
    lightSelection: o.lightTreeImportanceSampling === false ? 1 : 0,
    bvhTraversal: o.bvhTraversal === 'cwbvh-closest' ? 1 : 0,
  });
  const engine = new HipPathTracerEngine(native, handle, o);
  if (typeof o.onWarning === 'function') engine.onWarning(o.onWarning);
  if (o.nrcEnabled) engine.setNeuralCache(true, o.nrcConfig);
  return engine;
}

// BackendConstructor (createEngineInternals.ts:452): (opts, scene, aabb,
// needsTlas) -> engine carrying backendId.  The TLAS is always built here.
async function constructPathTracerHipForDispatch(opts, scene /* , aabb, needsTlas */) {
  const engine = await createPTEngine_HIP(opts);
  if (scene) engine.setScene(scene);
  return engine;
}

// Batched frames (no counterpart in the reference): one renderFrame of each
// engine, with every bounce's rays of all of them traced in one persistent
// launch (vt_render_batch_*).  Each engine's result is what its own
// renderFrame would give.  Engines with the radiance cache, the
// cwbvh-closest traversal or nothing to render this frame cannot join.
function renderFrames(engines, inputs) {
  if (!Array.isArray(engines) || !Array.isArray(inputs) || engines.length === 0 || engines.length !== inputs.length) {
    throw new TypeError('renderFrames(engines, inputs) takes equally long, non-empty arrays');
  }
  const args = engines.map((eng, k) => {
    if (!(eng instanceof HipPathTracerEngine)) throw new TypeError(`engines[${k}] is not a pt-hip engine`);
    const a = eng._frameArgs(inputs[k]);
    if (a.done) throw new RangeError(`engines[${k}] has nothing to render this frame`);
    return a;
  });
  try {
    engines[0]._n.renderFrames(engines.map((e) => e._h), args.map((a) => a.frame));
  } catch (err) {
    for (const eng of engines) eng._fail(err);
    throw err;
  }
  return engines.map((eng, k) => eng._frameDone(args[k].target));
}

module.exports = {
  createPTEngine_HIP,
  constructPathTracerHipForDispatch,
  multiplyMat4,
  sceneBlob,
  loadAddon,
  renderFrames,
};
