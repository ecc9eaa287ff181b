'use strict';
// Contract checks of the pt-hip JavaScript backend that need no GPU: the
// addon loads and exports its functions, scene and option validation throw
// the reference's error classes before any device work, and multiplyMat4
// matches the binary64-sum-then-binary32 rule.  With VITRUM_HIP_GPU=1 it also
// renders a small scene on device 0 and checks the capture contract.
const assert = require('assert');
const pt = require('./index.js');

const addon = pt.loadAddon();
for (const k of ['create', 'destroy', 'setScene', 'renderFrame', 'capture', 'reset', 'stats',
  'traceClosest', 'traceAny', 'invertMat4', 'nrcConfig', 'version']) {
  assert.strictEqual(typeof addon[k], 'function', `addon.${k}`);
}
assert.ok(/gfx950/.test(addon.version()));

const I = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1];
const T = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 2, 3, 4, 1];
const m = pt.multiplyMat4(T, T);
assert.deepStrictEqual(Array.from(m.slice(12)), [4, 6, 8, 1]);
const inv = addon.invertMat4(new Float32Array(T));
assert.deepStrictEqual(Array.from(inv.slice(12)), [-2, -3, -4, 1]);
assert.strictEqual(addon.invertMat4(new Float32Array(16)), null);
assert.throws(() => addon.invertMat4(new Float32Array(3)), TypeError);

// Scene conversion: unsupported kinds and bad arrays are rejected up front.
const quad = {
  kind: 'mesh', id: 'q',
  positions: new Float32Array([0, 0, 0, 1, 0, 0, 1, 1, 0]),
  normals: new Float32Array([0, 0, 1, 0, 0, 1, 0, 0, 1]),
  indices: new Uint16Array([0, 1, 2]),
  material: { baseColor: [0.5, 0.5, 0.5], roughness: 1, metallic: 0 },
};
const blob = pt.sceneBlob({ primitives: [quad], emitters: [], environment: { kind: 'none' } });
assert.ok(blob.primitives[0].indices instanceof Uint32Array);
assert.strictEqual(blob.environment.kind, 0);
assert.throws(() => pt.sceneBlob({ primitives: [{ kind: 'sphere' }] }), RangeError);
assert.throws(() => pt.sceneBlob({ primitives: [Object.assign({}, quad, { transform: [1, 2] })] }), TypeError);
// Scene.emitters: analytic kinds become emitter records, mesh-area entries
// light through their mesh's emissive material, unknown kinds are rejected.
const lit = pt.sceneBlob({
  primitives: [quad],
  emitters: [
    { kind: 'point', id: 'p', color: [1, 1, 1], intensity: 5, position: [0, 2, 0] },
    { kind: 'rect-area', id: 'r', color: [1, 1, 1], intensity: 2, position: [0, 1, 0], uAxis: [1, 0, 0], vAxis: [0, 0, 1] },
    { kind: 'mesh-area', id: 'm', meshId: 'q', color: [1, 1, 1], intensity: 1 },
    { kind: 'spot', id: 's', color: [1, 1, 1], intensity: 1, position: [0, 1, 0], direction: [0, -1, 0], angle: 0.5, castShadow: false },
  ],
  environment: { kind: 'none' },
});
assert.strictEqual(lit.emitters.length, 3);
assert.deepStrictEqual(lit.emitters.map((e) => e.kind), [1, 3, 2]);
assert.strictEqual(lit.emitters[0].decay, 2);
assert.strictEqual(lit.emitters[2].castShadow, false);
assert.throws(() => pt.sceneBlob({ primitives: [], emitters: [{ kind: 'laser' }] }), RangeError);
assert.throws(() => pt.sceneBlob({ primitives: [], emitters: [{ kind: 'point', position: [0, NaN, 0] }] }), TypeError);
const sky = pt.sceneBlob({ primitives: [], environment: { kind: 'hdri', hdri: { constant: [1, 0.5, 0.25] }, intensity: 2 } });
assert.deepStrictEqual(sky.environment, { kind: 1, radiance: [2, 1, 0.5] });
// HDRI payloads decode to linear RGB binary32 texels (environmentMapPixels.ts).
const rgba = pt.sceneBlob({ primitives: [], environment: { kind: 'hdri', hdri: { width: 2, height: 1, data: new Float32Array([1, 2, 3, 9, 0, 0.5, 0, 1]) }, intensity: 2, rotationY: 0.5 } });
assert.strictEqual(rgba.environment.kind, 2);
assert.deepStrictEqual(Array.from(rgba.environment.texels), [1, 2, 3, 0, 0.5, 0]);
assert.strictEqual(rgba.environment.intensity, 2);
assert.strictEqual(rgba.environment.rotationY, 0.5);
const u8 = pt.sceneBlob({ primitives: [], environment: { kind: 'hdri', hdri: { width: 1, height: 1, data: new Uint8Array([255, 0, 51]), colorSpace: 'srgb' } } });
assert.deepStrictEqual(Array.from(u8.environment.texels), [1, 0, Math.fround(Math.pow((0.2 + 0.055) / 1.055, 2.4))]);
const grey = pt.sceneBlob({ primitives: [], environment: { kind: 'hdri', hdri: { image: { width: 2, height: 1, data: new Uint16Array([65535, 0]) }, __vitrum_hint__: { channels: 1 } } } });
assert.deepStrictEqual(Array.from(grey.environment.texels), [1, 1, 1, 0, 0, 0]);
const half = pt.sceneBlob({ primitives: [], environment: { kind: 'hdri', hdri: { width: 1, height: 1, data: new Uint16Array([0x3c00, 0x4000, 0x3800]), dataType: 'float16' } } });
assert.deepStrictEqual(Array.from(half.environment.texels), [1, 2, 0.5]);
assert.throws(() => pt.sceneBlob({ primitives: [], environment: { kind: 'hdri', hdri: { width: 1, height: 1, data: new Float32Array([1, -1, 0]) } } }), RangeError);
assert.throws(() => pt.sceneBlob({ primitives: [], environment: { kind: 'hdri', hdri: { width: 1, height: 1, data: [1, 1, 1] }, rotationY: Infinity } }), RangeError);
const skyBlob = pt.sceneBlob({ primitives: [], environment: { kind: 'procedural-sky', sunDirection: [0.3, 0.8, 0.1], turbidity: 3 } });
assert.deepStrictEqual(skyBlob.environment, { kind: 3, radiance: [0, 0, 0], intensity: 1, turbidity: 3, sunDirection: [0.3, 0.8, 0.1] });
assert.throws(() => pt.sceneBlob({ primitives: [], environment: { kind: 'procedural-sky', rayleigh: NaN } }), RangeError);
assert.throws(() => pt.sceneBlob({ primitives: [], environment: { kind: 'hdri', hdri: { constant: [1, 1, 1] }, intensity: 1e-50 } }), RangeError);
assert.throws(() => pt.sceneBlob({ primitives: [], environment: { kind: 'vacuum' } }), RangeError);

(async () => {
  await assert.rejects(pt.createPTEngine_HIP({ bogus: 1 }), TypeError);
  await assert.rejects(pt.createPTEngine_HIP({ denoiser: 'oidn-final' }), RangeError);
  await assert.rejects(pt.createPTEngine_HIP({ nrcEnabled: true, nrcConfig: { levels: 16 } }), TypeError);
  await assert.rejects(pt.createPTEngine_HIP({ nrcEnabled: true, nrcConfig: { recordCap: 2 } }), RangeError);
  await assert.rejects(pt.createPTEngine_HIP({ nrcEnabled: true, nrcConfig: { teacherVertices: 5 } }), RangeError);
  await assert.rejects(pt.createPTEngine_HIP({ bvhTraversal: 'cwbvh' }), RangeError);
  if (process.env.VITRUM_HIP_GPU !== '1') {
    await assert.rejects(pt.createPTEngine_HIP({}), (e) => e.code === 'VT_ERR_NO_DEVICE' || e.code === 'VT_ERR_HIP');
    console.log('pt-hip contract ok (no device)');
    return;
  }
  const eng = await pt.createPTEngine_HIP({ device: 0 });
  assert.strictEqual(eng.state, 'ready');
  const frames = [];
  const off = eng.onFrame((f) => frames.push(f));
  eng.setScene({ primitives: [quad], emitters: [], environment: { kind: 'hdri', hdri: { constant: [1, 1, 1] } } });
  const view = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -0.5, -0.5, -3, 1];
  const f = 1 / Math.tan(Math.PI / 12), near = 0.01, far = 100;
  const proj = [f, 0, 0, 0, 0, f, 0, 0, 0, 0, (far + near) / (near - far), -1, 0, 0, (2 * far * near) / (near - far), 0];
  for (let i = 0; i < 4; i += 1) {
    const out = eng.renderFrame({ viewMatrix: view, projMatrix: proj, viewport: { width: 16, height: 16, devicePixelRatio: 1 },
      frameIndex: i, frameSeed: 0x9e3779b9, quality: { samplesTarget: 3, bounces: 4 } });
    assert.strictEqual(out.kind, 'rendered');
    assert.strictEqual(out.samplesAccumulated, Math.min(i + 1, 3));
    assert.strictEqual(out.isConverged, i >= 2);
  }
  off();
  assert.strictEqual(frames.length, 3);
  const cap = await eng.captureFrame();
  assert.strictEqual(cap.width, 16);
  assert.strictEqual(cap.rgba.length, 16 * 16 * 4);
  const center = cap.rgba[(8 * 16 + 8) * 4];
  assert.ok(center > 0 && center < 1.5, `center radiance ${center}`);
  // colorSpace 'output': the present pass (aces + sRGB by default), in [0, 1]
  const outCap = await eng.captureFrame({ colorSpace: 'output' });
  assert.ok(outCap.rgba.every((x) => x >= 0 && x <= 1), 'output capture in [0, 1]');
  assert.ok(outCap.rgba[(8 * 16 + 8) * 4] > 0);
  const nd = eng.captureAux('normalDepth');
  assert.ok(Math.abs(nd.rgba[(8 * 16 + 8) * 4 + 2] - 1) < 1e-6, 'quad normal +z');
  assert.ok(nd.rgba[(8 * 16 + 8) * 4 + 3] > 2.5, 'hit distance ~3');
  // an analytic point light over the same quad
  const pl = await pt.createPTEngine_HIP({ device: 0 });
  pl.setScene({ primitives: [quad], environment: { kind: 'none' },
    emitters: [{ kind: 'point', id: 'p', color: [1, 1, 1], intensity: 4, position: [0.6, 0.3, 1] }] });
  pl.renderFrame({ viewMatrix: view, projMatrix: proj, viewport: { width: 16, height: 16, devicePixelRatio: 1 },
    frameIndex: 0, frameSeed: 7, quality: { bounces: 2 } });
  const pc = await pl.captureFrame();
  assert.ok(pc.rgba[(8 * 16 + 8) * 4] > 0, 'point light lights the quad');
  pl.dispose();
  // an HDRI map (rotated, uint8 sRGB payload) and a procedural sky light the quad
  const envs = [{ kind: 'hdri', hdri: { width: 4, height: 2, data: new Uint8Array(4 * 2 * 3).fill(200), colorSpace: 'srgb' }, rotationY: 1.2, intensity: 2 },
    { kind: 'procedural-sky', sunDirection: [0.2, 0.9, 0.4] }];
  for (const environment of envs) {
    const he = await pt.createPTEngine_HIP({ device: 0 });
    he.setScene({ primitives: [quad], emitters: [], environment });
    for (let i = 0; i < 2; i += 1) {
      he.renderFrame({ viewMatrix: view, projMatrix: proj, viewport: { width: 16, height: 16, devicePixelRatio: 1 },
        frameIndex: i, frameSeed: 7, quality: { bounces: 3 } });
    }
    const hc = await he.captureFrame();
    assert.ok(hc.rgba.every((x) => Number.isFinite(x) && x >= 0), `${environment.kind} output finite`);
    assert.ok(hc.rgba[(8 * 16 + 8) * 4] > 0, `${environment.kind} lights the quad`);
    he.dispose();
  }
  // batched frames: two engines' frames in one call equal their own renderFrame
  {
    const scenes = [{ primitives: [quad], emitters: [], environment: { kind: 'hdri', hdri: { constant: [1, 1, 1] } } },
      { primitives: [quad], environment: { kind: 'procedural-sky', sunDirection: [0.3, 0.8, 0.2] },
        emitters: [{ kind: 'point', color: [1, 1, 1], intensity: 2, position: [0.4, 0.4, 1] }] }];
    const solo = [], both = [];
    for (let k = 0; k < 2; k += 1) {
      solo.push(await pt.createPTEngine_HIP({ device: 0 }));
      both.push(await pt.createPTEngine_HIP({ device: 0 }));
      solo[k].setScene(scenes[k]);
      both[k].setScene(scenes[k]);
    }
    const inputs = [0, 1].map((k) => ({ viewMatrix: view, projMatrix: proj, viewport: { width: 16 + 8 * k, height: 16, devicePixelRatio: 1 },
      frameIndex: 0, frameSeed: 5 + k, quality: { bounces: 3 + k } }));
    for (let k = 0; k < 2; k += 1) solo[k].renderFrame(inputs[k]);
    const outs = pt.renderFrames(both, inputs);
    assert.strictEqual(outs.length, 2);
    for (let k = 0; k < 2; k += 1) {
      assert.strictEqual(outs[k].samplesAccumulated, 1);
      const a = await solo[k].captureFrame(), b = await both[k].captureFrame();
      assert.deepStrictEqual(Array.from(a.rgba), Array.from(b.rgba), `batched frame ${k} equals its solo frame`);
      solo[k].dispose();
      both[k].dispose();
    }
    assert.throws(() => pt.renderFrames([], []), TypeError);
  }
  // Scene mutations: each equals setScene of the patched scene; the refit
  // path of a one-triangle primitive builds the same trees as a rebuild.
  {
    const inp = { viewMatrix: view, projMatrix: proj, viewport: { width: 16, height: 16, devicePixelRatio: 1 },
      frameIndex: 0, frameSeed: 7, quality: { bounces: 2 } };
    const point = { kind: 'point', id: 'p', color: [1, 1, 1], intensity: 4, position: [0.6, 0.3, 1] };
    const base = { primitives: [quad], emitters: [point], environment: { kind: 'none' } };
    const moved = new Float32Array([0, 0, 0.2, 1, 0, 0.2, 1, 1, 0.2]);
    const image = async (e) => { e.renderFrame(inp); return Array.from((await e.captureFrame()).rgba); };
    const a = await pt.createPTEngine_HIP({ device: 0 });
    const b = await pt.createPTEngine_HIP({ device: 0 });
    const r = await pt.createPTEngine_HIP({ device: 0, primitiveUpdates: 'refit' });
    assert.ok(a.capabilities.supportsIncrementalScene && a.capabilities.supportsAddRemovePrimitive);
    a.setScene(base);
    r.setScene(base);
    const before = await image(a);
    a.updatePrimitive('q', { positions: moved });
    r.updatePrimitive('q', { positions: moved });
    b.setScene({ primitives: [Object.assign({}, quad, { positions: moved })], emitters: [point], environment: { kind: 'none' } });
    const want = await image(b);
    assert.deepStrictEqual(await image(a), want, 'updatePrimitive equals setScene of the patched scene');
    assert.deepStrictEqual(await image(r), want, 'the refit path equals it for a one-triangle primitive');
    assert.notDeepStrictEqual(want, before);
    assert.strictEqual(a.getScene().primitives[0].positions, moved);
    assert.throws(() => a.updatePrimitive('nope', { positions: moved }), /not found/);
    assert.throws(() => a.updatePrimitive('q', { kind: 'instanced-mesh' }), TypeError);
    assert.throws(() => a.updatePrimitive('q', { bogus: 1 }), TypeError);
    a.updatePrimitive('q', { material: { roughness: 0.3 } });
    assert.strictEqual(a.getScene().primitives[0].material.baseColor, quad.material.baseColor, 'material patch merges');
    const q2 = Object.assign({}, quad, { id: 'q2', positions: new Float32Array([0, 0, -0.5, 1, 0, -0.5, 1, 1, -0.5]) });
    a.addPrimitive(q2);
    assert.strictEqual(a.getScene().primitives.length, 2);
    assert.throws(() => a.addPrimitive(q2), /already exists/);
    a.removePrimitive('q2');
    assert.strictEqual(a.getScene().primitives.length, 1);
    assert.throws(() => a.removePrimitive('q2'), /not found/);
    a.updatePrimitive('q', { material: quad.material });
    a.updateEmitter('p', { intensity: 8 });
    b.setScene({ primitives: [Object.assign({}, quad, { positions: moved })], emitters: [Object.assign({}, point, { intensity: 8 })],
      environment: { kind: 'none' } });
    assert.deepStrictEqual(await image(a), await image(b), 'updateEmitter equals setScene of the patched scene');
    assert.throws(() => a.updateEmitter('p', { kind: 'spot' }), /cannot change/);
    a.updateEnvironment({ kind: 'hdri', hdri: { constant: [0.5, 0.5, 0.5] } });
    assert.strictEqual(a.getScene().environment.kind, 'hdri');
    for (const e of [a, b, r]) e.dispose();
    await assert.rejects(pt.createPTEngine_HIP({ device: 0, primitiveUpdates: 'sometimes' }), RangeError);
  }
  // traceClosest through the addon: a ray straight at the triangle.
  const hits = new Float32Array(addon.traceClosest(eng._h, new Float32Array([0.6, 0.3, 1, 0, 0, 0, -1, 1e30])));
  assert.ok(Math.abs(hits[2] - 1) < 1e-6, `t = ${hits[2]}`);
  // The radiance cache (nrcEnabled): trains online, finite non-negative output.
  const nrc = await pt.createPTEngine_HIP({ device: 0, nrcEnabled: true, nrcConfig: { recordCap: 4096, seed: 7, teacherVertices: 2 } });
  nrc.setScene({ primitives: [quad], emitters: [], environment: { kind: 'hdri', hdri: { constant: [1, 1, 1] } } });
  for (let i = 0; i < 4; i += 1) {
    nrc.renderFrame({ viewMatrix: view, projMatrix: proj, viewport: { width: 16, height: 16, devicePixelRatio: 1 },
      frameIndex: i, frameSeed: 7, quality: { bounces: 4 } });
  }
  const nc = await nrc.captureFrame();
  // uniform light selection (lightTreeImportanceSampling: false) renders too
  const uni = await pt.createPTEngine_HIP({ device: 0, lightTreeImportanceSampling: false });
  uni.setScene({ primitives: [quad], emitters: [], environment: { kind: 'hdri', hdri: { constant: [1, 1, 1] } } });
  uni.renderFrame({ viewMatrix: view, projMatrix: proj, viewport: { width: 16, height: 16, devicePixelRatio: 1 },
    frameIndex: 0, frameSeed: 7, quality: { bounces: 4 } });
  const uc = await uni.captureFrame();
  assert.ok(uc.rgba.every((x) => Number.isFinite(x) && x >= 0), 'uniform-selection output finite');
  uni.dispose();
  // the cwbvh-closest traversal renders the same quad
  const cw = await pt.createPTEngine_HIP({ device: 0, bvhTraversal: 'cwbvh-closest' });
  cw.setScene({ primitives: [quad], emitters: [], environment: { kind: 'hdri', hdri: { constant: [1, 1, 1] } } });
  cw.renderFrame({ viewMatrix: view, projMatrix: proj, viewport: { width: 16, height: 16, devicePixelRatio: 1 },
    frameIndex: 0, frameSeed: 7, quality: { bounces: 4 } });
  const cc = await cw.captureFrame();
  assert.ok(cc.rgba.every((x) => Number.isFinite(x) && x >= 0), 'cwbvh-closest output finite');
  cw.dispose();
  assert.ok(nc.rgba.every((x) => Number.isFinite(x) && x >= 0), 'cache output finite and non-negative');
  nrc.setNeuralCache(false);
  nrc.dispose();
  eng.dispose();
  assert.strictEqual(eng.state, 'disposed');
  assert.throws(() => eng.setScene({ primitives: [] }), Error);
  console.log('pt-hip contract ok (device 0)');
})().catch((e) => {
  console.error(e);
  process.exit(1);
});
