"""Batched persistent launch (vt_batch_*): scenes of very different sizes and
ray counts in one launch, its HIP-graph replay and the autotuner.

Contract: every scene's hit records equal vt_trace_closest on that scene
alone, bit for bit, for the direct launch, for each graph replay (with other
eager work on the stream between replays) and after autotuning.
"""
import numpy as np
import pytest

from vitrum_amd import native, workloads as W

pytestmark = pytest.mark.gpu


def scene_rays(sc, res, bounce):
    prim, st = W.primary_rays(sc["camera"], res, res, scene_radius=sc["scene_radius"])
    if not bounce:
        return prim, None
    pack = native.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
    return prim, (pack, st)


@pytest.fixture(scope="module")
def batch_case(engine):
    specs = [(W.cornell_scene(2048), 48, True), (W.cornell_scene(32768), 96, True),
             (W.spheres_scene(instances=16, lights=2), 128, True), (W.cornell_scene(8192), 0, False),
             (W.cornell_scene(4096), 64, False)]
    scenes, rays, refs = [], [], []
    for sc, res, bounce in specs:
        engine.set_scene(sc)
        if res == 0:  # a scene with no rays in this batch
            r = np.zeros(0, W.RAY_DTYPE)
        else:
            prim, extra = scene_rays(sc, res, bounce)
            r = prim
            if extra is not None:
                ph = engine.trace_closest(prim)
                r = np.concatenate([prim, W.bounce_rays(extra[0], prim, ph, extra[1], sc["scene_radius"])])
        refs.append(engine.trace_closest(r))
        scenes.append(sc)
        rays.append(r)
    offsets = np.concatenate([[0], np.cumsum([r.size for r in rays])]).astype(np.uint32)
    return scenes, np.concatenate(rays), offsets, np.concatenate(refs)


def compare(hits, ref):
    for k in ("prim", "inst", "t", "u", "v"):
        a, b = hits[k].view(np.uint32), ref[k].view(np.uint32)
        assert np.array_equal(a, b), (k, int(np.sum(a != b)))


def test_batch_equals_per_scene_runs(engine, batch_case):
    scenes, rays, offsets, ref = batch_case
    b = native.Batch(engine)
    try:
        for sc in scenes:
            b.add_scene(sc)
        n = rays.size
        d_r, d_h = engine.alloc(rays.nbytes), engine.alloc(n * W.HIT_DTYPE.itemsize)
        try:
            engine.h2d(d_r, rays)
            ms = b.trace_closest_device(d_r, offsets, d_h)
            assert ms > 0
            out = np.zeros(n, W.HIT_DTYPE)
            engine.d2h(out, d_h)
            compare(out, ref)
            # graph capture and replays, with eager work on the stream between
            b.capture(d_r, offsets, d_h)
            for rep in range(3):
                engine.h2d(d_h, np.zeros(n, W.HIT_DTYPE))
                engine.trace_closest(rays[:1000])  # eager launch on the same stream
                assert b.replay() > 0
                out = np.zeros(n, W.HIT_DTYPE)
                engine.d2h(out, d_h)
                compare(out, ref)
            # autotuner: picks valid parameters; results unchanged
            t = b.autotune(d_r, offsets, d_h, repeats=2)
            assert t["ms"] > 0 and t["near_first"] in (0, 1) and t["blocks_per_cu"] >= 1
            engine.h2d(d_h, np.zeros(n, W.HIT_DTYPE))
            b.trace_closest_device(d_r, offsets, d_h)
            out = np.zeros(n, W.HIT_DTYPE)
            engine.d2h(out, d_h)
            compare(out, ref)
        finally:
            engine.free(d_r)
            engine.free(d_h)
    finally:
        b.close()


def test_batch_rejects_bad_offsets(engine, batch_case):
    scenes = batch_case[0]
    b = native.Batch(engine)
    try:
        b.add_scene(scenes[0])
        d = engine.alloc(64)
        try:
            with pytest.raises(native.VtError):
                b.trace_closest_device(d, [1, 2], d)
            with pytest.raises(native.VtError):
                b.replay()
        finally:
            engine.free(d)
    finally:
        b.close()
