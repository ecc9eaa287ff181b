"""Batched frames (vt_render_batch_*): several engines, each with its own
scene, camera, viewport, seed and bounce limit, render together -- every
bounce's rays of all of them traced in one persistent launch -- and each
member's image and path signatures are bit-identical to rendering the same
frames on that engine alone.  The same holds for a frame recorded once into
a HIP graph and replayed for later frame indices (every replay checked
against the solo result, with eager work and synchronisation in between),
and for the batched-launch workload: 64 Cornell-style scenes with
log-uniform triangle counts and mixed resolutions at 4 spp, depth 6.
"""
import time

import numpy as np
import pytest

from vitrum_amd import native
from vitrum_amd import workloads as W

from render_cases import floor_with_emitters, hdri_sun_map

pytestmark = pytest.mark.gpu

SEED = 0x9E3779B9


def small_members():
    floor = floor_with_emitters([{"kind": "point", "position": [0.3, 1.5, 0.2], "color": [1, 1, 1], "intensity": 3.0}],
                                fov=20.0)
    floor["hdri"] = {"texels": hdri_sun_map(), "intensity": 0.5, "rotation_y": 0.8}
    return [(W.cornell_scene(8192), 64, 48, 8, 11),
            (W.spheres_scene(instances=24, lights=4), 40, 40, 6, 12),
            (floor, 32, 24, 4, 13)]


def frame_of(sc, w, h, bounces, seed, index):
    return {"inv_view_proj": W.inv_view_proj(sc["camera"], w, h), "width": w, "height": h,
            "frame_index": index, "frame_seed": seed, "max_bounces": bounces}


def solo(engines, specs, frames):
    """Per member, the image and signatures after each of `frames` frames."""
    out = []
    for eng, (sc, w, h, b, seed) in zip(engines, specs):
        eng.reset()
        steps = []
        for f in range(frames):
            fr = frame_of(sc, w, h, b, seed, f)
            eng.render_frame(fr["inv_view_proj"], w, h, f, seed, b)
            steps.append((eng.capture(w, h), eng.capture_signatures(w, h)))
        out.append(steps)
    return out


@pytest.fixture(scope="module")
def members():
    specs = small_members()
    engines = [native.Engine(0) for _ in specs]
    for eng, (sc, *_rest) in zip(engines, specs):
        eng.set_scene(sc)
    yield engines, specs
    for eng in engines:
        eng.close()


def test_batched_frames_equal_solo_frames(members):
    engines, specs = members
    ref = solo(engines, specs, 2)
    b = native.RenderBatch(engines[0])
    try:
        for eng in engines:
            b.add(eng)
        for eng in engines:
            eng.reset()
        for f in range(2):
            b.frame([frame_of(sc, w, h, bo, seed, f) for sc, w, h, bo, seed in specs])
            for k, (eng, (sc, w, h, _, _)) in enumerate(zip(engines, specs)):
                img, sig = eng.capture(w, h), eng.capture_signatures(w, h)
                assert np.array_equal(img, ref[k][f][0]), (k, f)
                assert np.array_equal(sig, ref[k][f][1]), (k, f)
                assert img[..., :3].mean() > 0
    finally:
        b.close()


def test_graph_replayed_frames_equal_solo_frames(members):
    engines, specs = members
    frames = 3
    ref = solo(engines, specs, frames)
    b = native.RenderBatch(engines[1])
    try:
        for eng in engines:
            b.add(eng)
        for eng in engines:
            eng.reset()
        b.capture([frame_of(sc, w, h, bo, seed, 0) for sc, w, h, bo, seed in specs])
        for f in range(frames):
            b.replay(f)
            for k, (eng, (sc, w, h, _, _)) in enumerate(zip(engines, specs)):
                assert np.array_equal(eng.capture(w, h), ref[k][f][0]), (k, f)
                assert np.array_equal(eng.capture_signatures(w, h), ref[k][f][1]), (k, f)
            engines[0].synchronize()
        # a member that changes its viewport invalidates the recording
        sc, w, h, bo, seed = specs[2]
        engines[2].render_frame(W.inv_view_proj(sc["camera"], w + 8, h), w + 8, h, 0, seed, bo)
        with pytest.raises(native.VtError) as ei:
            b.replay(frames)
        assert ei.value.status == -6
    finally:
        b.close()


def test_invalid_members_are_rejected(members):
    engines, specs = members
    b = native.RenderBatch(engines[0])
    try:
        fresh = native.Engine(0)
        try:
            with pytest.raises(native.VtError):
                b.add(fresh)  # no scene
        finally:
            fresh.close()
        with pytest.raises(native.VtError):
            b.frame([])  # no members
        b.add(engines[0])
        with pytest.raises(native.VtError):
            b.add(engines[0])  # twice
        with pytest.raises(native.VtError):
            b.replay(0)  # nothing captured
    finally:
        b.close()


def test_batched_frames_64_scenes_4spp_depth6():
    """The batched-launch workload: 64 scenes, triangle counts log-uniform in
    [1k, 2M], resolutions from {256^2, 512^2, 1024^2}, 4 spp at depth 6; every
    member's accumulated image equals its solo rendering bit for bit."""
    work = W.batch_workload()
    engines = []
    try:
        for sc, res in work:
            eng = native.Engine(0, build_threads=16)
            eng.set_scene(sc)
            engines.append(eng)
        specs = [(sc, res, res, 6, SEED + k) for k, (sc, res) in enumerate(work)]
        ref = []
        for eng, (sc, w, h, bo, seed) in zip(engines, specs):
            eng.reset()
        engines[0].synchronize()
        t0 = time.perf_counter()
        for eng, (sc, w, h, bo, seed) in zip(engines, specs):
            for f in range(4):
                eng.render_frame(W.inv_view_proj(sc["camera"], w, h), w, h, f, seed, bo)
        for eng in engines:
            eng.synchronize()
        t_solo = time.perf_counter() - t0
        for eng, (sc, w, h, bo, seed) in zip(engines, specs):
            ref.append(eng.capture(w, h))
            assert ref[-1][..., :3].mean() > 0
        b = native.RenderBatch(engines[0])
        try:
            for eng in engines:
                b.add(eng)
                eng.reset()
            b.capture([frame_of(sc, w, h, bo, seed, 0) for sc, w, h, bo, seed in specs])
            engines[0].synchronize()
            t0 = time.perf_counter()
            for f in range(4):
                b.replay(f)
            engines[0].synchronize()
            t_batch = time.perf_counter() - t0
            for k, (eng, (sc, w, h, _, _)) in enumerate(zip(engines, specs)):
                assert np.array_equal(eng.capture(w, h), ref[k]), k
        finally:
            b.close()
        print(f"{len(work)} scenes, {sum(w * h for _, w, h, _, _ in specs)} pixels per frame; 4 frames: "
              f"solo engines {t_solo * 1e3:.1f} ms, batched graph replays {t_batch * 1e3:.1f} ms", flush=True)
    finally:
        for eng in engines:
            eng.close()
