"""vt_trace_closest_device_noalloc (the allocation-free closest hit) against
vt_trace_closest_device: bit-identical hit records on two calls in a row on
the 1M-triangle workload's scene class, the second call with fewer rays (the
shared scratch keeps the first call's counters and spill contents)."""
import numpy as np
import pytest

from vitrum_amd import native
from vitrum_amd import workloads as W

pytestmark = pytest.mark.gpu


def _rays():
    sc = W.spheres_scene(instances=64, lights=4)
    prim, st = W.primary_rays(sc["camera"], 384, 384, scene_radius=sc["scene_radius"])
    return sc, prim, st


def test_noalloc_closest_hit_equals_device_entry_point():
    eng = native.Engine(0)
    try:
        sc, prim, st = _rays()
        eng.set_scene(sc)
        ph = eng.trace_closest(prim)
        pack = native.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
        rays = np.concatenate([prim, W.bounce_rays(pack, prim, ph, st, sc["scene_radius"])])
        # shuffled, so that the shorter second call also holds hits and misses
        rays = rays[np.random.default_rng(11).permutation(rays.size)]
        n = rays.size
        d_rays = eng.alloc(rays.nbytes)
        d_a = eng.alloc(n * W.HIT_DTYPE.itemsize)
        d_b = eng.alloc(n * W.HIT_DTYPE.itemsize)
        try:
            eng.h2d(d_rays, rays)
            for count in (n, 1000):
                print(f"noalloc vs device entry point: {count} rays", flush=True)
                a = np.zeros(count, W.HIT_DTYPE)
                b = np.full(count * W.HIT_DTYPE.itemsize, 0xAB, np.uint8).view(W.HIT_DTYPE)
                eng.h2d(d_b, b)  # poison the output so stale records cannot pass
                eng.trace_closest_device(d_rays, count, d_a)
                eng.trace_closest_device_noalloc(d_rays, count, d_b)
                eng.synchronize()
                eng.d2h(a, d_a)
                eng.d2h(b, d_b)
                for k in ("prim", "inst", "t", "u", "v"):
                    assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), (count, k)
                assert 0.05 < np.mean(a["inst"] != 0xFFFFFFFF) < 0.95
        finally:
            for d in (d_rays, d_a, d_b):
                eng.free(d)
    finally:
        eng.close()


def test_noalloc_rejects_bad_calls():
    eng = native.Engine(0)
    try:
        with pytest.raises(native.VtError) as ei:  # no scene yet
            eng.trace_closest_device_noalloc(0, 0, 0)
        assert ei.value.status == -6
        eng.set_scene(W.cornell_scene())
        with pytest.raises(native.VtError) as ei:
            eng.trace_closest_device_noalloc(0, 1 << 31, 0)
        assert ei.value.status in (-1, -2)
    finally:
        eng.close()
