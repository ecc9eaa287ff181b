"""Engine contract on the device: a failing vt_set_scene leaves the previous
scene bound (its traversal and its shading data), exceptions inside the
library come back as status codes, and a traversal error is reported once by
vt_capture."""
import numpy as np
import pytest

from vitrum_amd import native
from vitrum_amd import workloads as W

from render_cases import SEED

pytestmark = pytest.mark.gpu


def _frame(eng, sc, res=32):
    eng.reset()
    ivp = W.inv_view_proj(sc["camera"], res, res)
    for f in range(2):
        eng.render_frame(ivp, res, res, f, SEED, 4)
    return eng.capture(res, res)


def test_failed_set_scene_keeps_the_previous_scene(engine):
    cornell = W.cornell_scene()
    spheres = W.spheres_scene(instances=16, lights=4)
    rays, _ = W.primary_rays(cornell["camera"], 48, 48, scene_radius=cornell["scene_radius"])
    engine.set_scene(cornell)
    hits0 = engine.trace_closest(rays)
    info0 = engine.render_info()
    img0 = _frame(engine, cornell)
    bad_light = {**spheres, "primitives": [dict(p) for p in spheres["primitives"]]}
    bad_light["primitives"][-1] = dict(bad_light["primitives"][-1],
                                       material={"baseColor": [0, 0, 0], "emissive": [1e30, 1, 1],
                                                 "emissiveIntensity": 1e30})
    attempts = [(lambda: engine.inject_failure(1), -7, spheres),   # packing throws std::bad_alloc
                (lambda: engine.inject_failure(2), -7, spheres),   # shading-data upload fails
                (lambda: engine.inject_failure(4), -3, spheres),   # preparation throws
                (lambda: None, -2, bad_light)]                     # Le overflows f32
    for arm, status, scene in attempts:
        arm()
        with pytest.raises(native.VtError) as ei:
            engine.set_scene(scene)
        assert ei.value.status == status, str(ei.value)
        assert engine.render_info() == info0
        hits = engine.trace_closest(rays)
        for k in ("prim", "inst", "t", "u", "v"):
            assert np.array_equal(hits[k], hits0[k])
        assert np.array_equal(_frame(engine, cornell), img0)
    engine.set_scene(spheres)  # the hooks fired once each; a good scene binds
    assert engine.render_info()["emitter_count"] != info0["emitter_count"]
