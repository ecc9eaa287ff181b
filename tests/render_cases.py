"""Scenes, closed forms and oracle drivers shared by the integrator tests."""
import numpy as np

from oracle import oracle as O
from vitrum_amd import workloads as W

SEED = 0x9E3779B9


def materials_of(scene):
    return [p.get("material", {}) for p in scene["primitives"]]


def cpu_frames(scene, res, frames, bias, max_bounces=8, sampling=0, light_tree=True, first_frame=0):
    prims = [dict(p, material_id=i) for i, p in enumerate(scene["primitives"])]
    pack = O.pack_scene(prims)
    ivp = W.inv_view_proj(scene["camera"], res, res)
    sky = scene.get("sky", [0, 0, 0])
    out, sigs = [], []
    for f in range(first_frame, first_frame + frames):
        img, sig = O.render(pack, materials_of(scene), ivp, res, res, frame_seed=SEED, frame_index=f,
                            max_bounces=max_bounces, ray_bias=bias,
                            env_radiance=sky if any(sky) else None, sampling=sampling,
                            light_tree=light_tree)
        out.append(img)
        sigs.append(sig)
    return np.stack(out), np.stack(sigs)


def plane_under_sky(albedo, sky, roughness=1.0, metallic=0.0):
    """A 400x400 floor at y = 0 under a constant sky, seen from far above
    through a narrow field of view (view directions within 1.5 degrees of the
    normal)."""
    pos = np.array([-200, 0, -200, 200, 0, -200, 200, 0, 200, -200, 0, 200], np.float32)
    nrm = np.tile([0, 1, 0], 4).astype(np.float32)
    idx = np.array([0, 2, 1, 0, 3, 2], np.uint32)
    prim = {"kind": "mesh", "positions": pos, "normals": nrm, "indices": idx,
            "material": {"baseColor": [albedo] * 3, "roughness": roughness, "metallic": metallic}}
    cam = W.look_at_camera(eye=(0, 60, 0.01), target=(0, 0, 0), up=(0, 0, -1), vfov_deg=3)
    return {"primitives": [prim], "camera": cam, "scene_radius": 300.0, "sky": sky}


def directional_albedo_at_normal(albedo, roughness, metallic, n_theta=4000):
    """Quadrature of the hemispherical integral of f(n, wi) cos(theta) for the
    base BSDF (diffuse (1-F) kd/pi + GGX D G F / (4 cos cos) + Kulla-Conty
    multiscatter) with wo = n; the lobes are isotropic, so phi is trivial."""
    th = (np.arange(n_theta) + 0.5) / n_theta * (np.pi / 2)
    ct, st = np.cos(th), np.sin(th)
    ndl = ct
    h_ct = np.cos(th / 2)                      # n.h with wo = n
    vdh = h_ct
    f0 = 0.04 * (1 - metallic) + albedo * metallic
    F = f0 + (1 - f0) * np.clip(1 - vdh, 0, 1) ** 5
    alpha = max(roughness * roughness, 1e-3)
    d = alpha ** 2 / (np.pi * ((1 - h_ct ** 2) + h_ct ** 2 * alpha ** 2) ** 2)

    def g1(c):
        t2 = np.maximum(0, 1 - c * c) / np.maximum(c * c, 1e-10)
        return np.where(c <= 1e-6, 0.0, 2.0 / (1 + np.sqrt(1 + alpha * alpha * t2)))

    spec = d * g1(1.0) * g1(ndl) * F / np.maximum(4 * 1.0 * ndl, 1e-6)
    diff = (1 - F) * (1 - metallic) * albedo / np.pi
    E = np.array(O_LUT_E).reshape(8, 8)
    Eavg = np.array(O_LUT_EAVG)

    def lut(c, r):
        fr, fm = r * 7, np.clip(c, 0, 1) * 7
        r0, m0 = int(np.floor(fr)), np.floor(fm).astype(int)
        r1, m1 = min(r0 + 1, 7), np.minimum(m0 + 1, 7)
        tr, tm = fr - r0, fm - m0
        e0 = E[r0, m0] + (E[r0, m1] - E[r0, m0]) * tm
        e1 = E[r1, m0] + (E[r1, m1] - E[r1, m0]) * tm
        return np.clip(e0 + (e1 - e0) * tr, 0.02, 1)

    fr_ = roughness * 7
    r0 = int(np.floor(fr_))
    eavg = np.clip(Eavg[r0] + (Eavg[min(r0 + 1, 7)] - Eavg[r0]) * (fr_ - r0), 0.3, 1)
    favg = f0 + (1 - f0) / 21
    fms = favg * favg * eavg / max(1 - favg * (1 - eavg), 1e-4)
    ms = fms * (1 - lut(1.0, roughness)) * (1 - lut(ndl, roughness)) / max(np.pi * (1 - eavg), 1e-6)
    f = np.where(ndl > 1e-5, diff + spec + ms, 0.0)
    return float(np.sum(f * ct * st) * (np.pi / 2 / n_theta) * 2 * np.pi)


# Kulla-Conty tables of pt-webgpu/src/wgsl/pathTrace/material.wgsl.ts:2852-2866.
O_LUT_E = [0.1375, 0.5617, 0.7546, 0.8522, 0.9111, 0.9505, 0.9788, 1.0,
           0.2955, 0.515, 0.7091, 0.8192, 0.889, 0.937, 0.9721, 0.9988,
           0.5794, 0.5541, 0.6677, 0.7691, 0.8451, 0.9021, 0.9457, 0.98,
           0.7011, 0.6486, 0.6669, 0.7199, 0.7776, 0.8305, 0.8764, 0.9155,
           0.7335, 0.6901, 0.6696, 0.6756, 0.6972, 0.7262, 0.7578, 0.7893,
           0.7153, 0.6712, 0.6355, 0.6145, 0.6052, 0.6045, 0.6101, 0.6199,
           0.6669, 0.6137, 0.5657, 0.5286, 0.5, 0.478, 0.4611, 0.4483,
           0.6017, 0.537, 0.4773, 0.4296, 0.3905, 0.358, 0.3305, 0.3069]
O_LUT_EAVG = [0.9106, 0.8931, 0.8629, 0.8094, 0.725, 0.6147, 0.4931, 0.3766]
