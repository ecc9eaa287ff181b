"""Closest-hit traversal throughput on one MI355X.

A step is one pass of the closest-hit kernel over one batch of synthetic rays
resident in HBM: 1024x1024 pinhole primary rays plus one cosine-hemisphere
bounce ray per primary hit, on the ~1M-triangle scene (256 displaced
octa-spheres of 4,096 triangles with random rigid transforms in a 100^3 box,
plus 8 emissive light spheres).  The kernel time of every launch is taken
from HIP events recorded on the engine stream that runs the kernel.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "Mrays/s closest-hit; time-to-1e-2 RMSE, plain vs NRC, on 1 MI355X"
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(scene, rays, seconds):
    """The oracle (a CPU port of the reference traversal) on a bounded sample."""
    from oracle import oracle as O
    prims = [dict(p, material_id=i) for i, p in enumerate(scene["primitives"])]
    ps = O.PackedScene(O.pack_scene(prims))
    threads = min(16, os.cpu_count() or 1)
    probe = rays[: 20000]
    t0 = time.perf_counter()
    O.trace_closest(ps, probe, threads=threads)
    rate = probe.size / max(time.perf_counter() - t0, 1e-6)
    target = max(probe.size, int(rate * seconds))
    n, passes = 0, 0
    t0 = time.perf_counter()
    while n < target:
        chunk = rays[: min(rays.size, target - n)]
        O.trace_closest(ps, chunk, threads=threads)
        n += chunk.size
        passes += 1
    dt = time.perf_counter() - t0
    return {"value": round(n / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{n} rays = {passes} pass(es) over the same batch (primary rows first), "
                      f"{dt:.1f} s on {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this size first (HBM counter calibration)")
    ap.add_argument("--refill-min", type=int, default=0, help="traversal wave refill threshold (0: default)")
    ap.add_argument("--steps-per-refill", type=int, default=0, help="traversal steps between refills (0: default)")
    ap.add_argument("--near-first", type=int, default=-1, help="1: nearer child first, 0: left first (-1: default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path).", file=sys.stderr)
        sys.exit(2)

    from vitrum_amd import native, workloads as W

    eng = native.Engine(0)
    if args.refill_min or args.steps_per_refill or args.near_first >= 0:
        eng.set_trace_tuning(args.refill_min or 16, args.steps_per_refill or 16,
                             1 if args.near_first < 0 else args.near_first)
    if args.calibrate_copy_mib:
        nb = args.calibrate_copy_mib << 20
        a, b = eng.alloc(nb), eng.alloc(nb)
        eng.d2d(b, a, nb)
        eng.free(a)
        eng.free(b)
        log(f"calibration copy: {nb} bytes read + {nb} bytes written by copy16_kernel")
    t0 = time.perf_counter()
    sc = W.spheres_scene()
    log(f"scene generated in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    eng.set_scene(sc)
    build_s = time.perf_counter() - t0
    log(f"scene packed/built/uploaded in {build_s:.2f} s")
    prim, st = W.primary_rays(sc["camera"], args.width, args.height, scene_radius=sc["scene_radius"])
    ph = eng.trace_closest(prim)
    pack = native.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
    bounce = W.bounce_rays(pack, prim, ph, st, sc["scene_radius"])
    rays = np.concatenate([prim, bounce])
    n = rays.size
    log(f"{prim.size} primary + {bounce.size} bounce rays")

    # Work counters (untimed) -> algorithmic bytes per launch.
    _, counts = eng.trace_closest_counts(rays)
    c = counts.astype(np.float64).sum(axis=0)
    box_tests, tri_tests, tlas_tests, inst_entries = c
    # Reference layout: 32 B per binary node tested, 64 B per triangle tested
    # (index vec4u + 3 vertex vec4f; shared-bvh/src/strides.ts:24-71).
    ref_bytes = 32.0 * (box_tests + tlas_tests) + 64.0 * tri_tests
    # This kernel's layout: one 64 B Node64 per expansion (two boxes), 48 B per
    # triangle, 80 B per instance entry (W2L + metadata), 32 B ray + 20 B hit.
    own_bytes = 32.0 * (box_tests + tlas_tests) + 48.0 * tri_tests + 80.0 * inst_entries + 52.0 * n

    d_rays = eng.alloc(rays.nbytes)
    d_hits = eng.alloc(n * W.HIT_DTYPE.itemsize)
    eng.h2d(d_rays, rays)
    for _ in range(args.warmup):
        eng.trace_closest_device(d_rays, n, d_hits)
    eng.synchronize()
    kernel_ms = []
    eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.trace_closest_device(d_rays, n, d_hits)
        kernel_ms.append(eng.stats()["last_trace_ms"])
    eng.synchronize()
    wall = time.perf_counter() - t0
    out = np.zeros(n, W.HIT_DTYPE)
    eng.d2h(out, d_hits)
    eng.free(d_rays)
    eng.free(d_hits)
    avg_ms = float(np.mean(kernel_ms))
    value = n * args.steps / wall / 1e6
    achieved = ref_bytes / (avg_ms * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "spheres-1M closest-hit: 256 displaced octa-spheres x 4096 tris + 8 light "
                        "spheres (1,052,672 tris, 264 TLAS instances), 1024x1024 primary + 1 "
                        "cosine bounce per hit",
            "rays_per_step": int(n),
            "primary_rays": int(prim.size),
            "bounce_rays": int(bounce.size),
            "kernel_ms_mean": round(avg_ms, 4),
            "kernel_mrays_per_s": round(n / (avg_ms * 1e-3) / 1e6, 2),
            "scene_build_s": round(build_s, 3),
            "avg_box_tests_per_ray": round(box_tests / n, 3),
            "avg_tri_tests_per_ray": round(tri_tests / n, 3),
            "avg_tlas_tests_per_ray": round(tlas_tests / n, 3),
            "avg_instances_per_ray": round(inst_entries / n, 3),
            "own_layout_gbps": round(own_bytes / (avg_ms * 1e-3) / 1e9, 1),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": None,
            "bytes_per_launch": int(ref_bytes),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(sc, rays, args.cpu_seconds)
    eng.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
