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
import hashlib
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


def kernel_key(tuning):
    """Identifies the traversal kernel a traffic measurement belongs to: a
    hash of the kernel's source files and the launch tuning."""
    h = hashlib.sha256()
    for f in ("traverse.hip", "traverse.h"):
        with open(os.path.join(ROOT, "vitrum_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return f"src={h.hexdigest()[:16]} {tuning}"


def measured_traffic(key):
    """HBM bytes per launch of the closest-hit kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE in KiB per dispatch,
    each scaled by the reading of a plain copy of known size), or None when
    that measurement was made on other kernel sources or tuning."""
    path = os.path.join(ROOT, "profiles", "trace_closest_spheres1M.traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("kernel_key") != key:
        return None
    return int(round((t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024))


def _timed_passes(fn, rays, seconds):
    """Runs fn over leading chunks of `rays` for about `seconds`; returns
    (rays traced, passes, elapsed)."""
    probe = rays[: 20000]
    t0 = time.perf_counter()
    fn(probe)
    rate = probe.size / max(time.perf_counter() - t0, 1e-6)
    target = max(probe.size, int(rate * seconds))
    n, passes = 0, 0
    t0 = time.perf_counter()
    while n < target:
        chunk = rays[: min(rays.size, target - n)]
        fn(chunk)
        n += chunk.size
        passes += 1
    return n, passes, time.perf_counter() - t0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """The CPUs this process may run on (its affinity mask), at most 16: on a
    shared GPU host one GPU's process gets a 16-CPU share whatever the machine
    reports."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(scene, rays, seconds):
    """The reference's CPU closest hit -- the shared-bvh wide-BVH walk
    (intersectCompressedWideBvhFirstHit over tlasIntersect candidates, binary64),
    restated in C in the oracle -- on bounded samples of the same rays, on one
    host thread and on T threads; the oracle's f32 watertight traversal (the
    GPU kernel's own semantics) is timed the same way beside it."""
    from oracle import oracle as O
    prims = [dict(p, material_id=i) for i, p in enumerate(scene["primitives"])]
    ps = O.PackedScene(O.pack_scene(prims))
    threads = host_threads()
    t0 = time.perf_counter()
    cw, roots = O.scene_cwbvh(ps)
    collapse_s = time.perf_counter() - t0
    walk = lambda t: (lambda r: O.cwbvh_trace_closest(ps, cw, roots, r, threads=t))  # noqa: E731
    tight = lambda t: (lambda r: O.trace_closest(ps, r, threads=t))  # noqa: E731
    n, passes, dt = _timed_passes(walk(threads), rays, seconds * 0.3)
    n1, _, dt1 = _timed_passes(walk(1), rays, seconds * 0.2)
    n2, _, dt2 = _timed_passes(tight(threads), rays, seconds * 0.3)
    n3, _, dt3 = _timed_passes(tight(1), rays, seconds * 0.2)
    return {"value": round(n / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"shared-bvh CPU walk (binary64, 8-wide collapsed BVH, {collapse_s:.2f} s collapse): "
                      f"{n} rays = {passes} pass(es) over the leading rays of the batch (primary rows "
                      f"first), {dt:.1f} s on {threads} threads; 1 thread: {n1} rays in {dt1:.1f} s",
            "value_1_thread": round(n1 / dt1 / 1e6, 4),
            "watertight_f32_value": round(n2 / dt2 / 1e6, 4),
            "watertight_f32_value_1_thread": round(n3 / dt3 / 1e6, 4)}


# MLP FLOPs per sample: forward 2 (4*64*64 + 64*3); backward through
# W4..W1 2 (3*64 + 3*64*64); weight gradients 2 (4*64*64 + 3*64).
NRC_FWD_FLOPS = 2 * (4 * 64 * 64 + 64 * 3)
NRC_TRAIN_FLOPS = NRC_FWD_FLOPS + 2 * (3 * 64 + 3 * 64 * 64) + 2 * (4 * 64 * 64 + 3 * 64)
BF16_PEAK_TFLOPS = 2500.0


def blas_build(eng, native, triangles=1_000_000, reps=3):
    """Binned-SAH BLAS build of a synthetic triangle soup (standard-normal
    vertices, seeded): the device builder (vt_build_blas_device, best of
    `reps`, host buffers in and out) beside the host builder (vt_build_blas,
    one build); both produce the same nodes and permutation, checked here."""
    rng = np.random.default_rng(11)
    pos = np.concatenate([rng.standard_normal((3 * triangles, 3)).astype(np.float32),
                          np.zeros((3 * triangles, 1), np.float32)], 1)
    idx = np.concatenate([np.arange(3 * triangles, dtype=np.uint32).reshape(-1, 3),
                          np.zeros((triangles, 1), np.uint32)], 1)
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        dev = native.build_blas(pos, idx, engine=eng)
        best = min(best, time.perf_counter() - t0)
    t0 = time.perf_counter()
    host = native.build_blas(pos, idx)
    host_s = time.perf_counter() - t0
    same = bool(np.array_equal(dev["nodes"], host["nodes"]) and
                np.array_equal(dev["reordered_to_source"], host["reordered_to_source"]))
    return {"triangles": triangles, "device_ms": round(1e3 * best, 2), "host_ms": round(1e3 * host_s, 2),
            "speedup": round(host_s / best, 2), "identical": same, "nodes": int(dev["node_count"])}


def nrc_mlp(eng, reps=10):
    """Cache MLP kernels on their own: 1920x1080 inference queries and one
    training step of 65,536 paths x 4 recorded vertices, HIP-event timed."""
    rng = np.random.default_rng(5)
    n_inf, n_tr = 1920 * 1080, 65536 * 4
    f = np.zeros((n_inf, 16), np.float32)
    f[:, :14] = rng.random((n_inf, 14), dtype=np.float32)
    t = rng.random((n_tr, 3), dtype=np.float32)
    d_f, d_o, d_t = eng.alloc(f.nbytes), eng.alloc(n_inf * 12), eng.alloc(t.nbytes)
    try:
        eng.h2d(d_f, f)
        eng.h2d(d_t, t)
        eng.nrc_reset(seed=3)
        inf_ms, tr_ms = [], []
        for k in range(reps + 2):
            eng.nrc_infer_device(d_f, n_inf, d_o)
            eng.nrc_train_step_device(d_f, d_t, n_tr)
            a, b = eng.nrc_timing()
            if k >= 2:
                inf_ms.append(a)
                tr_ms.append(b)
    finally:
        for d in (d_f, d_o, d_t):
            eng.free(d)
    im, tm = float(np.median(inf_ms)), float(np.median(tr_ms))
    itf = n_inf * NRC_FWD_FLOPS / (im * 1e-3) / 1e12
    ttf = n_tr * NRC_TRAIN_FLOPS / (tm * 1e-3) / 1e12
    return {"infer_samples": n_inf, "infer_ms": round(im, 4), "infer_tflops": round(itf, 2),
            "train_samples": n_tr, "train_ms": round(tm, 4), "train_tflops": round(ttf, 2),
            "flops_per_sample": {"infer": NRC_FWD_FLOPS, "train": NRC_TRAIN_FLOPS},
            "mfma_util_infer": round(itf / BF16_PEAK_TFLOPS, 4),
            "mfma_util_train": round(ttf / BF16_PEAK_TFLOPS, 4),
            "peak_tflops": BF16_PEAK_TFLOPS, "dtype": "bf16 (fp32 accumulate)"}


def lum(img):
    return 0.2126 * img[..., 0] + 0.7152 * img[..., 1] + 0.0722 * img[..., 2]


def rel_rmse(img, ref):
    """Luminance RMSE (Rec. 709 weights) over the mean reference luminance."""
    return float(np.sqrt(np.mean((lum(img) - lum(ref)) ** 2)) / np.mean(lum(ref)))


def _accumulate(eng, ivp, width, height, first_frame, frames, seed):
    eng.reset()
    for f in range(frames):
        eng.render_frame(ivp, width, height, first_frame + f, seed, 8)
        if f % 256 == 255:
            eng.synchronize()
    return eng.capture(width, height)[..., :3].astype(np.float64)


def _tile_means(x, t=16):
    h, w = (x.shape[0] // t) * t, (x.shape[1] // t) * t
    return x[:h, :w].reshape(h // t, t, w // t, t).mean(axis=(1, 3))


def time_to_rmse(W, width, height, ref_spp, target, train_paths, max_frames, teacher_vertices=4):
    """Plain vs cache-terminated path tracer on the Cornell scene: wall time
    until the image's relative RMSE first reaches `target`, measured.

    The reference is two independent plain halves A and B (ref_spp/2 samples
    each).  Its own noise is removed from the error: with R = (A + B)/2,
    E[(x - R)^2] = MSE(x) + Var(R) and Var(R) = E[(A - B)^2] / 4, so
    MSE(x) = mean((x - R)^2) - mean((A - B)^2) / 4 (per-pixel luminance,
    averaged over the image; the correction's own sampling error over 2M
    pixels is reported as `error_uncertainty`).  Each estimator renders frames
    back to back; its error is checked on a geometric schedule and, once it is
    within 1.5x of the target, every `fine_step` frames; the reported time is
    the accumulated wall time at the first check at or below the target (no
    extrapolation).  Timed intervals end with a full engine synchronisation
    (the cache's training step included); image read-backs and error checks
    are outside them.  A least-squares fit e^2 = b^2 + v/N is reported beside
    the measurement, with b = 0 for the plain tracer (unbiased by
    construction)."""
    from vitrum_amd import native
    seed = 0x9E3779B9
    eng = native.Engine(0)
    sc = W.cornell_scene()
    eng.set_scene(sc)
    ivp = W.inv_view_proj(sc["camera"], width, height)
    eng.nrc_render_config(False)
    t0 = time.perf_counter()
    half = max(1, ref_spp // 2)
    ref_a = _accumulate(eng, ivp, width, height, 50_000_000, half, seed)
    ref_b = _accumulate(eng, ivp, width, height, 60_000_000, half, seed)
    log(f"reference {2 * half} spp in {time.perf_counter() - t0:.0f} s")
    ref = 0.5 * (ref_a + ref_b)
    la, lb, lr = lum(ref_a), lum(ref_b), lum(ref)
    mean_lum = float(np.mean(lr))
    dd = (la - lb) ** 2 / 4.0
    ref_var = float(np.mean(dd))
    # standard error of the noise correction (per-pixel terms are independent)
    corr_se = float(np.std(dd) / np.sqrt(dd.size))
    out = {"scene": "cornell (config-1 box, 32,768 tris)", "width": width, "height": height,
           "reference_spp": 2 * half, "target_rel_rmse": target, "train_paths_per_frame": train_paths,
           "teacher_vertices": teacher_vertices,
           "reference_rel_rmse": round(float(np.sqrt(ref_var)) / mean_lum, 5),
           "error_uncertainty": round(corr_se / (2 * target * mean_lum ** 2), 6),
           "method": "first measured crossing of the noise-corrected relative RMSE vs a split-half plain "
                     "reference; frame time = wall clock, frames back to back"}
    for name, nrc in (("plain", False), ("nrc", True)):
        eng.nrc_render_config(nrc, train_paths)
        if nrc:
            eng.nrc_reset(seed=11)
            eng.nrc_teacher_vertices(teacher_vertices)
        eng.reset()
        wall_ms, curve, crossing = 0.0, [], None
        check, fine = 1, None
        eng.synchronize()
        t_start = time.perf_counter()
        img = None
        for f in range(max_frames):
            eng.render_frame(ivp, width, height, f, seed, 8)
            if f + 1 == check or f + 1 == max_frames:
                eng.synchronize()
                wall_ms += (time.perf_counter() - t_start) * 1e3
                img = eng.capture(width, height)[..., :3].astype(np.float64)
                mse = float(np.mean((lum(img) - lr) ** 2)) - ref_var
                e = float(np.sqrt(max(mse, 0.0))) / mean_lum
                curve.append([f + 1, round(wall_ms, 2), round(e, 5)])
                if len(curve) % 8 == 1 or e <= target:
                    log(f"{name}: {f + 1} spp, {wall_ms:.0f} ms, rel RMSE {e:.4f}")
                if e <= target:
                    crossing = (f + 1, wall_ms)
                    break
                if fine is None and e < 1.5 * target:
                    fine = max(8, (f + 1) // 64)
                check = check + fine if fine else max(check + 1, int(check * 1.5))
                eng.synchronize()
                t_start = time.perf_counter()
        frames = curve[-1][0]
        entry = {"frames": frames, "frame_ms": round(wall_ms / frames, 3), "curve": curve[::max(1, len(curve) // 24)]}
        entry["spp_at_target"] = crossing[0] if crossing else None
        entry["time_to_target_ms"] = round(crossing[1], 1) if crossing else None
        pts = np.array([[n, e] for n, _, e in curve if n >= 8], np.float64)
        if pts.shape[0] >= 2:
            if nrc:
                A = np.stack([np.ones(pts.shape[0]), 1.0 / pts[:, 0]], axis=1)
                (b2, v), *_ = np.linalg.lstsq(A, pts[:, 1] ** 2, rcond=None)
                b2 = max(float(b2), 0.0)
            else:
                b2, v = 0.0, float(np.sum(pts[:, 1] ** 2 / pts[:, 0]) / np.sum(1.0 / pts[:, 0] ** 2))
            entry["fit_bias"] = round(float(np.sqrt(b2)), 5)
            entry["fit_sample_variance"] = round(float(v), 5)
        # bias at the crossing: signed image-mean luminance difference and the
        # RMS of 16x16-tile mean differences (noise shrinks 256x in the tiles;
        # the plain tracer's figure is the noise floor of this statistic)
        d = (lum(img) - lr) / mean_lum
        tiles = _tile_means(d)
        entry["mean_rel_bias"] = round(float(d.mean()), 6)
        entry["tile16_rel_diff_rms"] = round(float(np.sqrt(np.mean(tiles ** 2))), 6)
        entry["tile16_rel_diff_p99"] = round(float(np.quantile(np.abs(tiles), 0.99)), 6)
        out[name] = entry
    p, n = out["plain"].get("time_to_target_ms"), out["nrc"].get("time_to_target_ms")
    out["nrc_speedup"] = round(p / n, 3) if p and n else None
    eng.close()
    return out


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
    ap.add_argument("--near-first", type=int, default=-1,
                    help="2: nearest first in BLASes and the TLAS, 1: in BLASes only, 0: left first (-1: default)")
    ap.add_argument("--defer-min", type=int, default=0,
                    help="lanes that must need a leaf/instance operation before a wave runs it (0: default)")
    ap.add_argument("--node-width", type=int, default=4, choices=(2, 4),
                    help="traversal node width: 4-wide trees collapsed from the binary BVH (default) or binary")
    ap.add_argument("--no-nrc-mlp", action="store_true", help="skip the cache MLP kernel measurement")
    ap.add_argument("--no-blas-build", action="store_true", help="skip the BLAS build measurement")
    ap.add_argument("--no-render-rmse", action="store_true",
                    help="skip the plain vs cache time-to-RMSE measurement (about three minutes of GPU time)")
    ap.add_argument("--rmse-size", default="1920x1080")
    ap.add_argument("--rmse-ref-spp", type=int, default=4096)
    ap.add_argument("--rmse-target", type=float, default=1e-2)
    ap.add_argument("--nrc-teacher-vertices", type=int, default=4,
                    help="shaded vertices of a self-training path before its terminal cache query (0: full length)")
    ap.add_argument("--rmse-max-frames", type=int, default=16384,
                    help="frames after which an estimator that has not reached the target stops")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path).", file=sys.stderr)
        sys.exit(2)

    from vitrum_amd import native, workloads as W

    eng = native.Engine(0)
    eng.set_node_width(args.node_width)
    near_first = 2 if args.near_first < 0 else args.near_first
    if args.refill_min or args.steps_per_refill or args.near_first >= 0 or args.defer_min:
        eng.set_trace_tuning(args.refill_min or 8, args.steps_per_refill or 16, near_first, args.defer_min or 12)
    key = kernel_key(f"node_width={args.node_width} refill={args.refill_min or 8} "
                     f"steps={args.steps_per_refill or 16} near_first={near_first} "
                     f"defer={args.defer_min or 12}")
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
            "node_width": args.node_width,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": measured_traffic(key),
            "kernel_key": key,
            "bytes_per_launch": int(ref_bytes),
        },
        "cpu_baseline": None,
    }
    if not args.no_nrc_mlp:
        result["nrc_mlp"] = nrc_mlp(eng)
        log(f"cache MLP: {result['nrc_mlp']}")
    if not args.no_blas_build:
        result["blas_build"] = blas_build(eng, native)
        log(f"BLAS build: {result['blas_build']}")
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(sc, rays, args.cpu_seconds)
    eng.close()
    if not args.no_render_rmse:
        w, h = (int(x) for x in args.rmse_size.split("x"))
        try:
            result["time_to_rmse"] = time_to_rmse(W, w, h, args.rmse_ref_spp, args.rmse_target, 65536,
                                                  args.rmse_max_frames, args.nrc_teacher_vertices)
        except Exception as exc:  # the traversal line above stays valid
            result["time_to_rmse"] = {"error": repr(exc)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
