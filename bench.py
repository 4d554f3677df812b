"""Benchmark: Mpaths/s and ms/frame of the single-scattering tracer on one MI355X.

A step is one progressive pass (RenderEstimate: trace + 3x3 frame filter + running average +
tonemap + bilateral display filter) over the whole film. The workload is the CT-like plain
tracer scene: 256^3 uint16 phantom of 12 nested ellipsoids in 5 intensity bands with N(0, 20^2)
noise, 1024 x 768 film, 2 plane lights + a uniform environment sphere light, hybrid
BRDF/phase shading keyed on gradient magnitude, 8-node transfer function, 64 progressive
passes. Inputs (volume, transfer functions, seeds) are resident on the device before timing.

Printed JSON line: value = paths / wall time of the K timed passes; roofline of the trace
kernel (algorithmic bytes 16 B per trilinear sample + 64 B per path, exact sample counts from a
counting replay of the same passes with the same seeds, divided by the HIP-event time of the
trace kernel); RMSE of the tonemapped running estimate against the FP64 oracle's expectation
(tests/golden/fp64_ct.npz); the CPU oracle timed on a bounded sample on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "exposure-render_amd"))

METRIC = "Mpaths/s + ms/frame at RMSE≤1% vs FP64; batched & adjoint vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0


def tonemap(xyz, exposure):
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    rgb = np.stack([3.240479 * x - 1.537150 * y - 0.498535 * z,
                    -0.969256 * x + 1.875991 * y + 0.041556 * z,
                    0.055648 * x - 0.204043 * y + 1.057311 * z], axis=-1)
    return np.clip(1.0 - np.exp(-(rgb / exposure)), 0.0, 1.0)


def rmse_vs_golden(running, golden_path):
    if not os.path.exists(golden_path):
        return None
    g = np.load(golden_path)
    cx, cy = g["cx"], g["cy"]
    ref = g["expect"]
    exposure = float(g["exposure"])
    got = running[cy, cx, :].astype(np.float64)
    d = tonemap(got[:, :3], exposure) - tonemap(ref[:, :3], exposure)
    rmse = float(np.sqrt(np.mean(d * d)))
    rel = float(np.sqrt(np.mean(((got[:, :3] - ref[:, :3]) / np.maximum(np.abs(ref[:, :3]), 1e-3)) ** 2)))
    return {"display": rmse, "xyz_relative": rel, "pixels": int(len(cx)), "reference_spp": int(g["spp"])}


def traffic_from_profile(path):
    """HBM bytes per trace launch from the committed rocprofv3 --pmc summary, if present."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data.get("hbm_bytes_per_launch")


def cpu_baseline(scene_builder, passes, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    sc = scene_builder()
    orc = O.Oracle(sc, nthreads=threads)
    orc.compute_gm_max()
    W, H = sc.film
    s1, s2 = O.seeds_for(W, H, 99)
    t0 = time.perf_counter()
    for _ in range(passes):
        orc.trace(s1, s2)
    dt = time.perf_counter() - t0
    return W * H * passes / dt / 1e6, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-count", action="store_true", help="skip the counting replay (roofline)")
    ap.add_argument("--cpu-passes", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the tracer has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import exposure_render as er
    from exposure_render import scenes
    er.shutdown()
    sc = scenes.config_ct()
    r = er.Renderer(sc)
    r.seed(1234)
    lib = er.load()
    W, H = r.W, r.H
    if args.warmup > 0:
        r.render_passes(args.warmup)
    s1, s2 = r.get_seeds()
    n0 = r.no_iterations()

    lib.er_synchronize()
    t0 = time.perf_counter()
    timing = r.render_passes(args.steps)
    lib.er_synchronize()
    wall = time.perf_counter() - t0

    paths = W * H * args.steps
    value = paths / wall / 1e6
    ms_per_step = wall * 1e3 / args.steps
    running = r.running()
    rm = rmse_vs_golden(running, os.path.join(ROOT, "tests", "golden", "fp64_ct.npz"))

    roofline = None
    counts = None
    if not args.no_count:
        # replay the same passes with the same seeds and counters on (exact unit counts)
        r.set_seeds(s1, s2)
        r.reset_counters()
        r.set_counting(True)
        r.render_passes(args.steps)
        r.set_counting(False)
        counts = r.counters()
        samples = counts["primary_samples"] + counts["shadow_samples"] + counts["shading_samples"]
        alg_bytes = 16.0 * samples + 64.0 * counts["paths"]
        per_launch = alg_bytes / args.steps
        avg_ms = timing.trace_ms / args.steps
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic_from_profile(os.path.join(ROOT, "profiles", "trace_ct_traffic.json")),
                    "kernel": "k_trace<FULL>", "avg_launch_ms": round(avg_ms, 4),
                    "algorithmic_bytes_per_launch": int(per_launch)}

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        v, dt = cpu_baseline(scenes.config_ct, args.cpu_passes, threads)
        cpu = {"value": round(v, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_passes} full passes of the same scene ({W}x{H} paths each) by the FP32 "
                         f"oracle restatement, {dt:.1f} s"}

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "ct_phantom_256c_1024x768_64spp", "film": [W, H],
                   "volume": list(sc.voxels.shape[::-1]), "spp": args.steps, "estimator": "full",
                   "shading": "hybrid", "lights": "2 plane + environment sphere", "bilateral": True},
        "ms_per_frame": {"total": round(timing.total_ms / args.steps, 4),
                         "trace": round(timing.trace_ms / args.steps, 4),
                         "post": round(timing.post_ms / args.steps, 4)},
        "rmse": rm, "roofline": roofline, "cpu_baseline": cpu, "counts": counts,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    r.close()
    er.shutdown()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
