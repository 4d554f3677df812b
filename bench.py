"""Benchmark: Mpaths/s and ms/frame of the single-scattering tracer on one MI355X.

Headline (the JSON line's `value`): the CT-like plain tracer workload - 256^3 uint16 phantom
of 12 nested ellipsoids in 5 intensity bands with N(0, 20^2) noise, 1024 x 768 film, 2 plane
lights + a uniform environment sphere light, hybrid BRDF/phase shading keyed on gradient
magnitude, 8-node transfer function, bilateral display filter, 64 progressive passes. A step
is one progressive pass over the whole film (RenderEstimate: trace + 3x3 frame filter +
running average + tonemap + bilateral). Inputs are resident on the device before timing;
`value` = paths / wall time of the timed passes. Each workload is first autotuned
(er_autotune: brick, tile, workgroups and shadow waves per CU; outside the timed region,
which it leaves where it was; `--no-tune` keeps the defaults); `autotune` reports the choice.

Also reported, each on its own workload:
  roofline      the primary-march kernel (the dominant one) against the HBM roofline: the
                bytes it must read, 16 B per trilinear sample it fetches (8 uint16 corners) +
                32 B per path (seeds in/out, frame estimate), unit counts from a counting replay
                of the same passes with the same seeds, over its HIP-event-timed average
                duration. The bytes of every sample the reference would fetch (exact empty-space
                skipping removes almost all of them) are reported apart as
                reference_equivalent, never as the fraction;
  rmse          tonemapped running estimate vs the FP64 oracle's converged image (fixtures
                under tests/golden: every pixel for the small workload, 4096 for the others);
  cpu_baseline  the CPU oracle (the reference estimator restated, FP32, OpenMP over rows) on
                every core this process may use, after a warm-up pass; host and compiler
                recorded;
  small         the 64^3 blob workload (256 x 256, 16 passes, run on to the RMSE gate):
                Mpaths/s, RMSE, GPU / CPU;
  woodcock      the CT workload with Woodcock delta tracking (per-macrocell majorants);
  large         the 512 x 512 x 460 clipped volume at 1920 x 1080, 256 passes: Mpaths/s, RMSE;
  adjoint       the differentiation workload (128^3, 512 x 512, 32 passes): forward and
                path-replay adjoint wall times;
  batched       64 jobs of uneven size: one er_render_batch vs the same jobs one after another.
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
    """HBM bytes per primary-kernel launch from the committed rocprofv3 --pmc summary."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data.get("hbm_bytes_per_launch")


def copy_rate_gbs(nbytes=1 << 30, reps=10):
    """Device-to-device copy rate (bytes read + written per second): the HBM rate a plain
    streaming kernel reaches on this GPU, reported beside the published peak. It runs in a
    child process: torch's HIP runtime does not see the GPU in a process where the tracer
    library's runtime already holds it."""
    import subprocess
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-rate", str(nbytes), str(reps)],
                         capture_output=True, text=True, timeout=300, check=True).stdout
    return float(out.strip().splitlines()[-1])


def _copy_rate_child(nbytes, reps):
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    secs = t0.elapsed_time(t1) * 1e-3 / reps
    print(2.0 * nbytes / secs / 1e9, flush=True)


def cpu_quota():
    """CPUs granted by the cgroup's CPU bandwidth limit (cgroup v2 cpu.max, v1 cfs quota), or
    None without a limit."""
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        return None if q == "max" else max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    try:
        with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
            q = int(f.read())
        with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
            per = int(f.read())
        return None if q <= 0 else max(1, q // per)
    except (OSError, ValueError):
        return None


def cpu_threads():
    """Every core this process may run on: its affinity mask, capped by the cgroup's CPU quota
    (more threads than the quota grants would only be throttled)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    q = cpu_quota()
    return min(n, q) if q else n


def cpu_host():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"nproc": os.cpu_count(), "affinity": affinity, "cgroup_quota_cpus": cpu_quota(), "threads": cpu_threads(),
            "model": model, "compiler": O.build_info()}


def cpu_rate(scene_builder, threads, target_s=15.0, max_passes=12):
    """The CPU oracle on full passes of the scene after one untimed warm-up pass: as many
    passes as fit about target_s (at least one, at most max_passes). Returns Mpaths/s, the
    timed seconds and the pass count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    sc = scene_builder()
    orc = O.Oracle(sc, nthreads=threads)
    if sc.tracer.settings.shading_type not in (0, 1):
        orc.compute_gm_max()
    W, H = sc.film
    s1, s2 = O.seeds_for(W, H, 99)
    t0 = time.perf_counter()
    orc.trace(s1, s2)                        # warm-up (caches, OpenMP pool, page faults)
    one = time.perf_counter() - t0
    passes = int(max(1, min(max_passes, target_s // max(one, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(passes):
        orc.trace(s1, s2)
    dt = time.perf_counter() - t0
    return W * H * passes / dt / 1e6, dt, passes


def timed_passes(er, r, steps):
    lib = er.load()
    lib.er_synchronize()
    t0 = time.perf_counter()
    timing = r.render_passes(steps)
    lib.er_synchronize()
    return time.perf_counter() - t0, timing


def tune_summary(r):
    """er_autotune on the scene (outside the timed region; the render state is restored):
    brick, tile, primary workgroups and shadow waves per CU, heavy tiles, rays per wave."""
    t0 = time.perf_counter()
    t = r.autotune(passes=0)
    return {"brick": t["brick"], "tile": [t["tile_w"], t["tile_h"]], "blocks_per_cu": t["blocks_per_cu"],
            "shadow_waves_per_cu": t["shadow_waves_per_cu"], "heavy_tiles": t["heavy_tiles"],
            "heavy_rays_per_wave": t["heavy_rays_per_wave"], "rays_per_wave": t["rays_per_wave"],
            "shadow_rays_per_wave": t["shadow_rays_per_wave"], "trials": t["trials"],
            "ms_per_pass": round(t["ms_per_pass"], 4), "default_ms_per_pass": round(t["default_ms_per_pass"], 4),
            "tune_s": round(time.perf_counter() - t0, 3)}


def bench_small(er, scenes, threads, no_cpu, tune=True):
    """The 64^3 blob workload: 16 passes; more if the RMSE gate needs them."""
    sc = scenes.config_small()
    r = er.Renderer(sc)
    r.seed(4321)
    tn = tune_summary(r) if tune else None
    r.render_passes(2)
    r.rebind()
    passes, wall = 0, 0.0
    rm = None
    for block in (16, 16, 32, 64, 128, 256, 512):
        dt, _ = timed_passes(er, r, block)
        wall += dt
        passes += block
        rm = rmse_vs_golden(r.running(), os.path.join(ROOT, "tests", "golden", "fp64_small.npz"))
        if rm is None or rm["display"] <= 0.01:
            break
    W, H = r.W, r.H
    value = W * H * passes / wall / 1e6
    out = {"workload": "blobs_64c_256x256", "passes": passes, "mpaths_per_s": round(value, 2),
           "ms_per_frame": round(wall * 1e3 / passes, 4), "rmse": rm, "autotune": tn}
    if not no_cpu:
        v, dt, n = cpu_rate(scenes.config_small, threads)
        out["cpu_mpaths_per_s"] = round(v, 4)
        out["cpu_threads"] = threads
        out["cpu_sample"] = f"{n} full passes of 256x256 paths after a warm-up pass, {dt:.1f} s"
        out["gpu_over_cpu"] = round(value / v, 1)
    r.close()
    return out


def bench_large(er, scenes, tune=True):
    """The large anisotropic workload (512 x 512 x 460, 0.7 x 0.7 x 1.0 spacing, clipping box,
    1920 x 1080, 3 plane lights, hybrid shading): throughput over its 256 passes and the RMSE
    of the result against the FP64 fixture (4096 pixels at 16384 spp)."""
    sc = scenes.config_large()
    r = er.Renderer(sc)
    r.seed(99)
    tn = tune_summary(r) if tune else None
    r.render_passes(2)
    r.rebind()
    dt, t = timed_passes(er, r, sc.passes)
    out = {"workload": "tubes_512x512x460_1920x1080_clip", "passes": sc.passes,
           "mpaths_per_s": round(r.W * r.H * sc.passes / dt / 1e6, 2), "ms_per_frame": round(dt * 1e3 / sc.passes, 4),
           "rmse": rmse_vs_golden(r.running(), os.path.join(ROOT, "tests", "golden", "fp64_large.npz")),
           "autotune": tn}
    r.close()
    return out


def bench_woodcock(er, scenes, passes, tune=True):
    """The CT workload with Woodcock delta tracking (full estimator, per-macrocell majorants)
    instead of the ray march: throughput over `passes` passes and the RMSE against the FP64
    fixture of the same sampler (the two samplers converge to different images: the ray
    march is a jittered fixed-step quadrature of the free-flight integral)."""
    from exposure_render import capi
    sc = scenes.config_ct()
    sc.tracer.settings.sampler = capi.SAMPLER_WOODCOCK
    r = er.Renderer(sc)
    r.seed(77)
    tn = tune_summary(r) if tune else None
    r.render_passes(2)
    r.rebind()
    dt, t = timed_passes(er, r, passes)
    out = {"workload": "ct_phantom_256c_1024x768_woodcock", "passes": passes,
           "mpaths_per_s": round(r.W * r.H * passes / dt / 1e6, 2), "ms_per_frame": round(dt * 1e3 / passes, 4),
           "rmse": rmse_vs_golden(r.running(), os.path.join(ROOT, "tests", "golden", "fp64_ct_woodcock.npz")),
           "autotune": tn}
    r.close()
    return out


def bench_adjoint(er, scenes):
    """Differentiation workload: forward passes vs the path-replay adjoint of the same passes."""
    sc = scenes.config_adjoint()
    r = er.Renderer(sc)
    lib = er.load()
    s1, s2 = er.seeds_for(r.W, r.H, 1)
    out = {"workload": "blobs_128c_512x512_32spp_lut256x4", "passes": sc.passes}
    for name, lut in (("initial_lut", scenes.initial_tf_lut()), ("target_lut", scenes.random_tf_lut(4))):
        r.set_tf_lut(lut, sc.tf_range)
        r.set_seeds(s1, s2)
        r.render_passes(2)                      # warm-up
        r.adjoint_lut(s1, s2, 2, np.ones((r.H, r.W, 3), np.float32))
        r.set_tf_lut(lut, sc.tf_range)
        r.set_seeds(s1, s2)
        fwd, _ = timed_passes(er, r, sc.passes)
        img = r.running()[..., :3]
        g_img = (2.0 * img).astype(np.float32)  # dL/dimage of L = sum(image^2)
        lib.er_synchronize()
        t0 = time.perf_counter()
        r.adjoint_lut(s1, s2, sc.passes, g_img)
        adj = time.perf_counter() - t0
        out[name] = {"forward_ms": round(fwd * 1e3, 3), "adjoint_ms": round(adj * 1e3, 3),
                     "adjoint_over_forward": round(adj / fwd, 3)}
    r.close()
    return out


def bench_batched(er, scenes, n_jobs):
    """64 jobs: er_render_batch vs the same jobs rendered one after another, with er_render_passes
    (the primary comparison: each job's passes already pipelined on two streams) and with
    RenderEstimate-style single synchronous passes. The batch runs with the brick size, tile
    shape and workgroups per CU er_autotune_batch picks; the sequential runs use the defaults.
    The results are bit-identical (same seeds, same paths, and no launch configuration changes
    the image), so the RMSE is equal."""
    lib = er.load()
    lib.er_set_list_capacity(4096)
    jobs = scenes.config_batch(n_jobs=n_jobs)
    rs = [er.Renderer(sc) for sc in jobs]
    passes = [sc.passes for sc in jobs]
    seeds = [er.seeds_for(r.W, r.H, 40 + k) for k, r in enumerate(rs)]
    for r in rs:
        r.render_passes(1)

    def reset():
        for r, (s1, s2) in zip(rs, seeds):
            r.rebind()
            r.set_seeds(s1, s2)
        lib.er_synchronize()

    reset()
    t0 = time.perf_counter()
    for r, p in zip(rs, passes):
        for _ in range(p):
            r.render_estimate()
    seq_est = time.perf_counter() - t0
    ref = [r.frame() for r in rs]
    reset()
    t0 = time.perf_counter()
    for r, p in zip(rs, passes):
        r.render_passes(p)
    lib.er_synchronize()
    seq_passes = time.perf_counter() - t0
    # launch configuration for the batch from the autotuner (a one-time cost, reported apart)
    reset()
    t0 = time.perf_counter()
    tune = er.autotune_batch(rs, rounds=4)
    tune_s = time.perf_counter() - t0
    # the first call captures and instantiates the batch graph, later calls on the same jobs
    # reuse it (er_render_batch's graph cache); both are reported, the second is the timed one
    reset()
    t0 = time.perf_counter()
    er.render_batch(rs, passes)
    lib.er_synchronize()
    bat_first = time.perf_counter() - t0
    reset()
    t0 = time.perf_counter()
    er.render_batch(rs, passes)
    lib.er_synchronize()
    bat = time.perf_counter() - t0
    same = all(np.array_equal(a, r.frame()) for a, r in zip(ref, rs))
    paths = sum(r.W * r.H * p for r, p in zip(rs, passes))
    for r in rs:
        r.close()
    lib.er_set_list_capacity(32)
    return {"workload": f"{n_jobs}_jobs_edges_32-256_films_128-1024_spp_8-128", "paths": int(paths),
            "sequential_estimate_s": round(seq_est, 4), "sequential_passes_s": round(seq_passes, 4),
            "batch_s": round(bat, 4), "batch_first_call_s": round(bat_first, 4),
            "batch_over_sequential_passes": round(bat / seq_passes, 4),
            "batch_over_sequential_estimate": round(bat / seq_est, 4),
            "batch_mpaths_per_s": round(paths / bat / 1e6, 2), "identical_frames": bool(same),
            "autotune": {"brick": tune["brick"], "tile": [tune["tile_w"], tune["tile_h"]],
                         "blocks_per_cu": tune["blocks_per_cu"], "rays_per_wave": tune["rays_per_wave"],
                         "shadow_rays_per_wave": tune["shadow_rays_per_wave"], "trials": tune["trials"],
                         "round_ms": round(tune["ms_per_pass"], 3),
                         "default_round_ms": round(tune["default_ms_per_pass"], 3), "tune_s": round(tune_s, 3)}}


def main():
    if len(sys.argv) == 4 and sys.argv[1] == "--copy-rate":
        _copy_rate_child(int(sys.argv[2]), int(sys.argv[3]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-count", action="store_true", help="skip the counting replay (roofline)")
    ap.add_argument("--no-extra", action="store_true", help="headline workload only")
    ap.add_argument("--no-tune", action="store_true",
                    help="render with the default launch configuration instead of the autotuned one")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (default: every core)")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched section (its own JSON line; for kernel profiles of the batch)")
    ap.add_argument("--large-only", action="store_true",
                    help="only the large section (its own JSON line; for kernel profiles of that workload)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint section (its own JSON line; for kernel profiles of the adjoint)")
    ap.add_argument("--woodcock-only", action="store_true",
                    help="only the Woodcock section (its own JSON line; for kernel profiles of that sampler)")
    ap.add_argument("--small-only", action="store_true",
                    help="only the small section, without its CPU leg (its own JSON line; for kernel profiles)")
    ap.add_argument("--profile", choices=["small", "ct", "large", "woodcock", "adjoint"], default=None,
                    help="render --steps passes of one workload (autotuned unless --no-tune) and nothing else, "
                         "for rocprofv3 kernel statistics and counter runs")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy a device buffer of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the tracer has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    if args.batched_only:
        import exposure_render as er
        from exposure_render import scenes
        print(json.dumps({"batched": bench_batched(er, scenes, args.batch_jobs)}))
        return
    if args.large_only:
        import exposure_render as er
        from exposure_render import scenes
        print(json.dumps({"large": bench_large(er, scenes, not args.no_tune)}))
        return
    if args.adjoint_only:
        import exposure_render as er
        from exposure_render import scenes
        print(json.dumps({"adjoint": bench_adjoint(er, scenes)}))
        return
    if args.profile:
        import exposure_render as er
        from exposure_render import capi, scenes
        name = args.profile
        sc = {"small": scenes.config_small, "ct": scenes.config_ct, "large": scenes.config_large,
              "woodcock": scenes.config_ct, "adjoint": scenes.config_adjoint}[name]()
        if name == "woodcock":
            sc.tracer.settings.sampler = capi.SAMPLER_WOODCOCK
        r = er.Renderer(sc)
        r.seed(11)
        if name == "adjoint":
            r.set_tf_lut(scenes.initial_tf_lut(), sc.tf_range)
        tn = None if args.no_tune else tune_summary(r)
        if args.warmup > 0:
            r.render_passes(args.warmup)
        dt, t = timed_passes(er, r, args.steps)
        print(json.dumps({"profile": name, "film": [r.W, r.H], "passes": args.steps,
                          "ms_per_pass": round(dt * 1e3 / args.steps, 4),
                          "primary_ms_per_pass": round(t.primary_ms / args.steps, 4), "autotune": tn}))
        r.close()
        return
    if args.woodcock_only:
        import exposure_render as er
        from exposure_render import scenes
        print(json.dumps({"woodcock": bench_woodcock(er, scenes, 64, not args.no_tune)}))
        return
    if args.small_only:
        import exposure_render as er
        from exposure_render import scenes
        print(json.dumps({"small": bench_small(er, scenes, 1, True, not args.no_tune)}))
        return
    if args.copy_probe:
        import torch
        a = torch.empty(args.copy_probe // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe}))
        return

    import exposure_render as er
    from exposure_render import scenes
    er.shutdown()
    sc = scenes.config_ct()
    r = er.Renderer(sc)
    r.seed(1234)
    W, H = r.W, r.H
    tune = None if args.no_tune else tune_summary(r)
    if args.warmup > 0:
        r.render_passes(args.warmup)
    s1, s2 = r.get_seeds()

    wall, timing = timed_passes(er, r, args.steps)
    paths = W * H * args.steps
    value = paths / wall / 1e6
    ms_per_step = wall * 1e3 / args.steps
    running = r.running()
    golden_ct = os.path.join(ROOT, "tests", "golden", "fp64_ct.npz")
    rm = rmse_vs_golden(running, golden_ct)
    if rm is not None:
        # the passes this workload needs to reach the gate (rendered after the timed region)
        total, g = args.steps, rm
        while g["display"] > 0.01 and total < 2048:
            r.render_passes(total)
            total *= 2
            g = rmse_vs_golden(r.running(), golden_ct)
        rm["gate"] = {"passes": total, "display": g["display"], "met": g["display"] <= 0.01}

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
        avg_ms = timing.primary_ms / args.steps
        # the bytes k_primary must read: 16 B per trilinear sample it fetches + 32 B per path
        per_launch = (16.0 * counts["primary_fetches"] + 32.0 * counts["paths"]) / args.steps
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic_from_profile(os.path.join(ROOT, "profiles", "k_primary_ct_traffic.json")),
                    "kernel": "k_primary<1,0,1,%d>" % (64 // tune["rays_per_wave"] if tune else 1),
                    "avg_launch_ms": round(avg_ms, 4),
                    "bytes_per_launch": int(per_launch),
                    "units_per_launch": {"fetched_samples": round(counts["primary_fetches"] / args.steps),
                                         "paths": W * H}}
        if roofline["traffic"]:
            roofline["hbm_GBs"] = round(roofline["traffic"] / (avg_ms * 1e-3) / 1e9, 1)
        copy = copy_rate_gbs()
        roofline["measured_copy_GBs"] = round(copy, 1)
        roofline["frac_of_measured_copy"] = round(achieved / copy, 4)
        # every sample of the reference's march, skipped or not (not a memory rate of this kernel)
        ref_eq = (16.0 * counts["primary_samples"] + 32.0 * counts["paths"]) / args.steps
        samples = counts["primary_samples"] + counts["shadow_samples"] + counts["shading_samples"]
        whole = (16.0 * samples + 64.0 * counts["paths"]) / args.steps
        roofline["reference_equivalent"] = {
            "bytes_per_launch": int(ref_eq), "GBs": round(ref_eq / (avg_ms * 1e-3) / 1e9, 1),
            "skipped_fraction": round(1.0 - counts["primary_fetches"] / max(counts["primary_samples"], 1), 5),
            "whole_pass_bytes": int(whole), "whole_pass_GBs": round(whole / (ms_per_step * 1e-3) / 1e9, 1)}

    threads = args.cpu_threads or cpu_threads()
    cpu = None
    if not args.no_cpu:
        v, dt, n = cpu_rate(scenes.config_ct, threads)
        cpu = {"value": round(v, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
               "sample": f"{n} full passes of the same scene ({W}x{H} paths each) by the FP32 oracle restatement "
                         f"after one warm-up pass, {dt:.1f} s",
               "host": cpu_host()}
    r.close()

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"ct_phantom_256c_1024x768_{args.steps}spp", "film": [W, H],
                   "volume": list(sc.voxels.shape[::-1]), "spp": args.steps, "estimator": "full",
                   "shading": "hybrid", "lights": "2 plane + environment sphere", "bilateral": True},
        "ms_per_frame": {"total": round(timing.total_ms / args.steps, 4),
                         "primary": round(timing.primary_ms / args.steps, 4),
                         "trace_overlapping": round(timing.trace_ms / args.steps, 4),
                         "post": round(timing.post_ms / args.steps, 4)},
        "rmse": rm, "roofline": roofline, "cpu_baseline": cpu, "counts": counts, "autotune": tune,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    if not args.no_extra:
        out["small"] = bench_small(er, scenes, threads, args.no_cpu, not args.no_tune)
        out["woodcock"] = bench_woodcock(er, scenes, 64, not args.no_tune)
        out["large"] = bench_large(er, scenes, not args.no_tune)
        out["adjoint"] = bench_adjoint(er, scenes)
        out["batched"] = bench_batched(er, scenes, args.batch_jobs)
    er.shutdown()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
