"""The launch autotuner (er_autotune / er_set_tune).

No launch configuration may change the image: paths are independent per pixel, and the
empty-space skipping is exact for every brick size. So every configuration must render
bit-identical frames, running estimates, RNG states and reference sample counts, and tuning
must leave a render exactly where it was.
"""
import numpy as np
import pytest

import exposure_render as er
from exposure_render import capi, scenes

pytestmark = pytest.mark.gpu

COUNTERS = ["paths", "primary_samples", "shadow_samples", "shading_samples", "scatter_events", "shadow_rays"]


def _render(sc, passes, tune=None, seed=9):
    r = er.Renderer(sc, seed=seed)
    if tune is not None:
        r.set_tune(*tune)
    r.set_counting(True)
    r.reset_counters()
    for _ in range(passes):
        r.render_estimate()
    out = (r.frame(), r.running(), r.get_seeds(), r.get_estimate(), r.counters(), r.get_tune())
    r.close()
    return out


def _same(a, b, what):
    fa, ra, (sa1, sa2), da, ca, _ = a
    fb, rb, (sb1, sb2), db, cb, tb = b
    assert np.array_equal(fa, fb), f"frame differs with {what} {tb}"
    assert np.array_equal(ra, rb), f"running estimate differs with {what} {tb}"
    assert np.array_equal(sa1, sb1) and np.array_equal(sa2, sb2), f"RNG state differs with {what} {tb}"
    assert np.array_equal(da, db), f"display differs with {what} {tb}"
    for k in COUNTERS:
        assert ca[k] == cb[k], (what, k, ca[k], cb[k])


@pytest.mark.parametrize("estimator", [capi.ESTIMATOR_COMPAT, capi.ESTIMATOR_FULL])
def test_every_configuration_renders_the_same_image(fresh, estimator):
    sc = scenes.config_small(scale=0.5, estimator=estimator)
    print(f"[tune] small estimator={estimator}", flush=True)
    base = _render(sc, 3)
    assert base[5]["brick"] == 8 and base[5]["tile_w"] == 8
    for tune in [(2, 8, 7), (4, 4, 7), (16, 16, 5), (32, 32, 3), (8, 8, 9), (4, 32, 4)]:
        _same(base, _render(sc, 3, tune), "tune")


def test_ct_brick_sizes_render_the_same_image(fresh):
    sc = scenes.config_ct(scale=0.25)
    print("[tune] ct 64^3 bricks", flush=True)
    base = _render(sc, 2)
    for tune in [(2, 8, 7), (4, 16, 6), (16, 4, 8)]:
        _same(base, _render(sc, 2, tune), "tune")


def test_large_volume_bricks_render_the_same_image(fresh):
    """The full 512 x 512 x 460 clipped volume (1.9 GB of cell records) at 960 x 540: the
    default 8-cell brick (its occupancy bits take 29.5 KB of LDS) and the 16- and 32-cell
    bricks render the same passes bit for bit, with the reference sample counts."""
    sc = scenes.config_large()
    sc.tracer.camera.film_size[0], sc.tracer.camera.film_size[1] = 960, 540
    print("[tune] large 512x512x460 bricks", flush=True)
    base = _render(sc, 2)
    assert base[5]["brick"] == 8
    for tune in [(16, 8, 7), (32, 16, 5)]:
        _same(base, _render(sc, 2, tune), "tune")
    # 4-cell bricks would need 251 KB of occupancy bits: beyond the 52.5 KB LDS budget
    r = er.Renderer(sc, seed=9)
    with pytest.raises(er.ExposureRenderError):
        r.set_tune(4, 8, 7)
    assert r.get_tune()["brick"] == 8
    r.close()


def test_autotune_restores_the_render_and_caches(fresh):
    sc = scenes.config_small(scale=0.5)
    print("[tune] autotune restore", flush=True)
    ref = er.Renderer(sc, seed=3)
    for _ in range(2):
        ref.render_estimate()
    want = (ref.running().copy(), ref.get_seeds(), ref.get_estimate().copy())
    ref.close()

    r = er.Renderer(sc, seed=3)
    r.set_counting(True)
    r.reset_counters()
    r.render_estimate()
    before = (r.running().copy(), r.get_seeds(), r.no_iterations(), r.counters())
    t = r.autotune(passes=2)
    assert t["trials"] >= 10 and t["ms_per_pass"] > 0 and t["ms_per_pass"] <= t["default_ms_per_pass"]
    assert t["brick"] in (2, 4, 8, 16, 32) and t["tile_w"] * t["tile_h"] == 64 and 3 <= t["blocks_per_cu"] <= 9
    after = (r.running(), r.get_seeds(), r.no_iterations(), r.counters())
    assert np.array_equal(before[0], after[0])
    assert np.array_equal(before[1][0], after[1][0]) and np.array_equal(before[1][1], after[1][1])
    assert before[2] == after[2] == 1
    assert before[3] == after[3]
    assert r.get_tune()["brick"] == t["brick"] and r.get_tune()["tile_w"] == t["tile_w"]
    # the render continues as if no tuning had happened
    r.render_estimate()
    assert np.array_equal(r.running(), want[0])
    assert np.array_equal(r.get_seeds()[0], want[1][0])
    assert np.array_equal(r.get_estimate(), want[2])
    # a second call takes the cached result
    t2 = r.autotune(passes=2)
    assert t2["trials"] == 0 and (t2["brick"], t2["tile_w"], t2["blocks_per_cu"]) == (t["brick"], t["tile_w"],
                                                                                        t["blocks_per_cu"])
    r.close()


def test_set_tune_rejects_invalid_configurations(fresh):
    sc = scenes.config_small(scale=0.25)
    r = er.Renderer(sc, seed=1)
    for bad in [(3, 8, 7), (8, 6, 7), (8, 64, 7), (8, 8, 0), (8, 8, 17), (128, 8, 7), (8, 8, 7, 0), (8, 8, 7, 65),
                (8, 8, 7, 4, -1), (8, 8, 7, 4, 4097), (8, 8, 7, 4, 0, 0), (8, 8, 7, 4, 0, 8), (8, 8, 7, 4, 0, 128), (8, 8, 7, 4, 0, 64, 1),
                (8, 8, 7, 4, 0, 64, 32), (8, 8, 7, 4, 16, 64, 64, 64), (8, 8, 7, 4, 16, 64, 64, 2)]:
        with pytest.raises(er.ExposureRenderError):
            r.set_tune(*bad)
    r.close()


def test_shadow_grid_does_not_change_the_image(fresh):
    """The shadow marches are independent and each writes its own visibility: any number of
    persistent shadow waves (which changes which lane marches which ray, and when) renders the
    same image bit for bit."""
    sc = scenes.config_ct(scale=0.5)   # enough shadow jobs to outrun the static chunks at 1-2 waves per CU
    out = []
    for waves in (1, 2, 4, 8, 64):
        print(f"[tune] shadow waves per CU {waves}", flush=True)
        r = er.Renderer(sc, seed=9)
        r.set_tune(8, 8, 7, waves)
        assert r.get_tune()["shadow_waves_per_cu"] == waves
        r.set_counting(True)
        r.reset_counters()
        r.render_passes(3)
        out.append((r.running().copy(), r.get_estimate().copy(), r.counters()))
        r.close()
    print(f"[tune] shadow rays per pass {out[0][2]['shadow_rays'] // 3}", flush=True)
    for run, disp, cnt in out[1:]:
        assert np.array_equal(run, out[0][0]) and np.array_equal(disp, out[0][1])
        for k in COUNTERS:
            assert cnt[k] == out[0][2][k], (k, cnt[k], out[0][2][k])


@pytest.mark.parametrize("scale", [0.5, 1.0])
def test_tile_order_and_heavy_rays_do_not_change_the_image(fresh, scale):
    """From its second pass on, a pass claims its tiles most expensive first (by the previous
    pass's costs), and with heavy_tiles > 0 marches the rays of the most expensive tiles with
    16 or 4 lanes per ray (group_march) or one whole wave per ray (coop_march). None of it
    changes any pixel, RNG state or sample count."""
    sc = scenes.config_ct(scale=scale)
    out = []
    for heavy, hrpw in ((0, 4), (16, 4), (256, 16), (256, 1), (4096, 4)):
        print(f"[tune] heavy tiles {heavy} rays per wave {hrpw} scale {scale}", flush=True)
        r = er.Renderer(sc, seed=9)
        r.set_tune(8, 8, 7, 4, heavy, 64, 64, hrpw)
        assert r.get_tune()["heavy_tiles"] == heavy and r.get_tune()["heavy_rays_per_wave"] == hrpw
        r.set_counting(True)
        r.reset_counters()
        r.render_passes(4)
        out.append((r.running().copy(), r.get_estimate().copy(), r.get_seeds(), r.counters()))
        r.close()
    for run, disp, (s1, s2), cnt in out[1:]:
        assert np.array_equal(run, out[0][0]) and np.array_equal(disp, out[0][1])
        assert np.array_equal(s1, out[0][2][0]) and np.array_equal(s2, out[0][2][1])
        # the reference's sample counts (the march's own iteration and skip tallies differ: a
        # whole-wave march reads the empty samples inside a block instead of skipping them)
        for k in COUNTERS:
            assert cnt[k] == out[0][3][k], (k, cnt[k], out[0][3][k])


@pytest.mark.parametrize("case", ["small_compat", "small_full", "ct", "large_clipped"])
def test_rays_per_wave_do_not_change_the_image(fresh, case):
    """Primary and shadow rays per wave: 64 (a lane per ray), 16 and 4 (groups of 4 and 16
    lanes per ray, each round trip reading 24 and 32 consecutive samples of the ray in
    group_step) and, for primary rays, 1 (the whole wave per ray) render the same frames,
    running estimates, RNG states and reference sample counts bit for bit: every position is
    the sequential march's chain of additions and the optical depth is summed in sample order.
    The large case has a clipping box."""
    if case == "small_compat":
        sc = scenes.config_small(scale=0.5, estimator=capi.ESTIMATOR_COMPAT)
    elif case == "small_full":
        sc = scenes.config_small(scale=0.5)
    elif case == "ct":
        sc = scenes.config_ct(scale=0.5)
    else:
        sc = scenes.config_large(scale=0.5)
        sc.tracer.camera.film_size[0], sc.tracer.camera.film_size[1] = 480, 270
    print(f"[tune] rays per wave {case}", flush=True)
    base = _render(sc, 3)
    assert base[5]["rays_per_wave"] == 64 and base[5]["shadow_rays_per_wave"] == 64
    for rpw, srpw in ((16, 64), (4, 64), (1, 64), (64, 16), (64, 4), (16, 16)):
        print(f"[tune] rays per wave {case} {rpw} shadow {srpw}", flush=True)
        r = er.Renderer(sc, seed=9)
        t = r.get_tune()
        r.set_tune(t["brick"], t["tile_w"], t["blocks_per_cu"], t["shadow_waves_per_cu"], 0, rpw, srpw)
        assert r.get_tune()["rays_per_wave"] == rpw and r.get_tune()["shadow_rays_per_wave"] == srpw
        r.set_counting(True)
        r.reset_counters()
        for _ in range(3):
            r.render_estimate()
        got = (r.frame(), r.running(), r.get_seeds(), r.get_estimate(), r.counters(), r.get_tune())
        r.close()
        _same(base, got, f"{rpw} / {srpw} rays per wave")


@pytest.mark.parametrize("rpw,srpw", [(16, 16), (4, 4), (1, 64)])
def test_batch_rays_per_wave_matches_sequential(fresh, rpw, srpw):
    """A batch whose jobs share a rays-per-wave setting runs its primary and shadow marches with
    that many lanes per ray (k_primary_b / k_shadow_b); it renders what the same jobs render
    one after another with one lane per ray."""
    lib = capi.load()
    lib.er_set_list_capacity(64)
    scs = [scenes.config_small(scale=0.25 + 0.125 * k, seed=k + 1) for k in range(3)]
    seq = []
    for sc in scs:
        r = er.Renderer(sc, seed=5)
        r.render_passes(3)
        seq.append((r.running().copy(), r.get_estimate().copy(), r.get_seeds()))
        r.close()
    rs = []
    for sc in scs:
        r = er.Renderer(sc, seed=5)
        t = r.get_tune()
        r.set_tune(t["brick"], t["tile_w"], t["blocks_per_cu"], t["shadow_waves_per_cu"], 0, rpw, srpw)
        rs.append(r)
    print(f"[tune] batch rays per wave {rpw} shadow {srpw}", flush=True)
    er.render_batch(rs, [3, 3, 3])
    for r, (run, disp, (s1, s2)) in zip(rs, seq):
        assert np.array_equal(r.running(), run)
        assert np.array_equal(r.get_estimate(), disp)
        g1, g2 = r.get_seeds()
        assert np.array_equal(g1, s1) and np.array_equal(g2, s2)
        r.close()
    lib.er_set_list_capacity(32)


def test_batch_with_tuned_jobs_matches_sequential(fresh):
    """Jobs of one batch with different tiles and bricks (the batch kernels read each job's
    configuration) render what the same jobs render one after another."""
    lib = capi.load()
    lib.er_set_list_capacity(64)
    tunes = [(2, 4, 7), (16, 32, 7), (4, 16, 7)]
    scs = [scenes.config_small(scale=0.25 + 0.125 * k, seed=k + 1) for k in range(3)]
    seq = []
    for sc, tune in zip(scs, tunes):
        r = er.Renderer(sc, seed=5)
        r.set_tune(*tune)
        r.render_passes(3)
        seq.append((r.running().copy(), r.get_estimate().copy()))
        r.close()
    rs = []
    for sc, tune in zip(scs, tunes):
        r = er.Renderer(sc, seed=5)
        r.set_tune(*tune)
        rs.append(r)
    er.render_batch(rs, [3, 3, 3])
    for r, (run, disp) in zip(rs, seq):
        assert np.array_equal(r.running(), run)
        assert np.array_equal(r.get_estimate(), disp)
        r.close()
    lib.er_set_list_capacity(32)


def test_autotune_batch_restores_and_matches_sequential(fresh):
    lib = capi.load()
    lib.er_set_list_capacity(64)
    scs = scenes.config_batch(n_jobs=4, scale=0.1, seed=6)
    passes = [min(sc.passes, 6) for sc in scs]
    seq = []
    for sc, n in zip(scs, passes):
        r = er.Renderer(sc, seed=8)
        r.render_passes(n)
        seq.append((r.running().copy(), r.get_estimate().copy(), r.get_seeds()))
        r.close()
    rs = [er.Renderer(sc, seed=8) for sc in scs]
    t = er.autotune_batch(rs, rounds=2)
    print(f"[tune] batch {t}", flush=True)
    assert t["trials"] >= 10 and t["tile_w"] * t["tile_h"] == 64
    for r in rs:
        assert r.no_iterations() == 0
        assert r.get_tune()["tile_w"] == t["tile_w"] and r.get_tune()["blocks_per_cu"] == t["blocks_per_cu"]
    er.render_batch(rs, passes)
    for r, (run, disp, (s1, s2)) in zip(rs, seq):
        assert np.array_equal(r.running(), run)
        assert np.array_equal(r.get_estimate(), disp)
        g1, g2 = r.get_seeds()
        assert np.array_equal(g1, s1) and np.array_equal(g2, s2)
        r.close()
    lib.er_set_list_capacity(32)


def test_timed_calls_of_different_lengths_match_single_passes(fresh):
    """er_render_passes with timing over calls of 2, 7 and 3 passes (the tracer keeps its
    timing events from call to call, growing them for the longer call): the same frame,
    running estimate, RNG state, display and sample counts as 12 single passes, and every
    call's timing covers its own passes."""
    sc = scenes.config_small(scale=0.5)
    print("[timing] small, calls of 2, 7 and 3 passes", flush=True)
    base = _render(sc, 12)
    r = er.Renderer(sc, seed=9)
    r.set_counting(True)
    r.reset_counters()
    for n in (2, 7, 3):
        t = r.render_passes(n)
        assert t.passes == n
        assert t.total_ms > 0.0 and t.trace_ms > 0.0 and t.post_ms > 0.0 and t.primary_ms > 0.0
        # per pass the primary kernel lies within the pass's trace (events of microsecond resolution)
        assert t.primary_ms <= t.trace_ms + 0.01 * n
    out = (r.frame(), r.running(), r.get_seeds(), r.get_estimate(), r.counters(), r.get_tune())
    assert r.no_iterations() == 12
    r.close()
    _same(base, out, "timed calls")
