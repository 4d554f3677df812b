"""er_render_batch: many tracers per launch sequence, graph-captured per pass round.

Paths are independent and seeded per pixel, so a batch must reproduce running each tracer's
passes on its own bit for bit: frame, running estimate, RNG state, display and pass count.
"""
import numpy as np
import pytest

import exposure_render as er
from exposure_render import capi, scenes
import oracle as O

pytestmark = pytest.mark.gpu


def state(r):
    s1, s2 = r.get_seeds()
    return r.frame(), r.running(), s1, s2, r.get_estimate(), r.no_iterations()


@pytest.mark.parametrize("n_jobs", [1, 7])
def test_batch_matches_sequential(fresh, n_jobs):
    jobs = scenes.config_batch(n_jobs=n_jobs, scale=0.1, seed=3)
    passes = [min(sc.passes, 12) for sc in jobs]
    print(f"[batch] jobs={n_jobs} films={[sc.film for sc in jobs]} passes={passes}", flush=True)
    # two independent sets of freshly bound tracers: the running estimate of a rebound tracer
    # keeps its old content (framebuffer.h:45-49), so both runs start from fresh buffers
    seq = [er.Renderer(sc) for sc in jobs]
    rs = [er.Renderer(sc) for sc in jobs]
    seeds = [O.seeds_for(r.W, r.H, 40 + k) for k, r in enumerate(rs)]
    for r, (s1, s2), n in zip(seq, seeds, passes):
        r.set_seeds(s1, s2)
        r.render_passes(n)
    ref = [state(r) for r in seq]
    for r, (s1, s2) in zip(rs, seeds):
        r.set_seeds(s1, s2)
    er.render_batch(rs, passes)
    for k, r in enumerate(rs):
        got = state(r)
        for name, a, b in zip(("frame", "running", "seeds1", "seeds2", "display"), ref[k][:5], got[:5]):
            assert np.array_equal(a, b), f"job {k}: {name} differs"
        assert got[5] == ref[k][5] == passes[k]
    for r in rs + seq:
        r.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_woodcock_batch_matches_sequential(fresh, mixed):
    """Jobs tracked with Woodcock delta tracking (every job, or every other job beside ray
    marching ones) render in a batch what they render one after another: the batch kernels
    defer the Woodcock shadow marches like the single-tracer ones (on their own RNG streams)
    and then march one lane per ray."""
    jobs = scenes.config_batch(n_jobs=5, scale=0.1, seed=6)
    for k, sc in enumerate(jobs):
        if not mixed or k % 2 == 0:
            sc.tracer.settings.sampler = capi.SAMPLER_WOODCOCK
    passes = [min(sc.passes, 6) for sc in jobs]
    print(f"[batch] woodcock mixed={mixed} films={[sc.film for sc in jobs]} passes={passes}", flush=True)
    seq = [er.Renderer(sc) for sc in jobs]
    rs = [er.Renderer(sc) for sc in jobs]
    seeds = [O.seeds_for(r.W, r.H, 60 + k) for k, r in enumerate(rs)]
    for r, (s1, s2), n in zip(seq, seeds, passes):
        r.set_seeds(s1, s2)
        r.render_passes(n)
    ref = [state(r) for r in seq]
    for r, (s1, s2) in zip(rs, seeds):
        r.set_seeds(s1, s2)
    er.render_batch(rs, passes)
    for k, r in enumerate(rs):
        got = state(r)
        for name, a, b in zip(("frame", "running", "seeds1", "seeds2", "display"), ref[k][:5], got[:5]):
            assert np.array_equal(a, b), f"job {k}: {name} differs"
        assert got[5] == ref[k][5] == passes[k]
    for r in rs + seq:
        r.close()


def test_batch_rejects_duplicates(fresh):
    sc = scenes.config_batch(n_jobs=1, scale=0.1, seed=4)[0]
    r = er.Renderer(sc)
    with pytest.raises(er.ExposureRenderError):
        er.render_batch([r, r], [1, 1])
    r.close()


def test_batch_with_objects_and_bitmaps_matches_sequential(fresh):
    """Jobs with surface objects and bitmap textures (full estimator, where ObjectIDs name the
    objects) render in a batch what they render one after another."""
    lib = capi.load()
    lib.er_set_list_capacity(64)
    jobs = scenes.config_batch(n_jobs=3, scale=0.1, seed=8)
    rng = np.random.default_rng(5)
    for sc in jobs:
        b = sc.add_bitmap(rng.integers(0, 256, size=(8, 8, 4), dtype=np.uint8))
        tex = sc.add_texture(scenes.bitmap_texture(b))
        spc = sc.add_texture(scenes.uniform_texture((0.1, 0.1, 0.1)))
        tm, inv = scenes.look_at((0.0, -0.7, 0.0), (0.0, 1.0, 0.0))
        sc.add_object(scenes.object_desc(capi.SHAPE_PLANE, tm, inv, size=(3.0, 3.0, 0.0), diffuse_tex=tex,
                                         specular_tex=spc, glossiness_tex=spc))
    passes = [min(sc.passes, 6) for sc in jobs]
    print(f"[batch] objects jobs={len(jobs)} films={[sc.film for sc in jobs]} passes={passes}", flush=True)
    seq = [er.Renderer(sc) for sc in jobs]
    rs = [er.Renderer(sc) for sc in jobs]
    seeds = [O.seeds_for(r.W, r.H, 70 + k) for k, r in enumerate(rs)]
    for r, (s1, s2), n in zip(seq, seeds, passes):
        r.set_seeds(s1, s2)
        r.render_passes(n)
    ref = [state(r) for r in seq]
    for r, (s1, s2) in zip(rs, seeds):
        r.set_seeds(s1, s2)
    er.render_batch(rs, passes)
    for k, r in enumerate(rs):
        got = state(r)
        for name, a, b in zip(("frame", "running", "seeds1", "seeds2", "display"), ref[k][:5], got[:5]):
            assert np.array_equal(a, b), f"job {k}: {name} differs"
    for r in rs + seq:
        r.close()
    lib.er_set_list_capacity(32)


def test_full_batch_first_pass_matches_the_oracle(fresh):
    """The 64-job batching workload at full size (volumes of 32-256 cells, films of 128-1024
    pixels, 1-3 lights, phase or hybrid shading): one batch round, then a seeded subset of
    pixels of eight jobs against the CPU oracle's trace from the same seeds - the frame within
    1e-4 relative, the alpha (scatter / no scatter) and the RNG state after the pass exactly."""
    lib = capi.load()
    lib.er_set_list_capacity(4096)
    jobs = scenes.config_batch(n_jobs=64)
    rs = [er.Renderer(sc) for sc in jobs]
    seeds = [er.seeds_for(r.W, r.H, 40 + k) for k, r in enumerate(rs)]
    for r, (s1, s2) in zip(rs, seeds):
        r.set_seeds(s1, s2)
    print(f"[batch] 64 jobs, films {sorted(r.W for r in rs)[::8]}..., one round", flush=True)
    er.render_batch(rs, [1] * len(rs))
    rng = np.random.default_rng(17)
    exact = total = 0
    for k in sorted(int(v) for v in rng.choice(len(rs), size=8, replace=False)):
        r, sc = rs[k], jobs[k]
        orc = O.Oracle(sc, nthreads=8)
        if sc.tracer.settings.shading_type not in (capi.SHADING_BRDF_ONLY, capi.SHADING_PHASE_ONLY):
            assert r.gm_max() == orc.compute_gm_max()
        idx = rng.choice(r.W * r.H, size=256, replace=False)
        px, py = (idx % r.W).astype(np.int32), (idx // r.W).astype(np.int32)
        a1, a2 = seeds[k][0][idx].copy(), seeds[k][1][idx].copy()
        want, _, _ = orc.trace_pixels(px, py, a1, a2)
        got = r.frame()[py, px]
        g1, g2 = r.get_seeds()
        print(f"[batch] job {k}: film {r.W}x{r.H} vol {sc.voxels.shape} shading {sc.tracer.settings.shading_type} "
              f"events {int(want[:, 3].sum())}/256", flush=True)
        assert np.array_equal(g1[idx], a1) and np.array_equal(g2[idx], a2), f"job {k}: RNG state"
        assert np.array_equal(got[:, 3], want[:, 3]), f"job {k}: alpha"
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-6, err_msg=f"job {k}")
        exact += int(np.sum(got == want))
        total += got.size
    print(f"[batch] exact-match fraction {exact / total:.6f}", flush=True)
    for r in rs:
        r.close()
    lib.er_set_list_capacity(32)


def test_cached_batch_graph_reproduces_the_batch(fresh):
    """A second er_render_batch of the same jobs reuses the instantiated graph (only the job
    tables and parameter blocks are uploaded again) and renders the same bits; other pass
    counts need another graph and still match sequential rendering."""
    jobs = scenes.config_batch(n_jobs=5, scale=0.1, seed=9)
    passes = [min(sc.passes, 9) for sc in jobs]
    rs = [er.Renderer(sc) for sc in jobs]
    seeds = [er.seeds_for(r.W, r.H, 90 + k) for k, r in enumerate(rs)]

    def run(npasses):
        for r, (s1, s2) in zip(rs, seeds):
            r.rebind()
            r.set_seeds(s1, s2)
        er.render_batch(rs, npasses)
        return [state(r) for r in rs]

    # (a rebind keeps the running estimate's content, framebuffer.h:45-49, so the running
    # estimates of the two runs differ in rounding; the last frame and the RNG states do not)
    first = run(passes)
    second = run(passes)
    for k in range(len(rs)):
        for i, name in ((0, "frame"), (2, "seeds1"), (3, "seeds2")):
            assert np.array_equal(first[k][i], second[k][i]), f"job {k}: {name} differs on the cached graph"
        assert first[k][5] == second[k][5] == passes[k]
    other = [max(1, p - 3) for p in passes]
    got = run(other)
    for k, (sc, (s1, s2)) in enumerate(zip(jobs, seeds)):
        ref = er.Renderer(sc)
        ref.set_seeds(s1, s2)
        ref.render_passes(other[k])
        want = state(ref)
        ref.close()
        assert np.array_equal(want[0], got[k][0]) and np.array_equal(want[2], got[k][2]), f"job {k}"
    for r in rs:
        r.close()
