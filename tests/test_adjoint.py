"""The path-replay adjoint of the node-table transfer function (er_adjoint_lut).

The adjoint differentiates the smooth form of the estimator (expected shadow transmittance,
sampled free flight weighted by its probability ratio) with the forward pass's decisions.
er_adjoint_value evaluates that smooth estimator with another node table while keeping the
decisions, so central differences of it in the table must reproduce the adjoint's gradient
(up to O(h^2)), which pins the adjoint to the mathematics it claims. The smooth estimator's
expectation equals the forward estimator's, checked on the image mean. A short Adam run
checks the torch.autograd wrapper end to end.
"""
import numpy as np
import pytest
import torch

import exposure_render as er
from exposure_render import adjoint, capi, scenes
import oracle as O

pytestmark = pytest.mark.gpu

H_FD = 1e-3        # central-difference step in node-table units
FD_RTOL = 0.02     # gradients within 2% of central differences


def positive_lut(seed):
    """A table with every opacity and colour strictly positive, so that no max(0, .) clamp is
    active near it and every sample is differentiable."""
    rng = np.random.default_rng(seed)
    n = capi.TF_LUT_NODES
    x = np.linspace(0.0, 1.0, n)
    lut = np.zeros((n, 4), dtype=np.float32)
    lut[:, 0] = 0.05 + 0.5 * x + 0.1 * np.sin(6.0 * x + rng.uniform(0, 3))
    for c in range(3):
        lut[:, 1 + c] = 0.5 + 0.3 * np.sin(4.0 * x + rng.uniform(0, 6))
    return lut


def setup(scale, seed=3, lut=None):
    sc = scenes.config_adjoint(scale=scale)
    r = er.Renderer(sc)
    lut = positive_lut(seed) if lut is None else lut
    r.set_tf_lut(lut, sc.tf_range)
    s1, s2 = O.seeds_for(r.W, r.H, 100 + seed)
    return sc, r, lut, s1, s2


def test_adjoint_matches_central_differences(fresh):
    sc, r, lut, s1, s2 = setup(scale=0.25)
    passes = 4
    rng = np.random.default_rng(11)
    A = rng.standard_normal((r.H, r.W, 3)).astype(np.float32)
    print(f"[adjoint-fd] film={r.W}x{r.H} vol={sc.voxels.shape} passes={passes}", flush=True)
    g = r.adjoint_lut(s1, s2, passes, A)
    assert np.all(np.isfinite(g))

    def L(table):
        img = r.adjoint_value(s1, s2, passes, table)
        return float(np.sum(A.astype(np.float64) * img.astype(np.float64)))

    mag = np.abs(g)
    cand = np.argwhere(mag > 1e-2 * mag.max())
    pick = cand[rng.choice(len(cand), size=min(16, len(cand)), replace=False)]
    worst = 0.0
    for node, ch in pick:
        tp = lut.copy(); tp[node, ch] += H_FD
        tm = lut.copy(); tm[node, ch] -= H_FD
        fd = (L(tp) - L(tm)) / (2.0 * H_FD)
        rel = abs(fd - g[node, ch]) / max(abs(fd), 1e-12)
        worst = max(worst, rel)
        print(f"[adjoint-fd] node {node:3d} ch {ch}: adjoint {g[node, ch]: .6e}  fd {fd: .6e}  rel {rel:.2e}", flush=True)
        assert rel < FD_RTOL, (node, ch, g[node, ch], fd)
    print(f"[adjoint-fd] worst relative difference {worst:.2e}", flush=True)
    r.close()


def test_smooth_estimator_has_the_forward_expectation(fresh):
    sc, r, lut, s1, s2 = setup(scale=0.25, seed=5)
    passes = 48
    r.set_seeds(s1, s2)
    r.render_passes(passes)
    fwd = r.running()[..., :3].astype(np.float64)
    smooth = r.adjoint_value(s1, s2, passes, lut).astype(np.float64)
    m_f, m_s = fwd.mean(), smooth.mean()
    print(f"[adjoint-mean] forward {m_f:.6f} smooth {m_s:.6f}", flush=True)
    assert m_f > 0.0
    assert abs(m_s - m_f) / m_f < 0.02
    r.close()


def test_adjoint_leaves_tracer_state(fresh):
    sc, r, lut, s1, s2 = setup(scale=0.125)
    r.set_seeds(s1, s2)
    r.render_passes(2)
    before = (r.running().copy(), r.get_seeds(), r.no_iterations())
    r.adjoint_lut(s1, s2, 2, np.ones((r.H, r.W, 3), np.float32))
    after = (r.running(), r.get_seeds(), r.no_iterations())
    assert np.array_equal(before[0], after[0])
    assert np.array_equal(before[1][0], after[1][0]) and np.array_equal(before[1][1], after[1][1])
    assert before[2] == after[2]
    r.close()


def test_adjoint_rejects_unsupported_modes(fresh):
    sc = scenes.config_adjoint(scale=0.125)
    sc.tracer.settings.shading_type = capi.SHADING_HYBRID
    r = er.Renderer(sc)
    s1, s2 = O.seeds_for(r.W, r.H, 1)
    with pytest.raises(er.ExposureRenderError):
        r.adjoint_lut(s1, s2, 1, np.zeros((r.H, r.W, 3), np.float32))   # no node table set
    r.set_tf_lut(positive_lut(1), sc.tf_range)
    with pytest.raises(er.ExposureRenderError):
        r.adjoint_lut(s1, s2, 1, np.zeros((r.H, r.W, 3), np.float32))   # hybrid shading
    r.close()


def test_autograd_adam_reduces_loss(fresh):
    sc = scenes.config_adjoint(scale=0.125)
    r = er.Renderer(sc)
    passes = 8
    target_lut = scenes.random_tf_lut(21)
    t1, t2 = O.seeds_for(r.W, r.H, 999)
    with torch.no_grad():
        target = adjoint.render_tf(torch.from_numpy(target_lut), r, 32, (t1, t2), sc.tf_range).clone()
    lut = torch.tensor(scenes.initial_tf_lut(), requires_grad=True)
    opt = torch.optim.Adam([lut], lr=0.02)
    losses = []
    for step in range(25):
        s1, s2 = O.seeds_for(r.W, r.H, step)
        opt.zero_grad()
        img = adjoint.render_tf(lut, r, passes, (s1, s2), sc.tf_range)
        loss = ((img.double() - target.double()) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            lut.clamp_(min=0.0)
        losses.append(float(loss.detach()))
    print(f"[adjoint-adam] loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
    r.close()
