"""The path-replay adjoint at the differentiation workload's full size: 128^3 blob volume,
512 x 512 film, 32 passes, 256 x 4 node table.

* Central differences (h = 1e-3) of the smooth estimator on 16 random table entries must
  match the adjoint within 2% (the adjoint is that estimator's exact derivative for fixed seeds).
* A 4096-spp reference gradient: the mean of the adjoint over 128 independent 32-pass seed
  sets estimates d E[L] / d(table) for the linear loss L = <A, running estimate>. Forward
  renders (the plain tracer, discrete decisions) at table +- h with common seeds estimate the
  same derivative independently of the smooth construction; the two must agree within a
  central-limit bound of their standard errors plus the central difference's O(h^2) term, and
  the single 32-pass adjoint must lie within its own bound of the reference.
* 200 Adam steps of node-table recovery (L2 loss against a target rendered from a hidden
  table, summed in FP64) through the torch.autograd wrapper reduce the loss and the table error.
"""
import numpy as np
import pytest
import torch

import exposure_render as er
from exposure_render import adjoint, capi, scenes

pytestmark = pytest.mark.gpu

H_FD = 1e-3
FD_RTOL = 0.02


def positive_lut(seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, 1.0, capi.TF_LUT_NODES)
    lut = np.zeros((capi.TF_LUT_NODES, 4), dtype=np.float32)
    lut[:, 0] = 0.05 + 0.5 * x + 0.1 * np.sin(6.0 * x + rng.uniform(0, 3))
    for c in range(3):
        lut[:, 1 + c] = 0.5 + 0.3 * np.sin(4.0 * x + rng.uniform(0, 6))
    return lut


def full_size(seed):
    sc = scenes.config_adjoint()
    r = er.Renderer(sc)
    lut = positive_lut(seed)
    r.set_tf_lut(lut, sc.tf_range)
    print(f"[adjoint-full] film {r.W}x{r.H} volume {sc.voxels.shape} passes {sc.passes}", flush=True)
    return sc, r, lut


def test_full_size_adjoint_matches_central_differences(fresh):
    sc, r, lut = full_size(3)
    passes = sc.passes
    s1, s2 = er.seeds_for(r.W, r.H, 103)
    rng = np.random.default_rng(11)
    A = rng.standard_normal((r.H, r.W, 3)).astype(np.float32)
    g = r.adjoint_lut(s1, s2, passes, A)
    assert np.all(np.isfinite(g))

    def L(table):
        img = r.adjoint_value(s1, s2, passes, table)
        return float(np.sum(A.astype(np.float64) * img.astype(np.float64)))

    mag = np.abs(g)
    cand = np.argwhere(mag > 1e-2 * mag.max())
    pick = cand[rng.choice(len(cand), size=min(16, len(cand)), replace=False)]
    assert len(pick) == 16
    worst = 0.0
    for node, ch in pick:
        tp = lut.copy(); tp[node, ch] += H_FD
        tm = lut.copy(); tm[node, ch] -= H_FD
        fd = (L(tp) - L(tm)) / (2.0 * H_FD)
        rel = abs(fd - g[node, ch]) / max(abs(fd), 1e-12)
        worst = max(worst, rel)
        print(f"[adjoint-full] node {node:3d} ch {ch}: adjoint {g[node, ch]: .6e}  fd {fd: .6e}  rel {rel:.2e}", flush=True)
        assert rel < FD_RTOL, (node, ch, g[node, ch], fd)
    print(f"[adjoint-full] worst relative difference {worst:.2e}", flush=True)
    r.close()


def test_full_size_adjoint_agrees_with_a_high_spp_reference(fresh):
    sc, r, lut = full_size(5)
    passes, sets = sc.passes, 128
    A = np.random.default_rng(12).standard_normal((r.H, r.W, 3)).astype(np.float32)
    A64 = A.astype(np.float64)
    gs = []
    for k in range(sets):
        s1, s2 = er.seeds_for(r.W, r.H, 5000 + k)
        gs.append(r.adjoint_lut(s1, s2, passes, A))
    gs = np.array(gs)
    g_ref = gs.mean(axis=0)
    se_ref = gs.std(axis=0, ddof=1) / np.sqrt(sets)
    print(f"[adjoint-ref] {sets} x {passes} spp: max |g| {np.abs(g_ref).max():.4e}", flush=True)
    # a single 32-pass gradient lies within its own spread of the reference
    z1 = np.abs(gs[0] - g_ref) / np.maximum(gs.std(axis=0, ddof=1), 1e-30)
    assert np.mean(z1 < 3.0) > 0.95
    # forward finite differences with common seeds on the largest entries
    h = 0.02
    order = np.argsort(-np.abs(g_ref).ravel())
    checked = 0
    for flat in order[:4]:
        node, ch = divmod(int(flat), 4)
        diffs = []
        for k in range(sets):
            s1, s2 = er.seeds_for(r.W, r.H, 5000 + k)
            vals = []
            for sign in (1.0, -1.0):
                t = lut.copy()
                t[node, ch] += sign * h
                r.set_tf_lut(t, sc.tf_range)
                r.set_seeds(s1, s2)
                r.render_passes(passes)
                vals.append(float(np.sum(A64 * r.running()[..., :3].astype(np.float64))))
            diffs.append((vals[0] - vals[1]) / (2.0 * h))
        r.set_tf_lut(lut, sc.tf_range)
        diffs = np.array(diffs)
        fd, se_fd = diffs.mean(), diffs.std(ddof=1) / np.sqrt(sets)
        bound = 4.0 * np.hypot(se_ref[node, ch], se_fd) + 0.02 * abs(g_ref[node, ch])
        print(f"[adjoint-ref] node {node:3d} ch {ch}: reference {g_ref[node, ch]: .5e} (se {se_ref[node, ch]:.2e})  "
              f"forward fd {fd: .5e} (se {se_fd:.2e})  bound {bound:.2e}", flush=True)
        assert abs(fd - g_ref[node, ch]) <= bound
        checked += 1
    assert checked == 4
    r.close()


def test_full_size_adam_recovers_the_table(fresh):
    sc = scenes.config_adjoint()
    r = er.Renderer(sc)
    passes = sc.passes
    target_lut = scenes.random_tf_lut(21)
    t1, t2 = er.seeds_for(r.W, r.H, 999)
    with torch.no_grad():
        target = adjoint.render_tf(torch.from_numpy(target_lut), r, 4 * passes, (t1, t2), sc.tf_range).clone()
    lut = torch.tensor(scenes.initial_tf_lut(), requires_grad=True)
    lut0 = lut.detach().clone()
    opt = torch.optim.Adam([lut], lr=0.02)
    losses = []
    for step in range(200):
        s1, s2 = er.seeds_for(r.W, r.H, step)
        opt.zero_grad()
        img = adjoint.render_tf(lut, r, passes, (s1, s2), sc.tf_range)
        loss = ((img.double() - target.double()) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            lut.clamp_(min=0.0)
        losses.append(float(loss.detach()))
        if step % 20 == 0:
            print(f"[adam-full] step {step:3d} loss {losses[-1]:.5f}", flush=True)
    # the table error over the intensities the volume holds (nodes below its maximum)
    vmax = int(sc.voxels.max())
    n_used = min(int(np.ceil(vmax * (capi.TF_LUT_NODES - 1) / sc.tf_range)) + 1, capi.TF_LUT_NODES)
    tgt = torch.from_numpy(target_lut)
    err0 = float((lut0[:n_used] - tgt[:n_used]).abs().mean())
    err1 = float((lut.detach()[:n_used] - tgt[:n_used]).abs().mean())
    first, last = np.mean(losses[:5]), np.mean(losses[-10:])
    print(f"[adam-full] loss {first:.5f} -> {last:.5f}; table error {err0:.4f} -> {err1:.4f} over {n_used} nodes",
          flush=True)
    assert last < 0.2 * first
    assert err1 < err0
    r.close()
