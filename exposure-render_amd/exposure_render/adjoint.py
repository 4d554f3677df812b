"""torch.autograd wrapper of the node-table transfer function and its path-replay adjoint.

    image = render_tf(lut, renderer, passes, seeds)   # (H, W, 3) running XYZ estimate
    loss = ((image - target) ** 2).sum(); loss.backward()   # lut.grad via er_adjoint_lut

The forward pass sets the node table (which restarts accumulation), loads the given
per-pixel seeds and renders `passes` progressive passes through the C-ABI; the backward pass
replays the same paths from the same seeds on the GPU (er_adjoint_lut). Tensors live on the
host: the renderer owns its device memory, and the boundary is the C-ABI.
"""
import numpy as np
import torch


class TFRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lut, renderer, passes, s1, s2, range_):
        renderer.set_tf_lut(lut.detach().cpu().numpy().astype(np.float32), range_)
        renderer.set_seeds(s1, s2)
        renderer.render_passes(passes)
        img = renderer.running()[..., :3].copy()
        ctx.renderer, ctx.passes, ctx.s1, ctx.s2 = renderer, passes, s1, s2
        ctx.lut_dtype = lut.dtype
        return torch.from_numpy(img)

    @staticmethod
    def backward(ctx, grad_img):
        g = ctx.renderer.adjoint_lut(ctx.s1, ctx.s2, ctx.passes, grad_img.detach().cpu().numpy())
        return torch.from_numpy(g).to(ctx.lut_dtype), None, None, None, None, None


def render_tf(lut, renderer, passes, seeds, range_=4095.0):
    s1, s2 = seeds
    return TFRender.apply(lut, renderer, passes, s1, s2, range_)
