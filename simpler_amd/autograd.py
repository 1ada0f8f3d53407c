"""torch.autograd binding of the path-replay adjoint.

    image = simpler_amd.autograd.render(params, scene, res, n, seed=1)
    loss = (image * weights).sum(); loss.backward()   # params.grad

`params` is the scene's differentiable field parameter vector
(Scene.params(): spline coefficients, or source amplitudes, phases and the
wavenumber).  The forward pass uploads it and renders on the GPU; the
backward pass runs the adjoint kernel with dloss/dimage as the objective
weights, re-tracing the same photons (same seed and photon range)."""
import torch


class FieldRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, scene, res, n, seed, offset, n_total):
        scene.set_params(params.detach().cpu().double().numpy())
        img, _ = scene.render(n, res, seed=seed, offset=offset, n_total=n_total, stats=False)
        ctx.scene = scene
        ctx.args = (n, seed, offset, n_total)
        ctx.save_for_backward(params)
        return img

    @staticmethod
    def backward(ctx, grad_img):
        (params,) = ctx.saved_tensors
        scene = ctx.scene
        n, seed, offset, n_total = ctx.args
        scene.set_params(params.detach().cpu().double().numpy())
        g, _, _ = scene.adjoint_render(grad_img.contiguous().float(), n, seed=seed, offset=offset, n_total=n_total)
        return g.to(device=params.device, dtype=params.dtype).view_as(params), None, None, None, None, None, None


def render(params, scene, res, n, seed=1, offset=0, n_total=None):
    """Differentiable render of photons [offset, offset + n) of an n_total
    render (float32 CUDA image (Z, Y, X))."""
    return FieldRender.apply(params, scene, res, n, seed, offset, n_total)
