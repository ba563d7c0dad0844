"""torch.autograd.Function over the HIP renderer and its adjoint, for the
verification harness only: forward = pt_render of the mean image for a
material-parameter tensor, backward = pt_adjoint with the incoming image
gradient. Parameters are (num_materials, 7) in pt_adjoint order."""
import numpy as np
import torch

from glslpt import scenes


class RenderMaterials(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, renderer, scene, base_materials, first_sample, num_samples):
        mats = scenes.apply_params(base_materials, params.detach().cpu().numpy())
        renderer.upload_scene(scene, materials=mats)
        renderer.clear()
        renderer.render(first_sample, num_samples)
        renderer.synchronize()
        img = renderer.read_accum()[..., :3] / np.float32(num_samples)
        ctx.renderer, ctx.first, ctx.n = renderer, first_sample, num_samples
        return torch.from_numpy(np.ascontiguousarray(img))

    @staticmethod
    def backward(ctx, grad_img):
        g = ctx.renderer.adjoint(ctx.first, ctx.n, grad_img.detach().cpu().numpy().astype(np.float32))
        return torch.from_numpy(g), None, None, None, None, None


def clamp_params(p):
    """Keep parameters in their valid ranges (baseColor, metallic, clearcoat
    in [0, 1], roughness in [0.01, 1], ior in [1.01, 3])."""
    with torch.no_grad():
        p[:, 0:3].clamp_(0.0, 1.0)
        p[:, 3].clamp_(0.01, 1.0)
        p[:, 4].clamp_(0.0, 1.0)
        p[:, 5].clamp_(1.01, 3.0)
        p[:, 6].clamp_(0.0, 1.0)
