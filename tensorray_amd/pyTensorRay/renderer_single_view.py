"""torch.autograd.Function around the MI355X integrators
(pyTensorRay/renderer_single_view.py:43-115 of the reference).

forward: set the seed, renderC.  backward: the same seed, renderD of the
interior integrator and of every edge integrator (the counter-based RNG
gives the adjoint exactly the forward's paths), then the gradient of
parameter group 0 with NaNs zeroed.  With options.checkpoint = True the
forward keeps its path checkpoint and the interior adjoint re-evaluates the
paths from it instead of re-tracing them.
"""
import torch

from tensorray_amd import TensorRay as TR
from tensorray_amd.pyTensorRay.utils import image_tensor_to_torch, image_torch_to_tensor


class RenderFunctionSingleView(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, context, params):
        ctx.context = context
        context.update(params)
        scene = context["scene"]
        height, width = scene.cameras[0].height, scene.cameras[0].width
        TR.set_rnd_seed(context["seed"])
        img = context["integrator"].renderC(scene, context["options"])
        return image_tensor_to_torch(img, height, width).clone().detach()

    @staticmethod
    def backward(ctx, grad_out):
        c = ctx.context
        scene = c["scene"]
        height, width = scene.cameras[0].height, scene.cameras[0].width
        dLdI = image_torch_to_tensor(grad_out, height, width)
        TR.set_rnd_seed(c["seed"])
        c["integrator"].renderD(scene, c["options"], dLdI)
        for e in c["edge_integrators"]:
            e.renderD(scene, c["options"], dLdI)
        grad_vertex = c["scene_param_manager"].get_grads()[0]
        grad_vertex[grad_vertex.isnan()] = 0.0
        return grad_vertex, None, None


class RenderSingleView(torch.nn.Module):
    def __init__(self, scene, scene_param_manager, integrator, edge_integrators, options, guiding_options=None):
        super().__init__()
        self.context = dict(scene=scene, scene_param_manager=scene_param_manager, integrator=integrator,
                            edge_integrators=edge_integrators, options=options, guiding_options=guiding_options,
                            seed=1234)

    def set_state(self, state):
        self.context.update(state)

    def forward(self, V, params={}):
        return RenderFunctionSingleView.apply(V, self.context, params)
