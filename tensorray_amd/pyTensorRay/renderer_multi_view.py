"""Per-view autograd Function for multi-view optimization
(pyTensorRay/renderer_multi_view.py:41-105 of the reference): each call
switches scene.cameras[0] to one of the context's cameras, renders it, and
its backward runs the interior and edge adjoints of that view with the
same seed.  For all views of a step in one launch use
tensorray_amd.multiview.MultiView instead."""
import torch

from tensorray_amd import TensorRay as TR
from tensorray_amd.pyTensorRay.utils import image_tensor_to_torch, image_torch_to_tensor


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, camera_id, context, params):
        ctx.context = context
        context.update(params)
        scene = context["scene"]
        ctx.camera_id = camera_id
        scene.cameras[0].update(context["cameras"][camera_id])
        height, width = scene.cameras[0].height, scene.cameras[0].width
        TR.set_rnd_seed(context["seed"])
        img = context["integrator"].renderC(scene, context["options"])
        return image_tensor_to_torch(img, height, width).clone().detach()

    @staticmethod
    def backward(ctx, grad_out):
        c = ctx.context
        scene = c["scene"]
        scene.cameras[0].update(c["cameras"][ctx.camera_id])
        height, width = scene.cameras[0].height, scene.cameras[0].width
        dLdI = image_torch_to_tensor(grad_out, height, width)
        TR.set_rnd_seed(c["seed"])
        c["integrator"].renderD(scene, c["options"], dLdI)
        for e in c["edge_integrators"]:
            e.renderD(scene, c["options"], dLdI)
        grad_vertex = c["scene_param_manager"].get_grads()[0]
        grad_vertex[grad_vertex.isnan()] = 0.0
        return grad_vertex, None, None, None


class Render(torch.nn.Module):
    def __init__(self, scene, cameras, scene_param_manager, integrator, edge_integrators, options, guiding_options):
        super().__init__()
        self.context = dict(scene=scene, cameras=cameras, scene_param_manager=scene_param_manager,
                            integrator=integrator, edge_integrators=edge_integrators, options=options,
                            guiding_options=guiding_options, seed=1234)

    def set_state(self, state):
        self.context.update(state)

    def forward(self, V, camera_id, params={}):
        return RenderFunction.apply(V, camera_id, self.context, params)
