"""Per-pose autograd Function for multi-pose optimization
(pyTensorRay/renderer_multi_pose.py:40-133 of the reference): each call
rotates shape `shape_id` about its centre into one of the context's poses,
configures the scene (a refit when only positions change) and renders; its
backward runs the interior and edge adjoints of that pose with the same
seed."""
import torch

from tensorray_amd import TensorRay as TR
from tensorray_amd.pyTensorRay.utils import image_tensor_to_torch, image_torch_to_tensor


def _set_pose(context, pose_id):
    scene = context["scene"]
    shape = scene.shapes[context["shape_id"]]
    shape.rotate_matrix = context["rotation_tensors"][pose_id]
    shape.obj_center = context["obj_center"]
    scene.configure()


class RenderFunctionMultiPose(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, pose_id, context, params):
        ctx.context = context
        context.update(params)
        ctx.pose_id = pose_id
        _set_pose(context, pose_id)
        scene = context["scene"]
        height, width = scene.cameras[0].height, scene.cameras[0].width
        TR.set_rnd_seed(context["seed"])
        img = context["integrator"].renderC(scene, context["options"])
        return image_tensor_to_torch(img, height, width).clone().detach()

    @staticmethod
    def backward(ctx, grad_out):
        c = ctx.context
        _set_pose(c, ctx.pose_id)
        scene = c["scene"]
        height, width = scene.cameras[0].height, scene.cameras[0].width
        dLdI = image_torch_to_tensor(grad_out, height, width)
        TR.set_rnd_seed(c["seed"])
        c["integrator"].renderD(scene, c["options"], dLdI)
        for e in c["edge_integrators"]:
            e.renderD(scene, c["options"], dLdI)
        grad_vertex = c["scene_param_manager"].get_grads()[0]
        grad_vertex[grad_vertex.isnan()] = 0.0
        return grad_vertex, None, None, None


class RenderMultiPose(torch.nn.Module):
    def __init__(self, scene, shape_id, rotation_tensors, obj_center, scene_param_manager, integrator,
                 edge_integrators, options, guiding_options):
        super().__init__()
        self.context = dict(scene=scene, shape_id=shape_id, rotation_tensors=rotation_tensors, obj_center=obj_center,
                            scene_param_manager=scene_param_manager, integrator=integrator,
                            edge_integrators=edge_integrators, options=options, guiding_options=guiding_options,
                            seed=1234)

    def set_state(self, state):
        self.context.update(state)

    def forward(self, V, pose_id, params={}):
        return RenderFunctionMultiPose.apply(V, pose_id, self.context, params)
