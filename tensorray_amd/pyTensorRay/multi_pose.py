"""Pose generation for multi-pose rendering (pyTensorRay/multi_pose.py:
40-82 of the reference), without its gin configuration layer."""
import os

import numpy as np

from tensorray_amd import TensorRay as TR
from tensorray_amd.pyTensorRay.utils import image_tensor_to_torch, save_torch_image


def gen_random_rotations(batch, seed):
    """`batch` random rotations (scipy Rotation.random(batch, seed)), the
    first replaced by the identity; rows of 9 row-major entries."""
    from scipy.spatial.transform import Rotation as R
    matrices = R.random(batch, seed).as_matrix()
    matrices[0] = np.eye(3)
    return np.reshape(matrices, (batch, 9))


def gen_z_axis_rotations(batch):
    """Rotations about z by 2 pi k / batch."""
    from scipy.spatial.transform import Rotation as R
    matrices = R.from_euler('z', np.linspace(0, 2 * np.pi, batch, endpoint=False)).as_matrix()
    return np.reshape(matrices, (batch, 9))


def matrices_to_tensor(matrices):
    """Rows of 9 entries -> 4x4 TR.Tensorf rotation matrices."""
    res = []
    for mat in matrices:
        m = [float(x) for x in mat]
        res.append(TR.Tensorf(m[0], m[1], m[2], 0.0, m[3], m[4], m[5], 0.0, m[6], m[7], m[8], 0.0,
                              0.0, 0.0, 0.0, 1.0))
    return res


def rotation_matrices(filename):
    return np.loadtxt(filename)


def render_save_multi_pose(scene, shape_id, rotation_tensors, options, integrator, out_dir):
    """Renders shape `shape_id` in every pose (rotated about its centre) and
    saves sensor_<i>.exr images."""
    obj = scene.shapes[shape_id]
    obj_center = obj.get_obj_center()
    for i, rot in enumerate(rotation_tensors):
        obj.obj_center = obj_center
        obj.rotate_matrix = rot
        scene.configure()
        img_tensor = integrator.renderC(scene, options)
        img = image_tensor_to_torch(img_tensor, scene.get_height(0), scene.get_width(0))
        save_torch_image(os.path.join(out_dir, "sensor_{:d}.exr".format(i)), img)
