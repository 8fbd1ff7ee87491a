"""The `TensorRay` module's host side, without a GPU: scene XML and OBJ
loading (SceneLoader.cpp:176-410 subset), shape queries, the parameter pool
(get_size_param / get_var / set_var in the reference's layouts) and mesh
export."""
import os

import numpy as np
import pytest
import torch

TR = pytest.importorskip("tensorray_amd.TensorRay")
from tensorray_amd.pyTensorRay.common import get_int_ptr, get_ptr  # noqa: E402
from tensorray_amd.pyTensorRay.parameter_manager import SceneParameterManager  # noqa: E402

CUBE = """# unit cube, quads
v -1 -1 -1
v 1 -1 -1
v 1 1 -1
v -1 1 -1
v -1 -1 1
v 1 -1 1
v 1 1 1
v -1 1 1
vt 0 0
vt 1 0
vt 1 1
vt 0 1
f 1/1 4/4 3/3 2/2
f 5/1 6/2 7/3 8/4
f 1/1 2/2 6/3 5/4
f 2/1 3/2 7/3 6/4
f 3/1 4/2 8/3 7/4
f 4/1 1/2 5/3 8/4
"""

QUAD = """v -0.5 0 -0.5
v 0.5 0 -0.5
v 0.5 0 0.5
v -0.5 0 0.5
f -4 -1 -2
f -4 -2 -3
"""

XML = """<?xml version="1.0"?>
<!-- a cube on a light -->
<scene version="0.6.0">
  <bsdf type="diffuse" id="white"><rgb name="reflectance" value="0.5, 0.6, 0.7"/></bsdf>
  <bsdf type="microfacet" id="glossy">
    <float name="roughness" value="0.25"/>
    <spectrum name="diffuseReflectance" value="0.1"/>
    <spectrum name="specularReflectance" value="0.04, 0.05, 0.06"/>
  </bsdf>
  <bsdf type="roughconductor" id="gold">
    <float name="alpha" value="0.3"/>
    <spectrum name="eta" value="0.143, 0.374, 1.44"/>
    <spectrum name="k" value="3.98, 2.38, 1.6"/>
  </bsdf>
  <shape type="obj" id="cube">
    <string name="filename" value="cube.obj"/>
    <ref id="glossy"/>
    <transform name="to_world">
      <scale x="0.5" y="0.5" z="0.5"/>
      <rotate y="1" angle="90"/>
      <translate x="0" y="1" z="0"/>
    </transform>
    <boolean name="face_normals" value="false"/>
  </shape>
  <shape type="obj" id="light">
    <string name="filename" value="quad.obj"/>
    <ref id="white"/>
    <transform name="toWorld"><translate y="3"/></transform>
    <emitter type="area"><rgb name="radiance" value="10, 9, 8"/></emitter>
  </shape>
  <sensor type="perspective">
    <float name="fov" value="40"/>
    <transform name="to_world"><lookat origin="0, 1, 4" target="0, 1, 0" up="0, 1, 0"/></transform>
    <film type="hdrfilm"><integer name="width" value="64"/><integer name="height" value="48"/></film>
  </sensor>
</scene>
"""


@pytest.fixture
def scene_dir(tmp_path):
    (tmp_path / "cube.obj").write_text(CUBE)
    (tmp_path / "quad.obj").write_text(QUAD)
    (tmp_path / "scene.xml").write_text(XML)
    TR.env_release()  # empties the parameter pool
    yield tmp_path
    TR.env_release()


def test_load_scene_xml(scene_dir):
    s = TR.Scene()
    s.load_file(str(scene_dir / "scene.xml"))
    assert len(s.shapes) == 2 and s.num_bsdf == 3 and len(s.emitters) == 1 and len(s.cameras) == 1
    assert s.get_width() == 64 and s.get_height() == 48
    cube, light = s.shapes
    assert cube.get_vertex_count() == 8 and cube.get_face_count() == 12
    # 12 cube edges + 6 face diagonals
    assert cube.get_edge_count() == 18
    assert light.get_face_count() == 2 and light.get_edge_count() == 5
    F = torch.zeros(12, 3, dtype=torch.int32)
    cube.get_face_indices(get_int_ptr(F))
    assert F.min().item() == 0 and F.max().item() == 7
    # fan triangulation of the first quad (1 4 3 2)
    assert F[0].tolist() == [0, 3, 2] and F[1].tolist() == [0, 2, 1]
    E = torch.zeros(2, 18, dtype=torch.int32)
    cube.get_edge_data(get_int_ptr(E))
    assert (E[0] < E[1]).all()  # v0 < v1
    c = cube.get_obj_center()
    out = torch.zeros(3)
    c.value_to_torch(get_ptr(out))
    assert torch.allclose(out, torch.zeros(3))


def test_unsupported_elements_raise(tmp_path):
    (tmp_path / "bad.xml").write_text('<scene><bsdf type="plastic" id="p"/></scene>')
    with pytest.raises(RuntimeError, match="Unsupported BSDF"):
        TR.Scene().load_file(str(tmp_path / "bad.xml"))
    (tmp_path / "bad2.xml").write_text('<scene><shape type="sphere"/></scene>')
    with pytest.raises(RuntimeError, match="Unsupported shape"):
        TR.Scene().load_file(str(tmp_path / "bad2.xml"))
    (tmp_path / "bad3.xml").write_text('<scene><bsdf type="diffuse" id="a"></scene>')
    with pytest.raises(RuntimeError, match="XML parsing failed"):
        TR.Scene().load_file(str(tmp_path / "bad3.xml"))


def test_parameter_pool_layout(scene_dir):
    s = TR.Scene()
    s.load_file(str(scene_dir / "scene.xml"))
    cube = s.shapes[0]
    cube.diff_all_vertex_pos()
    s.bsdfs[1].diff_texture("roughness")
    s.shapes[1].diff_translation()
    s.bsdfs[2].diff_texture("eta")
    with pytest.raises(RuntimeError, match="not supported"):
        s.bsdfs[0].diff_texture("roughness")
    sizes = []
    assert TR.get_size_param(sizes) == 24 + 1 + 3 + 3
    assert sizes == [24, 1, 3, 3] and TR.get_num_param() == 4
    pm = SceneParameterManager([0, 1])
    assert pm.param_counts == [24, 7]
    pm.get_values()
    V = pm.params[0].detach().clone().view(3, -1)  # [3][V] planes, as the reference's Vec3 tensors
    assert V.shape == (3, 8)
    assert torch.equal(V[:, 0], torch.tensor([-1.0, -1.0, -1.0]))
    assert pm.params[1].detach().tolist() == pytest.approx([0.25, 0, 0, 0, 0.143, 0.374, 1.44], rel=1e-6)
    with torch.no_grad():
        pm.params[0].view(3, -1)[1, :] += 2.0
        pm.params[1][0] = 0.5
    pm.set_values()
    pm2 = SceneParameterManager([0])
    pm2.get_values()
    assert torch.equal(pm2.params[0][:24].view(3, -1)[1], V[1] + 2.0)
    assert pm2.params[0][24].item() == 0.5
    g = pm2.get_grads()[0]  # nothing rendered yet
    assert g.shape == (31,) and (g == 0).all()


def test_export_mesh_roundtrip(scene_dir):
    s = TR.Scene()
    s.load_file(str(scene_dir / "scene.xml"))
    out = scene_dir / "out.obj"
    s.shapes[0].export_mesh(str(out))
    s2 = TR.Scene()
    (scene_dir / "s2.xml").write_text(
        '<scene><bsdf type="diffuse" id="w"><rgb name="reflectance" value="0.5"/></bsdf>'
        '<shape type="obj"><string name="filename" value="out.obj"/><ref id="w"/></shape></scene>')
    s2.load_file(str(scene_dir / "s2.xml"))
    assert s2.shapes[0].get_face_count() == 12 and s2.shapes[0].get_vertex_count() == 8
    F1 = torch.zeros(12, 3, dtype=torch.int32)
    F2 = torch.zeros(12, 3, dtype=torch.int32)
    s.shapes[0].get_face_indices(get_int_ptr(F1))
    s2.shapes[0].get_face_indices(get_int_ptr(F2))
    assert torch.equal(F1, F2)


def test_render_options_defaults():
    o = TR.RenderOptions(1234, 3, 16, 4, 4, 2, 1)
    assert (o.seed, o.max_bounces, o.spp, o.sppe, o.sppse0, o.sppse1, o.sppe0) == (1234, 3, 16, 4, 4, 2, 1)
    assert o.spp_batch == 16 and o.sppe_batch == 4
    assert o.g_direct is False and o.g_direct_max_size == 100000 and o.g_direct_spp == 16
    assert o.checkpoint is False


def test_multi_view_cameras():
    """Camera generation of the multi-view helpers (multi_view.py:41-90 of
    the reference) and Camera.update, which copies another camera's pose,
    resolution and fov into the scene's sensor."""
    from tensorray_amd.pyTensorRay import multi_view
    a = multi_view.gen_camera_positions(12, 3.0, False)
    b = multi_view.gen_camera_positions(12, 3.0, False)
    assert a.shape == (12, 3) and np.array_equal(a, b)          # seeded by the batch size
    np.testing.assert_allclose(np.linalg.norm(a, axis=1), 3.0, rtol=1e-12)
    np.testing.assert_allclose(a[:, 1], 3.0 * np.linspace(0.99, -0.99, 12), rtol=1e-12)
    h = multi_view.gen_camera_positions(8, 2.0, True)
    assert (h[:, 2] > 0).all()
    cams = multi_view.gen_cameras(a, [0, 0, 0], [0, 1, 0], 45.0, [64, 48])
    assert len(cams) == 12 and (cams[0].width, cams[0].height) == (64, 48)
    s = TR.Camera(TR.Vector3f(0, 0, 5), TR.Vector3f(0, 0, 0), TR.Vector3f(0, 1, 0), 8, 8, 30.0)
    s.update(cams[3])
    assert (s.width, s.height) == (64, 48)
    p = torch.zeros(3)
    s.posTensor.value_to_torch(get_ptr(p))
    np.testing.assert_allclose(p.numpy(), a[3], rtol=1e-6)


def test_exr_writer(tmp_path):
    """OpenEXR output as the reference's SaveEXR stores it (B, G, R halves,
    uncompressed scanlines): fixed header bytes from the file-format
    specification, one offset per scanline pointing at that scanline's y,
    and pixel values equal to their IEEE half rounding."""
    import struct
    from tensorray_amd.pyTensorRay import utils
    rng = np.random.default_rng(3)
    img = (rng.uniform(0, 4, (5, 7, 3)) * [1.0, 0.5, 0.25]).astype(np.float32)
    f = str(tmp_path / "a.exr")
    utils.save_torch_image(f, torch.from_numpy(img))
    data = open(f, "rb").read()
    assert data[:8] == bytes([0x76, 0x2F, 0x31, 0x01, 2, 0, 0, 0])
    chl = data.index(b"channels\0chlist\0")
    (size,) = struct.unpack_from("<i", data, chl + 16)
    assert size == 3 * (2 + 16) + 1
    assert data[chl + 20:chl + 22] == b"B\0" and struct.unpack_from("<i", data, chl + 22)[0] == 1  # HALF
    assert b"compression\0compression\0\x01\0\0\0\0" in data
    hdr_end = data.index(b"screenWindowWidth\0float\0") + len(b"screenWindowWidth\0float\0") + 8 + 1
    offs = struct.unpack_from("<5Q", data, hdr_end)
    for y, off in enumerate(offs):
        assert struct.unpack_from("<ii", data, off) == (y, 3 * 2 * 7)
    assert len(data) == offs[-1] + 8 + 3 * 2 * 7
    back = utils.load_exr(f)
    assert back.shape == (5, 7, 3)
    assert np.array_equal(back, img.astype(np.float16).astype(np.float32))


def test_multi_pose_helpers(scene_dir):
    """Multi-pose helpers (multi_pose.py:40-82 of the reference): the first
    random pose is the identity, every pose is a proper rotation, z-axis
    poses step by 2 pi / n, and the 16-value Tensorf holds the 4x4 matrix."""
    from tensorray_amd.pyTensorRay import multi_pose
    rots = multi_pose.gen_z_axis_rotations(4)
    assert np.allclose(rots[0], np.eye(3).ravel())
    np.testing.assert_allclose(rots[1].reshape(3, 3), [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-12)
    tensors = multi_pose.matrices_to_tensor(rots)
    assert len(tensors) == 4 and tensors[1].size() == 16
    r = multi_pose.gen_random_rotations(5, 3)
    assert np.allclose(r[0], np.eye(3).ravel())
    for m in r:
        M = m.reshape(3, 3)
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-12) and np.isclose(np.linalg.det(M), 1.0)
    s = TR.Scene()
    s.load_file(str(scene_dir / "scene.xml"))
    c = s.shapes[0].get_obj_center()
    cv = torch.zeros(3)
    c.value_to_torch(get_ptr(cv))
    assert np.allclose(cv.numpy(), 0.0, atol=1e-6)   # the unit cube is centred at the origin


def test_pytensorray_helpers(scene_dir):
    """Image readers and the uniform Laplacian of pyTensorRay.utils, the
    option constructors of pyTensorRay.fwd."""
    from tensorray_amd.pyTensorRay import fwd, utils
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 2, (4, 6, 3)).astype(np.float32)
    utils.save_torch_image(str(scene_dir / "i_2.pfm"), torch.from_numpy(img))
    utils.save_torch_image(str(scene_dir / "i_10.pfm"), torch.from_numpy(2 * img))
    assert torch.equal(utils.load_torch_image(str(scene_dir / "i_2.pfm")), torch.from_numpy(img))
    imgs = utils.read_images(str(scene_dir), ".pfm")
    assert len(imgs) == 2 and torch.equal(imgs[1], torch.from_numpy(2 * img))  # numeric order: 2 before 10
    edges = torch.tensor([[0, 1, 2, 0], [1, 2, 0, 3]])
    L = utils.uniform_laplacian(4, edges).to_dense()
    np.testing.assert_array_equal(L.numpy(), [[3, -1, -1, -1], [-1, 2, -1, 0], [-1, -1, 2, 0], [-1, 0, 0, 1]])
    o = fwd.RenderOptions(7, 2, 4, 1, 1, 0, 0, g_direct=True, checkpoint=True)
    b = fwd.RenderBatchOptions(spp_batch=2)
    utils.update_render_batch_options(o, b)
    g = fwd.GuidingOptions(g_direct_spp=8)
    utils.update_guiding_options(o, g)
    assert o.spp_batch == 2 and o.g_direct_spp == 8 and o.g_direct is False and o.checkpoint is True
    s = fwd.Scene(str(scene_dir / "scene.xml"))
    assert len(s.shapes) == 2


def test_scalar_parameter_and_linked_translation(scene_dir):
    """Tensorf(v, True) registers a scalar parameter (Tensor.h:1675-1685) and
    `Tensorf(x, y, z) * param` stays linked to it, the way the reference's
    validation scripts move a shape (validate_gpu.py:69-80)."""
    x = TR.Tensorf(0.5, True)
    assert TR.get_num_param() == 1
    sizes = []
    assert TR.get_size_param(sizes) == 1 and sizes == [1]
    t = TR.Tensorf(0.0, 0.0, 2.0) * x
    out = torch.zeros(3)
    t.value_to_torch(get_ptr(out))
    assert out.tolist() == [0.0, 0.0, 1.0]
    val = torch.tensor([-1.5])
    TR.set_var(0, get_ptr(val), 0)
    t.value_to_torch(get_ptr(out))
    assert out.tolist() == [0.0, 0.0, -3.0]   # the link follows the parameter
    got = torch.zeros(1)
    TR.get_var(0, get_ptr(got), 0)
    assert got.item() == -1.5
    # constants multiply elementwise, a single value broadcasts
    c = TR.Tensorf(1.0, 2.0, 3.0) * TR.Tensorf(2.0, 3.0, 4.0)
    c.value_to_torch(get_ptr(out))
    assert out.tolist() == [2.0, 6.0, 12.0]
    c = 0.5 * TR.Tensorf(2.0, 4.0, 6.0)
    c.value_to_torch(get_ptr(out))
    assert out.tolist() == [1.0, 2.0, 3.0]
    with pytest.raises(RuntimeError, match="Tensorf product"):
        TR.Tensorf(1.0, True) * x
    # a shape's translation may hold the linked product
    s = TR.Scene()
    s.load_file(str(scene_dir / "scene.xml"))
    s.shapes[0].translation = TR.Tensorf(1.0, 0.0, 0.0) * x
    s.shapes[0].translation.value_to_torch(get_ptr(out))
    assert out.tolist() == [-1.5, 0.0, 0.0]
