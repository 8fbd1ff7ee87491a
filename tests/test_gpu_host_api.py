"""The `TensorRay` module end to end on the GPU: renderC / renderD through
the reference-style classes equal the C ABI's results for the same seed,
the autograd Function returns the summed interior + edge gradient of
parameter group 0, and the checkpoint option changes nothing but speed."""
import numpy as np
import pytest
import torch

from tensorray_amd import _capi, scenes

pytestmark = pytest.mark.gpu

TR = pytest.importorskip("tensorray_amd.TensorRay")


def _tr_scene(spec, cam_d, W, H):
    """The scenes.SceneSpec built through the module's programmatic API."""
    return spec.to_module(cam_d, W, H)


@pytest.fixture
def tr_env():
    TR.env_release()
    TR.env_create(0)
    yield
    TR.env_release()


def test_module_matches_c_abi(tr_env, gpu_ctx):
    from tensorray_amd.pyTensorRay.common import get_ptr
    from tensorray_amd.pyTensorRay.parameter_manager import SceneParameterManager
    from tensorray_amd.pyTensorRay.utils import image_tensor_to_torch, image_torch_to_tensor
    spec = scenes.sphere_scene(subdiv=2)
    W = H = 32
    cam_d = scenes.CAMERA_DEFAULT
    s = _tr_scene(spec, cam_d, W, H)
    s.shapes[0].diff_all_vertex_pos()
    s.configure()
    pm = SceneParameterManager([0])
    opts = TR.RenderOptions(1234, 2, 8, 4, 4, 4, 4)
    TR.set_rnd_seed(77)
    img = image_tensor_to_torch(TR.PathTracer().renderC(s, opts), H, W)
    # the same render through the C ABI (seed = the module's rnd seed)
    ds = _capi.DeviceScene(gpu_ctx, spec)
    cam = _capi.make_camera(W, H, **cam_d)
    opt = _capi.make_options(seed=77, max_bounces=2, spp=8, sppe=4, sppse0=4, sppse1=4, sppe0=4)
    film = torch.empty(3, W * H, device="cuda")
    ds.render_fwd(cam, opt, film.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(img, film.cpu().t().reshape(H, W, 3))
    # renderD of the interior + four edge integrators vs the C ABI
    dl = torch.rand(H, W, 3, generator=torch.Generator().manual_seed(0))
    dLdI = image_torch_to_tensor(dl, H, W)
    TR.set_rnd_seed(77)
    TR.PathTracer().renderD(s, opts, dLdI)
    for e in (TR.PrimaryEdgeIntegrator(), TR.PixelBoundaryIntegrator(), TR.DirectEdgeIntegrator(),
              TR.IndirectEdgeIntegrator()):
        e.renderD(s, opts, dLdI)
    g = pm.get_grads()[0].view(3, -1)  # [3][V] planes
    dld = dl.reshape(H * W, 3).t().contiguous().cuda()
    dpos = torch.zeros(ds.info().num_vertices, 3, device="cuda")
    ds.render_adj(cam, opt, dld.data_ptr(), dpos.data_ptr())
    for k in ("primary", "pixel", "direct", "indirect"):
        ds.render_boundary(cam, opt, k, dld.data_ptr(), dpos.data_ptr())
    torch.cuda.synchronize()
    nv = spec.meshes[0].pos.shape[0]
    ref = dpos.cpu()[:nv].t()
    scale = ref.abs().max().item()
    assert scale > 0
    assert (g - ref).abs().max().item() <= 1e-4 * scale
    ds.close()


def test_autograd_function_and_checkpoint(tr_env):
    from tensorray_amd.pyTensorRay.parameter_manager import SceneParameterManager
    from tensorray_amd.pyTensorRay.renderer_single_view import RenderSingleView
    spec = scenes.sphere_scene(subdiv=2)
    W = H = 32
    s = _tr_scene(spec, scenes.CAMERA_DEFAULT, W, H)
    s.shapes[0].diff_all_vertex_pos()
    s.configure()
    pm = SceneParameterManager([0])
    pm.get_values()
    target = torch.full((H, W, 3), 0.2)
    grads, imgs = [], []
    for ck in (False, True):
        opts = TR.RenderOptions(1234, 2, 8, 2, 2, 2, 2)
        opts.checkpoint = ck
        r = RenderSingleView(s, pm, TR.PathTracer(), [TR.PrimaryEdgeIntegrator(), TR.DirectEdgeIntegrator()], opts)
        img = r(pm.params[0], {"seed": 5})
        loss = ((img - target) ** 2).sum()
        pm.params[0].grad = None
        loss.backward()
        grads.append(pm.params[0].grad.clone())
        imgs.append(img.detach())
    assert torch.equal(imgs[0], imgs[1])
    scale = grads[0].abs().max().item()
    assert scale > 0
    assert (grads[0] - grads[1]).abs().max().item() <= 1e-4 * scale
    # an optimizer step moves the mesh through set_values + configure (BVH refit)
    with torch.no_grad():
        pm.params[0] -= 1e-3 * grads[0] / scale
    pm.set_values()
    s.configure()
    img2 = RenderSingleView(s, pm, TR.PathTracer(), [], TR.RenderOptions(1234, 2, 8, 0, 0, 0, 0))(pm.params[0], {"seed": 5})
    assert not torch.equal(img2, imgs[0])


def test_multi_pose_function(tr_env):
    """RenderMultiPose (renderer_multi_pose.py of the reference): the identity
    pose renders the unrotated scene, another pose renders a different
    image, and its backward returns a finite, non-zero vertex gradient."""
    from tensorray_amd.pyTensorRay import multi_pose
    from tensorray_amd.pyTensorRay.parameter_manager import SceneParameterManager
    from tensorray_amd.pyTensorRay.renderer_multi_pose import RenderMultiPose
    spec = scenes.sphere_scene(subdiv=2, amplitude=0.1)
    W = H = 32
    s = _tr_scene(spec, scenes.CAMERA_DEFAULT, W, H)
    s.shapes[0].diff_all_vertex_pos()
    s.configure()
    pm = SceneParameterManager([0])
    pm.get_values()
    opts = TR.RenderOptions(1234, 2, 8, 2, 0, 0, 0)
    plain = RenderSingleViewImage(s, opts)
    rots = multi_pose.matrices_to_tensor(multi_pose.gen_z_axis_rotations(4))
    r = RenderMultiPose(s, 0, rots, s.shapes[0].get_obj_center(), pm, TR.PathTracer(), [TR.PrimaryEdgeIntegrator()],
                        opts, None)
    img0 = r(pm.params[0], 0, {"seed": 1234})
    assert (img0 - plain).abs().mean().item() <= 1e-4 * plain.abs().mean().item()
    img1 = r(pm.params[0], 1, {"seed": 1234})
    assert (img1 - img0).abs().mean().item() > 1e-3 * img0.abs().mean().item()
    pm.params[0].grad = None
    ((img1 - 0.2) ** 2).sum().backward()
    g = pm.params[0].grad
    assert torch.isfinite(g).all() and g.abs().max().item() > 0


def RenderSingleViewImage(scene, opts):
    h, w = scene.cameras[0].height, scene.cameras[0].width
    from tensorray_amd.pyTensorRay.utils import image_tensor_to_torch
    TR.set_rnd_seed(1234)
    return image_tensor_to_torch(TR.PathTracer().renderC(scene, opts), h, w)


LOADED_XML = """<?xml version="1.0"?>
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
    <transform name="to_world"><scale x="0.5" y="0.5" z="0.5"/><translate x="0" y="1" z="0"/></transform>
    <boolean name="face_normals" value="false"/>
  </shape>
  <shape type="obj" id="floor">
    <string name="filename" value="floor.obj"/>
    <ref id="gold"/>
  </shape>
  <shape type="obj" id="light">
    <string name="filename" value="light.obj"/>
    <ref id="white"/>
    <transform name="toWorld"><translate y="3"/></transform>
    <emitter type="area"><rgb name="radiance" value="10, 9, 8"/></emitter>
  </shape>
  <sensor type="perspective">
    <float name="fov" value="40"/>
    <transform name="to_world"><lookat origin="0, 1.5, 4" target="0, 0.75, 0" up="0, 1, 0"/></transform>
    <film type="hdrfilm"><integer name="width" value="64"/><integer name="height" value="48"/></film>
  </sensor>
</scene>
"""

# the emitter faces down (-y): triangles (1 2 3) and (1 3 4)
LIGHT = """v -0.5 0 -0.5
v 0.5 0 -0.5
v 0.5 0 0.5
v -0.5 0 0.5
f 1 2 3
f 1 3 4
"""

FLOOR = """v -3 0 -3
v 3 0 -3
v 3 0 3
v -3 0 3
f 1 4 3 2
"""


def test_loaded_scene_matches_oracle(tr_env, oracle_mod, tmp_path):
    """A scene read from disk (Mitsuba-style XML + OBJ files: quads, a
    to_world scale + translate, smooth shading, the three BSDF classes, an
    area emitter, a perspective sensor) and rendered through the module's
    PathTracer.renderC equals the oracle's film of the same geometry bit for
    bit; the geometry is rebuilt from the OBJ text and the transforms (exact
    in float) and the loader's own face lists."""
    from test_host_api import CUBE
    from tensorray_amd.pyTensorRay.common import get_int_ptr
    from tensorray_amd.pyTensorRay.utils import image_tensor_to_torch
    (tmp_path / "cube.obj").write_text(CUBE)
    (tmp_path / "light.obj").write_text(LIGHT)
    (tmp_path / "floor.obj").write_text(FLOOR)
    (tmp_path / "scene.xml").write_text(LOADED_XML)
    s = TR.Scene()
    s.load_file(str(tmp_path / "scene.xml"))
    s.configure()
    W, H = s.get_width(), s.get_height()
    assert (W, H) == (64, 48)
    TR.set_rnd_seed(77)
    img = image_tensor_to_torch(TR.PathTracer().renderC(s, TR.RenderOptions(77, 3, 8, 0, 0, 0, 0)), H, W)

    def obj_verts(text):
        return np.array([[float(t) for t in ln.split()[1:4]] for ln in text.splitlines() if ln.startswith("v ")],
                        np.float32)

    def faces(shape):
        F = torch.zeros(shape.get_face_count(), 3, dtype=torch.int32)
        shape.get_face_indices(get_int_ptr(F))
        return F.numpy().astype(np.uint32)

    spec = scenes.SceneSpec()
    spec.materials += [scenes.Material(scenes.MAT_DIFFUSE, albedo=(0.5, 0.6, 0.7)),
                       scenes.Material(scenes.MAT_MICROFACET, albedo=(0.1, 0.1, 0.1), spec=(0.04, 0.05, 0.06),
                                       roughness=0.25),
                       scenes.Material(scenes.MAT_ROUGHCONDUCTOR, albedo=(0.143, 0.374, 1.44), spec=(3.98, 2.38, 1.6),
                                       roughness=0.3)]
    cube = obj_verts(CUBE) * np.float32(0.5) + np.array([0, 1, 0], np.float32)
    light = obj_verts(LIGHT) + np.array([0, 3, 0], np.float32)
    spec.meshes += [scenes.Mesh(cube, faces(s.shapes[0]), material=1, flags=scenes.MESH_SMOOTH),
                    scenes.Mesh(obj_verts(FLOOR), faces(s.shapes[1]), material=2, flags=0),
                    scenes.Mesh(light, faces(s.shapes[2]), material=0, emit=(10.0, 9.0, 8.0), flags=0)]
    cam = dict(pos=(0.0, 1.5, 4.0), target=(0.0, 0.75, 0.0), up=(0.0, 1.0, 0.0), fov=40.0)
    want = oracle_mod.Scene(spec).render_film(cam, W, H, 77, 3, 8)
    got = img.reshape(H * W, 3).t().numpy()
    assert (want > 0).mean() > 0.3
    assert np.array_equal(got, want), float(np.abs(got - want).max())
