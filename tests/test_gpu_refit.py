"""Scene updates on the device (trc_scene_set_vertices with refit): after
the vertices move, the face and vertex normals, areas, emitter table,
secondary-edge list, triangle records and BVH boxes are recomputed on the
GPU; renders and boundary gradients then equal the oracle's for the moved
scene (films bit-exact), the per-camera primary-edge list selected on the
device gives the oracle's primary-edge gradient, the positions read back
are the ones set, and a tree whose SAH cost grows past 1.25x its build's is
rebuilt."""
import numpy as np
import pytest

from tensorray_amd import _capi, scenes

pytestmark = pytest.mark.gpu


def _allpos(spec):
    return np.ascontiguousarray(np.concatenate([np.asarray(m.pos, np.float32) for m in spec.meshes]))


def _moved(make, scale_mesh=1.0, jitter=0.0, light_scale=1.0, seed=0, shuffle=False):
    """The scene of `make` with its first mesh's vertices scaled, jittered
    (uniform in +-jitter per coordinate) or shuffled among themselves, and
    its light mesh scaled."""
    spec = make()
    rng = np.random.default_rng(seed)
    p = np.asarray(spec.meshes[0].pos, np.float64) * scale_mesh
    p = p + rng.uniform(-jitter, jitter, p.shape)
    if shuffle:
        p = p[rng.permutation(len(p))]
    spec.meshes[0].pos = p.astype(np.float32)
    for m in spec.meshes:
        if m.emit is not None and light_scale != 1.0:
            q = np.asarray(m.pos, np.float64)
            c = q.mean(0)
            m.pos = (c + (q - c) * light_scale).astype(np.float32)
    return spec


def _check_forward(gpu_ctx, oracle_mod, ds, spec, W, H, spp, B, seed):
    import torch
    cam = _capi.make_camera(W, H, **scenes.CAMERA_DEFAULT)
    opt = _capi.make_options(seed=seed, max_bounces=B, spp=spp)
    film = torch.empty(3, W * H, device="cuda")
    ds.render_fwd(cam, opt, film.data_ptr())
    torch.cuda.synchronize()
    want = oracle_mod.Scene(spec).render_film(scenes.CAMERA_DEFAULT, W, H, seed, B, spp)
    got = film.cpu().numpy()
    assert np.array_equal(got, want), float(np.abs(got - want).max())
    return got


def _check_boundary(gpu_ctx, oracle_mod, ds, spec, W, H, B, seed, kind, n):
    import torch
    cam = _capi.make_camera(W, H, **scenes.CAMERA_DEFAULT)
    opt = _capi.make_options(seed=seed, max_bounces=B, spp=1, sppe=n, sppse0=n, sppse1=n, sppe0=n)
    dl = np.random.default_rng(4).uniform(0.2, 1.0, (3, W * H)).astype(np.float32)
    dpos = torch.zeros(ds.info().num_vertices, 3, device="cuda")
    ds.render_boundary(cam, opt, kind, torch.from_numpy(dl).cuda().data_ptr(), dpos.data_ptr())
    torch.cuda.synchronize()
    ref = oracle_mod.boundary_adj(oracle_mod.Scene(spec), scenes.CAMERA_DEFAULT, W, H, seed, B, kind, n, dl)
    m = spec.diff_vertex_mask()
    got = dpos.cpu().numpy().astype(np.float64)[m]
    want = ref[m]
    scale = np.abs(want).max()
    assert scale > 0, kind
    err = np.abs(got - want)
    assert (err <= 1e-3 * np.abs(want) + 1e-5 * scale).all(), (kind, float((err / (1e-3 * np.abs(want) + 1e-5 * scale)).max()))


@pytest.mark.parametrize("case", ["flat", "smooth", "light"])
def test_device_refit_matches_oracle(gpu_ctx, oracle_mod, case):
    make = {"flat": lambda: scenes.sphere_scene(subdiv=3, amplitude=0.05),
            "smooth": lambda: scenes.sphere_scene(subdiv=3, amplitude=0.05, smooth=True),
            "light": lambda: scenes.sphere_scene(subdiv=3, amplitude=0.05)}[case]
    ds = _capi.DeviceScene(gpu_ctx, make())
    moved = _moved(make, scale_mesh=1.03, jitter=0.004, light_scale=1.3 if case == "light" else 1.0, seed=1)
    pos = _allpos(moved)
    ds.set_vertices(pos.ctypes.data, len(pos), True)
    info = ds.info()
    assert info.bvh_refits == 1 and info.bvh_rebuilds == 0
    assert np.array_equal(ds.vertices(), pos)
    _check_forward(gpu_ctx, oracle_mod, ds, moved, 48, 40, 4, 3, 11)
    for kind in ("primary", "direct", "indirect"):
        _check_boundary(gpu_ctx, oracle_mod, ds, moved, 48, 40, 2, 12, kind, 4)
    # a second refit from a device pointer
    import torch
    moved2 = _moved(make, scale_mesh=0.98, jitter=0.003, seed=2)
    pos2 = torch.from_numpy(_allpos(moved2)).cuda()
    ds.set_vertices(pos2.data_ptr(), pos2.shape[0], True)
    assert ds.info().bvh_refits == 2
    _check_forward(gpu_ctx, oracle_mod, ds, moved2, 40, 32, 2, 2, 13)
    _check_boundary(gpu_ctx, oracle_mod, ds, moved2, 40, 32, 1, 14, "direct", 4)
    ds.close()


def test_sah_growth_triggers_rebuild(gpu_ctx, oracle_mod):
    """Shuffling the sphere's vertices among themselves stretches every
    triangle across the sphere, so the refit tree's boxes overlap and its SAH
    cost grows past 1.25x its build's; the update rebuilds it, and the scene
    still renders as the oracle's.  A small move does not rebuild."""
    make = lambda: scenes.sphere_scene(subdiv=3)  # noqa: E731
    ds = _capi.DeviceScene(gpu_ctx, make())
    small = _allpos(_moved(make, jitter=0.01, seed=4))
    ds.set_vertices(small.ctypes.data, len(small), True)
    assert ds.info().bvh_rebuilds == 0
    moved = _moved(make, shuffle=True, seed=3)
    pos = _allpos(moved)
    ds.set_vertices(pos.ctypes.data, len(pos), True)
    info = ds.info()
    assert info.bvh_refits == 2 and info.bvh_rebuilds == 1, info.bvh_sah_cost
    assert np.array_equal(ds.vertices(), pos)
    _check_forward(gpu_ctx, oracle_mod, ds, moved, 48, 40, 4, 2, 21)
    ds.close()
