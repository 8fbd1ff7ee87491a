"""GPU parity of the HIP path against the oracle, through the C ABI.

Bars: hit records bit-exact (integer / index work and the same FP32 watertight
test); per-path radiance and films bit-exact (both sides evaluate the same
FP32 expressions in the same order, contraction off); adjoint within
rel 1e-3 of the oracle's reverse-mode AD, or abs 1e-3 of the largest
gradient component (float atomics change the summation order).
"""
import numpy as np
import pytest

from tensorray_amd import _capi, scenes

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-2.5, 2.5, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3] = 1e-3
    r[:, 4:7] = d
    r[:, 7] = 1e30
    return r


@pytest.mark.parametrize("which", ["config1", "metric"])
def test_intersect_bit_exact(gpu_ctx, oracle_mod, which):
    torch = _torch()
    spec = scenes.config1_scene() if which == "config1" else scenes.metric_scene()
    ds = _capi.DeviceScene(gpu_ctx, spec)
    S = oracle_mod.Scene(spec)
    n = 50000
    rays = _rays(n, 1)
    ref_t, ref_uvt = S.intersect(rays, brute=(which == "config1"))
    rd = torch.from_numpy(rays).cuda()
    tri = torch.empty(n, dtype=torch.int32, device="cuda")
    uvt = torch.empty(n, 3, device="cuda")
    ds.intersect(rd.data_ptr(), n, tri.data_ptr(), uvt.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(tri.cpu().numpy(), ref_t)
    assert np.array_equal(uvt.cpu().numpy().view(np.uint32), ref_uvt.view(np.uint32))
    # occlusion queries agree with closest hits
    ds.intersect(rd.data_ptr(), n, tri.data_ptr(), uvt.data_ptr(), any_hit=True)
    torch.cuda.synchronize()
    assert np.array_equal(tri.cpu().numpy() >= 0, ref_t >= 0)
    ds.close()


def test_intersect_empty_and_miss(gpu_ctx):
    torch = _torch()
    ds = _capi.DeviceScene(gpu_ctx, scenes.config1_scene())
    ds.intersect(0, 0, 0, 0)  # empty batch is a no-op
    r = torch.tensor([[0, 10, 0, 1e-3, 0, 1, 0, 1e30]], dtype=torch.float32, device="cuda")
    tri = torch.empty(1, dtype=torch.int32, device="cuda")
    uvt = torch.empty(1, 3, device="cuda")
    ds.intersect(r.data_ptr(), 1, tri.data_ptr(), uvt.data_ptr())
    torch.cuda.synchronize()
    assert tri.item() == -1 and uvt.cpu().tolist() == [[0.0, 0.0, -1.0]]
    ds.close()


CASES = {
    # name: (scene factory, W, H, spp, B)
    "config1": (scenes.config1_scene, 64, 64, 16, 1),
    "metric_small": (scenes.metric_scene, 96, 96, 4, 3),
    "textured": (lambda: scenes.sphere_scene(subdiv=3, amplitude=0.05, texture_res=64), 48, 48, 4, 2),
    "smooth": (lambda: scenes.sphere_scene(subdiv=2, smooth=True), 48, 48, 4, 2),
    "conductor": (lambda: scenes.sphere_scene(subdiv=3, amplitude=0.05, material=scenes.Material(
        scenes.MAT_ROUGHCONDUCTOR, albedo=(0.143, 0.374, 1.44), spec=(3.98, 2.38, 1.60), roughness=0.35)),
        48, 48, 4, 3),
    # the deepest paths the path record holds (max_bounces 8) in the closed room
    "deepest": (lambda: scenes.box_scene(subdiv=2), 24, 24, 2, 8),
}


@pytest.mark.parametrize("case", list(CASES))
def test_forward_bit_exact(gpu_ctx, oracle_mod, case):
    torch = _torch()
    make, W, H, spp, B = CASES[case]
    spec = make()
    ds = _capi.DeviceScene(gpu_ctx, spec)
    cam_d = scenes.CAMERA_DEFAULT
    cam = _capi.make_camera(W, H, **cam_d)
    opt = _capi.make_options(seed=1234, max_bounces=B, spp=spp)
    paths = torch.empty(spp * W * H, 3, device="cuda")
    ds.render_paths(cam, opt, paths.data_ptr())
    film = torch.empty(3, W * H, device="cuda")
    ds.render_fwd(cam, opt, film.data_ptr())
    torch.cuda.synchronize()
    S = oracle_mod.Scene(spec)
    ref = S.render_paths(cam_d, W, H, 1234, B, 0, spp * W * H)
    got = paths.cpu().numpy()
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    ref_film = S.render_film(cam_d, W, H, 1234, B, spp)
    assert np.array_equal(film.cpu().numpy(), ref_film)
    assert (ref_film > 0).mean() > 0.2  # not a degenerate (black) image
    ds.close()


def _grad_close(g, r, rel=1e-3, floor=1e-5):
    """Per entry: |g - r| <= rel |r| + floor max|r|.  The floor covers the
    reassociation of the float atomics that accumulate each entry on the GPU
    (the oracle sums in double); returns (ok, worst error / bound)."""
    scale = np.abs(r).max()
    err = np.abs(g - r)
    bound = rel * np.abs(r) + floor * scale
    return bool((err <= bound).all()), float((err / np.maximum(bound, 1e-300)).max())


@pytest.mark.parametrize("case", ["config1", "metric_small", "textured", "conductor", "smooth", "deepest"])
def test_adjoint_matches_oracle(gpu_ctx, oracle_mod, case):
    torch = _torch()
    make, W, H, spp, B = CASES[case]
    spec = make()
    ds = _capi.DeviceScene(gpu_ctx, spec)
    info = ds.info()
    cam_d = scenes.CAMERA_DEFAULT
    cam = _capi.make_camera(W, H, **cam_d)
    opt = _capi.make_options(seed=99, max_bounces=B, spp=spp)
    dl = np.random.default_rng(0).uniform(-1, 1, (3, W * H)).astype(np.float32)
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    dmat = torch.zeros(info.num_materials, 8, device="cuda")
    dtex = torch.zeros(max(1, info.texture_floats), device="cuda")
    ds.render_adj(cam, opt, torch.from_numpy(dl).cuda().data_ptr(), dpos.data_ptr(), dmat.data_ptr(),
                  dtex_ptr=dtex.data_ptr() if info.texture_floats else None)
    torch.cuda.synchronize()
    S = oracle_mod.Scene(spec)
    rp, rm, rt = S.render_adj(cam_d, W, H, 99, B, spp, dl, with_tex=True)
    if info.texture_floats:
        assert np.abs(rt).max() > 0
        ok, e = _grad_close(dtex.cpu().numpy().astype(np.float64), rt)
        assert ok, e
    mask = spec.diff_vertex_mask()
    ok, e = _grad_close(dpos.cpu().numpy().astype(np.float64)[mask], rp[mask])
    assert ok, e
    # non-differentiable shapes receive nothing
    assert (dpos.cpu().numpy()[~mask] == 0).all()
    gm = dmat.cpu().numpy().astype(np.float64)
    ok, e = _grad_close(gm[:, :7], rm[:, :7])
    assert ok, e
    ds.close()


def test_full_size_forward_deterministic_and_sampled_parity(gpu_ctx, oracle_mod):
    """Metric workload at full size (512x512, 16 spp, depth 4): two launches
    give identical films; a strided sample of paths matches the oracle."""
    torch = _torch()
    spec = scenes.metric_scene()
    ds = _capi.DeviceScene(gpu_ctx, spec)
    W = H = 512
    spp, B = 16, 3
    cam_d = scenes.CAMERA_DEFAULT
    cam = _capi.make_camera(W, H, **cam_d)
    opt = _capi.make_options(seed=1234, max_bounces=B, spp=spp)
    paths = torch.empty(spp * W * H, 3, device="cuda")
    ds.render_paths(cam, opt, paths.data_ptr())
    f1 = torch.empty(3, W * H, device="cuda")
    f2 = torch.empty(3, W * H, device="cuda")
    ds.render_fwd(cam, opt, f1.data_ptr())
    ds.render_fwd(cam, opt, f2.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(f1, f2)
    assert torch.isfinite(f1).all() and (f1 >= 0).all()
    S = oracle_mod.Scene(spec)
    got = paths.cpu().numpy()
    for first in (0, 1_000_003, 3_000_017):
        ref = S.render_paths(cam_d, W, H, 1234, B, first, 20000)
        assert np.array_equal(got[first:first + 20000], ref)
    ds.close()


def test_adjoint_is_linear_in_dLdI(gpu_ctx):
    """Size-independent property at the full metric size: the adjoint of
    a*g1 + b*g2 equals a*adj(g1) + b*adj(g2) up to float atomics."""
    torch = _torch()
    spec = scenes.metric_scene()
    ds = _capi.DeviceScene(gpu_ctx, spec)
    info = ds.info()
    W = H = 512
    cam = _capi.make_camera(W, H, **scenes.CAMERA_DEFAULT)
    opt = _capi.make_options(seed=7, max_bounces=3, spp=4)
    g = torch.Generator(device="cuda").manual_seed(0)
    g1 = torch.rand(3, W * H, device="cuda", generator=g) - 0.5
    g2 = torch.rand(3, W * H, device="cuda", generator=g) - 0.5

    def adj(dl):
        d = torch.zeros(info.num_vertices, 3, device="cuda")
        ds.render_adj(cam, opt, dl.contiguous().data_ptr(), d.data_ptr(), None)
        torch.cuda.synchronize()
        return d.double()

    a1, a2, a12 = adj(g1), adj(g2), adj(2.0 * g1 - 0.5 * g2)
    ref = 2.0 * a1 - 0.5 * a2
    scale = ref.abs().max().item()
    assert scale > 0
    assert ((a12 - ref).abs().max().item()) < 1e-4 * scale
    ds.close()


def test_invalid_options_rejected(gpu_ctx):
    torch = _torch()
    ds = _capi.DeviceScene(gpu_ctx, scenes.config1_scene())
    cam = _capi.make_camera(8, 8)
    film = torch.empty(3, 64, device="cuda")
    with pytest.raises(_capi.TRCError, match="max_bounces"):
        ds.render_fwd(cam, _capi.make_options(max_bounces=9, spp=1), film.data_ptr())
    with pytest.raises(_capi.TRCError, match="spp"):
        ds.render_fwd(cam, _capi.make_options(max_bounces=1, spp=0), film.data_ptr())
    ds.close()


BOUNDARY_CASES = {
    # kind: (scene factory, camera, W, H, spp, B)
    "primary": (lambda: scenes.sphere_scene(subdiv=2), scenes.CAMERA_DEFAULT, 48, 48, 4, 1),
    "pixel": (lambda: scenes.sphere_scene(subdiv=2), scenes.CAMERA_DEFAULT, 48, 48, 4, 1),
    "direct": (lambda: scenes.sphere_scene(subdiv=2), scenes.CAMERA_DEFAULT, 48, 48, 8, 2),
    "indirect": (lambda: scenes.box_scene(subdiv=2), scenes.BOX_CAMERA, 32, 32, 16, 3),
}


@pytest.mark.parametrize("kind", list(BOUNDARY_CASES))
def test_boundary_gradients_match_oracle(gpu_ctx, oracle_mod, kind):
    torch = _torch()
    make, cam_d, W, H, spp, B = BOUNDARY_CASES[kind]
    spec = make()
    ds = _capi.DeviceScene(gpu_ctx, spec)
    info = ds.info()
    cam = _capi.make_camera(W, H, **cam_d)
    opt = _capi.make_options(seed=31, max_bounces=B, spp=1, sppe=spp, sppe0=spp, sppse0=spp, sppse1=spp)
    dl = np.random.default_rng(2).uniform(0.2, 1.0, (3, W * H)).astype(np.float32)
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    ds.render_boundary(cam, opt, kind, torch.from_numpy(dl).cuda().data_ptr(), dpos.data_ptr())
    torch.cuda.synchronize()
    S = oracle_mod.Scene(spec)
    ref = oracle_mod.boundary_adj(S, cam_d, W, H, 31, B, kind, spp, dl)
    got = dpos.cpu().numpy().astype(np.float64)
    mask = spec.diff_vertex_mask()
    assert np.abs(ref[mask]).max() > 0, "oracle boundary term is zero on this scene"
    ok, e = _grad_close(got[mask], ref[mask])
    assert ok, e
    ds.close()


@pytest.mark.parametrize("kind", ["primary", "direct", "indirect"])
def test_boundary_multi_chunk_matches_oracle(gpu_ctx, oracle_mod, kind):
    """More samples than one survivor-queue chunk (2^22): 256 x 256 x 65 =
    4,259,840 edge samples run as two select / eval chunks that share the
    queue; the second chunk's sample indices continue the first's."""
    torch = _torch()
    make, cam_d, _, _, _, B = BOUNDARY_CASES[kind]
    W = H = 256
    spp = 65
    assert W * H * spp > (1 << 22)
    spec = make()
    ds = _capi.DeviceScene(gpu_ctx, spec)
    info = ds.info()
    cam = _capi.make_camera(W, H, **cam_d)
    opt = _capi.make_options(seed=31, max_bounces=B, spp=1, sppe=spp, sppse0=spp, sppse1=spp)
    dl = np.random.default_rng(2).uniform(0.2, 1.0, (3, W * H)).astype(np.float32)
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    ds.render_boundary(cam, opt, kind, torch.from_numpy(dl).cuda().data_ptr(), dpos.data_ptr())
    torch.cuda.synchronize()
    S = oracle_mod.Scene(spec)
    ref = oracle_mod.boundary_adj(S, cam_d, W, H, 31, B, kind, spp, dl)
    got = dpos.cpu().numpy().astype(np.float64)
    mask = spec.diff_vertex_mask()
    assert np.abs(ref[mask]).max() > 0
    ok, e = _grad_close(got[mask], ref[mask])
    assert ok, e
    ds.close()


@pytest.mark.parametrize("subdiv", [2, 4])
def test_guided_direct_matches_oracle(gpu_ctx, oracle_mod, subdiv):
    """Guided sampling of the direct boundary term (AQ_distrb.h): guide build
    and guided samples against the oracle; subdivision 4 has several
    thousand secondary edges, so the guide's blocked reductions span
    several chunks."""
    torch = _torch()
    spec = scenes.sphere_scene(subdiv=subdiv)
    W = H = 48
    ds = _capi.DeviceScene(gpu_ctx, spec)
    info = ds.info()
    cam_d = scenes.CAMERA_DEFAULT
    cam = _capi.make_camera(W, H, **cam_d)
    opt = _capi.make_options(seed=31, max_bounces=2, spp=1, sppse0=8, g_direct=True, g_direct_spp=8, g_eps=0.1)
    dl = np.random.default_rng(2).uniform(0.2, 1.0, (3, W * H)).astype(np.float32)
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    ds.render_boundary(cam, opt, "direct", torch.from_numpy(dl).cuda().data_ptr(), dpos.data_ptr())
    torch.cuda.synchronize()
    S = oracle_mod.Scene(spec)
    ref = oracle_mod.boundary_adj(S, cam_d, W, H, 31, 2, "direct", 8, dl, guide=(8, 0.1))
    unguided = oracle_mod.boundary_adj(S, cam_d, W, H, 31, 2, "direct", 8, dl)
    mask = spec.diff_vertex_mask()
    got = dpos.cpu().numpy().astype(np.float64)
    assert not np.allclose(ref[mask], unguided[mask])  # guiding changes the samples
    ok, e = _grad_close(got[mask], ref[mask])
    assert ok, e
    # guide limits
    small = _capi.make_options(seed=31, max_bounces=2, spp=1, sppse0=1, g_direct=True, g_direct_max_size=10)
    with pytest.raises(_capi.TRCError, match="g_direct_max_size"):
        ds.render_boundary(cam, small, "direct", torch.from_numpy(dl).cuda().data_ptr(), dpos.data_ptr())
    bad = _capi.make_options(seed=31, max_bounces=2, spp=1, sppse0=1, g_direct=True, g_eps=1.5)
    with pytest.raises(_capi.TRCError, match="g_eps"):
        ds.render_boundary(cam, bad, "direct", torch.from_numpy(dl).cuda().data_ptr(), dpos.data_ptr())
    ds.close()


def test_guide_kept_across_views(gpu_ctx, oracle_mod):
    """g_keep: the direct-term guide is built from g_seed, kept on the scene
    and shared by later calls with other seeds and cameras (the views of one
    optimisation step), then rebuilt once the vertices move."""
    torch = _torch()
    spec = scenes.sphere_scene(subdiv=2)
    W = H = 48
    ds = _capi.DeviceScene(gpu_ctx, spec)
    info = ds.info()
    S = oracle_mod.Scene(spec)
    mask = spec.diff_vertex_mask()
    dl = np.random.default_rng(2).uniform(0.2, 1.0, (3, W * H)).astype(np.float32)
    dld = torch.from_numpy(dl).cuda()
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    side = dict(pos=(1.5, 1.2, 2.5), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0)

    def check(scene_ref, cam_d, seed):
        opt = _capi.make_options(seed=seed, max_bounces=2, spp=1, sppse0=8, g_direct=True, g_direct_spp=8,
                                 g_eps=0.1, g_keep=True, g_seed=5)
        dpos.zero_()
        ds.render_boundary(_capi.make_camera(W, H, **cam_d), opt, "direct", dld.data_ptr(), dpos.data_ptr())
        torch.cuda.synchronize()
        ref = oracle_mod.boundary_adj(scene_ref, cam_d, W, H, seed, 2, "direct", 8, dl, guide=(8, 0.1, 5))
        own = oracle_mod.boundary_adj(scene_ref, cam_d, W, H, seed, 2, "direct", 8, dl, guide=(8, 0.1))
        assert not np.allclose(ref[mask], own[mask])  # the guide seed matters
        ok, e = _grad_close(dpos.cpu().numpy().astype(np.float64)[mask], ref[mask])
        assert ok, (seed, e)

    check(S, scenes.CAMERA_DEFAULT, 31)   # builds the guide from g_seed 5
    check(S, side, 77)                    # reuses it
    moved = scenes.sphere_scene(subdiv=2)
    moved.meshes[0].pos = (np.asarray(moved.meshes[0].pos, np.float64) * 1.02).astype(np.float32)
    allpos = np.ascontiguousarray(np.concatenate([np.asarray(m.pos, np.float32) for m in moved.meshes]))
    ds.set_vertices(allpos.ctypes.data, len(allpos), True)
    check(oracle_mod.Scene(moved), side, 77)   # the scene changed: rebuilt
    ds.close()


@pytest.mark.parametrize("case", ["config1", "metric_small", "textured", "smooth", "conductor", "deepest"])
def test_checkpointed_adjoint(gpu_ctx, oracle_mod, case):
    """trc_render_fwd_ckpt writes the same film as trc_render_fwd; the
    adjoint from the checkpoint (no traversal) matches the oracle's AD like
    the replay adjoint does."""
    torch = _torch()
    make, W, H, spp, B = CASES[case]
    spec = make()
    ds = _capi.DeviceScene(gpu_ctx, spec)
    info = ds.info()
    cam_d = scenes.CAMERA_DEFAULT
    cam = _capi.make_camera(W, H, **cam_d)
    opt = _capi.make_options(seed=99, max_bounces=B, spp=spp)
    nb = ds.checkpoint_bytes(cam, opt)
    assert nb == spp * W * H * (4 + 12 * (B + 1))
    ck = torch.empty(nb, dtype=torch.uint8, device="cuda")
    f1 = torch.empty(3, W * H, device="cuda")
    f2 = torch.empty(3, W * H, device="cuda")
    ds.render_fwd(cam, opt, f1.data_ptr())
    ds.render_fwd_ckpt(cam, opt, f2.data_ptr(), ck.data_ptr(), nb)
    torch.cuda.synchronize()
    assert torch.equal(f1, f2)
    dl = np.random.default_rng(0).uniform(-1, 1, (3, W * H)).astype(np.float32)
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    dmat = torch.zeros(info.num_materials, 8, device="cuda")
    dtex = torch.zeros(max(1, info.texture_floats), device="cuda")
    ds.render_adj_ckpt(cam, opt, ck.data_ptr(), nb, torch.from_numpy(dl).cuda().data_ptr(), dpos.data_ptr(),
                       dmat.data_ptr(), dtex.data_ptr() if info.texture_floats else None)
    torch.cuda.synchronize()
    S = oracle_mod.Scene(spec)
    rp, rm, rt = S.render_adj(cam_d, W, H, 99, B, spp, dl, with_tex=True)
    mask = spec.diff_vertex_mask()
    ok, e = _grad_close(dpos.cpu().numpy().astype(np.float64)[mask], rp[mask])
    assert ok, e
    ok, e = _grad_close(dmat.cpu().numpy().astype(np.float64)[:, :7], rm[:, :7])
    assert ok, e
    if info.texture_floats:
        ok, e = _grad_close(dtex.cpu().numpy().astype(np.float64), rt)
        assert ok, e
    with pytest.raises(_capi.TRCError, match="too small"):
        ds.render_adj_ckpt(cam, opt, ck.data_ptr(), nb - 1, torch.from_numpy(dl).cuda().data_ptr(), dpos.data_ptr())
    ds.close()
