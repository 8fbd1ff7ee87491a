"""Multi-view rendering through prepared batches: every view's film equals
its own single-view render bit for bit, the summed gradients equal the sum
of the single-view adjoints (float atomics aside), and a captured step gives
the eager result on every replay, with eager work between replays."""
import pytest

from tensorray_amd import _capi, multiview, scenes

pytestmark = pytest.mark.gpu


def _close(a, ref, tol=1e-4):
    a, ref = a.double(), ref.double()
    scale = ref.abs().max().item()
    if scale == 0:
        return bool((a == 0).all())
    return (a - ref).abs().max().item() <= tol * scale


def _setup(ctx, smooth=False):
    import torch
    spec = scenes.multiview_scene(freq=6, amplitude=0.05, texture_res=32)
    if smooth:
        spec.meshes[0].flags |= scenes.MESH_SMOOTH
    ds = _capi.DeviceScene(ctx, spec)
    cams = scenes.fibonacci_cameras(5, radius=3.0, seed=3)
    views = multiview.random_views(cams, seed=1, res=(16, 40), spp_choices=(2, 4), depth=3)
    g = torch.Generator(device="cuda").manual_seed(5)
    targets = [torch.rand(3, v.width * v.height, device="cuda", generator=g) for v in views]
    return ds, views, targets, spec


def _single_view_reference(ds, views, targets):
    import torch
    info = ds.info()
    films = []
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    dmat = torch.zeros(info.num_materials, 8, device="cuda")
    dtex = torch.zeros(max(1, info.texture_floats), device="cuda")
    for v, t in zip(views, targets):
        cam = _capi.make_camera(v.width, v.height, **v.cam)
        opt = _capi.make_options(seed=v.seed, max_bounces=v.max_bounces, spp=v.spp)
        nb = ds.checkpoint_bytes(cam, opt)
        ck = torch.empty(nb, dtype=torch.uint8, device="cuda")
        film = torch.empty(3, v.width * v.height, device="cuda")
        ds.render_fwd_ckpt(cam, opt, film.data_ptr(), ck.data_ptr(), nb)
        dl = film - t
        ds.render_adj_ckpt(cam, opt, ck.data_ptr(), nb, dl.data_ptr(), dpos.data_ptr(), dmat.data_ptr(),
                           dtex.data_ptr())
        films.append(film)
    torch.cuda.synchronize()
    return films, dpos, dmat, dtex


@pytest.mark.parametrize("smooth", [False, True])
def test_multiview_matches_single_views(gpu_ctx, oracle_mod, smooth):
    import numpy as np
    import torch
    ds, views, targets, spec = _setup(gpu_ctx, smooth)
    films, dpos, dmat, dtex = _single_view_reference(ds, views, targets)
    assert dpos.abs().max().item() > 0 and dtex.abs().max().item() > 0
    mv = multiview.MultiView(gpu_ctx, ds, views, targets)
    mv.tune((64, 128, 256), reps=1)
    mv.step_eager()
    torch.cuda.synchronize()
    for k in range(len(views)):
        assert torch.equal(mv.film[k], films[k]), k
    assert _close(mv.dpos, dpos) and _close(mv.dmat, dmat) and _close(mv.dtex, dtex)
    assert (mv.vn == 0).all()
    # every view's film equals the oracle's, and the summed gradients the
    # oracle's sum over views
    S = oracle_mod.Scene(spec)
    ref = np.zeros((ds.info().num_vertices, 3))
    for k, v in enumerate(views):
        want = S.render_film(v.cam, v.width, v.height, v.seed, v.max_bounces, v.spp)
        assert np.array_equal(mv.film[k].cpu().numpy(), want), k
        dl = want - targets[k].cpu().numpy()
        ref += S.render_adj(v.cam, v.width, v.height, v.seed, v.max_bounces, v.spp, dl)[0]
    m = spec.diff_vertex_mask()
    got = mv.dpos.cpu().numpy().astype(np.float64)[m]
    scale = np.abs(ref[m]).max()
    assert scale > 0
    assert (np.abs(got - ref[m]) <= 1e-3 * np.abs(ref[m]) + 1e-5 * scale).all()
    mv.close()
    ds.close()


def test_multiview_graph_replays(gpu_ctx):
    import torch
    ds, views, targets, _ = _setup(gpu_ctx)
    films, dpos, dmat, dtex = _single_view_reference(ds, views, targets)
    mv = multiview.MultiView(gpu_ctx, ds, views, targets)
    mv.capture()
    for rep in range(3):
        # eager work between replays: scribble over every output
        for k in range(len(views)):
            mv.film[k].fill_(-1.0)
            mv.paths[k].fill_(-1.0)
        mv.dpos.fill_(7.0)
        torch.cuda.synchronize()
        mv.replay()
        torch.cuda.synchronize()
        for k in range(len(views)):
            assert torch.equal(mv.film[k], films[k]), (rep, k)
        assert _close(mv.dpos, dpos) and _close(mv.dmat, dmat) and _close(mv.dtex, dtex), rep
    mv.close()
    ds.close()


def test_config3_workload_matches_single_views_and_oracle(gpu_ctx, oracle_mod):
    """The multi-view workload at full size: 32 views on a Fibonacci sphere
    of radius 3 (seed 0) around the 200,000-triangle displaced geodesic
    sphere with a 1024^2 Perlin albedo texture, square resolutions uniform
    over 128..512, spp from {8, 16, 32}, depth 4, tuned tiles and a captured
    graph.  Every view's film equals its single-view render bit for bit, the
    summed gradients equal the sum of the single-view adjoints up to
    float-atomic order, and 20,000 paths of two views (four ranges of 5,000
    spread over each view) equal the oracle's per-path radiance bit for bit."""
    import numpy as np
    import torch
    spec = scenes.multiview_scene(freq=100, amplitude=0.05, noise_seed=1, texture_res=1024, texture_seed=7)
    ds = _capi.DeviceScene(gpu_ctx, spec)
    assert ds.info().num_triangles == 200002
    cams = scenes.fibonacci_cameras(32, radius=3.0, seed=0)
    views = multiview.random_views(cams, seed=0, res=(128, 512), spp_choices=(8, 16, 32), depth=4)
    g = torch.Generator(device="cuda").manual_seed(11)
    targets = [torch.rand(3, v.width * v.height, device="cuda", generator=g) for v in views]
    print(f"config3: {sum(v.paths for v in views)} paths over {len(views)} views", flush=True)
    mv = multiview.MultiView(gpu_ctx, ds, views, targets)
    mv.tune((128, 256, 1024), reps=1)
    mv.capture()
    mv.replay()
    torch.cuda.synchronize()
    films, dpos, dmat, dtex = _single_view_reference(ds, views, targets)
    for k in range(len(views)):
        assert torch.equal(mv.film[k], films[k]), k
    assert _close(mv.dpos, dpos) and _close(mv.dmat, dmat) and _close(mv.dtex, dtex)
    S = oracle_mod.Scene(spec)
    for k in (0, len(views) // 2):
        v = views[k]
        for q in range(4):
            first = (v.paths - 5000) * q // 3
            ref = S.render_paths(v.cam, v.width, v.height, v.seed, v.max_bounces, first, 5000)
            got = mv.paths[k][first:first + 5000].cpu().numpy()
            assert np.array_equal(got, ref), (k, first, float(np.abs(got - ref).max()))
    mv.close()
    ds.close()
