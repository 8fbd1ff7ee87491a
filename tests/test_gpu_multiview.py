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
    return ds, views, targets


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
def test_multiview_matches_single_views(gpu_ctx, smooth):
    import torch
    ds, views, targets = _setup(gpu_ctx, smooth)
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
    mv.close()
    ds.close()


def test_multiview_graph_replays(gpu_ctx):
    import torch
    ds, views, targets = _setup(gpu_ctx)
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
