"""Serving loop: every request's result is the same alone and in the mix
(film bit-exact, gradients up to float-atomic order), admission never
exceeds the HBM budget, and requests arriving while others run overlap with
them in the same batches."""
import pytest

from tensorray_amd import _capi, scenes, serving

pytestmark = pytest.mark.gpu


def _requests():
    cam = dict(scenes.CAMERA_DEFAULT)
    out = []
    for i, (scene, side, spp, B) in enumerate([("a", 40, 4, 2), ("b", 24, 8, 3), ("a", 16, 2, 0), ("b", 48, 3, 1),
                                               ("a", 32, 5, 3), ("b", 20, 6, 2)]):
        out.append(serving.Request(scene=scene, width=side, height=side, spp=spp, max_bounces=B, seed=100 + i,
                                   cam=cam, arrival=0.0005 * i, rid=i))
    return out


@pytest.mark.parametrize("order", ["fifo", "srpt"])
def test_served_results_equal_standalone(gpu_ctx, order):
    import torch
    srv = serving.Server(gpu_ctx, budget_bytes=1 << 30, batch_paths=9000, chunk_paths=4000, order=order, aging=1e5)
    srv.add_scene("a", scenes.sphere_scene(subdiv=2))
    srv.add_scene("b", scenes.sphere_scene(subdiv=3, amplitude=0.05, smooth=True))
    reqs = _requests()
    for r in reqs:
        r.target = torch.full((3, r.width * r.height), 0.1, device="cuda")
    srv.run(reqs, realtime=True)
    assert srv.steps > 3  # chunked over several steps
    for r in reqs:
        assert r.t_done is not None and r.latency >= 0
        ds = srv.scenes[r.scene]
        cam = _capi.make_camera(r.width, r.height, **r.cam)
        opt = _capi.make_options(seed=r.seed, max_bounces=r.max_bounces, spp=r.spp)
        nb = ds.checkpoint_bytes(cam, opt)
        ck = torch.empty(nb, dtype=torch.uint8, device="cuda")
        film = torch.empty(3, r.width * r.height, device="cuda")
        ds.render_fwd_ckpt(cam, opt, film.data_ptr(), ck.data_ptr(), nb)
        dl = film - r.target
        info = ds.info()
        dpos = torch.zeros(info.num_vertices, 3, device="cuda")
        dmat = torch.zeros(info.num_materials, 8, device="cuda")
        ds.render_adj_ckpt(cam, opt, ck.data_ptr(), nb, dl.data_ptr(), dpos.data_ptr(), dmat.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(r.film, film), r.rid
        scale = dpos.abs().max().item()
        if r.max_bounces > 0:
            assert scale > 0, r.rid
        assert (r.dpos - dpos).abs().max().item() <= 1e-4 * scale, r.rid
        mscale = dmat.abs().max().item()
        assert (r.dmat - dmat).abs().max().item() <= 1e-4 * max(mscale, 1e-12), r.rid


def test_admission_respects_budget(gpu_ctx):
    srv = serving.Server(gpu_ctx, budget_bytes=1 << 30, batch_paths=1 << 20, chunk_paths=1 << 20)
    srv.add_scene("a", scenes.sphere_scene(subdiv=2))
    srv.add_scene("b", scenes.sphere_scene(subdiv=2))
    reqs = _requests()
    need = [serving._Active.bytes_needed(r, srv.scenes[r.scene]) for r in reqs]
    srv.budget = max(need) + min(need) - 1  # at most the largest request plus less than the smallest
    srv.run(reqs, realtime=False)
    assert all(r.t_done is not None for r in reqs)
    assert srv.peak_in_use <= srv.budget
    big = serving.Request(scene="a", width=4096, height=4096, spp=64, max_bounces=3, seed=1, cam=dict(scenes.CAMERA_DEFAULT))
    with pytest.raises(ValueError, match="budget"):
        srv.submit(big)
