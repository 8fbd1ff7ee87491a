"""Batched launches (trc_render_batch): requests of different scenes, sizes
and depths, split into path ranges over several batches, give the same
per-path radiance and checkpoints (bit for bit) and the same gradients
(float atomics aside) as one trc_render_fwd_ckpt / trc_render_adj_ckpt call
per request."""
import numpy as np
import pytest

from tensorray_amd import _capi, scenes

pytestmark = pytest.mark.gpu

REQUESTS = [
    # scene factory, W, H, spp, max_bounces, seed
    (lambda: scenes.sphere_scene(subdiv=2), 48, 40, 4, 2, 3),
    (lambda: scenes.sphere_scene(subdiv=3, amplitude=0.05, texture_res=32), 32, 32, 8, 3, 4),
    (lambda: scenes.sphere_scene(subdiv=2, smooth=True), 24, 24, 4, 1, 5),
    (lambda: scenes.sphere_scene(subdiv=2), 17, 9, 3, 0, 6),
]


def test_batch_matches_single_requests(gpu_ctx):
    import torch
    reqs = []
    for make, W, H, spp, B, seed in REQUESTS:
        spec = make()
        ds = _capi.DeviceScene(gpu_ctx, spec)
        info = ds.info()
        cam = _capi.make_camera(W, H, **scenes.CAMERA_DEFAULT)
        opt = _capi.make_options(seed=seed, max_bounces=B, spp=spp)
        P = spp * W * H
        nb = ds.checkpoint_bytes(cam, opt)
        r = dict(ds=ds, info=info, cam=cam, opt=opt, P=P, nb=nb, W=W, H=H, spp=spp)
        # single-request reference
        r["paths_ref"] = torch.empty(P, 3, device="cuda")
        ds.render_paths(cam, opt, r["paths_ref"].data_ptr())
        r["ck_ref"] = torch.zeros(nb, dtype=torch.uint8, device="cuda")
        film = torch.empty(3, W * H, device="cuda")
        ds.render_fwd_ckpt(cam, opt, film.data_ptr(), r["ck_ref"].data_ptr(), nb)
        r["dl"] = torch.rand(3, W * H, device="cuda", generator=torch.Generator(device="cuda").manual_seed(seed))
        r["dpos_ref"] = torch.zeros(info.num_vertices, 3, device="cuda")
        r["dmat_ref"] = torch.zeros(info.num_materials, 8, device="cuda")
        r["dtex_ref"] = torch.zeros(max(1, info.texture_floats), device="cuda")
        ds.render_adj_ckpt(cam, opt, r["ck_ref"].data_ptr(), nb, r["dl"].data_ptr(), r["dpos_ref"].data_ptr(),
                           r["dmat_ref"].data_ptr(), r["dtex_ref"].data_ptr() if info.texture_floats else None)
        # batched buffers
        r["paths"] = torch.full((P, 3), -1.0, device="cuda")
        r["ck"] = torch.zeros(nb, dtype=torch.uint8, device="cuda")
        r["dpos"] = torch.zeros(info.num_vertices, 3, device="cuda")
        r["dmat"] = torch.zeros(info.num_materials, 8, device="cuda")
        r["dtex"] = torch.zeros(max(1, info.texture_floats), device="cuda")
        r["vn"] = torch.zeros(info.num_vertices, 3, device="cuda")
        reqs.append(r)
    torch.cuda.synchronize()

    def jobs(cuts, adjoint):
        """cuts[k]: fraction boundaries of request k; one batch per interval."""
        nbatches = len(cuts[0]) - 1
        for b in range(nbatches):
            js = []
            for r, c in zip(reqs, cuts):
                lo, hi = int(c[b] * r["P"]), int(c[b + 1] * r["P"])
                if hi <= lo:
                    continue
                ds = r["ds"]
                if adjoint:
                    js.append(ds.job(r["cam"], r["opt"], lo, hi - lo, ckpt=r["ck"].data_ptr(),
                                     dldi=r["dl"].data_ptr(), dpos=r["dpos"].data_ptr(), dmat=r["dmat"].data_ptr(),
                                     dtex=r["dtex"].data_ptr() if r["info"].texture_floats else None,
                                     vn_scratch=r["vn"].data_ptr()))
                else:
                    js.append(ds.job(r["cam"], r["opt"], lo, hi - lo, paths=r["paths"].data_ptr(),
                                     ckpt=r["ck"].data_ptr()))
            gpu_ctx.render_batch(js, adjoint)

    jobs([[0, 0.3, 1.0], [0, 0.7, 1.0], [0, 0.0, 1.0], [0, 0.5, 1.0]], adjoint=False)
    jobs([[0, 0.5, 1.0], [0, 0.25, 1.0], [0, 1.0, 1.0], [0, 0.1, 1.0]], adjoint=True)
    torch.cuda.synchronize()
    for r in reqs:
        assert torch.equal(r["paths"], r["paths_ref"])
        assert torch.equal(r["ck"], r["ck_ref"])
        for k in ("dpos", "dmat", "dtex"):
            ref = r[k + "_ref"].double()
            scale = ref.abs().max().item()
            if scale == 0:
                assert (r[k] == 0).all()
                continue
            assert (r[k].double() - ref).abs().max().item() <= 1e-4 * scale, k
        assert (r["vn"] == 0).all()  # scratch left zeroed
        r["ds"].close()


def test_batch_rejects_bad_jobs(gpu_ctx):
    import torch
    ds = _capi.DeviceScene(gpu_ctx, scenes.config1_scene())
    cam = _capi.make_camera(8, 8)
    opt = _capi.make_options(seed=1, max_bounces=1, spp=2)
    ck = torch.zeros(ds.checkpoint_bytes(cam, opt), dtype=torch.uint8, device="cuda")
    paths = torch.empty(128, 3, device="cuda")
    with pytest.raises(_capi.TRCError, match="path range"):
        gpu_ctx.render_batch([ds.job(cam, opt, 100, 100, paths=paths.data_ptr(), ckpt=ck.data_ptr())], False)
    with pytest.raises(_capi.TRCError, match="dLdI"):
        gpu_ctx.render_batch([ds.job(cam, opt, 0, 128, ckpt=ck.data_ptr())], True)
    gpu_ctx.render_batch([], False)  # empty batch is a no-op
    ds.close()


def test_batch_with_solo_jobs(gpu_ctx):
    """In a batch mixing scenes, jobs of at least 4M paths run through the
    single-request kernels over their path range and the rest through the
    tiled launch; results are the same either way."""
    import torch
    big = _capi.DeviceScene(gpu_ctx, scenes.sphere_scene(subdiv=3, amplitude=0.05))
    small = _capi.DeviceScene(gpu_ctx, scenes.sphere_scene(subdiv=2, smooth=True))
    reqs = []
    for ds, W, H, spp, B, seed in ((big, 512, 512, 32, 2, 21), (small, 40, 40, 4, 2, 22)):
        info = ds.info()
        cam = _capi.make_camera(W, H, **scenes.CAMERA_DEFAULT)
        opt = _capi.make_options(seed=seed, max_bounces=B, spp=spp)
        P = spp * W * H
        nb = ds.checkpoint_bytes(cam, opt)
        r = dict(ds=ds, cam=cam, opt=opt, P=P, nb=nb, info=info)
        r["ck_ref"] = torch.zeros(nb, dtype=torch.uint8, device="cuda")
        film = torch.empty(3, W * H, device="cuda")
        ds.render_fwd_ckpt(cam, opt, film.data_ptr(), r["ck_ref"].data_ptr(), nb)
        r["paths_ref"] = torch.empty(P, 3, device="cuda")
        ds.render_paths(cam, opt, r["paths_ref"].data_ptr())
        r["dl"] = torch.rand(3, W * H, device="cuda", generator=torch.Generator(device="cuda").manual_seed(seed))
        r["dpos_ref"] = torch.zeros(info.num_vertices, 3, device="cuda")
        r["dmat_ref"] = torch.zeros(info.num_materials, 8, device="cuda")
        ds.render_adj_ckpt(cam, opt, r["ck_ref"].data_ptr(), nb, r["dl"].data_ptr(), r["dpos_ref"].data_ptr(),
                           r["dmat_ref"].data_ptr())
        r["paths"] = torch.full((P, 3), -1.0, device="cuda")
        r["ck"] = torch.zeros(nb, dtype=torch.uint8, device="cuda")
        r["dpos"] = torch.zeros(info.num_vertices, 3, device="cuda")
        r["dmat"] = torch.zeros(info.num_materials, 8, device="cuda")
        r["vn"] = torch.zeros(info.num_vertices, 3, device="cuda")
        reqs.append(r)
    rb, rs = reqs
    cut = 5_000_000  # [0, 5M) of the big request runs solo, [5M, P) is tiled; the small request is tiled
    half = rs["P"] // 2
    for (lo, hi), (slo, shi) in (((0, cut), (0, half)), ((cut, rb["P"]), (half, rs["P"]))):
        for adjoint in (False, True):
            js = []
            for r, a, b in ((rb, lo, hi), (rs, slo, shi)):
                if adjoint:
                    js.append(r["ds"].job(r["cam"], r["opt"], a, b - a, ckpt=r["ck"].data_ptr(), dldi=r["dl"].data_ptr(),
                                          dpos=r["dpos"].data_ptr(), dmat=r["dmat"].data_ptr(),
                                          vn_scratch=r["vn"].data_ptr()))
                else:
                    js.append(r["ds"].job(r["cam"], r["opt"], a, b - a, paths=r["paths"].data_ptr(),
                                          ckpt=r["ck"].data_ptr()))
            if adjoint:
                torch.cuda.synchronize()
            gpu_ctx.render_batch(js, adjoint)
    torch.cuda.synchronize()
    for r in reqs:
        assert torch.equal(r["paths"], r["paths_ref"])
        assert torch.equal(r["ck"], r["ck_ref"])
        for k in ("dpos", "dmat"):
            ref = r[k + "_ref"].double()
            scale = ref.abs().max().item()
            assert scale > 0
            assert (r[k].double() - ref).abs().max().item() <= 1e-4 * scale, k
        assert (r["vn"] == 0).all()
    big.close()
    small.close()
