"""Prepared batches (trc_plan_*): a plan's launch gives the same per-path
radiance and checkpoints (bit for bit) and gradients (float atomics aside)
as trc_render_batch over the same jobs, for any tile size; relaunching an
adjoint plan accumulates again (its gradient copies are re-zeroed); the
tuner returns one of its candidates; and a launch captured into a graph
gives the eager result on every replay, with eager work between replays."""
import pytest

from tensorray_amd import _capi, scenes

pytestmark = pytest.mark.gpu

VIEWS = [
    # W, H, spp, max_bounces, seed, camera position
    (40, 32, 4, 3, 11, (0.0, 1.0, 3.0)),
    (24, 24, 8, 3, 12, (2.0, 1.5, 2.0)),
    (56, 20, 2, 3, 13, (-2.5, 0.5, 1.5)),
]


def _setup(ctx):
    import torch
    ds = _capi.DeviceScene(ctx, scenes.sphere_scene(subdiv=2, amplitude=0.05, texture_res=16))
    info = ds.info()
    views = []
    for W, H, spp, B, seed, pos in VIEWS:
        cam = _capi.make_camera(W, H, pos=pos)
        opt = _capi.make_options(seed=seed, max_bounces=B, spp=spp)
        P = spp * W * H
        nb = ds.checkpoint_bytes(cam, opt)
        g = torch.Generator(device="cuda").manual_seed(seed)
        v = dict(cam=cam, opt=opt, P=P, nb=nb, W=W, H=H,
                 paths=torch.full((P, 3), -1.0, device="cuda"), ck=torch.zeros(nb, dtype=torch.uint8, device="cuda"),
                 dl=torch.rand(3, W * H, device="cuda", generator=g),
                 dpos=torch.zeros(info.num_vertices, 3, device="cuda"),
                 dmat=torch.zeros(info.num_materials, 8, device="cuda"),
                 dtex=torch.zeros(max(1, info.texture_floats), device="cuda"))
        views.append(v)
    return ds, info, views


def _jobs(ds, info, views, adjoint):
    out = []
    for v in views:
        if adjoint:
            out.append(ds.job(v["cam"], v["opt"], 0, v["P"], ckpt=v["ck"].data_ptr(), dldi=v["dl"].data_ptr(),
                              dpos=v["dpos"].data_ptr(), dmat=v["dmat"].data_ptr(),
                              dtex=v["dtex"].data_ptr() if info.texture_floats else None))
        else:
            out.append(ds.job(v["cam"], v["opt"], 0, v["P"], paths=v["paths"].data_ptr(), ckpt=v["ck"].data_ptr()))
    return out


def _close(a, ref, tol=1e-4):
    a, ref = a.double(), ref.double()
    scale = ref.abs().max().item()
    if scale == 0:
        return bool((a == 0).all())
    return (a - ref).abs().max().item() <= tol * scale


def _reference(ctx, ds, info, views):
    """trc_render_batch results of the same jobs."""
    import torch
    ctx.render_batch(_jobs(ds, info, views, False), adjoint=False)
    ctx.render_batch(_jobs(ds, info, views, True), adjoint=True)
    torch.cuda.synchronize()
    ref = [{k: v[k].clone() for k in ("paths", "ck", "dpos", "dmat", "dtex")} for v in views]
    for v in views:
        v["paths"].fill_(-1.0)
        v["ck"].zero_()
        for k in ("dpos", "dmat", "dtex"):
            v[k].zero_()
    torch.cuda.synchronize()
    return ref


def test_plan_matches_batch_and_tiles(gpu_ctx):
    import torch
    ds, info, views = _setup(gpu_ctx)
    ref = _reference(gpu_ctx, ds, info, views)
    fwd = gpu_ctx.plan(_jobs(ds, info, views, False), adjoint=False)
    adj = gpu_ctx.plan(_jobs(ds, info, views, True), adjoint=True)
    assert fwd.tile % 64 == 0 and fwd.tile > 0
    for tile in (0, 64, 256, 4096):
        fwd.tile = tile
        adj.tile = tile
        for v in views:
            v["paths"].fill_(-1.0)
            v["ck"].zero_()
            for k in ("dpos", "dmat", "dtex"):
                v[k].zero_()
        fwd.launch()
        adj.launch()
        torch.cuda.synchronize()
        for v, r in zip(views, ref):
            assert torch.equal(v["paths"], r["paths"]), tile
            assert torch.equal(v["ck"], r["ck"]), tile
            for k in ("dpos", "dmat", "dtex"):
                assert _close(v[k], r[k]), (tile, k)
    # a second adjoint launch adds the same gradient again
    adj.launch()
    torch.cuda.synchronize()
    for v, r in zip(views, ref):
        assert r["dpos"].abs().max().item() > 0
        assert _close(v["dpos"], 2 * r["dpos"])
    with pytest.raises(_capi.TRCError, match="multiple of 64"):
        fwd.tile = 100
    fwd.close()
    adj.close()
    ds.close()


def test_plan_tuner(gpu_ctx):
    import torch
    ds, info, views = _setup(gpu_ctx)
    ref = _reference(gpu_ctx, ds, info, views)
    fwd = gpu_ctx.plan(_jobs(ds, info, views, False), adjoint=False)
    cands = [128, 512, 2048]
    best = fwd.tune(cands, reps=2)
    assert best in cands and fwd.tile == best
    torch.cuda.synchronize()
    for v, r in zip(views, ref):
        assert torch.equal(v["paths"], r["paths"])
    with pytest.raises(_capi.TRCError, match="multiples of 64"):
        fwd.tune([0, 96], reps=1)
    fwd.close()
    with pytest.raises(_capi.TRCError, match="at least one job"):
        gpu_ctx.plan([], adjoint=False)
    ds.close()


def test_plan_graph_replays(gpu_ctx):
    import torch
    ds, info, views = _setup(gpu_ctx)
    ref = _reference(gpu_ctx, ds, info, views)
    fwd = gpu_ctx.plan(_jobs(ds, info, views, False), adjoint=False)
    adj = gpu_ctx.plan(_jobs(ds, info, views, True), adjoint=True)
    # warm the launch path (occupancy queries are cached on first use)
    fwd.launch()
    adj.launch()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    with torch.cuda.graph(g, stream=side):
        fwd.launch(stream=torch.cuda.current_stream())
        adj.launch(stream=torch.cuda.current_stream())
    for rep in range(3):
        # eager work between replays: clear the outputs with fill kernels
        for v in views:
            v["paths"].fill_(-1.0)
            v["ck"].zero_()
            for k in ("dpos", "dmat", "dtex"):
                v[k].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for v, r in zip(views, ref):
            assert torch.equal(v["paths"], r["paths"]), rep
            assert torch.equal(v["ck"], r["ck"]), rep
            for k in ("dpos", "dmat", "dtex"):
                assert _close(v[k], r[k]), (rep, k)
    del g
    fwd.close()
    adj.close()
    ds.close()
