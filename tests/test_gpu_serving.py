"""Serving loop: every request's result is the same alone and in the mix
(film bit-exact, gradients up to float-atomic order), interior and edge
gradients included; served films equal the oracle's bit for bit and served
gradients the oracle's interior + boundary gradients; admission never
exceeds the HBM budget; and on the four-scene serving workload (1k / 20k /
200k / 1M triangles, res 64..1024, spp 4..64, depth 2..6) every film is
bit-identical alone and in the mix."""
import numpy as np
import pytest

from tensorray_amd import _capi, scenes, serving

pytestmark = pytest.mark.gpu

SPECS = {"a": lambda: scenes.sphere_scene(subdiv=2),
         "b": lambda: scenes.sphere_scene(subdiv=3, amplitude=0.05, smooth=True)}


def _requests(edges=False):
    cam = dict(scenes.CAMERA_DEFAULT)
    out = []
    for i, (scene, side, spp, B) in enumerate([("a", 40, 4, 2), ("b", 24, 8, 3), ("a", 16, 2, 0), ("b", 48, 3, 1),
                                               ("a", 32, 5, 3), ("b", 20, 6, 2)]):
        r = serving.Request(scene=scene, width=side, height=side, spp=spp, max_bounces=B, seed=100 + i, cam=cam,
                            arrival=0.0005 * i, rid=i)
        if edges and i % 2 == 0:
            r.sppe, r.sppe0, r.sppse0, r.sppse1 = 3, 4, 2, 2
        out.append(r)
    return out


def _standalone(ds, r):
    """One request through the plain path: forward with checkpoint, dL/dI,
    interior adjoint, then each boundary term asked for."""
    import torch
    cam = _capi.make_camera(r.width, r.height, **r.cam)
    opt = r.options()
    nb = ds.checkpoint_bytes(cam, opt)
    ck = torch.empty(nb, dtype=torch.uint8, device="cuda")
    film = torch.empty(3, r.width * r.height, device="cuda")
    ds.render_fwd_ckpt(cam, opt, film.data_ptr(), ck.data_ptr(), nb)
    dl = film - r.target if r.target is not None else film.clone()
    info = ds.info()
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    dmat = torch.zeros(info.num_materials, 8, device="cuda")
    ds.render_adj_ckpt(cam, opt, ck.data_ptr(), nb, dl.data_ptr(), dpos.data_ptr(), dmat.data_ptr())
    serving.run_edge_terms(ds, r, cam, opt, dl.data_ptr(), dpos.data_ptr(), None)
    torch.cuda.synchronize()
    return film, dpos, dmat


def _serve(gpu_ctx, order, edges):
    import torch
    srv = serving.Server(gpu_ctx, budget_bytes=1 << 30, batch_paths=9000, chunk_paths=4000, order=order, aging=1e5)
    for k, make in SPECS.items():
        srv.add_scene(k, make())
    reqs = _requests(edges)
    for r in reqs:
        r.target = torch.full((3, r.width * r.height), 0.1, device="cuda")
    srv.run(reqs, realtime=True)
    assert srv.steps > 3  # chunked over several steps
    return srv, reqs


@pytest.mark.parametrize("order", ["fifo", "srpt"])
def test_served_results_equal_standalone(gpu_ctx, order):
    srv, reqs = _serve(gpu_ctx, order, edges=True)
    for r in reqs:
        assert r.t_done is not None and r.latency >= 0
        film, dpos, dmat = _standalone(srv.scenes[r.scene], r)
        assert torch_equal(r.film, film), r.rid
        scale = dpos.abs().max().item()
        if r.max_bounces > 0:
            assert scale > 0, r.rid
        assert (r.dpos - dpos).abs().max().item() <= 1e-4 * scale, r.rid
        mscale = dmat.abs().max().item()
        assert (r.dmat - dmat).abs().max().item() <= 1e-4 * max(mscale, 1e-12), r.rid
    for ds in srv.scenes.values():
        ds.close()


def torch_equal(a, b):
    import torch
    return torch.equal(a, b)


def test_served_results_match_oracle(gpu_ctx, oracle_mod):
    """Served films equal the oracle's films bit for bit; served gradients
    (interior adjoint + the boundary terms of the requests that ask for them)
    equal the oracle's sum of the same estimators within 1e-3 relative."""
    srv, reqs = _serve(gpu_ctx, "fifo", edges=True)
    oracle_scenes = {k: oracle_mod.Scene(make()) for k, make in SPECS.items()}
    masks = {k: make().diff_vertex_mask() for k, make in SPECS.items()}
    for r in reqs:
        S = oracle_scenes[r.scene]
        film = S.render_film(r.cam, r.width, r.height, r.seed, r.max_bounces, r.spp)
        got = r.film.cpu().numpy()
        assert np.array_equal(got, film), (r.rid, np.abs(got - film).max())
        dl = film - np.float32(0.1)
        ref, _ = S.render_adj(r.cam, r.width, r.height, r.seed, r.max_bounces, r.spp, dl)
        for kind, n in r.edge_terms():
            ref = ref + oracle_mod.boundary_adj(S, r.cam, r.width, r.height, r.seed, r.max_bounces, kind, n, dl)
        m = masks[r.scene]
        g = r.dpos.cpu().numpy().astype(np.float64)[m]
        want = ref[m]
        scale = np.abs(want).max()
        if scale == 0:
            assert np.abs(g).max() == 0, r.rid
            continue
        err = np.abs(g - want)
        bound = 1e-3 * np.abs(want) + 1e-5 * scale
        assert (err <= bound).all(), (r.rid, float((err / np.maximum(bound, 1e-30)).max()))
    for ds in srv.scenes.values():
        ds.close()


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
    big = serving.Request(scene="a", width=4096, height=4096, spp=64, max_bounces=3, seed=1,
                          cam=dict(scenes.CAMERA_DEFAULT))
    with pytest.raises(ValueError, match="budget"):
        srv.submit(big)
    for ds in srv.scenes.values():
        ds.close()


CONFIG4_SPECS = {"s1k": lambda: scenes.sphere_scene(subdiv=3, amplitude=0.05, noise_seed=2),
                 "s20k": lambda: scenes.sphere_scene(subdiv=5, amplitude=0.05, noise_seed=1),
                 "s200k": lambda: scenes.sphere_scene(freq=100, amplitude=0.05, noise_seed=3),
                 "s1m": lambda: scenes.sphere_scene(freq=224, amplitude=0.05, noise_seed=4)}


def test_config4_workload_alone_equals_mix(gpu_ctx, oracle_mod):
    """The serving workload at full size: four resident scenes of 1,280 /
    20,480 / 200,000 / 1,003,520 triangles, 16 requests with resolution
    uniform over 64..1024, spp over 4..64, depth over 2..6 (numpy PCG64 seed
    2024), Poisson arrivals at 100 requests / s so that they overlap.  Every
    served film equals the plain path's bit for bit and every gradient
    agrees up to float-atomic order; the smallest request's film equals the
    oracle's."""
    import torch
    specs = {k: make() for k, make in CONFIG4_SPECS.items()}
    srv = serving.Server(gpu_ctx, budget_bytes=64 << 30, batch_paths=16 << 20, chunk_paths=16 << 20)
    for k, spec in specs.items():
        srv.add_scene(k, spec)
    names = list(specs)
    served = serving.make_requests(names, 16, 100.0, seed=2024, res=(64, 1024), spp=(4, 64), depth=(2, 6))
    alone = serving.make_requests(names, 16, 100.0, seed=2024, res=(64, 1024), spp=(4, 64), depth=(2, 6))
    print(f"config4: {sum(r.paths for r in served)} paths in {len(served)} requests", flush=True)
    srv.run(served, realtime=True)
    serving.run_plain(gpu_ctx, srv.scenes, alone, realtime=False)
    torch.cuda.synchronize()
    assert {r.width for r in served} == {r.width for r in alone}
    for a, b in zip(served, alone):
        assert a.rid == b.rid and a.paths == b.paths
        assert torch.equal(a.film, b.film), a.rid
        scale = b.dpos.abs().max().item()
        assert scale > 0, a.rid
        assert (a.dpos - b.dpos).abs().max().item() <= 1e-4 * scale, a.rid
    small = min(served, key=lambda r: r.paths)
    S = oracle_mod.Scene(specs[small.scene])
    film = S.render_film(small.cam, small.width, small.height, small.seed, small.max_bounces, small.spp)
    assert np.array_equal(small.film.cpu().numpy(), film), small.rid
    for ds in srv.scenes.values():
        ds.close()
