"""Benchmark: Mpaths/s of forward + adjoint on one MI355X.

Workload (metric workload M): 512x512 pinhole camera (fov 45) at (0, 1, 3)
looking at the origin; an icosphere of subdivision 5 (20,480 triangles)
displaced along its normals by seeded 3-octave value noise (amplitude 0.05,
noise seed 1) with a GGX microfacet BSDF (roughness 0.3, F0 0.04, kd 0) on a
Lambertian ground quad under a 1x1 area light (radiance 10) at y = 2.
16 spp, max depth 4 (3 NEE + BSDF bounces after the primary hit), FP32.

One step = forward render of all W*H*spp paths (per-path radiance + film
reduction) + dL/dI = I - I_target on the device + the adjoint (backward
sweep scattering into vertex and material gradients).  By default the
forward writes a path checkpoint (hits and visibility of every ray query)
and the adjoint re-evaluates the paths from it without tracing;
--adjoint replay re-traces every path instead.
I_target is the same scene with displacement amplitude 0.06, rendered once
with seed 99.  Inputs are resident in HBM before the timed region.

Prints one JSON line (the contract of the driver).
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

MI355X_HBM_PEAK = 8.0e12  # B/s, published


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--width", type=int, default=512)
    p.add_argument("--height", type=int, default=512)
    p.add_argument("--spp", type=int, default=16)
    p.add_argument("--depth", type=int, default=4)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-spp", type=int, default=16, help="samples per pixel of the CPU baseline sample")
    p.add_argument("--cpu-runs", type=int, default=5, help="timed CPU baseline runs (median reported)")
    p.add_argument("--adjoint", choices=["checkpoint", "replay"], default="checkpoint",
                   help="adjoint from the forward's path checkpoint (no re-tracing) or by replaying the paths")
    p.add_argument("--serve-requests", type=int, default=48,
                   help="requests of the served-vs-plain comparison (0 skips it)")
    p.add_argument("--serve-sweep", action="store_true",
                   help="serving workload over all load points: 4 resident scenes (1k / 20k / 200k / 1M "
                        "triangles), res U(64..1024), spp U(4..64), depth U(2..6), Poisson rates 0.25..1.25 x "
                        "plain capacity (default: saturation and the 0.75 load point only)")
    p.add_argument("--no-serve-mix", action="store_true",
                   help="skip the serving workload (4 scenes, mixed sizes) entirely")
    p.add_argument("--serve-large-scenes", action="store_true",
                   help="serve over 1k / 20k / 200k / 1M-triangle scenes (slower set-up) instead of 1k / 20k")
    p.add_argument("--multiview-views", type=int, default=32,
                   help="views of the multi-view workload (0 skips it)")
    p.add_argument("--boundary-spp", type=int, default=16,
                   help="edge samples per pixel of each boundary class in the boundary section (0: skip it)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="after the timed region, run an elementwise kernel over a MIB-sized device buffer "
                        "(reads and writes MIB: the reference of known traffic for scaling the FETCH_SIZE / "
                        "WRITE_SIZE counters)")
    return p.parse_args()


def algorithmic_bytes(stats, spp_paths, B, n_light_tris):
    """Algorithmic HBM bytes per path of the forward and adjoint passes.

    Per ray query: 32 B ray + 16 B hit + 128 B per BVH node visited + 48 B
    per triangle tested (averages measured on the same rays and BVH by the
    statistics kernel).  Per bounce: 56 B shading data + NEE light-table
    search 4*ceil(log2 T_L) + 48 B light triangle; film read-modify-write 24 B.
    The replay adjoint re-traces the forward (same queries) and adds 12 B of
    dL/dI and per path vertex the vertex-gradient RMW (3 vertices * 12 B * 2)
    and the material-gradient RMW (2 * 4 B).  With a checkpoint the forward
    also writes 4 B + 12 B per vertex, and the adjoint reads them back in
    place of every ray query.
    """
    paths, queries, nodes, tris = stats
    q_per_path = queries / paths
    b_query = 32 + 16 + 128 * (nodes / queries) + 48 * (tris / queries)
    bounces = max(0.0, (q_per_path - 1.0) / 2.0)  # 1 primary + (shadow + bsdf) per bounce
    b_shade = 56 * (1 + bounces)
    b_nee = bounces * (4 * math.ceil(math.log2(max(2, n_light_tris))) + 48)
    b_grad = 12 + (1 + bounces) * (3 * 12 * 2 + 2 * 4)
    b_ckpt = 4 + 12 * (1 + bounces)
    b_fwd = q_per_path * b_query + b_shade + b_nee + 24
    return dict(fwd=b_fwd, adj=b_fwd + b_grad, fwd_ckpt=b_fwd + b_ckpt, adj_ckpt=b_ckpt + b_shade + b_nee + b_grad), \
        dict(queries_per_path=q_per_path, nodes_per_query=nodes / queries, tris_per_query=tris / queries,
             bytes_per_query=b_query)


def host_cpu_info():
    """What the host offers: logical CPUs, the CPUs this process may run on,
    the CPU model, and the thread share granted to one GPU's job (the
    OMP_NUM_THREADS the job runs under; all usable CPUs when unset)."""
    info = dict(os_cpu_count=os.cpu_count())
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["usable_cpus"] = os.cpu_count()
    try:
        import subprocess
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        info["nproc"] = None
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info["cpu_model"] = model
    share = os.environ.get("OMP_NUM_THREADS")
    usable = info["usable_cpus"] or 1
    info["thread_share"] = min(usable, int(share)) if share and share.isdigit() and int(share) > 0 else usable
    return info


def cpu_baseline(spec, cam, W, H, B, spp, cpu_spp, runs):
    """The oracle (scalar C++ restatement, one std::thread per host CPU of
    this job's share) on the same workload: the first cpu_spp samples of
    every pixel, forward + adjoint; median of `runs` timed runs after one
    warm-up."""
    import numpy as np
    import oracle
    host = host_cpu_info()
    threads = host["thread_share"]
    S = oracle.Scene(spec)
    n = cpu_spp * W * H
    dl = np.full((3, W * H), 0.01, np.float32)
    times = []
    for r in range(runs + 1):
        t0 = time.perf_counter()
        S.render_paths(cam, W, H, 1234, B, 0, n, threads=threads)
        t1 = time.perf_counter()
        S.render_adj(cam, W, H, 1234, B, spp, dl, first=0, count=n, threads=threads)
        t2 = time.perf_counter()
        if r > 0:
            times.append((t2 - t0, t1 - t0, t2 - t1))
    times.sort()
    tot, tf, ta = times[len(times) // 2]
    return dict(value=n / tot / 1e6, unit="Mpaths/s", cores=threads, kind="port",
                sample=f"{cpu_spp} spp x {W}x{H} = {n} paths of the same scene, depth {B + 1}, forward + adjoint "
                       f"(replay); median of {runs} runs after a warm-up: fwd {tf:.2f} s, adj {ta:.2f} s; "
                       f"{threads} threads = this job's CPU share (OMP_NUM_THREADS) of {host['usable_cpus']} "
                       f"usable CPUs",
                host=host)


def served_comparison(ctx, a):
    """Served (continuous batching) vs plain (one request per launch
    sequence) on the same request list: 512x512, depth 4, spp uniform in
    1..16, over resident scenes.  Saturation throughput (all requests queued
    at t = 0) for both, then p50 / p99 latency under Poisson arrivals at 0.75x
    the plain path's measured request rate."""
    import torch
    from tensorray_amd import scenes, serving
    specs = {"s1k": scenes.sphere_scene(subdiv=3, amplitude=0.05, noise_seed=2),
             "s20k": scenes.metric_scene(amplitude=0.05, noise_seed=1)}
    if a.serve_large_scenes:
        specs["s200k"] = scenes.sphere_scene(subdiv=7, amplitude=0.05, noise_seed=3)
        specs["s1m"] = scenes.sphere_scene(subdiv=8, amplitude=0.05, noise_seed=4)
    srv = serving.Server(ctx, budget_bytes=64 << 30, batch_paths=8 << 20, chunk_paths=4 << 20)
    for k, v in specs.items():
        srv.add_scene(k, v)
    names = list(specs)

    def reqs(rate, seed):
        return serving.make_requests(names, a.serve_requests, rate, seed=seed, res=(512, 512), spp=(1, 16),
                                     depth=(4, 4))

    # warm-up of both paths
    serving.run_plain(ctx, srv.scenes, reqs(0, 7)[:4], realtime=False)
    srv.run(reqs(0, 7)[:4], realtime=False)
    out = {}
    for mode in ("plain", "served"):
        rs = reqs(0, 2024)
        t0 = time.perf_counter()
        if mode == "plain":
            serving.run_plain(ctx, srv.scenes, rs, realtime=False)
        else:
            srv.run(rs, realtime=False)
        wall = time.perf_counter() - t0
        paths = sum(r.paths for r in rs)
        out[mode] = dict(mpaths_per_s=paths / wall / 1e6, wall_s=wall, requests=len(rs), paths=paths)
    rate = 0.75 * a.serve_requests / out["plain"]["wall_s"]
    for mode in ("plain", "served"):
        rs = reqs(rate, 2025)
        if mode == "plain":
            serving.run_plain(ctx, srv.scenes, rs, realtime=True)
        else:
            srv.run(rs, realtime=True)
        lat = serving.percentiles([r.latency * 1e3 for r in rs])
        out[mode].update(latency_ms_at_075_load=lat)
    out["poisson_rate_per_s"] = rate
    out["served_over_plain"] = out["served"]["mpaths_per_s"] / out["plain"]["mpaths_per_s"]
    out["workload"] = (f"{a.serve_requests} requests, 512x512, depth 4, spp U(1..16) (numpy PCG64 seed 2024), "
                       f"scenes {', '.join(f'{k}={v.num_triangles()} tris' for k, v in specs.items())}; "
                       f"HBM budget 64 GiB, batches of <= 8M paths")
    for ds in srv.scenes.values():
        ds.close()
    return out


def served_sweep(ctx, a, loads=(0.25, 0.5, 0.75, 1.0, 1.25)):
    """Serving over four resident scenes of 1,280 / 20,480 / 200,000 /
    1,003,520 triangles (displaced spheres on a ground plane): requests with
    resolution uniform over 64..1024, spp over 4..64, depth over 2..6
    (numpy PCG64 seed 2024).  Saturation throughput of plain and served, then
    p50 / p99 latency under Poisson arrivals at 0.25 .. 1.25 x the plain
    path's measured request rate."""
    from tensorray_amd import scenes, serving
    t0 = time.perf_counter()
    specs = {"s1k": scenes.sphere_scene(subdiv=3, amplitude=0.05, noise_seed=2),
             "s20k": scenes.sphere_scene(subdiv=5, amplitude=0.05, noise_seed=1),
             "s200k": scenes.sphere_scene(freq=100, amplitude=0.05, noise_seed=3),
             "s1m": scenes.sphere_scene(freq=224, amplitude=0.05, noise_seed=4)}
    srvs = {"fifo": serving.Server(ctx, budget_bytes=64 << 30, batch_paths=16 << 20, chunk_paths=16 << 20,
                                   small_paths=4 << 20),
            "srpt": serving.Server(ctx, budget_bytes=64 << 30, batch_paths=8 << 20, chunk_paths=8 << 20,
                                   order="srpt")}
    srv = srvs["fifo"]
    for k, v in specs.items():
        srv.add_scene(k, v)
    srvs["srpt"].scenes = srv.scenes
    build_s = time.perf_counter() - t0
    names = list(specs)
    n = a.serve_requests

    def reqs(rate, seed):
        return serving.make_requests(names, n, rate, seed=seed, res=(64, 1024), spp=(4, 64), depth=(2, 6))

    serving.run_plain(ctx, srv.scenes, reqs(0, 7)[:4], realtime=False)
    for sv in srvs.values():
        sv.run(reqs(0, 7)[:4], realtime=False)
    out = dict(scenes={k: v.num_triangles() for k, v in specs.items()}, scene_build_s=build_s,
               policies={"served_fifo": "oldest first; a request of more than 4M paths runs alone and whole, "
                                        "smaller ones are batched (up to 16M paths) with the small ones behind "
                                        "them", "served_srpt": "shortest remaining work first, 8M-path chunks "
                                                               "and steps"})
    modes = ("plain", "served_fifo", "served_srpt")

    def run_mode(mode, rs, realtime):
        if mode == "plain":
            serving.run_plain(ctx, srv.scenes, rs, realtime=realtime)
        else:
            srvs[mode.split("_")[1]].run(rs, realtime=realtime)

    for mode in modes:
        rs = reqs(0, 2024)
        t1 = time.perf_counter()
        run_mode(mode, rs, False)
        wall = time.perf_counter() - t1
        paths = sum(r.paths for r in rs)
        out[mode] = dict(mpaths_per_s=paths / wall / 1e6, wall_s=wall, requests=n, paths=paths)
    cap = n / out["plain"]["wall_s"]
    out["plain_capacity_req_per_s"] = cap
    for mode in modes[1:]:
        out[mode + "_over_plain"] = out[mode]["mpaths_per_s"] / out["plain"]["mpaths_per_s"]
    sweep = []
    for load in loads:
        row = dict(load=load, rate_per_s=load * cap)
        for mode in modes:
            rs = reqs(load * cap, 2025)
            t1 = time.perf_counter()
            run_mode(mode, rs, True)
            wall = time.perf_counter() - t1
            lat = serving.percentiles([r.latency * 1e3 for r in rs])
            row[mode] = dict(p50_ms=lat["p50"], p99_ms=lat["p99"],
                             mpaths_per_s=sum(r.paths for r in rs) / wall / 1e6)
        sweep.append(row)
    out["latency_sweep"] = sweep
    out["workload"] = (f"{n} requests per run over scenes {', '.join(f'{k}={v}' for k, v in out['scenes'].items())} "
                       f"tris; res U(64..1024), spp U(4..64), depth U(2..6), numpy PCG64 seed 2024; HBM budget 64 GiB")
    for ds in srv.scenes.values():
        ds.close()
    return out


def multiview_comparison(ctx, a):
    """Multi-view workload: views on a seeded Fibonacci sphere of radius 3
    around a 200,000-triangle displaced geodesic sphere with a 1024^2 Perlin
    albedo texture; resolutions uniform over 128..512, spp from {8, 16, 32},
    depth 4.  Tuned prepared batches captured in a graph vs the plain path
    (one forward + adjoint launch sequence per view, stream-ordered)."""
    import torch
    from tensorray_amd import _capi, multiview, scenes
    t0 = time.perf_counter()
    spec = scenes.multiview_scene(freq=100, amplitude=0.05, noise_seed=1, texture_res=1024, texture_seed=7)
    ds = _capi.DeviceScene(ctx, spec)
    build_s = time.perf_counter() - t0
    cams = scenes.fibonacci_cameras(a.multiview_views, radius=3.0, seed=0)
    views = multiview.random_views(cams, seed=0, res=(128, 512), spp_choices=(8, 16, 32), depth=4)
    mv = multiview.MultiView(ctx, ds, views)
    P = mv.total_paths
    tiles = mv.tune((64, 128, 256, 512, 1024), reps=2)
    mv.capture()
    stream = mv.stream
    for _ in range(2):
        mv.replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        mv.replay()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t1) * 1e3 / a.steps

    def plain_step():
        mv.clear_grads()
        for k, v in enumerate(views):
            cam, opt = mv.cams[k], mv.opts[k]
            ds.render_fwd_ckpt(cam, opt, mv.film[k].data_ptr(), mv.ckpt[k].data_ptr(), mv.ckpt[k].numel(), stream)
            mv.dldi[k].copy_(mv.film[k])
            ds.render_adj_ckpt(cam, opt, mv.ckpt[k].data_ptr(), mv.ckpt[k].numel(), mv.dldi[k].data_ptr(),
                               mv.dpos.data_ptr(), mv.dmat.data_ptr(), mv.dtex.data_ptr(), stream=stream)

    plain_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        plain_step()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t1) * 1e3 / a.steps
    out = dict(views=len(views), paths=P, triangles=ds.info().num_triangles, scene_build_s=build_s,
               tile_fwd=tiles[0], tile_adj=tiles[1],
               graph=dict(ms_per_step=graph_ms, mpaths_per_s=P / graph_ms / 1e3),
               plain=dict(ms_per_step=plain_ms, mpaths_per_s=P / plain_ms / 1e3),
               graph_over_plain=plain_ms / graph_ms,
               workload=f"{len(views)} views (Fibonacci sphere r=3, seed 0), res U(128..512), spp {{8,16,32}}, "
                        f"depth 4, 200,000-tri displaced geodesic sphere, 1024^2 Perlin albedo; forward + adjoint "
                        f"(checkpoint) of all views per step")
    mv.close()
    ds.close()
    return out


def streaming_copy_gbs(mib=2048, reps=5):
    """Achievable HBM bandwidth on this box: a device-to-device copy of
    `mib` MiB (read + write bytes over the best of `reps` HIP-event timings)."""
    import torch
    src = torch.empty(mib * (1 << 20) // 4, device="cuda")
    src.fill_(1.0)
    dst = torch.empty_like(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        e0.record()
        dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del src, dst
    return 2.0 * mib * (1 << 20) / (best * 1e-3) / 1e9


def boundary_terms(scene, cam, W, H, B, dldi, a):
    """Edge-sample throughput of the boundary integrators on the metric
    workload (512^2, depth 4, seed 1234): each class at `boundary_spp`
    samples per pixel (W * H * boundary_spp samples), dL/dI = the plain
    step's film minus the target.  Guided direct includes building its guide
    (g_direct_spp = 16 samples per secondary edge)."""
    import torch
    from tensorray_amd import _capi
    n = W * H * a.boundary_spp
    info = scene.info()
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    classes = [("primary", "primary", dict(sppe=a.boundary_spp)),
               ("pixel", "pixel", dict(sppe0=a.boundary_spp)),
               ("direct", "direct", dict(sppse0=a.boundary_spp)),
               ("direct_guided", "direct", dict(sppse0=a.boundary_spp, g_direct=True)),
               ("indirect", "indirect", dict(sppse1=a.boundary_spp))]
    out = dict(samples_per_class=n, workload=f"{W}x{H}, {a.boundary_spp} edge samples per pixel per class, "
                                             f"depth {B + 1}, metric scene")
    for name, kind, kw in classes:
        opt = _capi.make_options(seed=1234, max_bounces=B, **kw)
        scene.render_boundary(cam, opt, kind, dldi.data_ptr(), dpos.data_ptr(), stream)
        torch.cuda.synchronize()
        ms, wall = [], []
        for _ in range(a.steps):
            t0 = time.perf_counter()
            ev[0].record(stream)
            scene.render_boundary(cam, opt, kind, dldi.data_ptr(), dpos.data_ptr(), stream)
            ev[1].record(stream)
            torch.cuda.synchronize()
            wall.append((time.perf_counter() - t0) * 1e3)
            ms.append(ev[0].elapsed_time(ev[1]))
        t = sum(ms) / len(ms)
        out[name] = dict(ms=t, wall_ms=sum(wall) / len(wall), msamples_per_s=n / t / 1e3)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): the integrator renders on a single "
              f"MI355X; sharding paths or views over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    import torch
    from tensorray_amd import _capi, scenes

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    W, H, spp, B = a.width, a.height, a.spp, a.depth - 1
    cam_d = scenes.CAMERA_DEFAULT
    ctx = _capi.Context(0)
    spec = scenes.metric_scene(amplitude=0.05, noise_seed=1)
    tspec = scenes.metric_scene(amplitude=0.06, noise_seed=1)
    scene = _capi.DeviceScene(ctx, spec)
    tscene = _capi.DeviceScene(ctx, tspec)
    info = scene.info()
    cam = _capi.make_camera(W, H, **cam_d)
    opt = _capi.make_options(seed=1234, max_bounces=B, spp=spp)
    topt = _capi.make_options(seed=99, max_bounces=B, spp=spp)
    stream = torch.cuda.current_stream()
    P = W * H * spp
    target = torch.zeros(3, W * H, device="cuda")
    tscene.render_fwd(cam, topt, target.data_ptr(), stream)
    film = torch.zeros(3, W * H, device="cuda")
    paths = torch.empty(P, 3, device="cuda") if a.adjoint == "replay" else None
    nb = scene.checkpoint_bytes(cam, opt)
    ckpt = torch.empty(nb, dtype=torch.uint8, device="cuda") if a.adjoint == "checkpoint" else None
    dpos = torch.zeros(info.num_vertices, 3, device="cuda")
    dmat = torch.zeros(info.num_materials, 8, device="cuda")
    lib = ctx.lib
    torch.cuda.synchronize()
    stats = scene.render_stats(cam, opt, stream)

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    fwd_ms, adj_ms = [], []

    def step(timed):
        if timed:
            ev[0].record(stream)
        if a.adjoint == "checkpoint":
            scene.render_fwd_ckpt(cam, opt, film.data_ptr(), ckpt.data_ptr(), nb, stream)
        else:
            scene.render_paths(cam, opt, paths.data_ptr(), stream)
        if timed:
            ev[1].record(stream)
        if a.adjoint == "replay":
            # film = mean over samples (deterministic reduction kernel)
            scene.render_fwd_reduce(paths.data_ptr(), spp, W * H, film.data_ptr(), stream)
        dldi = film - target
        dpos.zero_()
        dmat.zero_()
        if timed:
            ev[2].record(stream)
        if a.adjoint == "checkpoint":
            scene.render_adj_ckpt(cam, opt, ckpt.data_ptr(), nb, dldi.data_ptr(), dpos.data_ptr(), dmat.data_ptr(),
                                  stream=stream)
        else:
            scene.render_adj(cam, opt, dldi.data_ptr(), dpos.data_ptr(), dmat.data_ptr(), stream)
        if timed:
            ev[3].record(stream)

    for _ in range(a.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step(True)
        torch.cuda.synchronize()
        fwd_ms.append(ev[0].elapsed_time(ev[1]))
        adj_ms.append(ev[2].elapsed_time(ev[3]))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / a.steps
    mpaths = P / (ms_per_step * 1e-3) / 1e6

    bb, st = algorithmic_bytes(stats, P, B, info.num_emitter_triangles)
    fwd_avg = sum(fwd_ms) / len(fwd_ms)
    adj_avg = sum(adj_ms) / len(adj_ms)
    ck = a.adjoint == "checkpoint"
    if adj_avg >= fwd_avg:
        kname, kms = ("k_render_adj_ckpt", adj_avg) if ck else ("k_render_adj", adj_avg)
        kbytes = (bb["adj_ckpt"] if ck else bb["adj"]) * P
    else:
        # the forward kernel with (true) or without (false) checkpoint writes
        kname, kms = f"k_render_paths<{'true' if ck else 'false'}>", fwd_avg
        kbytes = (bb["fwd_ckpt"] if ck else bb["fwd"]) * P
    achieved = kbytes / (kms * 1e-3)
    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                traffic = json.load(f).get(kname)
        except (OSError, ValueError):
            traffic = None
    copy_gbs = streaming_copy_gbs()
    # `achieved` / `frac` are the modeled (algorithmic) bytes over the kernel
    # time; `dram_gbs` / `dram_frac` are the DRAM bytes the counters measured
    # for the same kernel (FETCH_SIZE + WRITE_SIZE per launch, profiles/)
    # over the same time.  Most modeled bytes are BVH node reads that hit
    # in L2 / MALL, so the kernel is latency and issue bound, not DRAM bound.
    dram_gbs = traffic / (kms * 1e-3) / 1e9 if traffic else None
    roofline = dict(bound="hbm", achieved=achieved / 1e9, peak=MI355X_HBM_PEAK / 1e9, unit="GB/s",
                    achievable_copy=copy_gbs, frac_of_achievable=achieved / 1e9 / copy_gbs,
                    frac=achieved / MI355X_HBM_PEAK, traffic=traffic, dram_gbs=dram_gbs,
                    dram_frac=dram_gbs * 1e9 / MI355X_HBM_PEAK if dram_gbs else None,
                    achieved_is="modeled algorithmic bytes / kernel time; dram_gbs = measured DRAM bytes / "
                                "kernel time",
                    limiter="VALU issue under divergence and memory latency, not DRAM bandwidth: the "
                            "modeled bytes are mostly BVH reads served by L2 / MALL (dram_frac); lane "
                            "utilisation and issue counters in profiles/k_render_metric512.md",
                    kernel=kname, kernel_ms=kms,
                    algorithmic_bytes_per_launch=kbytes, fwd_kernel_ms=fwd_avg, adj_kernel_ms=adj_avg, **st)
    if a.copy_calibration:
        # a shader kernel of known traffic (reads MIB, writes MIB): torch's
        # elementwise add; a plain copy_ may run on a DMA engine, which the
        # cache counters do not see
        src = torch.ones(a.copy_calibration * (1 << 20) // 4, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.add(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    def optional(section, fn):
        """The side measurements never cost the headline line: a failure is
        reported in its field."""
        try:
            return fn(ctx, a)
        except Exception as e:  # noqa: BLE001
            print(f"bench.py: {section} failed: {e!r}", file=sys.stderr)
            return {"error": repr(e)}

    dldi = film - target
    boundary = (optional("boundary", lambda c, aa: boundary_terms(scene, cam, W, H, B, dldi, aa))
                if a.boundary_spp > 0 else None)
    served = optional("served", served_comparison) if a.serve_requests > 0 else None
    multi = optional("multiview", multiview_comparison) if a.multiview_views > 0 else None
    if a.no_serve_mix:
        sweep = None
    elif a.serve_sweep:
        sweep = optional("served_sweep", served_sweep)
    else:
        sweep = optional("served_sweep", lambda c, aa: served_sweep(c, aa, loads=(0.75,)))
    cpu = None
    if not a.no_cpu_baseline:
        try:
            cpu = cpu_baseline(spec, cam_d, W, H, B, spp, a.cpu_spp, a.cpu_runs)
        except Exception as e:  # noqa: BLE001
            print(f"bench.py: cpu baseline failed: {e!r}", file=sys.stderr)
    out = {
        "metric": "Mpaths/s fwd+adjoint (512², depth 4) & p99 latency, served vs plain, 1 GPU",
        "value": mpaths,
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded displaced icosphere scene)",
        "config": {"workload": f"plain: one request per launch sequence, {W}x{H}, {spp} spp, depth {B + 1}, "
                               f"{info.num_triangles} triangles (20,480-tri displaced icosphere + ground + light), "
                               f"forward + adjoint ({a.adjoint})", "mode": "plain", "adjoint": a.adjoint,
                   "paths_per_step": P},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "served": served,
        "multiview": multi,
        "boundary": boundary,
    }
    if sweep is not None:
        out["served_sweep"] = sweep
    print(json.dumps(out))
    scene.close()
    tscene.close()
    ctx.close()


if __name__ == "__main__":
    main()
