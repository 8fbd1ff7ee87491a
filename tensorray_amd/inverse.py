"""Inverse rendering: recover a mesh's vertex positions from images.

The reference's loop (optimize_multi_view.py:116-160) renders one view at a
time through its autograd Function (renderer_multi_view.py:41-105), steps
its optimizer and calls scene.configure(), which rebuilds the whole OptiX
acceleration structure every iteration (Scene.cpp:522-533); its C++
examples use Adam (Optimizer.h:66-123).  Here every step

* renders each view forward with a path checkpoint (interior term),
* forms the L2 image loss and dL/dI = 2 (I - T) / (3 W H) on the device,
* runs the interior adjoint from the checkpoint and the boundary terms
  (primary edges and the guided direct secondary edges by default, the
  reference's edge integrators in its examples), all accumulating into one
  [V][3] gradient,
* takes one device Adam step on the vertex positions (trc_adam_step), and
* updates the scene on the device (trc_scene_set_vertices with refit: face
  and vertex normals, emitter table, secondary edges and the BVH's node
  bounds recomputed on the GPU, same topology); the library rebuilds the
  tree once its SAH cost has grown past 1.25x its cost at the last build.

Each phase is timed (HIP events on the stream for device work, the host
clock around the scene update, which ends with a synchronisation), the
refit and rebuild counts are reported, and so is the vertex error against
the target mesh.
"""
import argparse
import dataclasses
import json
import sys
import time
from typing import List

from tensorray_amd import _capi, scenes


@dataclasses.dataclass
class Setup:
    freq: int = 316                 # geodesic frequency: 20 * freq^2 triangles (316: 1,997,120)
    amplitude: float = 0.1          # target displacement
    noise_seed: int = 1
    views: int = 16
    res: int = 512
    spp: int = 256
    depth: int = 8                  # max_bounces = depth - 1
    sppe: int = 16                  # primary-edge samples per pixel (0 skips)
    sppse0: int = 16                # direct secondary-edge samples per pixel (0 skips)
    sppse1: int = 0                 # indirect secondary-edge samples per pixel
    sppe0: int = 0                  # pixel-boundary samples per pixel
    guiding: bool = True            # guiding of the direct secondary-edge term
    share_guide: bool = True        # one guide per step for all views (trc_options.g_keep); 0: one per view
    target_spp: int = 256
    lr: float = 1e-3
    steps: int = 100
    optimizer: str = "adam"         # "adam" on the positions, or "largesteps" (uniform Adam on u = (I + lmbda L) x)
    lmbda: float = 20.0             # largesteps: weight of the uniform Laplacian


def shape_scene(pos, idx):
    """Lambertian (albedo 0.5) mesh lit by two 1x1 area lights (radiance
    10) at y = +2 and y = -2 facing it, so that views from every side see a
    lit surface."""
    spec = scenes.SceneSpec()
    spec.materials.append(scenes.Material(scenes.MAT_DIFFUSE, albedo=(0.5, 0.5, 0.5)))
    spec.materials.append(scenes.Material(scenes.MAT_DIFFUSE, albedo=(0.0, 0.0, 0.0)))
    spec.meshes.append(scenes.Mesh(pos, idx, material=0, flags=scenes.MESH_DIFF))
    top, ti, _ = scenes.quad_y(2.0, -0.5, 0.5, -0.5, 0.5, up=False)
    bot, bi, _ = scenes.quad_y(-2.0, -0.5, 0.5, -0.5, 0.5, up=True)
    spec.meshes.append(scenes.Mesh(top, ti, material=1, emit=(10.0, 10.0, 10.0)))
    spec.meshes.append(scenes.Mesh(bot, bi, material=1, emit=(10.0, 10.0, 10.0)))
    return spec


class InverseRendering:
    """Adam on the vertex positions of `init_spec`'s differentiable mesh
    against target images of `target_spec` (same topology)."""

    def __init__(self, ctx, init_spec, target_spec, cams: List[dict], setup: Setup, stream=None, log=None):
        import torch
        say = log or (lambda msg: None)
        self.torch = torch
        self.ctx = ctx
        self.s = setup
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self.ds = _capi.DeviceScene(ctx, init_spec)
        say("initial scene resident")
        info = self.ds.info()
        self.V = info.num_vertices
        self.max_edges = 3 * info.num_triangles
        W = H = setup.res
        self.cams = [_capi.make_camera(W, H, **c) for c in cams]
        # target images (rendered once from the target mesh, own seed)
        tds = _capi.DeviceScene(ctx, target_spec)
        say("target scene resident")
        self.targets = []
        for k, cam in enumerate(self.cams):
            t = torch.empty(3, W * H, device="cuda")
            topt = _capi.make_options(seed=99_000 + k, max_bounces=setup.depth - 1, spp=setup.target_spp)
            tds.render_fwd(cam, topt, t.data_ptr(), self.stream)
            self.targets.append(t)
            torch.cuda.synchronize()
            say(f"target view {k} rendered")
        self.n_mesh = len(init_spec.meshes[0].pos)   # the optimized mesh comes first
        self.target_pos = torch.from_numpy(tds.vertices()[:self.n_mesh]).cuda()
        torch.cuda.synchronize()
        tds.close()
        self.x = torch.from_numpy(self.ds.vertices()).cuda().contiguous()   # [V][3]
        self.grad = torch.zeros_like(self.x)
        if setup.optimizer == "adam":
            self.m = torch.zeros_like(self.x)
            self.v = torch.zeros_like(self.x)
        elif setup.optimizer == "largesteps":
            # LargeSteps (pyTensorRay/optimizer.py:147-180): the optimized variable
            # is u = M x, M = I + lmbda L on the mesh; gradients map as M^-1 g and
            # positions come back as x = M^-1 u
            self.solver = ctx.lsolver(init_spec.meshes[0].idx, self.n_mesh, setup.lmbda)
            self.u = torch.empty(self.n_mesh, 3, device="cuda")
            self.solver.apply(self.x.data_ptr(), self.u.data_ptr(), self.stream)
            self.gu = torch.zeros_like(self.u)
            self.m = torch.zeros_like(self.u)
            self.v = torch.zeros_like(self.u)
        else:
            raise ValueError(f"unknown optimizer {setup.optimizer!r}")
        opt0 = self._options(0, 0)
        self.ckpt = torch.empty(self.ds.checkpoint_bytes(self.cams[0], opt0), dtype=torch.uint8, device="cuda")
        self.film = torch.empty(3, W * H, device="cuda")
        self.dldi = torch.empty(3, W * H, device="cuda")
        self.steps_done = 0
        self.log = []

    def _options(self, step, view):
        s = self.s
        # one guiding leaf per secondary edge: room for every edge of the mesh.
        # The guide does not depend on the camera: with share_guide the views
        # of a step share one guide, built from the step's seed by the first.
        return _capi.make_options(seed=1234 + 1000 * step + view, max_bounces=s.depth - 1, spp=s.spp, sppe=s.sppe,
                                  sppse0=s.sppse0, sppse1=s.sppse1, sppe0=s.sppe0, g_direct=s.guiding,
                                  g_direct_max_size=self.max_edges, g_keep=s.share_guide,
                                  g_seed=1234 + 1000 * step)

    def vertex_error(self):
        """Mean Euclidean distance of the mesh vertices to the target's."""
        return (self.x[:self.n_mesh] - self.target_pos).norm(dim=1).mean().item()

    def radial_error(self):
        """Mean distance along the target's radial direction (the direction
        of its displacement): the shape error without the tangential drift
        of vertices, which moves no surface."""
        t = self.target_pos
        n = t / t.norm(dim=1, keepdim=True)
        return ((self.x[:self.n_mesh] - t) * n).sum(dim=1).abs().mean().item()

    def step(self):
        torch = self.torch
        st = self.stream
        t = self.steps_done + 1
        ev = {k: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for k in ("forward", "loss", "adjoint", "boundary", "adam")}
        acc = {k: 0.0 for k in ev}
        loss = torch.zeros((), device="cuda")
        W = H = self.s.res
        scale = 2.0 / (3 * W * H)
        with torch.cuda.stream(st):
            self.grad.zero_()
        t0 = time.perf_counter()
        for k, cam in enumerate(self.cams):
            opt = self._options(t, k)
            ev["forward"][0].record(st)
            self.ds.render_fwd_ckpt(cam, opt, self.film.data_ptr(), self.ckpt.data_ptr(), self.ckpt.numel(), st)
            ev["forward"][1].record(st)
            ev["loss"][0].record(st)
            with torch.cuda.stream(st):
                torch.sub(self.film, self.targets[k], out=self.dldi)
                loss += self.dldi.square().sum() / (3 * W * H)
                self.dldi.mul_(scale)
            ev["loss"][1].record(st)
            ev["adjoint"][0].record(st)
            self.ds.render_adj_ckpt(cam, opt, self.ckpt.data_ptr(), self.ckpt.numel(), self.dldi.data_ptr(),
                                    self.grad.data_ptr(), stream=st)
            ev["adjoint"][1].record(st)
            ev["boundary"][0].record(st)
            for kind, n in (("primary", self.s.sppe), ("pixel", self.s.sppe0), ("direct", self.s.sppse0),
                            ("indirect", self.s.sppse1)):
                if n > 0:
                    self.ds.render_boundary(cam, opt, kind, self.dldi.data_ptr(), self.grad.data_ptr(), st)
            ev["boundary"][1].record(st)
            torch.cuda.synchronize()
            for name, (a, b) in ev.items():
                if name != "adam":
                    acc[name] += a.elapsed_time(b)
        ev["adam"][0].record(st)
        if self.s.optimizer == "adam":
            self.ctx.adam_step(self.x.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                               self.x.numel(), t, lr=self.s.lr, stream=st)
        else:
            # mesh vertices are the first n_mesh rows of x and grad
            self.solver.solve(self.grad.data_ptr(), self.gu.data_ptr(), max_iter=400, tol=1e-5, stream=st)
            self.ctx.adam_uniform_step(self.u.data_ptr(), self.gu.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                                       self.u.numel(), t, lr=self.s.lr, stream=st)
            self.solver.solve(self.u.data_ptr(), self.x.data_ptr(), max_iter=400, tol=1e-6, stream=st)
        ev["adam"][1].record(st)
        torch.cuda.synchronize()
        acc["adam"] = ev["adam"][0].elapsed_time(ev["adam"][1])
        r0 = time.perf_counter()
        self.ds.set_vertices(self.x.data_ptr(), self.V, refit=True, stream=st)
        torch.cuda.synchronize()
        acc["refit"] = (time.perf_counter() - r0) * 1e3
        info = self.ds.info()
        acc["total"] = (time.perf_counter() - t0) * 1e3
        self.steps_done = t
        rec = dict(step=t, loss=loss.item(), vertex_error=self.vertex_error(), radial_error=self.radial_error(),
                   bvh_sah=round(info.bvh_sah_cost, 4), bvh_rebuilds=info.bvh_rebuilds,
                   ms={k: round(v, 3) for k, v in acc.items()})
        self.log.append(rec)
        return rec

    def close(self):
        if getattr(self, "solver", None) is not None:
            self.solver.close()
        self.ds.close()


def make_problem(ctx, s: Setup, stream=None, log=None):
    say = log or (lambda msg: None)
    pos, idx = scenes.geodesic_sphere(s.freq)
    say(f"mesh: {len(idx)} triangles, {len(pos)} vertices")
    target = scenes.displace(pos, s.amplitude, s.noise_seed)
    say("target displaced")
    cams = scenes.fibonacci_cameras(s.views, radius=3.0, seed=0)
    return InverseRendering(ctx, shape_scene(pos, idx), shape_scene(target, idx), cams, s, stream, log=say)


def main(argv=None):
    p = argparse.ArgumentParser(description="Inverse rendering of a displaced geodesic sphere from multi-view images")
    d = Setup()
    for f in dataclasses.fields(Setup):
        if f.type is bool or isinstance(getattr(d, f.name), bool):
            p.add_argument("--" + f.name.replace("_", "-"), type=int, default=int(getattr(d, f.name)))
        else:
            p.add_argument("--" + f.name.replace("_", "-"), type=type(getattr(d, f.name)), default=getattr(d, f.name))
    a = p.parse_args(argv)
    s = Setup(**{f.name: (bool(getattr(a, f.name)) if isinstance(getattr(d, f.name), bool) else getattr(a, f.name))
                 for f in dataclasses.fields(Setup)})
    import torch
    ctx = _capi.Context(0)
    t0 = time.perf_counter()
    prob = make_problem(ctx, s, log=lambda msg: print(json.dumps(dict(setup_progress=msg,
                                                                      t_s=round(time.perf_counter() - t0, 2))),
                                                     flush=True))
    setup_s = time.perf_counter() - t0
    e0, r0 = prob.vertex_error(), prob.radial_error()
    print(json.dumps(dict(setup=dataclasses.asdict(s), triangles=int(prob.ds.info().num_triangles),
                          vertices=prob.V, setup_s=round(setup_s, 2), initial_vertex_error=e0,
                          initial_radial_error=r0)), flush=True)
    for _ in range(s.steps):
        print(json.dumps(prob.step()), flush=True)
    steps = prob.log
    keys = steps[0]["ms"].keys()
    mean = {k: sum(r["ms"][k] for r in steps) / len(steps) for k in keys}
    paths = s.views * s.res * s.res * s.spp
    print(json.dumps(dict(summary=True, steps=len(steps), initial_vertex_error=e0,
                          final_vertex_error=steps[-1]["vertex_error"], initial_radial_error=r0,
                          final_radial_error=steps[-1]["radial_error"], final_loss=steps[-1]["loss"],
                          mean_ms=mean, paths_per_step=paths,
                          interior_mpaths_per_s=paths / ((mean["forward"] + mean["adjoint"]) * 1e-3) / 1e6)),
          flush=True)
    prob.close()
    torch.cuda.synchronize()
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
