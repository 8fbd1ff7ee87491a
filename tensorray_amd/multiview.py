"""Multi-view rendering: many camera views of one scene, of uneven
resolution and sample count, traced forward and adjoint in one persistent
launch each.

The reference renders views one at a time, one forward + adjoint launch
sequence per view (renderer_multi_view.py:55-78, RenderFunctionMultiView).
Here all views are the jobs of two prepared batches (trc_plan_*): the
forward plan writes every view's per-path radiance and path checkpoint, a
film reduction per view and dL/dI follow, and the adjoint plan accumulates
every view's gradient into one [V][3] / [M][8] / texel buffer.  Tile sizes
are picked by the library's tuner, and the whole step (forward, films,
dL/dI, gradient clear, adjoint) is captured into one graph and replayed.
"""
import dataclasses
from typing import List, Optional

from tensorray_amd import _capi


@dataclasses.dataclass
class View:
    width: int
    height: int
    spp: int
    max_bounces: int
    seed: int
    cam: dict                        # pos, target, up, fov

    @property
    def paths(self):
        return self.spp * self.width * self.height


def random_views(cams, seed=0, res=(128, 512), spp_choices=(8, 16, 32), depth=4):
    """Views for camera dicts `cams`: square resolution uniform over `res`
    (inclusive), spp drawn from `spp_choices`, depth d = max_bounces + 1
    (numpy PCG64(seed))."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k, c in enumerate(cams):
        side = int(rng.integers(res[0], res[1] + 1))
        out.append(View(width=side, height=side, spp=int(spp_choices[int(rng.integers(len(spp_choices)))]),
                        max_bounces=depth - 1, seed=1000 + k, cam=dict(c)))
    return out


class MultiView:
    """All views of one resident scene in one forward and one adjoint launch.

    targets: optional list of [3][H*W] device tensors; dL/dI = I - target
    (I itself when no target is given).  Gradients of all views sum into
    `dpos` ([V][3]), `dmat` ([M][8]) and `dtex` (texels)."""

    def __init__(self, ctx, ds, views: List[View], targets: Optional[list] = None, stream=None):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.ds = ds
        self.views = list(views)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        info = ds.info()
        self.info = info
        dev = "cuda"
        self.cams = [_capi.make_camera(v.width, v.height, **v.cam) for v in self.views]
        self.opts = [_capi.make_options(seed=v.seed, max_bounces=v.max_bounces, spp=v.spp) for v in self.views]
        self.paths, self.ckpt, self.film, self.dldi, self.target = [], [], [], [], []
        for k, v in enumerate(self.views):
            nb = ds.checkpoint_bytes(self.cams[k], self.opts[k])
            self.paths.append(torch.empty(v.paths, 3, device=dev))
            self.ckpt.append(torch.empty(nb, dtype=torch.uint8, device=dev))
            self.film.append(torch.empty(3, v.width * v.height, device=dev))
            self.dldi.append(torch.empty(3, v.width * v.height, device=dev))
            self.target.append(targets[k] if targets is not None else None)
        self.dpos = torch.zeros(info.num_vertices, 3, device=dev)
        self.dmat = torch.zeros(info.num_materials, 8, device=dev)
        self.dtex = torch.zeros(max(1, info.texture_floats), device=dev)
        self.vn = torch.zeros(info.num_vertices, 3, device=dev)
        fjobs, ajobs = [], []
        for k, v in enumerate(self.views):
            cam, opt = self.cams[k], self.opts[k]
            fjobs.append(ds.job(cam, opt, 0, v.paths, paths=self.paths[k].data_ptr(), ckpt=self.ckpt[k].data_ptr()))
            ajobs.append(ds.job(cam, opt, 0, v.paths, ckpt=self.ckpt[k].data_ptr(), dldi=self.dldi[k].data_ptr(),
                                dpos=self.dpos.data_ptr(), dmat=self.dmat.data_ptr(),
                                dtex=self.dtex.data_ptr() if info.texture_floats else None,
                                vn_scratch=self.vn.data_ptr()))
        self.fwd = ctx.plan(fjobs, adjoint=False)
        self.adj = ctx.plan(ajobs, adjoint=True)
        self.graph = None

    @property
    def total_paths(self):
        return sum(v.paths for v in self.views)

    def tune(self, candidates=(64, 128, 256, 512, 1024), reps=3):
        """Picks the forward and adjoint tile sizes (the adjoint's tuning
        launches accumulate into the gradients, which are cleared after)."""
        bf = self.fwd.tune(candidates, reps=reps, stream=self.stream)
        self.films()  # the adjoint tunes on this step's dL/dI
        ba = self.adj.tune(candidates, reps=reps, stream=self.stream)
        with self.torch.cuda.stream(self.stream):
            self.clear_grads()
        return bf, ba

    def films(self):
        """Film reduction and dL/dI of every view (stream-ordered)."""
        torch = self.torch
        with torch.cuda.stream(self.stream):
            for k, v in enumerate(self.views):
                self.ds.render_fwd_reduce(self.paths[k].data_ptr(), v.spp, v.width * v.height,
                                          self.film[k].data_ptr(), self.stream)
                if self.target[k] is not None:
                    torch.sub(self.film[k], self.target[k], out=self.dldi[k])
                else:
                    self.dldi[k].copy_(self.film[k])

    def clear_grads(self):
        self.dpos.zero_()
        self.dmat.zero_()
        self.dtex.zero_()

    def step_eager(self):
        """Forward of all views, films, dL/dI, cleared gradients, adjoint."""
        self.fwd.launch(stream=self.stream)
        self.films()
        with self.torch.cuda.stream(self.stream):
            self.clear_grads()
        self.adj.launch(stream=self.stream)

    def capture(self):
        """Captures step_eager into a graph (launch paths are warmed first)."""
        torch = self.torch
        self.step_eager()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream()
        saved = self.stream
        with torch.cuda.graph(g, stream=cap):
            self.stream = torch.cuda.current_stream()
            try:
                self.step_eager()
            finally:
                self.stream = saved
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()

    def close(self):
        self.graph = None
        self.fwd.close()
        self.adj.close()
