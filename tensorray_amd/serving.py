"""Serving loop: continuous batching of render-and-gradient requests on one
GPU.

The reference answers one synchronous request per call (TensorRay.cpp:
192-221); here a server keeps scenes resident in HBM and runs many requests
of different scenes, resolutions, sample counts and depths through the same
kernel launches (trc_render_batch):

* admission: a request is admitted when its transient buffers (per-path
  radiance, path checkpoint, images, gradient buffers) fit the HBM budget
  next to the requests already in flight; otherwise it waits in FIFO order;
* iteration-level scheduling: every step hands out a budget of
  `batch_paths` paths as chunks (path ranges) of the admitted requests'
  current phase -- forward, or adjoint once the forward is complete -- and
  launches one forward batch and one adjoint batch; requests are taken
  oldest first (order="fifo"; a request of more than `small_paths` paths
  at the head of the queue then runs alone and whole, small ones are
  batched with the small ones queued behind them) or
  least remaining work first (order="srpt": shortest remaining processing
  time, which lets small requests pass large ones; with `aging` > 0 the
  remaining work is discounted by the time waited, which bounds how long a
  large request can be passed); a request whose forward finishes in
  a step gets its film reduction and dL/dI = I - target on the stream and
  joins the adjoint batch of the same step; requests arriving meanwhile are
  admitted at the next step, so their work overlaps with requests already on
  the device;
* the loop never blocks on the GPU: at most `max_inflight` steps are queued
  on the stream, and while they run it keeps polling for arrivals and
  completions, so new requests join the next step;
* completion: an event per request; latency = completion - arrival.

Each path's result does not depend on the batch it ran in (counter-based
RNG, per-request buffers), so a request's film is bit-identical alone and in
a mix, and its gradients agree up to the order of float atomics.
"""
import collections
import dataclasses
import time
from typing import Dict, List, Optional

from tensorray_amd import _capi


@dataclasses.dataclass
class Request:
    scene: str
    width: int
    height: int
    spp: int
    max_bounces: int
    seed: int
    cam: dict                       # pos, target, up, fov
    target: object = None           # torch [3][H*W] on the device; dL/dI = I - target (zeros if None)
    arrival: float = 0.0            # seconds after the start of the run
    rid: int = -1
    # edge samples per pixel of the four boundary integrators (RenderOptions
    # sppe / sppe0 / sppse0 / sppse1); their gradients add into dpos next to
    # the interior adjoint, as the reference's backward sums every
    # integrator's renderD (renderer_single_view.py:74-76)
    sppe: int = 0
    sppe0: int = 0
    sppse0: int = 0
    sppse1: int = 0
    # results
    film: object = None             # torch [3][H*W]
    dpos: object = None             # torch [V][3]
    dmat: object = None             # torch [M][8]
    t_admit: Optional[float] = None
    t_done: Optional[float] = None

    @property
    def paths(self):
        return self.spp * self.width * self.height

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival

    def options(self):
        return _capi.make_options(seed=self.seed, max_bounces=self.max_bounces, spp=self.spp, sppe=self.sppe,
                                  sppse0=self.sppse0, sppse1=self.sppse1, sppe0=self.sppe0)

    def edge_terms(self):
        """(integrator, samples per pixel) of the boundary terms asked for."""
        return [(k, n) for k, n in (("primary", self.sppe), ("pixel", self.sppe0), ("direct", self.sppse0),
                                    ("indirect", self.sppse1)) if n > 0]


def run_edge_terms(ds, req, cam, opt, dldi_ptr, dpos_ptr, stream):
    """Adds the request's boundary-term gradients (each integrator's
    renderD) into dpos, stream-ordered after dL/dI."""
    for kind, _ in req.edge_terms():
        ds.render_boundary(cam, opt, kind, dldi_ptr, dpos_ptr, stream)


class _Active:
    """Device state of an admitted request."""

    def __init__(self, req, ds, torch, stream):
        self.req = req
        self.ds = ds
        info = ds.info()
        self.cam = _capi.make_camera(req.width, req.height, **req.cam)
        self.opt = req.options()
        P = req.paths
        self.nb = ds.checkpoint_bytes(self.cam, self.opt)
        with torch.cuda.stream(stream):
            self.paths = torch.empty(P, 3, device="cuda")
            self.ckpt = torch.empty(self.nb, dtype=torch.uint8, device="cuda")
            self.film = torch.empty(3, req.width * req.height, device="cuda")
            self.dldi = torch.empty(3, req.width * req.height, device="cuda")
            req.dpos = torch.zeros(info.num_vertices, 3, device="cuda")
            req.dmat = torch.zeros(info.num_materials, 8, device="cuda")
            self.vn = torch.zeros(info.num_vertices, 3, device="cuda")
        self.fwd_next = 0
        self.adj_next = 0
        self.done_event = None

    @staticmethod
    def bytes_needed(req, ds):
        info = ds.info()
        P = req.paths
        ck = P * (4 + 12 * (req.max_bounces + 1))
        img = 4 * 3 * req.width * req.height * 3          # film, dL/dI, target
        grads = 4 * (3 * info.num_vertices * 2 + 8 * info.num_materials)
        return 12 * P + ck + img + grads


class Server:
    """Continuous-batching server over resident scenes (one GPU, one stream)."""

    def __init__(self, ctx, budget_bytes=64 << 30, batch_paths=16 << 20, chunk_paths=16 << 20, max_inflight=2,
                 order="fifo", aging=0.0, small_paths=4 << 20):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.budget = int(budget_bytes)
        self.batch_paths = int(batch_paths)
        self.chunk_paths = int(chunk_paths)
        self.scenes: Dict[str, _capi.DeviceScene] = {}
        self.waiting = collections.deque()
        self.active: List[_Active] = []
        self.finishing: List[_Active] = []
        self.in_use = 0
        self.stream = torch.cuda.current_stream()
        self.t0 = None
        self.steps = 0
        self.peak_in_use = 0
        # at most max_inflight steps queued on the GPU, so that arrivals join
        # the next batch instead of queueing behind many launched ones
        self.max_inflight = max_inflight
        self.inflight = collections.deque()
        if order not in ("fifo", "srpt"):
            raise ValueError(f"unknown order {order!r}")
        self.order = order
        # srpt with aging: a request's remaining work is discounted by
        # `aging` paths per second it has waited since arrival, so that
        # large requests are not starved (0: pure shortest-remaining-first)
        self.aging = float(aging)
        # small_paths > 0: a step never mixes sizes.  A request of more than
        # small_paths paths at the head of the queue runs alone, whole (one
        # forward and one adjoint launch, as the plain path would run it);
        # a small one at the head is batched with the small requests queued
        # right behind it, up to batch_paths, stopping at the first large
        # one.  No request waits for a later one, and small requests share
        # launches instead of each paying a launch sequence that cannot fill
        # the GPU.  FIFO only (srpt preempts large requests by chunks); 0
        # hands one path budget per step to requests of any size.
        self.small_paths = int(small_paths) if order == "fifo" else 0

    def add_scene(self, name, spec):
        """Uploads a scene (BVH, triangles, textures) once; it stays resident."""
        self.scenes[name] = _capi.DeviceScene(self.ctx, spec)
        return self.scenes[name]

    def now(self):
        return time.perf_counter() - self.t0

    def submit(self, req):
        if req.scene not in self.scenes:
            raise KeyError(f"unknown scene {req.scene!r}")
        if req.sppe0 and (req.sppe0 * req.width * req.height) % 4:
            raise ValueError(f"request {req.rid}: sppe0 * W * H must be a multiple of 4")
        need = _Active.bytes_needed(req, self.scenes[req.scene])
        if need > self.budget:
            raise ValueError(f"request {req.rid} needs {need} bytes, more than the whole budget {self.budget}")
        self.waiting.append(req)

    def _admit(self):
        while self.waiting:
            req = self.waiting[0]
            ds = self.scenes[req.scene]
            need = _Active.bytes_needed(req, ds)
            if self.in_use + need > self.budget:
                break
            self.waiting.popleft()
            a = _Active(req, ds, self.torch, self.stream)
            a.bytes = need
            self.in_use += need
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            req.t_admit = self.now()
            self.active.append(a)

    def _collect(self):
        """Retires requests whose adjoint has completed."""
        still = []
        for a in self.finishing:
            if a.done_event.query():
                a.req.t_done = self.now()
                self.in_use -= a.bytes
                a.paths = a.ckpt = a.dldi = a.vn = None
            else:
                still.append(a)
        self.finishing = still

    def step(self):
        """One scheduling iteration; never blocks on the GPU.  Returns False
        when there is nothing to do, "wait" when the GPU already holds
        max_inflight steps (the caller polls again), True after launching."""
        torch = self.torch
        self._collect()
        self._admit()
        while self.inflight and self.inflight[0].query():
            self.inflight.popleft()
        if not self.active:
            return bool(self.finishing or self.waiting)
        if len(self.inflight) >= self.max_inflight:
            return "wait"
        self.steps += 1
        if self.order == "srpt":
            now = self.now()
            queue = sorted(self.active, key=lambda a: 2 * a.req.paths - a.fwd_next - a.adj_next
                           - self.aging * (now - a.req.arrival))
        else:
            queue = self.active
        # one budget of batch_paths per step, handed out in queue order: each
        # request takes a chunk of its current phase -- forward, or adjoint
        # once its forward is complete (a request whose forward completes in
        # this step can start its adjoint in the same step: the film and
        # dL/dI are computed on the stream between the two launches)
        fjobs, ajobs, fdone, adone, budget = [], [], [], [], self.batch_paths
        chunk = self.chunk_paths
        head_large = self.small_paths > 0 and queue[0].req.paths > self.small_paths
        if head_large:
            budget = chunk = 1 << 62
        for i, a in enumerate(queue):
            if budget <= 0:
                break
            if self.small_paths > 0 and i > 0 and (head_large or a.req.paths > self.small_paths):
                break
            P = a.req.paths
            if a.fwd_next < P:
                n = min(P - a.fwd_next, chunk, budget)
                fjobs.append(a.ds.job(a.cam, a.opt, a.fwd_next, n, paths=a.paths.data_ptr(),
                                      ckpt=a.ckpt.data_ptr()))
                a.fwd_next += n
                budget -= n
                if a.fwd_next < P:
                    continue
                fdone.append(a)
            if a.adj_next < P and budget > 0:
                n = min(P - a.adj_next, chunk, budget)
                ajobs.append(a.ds.job(a.cam, a.opt, a.adj_next, n, ckpt=a.ckpt.data_ptr(), dldi=a.dldi.data_ptr(),
                                      dpos=a.req.dpos.data_ptr(), dmat=a.req.dmat.data_ptr(),
                                      vn_scratch=a.vn.data_ptr()))
                a.adj_next += n
                budget -= n
                if a.adj_next >= P:
                    adone.append(a)
        if fjobs:
            self.ctx.render_batch(fjobs, adjoint=False, stream=self.stream)
        with torch.cuda.stream(self.stream):
            for a in fdone:
                a.ds.render_fwd_reduce(a.paths.data_ptr(), a.req.spp, a.req.width * a.req.height,
                                       a.film.data_ptr(), self.stream)
                if a.req.target is not None:
                    torch.sub(a.film, a.req.target, out=a.dldi)
                else:
                    a.dldi.copy_(a.film)
                a.req.film = a.film
                # boundary terms need only dL/dI: they run in this step,
                # beside the first chunk of the interior adjoint
                run_edge_terms(a.ds, a.req, a.cam, a.opt, a.dldi.data_ptr(), a.req.dpos.data_ptr(), self.stream)
        if ajobs:
            self.ctx.render_batch(ajobs, adjoint=True, stream=self.stream)
        for a in adone:
            a.done_event = torch.cuda.Event()
            a.done_event.record(self.stream)
            self.active.remove(a)
            self.finishing.append(a)
        ev = torch.cuda.Event()
        ev.record(self.stream)
        self.inflight.append(ev)
        return True

    def run(self, requests, realtime=True):
        """Serves `requests` (sorted by arrival); with realtime, a request is
        submitted once its arrival time has passed.  Returns the requests."""
        torch = self.torch
        reqs = sorted(requests, key=lambda r: r.arrival)
        torch.cuda.synchronize()
        self.t0 = time.perf_counter()
        i = 0
        while i < len(reqs) or self.waiting or self.active or self.finishing:
            t = self.now()
            while i < len(reqs) and (not realtime or reqs[i].arrival <= t):
                self.submit(reqs[i])
                i += 1
            busy = self.step()
            if not busy and i < len(reqs):
                # idle until the next arrival
                time.sleep(max(0.0, min(reqs[i].arrival - self.now(), 0.001)))
            elif busy == "wait" or (self.finishing and not self.active and not self.waiting):
                time.sleep(0.00005)
        torch.cuda.synchronize()
        self._collect()
        return reqs


def run_plain(ctx, scenes, requests, realtime=True):
    """The plain path: one request at a time, each a forward (with
    checkpoint) + dL/dI + adjoint launch sequence, completed before the next
    starts."""
    import torch
    reqs = sorted(requests, key=lambda r: r.arrival)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in reqs:
        if realtime:
            while time.perf_counter() - t0 < r.arrival:
                time.sleep(0.0002)
        ds = scenes[r.scene]
        info = ds.info()
        cam = _capi.make_camera(r.width, r.height, **r.cam)
        opt = r.options()
        nb = ds.checkpoint_bytes(cam, opt)
        ck = torch.empty(nb, dtype=torch.uint8, device="cuda")
        r.film = torch.empty(3, r.width * r.height, device="cuda")
        ds.render_fwd_ckpt(cam, opt, r.film.data_ptr(), ck.data_ptr(), nb, stream)
        dldi = r.film - r.target if r.target is not None else r.film.clone()
        r.dpos = torch.zeros(info.num_vertices, 3, device="cuda")
        r.dmat = torch.zeros(info.num_materials, 8, device="cuda")
        ds.render_adj_ckpt(cam, opt, ck.data_ptr(), nb, dldi.data_ptr(), r.dpos.data_ptr(), r.dmat.data_ptr(),
                           stream=stream)
        run_edge_terms(ds, r, cam, opt, dldi.data_ptr(), r.dpos.data_ptr(), stream)
        torch.cuda.synchronize()
        r.t_done = time.perf_counter() - t0
    return reqs


def percentiles(values, qs=(50, 99)):
    import numpy as np
    v = np.asarray(sorted(values), float)
    return {f"p{q}": float(np.percentile(v, q)) for q in qs}


def make_requests(scene_names, n, rate, seed=2024, res=(64, 1024), spp=(4, 64), depth=(2, 6), cams=None,
                  edge_spp=0):
    """Poisson arrivals (mean `rate` requests / s) of requests with
    resolution, spp and depth drawn uniformly (numpy PCG64(seed)); depth d
    means max_bounces d - 1.  edge_spp > 0 asks every request for the four
    boundary terms at that many edge samples per pixel (the pixel term
    rounded up so that sppe0 * W * H is a multiple of 4)."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        side = int(rng.integers(res[0], res[1] + 1))
        out.append(Request(scene=scene_names[int(rng.integers(len(scene_names)))], width=side, height=side,
                           spp=int(rng.integers(spp[0], spp[1] + 1)),
                           max_bounces=int(rng.integers(depth[0], depth[1] + 1)) - 1, seed=int(rng.integers(1 << 30)),
                           cam=dict(cams[0]) if cams else dict(pos=(0.0, 1.0, 3.0), target=(0.0, 0.0, 0.0),
                                                               up=(0.0, 1.0, 0.0), fov=45.0),
                           arrival=t, rid=i))
        if edge_spp > 0:
            r = out[-1]
            r.sppe = r.sppse0 = r.sppse1 = edge_spp
            r.sppe0 = edge_spp
            while (r.sppe0 * side * side) % 4:
                r.sppe0 += 1
    return out
