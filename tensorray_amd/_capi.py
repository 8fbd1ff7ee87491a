"""ctypes binding of the C ABI in include/trc.h (libtrc.so).

This is the thin Python edge over the native library: device buffers come
from PyTorch-ROCm tensors (passed by `data_ptr()`), streams from
`torch.cuda.current_stream().cuda_stream`.  There is no CPU fallback: if the
library or a GPU is missing every entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrc.so")

TRC_OK = 0
STATUS_NAMES = {1: "TRC_ERR_INVALID", 2: "TRC_ERR_HIP", 3: "TRC_ERR_NOMEM", 4: "TRC_ERR_STATE",
                5: "TRC_ERR_OVERFLOW"}

c_float_p = ctypes.POINTER(ctypes.c_float)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)
c_i32_p = ctypes.POINTER(ctypes.c_int32)
vp = ctypes.c_void_p


class trc_camera(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("target", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("fov", ctypes.c_float)]


class trc_options(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint32), ("max_bounces", ctypes.c_int32), ("spp", ctypes.c_int32),
                ("sppe", ctypes.c_int32), ("sppse0", ctypes.c_int32), ("sppse1", ctypes.c_int32),
                ("sppe0", ctypes.c_int32), ("quiet", ctypes.c_int32), ("g_direct", ctypes.c_int32),
                ("g_direct_depth", ctypes.c_int32), ("g_direct_max_size", ctypes.c_int32),
                ("g_direct_spp", ctypes.c_int32), ("g_direct_thold", ctypes.c_float), ("g_eps", ctypes.c_float),
                ("g_keep", ctypes.c_int32), ("g_seed", ctypes.c_uint32)]


class trc_job(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("cam", trc_camera), ("opt", trc_options),
                ("first_path", ctypes.c_uint64), ("num_paths", ctypes.c_uint64), ("paths", ctypes.c_void_p),
                ("ckpt", ctypes.c_void_p), ("dLdI", ctypes.c_void_p), ("dpos", ctypes.c_void_p),
                ("dmat", ctypes.c_void_p), ("dtex", ctypes.c_void_p), ("vn_scratch", ctypes.c_void_p)]


class trc_scene_info(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_int32), ("num_triangles", ctypes.c_int32),
                ("num_materials", ctypes.c_int32), ("num_emitter_triangles", ctypes.c_int32),
                ("num_edges", ctypes.c_int32), ("bvh_nodes", ctypes.c_int32), ("bvh_leaves", ctypes.c_int32),
                ("bvh_max_depth", ctypes.c_int32), ("bvh_stack_need", ctypes.c_int32),
                ("bvh_sah_cost", ctypes.c_double), ("device_bytes", ctypes.c_int64),
                ("texture_floats", ctypes.c_int64), ("bvh_refits", ctypes.c_int32),
                ("bvh_rebuilds", ctypes.c_int32)]


# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("trc_context_create", ctypes.c_int32, [ctypes.c_int32, ctypes.POINTER(vp)]),
    ("trc_context_create_host", ctypes.c_int32, [ctypes.POINTER(vp)]),
    ("trc_context_destroy", ctypes.c_int32, [vp]),
    ("trc_last_error", ctypes.c_char_p, [vp]),
    ("trc_device_cus", ctypes.c_int32, [vp]),
    ("trc_scene_create", ctypes.c_int32, [vp, ctypes.POINTER(vp)]),
    ("trc_scene_destroy", ctypes.c_int32, [vp]),
    ("trc_scene_add_material", ctypes.c_int32, [vp, ctypes.c_int32, c_float_p, c_float_p, ctypes.c_float,
                                                ctypes.c_int32]),
    ("trc_scene_add_texture", ctypes.c_int32, [vp, c_float_p, ctypes.c_int32, ctypes.c_int32]),
    ("trc_scene_add_mesh", ctypes.c_int32, [vp, c_float_p, ctypes.c_int32, c_u32_p, ctypes.c_int32, c_float_p,
                                            ctypes.c_int32, c_u32_p, ctypes.c_int32, c_float_p, ctypes.c_int32]),
    ("trc_scene_configure", ctypes.c_int32, [vp, vp]),
    ("trc_scene_prepare", ctypes.c_int32, [vp]),
    ("trc_scene_set_vertices", ctypes.c_int32, [vp, vp, ctypes.c_int32, ctypes.c_int32, vp]),
    ("trc_scene_set_material", ctypes.c_int32, [vp, ctypes.c_int32, c_float_p, c_float_p, ctypes.c_float]),
    ("trc_scene_get_info", ctypes.c_int32, [vp, ctypes.POINTER(trc_scene_info)]),
    ("trc_scene_get_vertices", ctypes.c_int32, [vp, c_float_p]),
    ("trc_scene_get_indices", ctypes.c_int32, [vp, c_u32_p]),
    ("trc_scene_get_bvh", ctypes.c_int32, [vp, c_u32_p, c_u32_p]),
    ("trc_scene_get_edges", ctypes.c_int32, [vp, c_i32_p]),
    ("trc_intersect", ctypes.c_int32, [vp, vp, ctypes.c_int32, vp, vp, ctypes.c_int32, vp]),
    ("trc_render_fwd", ctypes.c_int32, [vp, ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options), vp, vp]),
    ("trc_render_fwd_noalloc", ctypes.c_int32, [vp, ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options), vp,
                                                vp]),
    ("trc_render_paths", ctypes.c_int32, [vp, ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options), vp, vp]),
    ("trc_film_from_paths", ctypes.c_int32, [vp, vp, ctypes.c_int32, ctypes.c_int32, vp, vp]),
    ("trc_render_stats", ctypes.c_int32, [vp, ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options),
                                          ctypes.POINTER(ctypes.c_uint64), vp]),
    ("trc_render_boundary", ctypes.c_int32, [vp, ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options),
                                             ctypes.c_int32, vp, vp, vp]),
    ("trc_render_deriv", ctypes.c_int32, [vp, ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options),
                                          ctypes.c_int32, vp, vp, vp, vp]),
    ("trc_render_adj", ctypes.c_int32, [vp, ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options), vp, vp, vp,
                                        vp, vp]),
    ("trc_checkpoint_bytes", ctypes.c_int64, [ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options)]),
    ("trc_render_batch", ctypes.c_int32, [vp, ctypes.POINTER(trc_job), ctypes.c_int32, ctypes.c_int32, vp]),
    ("trc_render_fwd_ckpt", ctypes.c_int32, [vp, ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options), vp, vp,
                                             ctypes.c_int64, vp]),
    ("trc_render_adj_ckpt", ctypes.c_int32, [vp, ctypes.POINTER(trc_camera), ctypes.POINTER(trc_options), vp,
                                             ctypes.c_int64, vp, vp, vp, vp, vp]),
    ("trc_plan_create", ctypes.c_int32, [vp, ctypes.POINTER(trc_job), ctypes.c_int32, ctypes.c_int32,
                                         ctypes.POINTER(vp)]),
    ("trc_plan_set_tile", ctypes.c_int32, [vp, ctypes.c_uint32]),
    ("trc_plan_tile", ctypes.c_uint32, [vp]),
    ("trc_plan_launch", ctypes.c_int32, [vp, vp]),
    ("trc_plan_tune", ctypes.c_int32, [vp, c_u32_p, ctypes.c_int32, ctypes.c_int32, c_u32_p, vp]),
    ("trc_plan_destroy", ctypes.c_int32, [vp]),
    ("trc_adam_step", ctypes.c_int32, [vp, vp, vp, vp, vp, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float, vp]),
    ("trc_adam_uniform_step", ctypes.c_int32, [vp, vp, vp, vp, vp, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float,
                                               ctypes.c_float, ctypes.c_float, vp]),
    ("trc_lsolver_create", ctypes.c_int32, [vp, c_u32_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                            ctypes.POINTER(vp)]),
    ("trc_lsolver_apply", ctypes.c_int32, [vp, vp, vp, vp]),
    ("trc_lsolver_solve", ctypes.c_int32, [vp, vp, vp, ctypes.c_int32, ctypes.c_float, vp]),
    ("trc_lsolver_status", ctypes.c_int32, [vp, ctypes.POINTER(ctypes.c_int32), c_float_p]),
    ("trc_lsolver_destroy", ctypes.c_int32, [vp]),
]

_lib = None


def load_library(path=LIB_PATH):
    """Load libtrc.so and declare every entry point (no device calls)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"native library {path} is missing: run tensorray_amd.build.build() first")
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _fp(a):
    return a.ctypes.data_as(c_float_p) if a is not None else None


def _up(a):
    return a.ctypes.data_as(c_u32_p) if a is not None else None


class TRCError(RuntimeError):
    pass


class Context:
    """A device context (EnvCreate of the reference)."""

    def __init__(self, device=0, host_only=False):
        """host_only: no device (scene assembly and validation only)."""
        self.lib = load_library()
        self.host_only = host_only
        h = vp()
        if host_only:
            st = self.lib.trc_context_create_host(ctypes.byref(h))
        else:
            st = self.lib.trc_context_create(device, ctypes.byref(h))
        self.h = h
        if st != TRC_OK:
            msg = self.lib.trc_last_error(h).decode() if h.value else "no context"
            raise TRCError(f"trc_context_create: {STATUS_NAMES.get(st, st)}: {msg}")

    def check(self, st, what):
        if st != TRC_OK:
            raise TRCError(f"{what}: {STATUS_NAMES.get(st, st)}: {self.lib.trc_last_error(self.h).decode()}")

    @property
    def num_cus(self):
        return self.lib.trc_device_cus(self.h)

    def render_batch(self, jobs, adjoint, stream=None):
        """jobs: list of trc_job."""
        arr = (trc_job * len(jobs))(*jobs)
        self.check(self.lib.trc_render_batch(self.h, arr, len(jobs), 1 if adjoint else 0, _stream_ptr(stream)),
                   "trc_render_batch")

    def adam_step(self, x_ptr, grad_ptr, m_ptr, v_ptr, n, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-6,
                  stream=None):
        """One Adam step on n device floats (trc_adam_step)."""
        self.check(self.lib.trc_adam_step(self.h, x_ptr, grad_ptr, m_ptr, v_ptr, int(n), int(step), lr, beta1, beta2,
                                          eps, _stream_ptr(stream)), "trc_adam_step")

    def adam_uniform_step(self, x_ptr, grad_ptr, m_ptr, v_ptr, n, step, lr=1e-3, beta1=0.9, beta2=0.999,
                          stream=None):
        """One uniform-Adam step on n device floats (trc_adam_uniform_step)."""
        self.check(self.lib.trc_adam_uniform_step(self.h, x_ptr, grad_ptr, m_ptr, v_ptr, int(n), int(step), lr, beta1,
                                                  beta2, _stream_ptr(stream)), "trc_adam_uniform_step")

    def lsolver(self, faces, nv, lam):
        """Solver for I + lam * L of the mesh (trc_lsolver_create)."""
        return LSolver(self, faces, nv, lam)

    def plan(self, jobs, adjoint):
        """A prepared batch (trc_plan_create) over `jobs`."""
        return Plan(self, jobs, adjoint)

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.trc_context_destroy(self.h)
            self.h = None


class LSolver:
    """M = I + lam * L (uniform Laplacian of a triangle mesh) on the device:
    apply (y = M x) and conjugate-gradient solve, on [nv][3] vectors."""

    def __init__(self, ctx, faces, nv, lam):
        self.ctx = ctx
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        h = vp()
        ctx.check(ctx.lib.trc_lsolver_create(ctx.h, _up(f), len(f), int(nv), float(lam), ctypes.byref(h)),
                  "trc_lsolver_create")
        self.h = h
        self.nv = int(nv)

    def apply(self, x_ptr, y_ptr, stream=None):
        self.ctx.check(self.ctx.lib.trc_lsolver_apply(self.h, x_ptr, y_ptr, _stream_ptr(stream)), "trc_lsolver_apply")

    def solve(self, b_ptr, x_ptr, max_iter=200, tol=1e-6, stream=None):
        self.ctx.check(self.ctx.lib.trc_lsolver_solve(self.h, b_ptr, x_ptr, int(max_iter), float(tol),
                                                      _stream_ptr(stream)), "trc_lsolver_solve")

    def status(self):
        """(iterations, relative residual per column) of the last solve."""
        it = ctypes.c_int32(0)
        rr = np.zeros(3, np.float32)
        self.ctx.check(self.ctx.lib.trc_lsolver_status(self.h, ctypes.byref(it), _fp(rr)), "trc_lsolver_status")
        return int(it.value), rr

    def close(self):
        if self.h is not None and self.h.value:
            self.ctx.lib.trc_lsolver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """A prepared batch: the job table is uploaded once; launch() only
    launches kernels (capturable into a graph).  The jobs' buffers must
    outlive the plan."""

    def __init__(self, ctx, jobs, adjoint):
        self.ctx = ctx
        self.adjoint = bool(adjoint)
        self.jobs = list(jobs)
        arr = (trc_job * len(self.jobs))(*self.jobs)
        h = vp()
        ctx.check(ctx.lib.trc_plan_create(ctx.h, arr, len(self.jobs), 1 if adjoint else 0, ctypes.byref(h)),
                  "trc_plan_create")
        self.h = h

    @property
    def tile(self):
        return int(self.ctx.lib.trc_plan_tile(self.h))

    @tile.setter
    def tile(self, paths):
        self.ctx.check(self.ctx.lib.trc_plan_set_tile(self.h, int(paths)), "trc_plan_set_tile")

    def launch(self, stream=None):
        self.ctx.check(self.ctx.lib.trc_plan_launch(self.h, _stream_ptr(stream)), "trc_plan_launch")

    def tune(self, candidates, reps=3, stream=None):
        """Times each candidate tile (median of `reps` launches after one
        warm-up), keeps the fastest and returns it."""
        import numpy as np
        c = np.ascontiguousarray(candidates, dtype=np.uint32)
        best = ctypes.c_uint32(0)
        self.ctx.check(self.ctx.lib.trc_plan_tune(self.h, _up(c), len(c), int(reps), ctypes.byref(best),
                                                  _stream_ptr(stream)), "trc_plan_tune")
        return int(best.value)

    def close(self):
        if self.h is not None and self.h.value:
            self.ctx.lib.trc_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_camera(width, height, pos=(0.0, 1.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0):
    c = trc_camera()
    c.pos[:] = pos
    c.target[:] = target
    c.up[:] = up
    c.width = width
    c.height = height
    c.fov = fov
    return c


def make_options(seed=1234, max_bounces=1, spp=1, sppe=0, sppse0=0, sppse1=0, sppe0=0, quiet=1, g_direct=False,
                 g_direct_depth=0, g_direct_max_size=100000, g_direct_spp=16, g_direct_thold=0.01, g_eps=0.01,
                 g_keep=False, g_seed=0):
    """RenderOptions with the reference's defaults (Integrator.h:47-78).
    g_keep / g_seed: build the direct-term guide from g_seed and keep it on
    the scene for later calls (views) until the scene changes."""
    o = trc_options()
    o.g_keep = 1 if g_keep else 0
    o.g_seed = g_seed
    o.g_direct = 1 if g_direct else 0
    o.g_direct_depth = g_direct_depth
    o.g_direct_max_size = g_direct_max_size
    o.g_direct_spp = g_direct_spp
    o.g_direct_thold = g_direct_thold
    o.g_eps = g_eps
    o.seed = seed
    o.max_bounces = max_bounces
    o.spp = spp
    o.sppe = sppe
    o.sppse0 = sppse0
    o.sppse1 = sppse1
    o.sppe0 = sppe0
    o.quiet = quiet
    return o


def _stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class DeviceScene:
    """A scene resident in HBM (Scene + Scene::Configure of the reference)."""

    def __init__(self, ctx, spec, stream=None):
        self.ctx = ctx
        lib = ctx.lib
        h = vp()
        ctx.check(lib.trc_scene_create(ctx.h, ctypes.byref(h)), "trc_scene_create")
        self.h = h
        self.spec = spec
        for t in spec.textures:
            t = np.ascontiguousarray(t, np.float32)
            if lib.trc_scene_add_texture(h, _fp(t), t.shape[1], t.shape[0]) < 0:
                ctx.check(1, "trc_scene_add_texture")
        for m in spec.materials:
            a = np.asarray(m.albedo, np.float32)
            s = np.asarray(m.spec, np.float32)
            if lib.trc_scene_add_material(h, m.type, _fp(a), _fp(s), m.roughness, m.texture) < 0:
                ctx.check(1, "trc_scene_add_material")
        self._keep = []
        for m in spec.meshes:
            pos = np.ascontiguousarray(m.pos, np.float32)
            idx = np.ascontiguousarray(m.idx, np.uint32)
            uv = np.ascontiguousarray(m.uv, np.float32) if m.uv is not None else None
            uvi = np.ascontiguousarray(m.uv_idx, np.uint32) if m.uv_idx is not None else None
            em = np.asarray(m.emit, np.float32) if m.emit is not None else None
            r = lib.trc_scene_add_mesh(h, _fp(pos), len(pos), _up(idx), len(idx), _fp(uv),
                                       len(uv) if uv is not None else 0, _up(uvi), m.material, _fp(em), m.flags)
            if r < 0:
                ctx.check(1, "trc_scene_add_mesh")
        if ctx.host_only:
            ctx.check(lib.trc_scene_prepare(h), "trc_scene_prepare")
        else:
            ctx.check(lib.trc_scene_configure(h, _stream_ptr(stream)), "trc_scene_configure")

    def info(self):
        i = trc_scene_info()
        self.ctx.check(self.ctx.lib.trc_scene_get_info(self.h, ctypes.byref(i)), "trc_scene_get_info")
        return i

    def vertices(self):
        i = self.info()
        out = np.zeros((i.num_vertices, 3), np.float32)
        self.ctx.check(self.ctx.lib.trc_scene_get_vertices(self.h, _fp(out)), "trc_scene_get_vertices")
        return out

    def bvh(self):
        i = self.info()
        nodes = np.zeros((i.bvh_nodes, 32), np.uint32)
        order = np.zeros(i.num_triangles, np.uint32)
        self.ctx.check(self.ctx.lib.trc_scene_get_bvh(self.h, _up(nodes), _up(order)), "trc_scene_get_bvh")
        return nodes, order

    def edges(self):
        i = self.info()
        out = np.zeros((i.num_edges, 5), np.int32)
        self.ctx.check(self.ctx.lib.trc_scene_get_edges(self.h, out.ctypes.data_as(c_i32_p)), "trc_scene_get_edges")
        return out

    def set_vertices(self, pos_ptr, nv, refit=True, stream=None):
        self.ctx.check(self.ctx.lib.trc_scene_set_vertices(self.h, pos_ptr, nv, int(refit), _stream_ptr(stream)),
                       "trc_scene_set_vertices")

    def set_material(self, mi, albedo=None, spec=None, roughness=0.0):
        a = np.asarray(albedo, np.float32) if albedo is not None else None
        s = np.asarray(spec, np.float32) if spec is not None else None
        self.ctx.check(self.ctx.lib.trc_scene_set_material(self.h, mi, _fp(a), _fp(s), roughness),
                       "trc_scene_set_material")

    # ---- device work (pointers are device addresses) ----
    def intersect(self, rays_ptr, n, tri_ptr, uvt_ptr, any_hit=False, stream=None):
        self.ctx.check(self.ctx.lib.trc_intersect(self.h, rays_ptr, n, tri_ptr, uvt_ptr, int(any_hit),
                                                  _stream_ptr(stream)), "trc_intersect")

    def render_fwd(self, cam, opt, film_ptr, stream=None):
        self.ctx.check(self.ctx.lib.trc_render_fwd(self.h, ctypes.byref(cam), ctypes.byref(opt), film_ptr,
                                                   _stream_ptr(stream)), "trc_render_fwd")

    def render_fwd_noalloc(self, cam, opt, film_ptr, stream=None):
        self.ctx.check(self.ctx.lib.trc_render_fwd_noalloc(self.h, ctypes.byref(cam), ctypes.byref(opt), film_ptr,
                                                           _stream_ptr(stream)), "trc_render_fwd_noalloc")

    def render_paths(self, cam, opt, paths_ptr, stream=None):
        self.ctx.check(self.ctx.lib.trc_render_paths(self.h, ctypes.byref(cam), ctypes.byref(opt), paths_ptr,
                                                     _stream_ptr(stream)), "trc_render_paths")

    def render_fwd_reduce(self, paths_ptr, spp, wh, film_ptr, stream=None):
        self.ctx.check(self.ctx.lib.trc_film_from_paths(self.ctx.h, paths_ptr, spp, wh, film_ptr,
                                                        _stream_ptr(stream)), "trc_film_from_paths")

    def render_stats(self, cam, opt, stream=None):
        """(paths, queries, node visits, triangle tests) of the forward paths."""
        out = (ctypes.c_uint64 * 4)()
        self.ctx.check(self.ctx.lib.trc_render_stats(self.h, ctypes.byref(cam), ctypes.byref(opt), out,
                                                     _stream_ptr(stream)), "trc_render_stats")
        return tuple(int(x) for x in out)

    def render_adj(self, cam, opt, dldi_ptr, dpos_ptr, dmat_ptr=None, stream=None, dtex_ptr=None):
        self.ctx.check(self.ctx.lib.trc_render_adj(self.h, ctypes.byref(cam), ctypes.byref(opt), dldi_ptr, dpos_ptr,
                                                   dmat_ptr, dtex_ptr, _stream_ptr(stream)), "trc_render_adj")

    def checkpoint_bytes(self, cam, opt):
        return int(self.ctx.lib.trc_checkpoint_bytes(ctypes.byref(cam), ctypes.byref(opt)))

    def render_fwd_ckpt(self, cam, opt, film_ptr, ckpt_ptr, ckpt_bytes, stream=None):
        self.ctx.check(self.ctx.lib.trc_render_fwd_ckpt(self.h, ctypes.byref(cam), ctypes.byref(opt), film_ptr,
                                                        ckpt_ptr, ckpt_bytes, _stream_ptr(stream)),
                       "trc_render_fwd_ckpt")

    def render_adj_ckpt(self, cam, opt, ckpt_ptr, ckpt_bytes, dldi_ptr, dpos_ptr, dmat_ptr=None, dtex_ptr=None,
                        stream=None):
        self.ctx.check(self.ctx.lib.trc_render_adj_ckpt(self.h, ctypes.byref(cam), ctypes.byref(opt), ckpt_ptr,
                                                        ckpt_bytes, dldi_ptr, dpos_ptr, dmat_ptr, dtex_ptr,
                                                        _stream_ptr(stream)), "trc_render_adj_ckpt")

    def job(self, cam, opt, first, count, paths=None, ckpt=None, dldi=None, dpos=None, dmat=None, dtex=None,
            vn_scratch=None):
        """A trc_job of this scene (pointers are device addresses or None)."""
        j = trc_job()
        j.scene = self.h.value
        j.cam = cam
        j.opt = opt
        j.first_path = first
        j.num_paths = count
        j.paths, j.ckpt, j.dLdI, j.dpos = paths, ckpt, dldi, dpos
        j.dmat, j.dtex, j.vn_scratch = dmat, dtex, vn_scratch
        return j

    BOUNDARY_KINDS = {"primary": 0, "pixel": 1, "direct": 2, "indirect": 3}

    def render_boundary(self, cam, opt, kind, dldi_ptr, dpos_ptr, stream=None):
        k = self.BOUNDARY_KINDS[kind] if isinstance(kind, str) else int(kind)
        self.ctx.check(self.ctx.lib.trc_render_boundary(self.h, ctypes.byref(cam), ctypes.byref(opt), k, dldi_ptr,
                                                        dpos_ptr, _stream_ptr(stream)), "trc_render_boundary")

    DERIV_KINDS = {"interior": -1, "primary": 0, "pixel": 1, "direct": 2, "indirect": 3}

    def render_deriv(self, cam, opt, kind, tpos_ptr, tmat_ptr, dimg_ptr, stream=None):
        """dimg [3][H*W] += derivative image of integrator `kind` along the
        tangents tpos [V][3] (and tmat [M][8] or None) -- device pointers."""
        k = self.DERIV_KINDS[kind] if isinstance(kind, str) else int(kind)
        self.ctx.check(self.ctx.lib.trc_render_deriv(self.h, ctypes.byref(cam), ctypes.byref(opt), k, tpos_ptr,
                                                     tmat_ptr, dimg_ptr, _stream_ptr(stream)), "trc_render_deriv")

    def close(self):
        if self.h is not None and self.h.value:
            self.ctx.lib.trc_scene_destroy(self.h)
            self.h = None
