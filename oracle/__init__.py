"""ORACLE -- test infrastructure only (the checker for the HIP path).

ctypes wrapper of oracle/oracle.cpp, a scalar CPU restatement of the
reference integrator with its own reverse-mode AD.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package; the product (tensorray_amd) never does.

What pins each oracle function is stated in the docstrings of
tests/test_oracle.py.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "oracle.cpp")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "liboracle.so")

_lib = None


def build():
    """Compile oracle.cpp with g++ (-ffp-contract=off: expressions are
    evaluated exactly as written)."""
    os.makedirs(BUILD, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
           "-o", tmp, SRC, "-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        P = ctypes.POINTER
        f32, u32, i32, f64 = ctypes.c_float, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double
        L.oracle_philox.argtypes = [P(u32), P(u32), P(u32)]
        L.oracle_sincos.argtypes = [P(f32), ctypes.c_int, P(f32), P(f32)]
        L.oracle_exp2.argtypes = [P(f32), ctypes.c_int, P(f32)]
        L.oracle_scene_create.restype = vp
        L.oracle_scene_create.argtypes = [P(f32), ctypes.c_int, P(u32), ctypes.c_int, P(i32), P(i32), P(i32), P(f32),
                                          ctypes.c_int, P(f32), ctypes.c_int, P(f32), ctypes.c_int, P(u32)]
        L.oracle_scene_add_texture.argtypes = [vp, P(f32), ctypes.c_int, ctypes.c_int]
        L.oracle_scene_destroy.argtypes = [vp]
        L.oracle_scene_derived.argtypes = [vp, P(f32), P(f32), P(f32), P(f32), P(f32), P(f32)]
        L.oracle_intersect.argtypes = [vp, P(f32), ctypes.c_int, ctypes.c_int, ctypes.c_int, P(i32), P(f32),
                                       ctypes.c_int]
        L.oracle_camera.argtypes = [P(f32), ctypes.c_int, ctypes.c_int, P(f32)]
        L.oracle_camera_dirs.argtypes = [P(f32), ctypes.c_int, ctypes.c_int, P(f32), ctypes.c_int, P(f32)]
        L.oracle_render_paths.argtypes = [vp, P(f32), ctypes.c_int, ctypes.c_int, u32, ctypes.c_int, ctypes.c_uint64,
                                          ctypes.c_uint64, P(f32), ctypes.c_int]
        L.oracle_render_film.argtypes = [vp, P(f32), ctypes.c_int, ctypes.c_int, u32, ctypes.c_int, ctypes.c_int,
                                         P(f32), ctypes.c_int]
        L.oracle_path_twin.argtypes = [vp, P(f32), ctypes.c_int, ctypes.c_int, u32, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, P(f64)]
        L.oracle_render_adj.argtypes = [vp, P(f32), ctypes.c_int, ctypes.c_int, u32, ctypes.c_int, ctypes.c_int,
                                        P(f32), ctypes.c_uint64, ctypes.c_uint64, P(f64), P(f64), ctypes.c_int,
                                        P(f64)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox(_p(c, ctypes.c_uint32), _p(k, ctypes.c_uint32), _p(out, ctypes.c_uint32))
    return out


def sincos(x):
    x = np.ascontiguousarray(x, np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    lib().oracle_sincos(_p(x, ctypes.c_float), len(x), _p(s, ctypes.c_float), _p(c, ctypes.c_float))
    return s, c


def exp2(x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x)
    lib().oracle_exp2(_p(x, ctypes.c_float), len(x), _p(y, ctypes.c_float))
    return y


def ggx_d(z, alpha):
    """BSDF::GGX_D at half vectors with cos(theta_h) = z."""
    z = np.ascontiguousarray(z, np.float32)
    y = np.zeros_like(z)
    L = lib()
    L.oracle_ggx_d.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_float,
                               ctypes.POINTER(ctypes.c_float)]
    L.oracle_ggx_d(_p(z, ctypes.c_float), len(z), alpha, _p(y, ctypes.c_float))
    return y


def fresnel_conductor(ci, eta, k):
    """BSDF::FresnelConductor for one channel."""
    ci = np.ascontiguousarray(ci, np.float32)
    y = np.zeros_like(ci)
    L = lib()
    L.oracle_fresnel_conductor.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_float,
                                           ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
    L.oracle_fresnel_conductor(_p(ci, ctypes.c_float), len(ci), eta, k, _p(y, ctypes.c_float))
    return y


def camera_params(cam, w, h):
    """cam = dict(pos, target, up, fov).  Returns 19 floats: eye, right, up,
    dir, ax, bx, ay, by, tan_half, ratio, inv_area."""
    c = np.array(list(cam["pos"]) + list(cam["target"]) + list(cam["up"]) + [cam["fov"]], np.float32)
    out = np.zeros(19, np.float32)
    lib().oracle_camera(_p(c, ctypes.c_float), w, h, _p(out, ctypes.c_float))
    return out


def camera_dirs(cam, w, h, raster):
    c = np.array(list(cam["pos"]) + list(cam["target"]) + list(cam["up"]) + [cam["fov"]], np.float32)
    r = np.ascontiguousarray(raster, np.float32)
    out = np.zeros((len(r), 3), np.float32)
    lib().oracle_camera_dirs(_p(c, ctypes.c_float), w, h, _p(r, ctypes.c_float), len(r), _p(out, ctypes.c_float))
    return out


class Scene:
    """The oracle's own copy of a scene (derived data and a simple BVH are
    computed by the oracle itself)."""

    def __init__(self, spec):
        f = spec.flat()
        self.spec = spec
        self.flat = f
        mats = spec.material_table()
        uv = f.get("uv")
        uvi = f.get("uv_idx")
        self.V = len(f["pos"])
        self.T = len(f["idx"])
        self.M = len(mats)
        self.h = lib().oracle_scene_create(
            _p(f["pos"], ctypes.c_float), self.V, _p(f["idx"], ctypes.c_uint32), self.T,
            _p(f["tri_mat"], ctypes.c_int32), _p(f["tri_emitter"], ctypes.c_int32), _p(f["tri_flags"], ctypes.c_int32),
            _p(f["emit_rad"], ctypes.c_float), len(f["emit_rad"]), _p(mats, ctypes.c_float), len(mats),
            _p(uv, ctypes.c_float), len(uv) if uv is not None else 0, _p(uvi, ctypes.c_uint32))
        self.tex_floats = int(sum(np.asarray(t).size for t in spec.textures))
        for t in spec.textures:
            t = np.ascontiguousarray(t, np.float32)
            lib().oracle_scene_add_texture(self.h, _p(t, ctypes.c_float), t.shape[1], t.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_scene_destroy(self.h)
            self.h = None

    def derived(self):
        nl = int((self.flat["tri_emitter"] >= 0).sum())
        fn = np.zeros((self.T, 3), np.float32)
        area = np.zeros(self.T, np.float32)
        vn = np.zeros((self.V, 3), np.float32)
        cdf = np.zeros(nl, np.float32)
        lf = np.zeros(nl, np.float32)
        li = np.zeros(1, np.float32)
        fp = lambda a: _p(a, ctypes.c_float)
        lib().oracle_scene_derived(self.h, fp(fn), fp(area), fp(vn), fp(cdf), fp(lf), fp(li))
        return dict(face_normal=fn, area=area, vertex_normal=vn, light_cdf=cdf, light_f=lf, light_integral=li[0])

    def intersect(self, rays, brute=False, any_hit=False, threads=None):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        tri = np.zeros(n, np.int32)
        uvt = np.zeros((n, 3), np.float32)
        lib().oracle_intersect(self.h, _p(rays, ctypes.c_float), n, int(brute), int(any_hit), _p(tri, ctypes.c_int32),
                               _p(uvt, ctypes.c_float), threads or default_threads())
        return tri, uvt

    @staticmethod
    def _cam(cam):
        return np.array(list(cam["pos"]) + list(cam["target"]) + list(cam["up"]) + [cam["fov"]], np.float32)

    def render_paths(self, cam, w, h, seed, max_bounces, first, count, threads=None):
        out = np.zeros((count, 3), np.float32)
        c = self._cam(cam)
        lib().oracle_render_paths(self.h, _p(c, ctypes.c_float), w, h, seed, max_bounces, first, count,
                                  _p(out, ctypes.c_float), threads or default_threads())
        return out

    def render_film(self, cam, w, h, seed, max_bounces, spp, threads=None):
        film = np.zeros((3, w * h), np.float32)
        c = self._cam(cam)
        lib().oracle_render_film(self.h, _p(c, ctypes.c_float), w, h, seed, max_bounces, spp,
                                 _p(film, ctypes.c_float), threads or default_threads())
        return film

    PARAM_KINDS = {"vertex": 0, "roughness": 1, "albedo": 2, "spec": 3, "texel": 4}

    def path_twin_fd(self, cam, w, h, seed, max_bounces, path, kind, index, comp=0, delta=1e-4):
        """Central difference (float64) of one path's material-form estimator
        with respect to one parameter; returns an RGB vector."""
        out = np.zeros(6, np.float64)
        c = self._cam(cam)
        lib().oracle_path_twin(self.h, _p(c, ctypes.c_float), w, h, seed, max_bounces, int(path),
                               self.PARAM_KINDS[kind], int(index), int(comp), float(delta), _p(out, ctypes.c_double))
        return (out[:3] - out[3:]) / (2.0 * delta)

    def render_adj(self, cam, w, h, seed, max_bounces, spp, dLdI, first=0, count=None, threads=None,
                   with_tex=False):
        """Returns (dpos [V,3], dmat [M,8]) in float64, plus the texture
        gradients (textures' [h][w][3] texels concatenated) if with_tex."""
        dLdI = np.ascontiguousarray(dLdI, np.float32).reshape(3, w * h)
        if count is None:
            count = spp * w * h
        dpos = np.zeros((self.V, 3), np.float64)
        dmat = np.zeros((self.M, 8), np.float64)
        dtex = np.zeros(self.tex_floats, np.float64) if with_tex else None
        c = self._cam(cam)
        lib().oracle_render_adj(self.h, _p(c, ctypes.c_float), w, h, seed, max_bounces, spp, _p(dLdI, ctypes.c_float),
                                first, count, _p(dpos, ctypes.c_double), _p(dmat, ctypes.c_double),
                                threads or default_threads(), _p(dtex, ctypes.c_double))
        return (dpos, dmat, dtex) if with_tex else (dpos, dmat)


def _boundary_kinds():
    return {"primary": 0, "pixel": 1, "direct": 2, "indirect": 3}


def boundary_adj(scene, cam, w, h, seed, max_bounces, kind, spp, dLdI, first=0, count=None, threads=None,
                 guide=None):
    """Gradient [V,3] (float64) of one boundary estimator (primary, pixel,
    direct, indirect) for samples [first, first + count) of spp * w * h.
    guide = (g_spp, g_eps) enables guiding of the direct term, with the
    guide built from `seed`; (g_spp, g_eps, g_seed) builds it from g_seed."""
    L = lib()
    if not hasattr(L, "_bnd_decl"):
        P = ctypes.POINTER
        L.oracle_boundary_adj_gs.argtypes = [ctypes.c_void_p, P(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                             ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_float),
                                             ctypes.c_uint64, ctypes.c_uint64, P(ctypes.c_double), ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint32]
        L.oracle_guide.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float, P(ctypes.c_float),
                                   P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_int)]
        L._bnd_decl = True
    dLdI = np.ascontiguousarray(dLdI, np.float32).reshape(3, w * h)
    if count is None:
        count = spp * w * h
    dpos = np.zeros((scene.V, 3), np.float64)
    c = scene._cam(cam)
    g_seed = int(guide[2]) if guide and len(guide) > 2 else seed
    L.oracle_boundary_adj_gs(scene.h, _p(c, ctypes.c_float), w, h, seed, max_bounces, _boundary_kinds()[kind], spp,
                             _p(dLdI, ctypes.c_float), first, count, _p(dpos, ctypes.c_double),
                             threads or default_threads(), 1 if guide else 0, int(guide[0]) if guide else 0,
                             float(guide[1]) if guide else 0.0, g_seed)
    return dpos


def guide_weights(scene, seed, g_spp, g_eps):
    """Guiding leaves of the direct boundary term: (x0, x1, poly) per
    secondary edge, poly after the eps mix."""
    L = lib()
    n = ctypes.c_int(0)
    L.oracle_guide.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float,
                               ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                               ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)]
    L.oracle_guide(scene.h, seed, g_spp, g_eps, None, None, None, ctypes.byref(n))
    x0 = np.zeros(n.value, np.float32)
    x1 = np.zeros(n.value, np.float32)
    poly = np.zeros(n.value, np.float32)
    L.oracle_guide(scene.h, seed, g_spp, g_eps, _p(x0, ctypes.c_float), _p(x1, ctypes.c_float),
                   _p(poly, ctypes.c_float), ctypes.byref(n))
    return x0, x1, poly


def adam_step(x, g, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-6):
    """Adam::Step (LibTensorRay/Examples/InvRendering/Optimizer.h:95-122), in
    float32 with one rounding per operation, NaN / inf gradients as zero
    (renderer_single_view.py:82).  Returns new (x, m, v)."""
    f = np.float32
    x, g, m, v = (np.asarray(a, np.float32).copy() for a in (x, g, m, v))
    g[~np.isfinite(g)] = 0.0
    b1, b2, lr, eps = f(beta1), f(beta2), f(lr), f(eps)
    m = b1 * m + (f(1.0) - b1) * g
    v = b2 * v + (f(1.0) - b2) * (g * g)
    # 1 - beta^step of the float32 beta (the reference keeps beta in a float)
    bc1 = f(1.0 - float(b1) ** step)
    bc2 = f(1.0 - float(b2) ** step)
    mh = m / bc1
    vh = v / bc2
    x = x - (lr * mh) / (np.sqrt(vh) + eps)
    return x.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


def adam_uniform_step(x, g, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999):
    """Uniform Adam (pyTensorRay/optimizer.py:57-61, without its Laplacian
    solve): moments as adam_step, then x -= lr / ((1 - b1^t) sqrt(max(v) /
    (1 - b2^t))) m in float32, no update while max(v) = 0."""
    f = np.float32
    x, g, m, v = (np.asarray(a, np.float32).copy() for a in (x, g, m, v))
    g[~np.isfinite(g)] = 0.0
    b1, b2, lr = f(beta1), f(beta2), f(lr)
    m = b1 * m + (f(1.0) - b1) * g
    v = b2 * v + (f(1.0) - b2) * (g * g)
    bc1 = f(1.0 - float(b1) ** step)
    bc2 = f(1.0 - float(b2) ** step)
    vmax = v.max() if v.size else f(0.0)
    if vmax > 0:
        scale = lr / (bc1 * np.sqrt(vmax / bc2))
        x = x - scale * m
    return x.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


def laplacian_system(faces, nv, lam):
    """M = I + lam * L as a scipy CSR matrix (float64), L the uniform graph
    Laplacian of the triangle mesh: L_ii = number of neighbours of i, L_ij =
    -1 for every mesh edge (compute_matrix of the largesteps package used by
    pyTensorRay/optimizer.py:152)."""
    import scipy.sparse as sp
    f = np.asarray(faces, np.int64).reshape(-1, 3)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e = np.concatenate([e, e[:, ::-1]])
    e = e[e[:, 0] != e[:, 1]]
    e = np.unique(e, axis=0)
    A = sp.csr_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(nv, nv))
    deg = np.asarray(A.sum(axis=1)).ravel()
    L = sp.diags(deg) - A
    return (sp.identity(nv) + lam * L).tocsr()


# ---- BSDFs in a bare local frame (normal +z), for closed-form checks ----
def _decl_bsdf_local(L):
    if getattr(L, "_bsdf_local_decl", False):
        return
    P = ctypes.POINTER
    f32 = ctypes.c_float
    L.oracle_bsdf_local.argtypes = [ctypes.c_int, f32, P(f32), P(f32), P(f32), P(f32), ctypes.c_int, P(f32), P(f32)]
    L.oracle_bsdf_sample_local.argtypes = [ctypes.c_int, f32, P(f32), P(f32), ctypes.c_int, P(f32), P(f32)]
    L.oracle_vndf.argtypes = [P(f32), f32, P(f32), ctypes.c_int, P(f32), P(f32)]
    L.oracle_smith_g.argtypes = [P(f32), P(f32), f32, ctypes.c_int, P(f32)]
    L.oracle_mf_parts.argtypes = [P(f32), ctypes.c_int, f32, P(f32)]
    L._bsdf_local_decl = True


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    return a.reshape(shape) if shape is not None else a


def bsdf_local(mat_type, rough, albedo, spec, wo, wi):
    """BSDF::Eval (no shading correction) and BSDF::Pdf of one material for
    local directions wo, wi [n, 3].  Returns (f [n, 3], pdf [n])."""
    L = lib()
    _decl_bsdf_local(L)
    wo = _f32(wo, (-1, 3))
    wi = _f32(wi, (-1, 3))
    n = len(wo)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    a = _f32(albedo)
    s = _f32(spec)
    fp = lambda x: _p(x, ctypes.c_float)
    L.oracle_bsdf_local(int(mat_type), float(rough), fp(a), fp(s), fp(wo), fp(wi), n, fp(f), fp(pdf))
    return f, pdf


def bsdf_sample_local(mat_type, rough, wo, u):
    """BSDF::SampleOnly for local wo [n, 3] and samples u [n, 2]: (wi, pdf)."""
    L = lib()
    _decl_bsdf_local(L)
    wo = _f32(wo, (-1, 3))
    u = _f32(u, (-1, 2))
    n = len(wo)
    wi = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    fp = lambda x: _p(x, ctypes.c_float)
    L.oracle_bsdf_sample_local(int(mat_type), float(rough), fp(wo), fp(u), n, fp(wi), fp(pdf))
    return wi, pdf


def vndf(v, alpha, u):
    """BSDF::GGX_VNDF_Sample and GGX_VNDF_Pdf for local v [n, 3] and samples
    u [n, 2]: (wh [n, 3], pdf [n])."""
    L = lib()
    _decl_bsdf_local(L)
    v = _f32(v, (-1, 3))
    u = _f32(u, (-1, 2))
    n = len(v)
    wh = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    fp = lambda x: _p(x, ctypes.c_float)
    L.oracle_vndf(fp(v), float(alpha), fp(u), n, fp(wh), fp(pdf))
    return wh, pdf


def smith_g(v, wh, alpha):
    """BSDF::Smith_G(v, wh, alpha) for local v, wh [n, 3]."""
    L = lib()
    _decl_bsdf_local(L)
    v = _f32(v, (-1, 3))
    wh = _f32(wh, (-1, 3))
    g = np.zeros(len(v), np.float32)
    fp = lambda x: _p(x, ctypes.c_float)
    L.oracle_smith_g(fp(v), fp(wh), float(alpha), len(v), fp(g))
    return g


def microfacet_parts(cos_h, vh, cos_n, rough):
    """Microfacet::EvalInner's factors: GGX D (alpha = r^2) at cos_h, the
    Fresnel weight 2^{vh (-5.55473 vh - 6.8316)} at vh, Smith-Schlick G1
    (k = (r + 1)^2 / 8) at cos_n.  Returns three arrays."""
    L = lib()
    _decl_bsdf_local(L)
    a = np.stack(np.broadcast_arrays(*(np.asarray(x, np.float32) for x in (cos_h, vh, cos_n))), -1)
    a = _f32(a, (-1, 3))
    out = np.zeros_like(a)
    L.oracle_mf_parts(_p(a, ctypes.c_float), len(a), float(rough), _p(out, ctypes.c_float))
    return out[:, 0], out[:, 1], out[:, 2]
