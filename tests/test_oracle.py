"""Pins of the oracle (CPU restatement of the reference integrator).

Each oracle function is checked against something that does not retype its
own formula:
  philox            published known-answer vectors + rocRAND's independent
                    host implementation
  sincos / exp2     numpy float64
  intersect (BVH)   brute force over all triangles, watertightness on shared
                    vertices / edges, miss-record semantics (ptx.cu:92-99)
  camera            the reference's matrix pipeline (LookAt, Perspective,
                    ScreenToRaster, inverses; Matrix.cpp:181-220,
                    BaseCamera.cpp:73-88) restated in numpy float64
  radiance (NEE/MIS/BSDF sampling/light sampling)
                    Lambert's closed form for the irradiance of a polygonal
                    emitter on an unoccluded Lambertian plane
  adjoint (AD)      central finite differences of the per-path estimator
                    (reference tolerance: rel < 1e-3 or abs < 1e-3,
                    Tensor.h:3019)
"""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from tensorray_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def _kat():
    with open(os.path.join(HERE, "golden", "philox_kat.json")) as f:
        return json.load(f)["vectors"]


def test_philox_known_answers(oracle_mod):
    for v in _kat():
        out = oracle_mod.philox([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]])
        assert [f"{x:08x}" for x in out] == v["out"]


ROCRAND_PROG = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
struct E : rocrand_device::philox4x32_10_engine {
  E() : rocrand_device::philox4x32_10_engine(0, 0, 0) {}
  uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  E e; unsigned a, b, c, d, k0, k1;
  while (scanf("%x %x %x %x %x %x", &a, &b, &c, &d, &k0, &k1) == 6) {
    uint4 r = e.run(uint4{a, b, c, d}, uint2{k0, k1});
    printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
  }
}
"""


def test_philox_matches_rocrand(oracle_mod):
    """rocRAND ships an independent Philox4x32-10 (host-callable header)."""
    hdr = "/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"
    if not os.path.exists(hdr) or shutil.which("g++") is None:
        pytest.skip("rocRAND headers not available")
    rng = np.random.default_rng(5)
    vecs = [[int(x, 16) for x in v["ctr"] + v["key"]] for v in _kat()]
    vecs += rng.integers(0, 2 ** 32, size=(64, 6), dtype=np.uint64).tolist()
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "p.cpp"), os.path.join(d, "p")
        with open(src, "w") as f:
            f.write(ROCRAND_PROG)
        subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", src, "-o", exe],
                       check=True, capture_output=True)
        inp = "\n".join(" ".join(f"{x:x}" for x in v) for v in vecs) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for v, line in zip(vecs, out):
        got = oracle_mod.philox(v[:4], v[4:])
        assert " ".join(f"{x:08x}" for x in got) == line


def test_elementary_functions(oracle_mod):
    x = np.linspace(-4 * np.pi, 4 * np.pi, 200001).astype(np.float32)
    s, c = oracle_mod.sincos(x)
    xd = x.astype(np.float64)
    assert np.abs(s - np.sin(xd)).max() < 3e-7
    assert np.abs(c - np.cos(xd)).max() < 3e-7
    y = np.linspace(-24.0, 0.0, 100001).astype(np.float32)
    e = oracle_mod.exp2(y)
    ref = np.exp2(y.astype(np.float64))
    assert (np.abs(e - ref) / ref).max() < 4e-7


def _random_rays(n, seed, lo=-2.0, hi=2.0, tmin=1e-3):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = 1e30
    return r


@pytest.mark.parametrize("subdiv,amp", [(2, 0.0), (3, 0.05)])
def test_bvh_traversal_equals_brute_force(oracle_mod, subdiv, amp):
    spec = scenes.sphere_scene(subdiv=subdiv, amplitude=amp)
    S = oracle_mod.Scene(spec)
    rays = _random_rays(6000, 11 + subdiv)
    tb, ub = S.intersect(rays, brute=True)
    tv, uv = S.intersect(rays)
    assert (tb >= 0).sum() > 1000
    assert np.array_equal(tb, tv)
    assert np.array_equal(ub.view(np.uint32), uv.view(np.uint32))
    # any-hit agrees with closest-hit on occlusion
    ta, _ = S.intersect(rays, any_hit=True)
    assert np.array_equal(ta >= 0, tb >= 0)


def test_miss_record(oracle_mod):
    spec = scenes.config1_scene()
    S = oracle_mod.Scene(spec)
    r = np.array([[0, 10, 0, 1e-3, 0, 1, 0, 1e30]], np.float32)  # upwards, away from everything
    t, uvt = S.intersect(r, brute=True)
    assert t[0] == -1 and list(uvt[0]) == [0.0, 0.0, -1.0]


def test_watertight_vertices_and_edges(oracle_mod):
    """Rays aimed exactly at shared vertices and edge midpoints of a closed
    mesh never slip through (Woop et al. 2013)."""
    pos, idx = scenes.icosphere(3)
    spec = scenes.SceneSpec(meshes=[scenes.Mesh(pos, idx)], materials=[scenes.Material()])
    S = oracle_mod.Scene(spec)
    targets = [pos.astype(np.float64)]
    e = np.concatenate([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]])
    targets.append(0.5 * (pos[e[:, 0]].astype(np.float64) + pos[e[:, 1]].astype(np.float64)))
    tg = np.concatenate(targets)
    # origins outside the surface, roughly along the outward radial direction
    # (no grazing rays), so every ray must enter the closed mesh at the target
    rng = np.random.default_rng(3)
    radial = tg / np.linalg.norm(tg, axis=1, keepdims=True)
    orig = 3.0 * radial + 0.3 * rng.normal(size=tg.shape)
    d = tg - orig
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(tg), 8), np.float32)
    rays[:, :3] = orig
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = 1e30
    t, _ = S.intersect(rays, brute=True)
    assert (t >= 0).all()


# --- camera: the reference's matrix pipeline in float64 ---
def _look_at(eye, target, up):
    d = target - eye
    d /= np.linalg.norm(d)
    u = up / np.linalg.norm(up)
    r = np.cross(u, d)
    r /= np.linalg.norm(r)
    nu = np.cross(d, r)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = r, nu, d, eye
    return np.linalg.inv(m)  # Matrix::LookAt returns Inverse(camToWorld)


def _perspective(fov, ratio, near, far):
    p = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, far / (far - near), -far * near / (far - near)], [0, 0, 1, 0]])
    it = 1.0 / np.tan(np.radians(fov) / 2.0)
    return np.diag([it / ratio, it, 1, 1]) @ p


def _reference_dirs(cam, W, H, raster):
    eye, tgt, up = (np.array(cam[k], np.float64) for k in ("pos", "target", "up"))
    view = _look_at(eye, tgt, up)
    view_inv = np.linalg.inv(view)
    proj = _perspective(cam["fov"], W / H, 1.0, 1000.0)
    s2r = np.diag([W, H, 1, 1]) @ np.diag([-0.5, -0.5, 1, 1]) @ np.array(
        [[1, 0, 0, -1], [0, 1, 0, -1], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)
    r2c = np.linalg.inv(proj) @ np.linalg.inv(s2r)
    out = []
    for rx, ry in raster:
        h = r2c @ np.array([rx, ry, 0.0, 1.0])
        c = h[:3] / h[3]
        dv = c / np.linalg.norm(c)
        out.append((view_inv @ np.array([dv[0], dv[1], dv[2], 0.0]))[:3])
    return np.array(out)


@pytest.mark.parametrize("W,H", [(64, 64), (512, 512), (320, 200)])
def test_camera_matches_reference_matrices(oracle_mod, W, H):
    cam = dict(pos=(0.3, 1.2, 3.1), target=(0.0, -0.1, 0.2), up=(0.0, 1.0, 0.1), fov=45.0)
    rng = np.random.default_rng(2)
    raster = np.stack([rng.uniform(0, W, 200), rng.uniform(0, H, 200)], 1).astype(np.float32)
    ref = _reference_dirs(cam, W, H, raster)
    got = oracle_mod.camera_dirs(cam, W, H, raster).astype(np.float64)
    assert np.abs(got - ref).max() < 2e-6
    # centre of the image looks along the view direction
    c = oracle_mod.camera_dirs(cam, W, H, np.array([[W / 2, H / 2]], np.float32))[0]
    vd = np.array(cam["target"]) - np.array(cam["pos"])
    assert np.allclose(c, vd / np.linalg.norm(vd), atol=1e-6)


# --- analytic direct lighting ---
def _polygon_irradiance(x, n, verts, L):
    """Lambert's formula: E = L/2 * |sum_i theta_i (n . g_i)|, g_i the unit
    normal of the plane through x, v_i, v_{i+1}."""
    u = verts[None, :, :] - x[:, None, :]
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    E = np.zeros(len(x))
    k = len(verts)
    for i in range(k):
        a, b = u[:, i], u[:, (i + 1) % k]
        th = np.arccos(np.clip((a * b).sum(1), -1, 1))
        g = np.cross(a, b)
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        E += th * (g @ n)
    return 0.5 * L * np.abs(E)


def test_direct_lighting_matches_lambert_formula(oracle_mod):
    spec = scenes.SceneSpec()
    spec.materials = [scenes.Material(albedo=(0.5, 0.5, 0.5)), scenes.Material(albedo=(0.0, 0.0, 0.0))]
    g, gi, _ = scenes.quad_y(-1.0, -20.0, 20.0, -20.0, 20.0, up=True)
    spec.meshes.append(scenes.Mesh(g, gi, material=0))
    l, li, _ = scenes.quad_y(2.0, -0.5, 0.5, -0.5, 0.5, up=False)
    spec.meshes.append(scenes.Mesh(l, li, material=1, emit=(10.0, 10.0, 10.0)))
    S = oracle_mod.Scene(spec)
    cam = dict(pos=(0.0, 3.0, 2.5), target=(0.0, -1.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0)
    W = H = 48
    spp, seed = 64, 1234
    n = spp * W * H
    vals = S.render_paths(cam, W, H, seed, 1, 0, n)[:, 0].astype(np.float64)
    # recompute each path's camera ray from the same counter-based numbers
    p = np.arange(n, dtype=np.uint64)
    x0, x1, _, _ = scenes.philox4x32(p & 0xFFFFFFFF, p >> np.uint64(32), np.zeros(n), np.zeros(n), seed, 0x11)
    a = (x0 >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)
    b = (x1 >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)
    pix = (p % (W * H)).astype(np.int64)
    raster = np.stack([(pix % W).astype(np.float32) + a, (pix // W).astype(np.float32) + b], 1)
    d = oracle_mod.camera_dirs(cam, W, H, raster).astype(np.float64)
    o = np.array(cam["pos"], np.float64)
    t = (-1.0 - o[1]) / d[:, 1]
    x = o[None, :] + t[:, None] * d
    E = _polygon_irradiance(x, np.array([0.0, 1.0, 0.0]), l.astype(np.float64), 10.0)
    expect = 0.5 / np.pi * E
    diff = vals - expect
    se = diff.std() / np.sqrt(n)
    assert abs(diff.mean()) < 4 * se + 1e-6
    # per pixel: mean of the estimator vs mean of the closed form
    pv = vals.reshape(spp, W * H).mean(0)
    pe = expect.reshape(spp, W * H).mean(0)
    rel = np.abs(pv - pe) / np.maximum(pe, 1e-3)
    assert np.median(rel) < 0.02


# --- AD vs central finite differences ---
# The reference differentiates a material-form estimator: hit points keep
# their barycentrics and the primary / area weights are G / detach(G) and
# area / detach(area) (Camera.cpp:279-303, Scene.cpp:722-723).  Per sample
# this is a different function of the parameters than the ray-space
# estimator (their expectations agree once the pixel-boundary and edge terms
# are added), so the AD is pinned against central differences of the same
# material-form estimator: the oracle re-evaluates a path with one parameter
# shifted by +-delta in a float64 twin lane while every branch, hit record and
# detached factor keeps its unperturbed value.
@pytest.mark.parametrize("kind", ["diffuse", "microfacet", "conductor", "smooth"])
def test_adjoint_matches_finite_differences(oracle_mod, kind):
    if kind == "diffuse":
        spec, B = scenes.sphere_scene(subdiv=2), 2
    elif kind == "smooth":
        # smooth shading: the shading frame follows the averaged vertex normals
        mat = scenes.Material(scenes.MAT_MICROFACET, albedo=(0.2, 0.3, 0.1), spec=(0.04, 0.05, 0.06), roughness=0.4)
        spec, B = scenes.sphere_scene(subdiv=2, smooth=True, material=mat), 2
    elif kind == "conductor":
        # gold-like complex IOR in the albedo (eta) and spec (k) slots
        mat = scenes.Material(scenes.MAT_ROUGHCONDUCTOR, albedo=(0.143, 0.374, 1.44), spec=(3.98, 2.38, 1.60),
                              roughness=0.35)
        spec, B = scenes.sphere_scene(subdiv=2, material=mat), 3
    else:
        mat = scenes.Material(scenes.MAT_MICROFACET, albedo=(0.2, 0.3, 0.1), spec=(0.04, 0.05, 0.06), roughness=0.4)
        spec, B = scenes.sphere_scene(subdiv=2, material=mat), 3
    S = oracle_mod.Scene(spec)
    cam = scenes.CAMERA_DEFAULT
    W = H = 32
    seed = 77
    base = S.render_paths(cam, W, H, seed, B, 0, W * H)
    rng = np.random.default_rng(1)
    paths = rng.choice(np.where(base.sum(1) > 1e-3)[0], 16, replace=False)
    checked = 0
    for pth in paths:
        gw = rng.uniform(0.5, 1.5, 3).astype(np.float32)
        g = np.zeros((3, W * H), np.float32)
        g[:, pth % (W * H)] = gw
        dpos, dmat = S.render_adj(cam, W, H, seed, B, 1, g, first=int(pth), count=1)
        order = np.argsort(-np.abs(dpos).ravel())[:4]
        for flat in order:
            v, k = divmod(int(flat), 3)
            fd = S.path_twin_fd(cam, W, H, seed, B, pth, "vertex", v, k) @ gw.astype(np.float64)
            ad = dpos[v, k]
            assert abs(fd - ad) <= 1e-3 * max(abs(ad), abs(fd)) + 1e-6, (int(pth), v, k, fd, ad)
            checked += 1
        if kind not in ("diffuse", "smooth"):
            fd = S.path_twin_fd(cam, W, H, seed, B, pth, "roughness", 0) @ gw.astype(np.float64)
            assert abs(fd - dmat[0, 0]) <= 1e-3 * max(abs(fd), abs(dmat[0, 0])) + 1e-6
            for c in range(3):
                fd = S.path_twin_fd(cam, W, H, seed, B, pth, "spec", 0, c) @ gw.astype(np.float64)
                assert abs(fd - dmat[0, 4 + c]) <= 1e-3 * max(abs(fd), abs(dmat[0, 4 + c])) + 1e-6
        if kind == "conductor":
            for c in range(3):
                fd = S.path_twin_fd(cam, W, H, seed, B, pth, "albedo", 0, c) @ gw.astype(np.float64)
                assert abs(fd - dmat[0, 1 + c]) <= 1e-3 * max(abs(fd), abs(dmat[0, 1 + c])) + 1e-6
    assert checked >= 60


def test_albedo_gradient_matches_image_finite_differences(oracle_mod):
    """The image is linear in a Lambertian albedo and the sampling does not
    depend on it, so the AD gradient equals the central difference of the
    whole rendered loss (common random numbers)."""
    import copy
    spec = scenes.sphere_scene(subdiv=2)
    S = oracle_mod.Scene(spec)
    cam = scenes.CAMERA_DEFAULT
    W = H = 24
    spp, B, seed = 8, 2, 3
    w = np.random.default_rng(4).uniform(0.5, 1.5, (3, W * H)).astype(np.float32)
    _, dmat = S.render_adj(cam, W, H, seed, B, spp, w)
    for mi in (0, 1):
        for c in range(3):
            vals = []
            for sgn in (1, -1):
                sp = copy.deepcopy(spec)
                alb = list(sp.materials[mi].albedo)
                alb[c] += sgn * 1e-2
                sp.materials[mi].albedo = tuple(alb)
                f = oracle_mod.Scene(sp).render_film(cam, W, H, seed, B, spp)
                vals.append(float((f.astype(np.float64) * w).sum()))
            fd = (vals[0] - vals[1]) / 2e-2
            assert abs(fd - dmat[mi, 1 + c]) <= 2e-3 * abs(fd) + 1e-3, (mi, c, fd, dmat[mi, 1 + c])


def test_guided_direct_boundary_is_unbiased(oracle_mod):
    """Guiding changes only the sampling density of the direct boundary term
    (pdf_b = poly / sum(poly * volume), AQ_distrb.h:182-197): over 16 seeds
    the guided and unguided estimates of the translation gradient agree
    within 4 standard errors, and the guide weights satisfy
    sum(poly) = n * eps + (1 - eps) with leaves tiling [0, 1]."""
    spec = scenes.sphere_scene(subdiv=2)
    S = oracle_mod.Scene(spec)
    x0, x1, poly = oracle_mod.guide_weights(S, 31, 16, 0.1)
    n = len(poly)
    assert n > 0 and x0[0] == 0.0 and x1[-1] == 1.0 and np.array_equal(x0[1:], x1[:-1])
    assert abs(poly.sum() - (n * 0.1 + 0.9)) < 1e-3 * n
    assert (poly >= 0.1 - 1e-7).all() and poly.max() - poly.min() > 1e-3  # non-uniform weights
    _, _, flat = oracle_mod.guide_weights(S, 31, 16, 1.0)
    assert (flat == 1.0).all()
    W = H = 32
    dl = np.ones((3, W * H), np.float32)
    cam = scenes.CAMERA_DEFAULT
    mask = spec.diff_vertex_mask()
    est = {None: [], (16, 0.1): []}
    for g in est:
        for seed in range(16):
            a = oracle_mod.boundary_adj(S, cam, W, H, 100 + seed, 2, "direct", 64, dl, guide=g)
            est[g].append(a[mask].sum(0))
    a, b = np.array(est[None]), np.array(est[(16, 0.1)])
    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    assert np.abs(a[:, 1].mean()) > 10 * se[1]  # the term is not noise
    assert (np.abs(a.mean(0) - b.mean(0)) <= 4 * se + 1e-6).all(), (a.mean(0), b.mean(0), se)


def test_total_gradient_matches_finite_differences(oracle_mod):
    """The material-form estimator's total gradient (interior + primary edge
    + pixel boundary + secondary direct + indirect) against central finite
    differences of the rendered image for a rigid translation of the sphere
    (common random numbers within each FD pair): a z-score over 8
    independent seeds of each, |mean estimate - mean FD| <= 4 standard
    errors.  The interior term alone does not match; the sum does."""
    import copy
    spec = scenes.sphere_scene(subdiv=1)
    cam = scenes.CAMERA_DEFAULT
    W = H = 24
    B = 1
    w = np.full((3, W * H), 0.5, np.float32)
    S = oracle_mod.Scene(spec)
    nv = len(spec.meshes[0].pos)
    est, interior = [], []
    for seed in range(8):
        di, _ = S.render_adj(cam, W, H, 10 + seed, B, 512, w)
        parts = [oracle_mod.boundary_adj(S, cam, W, H, 10 + seed, B, k, 512, w)
                 for k in ("primary", "pixel", "direct", "indirect")]
        est.append((di[:nv].sum(0) + sum(p[:nv].sum(0) for p in parts))[1:])
        interior.append(di[:nv].sum(0)[1:])

    def loss(dy, dz, seed):
        s = copy.deepcopy(spec)
        s.meshes[0].pos = s.meshes[0].pos + np.array([0.0, dy, dz], np.float32)
        f = oracle_mod.Scene(s).render_film(cam, W, H, seed, B, 1024)
        return float((f.astype(np.float64) * w).sum())

    h = 1e-2
    fd = np.array([[(loss(h, 0, 1000 + sd) - loss(-h, 0, 1000 + sd)) / (2 * h),
                    (loss(0, h, 1000 + sd) - loss(0, -h, 1000 + sd)) / (2 * h)] for sd in range(8)])
    est, interior = np.array(est), np.array(interior)
    se = np.sqrt(est.var(0, ddof=1) / len(est) + fd.var(0, ddof=1) / len(fd))
    z = np.abs(est.mean(0) - fd.mean(0)) / se
    assert (z <= 4.0).all(), (est.mean(0), fd.mean(0), se)
    # the interior term alone misses the edge contributions along y
    se_i = np.sqrt(interior.var(0, ddof=1) / len(interior) + fd.var(0, ddof=1) / len(fd))
    assert abs(interior[:, 0].mean() - fd[:, 0].mean()) > 6 * se_i[0]


def test_ggx_and_conductor_fresnel_closed_forms(oracle_mod):
    """GGX_D (alpha = roughness, as RoughConductor uses it) is a normalized
    microfacet distribution: integral of D cos(theta_h) over the hemisphere
    = 1.  FresnelConductor: normal incidence ((eta-1)^2 + k^2) / ((eta+1)^2 +
    k^2); with k = 0 it is the unpolarized dielectric Fresnel reflectance;
    grazing incidence reflects everything."""
    for a in (0.1, 0.3, 0.7):
        th = (np.arange(200000) + 0.5) / 200000 * (np.pi / 2)
        D = oracle_mod.ggx_d(np.cos(th), a).astype(np.float64)
        integral = (D * np.cos(th) * np.sin(th)).sum() * 2 * np.pi * (np.pi / 2 / 200000)
        assert abs(integral - 1.0) < 2e-3, (a, integral)
    for eta, k in ((0.143, 3.98), (1.5, 0.0), (0.27, 2.78), (2.0, 1.0)):
        f0 = oracle_mod.fresnel_conductor([1.0], eta, k)[0]
        assert abs(f0 - ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)) < 1e-5
        assert abs(oracle_mod.fresnel_conductor([1e-4], eta, k)[0] - 1.0) < 2e-3
    ci = np.linspace(0.05, 1.0, 64)
    eta = 1.5
    ct = np.sqrt(1 - (1 - ci ** 2) / eta ** 2)
    rs = ((ci - eta * ct) / (ci + eta * ct)) ** 2
    rp = ((eta * ci - ct) / (eta * ci + ct)) ** 2
    got = oracle_mod.fresnel_conductor(ci, eta, 0.0)
    assert np.abs(got - 0.5 * (rs + rp)).max() < 1e-5


def test_texture_gradient_matches_finite_differences(oracle_mod):
    """Texel gradients of a textured albedo (BSDF::DiffTexture): reverse-mode
    AD of one path against the twin-lane central difference of that texel."""
    spec = scenes.sphere_scene(subdiv=2, texture_res=16)
    S = oracle_mod.Scene(spec)
    cam = scenes.CAMERA_DEFAULT
    W = H = 32
    seed, B = 5, 2
    base = S.render_paths(cam, W, H, seed, B, 0, W * H)
    rng = np.random.default_rng(3)
    checked = 0
    for pth in rng.choice(np.where(base.sum(1) > 1e-3)[0], 24, replace=False):
        g = np.zeros((3, W * H), np.float32)
        g[:, pth] = rng.uniform(0.5, 1.5, 3)
        _, _, dtex = S.render_adj(cam, W, H, seed, B, 1, g, first=int(pth), count=1, with_tex=True)
        for k in np.argsort(-np.abs(dtex))[:2]:
            if dtex[k] == 0.0:
                continue
            fd = S.path_twin_fd(cam, W, H, seed, B, pth, "texel", 0, int(k)) @ g[:, pth].astype(np.float64)
            assert abs(fd - dtex[k]) <= 1e-3 * max(abs(fd), abs(dtex[k])) + 1e-7, (int(pth), int(k), fd, dtex[k])
            checked += 1
    assert checked >= 10


def test_adam_oracle_closed_forms(oracle_mod):
    """Adam's first step moves each coordinate by lr * g / (|g| + eps)
    (m and v start at zero, so the bias-corrected moments are g and g^2);
    a constant gradient keeps that step size at every later step; a zero
    gradient leaves x unchanged; NaN gradients count as zero."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(5))
    x = rng.standard_normal(1000).astype(np.float32)
    g = (rng.standard_normal(1000) * 10.0 ** rng.uniform(-4, 2, 1000)).astype(np.float32)
    z = np.zeros_like(x)
    x1, m1, v1 = oracle_mod.adam_step(x, g, z, z, 1, lr=1e-2, eps=1e-6)
    g64 = g.astype(np.float64)
    np.testing.assert_allclose(x - x1, 1e-2 * g64 / (np.abs(g64) + 1e-6), rtol=2e-5, atol=1e-9)
    np.testing.assert_allclose(m1, 0.1 * g64, rtol=1e-6)
    np.testing.assert_allclose(v1, 0.001 * g64 * g64, rtol=5e-5)  # 1 - 0.999f = 0.00099999 in float32
    # constant gradient: mh = g, vh = g^2 at every step
    xs, ms, vs = x.copy(), z.copy(), z.copy()
    for t in range(1, 8):
        x_prev = xs
        xs, ms, vs = oracle_mod.adam_step(xs, g, ms, vs, t, lr=1e-2)
        np.testing.assert_allclose(x_prev - xs, 1e-2 * g64 / (np.abs(g64) + 1e-6), rtol=5e-4, atol=1e-8)
    # zero and NaN gradients
    gz = g.copy()
    gz[::3] = 0.0
    gz[1::3] = np.nan
    x2, m2, v2 = oracle_mod.adam_step(x, gz, z, z, 1)
    assert np.array_equal(x2[::3], x[::3]) and np.array_equal(x2[1::3], x[1::3])
    assert np.isfinite(m2).all() and np.isfinite(v2).all()


def test_uniform_adam_and_laplacian_oracle(oracle_mod):
    """Uniform Adam's first step moves every coordinate by lr * g / max|g|
    (bias-corrected moments g and g^2); the Laplacian system of an
    icosahedron has diagonal 1 + 5 lam, -lam on its 30 edges, zero row sums
    of L, and is symmetric positive definite."""
    import numpy as np
    from tensorray_amd import scenes
    rng = np.random.Generator(np.random.PCG64(9))
    x = rng.standard_normal(300).astype(np.float32)
    g = rng.standard_normal(300).astype(np.float32)
    z = np.zeros_like(x)
    x1, m1, v1 = oracle_mod.adam_uniform_step(x, g, z, z, 1, lr=1e-2)
    # x - x1 carries the float32 rounding of x1 (|x| up to ~3): atol of a few ulp
    np.testing.assert_allclose(x.astype(np.float64) - x1, 1e-2 * g.astype(np.float64) / np.abs(g).max(),
                               rtol=1e-5, atol=8 * np.finfo(np.float32).eps * np.abs(x).max())
    x2, _, _ = oracle_mod.adam_uniform_step(x, np.zeros_like(g), z, z, 1)
    assert np.array_equal(x2, x)
    pos, idx = scenes.icosphere(0)
    lam = 0.7
    M = oracle_mod.laplacian_system(idx, len(pos), lam).toarray()
    assert np.allclose(M, M.T)
    np.testing.assert_allclose(np.diag(M), 1 + 5 * lam)
    assert (np.isclose(M, -lam).sum()) == 2 * 30
    np.testing.assert_allclose((M - np.eye(12)).sum(axis=1), 0.0, atol=1e-12)
    assert np.linalg.eigvalsh(M).min() >= 1.0 - 1e-12   # I + lam L with L >= 0
