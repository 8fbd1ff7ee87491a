"""Pins of each boundary (edge) estimator of the oracle on its own, by scenes
built so that one class carries the whole derivative and its value is fixed
by geometry.  Reference: BoundaryPrimary.cpp:36-181, BoundaryPixel.cpp:
36-159, BoundaryDirect.cpp:38-143 with Edge.cpp:108-309, BoundaryIndirect.cpp:
38-192.  Every comparison is a z-score over independent seeds: |mean
estimate - expected| <= 4 standard errors (of the estimate, and of the
finite difference where the expectation is one).

* primary: a black triangle in front of an emitting background.  The image
  is Lb (1 - coverage), the triangle region contributes nothing, the
  background does not move, so the primary term alone must equal
  -Lb d(coverage)/d(vertex), where the coverage of every pixel by the
  projected triangle is computed exactly by polygon clipping (pinhole
  projection of the camera model, float64); interior and pixel terms are 0.
* pixel: the same with an emitting triangle (Lt): the primary term still
  equals -Lb d(coverage) (it carries the far side's radiance), and the
  pixel boundary plus the interior term (the triangle's domain is fixed in
  material form, so its change enters through the moving pixel footprint and
  the integrand) equals Lt d(coverage).  The interior term is pinned by
  twin-lane central differences in test_oracle.py, so this pins the pixel
  term.
* direct: a ground plane under a square light with a black square occluder
  halfway up, out of the camera's view, at one bounce: the only derivative is
  the moving shadow, so the direct term alone must equal the derivative of
  sum_pixels w (rho/pi) E, E from Lambert's formula on the visible part of
  the light (the light square minus the occluder's projection from the
  ground point, an axis-aligned rectangle), averaged over each pixel by a
  24 x 24 rule.  All other classes are exactly 0.
* indirect: a floor lit only by a wall that the light illuminates, and a
  black occluder out of view that shades the floor from the wall, at two
  bounces: every other class is exactly 0, and the indirect term equals
  central differences of the rendered images (common random numbers).
"""
import copy

import numpy as np
import pytest

from tensorray_amd import scenes

Z_MAX = 4.0


def _quad(c, u, v, n):
    """Quad with centre c, half-axes u, v, wound so that its normal is n."""
    c, u, v, n = (np.asarray(a, np.float64) for a in (c, u, v, n))
    p = [c - u - v, c + u - v, c + u + v, c - u + v]
    if np.dot(np.cross(p[1] - p[0], p[2] - p[0]), n) < 0:
        p = p[::-1]
    return np.array(p, np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _weights(W, H, seed=0):
    w = np.random.default_rng(seed).uniform(0.5, 1.5, W * H)
    return w, np.tile(w, (3, 1)).astype(np.float32)


def _z_ok(est, expect, se_expect=0.0, floor=1e-9):
    est = np.asarray(est, np.float64)
    se = np.sqrt(est.var(0, ddof=1) / len(est) + np.asarray(se_expect) ** 2)
    z = np.abs(est.mean(0) - expect) / (se + floor)
    return z, (z <= Z_MAX).all()


# ---- exact pixel coverage of a projected triangle ----
def _project(cam, W, H, p):
    """World points to raster coordinates: the camera of Matrix.cpp:181-220 /
    BaseCamera.cpp:73-88 maps raster (rx, ry) to (ratio t (1 - 2 rx / W),
    t (1 - 2 ry / H), 1) in the camera frame, t = tan(fov / 2)."""
    e = np.array(cam["pos"], np.float64)
    d = np.array(cam["target"], np.float64) - e
    d /= np.linalg.norm(d)
    u = np.array(cam["up"], np.float64)
    u /= np.linalg.norm(u)
    r = np.cross(u, d)
    r /= np.linalg.norm(r)
    nu = np.cross(d, r)
    t = np.tan(np.radians(cam["fov"]) / 2)
    q = p - e
    xc, yc = (q @ r) / (q @ d), (q @ nu) / (q @ d)
    return np.stack([W / 2 * (1 - xc / (W / H * t)), H / 2 * (1 - yc / t)], -1)


def _clip(poly, axis, val, keep_ge):
    out = []
    for i in range(len(poly)):
        a, b = poly[i], poly[(i + 1) % len(poly)]
        ina = a[axis] >= val if keep_ge else a[axis] <= val
        inb = b[axis] >= val if keep_ge else b[axis] <= val
        if ina:
            out.append(a)
        if ina != inb:
            out.append(a + (val - a[axis]) / (b[axis] - a[axis]) * (b - a))
    return out


def _area(poly):
    if len(poly) < 3:
        return 0.0
    P = np.array(poly)
    return 0.5 * abs(np.dot(P[:, 0], np.roll(P[:, 1], -1)) - np.dot(P[:, 1], np.roll(P[:, 0], -1)))


def coverage(tri2d, W, H):
    """Area of each unit pixel square covered by the 2D triangle (Sutherland-
    Hodgman clipping against the four sides)."""
    cov = np.zeros(W * H)
    lo = np.floor(tri2d.min(0)).astype(int)
    hi = np.ceil(tri2d.max(0)).astype(int)
    for py in range(max(lo[1], 0), min(hi[1], H)):
        for px in range(max(lo[0], 0), min(hi[0], W)):
            poly = list(tri2d)
            for ax, v, ge in ((0, px, True), (0, px + 1, False), (1, py, True), (1, py + 1, False)):
                poly = _clip(poly, ax, v, ge)
                if not poly:
                    break
            cov[py * W + px] = _area(poly)
    return cov


PRIM_CAM = dict(pos=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0)
PRIM_TRI = np.array([[-0.5, -0.4, 0.0], [0.6, -0.3, 0.1], [0.0, 0.5, -0.1]])


def _triangle_scene(Lt, Lb):
    """One differentiable triangle (normal towards the camera) in front of a
    6 x 6 background quad at z = -1 emitting Lb towards the camera; the
    triangle emits Lt (or nothing); every surface is black."""
    spec = scenes.SceneSpec()
    spec.materials = [scenes.Material(albedo=(0.0, 0.0, 0.0))]
    spec.meshes.append(scenes.Mesh(PRIM_TRI.astype(np.float32), np.array([[0, 1, 2]], np.uint32), material=0,
                                   emit=(Lt,) * 3 if Lt else None, flags=scenes.MESH_DIFF))
    bp, bi = _quad((0, 0, -1.0), (3.0, 0, 0), (0, 3.0, 0), (0, 0, 1))
    spec.meshes.append(scenes.Mesh(bp, bi, material=0, emit=(Lb,) * 3))
    return spec


def _coverage_gradient(W, H, wpix):
    """d(sum_pixels w coverage)/d(vertex coordinate) [3, 3] by central
    differences of the exact coverage (float64, h = 1e-6)."""
    g = np.zeros((3, 3))
    h = 1e-6
    for k in range(3):
        for c in range(3):
            tp, tm = PRIM_TRI.copy(), PRIM_TRI.copy()
            tp[k, c] += h
            tm[k, c] -= h
            dc = coverage(_project(PRIM_CAM, W, H, tp), W, H) - coverage(_project(PRIM_CAM, W, H, tm), W, H)
            g[k, c] = wpix @ dc / (2 * h)
    return g


def test_coverage_geometry():
    """The coverage helper: a projected triangle's pixel coverages sum to its
    shoelace area when it lies inside the image."""
    W = H = 12
    t2 = _project(PRIM_CAM, W, H, PRIM_TRI)
    assert (t2 > 0).all() and (t2[:, 0] < W).all() and (t2[:, 1] < H).all()
    assert abs(coverage(t2, W, H).sum() - _area(list(t2))) < 1e-9


def test_primary_edge_term_equals_coverage_derivative(oracle_mod):
    W = H = 12
    wpix, dl = _weights(W, H)
    S = oracle_mod.Scene(_triangle_scene(0.0, 1.0))
    expect = -3.0 * 1.0 * _coverage_gradient(W, H, wpix)  # 3 channels, Lb = 1
    est = [oracle_mod.boundary_adj(S, PRIM_CAM, W, H, 100 + s, 0, "primary", 512, dl)[:3] for s in range(32)]
    z, ok = _z_ok(est, expect)
    assert ok, (np.mean(est, 0), expect, z)
    assert np.abs(expect).max() > 20.0  # the derivative is not small
    for kind in ("pixel",):
        assert not oracle_mod.boundary_adj(S, PRIM_CAM, W, H, 7, 0, kind, 64, dl)[:3].any()
    assert not S.render_adj(PRIM_CAM, W, H, 7, 0, 16, dl)[0][:3].any()


def test_pixel_boundary_plus_interior_equals_emitter_coverage_derivative(oracle_mod):
    W = H = 12
    wpix, dl = _weights(W, H)
    Lt, Lb = 2.0, 1.0
    S = oracle_mod.Scene(_triangle_scene(Lt, Lb))
    dcov = 3.0 * _coverage_gradient(W, H, wpix)
    prim, near = [], []
    for s in range(32):
        prim.append(oracle_mod.boundary_adj(S, PRIM_CAM, W, H, 200 + s, 0, "primary", 512, dl)[:3])
        pix = oracle_mod.boundary_adj(S, PRIM_CAM, W, H, 300 + s, 0, "pixel", 512, dl)[:3]
        near.append(pix + S.render_adj(PRIM_CAM, W, H, 400 + s, 0, 64, dl)[0][:3])
    z, ok = _z_ok(prim, -Lb * dcov)
    assert ok, ("primary", z)
    z, ok = _z_ok(near, Lt * dcov)
    assert ok, ("pixel + interior", np.mean(near, 0), Lt * dcov, z)


# ---- direct: a moving occluder's shadow on a Lambertian ground ----
LY, OY, LH, OH, GH = 2.0, 1.0, 0.5, 0.3, 1.5
DIR_CAM = dict(pos=(0.0, 0.6, 3.5), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=25.0)


def _polygon_irradiance(x, verts, L):
    """Lambert's formula for a uniform polygonal emitter seen from points x on
    a plane with normal +y: E = L/2 |sum_i theta_i (n . g_i)|, verts [n, k, 3]
    per point."""
    u = verts - x[:, None, :]
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    E = np.zeros(len(x))
    k = verts.shape[1]
    for i in range(k):
        a, b = u[:, i], u[:, (i + 1) % k]
        th = np.arccos(np.clip((a * b).sum(1), -1, 1))
        g = np.cross(a, b)
        g /= np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-300)
        E += th * g[:, 1]
    return 0.5 * L * np.abs(E)


def _shadowed_irradiance(x, off):
    """Irradiance at ground points x from the light square minus the part the
    occluder square (centre off + (0, OY, 0), half side OH) hides: the
    occluder projects from x onto the light plane as an axis-aligned square,
    so the hidden part is an axis-aligned rectangle."""
    n = len(x)
    light = np.array([[-LH, LY, -LH], [LH, LY, -LH], [LH, LY, LH], [-LH, LY, LH]])
    E = _polygon_irradiance(x, np.broadcast_to(light, (n, 4, 3)).copy(), 10.0)
    oy = OY + off[1]
    s = (LY - x[:, 1]) / (oy - x[:, 1])
    px0 = x[:, 0] + (off[0] - OH - x[:, 0]) * s
    px1 = x[:, 0] + (off[0] + OH - x[:, 0]) * s
    pz0 = x[:, 2] + (off[2] - OH - x[:, 2]) * s
    pz1 = x[:, 2] + (off[2] + OH - x[:, 2]) * s
    ax0, ax1 = np.maximum(px0, -LH), np.minimum(px1, LH)
    az0, az1 = np.maximum(pz0, -LH), np.minimum(pz1, LH)
    ok = (ax1 > ax0) & (az1 > az0)
    if ok.any():
        m = ok.sum()
        ly = np.full(m, LY)
        R = np.stack([np.stack([ax0[ok], ly, az0[ok]], -1), np.stack([ax1[ok], ly, az0[ok]], -1),
                      np.stack([ax1[ok], ly, az1[ok]], -1), np.stack([ax0[ok], ly, az1[ok]], -1)], 1)
        E[ok] -= _polygon_irradiance(x[ok], R, 10.0)
    return E


def _shadow_scene():
    spec = scenes.SceneSpec()
    spec.materials = [scenes.Material(albedo=(0.5, 0.5, 0.5)), scenes.Material(albedo=(0.0, 0.0, 0.0))]
    g, gi, _ = scenes.quad_y(0.0, -GH, GH, -GH, GH, up=True)
    spec.meshes.append(scenes.Mesh(g, gi, material=0))
    l, li, _ = scenes.quad_y(LY, -LH, LH, -LH, LH, up=False)
    spec.meshes.append(scenes.Mesh(l, li, material=1, emit=(10.0, 10.0, 10.0)))
    o, oi, _ = scenes.quad_y(OY, -OH, OH, -OH, OH, up=True)
    spec.meshes.append(scenes.Mesh(o, oi, material=1, flags=scenes.MESH_DIFF))
    return spec


def _shadow_image(oracle_mod, W, H, off, sub=24):
    """Per-pixel mean of (rho / pi) E over a sub x sub midpoint rule."""
    fr = (np.arange(sub) + 0.5) / sub
    pix = np.arange(W * H)
    rx = ((pix % W)[:, None, None] + fr[None, :, None] + 0 * fr[None, None, :]).ravel()
    ry = ((pix // W)[:, None, None] + 0 * fr[None, :, None] + fr[None, None, :]).ravel()
    d = oracle_mod.camera_dirs(DIR_CAM, W, H, np.stack([rx, ry], 1).astype(np.float32)).astype(np.float64)
    o = np.array(DIR_CAM["pos"], np.float64)
    t = -o[1] / d[:, 1]
    x = o[None, :] + t[:, None] * d
    hit = (d[:, 1] < 0) & (np.abs(x[:, 0]) < GH) & (np.abs(x[:, 2]) < GH)
    rad = np.zeros(len(x))
    rad[hit] = 0.5 / np.pi * _shadowed_irradiance(x[hit], off)
    return rad.reshape(W * H, sub * sub).mean(1)


def test_direct_edge_term_equals_shadow_derivative(oracle_mod):
    W = H = 16
    wpix, dl = _weights(W, H)
    spec = _shadow_scene()
    S = oracle_mod.Scene(spec)
    occ = slice(8, 12)  # the occluder's four vertices
    # the closed form reproduces the rendered image (the estimator is unbiased)
    film = S.render_film(DIR_CAM, W, H, 7, 1, 2048)[0].astype(np.float64)
    img = _shadow_image(oracle_mod, W, H, np.zeros(3))
    assert abs(film.sum() - img.sum()) < 0.01 * img.sum()
    h = 1e-4
    expect = np.zeros(3)
    for ax in range(3):
        e = np.zeros(3)
        e[ax] = h
        expect[ax] = 3.0 * wpix @ (_shadow_image(oracle_mod, W, H, e) - _shadow_image(oracle_mod, W, H, -e)) / (2 * h)
    est = [oracle_mod.boundary_adj(S, DIR_CAM, W, H, 100 + s, 1, "direct", 4096, dl)[occ].sum(0) for s in range(24)]
    z, ok = _z_ok(est, expect)
    assert ok, (np.mean(est, 0), expect, z)
    assert abs(expect[0]) > 0.5 and abs(expect[1]) > 0.5
    # nothing else moves the image
    for kind in ("primary", "pixel", "indirect"):
        assert not oracle_mod.boundary_adj(S, DIR_CAM, W, H, 7, 1, kind, 64, dl)[occ].any(), kind
    assert not S.render_adj(DIR_CAM, W, H, 7, 1, 16, dl)[0][occ].any()


# ---- indirect: an occluder between a floor and the wall that lights it ----
IND_CAM = dict(pos=(-2.5, 0.3, 0.0), target=(-0.6, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=20.0)


def _indirect_scene(off=(0.0, 0.0, 0.0)):
    """A light at x = 0 facing +x, a white wall at x = 2 facing it, a white
    floor behind the light's plane (x < 0, never lit directly) and a black
    occluder above the floor, out of the camera's view and behind the light's
    plane: the floor sees the wall's light past the occluder only."""
    spec = scenes.SceneSpec()
    spec.materials = [scenes.Material(albedo=(0.8, 0.8, 0.8)), scenes.Material(albedo=(0.0, 0.0, 0.0))]
    l, li = _quad((0, 1.5, 0), (0, 0.5, 0), (0, 0, 0.5), (1, 0, 0))
    spec.meshes.append(scenes.Mesh(l, li, material=1, emit=(10.0, 10.0, 10.0)))
    w, wi = _quad((2, 1.5, 0), (0, 1.5, 0), (0, 0, 2), (-1, 0, 0))
    spec.meshes.append(scenes.Mesh(w, wi, material=0))
    f, fi = _quad((-1.5, 0, 0), (1.5, 0, 0), (0, 0, 2), (0, 1, 0))
    spec.meshes.append(scenes.Mesh(f, fi, material=0))
    o, oi = _quad(np.array((-0.6, 0.5, 0.0)) + np.asarray(off), (0.2, 0, 0), (0, 0, 0.2), (0, 1, 0))
    spec.meshes.append(scenes.Mesh(o, oi, material=1, flags=scenes.MESH_DIFF))
    return spec


def test_indirect_edge_term_equals_image_finite_differences(oracle_mod):
    W = H = 16
    B = 2
    _, dl = _weights(W, H)
    S = oracle_mod.Scene(_indirect_scene())
    occ = slice(12, 16)
    for kind in ("primary", "pixel", "direct"):
        assert not oracle_mod.boundary_adj(S, IND_CAM, W, H, 7, B, kind, 64, dl)[occ].any(), kind
    assert not S.render_adj(IND_CAM, W, H, 7, B, 16, dl)[0][occ].any()
    est = [oracle_mod.boundary_adj(S, IND_CAM, W, H, 100 + s, B, "indirect", 2048, dl)[occ].sum(0)
           for s in range(24)]
    h = 0.02
    fd = []
    for s in range(8):
        row = []
        for ax in range(2):  # along x and y; z is zero by symmetry
            e = np.zeros(3)
            e[ax] = h
            fp = oracle_mod.Scene(_indirect_scene(e)).render_film(IND_CAM, W, H, 500 + s, B, 2048)
            fm = oracle_mod.Scene(_indirect_scene(-e)).render_film(IND_CAM, W, H, 500 + s, B, 2048)
            row.append(((fp - fm).astype(np.float64) * dl).sum() / (2 * h))
        fd.append(row)
    fd = np.array(fd)
    se_fd = fd.std(0, ddof=1) / np.sqrt(len(fd))
    z, ok = _z_ok(np.array(est)[:, :2], fd.mean(0), se_fd)
    assert ok, (np.mean(est, 0), fd.mean(0), se_fd, z)
    assert (np.abs(fd.mean(0)) > 10 * se_fd).all()  # the term is far from noise
