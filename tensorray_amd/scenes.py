"""Synthetic scenes of the benchmark workloads (seeded, deterministic).

A `SceneSpec` is a plain description -- meshes (positions, triangles, uvs,
material, emission, flags), materials and textures -- that is uploaded to the
device through the C ABI (`tensorray_amd._capi.DeviceScene`).  The reference
loads Mitsuba-like XML (SceneLoader.cpp:176-410); the workloads here are
generated:

  * icosphere(subdiv): subdivided icosahedron, 20 * 4^k faces;
  * displace(): vertices moved along their normals by seeded 3-octave value
    noise (trilinear lattice noise with Philox lattice values);
  * quad_y(): axis-aligned quads (ground, 1x1 area light);
  * perlin_texture(): seeded gradient-noise RGB texture.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

MAT_DIFFUSE = 0
MAT_MICROFACET = 1
MAT_ROUGHCONDUCTOR = 2
MESH_SMOOTH = 1
MESH_DIFF = 2


# ---------------------------------------------------------------------------
# Philox4x32-10 (vectorized, used here only as a lattice hash)
# ---------------------------------------------------------------------------
def philox4x32(c0, c1, c2, c3, k0, k1):
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF for x in (c0, c1, c2, c3))
    k0 = np.uint64(k0) & np.uint64(0xFFFFFFFF)
    k1 = np.uint64(k1) & np.uint64(0xFFFFFFFF)
    M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    W0, W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        c0, c1, c2, c3 = ((p1 >> np.uint64(32)) ^ c1 ^ k0) & mask, p1 & mask, \
            ((p0 >> np.uint64(32)) ^ c3 ^ k1) & mask, p0 & mask
        k0 = (k0 + W0) & mask
        k1 = (k1 + W1) & mask
    return c0.astype(np.uint32), c1.astype(np.uint32), c2.astype(np.uint32), c3.astype(np.uint32)


@dataclass
class Mesh:
    pos: np.ndarray                 # [V,3] float32
    idx: np.ndarray                 # [T,3] uint32
    material: int = 0
    emit: Optional[tuple] = None    # RGB radiance of an area emitter
    flags: int = 0                  # MESH_SMOOTH | MESH_DIFF
    uv: Optional[np.ndarray] = None       # [nuv,2]
    uv_idx: Optional[np.ndarray] = None   # [T,3]


@dataclass
class Material:
    type: int = MAT_DIFFUSE
    albedo: tuple = (0.5, 0.5, 0.5)     # diffuse reflectance / microfacet kd / conductor eta
    spec: tuple = (0.04, 0.04, 0.04)    # microfacet F0 / conductor k
    roughness: float = 0.3
    texture: int = -1


@dataclass
class SceneSpec:
    meshes: List[Mesh] = field(default_factory=list)
    materials: List[Material] = field(default_factory=list)
    textures: List[np.ndarray] = field(default_factory=list)   # [h,w,3] float32

    # ---- flattened arrays (the layout of Scene::ExtractGeometry) ----
    def flat(self):
        pos, idx, mat, emitter, flags, rad = [], [], [], [], [], []
        uv, uv_idx = [], []
        any_uv = any(m.uv is not None for m in self.meshes)
        voff = uvoff = 0
        ne = 0
        for m in self.meshes:
            V, T = len(m.pos), len(m.idx)
            pos.append(np.asarray(m.pos, np.float32))
            idx.append(np.asarray(m.idx, np.uint32) + np.uint32(voff))
            mat.append(np.full(T, m.material, np.int32))
            if m.emit is not None:
                emitter.append(np.full(T, ne, np.int32))
                rad.append(np.asarray(m.emit, np.float32))
                ne += 1
            else:
                emitter.append(np.full(T, -1, np.int32))
            flags.append(np.full(T, m.flags, np.int32))
            if any_uv:
                if m.uv is not None:
                    uv.append(np.asarray(m.uv, np.float32))
                    uv_idx.append(np.asarray(m.uv_idx, np.uint32) + np.uint32(uvoff))
                    uvoff += len(m.uv)
                else:
                    uv.append(np.zeros((1, 2), np.float32))
                    uv_idx.append(np.full((T, 3), uvoff, np.uint32))
                    uvoff += 1
            voff += V
        out = dict(
            pos=np.ascontiguousarray(np.concatenate(pos)),
            idx=np.ascontiguousarray(np.concatenate(idx)),
            tri_mat=np.concatenate(mat), tri_emitter=np.concatenate(emitter),
            tri_flags=np.concatenate(flags),
            emit_rad=np.ascontiguousarray(np.stack(rad)) if rad else np.zeros((0, 3), np.float32),
        )
        if any_uv:
            out["uv"] = np.ascontiguousarray(np.concatenate(uv))
            out["uv_idx"] = np.ascontiguousarray(np.concatenate(uv_idx))
        return out

    def material_table(self):
        """[M,12] float32: type, texture, roughness, 0, albedo rgb, spec rgb, 0, 0."""
        t = np.zeros((len(self.materials), 12), np.float32)
        for i, m in enumerate(self.materials):
            t[i, 0] = m.type
            t[i, 1] = m.texture
            t[i, 2] = m.roughness
            t[i, 4:7] = m.albedo
            t[i, 7:10] = m.spec
        return t

    def diff_vertex_mask(self):
        mask = []
        for m in self.meshes:
            mask.append(np.full(len(m.pos), bool(m.flags & MESH_DIFF)))
        return np.concatenate(mask)

    def num_vertices(self):
        return sum(len(m.pos) for m in self.meshes)

    def num_triangles(self):
        return sum(len(m.idx) for m in self.meshes)

    def to_module(self, cam, W, H):
        """The same scene as a `TensorRay.Scene` (the reference-style module),
        built through its programmatic API, with one camera `cam` (a dict of
        pos, target, up, fov) at W x H.  Shapes and BSDFs keep this spec's
        order, so shapes[i] is meshes[i]."""
        from tensorray_amd import TensorRay as TR
        s = TR.Scene()
        for m in self.materials:
            tex = None if m.texture < 0 else self.textures[m.texture]
            s.add_bsdf(m.type, list(m.albedo), list(m.spec), m.roughness, tex)
        for mesh in self.meshes:
            s.add_mesh(np.asarray(mesh.pos, np.float32), np.asarray(mesh.idx, np.uint32), mesh.material,
                       None if mesh.uv is None else np.asarray(mesh.uv, np.float32),
                       None if mesh.uv_idx is None else np.asarray(mesh.uv_idx, np.uint32),
                       None if mesh.emit is None else list(mesh.emit), bool(mesh.flags & MESH_SMOOTH))
        s.add_camera(TR.Camera(TR.Vector3f(*cam["pos"]), TR.Vector3f(*cam["target"]), TR.Vector3f(*cam["up"]),
                               W, H, cam["fov"]))
        return s


# ---------------------------------------------------------------------------
# Generators
# ---------------------------------------------------------------------------
def icosphere(subdiv, radius=1.0, center=(0.0, 0.0, 0.0)):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        mid = {}

        def m(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[key] = len(v) - 1
            return mid[key]

        nf = []
        for (a, b, c) in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    pos = np.array(v, np.float64) * radius + np.asarray(center, np.float64)
    return pos.astype(np.float32), np.array(f, np.uint32)


def geodesic_sphere(freq, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Icosahedron with every face split into freq x freq triangles, projected
    onto the sphere: 20 * freq^2 triangles (freq 100: 200,000)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v0 = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                   (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    f0 = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
                   (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
                   (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)])
    n = int(freq)
    # lattice (i, j), i + j <= n, of one face: point = A + i/n (B - A) + j/n (C - A)
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    keep = ii + jj <= n
    ii, jj = ii[keep], jj[keep]
    lid = -np.ones((n + 1, n + 1), np.int64)
    lid[ii, jj] = np.arange(len(ii))
    a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    up = a + b <= n - 1                      # (i,j) (i+1,j) (i,j+1)
    dn = a + b <= n - 2                      # (i+1,j) (i+1,j+1) (i,j+1)
    tri_up = np.stack([lid[a[up], b[up]], lid[a[up] + 1, b[up]], lid[a[up], b[up] + 1]], 1)
    tri_dn = np.stack([lid[a[dn] + 1, b[dn]], lid[a[dn] + 1, b[dn] + 1], lid[a[dn], b[dn] + 1]], 1)
    local = np.concatenate([tri_up, tri_dn])
    pts, faces = [], []
    for k, (A, B, C) in enumerate(f0):
        P = v0[A] + (ii[:, None] / n) * (v0[B] - v0[A]) + (jj[:, None] / n) * (v0[C] - v0[A])
        pts.append(P / np.linalg.norm(P, axis=1, keepdims=True))
        faces.append(local + k * len(ii))
    P = np.concatenate(pts)
    F = np.concatenate(faces)
    # merge the copies of points on shared icosahedron edges
    key = np.round(P * 1e6).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)                # keep first-seen order
    remap = np.empty(len(order), np.int64)
    remap[order] = np.arange(len(order))
    pos = P[np.sort(first)] * radius + np.asarray(center, np.float64)
    idx = remap[inv.ravel()][F]
    return pos.astype(np.float32), idx.astype(np.uint32)


def sphere_uv(pos, center=(0.0, 0.0, 0.0)):
    """Longitude / latitude texture coordinates of points around `center`."""
    p = pos.astype(np.float64) - np.asarray(center, np.float64)
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    u = 0.5 + np.arctan2(p[:, 2], p[:, 0]) / (2.0 * np.pi)
    v = np.arccos(np.clip(p[:, 1], -1.0, 1.0)) / np.pi
    return np.stack([u, v], 1).astype(np.float32)


def _value_noise(p, seed, octave):
    fl = np.floor(p)
    i = fl.astype(np.int64)
    tt = p - fl
    tt = tt * tt * (3.0 - 2.0 * tt)
    acc = np.zeros(len(p))
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                x0, _, _, _ = philox4x32((i[:, 0] + a) & 0xFFFFFFFF, (i[:, 1] + b) & 0xFFFFFFFF,
                                         (i[:, 2] + c) & 0xFFFFFFFF, np.full(len(p), octave), seed, 0x5EED)
                val = (x0 >> np.uint32(8)).astype(np.float64) / 16777216.0
                wx = tt[:, 0] if a else 1.0 - tt[:, 0]
                wy = tt[:, 1] if b else 1.0 - tt[:, 1]
                wz = tt[:, 2] if c else 1.0 - tt[:, 2]
                acc += wx * wy * wz * val
    return acc


def displace(pos, amplitude, seed, octaves=3, base_freq=2.0, center=(0.0, 0.0, 0.0)):
    """p' = p + amplitude * noise(p) * n(p), n = radial direction, noise in [-1, 1)."""
    c = np.asarray(center, np.float64)
    p = pos.astype(np.float64) - c
    r = np.linalg.norm(p, axis=1, keepdims=True)
    n = p / r
    acc = np.zeros(len(p))
    wsum = 0.0
    for o in range(octaves):
        fr = base_freq * 2.0 ** o
        acc += 2.0 ** -o * _value_noise(p * fr + np.array([17.0, 31.0, 47.0]), seed, o)
        wsum += 2.0 ** -o
    d = amplitude * (2.0 * acc / wsum - 1.0)
    return (c + p + d[:, None] * n).astype(np.float32)


def quad_y(y, x0, x1, z0, z1, up=True):
    pos = np.array([[x0, y, z0], [x1, y, z0], [x1, y, z1], [x0, y, z1]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]] if up else [[0, 1, 2], [0, 2, 3]], np.uint32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    return pos, idx, uv


def perlin_texture(w, h, seed, octaves=4, base=8.0):
    """Gradient noise (Perlin 1985) with quintic fade; [h, w, 3] in [0, 1]."""
    tex = np.zeros((h, w, 3), np.float64)
    jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    for ch in range(3):
        acc = np.zeros((h, w))
        amp, freq = 0.5, base
        for o in range(octaves):
            x = (ii + 0.5) / w * freq
            y = (jj + 0.5) / h * freq
            fx, fy = np.floor(x), np.floor(y)
            ix, iy = fx.astype(np.int64), fy.astype(np.int64)
            tx, ty = x - fx, y - fy

            # gradient angles on the lattice points the octave touches
            nx, ny = int(fx.max()) + 2, int(fy.max()) + 2
            lx, ly = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
            r, _, _, _ = philox4x32(lx.ravel().astype(np.uint64), ly.ravel().astype(np.uint64),
                                    np.full(lx.size, ch), np.full(lx.size, o), seed, 0x7E47)
            ang = ((r >> np.uint32(8)).astype(np.float64) / 16777216.0 * 2.0 * np.pi).reshape(nx, ny)

            def grad(gx, gy, dx, dy):
                a = ang[gx, gy]
                return np.cos(a) * dx + np.sin(a) * dy

            n00 = grad(ix, iy, tx, ty)
            n10 = grad(ix + 1, iy, tx - 1, ty)
            n01 = grad(ix, iy + 1, tx, ty - 1)
            n11 = grad(ix + 1, iy + 1, tx - 1, ty - 1)
            u = tx * tx * tx * (tx * (tx * 6 - 15) + 10)
            v = ty * ty * ty * (ty * (ty * 6 - 15) + 10)
            nx0 = n00 + (n10 - n00) * u
            nx1 = n01 + (n11 - n01) * u
            acc += amp * (nx0 + (nx1 - nx0) * v)
            amp *= 0.5
            freq *= 2.0
        tex[:, :, ch] = np.clip(0.5 + 0.6 * acc, 0.0, 1.0)
    return tex.astype(np.float32)


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
CAMERA_DEFAULT = dict(pos=(0.0, 1.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0)


def sphere_scene(subdiv=2, amplitude=0.0, noise_seed=1, material=None, diff=True, smooth=False,
                 ground_albedo=(0.5, 0.5, 0.5), light_radiance=(10.0, 10.0, 10.0), texture_res=0,
                 texture_seed=7, freq=0):
    """Icosphere (radius 1 at the origin, optionally displaced) on a ground quad
    at y = -1.05 under a 1x1 area light at y = 2 facing down.

    material: Material for the sphere (default Lambertian albedo 0.5).
    texture_res > 0 adds a seeded Perlin albedo texture on the ground.
    freq > 0 uses a geodesic sphere of 20 * freq^2 triangles instead of the
    subdivided icosphere.
    """
    spec = SceneSpec()
    spec.materials.append(material or Material(MAT_DIFFUSE, albedo=(0.5, 0.5, 0.5)))
    gmat = Material(MAT_DIFFUSE, albedo=ground_albedo)
    if texture_res:
        spec.textures.append(perlin_texture(texture_res, texture_res, texture_seed))
        gmat.texture = 0
    spec.materials.append(gmat)
    spec.materials.append(Material(MAT_DIFFUSE, albedo=(0.0, 0.0, 0.0)))
    pos, idx = geodesic_sphere(freq) if freq else icosphere(subdiv)
    if amplitude:
        pos = displace(pos, amplitude, noise_seed)
    flags = (MESH_DIFF if diff else 0) | (MESH_SMOOTH if smooth else 0)
    spec.meshes.append(Mesh(pos, idx, material=0, flags=flags))
    g, gi, guv = quad_y(-1.05, -5.0, 5.0, -5.0, 5.0, up=True)
    spec.meshes.append(Mesh(g, gi, material=1, uv=guv * 4.0, uv_idx=gi.copy()))
    l, li, _ = quad_y(2.0, -0.5, 0.5, -0.5, 0.5, up=False)
    spec.meshes.append(Mesh(l, li, material=2, emit=tuple(light_radiance)))
    return spec


def _quad(p0, p1, p2, p3):
    """Quad p0 p1 p2 p3 (counter-clockwise seen from the side its normal
    points to): triangles (0, 1, 2), (0, 2, 3)."""
    return np.array([p0, p1, p2, p3], np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


BOX_CAMERA = dict(pos=(0.0, 0.0, 3.4), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0)


def box_scene(subdiv=2, amplitude=0.0, noise_seed=1, material=None):
    """Closed room [-1.5, 1.5]^3 (open towards +z) with a 0.8 x 0.8 area light
    under the ceiling, coloured side walls and an icosphere of radius 0.6.
    Inter-reflections make every boundary term (incl. the indirect secondary
    edges) non-zero.  Use with BOX_CAMERA."""
    spec = SceneSpec()
    spec.materials = [material or Material(MAT_DIFFUSE, albedo=(0.6, 0.6, 0.6)),
                      Material(MAT_DIFFUSE, albedo=(0.7, 0.7, 0.7)),
                      Material(MAT_DIFFUSE, albedo=(0.7, 0.2, 0.2)),
                      Material(MAT_DIFFUSE, albedo=(0.2, 0.7, 0.2)),
                      Material(MAT_DIFFUSE, albedo=(0.0, 0.0, 0.0))]
    pos, idx = icosphere(subdiv, radius=0.6, center=(0.0, -0.6, 0.0))
    if amplitude:
        pos = displace(pos, amplitude, noise_seed, center=(0.0, -0.6, 0.0))
    spec.meshes.append(Mesh(pos, idx, material=0, flags=MESH_DIFF))
    a = 1.5
    walls = [
        (_quad((-a, -a, -a), (-a, -a, a), (a, -a, a), (a, -a, -a)), 1),   # floor, +y
        (_quad((-a, a, -a), (a, a, -a), (a, a, a), (-a, a, a)), 1),       # ceiling, -y
        (_quad((-a, -a, -a), (a, -a, -a), (a, a, -a), (-a, a, -a)), 1),   # back, +z
        (_quad((-a, -a, -a), (-a, a, -a), (-a, a, a), (-a, -a, a)), 2),   # left, +x
        (_quad((a, -a, -a), (a, -a, a), (a, a, a), (a, a, -a)), 3),       # right, -x
    ]
    for (p, i), m in walls:
        spec.meshes.append(Mesh(p, i, material=m))
    l = 0.4
    lp, li = _quad((-l, a - 0.01, -l), (l, a - 0.01, -l), (l, a - 0.01, l), (-l, a - 0.01, l))
    spec.meshes.append(Mesh(lp, li, material=4, emit=(12.0, 12.0, 12.0)))
    return spec


def config1_scene():
    """64x64, 16 spp, depth 2: icosphere subdiv 2 (320 tris), Lambert 0.5."""
    return sphere_scene(subdiv=2, amplitude=0.0)


def metric_scene(amplitude=0.05, noise_seed=1):
    """Icosphere subdiv 5 (20,480 tris) displaced by 3-octave value noise
    (amplitude 0.05), GGX microfacet (roughness 0.3, F0 0.04, kd 0)."""
    mat = Material(MAT_MICROFACET, albedo=(0.0, 0.0, 0.0), spec=(0.04, 0.04, 0.04), roughness=0.3)
    return sphere_scene(subdiv=5, amplitude=amplitude, noise_seed=noise_seed, material=mat)


def fibonacci_cameras(n, radius=3.0, seed=0, target=(0.0, 0.0, 0.0), fov=45.0):
    """n camera positions on a sphere of `radius` around `target`: a
    Fibonacci lattice rotated by a seeded random angle (numpy PCG64(seed))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    phase = rng.uniform(0.0, 2.0 * np.pi)
    ga = np.pi * (3.0 - 5.0 ** 0.5)
    out = []
    for k in range(n):
        y = 1.0 - 2.0 * (k + 0.5) / n
        r = (1.0 - y * y) ** 0.5
        th = phase + ga * k
        d = np.array([r * np.cos(th), y, r * np.sin(th)])
        pos = np.asarray(target, np.float64) + radius * d
        up = (0.0, 1.0, 0.0) if abs(y) < 0.99 else (1.0, 0.0, 0.0)
        out.append(dict(pos=tuple(float(x) for x in pos), target=tuple(float(x) for x in target), up=up,
                        fov=float(fov)))
    return out


def multiview_scene(freq=100, amplitude=0.05, noise_seed=1, texture_res=1024, texture_seed=7):
    """Geodesic sphere of 20 * freq^2 triangles (freq 100: 200,000) displaced
    by 3-octave value noise, Lambertian with a seeded Perlin albedo texture
    (texture_res^2 RGB float32, longitude / latitude mapped), under a 1x1
    area light at y = 2.  No ground, so that cameras all around see it."""
    spec = SceneSpec()
    spec.textures.append(perlin_texture(texture_res, texture_res, texture_seed))
    spec.materials.append(Material(MAT_DIFFUSE, albedo=(1.0, 1.0, 1.0), texture=0))
    spec.materials.append(Material(MAT_DIFFUSE, albedo=(0.0, 0.0, 0.0)))
    pos, idx = geodesic_sphere(freq)
    uv = sphere_uv(pos)
    if amplitude:
        pos = displace(pos, amplitude, noise_seed)
    spec.meshes.append(Mesh(pos, idx, material=0, flags=MESH_DIFF, uv=uv, uv_idx=idx.copy()))
    l, li, _ = quad_y(2.0, -0.5, 0.5, -0.5, 0.5, up=False)
    spec.meshes.append(Mesh(l, li, material=1, emit=(10.0, 10.0, 10.0)))
    return spec
