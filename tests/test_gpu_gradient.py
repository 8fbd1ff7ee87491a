"""Total gradient of the HIP path against finite differences of its own
forward renders, at sample counts the CPU cannot reach.

The forward path is bit-exact with the oracle, so the finite differences
are those of the reference estimator.  The loss is <w, I> for a constant
adjoint image w; the parameter is a rigid translation of the differentiable
shape (every vertex of it moves by h along one axis), so dL/dh is the sum of
the vertex gradients over that shape.  Gradient = interior (trc_render_adj)
+ primary edges + pixel boundary + secondary direct + secondary indirect
(trc_render_boundary kinds 0-3).
"""
import numpy as np
import pytest

from tensorray_amd import _capi, scenes

pytestmark = pytest.mark.gpu


def _gradient_and_fd(ctx, spec, cam_d, W, H, B, spp, fd_spp, h, seed):
    import torch
    ds = _capi.DeviceScene(ctx, spec)
    info = ds.info()
    mask = spec.diff_vertex_mask()
    base = ds.vertices().copy()
    cam = _capi.make_camera(W, H, **cam_d)
    w = torch.full((3, W * H), 0.5, device="cuda")
    opt = _capi.make_options(seed=seed, max_bounces=B, spp=spp, sppe=spp, sppe0=spp, sppse0=spp, sppse1=spp)
    parts = {}
    d = torch.zeros(info.num_vertices, 3, device="cuda")
    ds.render_adj(cam, opt, w.data_ptr(), d.data_ptr(), None)
    torch.cuda.synchronize()
    parts["interior"] = d.cpu().numpy().astype(np.float64)[mask].sum(0)
    for k in ("primary", "pixel", "direct", "indirect"):
        d.zero_()
        ds.render_boundary(cam, opt, k, w.data_ptr(), d.data_ptr())
        torch.cuda.synchronize()
        parts[k] = d.cpu().numpy().astype(np.float64)[mask].sum(0)
    film = torch.empty(3, W * H, device="cuda")
    fopt = _capi.make_options(seed=seed + 1000, max_bounces=B, spp=fd_spp)

    def loss(pos):
        p = np.ascontiguousarray(pos, np.float32)
        ds.set_vertices(p.ctypes.data, len(p), True)
        ds.render_fwd(cam, fopt, film.data_ptr())
        torch.cuda.synchronize()
        return float((film.double() * w.double()).sum())

    fd = np.zeros(3)
    for k in range(3):
        dp = np.zeros(3, np.float32)
        dp[k] = h
        pp, pm = base.copy(), base.copy()
        pp[mask] += dp
        pm[mask] -= dp
        fd[k] = (loss(pp) - loss(pm)) / (2 * h)
    ds.close()
    return parts, fd


def test_total_gradient_matches_fd_with_interreflection(gpu_ctx):
    """Box scene, depth 3: the indirect edge term is needed (it carries a
    sizeable share of d/dz); total within 5% of max |FD|."""
    parts, fd = _gradient_and_fd(gpu_ctx, scenes.box_scene(subdiv=2), scenes.BOX_CAMERA, 64, 64, 3,
                                 2048, 8192, 1e-2, 1)
    total = sum(parts.values())
    scale = np.abs(fd).max()
    print("total", total, "fd", fd, {k: v.round(3).tolist() for k, v in parts.items()})
    assert scale > 10.0
    assert (np.abs(total - fd) <= 0.05 * scale).all(), (total, fd)
    # each edge class matters somewhere: removing it breaks the match
    for k in ("interior", "primary", "direct", "indirect"):
        assert (np.abs(total - parts[k] - fd) > 0.05 * scale).any(), k


def test_total_gradient_matches_fd_sphere_depth(gpu_ctx):
    """Sphere over a ground plane, depth 2, motion toward the camera."""
    parts, fd = _gradient_and_fd(gpu_ctx, scenes.sphere_scene(subdiv=2), scenes.CAMERA_DEFAULT, 64, 64, 2,
                                 2048, 8192, 1e-2, 1)
    total = sum(parts.values())
    print("total", total, "fd", fd)
    assert abs(fd[2]) > 100.0
    assert abs(total[2] - fd[2]) <= 0.02 * abs(fd[2]), (total, fd)
    assert abs(parts["interior"][2] - fd[2]) > 0.05 * abs(fd[2])
