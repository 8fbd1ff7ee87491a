"""Vertex and roughness gradients of the single-view workload with indirect
light, through the torch.autograd.Function, against central finite
differences on 32 random vertices.

Workload: 256 x 256, max depth 4 (max_bounces 3), 64 spp; an icosphere of
subdivision 5 (20,480 triangles) displaced along its normals by seeded
3-octave value noise (amplitude 0.05), GGX microfacet (roughness 0.3,
F0 0.04), over a ground quad under a 1 x 1 area light.  The gradient is the
sum of the interior integrator and the pixel, primary, direct and indirect
edge integrators (sppe = sppse0 = sppse1 = sppe0 = 64), taken by
`RenderSingleView` (renderer_single_view.py:43-85 of the reference) for the
loss <w, I> with a seeded weight image w.  The finite differences are
central, eps 1e-3, of the same loss rendered at 512 spp with one seed for
both sides, at four seeds.

Both sides carry Monte Carlo error, and in the glossy highlight the
per-sample vertex derivatives are heavy tailed (a 64 spp gradient of one
vertex there varies by tens of percent between seeds), so the check is
statistical: the gradient is the mean of K independent 64 spp forward +
adjoint passes, its standard error comes from their spread, the finite
difference's from the spread of its four seeds, and the test bounds the
normalised residuals z = (g - fd) / sqrt(se_g^2 + se_fd^2) besides the
correlation and the relative L2 error over the 96 coordinates.  With only
three degrees of freedom in se_fd, z has heavy tails, so the tails are
bounded by a count (at most 3 of 96 beyond 4) and not by the maximum.
Measured on an MI355X: relative L2 0.199, correlation 0.982, mean z^2 1.32,
max |z| 3.71; roughness -628.9 +- 7.2 against -614 +- 41.
"""
import numpy as np
import pytest
import torch

from tensorray_amd import scenes

pytestmark = pytest.mark.gpu

W = H = 256
MAX_BOUNCES = 3
SPP = 64
K = 64                 # independent gradient passes
FD_SPP = 512
FD_SEEDS = (7, 8, 9, 10)
EPS = 1e-3


@pytest.fixture
def tr_env():
    from tensorray_amd import TensorRay as TR
    TR.env_release()
    TR.env_create(0)
    yield TR
    TR.env_release()


def test_vertex_and_roughness_gradients_match_fd(tr_env):
    TR = tr_env
    from tensorray_amd.pyTensorRay.parameter_manager import SceneParameterManager
    from tensorray_amd.pyTensorRay.renderer_single_view import RenderSingleView
    spec = scenes.metric_scene()
    s = spec.to_module(scenes.CAMERA_DEFAULT, W, H)
    s.shapes[0].diff_all_vertex_pos()          # pool entry 0: [3 * V], x-, y-, z-planes
    s.bsdfs[0].diff_texture("roughness")       # pool entry 1: [1]
    s.configure()
    pm = SceneParameterManager([0])            # one group: vertices then roughness
    pm.get_values()
    x0 = pm.params[0].detach().clone()
    nv = (x0.numel() - 1) // 3
    assert nv == 10242 and abs(x0[-1].item() - 0.3) < 1e-6

    rng = np.random.Generator(np.random.PCG64(3))
    w = torch.tensor(rng.uniform(0.5, 1.5, (H, W, 3)), dtype=torch.float32)
    # 32 random vertices of the side the camera sees
    pos = x0[:3 * nv].view(3, nv).t().numpy().astype(np.float64)
    nrm = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    facing = np.nonzero(((np.array(scenes.CAMERA_DEFAULT["pos"]) - pos) * nrm).sum(1) > 0.2)[0]
    sel = np.sort(rng.choice(facing, 32, replace=False))
    coords = [(k * nv + v) for v in sel for k in range(3)] + [x0.numel() - 1]

    edges = [TR.PixelBoundaryIntegrator(), TR.PrimaryEdgeIntegrator(), TR.DirectEdgeIntegrator(),
             TR.IndirectEdgeIntegrator()]
    r = RenderSingleView(s, pm, TR.PathTracer(), edges, TR.RenderOptions(1234, MAX_BOUNCES, SPP, SPP, SPP, SPP, SPP))
    g = np.zeros((K, len(coords)))
    for i in range(K):
        img = r(pm.params[0], {"seed": 100 + i})
        loss = (img * w).sum()
        pm.params[0].grad = None
        loss.backward()
        g[i] = pm.params[0].grad.double().numpy()[coords]
    g_mean, g_se = g.mean(0), g.std(0, ddof=1) / np.sqrt(K)

    plain = RenderSingleView(s, pm, TR.PathTracer(), [], TR.RenderOptions(1234, MAX_BOUNCES, FD_SPP, 0, 0, 0, 0))
    wd = w.double()

    def loss_at(x, seed):
        with torch.no_grad():
            pm.params[0].copy_(x)
        pm.set_values()
        s.configure()
        return float((plain(pm.params[0], {"seed": seed}).detach().double() * wd).sum())

    fd = np.zeros((len(FD_SEEDS), len(coords)))
    for a, seed in enumerate(FD_SEEDS):
        for j, c in enumerate(coords):
            xp, xm = x0.clone(), x0.clone()
            xp[c] += EPS
            xm[c] -= EPS
            fd[a, j] = (loss_at(xp, seed) - loss_at(xm, seed)) / (2 * EPS)
    with torch.no_grad():
        pm.params[0].copy_(x0)
    pm.set_values()
    s.configure()
    fd_mean, fd_se = fd.mean(0), fd.std(0, ddof=1) / np.sqrt(len(FD_SEEDS))

    z = (g_mean - fd_mean) / np.sqrt(g_se ** 2 + fd_se ** 2 + 1e-12)
    gv, fv, zv = g_mean[:-1], fd_mean[:-1], z[:-1]
    rel = np.linalg.norm(gv - fv) / np.linalg.norm(fv)
    corr = np.corrcoef(gv, fv)[0, 1]
    print(f"vertices: rel L2 {rel:.3f}, corr {corr:.4f}, mean z^2 {np.mean(zv ** 2):.2f}, max |z| "
          f"{np.abs(zv).max():.2f}, |z| > 4: {np.sum(np.abs(zv) > 4)}; roughness: g {g_mean[-1]:.2f} +- {g_se[-1]:.2f}, fd {fd[:, -1].round(2).tolist()}")
    assert np.abs(fv).max() > 50.0           # the highlight makes some coordinates large
    assert corr >= 0.95
    assert rel <= 0.3
    assert np.mean(zv ** 2) <= 3.0           # 1 for unbiased estimates with well-estimated errors
    assert np.sum(np.abs(zv) > 4.0) <= 3
    assert abs(z[-1]) <= 4.0
    assert abs(g_mean[-1] - fd_mean[-1]) <= 0.1 * abs(fd_mean[-1])
