"""Device optimizer kernels against the oracle's float32 restatements:
Adam (trc_adam_step) and uniform Adam (trc_adam_uniform_step) bit for bit
over several steps with zero, NaN and infinite gradient entries; the
Laplacian system I + lam L (trc_lsolver) against scipy: its product to
float32 rounding and its conjugate-gradient solve to the tolerance asked."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grads(rng, n):
    g = (rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 1, n)).astype(np.float32)
    g[::97] = 0.0
    g[5::101] = np.nan
    g[7::103] = np.inf
    return g


@pytest.mark.parametrize("kind", ["adam", "uniform"])
def test_adam_matches_oracle(gpu_ctx, oracle_mod, kind):
    import torch
    rng = np.random.Generator(np.random.PCG64(11))
    n = 100_003
    x = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    dx, dm, dv = (torch.from_numpy(a.copy()).cuda() for a in (x, m, v))
    for t in range(1, 6):
        g = _grads(rng, n)
        dg = torch.from_numpy(g).cuda()
        if kind == "adam":
            gpu_ctx.adam_step(dx.data_ptr(), dg.data_ptr(), dm.data_ptr(), dv.data_ptr(), n, t, lr=1e-3)
            x, m, v = oracle_mod.adam_step(x, g, m, v, t, lr=1e-3)
        else:
            gpu_ctx.adam_uniform_step(dx.data_ptr(), dg.data_ptr(), dm.data_ptr(), dv.data_ptr(), n, t, lr=1e-3)
            x, m, v = oracle_mod.adam_uniform_step(x, g, m, v, t, lr=1e-3)
        torch.cuda.synchronize()
        for name, d, h in (("x", dx, x), ("m", dm, m), ("v", dv, v)):
            got = d.cpu().numpy()
            assert np.isfinite(got).all(), (t, name)
            assert np.array_equal(got, h), (t, name, np.abs(got - h).max())


def test_laplacian_solver(gpu_ctx, oracle_mod):
    import scipy.sparse.linalg as spla
    import torch
    from tensorray_amd import scenes
    pos, idx = scenes.geodesic_sphere(12)
    nv, lam = len(pos), 3.0
    M = oracle_mod.laplacian_system(idx, nv, lam)
    S = gpu_ctx.lsolver(idx, nv, lam)
    rng = np.random.Generator(np.random.PCG64(4))
    x = rng.standard_normal((nv, 3)).astype(np.float32)
    dx = torch.from_numpy(x).cuda()
    dy = torch.empty_like(dx)
    S.apply(dx.data_ptr(), dy.data_ptr())
    torch.cuda.synchronize()
    ref = M @ x.astype(np.float64)
    assert np.abs(dy.cpu().numpy() - ref).max() <= 1e-5 * np.abs(ref).max()
    b = rng.standard_normal((nv, 3)).astype(np.float32)
    db = torch.from_numpy(b).cuda()
    sol = torch.zeros_like(db)
    S.solve(db.data_ptr(), sol.data_ptr(), max_iter=500, tol=1e-6)
    iters, rel = S.status()
    assert 0 < iters < 500 and (rel <= 1e-6).all(), (iters, rel)
    xs = np.stack([spla.spsolve(M.tocsc(), b[:, c].astype(np.float64)) for c in range(3)], 1)
    assert np.abs(sol.cpu().numpy() - xs).max() <= 1e-4 * np.abs(xs).max()
    # a second solve warm-started from the answer stops at once (|r| <= tol |b| already)
    S.solve(db.data_ptr(), sol.data_ptr(), max_iter=500, tol=1e-4)
    assert S.status()[0] == 0
    with pytest.raises(Exception, match="out of range"):
        gpu_ctx.lsolver(np.array([[0, 1, nv]]), nv, lam)
    S.close()


def test_laplacian_solver_large_mesh_deterministic(gpu_ctx, oracle_mod):
    """A mesh with more rows than the solver's threads (289,002 vertices: the
    grid-stride loops and all block partials in use): the residual of the
    answer, checked in float64 on the host, is within the tolerance, and a
    second solve from the same start gives the same bits."""
    import torch
    from tensorray_amd import scenes
    pos, idx = scenes.geodesic_sphere(170)
    nv, lam = len(pos), 20.0
    assert nv > 1024 * 256
    M = oracle_mod.laplacian_system(idx, nv, lam)
    S = gpu_ctx.lsolver(idx, nv, lam)
    rng = np.random.Generator(np.random.PCG64(5))
    b = rng.standard_normal((nv, 3)).astype(np.float32)
    db = torch.from_numpy(b).cuda()
    sols = []
    for _ in range(2):
        sol = torch.zeros_like(db)
        S.solve(db.data_ptr(), sol.data_ptr(), max_iter=1000, tol=1e-5)
        torch.cuda.synchronize()
        iters, rel = S.status()
        assert 0 < iters < 1000 and (rel <= 1e-5).all(), (iters, rel)
        sols.append(sol.cpu().numpy())
    assert np.array_equal(sols[0], sols[1])
    res = M @ sols[0].astype(np.float64) - b
    rel_true = np.linalg.norm(res, axis=0) / np.linalg.norm(b.astype(np.float64), axis=0)
    assert (rel_true <= 1e-4).all(), rel_true
    S.close()


def test_pytensorray_largesteps(gpu_ctx, oracle_mod):
    """pyTensorRay's LargeSteps keeps the reference's flat [3 * n] layout:
    one step = M^-1 on the gradient, uniform Adam on u = M x, x = M^-1 u."""
    import scipy.sparse.linalg as spla
    import torch
    from tensorray_amd import scenes
    from tensorray_amd.pyTensorRay.optimizer import LargeSteps
    pos, idx = scenes.geodesic_sphere(5)
    n, lam, lr = len(pos), 2.0, 1e-2
    V = torch.from_numpy(pos.T.reshape(-1).copy()).requires_grad_(True)      # [3 * n], planes
    opt = LargeSteps([V], torch.from_numpy(idx.astype("int64")), None, lr=lr, lmbda=lam, ctx=gpu_ctx, tol=1e-7)
    rng = np.random.Generator(np.random.PCG64(8))
    g = rng.standard_normal((n, 3)).astype(np.float32)
    V.grad = torch.from_numpy(g.T.reshape(-1).copy())
    opt.step()
    torch.cuda.synchronize()
    M = oracle_mod.laplacian_system(idx, n, lam).tocsc()
    x0 = pos.astype(np.float64)
    gu = np.stack([spla.spsolve(M, g[:, c].astype(np.float64)) for c in range(3)], 1).astype(np.float32)
    u0 = (M @ x0).astype(np.float32)
    z = np.zeros_like(u0)
    u1, _, _ = oracle_mod.adam_uniform_step(u0, gu, z, z, 1, lr=lr)
    x1 = np.stack([spla.spsolve(M, u1[:, c].astype(np.float64)) for c in range(3)], 1)
    got = V.detach().numpy().reshape(3, n).T
    step = np.abs(x1 - x0).max()
    assert step > 0 and np.abs(got - x1).max() <= 1e-2 * step + 1e-6
