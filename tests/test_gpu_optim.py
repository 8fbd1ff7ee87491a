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
