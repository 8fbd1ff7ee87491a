"""Device Adam step (trc_adam_step) against the oracle's float32
restatement of Adam::Step: bit for bit over several steps, with zero, NaN
and infinite gradient entries."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_adam_matches_oracle(gpu_ctx, oracle_mod):
    import torch
    rng = np.random.Generator(np.random.PCG64(11))
    n = 100_003
    x = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    dx, dm, dv = (torch.from_numpy(a.copy()).cuda() for a in (x, m, v))
    for t in range(1, 6):
        g = (rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 1, n)).astype(np.float32)
        g[::97] = 0.0
        g[5::101] = np.nan
        g[7::103] = np.inf
        dg = torch.from_numpy(g).cuda()
        gpu_ctx.adam_step(dx.data_ptr(), dg.data_ptr(), dm.data_ptr(), dv.data_ptr(), n, t, lr=1e-3)
        x, m, v = oracle_mod.adam_step(x, g, m, v, t, lr=1e-3)
        torch.cuda.synchronize()
        for name, d, h in (("x", dx, x), ("m", dm, m), ("v", dv, v)):
            got = d.cpu().numpy()
            assert np.isfinite(got).all(), (t, name)
            assert np.array_equal(got, h), (t, name, np.abs(got - h).max())
