"""LargeSteps optimizer (pyTensorRay/optimizer.py:147-180 of the reference)
on the device.

The reference builds M = I + lmbda L (largesteps' compute_matrix, uniform
Laplacian), optimizes u = M x with largesteps' AdamUniform and maps back
with a Cholesky solve (cholespy).  Here M lives in the native library
(trc_lsolver: CSR matrix, conjugate-gradient solves on the three coordinate
columns at once) and the uniform Adam step is trc_adam_uniform_step; the
parameter keeps the reference's layout, a flat [3 * n] tensor of x-plane,
y-plane, z-plane (optimizer.py:170-172).
"""
import torch

from tensorray_amd import _capi


class LargeSteps(torch.optim.Optimizer):
    def __init__(self, V, F, scene_param_manager, lr=0.1, betas=(0.9, 0.999), lmbda=0.1, ctx=None,
                 max_iter=400, tol=1e-6):
        """V: [params] with params[0] the flat [3 * n] vertex positions;
        F: [T][3] face indices of the mesh (torch or numpy)."""
        self.scene_param_manager = scene_param_manager
        self.V = V[0]
        n = self.V.numel() // 3
        self.ctx = ctx or _capi.Context(torch.cuda.current_device())
        faces = F.detach().cpu().numpy() if torch.is_tensor(F) else F
        self.solver = self.ctx.lsolver(faces, n, lmbda)
        self.stream = torch.cuda.current_stream()
        self.x = self.V.detach().view(3, -1).t().contiguous().float().cuda()   # [n][3]
        self.u = torch.empty_like(self.x)
        self.solver.apply(self.x.data_ptr(), self.u.data_ptr(), self.stream)
        self.gu = torch.zeros_like(self.x)
        self.m1 = torch.zeros_like(self.x)
        self.m2 = torch.zeros_like(self.x)
        self.t = 0
        self.max_iter, self.tol = max_iter, tol
        super().__init__(V, dict(lr=lr, betas=betas))

    @torch.no_grad()
    def step(self, closure=None):
        loss = closure() if closure is not None else None
        if self.V.grad is None:
            return loss
        lr = self.param_groups[0]["lr"]
        b1, b2 = self.param_groups[0]["betas"]
        g = self.V.grad.view(3, -1).t().contiguous().float().cuda()
        # V = M^-1 u, so dL/du = M^-1 dL/dV (M is symmetric)
        self.solver.solve(g.data_ptr(), self.gu.data_ptr(), self.max_iter, self.tol, self.stream)
        self.t += 1
        self.ctx.adam_uniform_step(self.u.data_ptr(), self.gu.data_ptr(), self.m1.data_ptr(), self.m2.data_ptr(),
                                   self.u.numel(), self.t, lr, b1, b2, self.stream)
        self.solver.solve(self.u.data_ptr(), self.x.data_ptr(), self.max_iter, self.tol, self.stream)
        self.V.data.copy_(self.x.t().reshape(-1).to(self.V.device))
        return loss

    def zero_grad(self, set_to_none=False):
        super().zero_grad(set_to_none=set_to_none)
