"""Scene parameter groups over TensorRay's parameter pool
(pyTensorRay/parameter_manager.py of the reference).

Layout of the unknowns: | shape params | BSDF params | ..., one torch
tensor per group; group i covers pool entries [offsets[i], offsets[i+1]).
"""
import torch

from tensorray_amd import TensorRay as TR
from tensorray_amd.pyTensorRay.common import get_ptr


class SceneParameterManager:
    def __init__(self, param_id_offsets=(0,)):
        self.param_id_offsets = list(param_id_offsets)
        self.param_group_count = len(self.param_id_offsets)
        self.param_count_list = []
        self.total_param_count = TR.get_size_param(self.param_count_list)
        self.param_counts = []
        self.params = []
        for g in range(self.param_group_count):
            lo, hi = self._range(g)
            self.param_counts.append(sum(self.param_count_list[lo:hi]))
            self.params.append(torch.zeros(self.param_counts[g], requires_grad=True))

    def _range(self, g):
        lo = self.param_id_offsets[g]
        hi = self.param_id_offsets[g + 1] if g + 1 < self.param_group_count else len(self.param_count_list)
        return lo, hi

    def get_values(self):
        """Pool -> self.params (values)."""
        with torch.no_grad():
            for g in range(self.param_group_count):
                lo, hi = self._range(g)
                off = 0
                for i in range(lo, hi):
                    TR.get_var(i, get_ptr(self.params[g]), off)
                    off += self.param_count_list[i]

    def set_values(self):
        """self.params -> pool (call scene.configure() afterwards)."""
        for g in range(self.param_group_count):
            lo, hi = self._range(g)
            p = self.params[g].detach().contiguous()
            off = 0
            for i in range(lo, hi):
                TR.set_var(i, get_ptr(p), off)
                off += self.param_count_list[i]

    def get_grads(self):
        """Gradients accumulated by the renderD calls, one tensor per group
        (read and cleared)."""
        res = []
        for g in range(self.param_group_count):
            lo, hi = self._range(g)
            grad = torch.zeros(self.param_counts[g])
            off = 0
            for i in range(lo, hi):
                TR.get_grad(i, get_ptr(grad), off, True)
                off += self.param_count_list[i]
            res.append(grad)
        return res

    def __str__(self):
        return (f"Total param count = {self.total_param_count}\nParam list:{self.param_count_list}\n"
                f"Param group count = {self.param_group_count}\nGrouped params:{self.param_counts}")
