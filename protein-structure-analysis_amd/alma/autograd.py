"""torch.autograd.Function over the device energy and its backward kernel.

`EnergyModel(q, rstar, eps, topo_csr)(x)` returns the total energy of the
calculate_potential form (src/physics_engine.cpp:2242-2248: fresh neighbour
list, HCT Born radii, DH + GB + LJ + hard core within 12 A, SASA, dihedral
and disulfide terms) for coordinates x (n x 3, float64, any device), and its
backward returns dE/dx from the hand-written gradient kernels, including the
chain rule through the Born radii.  The Born radii are recomputed for every
call; nothing from one evaluation is reused by the next.

The kernels read coordinates through the C-ABI, which takes host arrays, so
x crosses PCIe once per call (n x 24 bytes) and the gradient comes back the
same way; the energy evaluation itself runs on the device.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native


class _EnergyFn(torch.autograd.Function):
    @staticmethod
    def forward(fctx, x, model):
        xyz = x.detach().to("cpu", torch.float64).contiguous().numpy()
        model.ctx.set_positions(xyz)
        e, g, gt = model.ctx.forces()
        model.last_dEdtheta = gt
        fctx.save_for_backward(torch.from_numpy(g).to(x.device, x.dtype))
        return x.new_tensor(e)

    @staticmethod
    def backward(fctx, grad_out):
        (g,) = fctx.saved_tensors
        return grad_out * g, None


class EnergyModel(torch.nn.Module):
    """Total energy of one system as a differentiable function of its
    coordinates.  `last_dEdtheta` holds the torsion gradients (one per
    rotatable bond) of the most recent call."""

    def __init__(self, q, rstar, eps, topo_csr=None, device: int = 0):
        super().__init__()
        self.ctx = _native.Context(device)
        if not _native.shares_torch_runtime():
            raise RuntimeError("libalma was loaded before torch: import torch before the alma "
                               "package so both use torch's HIP runtime")
        self.q = np.ascontiguousarray(q, np.float64)
        self.rstar = np.ascontiguousarray(rstar, np.float64)
        self.eps = np.ascontiguousarray(eps, np.float64)
        self.topo_csr = topo_csr
        self.ctx.set_system(np.zeros((len(self.q), 3)), self.q, self.rstar, self.eps, topo_csr)
        self.last_dEdtheta = None

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        if x.shape != (len(self.q), 3):
            raise ValueError("coordinates must have shape (%d, 3)" % len(self.q))
        return _EnergyFn.apply(x, self)
