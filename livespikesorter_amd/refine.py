"""Template refinement against drift: the residual energy of a fixed spike set as a differentiable
function of the templates (U) and the temporal basis (wPCA).

``residual_energy(sorter, x, spikes, U, wpca)`` returns L = sum R^2 with
R = X - sum_s amp_s * wave_{t_s} (wave = einsum(U, wPCA), placed at the spike's column), the
residual the pursuit leaves for its own spikes.  Its backward is liblss's adjoint kernel
(lss_residual_energy_grad): the spike-aligned residual windows correlated against U and wPCA.
Nothing of the forward pass is stored: the backward recomputes the residual from X
(checkpointing), so memory stays at the size of the inputs.  There is no CPU path; PyTorch only
supplies device memory and the autograd plumbing.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .sorter import Sorter


@dataclass
class SpikeSet:
    """A fixed spike set on the device, sorted by column (the order the kernels need)."""
    col: torch.Tensor    # int32 [n] peak column in the window
    tmp: torch.Tensor    # int32 [n]
    amp: torch.Tensor    # float32 [n]

    @staticmethod
    def from_arrays(col, tmp, amp, device="cuda:0") -> "SpikeSet":
        col = np.asarray(col, np.int64)
        o = np.argsort(col, kind="stable")
        return SpikeSet(torch.from_numpy(col[o].astype(np.int32)).to(device),
                        torch.from_numpy(np.asarray(tmp, np.int32)[o]).to(device),
                        torch.from_numpy(np.asarray(amp, np.float32)[o]).to(device))

    @staticmethod
    def from_spikes(sp, device="cuda:0") -> "SpikeSet":
        """From a SpikeList of the same window (its column, template, amplitude)."""
        return SpikeSet.from_arrays(sp.column, sp.template, sp.amplitude, device)

    def __len__(self):
        return int(self.col.numel())


class ResidualEnergy(torch.autograd.Function):
    """L(U, wPCA) for a fixed spike set; backward = the adjoint kernel (residual recomputed)."""

    @staticmethod
    def forward(ctx, U, wpca, sorter: Sorter, x_wc, spikes: SpikeSet):
        U = U.detach().contiguous()
        wpca = wpca.detach().contiguous()
        loss = sorter.residual_energy(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wpca)
        ctx.sorter, ctx.x, ctx.spikes = sorter, x_wc, spikes
        ctx.save_for_backward(U, wpca)
        return torch.tensor(loss, dtype=torch.float64, device=U.device)

    @staticmethod
    def backward(ctx, g):
        U, wpca = ctx.saved_tensors
        sp = ctx.spikes
        _, gU, gW = ctx.sorter.residual_energy_grad(ctx.x, sp.col, sp.tmp, sp.amp, U, wpca)
        s = g.to(torch.float32)
        return gU * s, gW * s, None, None, None


def residual_energy(sorter: Sorter, x_wc: torch.Tensor, spikes: SpikeSet, U: torch.Tensor,
                    wpca: torch.Tensor) -> torch.Tensor:
    """Differentiable residual energy (float64 scalar tensor on the device)."""
    return ResidualEnergy.apply(U, wpca, sorter, x_wc, spikes)


def refine_step(sorter: Sorter, x_wc, spikes: SpikeSet, U, wpca, lr: float):
    """One gradient step on (U, wPCA) with step size lr; returns (loss before, U', wpca')."""
    loss, gU, gW = sorter.residual_energy_grad(x_wc, spikes.col, spikes.tmp, spikes.amp,
                                               U.contiguous(), wpca.contiguous())
    return loss, U - lr * gU, wpca - lr * gW
