"""The slab energy of an evolved left-moving MV sheet as a differentiable function of its charge density,
behind ``torch.autograd.Function``.

forward:  rho (Nz, Ny, Nx, 8) float64 -> E_slab, computed by libglasma on the GPU (the MV sheet of
          InitialConditions::SingleSheet, SC/INITIALCONDITIONS/SUNCCURRENTS/RevSingleNucleus.cpp:7-210,
          built from rho; nsteps leapfrog steps of Dynamics::SUNc::Update, SC/DYNAMICS/SUNcDynamics.cpp:353;
          the energy of the layers |z - zc| < half);
backward: d E_slab / d rho from libglasma's reverse-mode derivative (checkpointed every ckpt steps).
torch is needed only here; the engine itself does not import it.
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import Glasma


class SlabEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho: torch.Tensor, glasma: Glasma, mv, nsteps: int, ckpt: int, zc: int, half: int):
        r = rho.detach().to("cpu", torch.float64).numpy()
        e = glasma.slab_energy_from_charge(mv, r, nsteps, zc, half)
        ctx.save_for_backward(rho)
        ctx.args = (glasma, mv, nsteps, ckpt, zc, half)
        return torch.tensor(e, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (rho,) = ctx.saved_tensors
        glasma, mv, nsteps, ckpt, zc, half = ctx.args
        _, g = glasma.slab_energy_gradient(mv, rho.detach().to("cpu", torch.float64).numpy(), nsteps, ckpt, zc, half)
        grad = torch.from_numpy(np.ascontiguousarray(g)).to(rho.device) * grad_out
        return grad, None, None, None, None, None, None


def slab_energy(rho: torch.Tensor, glasma: Glasma, mv, nsteps: int, ckpt: int, zc: int, half: int) -> torch.Tensor:
    return SlabEnergy.apply(rho, glasma, mv, nsteps, ckpt, zc, half)
