"""torch.autograd wrapper of the elastic P/SV modelling operator.

    uz, ux = seismograms(sim, lam, mu, rho)      # differentiable in lam, mu, rho
    loss = l2_misfit(uz, ux, obs_z, obs_x, dt)   # sum 0.5 dt r^2, the reference's L2
    loss.backward()

The forward pass is fwi_forward (kernel_PSV without gradient,
src/cpu/n_kernel_PSV.cpp:20-213) over the library's batched shots; the
backward pass is fwi_misfit_gradient: the adjoint propagation of the
upstream gradient dL/d(seismogram) / dt injected at the receivers in reversed
time (the residual injection of src/cpu/n_kernel_lib_PSV.cpp:511-522), the
fwi_grad2 correlation with the density term taken as adjoint velocity times
forward velocity, and the chain rule through mat_av2's staggered averages.

Tensors live wherever the caller keeps them; they cross the C ABI as host
FP64 arrays (the library owns its device copies).  Without the HIP library
the import works but every call raises (there is no CPU fallback).
"""
from __future__ import annotations

import numpy as np
import torch

from . import psv


def _host(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class Seismograms(torch.autograd.Function):
    """(lam, mu, rho) [nz][nx] -> (uz, ux) [rows][nt] for shots [shot0, shot0 + nshots)."""

    @staticmethod
    def forward(ctx, lam, mu, rho, sim: psv.Simulation, shot0: int, nshots: int):
        model = [_host(t) for t in (lam, mu, rho)]
        sim.set_model(*model)
        uz, ux = sim.forward(shot0, nshots)
        ctx.sim, ctx.model, ctx.shots = sim, model, (shot0, nshots)
        ctx.like = lam
        return (torch.from_numpy(uz).to(device=lam.device, dtype=lam.dtype),
                torch.from_numpy(ux).to(device=lam.device, dtype=lam.dtype))

    @staticmethod
    def backward(ctx, g_uz, g_ux):
        sim = ctx.sim
        sim.set_model(*ctx.model)   # the model of the forward pass, whatever ran since
        dt = float(sim.case["dt"])
        rows = sim._rows(*ctx.shots)
        shape = (rows, sim.nt)
        az = _host(g_uz) / dt if g_uz is not None else np.zeros(shape)
        ax = _host(g_ux) / dt if g_ux is not None else np.zeros(shape)
        d = sim.misfit_gradient(az, ax, *ctx.shots)
        like = ctx.like
        grads = [torch.from_numpy(a).to(device=like.device, dtype=like.dtype) for a in d]
        return grads[0], grads[1], grads[2], None, None, None


def seismograms(sim: psv.Simulation, lam, mu, rho, shot0: int = 0, nshots: int | None = None):
    nshots = sim.nshot - shot0 if nshots is None else nshots
    return Seismograms.apply(lam, mu, rho, sim, shot0, nshots)


def l2_misfit(uz, ux, obs_z, obs_x, dt: float):
    """The reference's L2 (adjsrc2, src/cpu/n_kernel_lib_PSV.cpp:614): sum 0.5 dt (syn - obs)^2."""
    return 0.5 * dt * (((uz - obs_z) ** 2).sum() + ((ux - obs_x) ** 2).sum())
