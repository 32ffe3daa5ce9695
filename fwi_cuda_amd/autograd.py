"""torch.autograd wrapper of the elastic P/SV modelling operator.

    uz, ux = seismograms(sim, lam, mu, rho)      # differentiable in lam, mu, rho
    loss = l2_misfit(uz, ux, obs_z, obs_x, dt)   # sum 0.5 dt r^2, the reference's L2
    loss.backward()

The forward pass is fwi_forward (kernel_PSV without gradient,
src/cpu/n_kernel_PSV.cpp:20-213) over the library's batched shots.  The
backward pass is one of two:

  adjoint="exact" (default)  fwi_misfit_gradient_exact: the discrete adjoint of
      the time step (reverse mode through every stencil, CPML memory update,
      free-surface image and source/receiver operation, the forward states of
      every step kept on the GPU), so the gradient is the derivative of the
      FP32-modelled seismograms to FP32 rounding;
  adjoint="correlation"      fwi_misfit_gradient: the reference's adjoint
      propagation of dL/d(seismogram) / dt injected at the receivers in
      reversed time (src/cpu/n_kernel_lib_PSV.cpp:511-522), the fwi_grad2
      correlation with the density term taken as adjoint velocity times forward
      velocity; it keeps only the checkpoints, but pairs forward and adjoint
      fields at the same time level (an O(dt) departure from the derivative).

Both end with the chain rule through mat_av2's staggered averages.

CUDA tensors stay on the GPU: the model, the seismograms, the upstream gradient
and the model gradient cross the C ABI as device pointers (fwi_*_device), which
the library reads and writes with its own kernels.  This needs torch loaded
before the library, so both use torch's HIP runtime (the library then binds to
the runtime already in the process).  CPU tensors cross as host FP64 arrays.
Without the HIP library the import works but every call raises (there is no CPU
fallback).
"""
from __future__ import annotations

import numpy as np
import torch

from . import psv


def _host(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


def _dev(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(torch.float64).contiguous()


class Seismograms(torch.autograd.Function):
    """(lam, mu, rho) [nz][nx] -> (uz, ux) [rows][nt] for shots [shot0, shot0 + nshots)."""

    @staticmethod
    def forward(ctx, lam, mu, rho, sim: psv.Simulation, shot0: int, nshots: int, exact: bool = True):
        ctx.sim, ctx.shots, ctx.like, ctx.exact = sim, (shot0, nshots), lam, exact
        rows = sim._rows(shot0, nshots)
        if lam.is_cuda:
            model = [_dev(t) for t in (lam, mu, rho)]
            uz = torch.empty((rows, sim.nt), dtype=torch.float64, device=lam.device)
            ux = torch.empty_like(uz)
            torch.cuda.synchronize(lam.device)   # the inputs' producers are done
            sim.set_model_device(*[t.data_ptr() for t in model])
            sim.forward_device(uz.data_ptr(), ux.data_ptr(), shot0, nshots)
            ctx.model = model
            return uz.to(lam.dtype), ux.to(lam.dtype)
        model = [_host(t) for t in (lam, mu, rho)]
        sim.set_model(*model)
        uz, ux = sim.forward(shot0, nshots)
        ctx.model = model
        return (torch.from_numpy(uz).to(device=lam.device, dtype=lam.dtype),
                torch.from_numpy(ux).to(device=lam.device, dtype=lam.dtype))

    @staticmethod
    def backward(ctx, g_uz, g_ux):
        sim, like = ctx.sim, ctx.like
        dt = float(sim.case["dt"])
        rows = sim._rows(*ctx.shots)
        shape = (rows, sim.nt)
        if like.is_cuda:
            z = torch.zeros(shape, dtype=torch.float64, device=like.device)
            az = _dev(g_uz) / dt if g_uz is not None else z
            ax = _dev(g_ux) / dt if g_ux is not None else z
            az, ax = az.contiguous(), ax.contiguous()
            d = [torch.empty(like.shape, dtype=torch.float64, device=like.device) for _ in range(3)]
            torch.cuda.synchronize(like.device)
            sim.set_model_device(*[t.data_ptr() for t in ctx.model])   # the forward's model, whatever ran since
            back = sim.misfit_gradient_exact_device if ctx.exact else sim.misfit_gradient_device
            back(az.data_ptr(), ax.data_ptr(), *[t.data_ptr() for t in d], *ctx.shots)
            grads = [t.to(like.dtype) for t in d]
        else:
            sim.set_model(*ctx.model)   # the model of the forward pass, whatever ran since
            az = _host(g_uz) / dt if g_uz is not None else np.zeros(shape)
            ax = _host(g_ux) / dt if g_ux is not None else np.zeros(shape)
            d = (sim.misfit_gradient_exact if ctx.exact else sim.misfit_gradient)(az, ax, *ctx.shots)
            grads = [torch.from_numpy(a).to(device=like.device, dtype=like.dtype) for a in d]
        return grads[0], grads[1], grads[2], None, None, None, None


def seismograms(sim: psv.Simulation, lam, mu, rho, shot0: int = 0, nshots: int | None = None,
                adjoint: str = "exact"):
    """Differentiable (uz, ux); adjoint is "exact" (the discrete adjoint) or "correlation"."""
    if adjoint not in ("exact", "correlation"):
        raise ValueError(f"adjoint must be 'exact' or 'correlation', not {adjoint!r}")
    nshots = sim.nshot - shot0 if nshots is None else nshots
    return Seismograms.apply(lam, mu, rho, sim, shot0, nshots, adjoint == "exact")


def l2_misfit(uz, ux, obs_z, obs_x, dt: float):
    """The reference's L2 (adjsrc2, src/cpu/n_kernel_lib_PSV.cpp:614): sum 0.5 dt (syn - obs)^2."""
    return 0.5 * dt * (((uz - obs_z) ** 2).sum() + ((ux - obs_x) ** 2).sum())
