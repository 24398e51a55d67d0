"""torch.autograd.Function over the device integrator and its discrete adjoint.

``integrate(y0, k, ...)`` integrates device-resident cells over one interval
with the adaptive Rosenbrock method and records every accepted step; its
backward pass runs the exact discrete adjoint of that step sequence
(mkpp_amd_device_adjoint) and returns dL/dy0 and dL/dk.  Inputs:

  y0  (num_species, num_cells) float64 CUDA, LayoutLeft like the C ABI
  k   (num_groups, num_cells) float64 CUDA, one rate constant per flux group
      and cell (``rate_constants`` gives the values the rate laws produce)

The gradient is that of the map with the step sizes frozen at the ones the
forward pass accepted (the step-size controller is not differentiated), which
is what central differences with the same step sequence measure.

Forward mode (torch.autograd.forward_ad dual tensors) runs the device
tangent-linear model along the same recorded steps
(mkpp_amd_device_tangent_linear) and returns dy(T) for a direction
(dy0, dk).  The torch.func transforms are not supported: the device calls
need tensors with storage, which their wrapped tensors lack.
"""

from __future__ import annotations

import torch

from . import native


def rate_constants(mechanism: str, conc, temp=None, air_density=None, rate_parameters=None):
    return native.device_rate_constants(mechanism, conc, temp, air_density, rate_parameters)


class RosenbrockIntegrate(torch.autograd.Function):
    """y(T) from (y0, k); the checkpoints, step sizes and step counts are
    returned beside it as non-differentiable outputs (they carry the recorded
    step sequence to the backward and forward-mode passes)."""

    @staticmethod
    def forward(y0, k, mechanism, temp, air_density, dt, method, atol, rtol, seed, max_steps):
        n, ncell = y0.shape
        ng, nv = native.diff_info(mechanism)
        if k.shape != (ng, ncell):
            raise ValueError(f"k must be ({ng}, {ncell})")
        y = y0.detach().contiguous().clone()
        kk = k.detach().contiguous()
        dev = y0.device
        ckpt = torch.empty((ncell, max_steps, nv), dtype=torch.float64, device=dev)
        hs = torch.empty((ncell, max_steps), dtype=torch.float64, device=dev)
        nsteps = torch.zeros(ncell, dtype=torch.int32, device=dev)
        status = torch.zeros(ncell, dtype=torch.int32, device=dev)
        at = torch.full((n,), float(atol), dtype=torch.float64, device=dev) if not torch.is_tensor(atol) else atol
        rt = torch.full((n,), float(rtol), dtype=torch.float64, device=dev) if not torch.is_tensor(rtol) else rtol
        native.device_integrate_checkpointed(mechanism, y, dt, method, temp, air_density, None, kk, at, rt, ckpt, hs,
                                             nsteps, seed=seed, status=status)
        bad = int((status != 0).sum())
        if bad:
            raise native.MkppError(-105, f"{bad} cells failed in the checkpointed forward pass "
                                         f"(status codes {sorted(set(status[status != 0].tolist()))})")
        return y, ckpt, hs, nsteps

    @staticmethod
    def setup_context(ctx, inputs, output):
        y0, k, mechanism, temp, air_density, dt, method, atol, rtol, seed, max_steps = inputs
        y, ckpt, hs, nsteps = output
        ctx.mechanism, ctx.method = mechanism, method
        ctx.mark_non_differentiable(ckpt, hs, nsteps)
        ctx.save_for_backward(y, k, temp, air_density, ckpt, hs, nsteps)
        ctx.save_for_forward(y0, k, temp, air_density, hs, nsteps)

    @staticmethod
    def backward(ctx, gy, gckpt, ghs, gnsteps):
        y, k, temp, air_density, ckpt, hs, nsteps = ctx.saved_tensors
        lam = gy.detach().contiguous().clone()
        kbar = native.device_adjoint(ctx.mechanism, y.detach(), ctx.method, temp, air_density, None,
                                     k.detach().contiguous(), ckpt, hs, nsteps, lam)
        return lam, kbar, None, None, None, None, None, None, None, None, None

    @staticmethod
    def jvp(ctx, dy0, dk, *unused):
        y0, k, temp, air_density, hs, nsteps = ctx.saved_tensors
        y = y0.detach().contiguous().clone()
        dy = torch.zeros_like(y) if dy0 is None else dy0.detach().contiguous().clone()
        dkk = None if dk is None else dk.detach().contiguous()
        native.device_tangent_linear(ctx.mechanism, y, ctx.method, temp, air_density, None, k.detach().contiguous(),
                                     hs, nsteps, dy, dkk)
        return dy, None, None, None


def integrate(y0, k, mechanism: str = "saprc99", temp=None, air_density=None, dt: float = 3600.0,
              method: str = "rodas3", atol=1.0e3, rtol=1.0e-3, seed: float = 1.0e-6, max_steps: int = 256):
    """Differentiable integration over one interval of length dt (reverse mode
    through the device adjoint, forward mode through the device tangent-linear
    model)."""
    y, _, _, _ = RosenbrockIntegrate.apply(y0, k, mechanism, temp, air_density, dt, method, atol, rtol, seed,
                                           max_steps)
    return y
