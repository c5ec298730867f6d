"""Differentiable energy boundary: the device energy as a `torch.autograd.Function`.

The reference obtains forces as `-jax.grad(energy_fn)` (deepmd_jax/md.py:292,
deepmd_jax/dpmodel.py:106-108).  Here the forward pass (`dpmd_compute`) already runs
the hand-written backward kernel, which recomputes the embedding activations instead
of storing per-neighbour tensors; the Function caches the resulting forces and virial,
and its backward returns `-F * grad_output` for the coordinates.  jax is not
available to register a `jax.custom_vjp`, so this is the one differentiable form.

    E = DeviceEnergy.apply(coord, system, 'fp32')   # coord: (N, 3) float64 CPU tensor
    E.backward()                                    # coord.grad = -F
"""
import numpy as np
import torch


class DeviceEnergy(torch.autograd.Function):
    """Energy of a single-segment DeviceSystem at `coord` (caller atom order)."""

    @staticmethod
    def forward(ctx, coord, system, precision='fp32'):
        x = coord.detach().to(torch.float64).cpu().numpy()
        system.set_positions(x.reshape((1,) + x.shape))
        if system.nlist_update():
            system.nlist_allocate(system.rc_skin_last, 1.2)
        e, f, v, _ = system.compute(precision)
        ctx.save_for_backward(torch.as_tensor(f[0], dtype=coord.dtype))
        ctx.virial = v[0]
        return torch.as_tensor(e[0], dtype=coord.dtype)

    @staticmethod
    def backward(ctx, grad_e):
        (force,) = ctx.saved_tensors
        return -force * grad_e, None, None


def energy_and_force(system, coord, precision='fp32'):
    """(E, F) through the autograd Function (F = -dE/dx)."""
    c = torch.as_tensor(np.asarray(coord, np.float64)).requires_grad_(True)
    e = DeviceEnergy.apply(c, system, precision)
    (g,) = torch.autograd.grad(e, [c])
    return float(e), -g.numpy()
