"""PyTorch binding (host mirror of python/generativeqc/torch.py:47-245).

``energy(coordinates, atomic_numbers, calculator)`` returns the DF-RHF energy
as a 0-d tensor in the input dtype/device; its backward is -forces * grad_out
with the analytic forces of the converged SCF (torch.py:47-84).
``batched_energy`` evaluates a list of molecules.  ``density_energy(plan, D)``
is the fixed-density energy 1/2 tr D (H + F(D)) + E_nuc of a prepared plan with
the backward dE/dD = F (the check of test_fock.py:56-60).  Higher-order
derivatives are refused as in the reference (torch.py:34-44).
"""

from __future__ import annotations

import numpy as np
import torch

from .calculator import Calculator


def _require_first_order_backward() -> None:
    if torch.is_grad_enabled():
        raise RuntimeError("energy supports only first-order coordinate derivatives")


class _EnergyFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coordinates, atomic_numbers, calculator, charge, multiplicity):
        values = coordinates.detach().cpu().double()
        atoms = [(int(atomic_numbers[i]), values[i].tolist()) for i in range(coordinates.shape[0])]
        want = ("energy", "forces") if coordinates.requires_grad else ("energy",)
        result = calculator.singlepoint(atoms, charge=charge, multiplicity=multiplicity, properties=want)
        if result.forces is not None:
            ctx.save_for_backward(torch.tensor(np.array(result.forces), dtype=coordinates.dtype, device=coordinates.device))
        return coordinates.new_tensor(result.energy)

    @staticmethod
    def backward(ctx, grad_output):
        _require_first_order_backward()
        (forces,) = ctx.saved_tensors
        # Forces are -dE/dR; autograd wants dE/dR.
        return -forces * grad_output, None, None, None, None


class _DensityEnergyFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, plan):
        d = density.detach().cpu().double().numpy()
        ev = plan.evaluate(0.5 * (d + d.T))
        ctx.save_for_backward(torch.as_tensor(ev.fock, dtype=density.dtype, device=density.device))
        return density.new_tensor(ev.energy)

    @staticmethod
    def backward(ctx, grad_output):
        _require_first_order_backward()
        (fock,) = ctx.saved_tensors
        return fock * grad_output, None


def density_energy(plan, density):
    """E(D) = 1/2 tr D (H + F(D)) + E_nuc at a fixed symmetric density, with
    dE/dD = F(D) as its backward."""
    if not isinstance(density, torch.Tensor) or not density.is_floating_point():
        raise TypeError("density must be a real floating-point tensor")
    if density.shape != (plan.nao, plan.nao):
        raise ValueError("density must be nbf x nbf")
    return _DensityEnergyFunction.apply(density, plan)


def _validated(coordinates, atomic_numbers):
    if not isinstance(coordinates, torch.Tensor) or not coordinates.is_floating_point():
        raise TypeError("coordinates must be a real floating-point tensor")
    if coordinates.ndim != 2 or coordinates.shape[1] != 3:
        raise ValueError("coordinates must have shape (natoms, 3)")
    z = [int(v) for v in atomic_numbers]
    if len(z) != coordinates.shape[0]:
        raise ValueError("atomic numbers and coordinates differ in length")
    return z


def energy(coordinates, atomic_numbers, calculator: Calculator | None = None, *, charge: int = 0, multiplicity: int = 1):
    """DF-RHF energy (Hartree) of one geometry (Bohr)."""
    z = _validated(coordinates, atomic_numbers)
    calculator = calculator or Calculator()
    return _EnergyFunction.apply(coordinates, z, calculator, int(charge), int(multiplicity))


def batched_energy(coordinates, atomic_numbers, calculator: Calculator | None = None, charges=None, multiplicities=None,
                   prepared_batch=None):
    """Energies of several molecules (one (natoms, 3) tensor each), stacked;
    the molecules run as one ragged batch (torch.py:187-245).  Backward uses
    the batch's analytic forces."""
    calculator = calculator or Calculator()
    coords = list(coordinates)
    zs = [_validated(c, z) for c, z in zip(coords, atomic_numbers)]
    if len(zs) != len(coords):
        raise ValueError("one atomic-number list per coordinate tensor is required")
    return _BatchedEnergyFunction.apply(calculator, zs, charges, multiplicities, prepared_batch, *coords)


class _BatchedEnergyFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, calculator, zs, charges, multiplicities, prepared_batch, *coords):
        systems = [
            [(z[i], c.detach().cpu().double()[i].tolist()) for i in range(c.shape[0])] for c, z in zip(coords, zs)
        ]
        want = ("energy", "forces") if any(c.requires_grad for c in coords) else ("energy",)
        if prepared_batch is None:
            result = calculator.batch_singlepoint(
                systems, charges=charges, multiplicities=multiplicities, strict=True, properties=want
            )
        else:
            xyz = [np.asarray([p for _, p in s]) for s in systems]
            result = prepared_batch.execute(xyz, strict=True, properties=want)
        forces = [
            torch.tensor(np.array(it.forces), dtype=c.dtype, device=c.device) if it.forces is not None else None
            for it, c in zip(result.items, coords)
        ]
        ctx.forces = forces
        ref = coords[0]
        return torch.as_tensor(result.energies, dtype=ref.dtype, device=ref.device)

    @staticmethod
    def backward(ctx, grad_output):
        _require_first_order_backward()
        grads = tuple(-f * grad_output[i] if f is not None else None for i, f in enumerate(ctx.forces))
        return (None, None, None, None, None, *grads)
