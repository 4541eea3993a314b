"""PyTorch binding (host mirror of python/generativeqc/torch.py:47-245).

``energy(coordinates, atomic_numbers, calculator)`` returns the DF-RHF energy
as a 0-d tensor in the input dtype/device.  ``batched_energy`` evaluates a list
of molecules.  The backward needs analytic nuclear forces, which this engine
does not provide yet: calling ``backward`` raises NotImplementedError rather
than returning a silent or approximate gradient.  Higher-order derivatives are
refused as in the reference (torch.py:34-44).
"""

from __future__ import annotations

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
        result = calculator.singlepoint(atoms, charge=charge, multiplicity=multiplicity, properties=("energy",))
        return coordinates.new_tensor(result.energy)

    @staticmethod
    def backward(ctx, grad_output):
        _require_first_order_backward()
        raise NotImplementedError("analytic DF-RHF forces are not yet provided by the HIP engine")


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


def batched_energy(coordinates, atomic_numbers, calculator: Calculator | None = None, charges=None, multiplicities=None):
    """Energies of several molecules (one tensor per molecule), stacked."""
    calculator = calculator or Calculator()
    charges = charges or [0] * len(coordinates)
    multiplicities = multiplicities or [1] * len(coordinates)
    return torch.stack(
        [energy(c, z, calculator, charge=q, multiplicity=m) for c, z, q, m in zip(coordinates, atomic_numbers, charges, multiplicities)]
    )
