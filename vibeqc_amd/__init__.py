"""MI355X-native density-fitted J/K Fock builds and RHF / UHF SCF.

Python host layer over the HIP engine ``libvibeqc_amd.so`` (C ABI in
``include/generativeqc``).  PyTorch is not required here; the optional
``vibeqc_amd.torch`` module provides the autograd binding.
"""

from .basis import (
    AOBasis,
    Atom,
    CanonicalBasis,
    Primitive,
    Shell,
    atoms_from,
    load_basis,
    load_canonical_basis,
    shells_for,
    write_canonical_basis,
)
from .calculator import Calculator, Result
from .checkpoint import CheckpointError, inspect_checkpoint
from .fock import FockBuildSpec, FockEvaluation, FockPlan, FockScfResult, FockTerm

__all__ = [
    "AOBasis",
    "Atom",
    "Calculator",
    "CanonicalBasis",
    "CheckpointError",
    "FockBuildSpec",
    "FockEvaluation",
    "FockPlan",
    "FockScfResult",
    "FockTerm",
    "Primitive",
    "Result",
    "Shell",
    "atoms_from",
    "inspect_checkpoint",
    "load_basis",
    "load_canonical_basis",
    "shells_for",
    "write_canonical_basis",
]
