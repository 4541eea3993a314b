"""MI355X-native density-fitted J/K Fock builds and RHF SCF.

Python host layer over the HIP engine ``libvibeqc_amd.so`` (C ABI in
``include/generativeqc``).  PyTorch is not required here; the optional
``vibeqc_amd.torch`` module provides the autograd binding.
"""

from .basis import AOBasis, Atom, Primitive, Shell, atoms_from, load_basis, shells_for
from .calculator import Calculator, Result
from .fock import FockBuildSpec, FockEvaluation, FockPlan, FockScfResult, FockTerm

__all__ = [
    "AOBasis",
    "Atom",
    "Calculator",
    "FockBuildSpec",
    "FockEvaluation",
    "FockPlan",
    "FockScfResult",
    "FockTerm",
    "Primitive",
    "Result",
    "Shell",
    "atoms_from",
    "load_basis",
    "shells_for",
]
