"""Shared test inputs built from the committed golden fixtures."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from conftest import golden_json, golden_npz
from vibeqc_amd.basis import NUMBERS, AOBasis, Atom, Primitive, Shell, shells_for


@dataclass
class Case:
    name: str
    atoms: tuple
    shells: tuple
    aux_shells: tuple
    representation: str
    charge: int
    multiplicity: int

    @property
    def z(self):
        return [a.atomic_number for a in self.atoms]

    @property
    def xyz(self):
        return np.array([a.position for a in self.atoms])

    @property
    def spherical(self):
        return self.representation == "real_spherical"

    def orbital(self) -> AOBasis:
        return AOBasis(self.atoms, self.shells, self.representation, self.charge, self.multiplicity)

    def auxiliary(self) -> AOBasis:
        return AOBasis(self.atoms, self.aux_shells, self.representation, self.charge, self.multiplicity)

    def oracle_shells(self, O, aux=False):
        src = self.aux_shells if aux else self.shells
        return [
            O.Shell(s.atom_index, s.l, tuple(p.exponent for p in s.primitives), tuple(p.coefficient for p in s.primitives))
            for s in src
        ]

    def oracle_problem(self, O):
        return O.build_problem(
            self.z, self.xyz, self.oracle_shells(O), self.spherical, self.oracle_shells(O, True), self.spherical, charge=self.charge
        )


def _shells(records):
    return tuple(
        Shell(r["atom_index"], r["angular_momentum"], tuple(Primitive(float(e), float(c)) for e, c in r["primitives"]))
        for r in records
    )


POSTHF = ("h2", "lih", "water", "f_heh")


def posthf_case(name: str):
    meta = golden_json(f"posthf_{name}.json")
    inp = meta["inputs"]
    atoms = tuple(Atom(int(z), tuple(map(float, p))) for z, p in zip(inp["atomic_numbers"], inp["coordinates"]))
    rep = "cartesian" if inp["basis_representation"] == "cartesian" else "real_spherical"
    case = Case(name, atoms, _shells(inp["shells"]), _shells(meta["auxiliary_shells"]), rep, inp["charge"], inp["multiplicity"])
    return case, meta, golden_npz(f"posthf_{name}.npz")


CCPVDZ = ("ccpvdz_water", "ccpvdz_water4", "ccpvdz_water16")


def ccpvdz_case(name: str):
    meta = golden_json(f"{name}.json")
    atoms = tuple(Atom(NUMBERS[sym], tuple(map(float, xyz))) for sym, xyz in meta["atoms_bohr"])
    case = Case(name, atoms, shells_for("cc-pvdz", atoms), shells_for("cc-pvdz-jkfit", atoms), "real_spherical", 0, 1)
    return case, meta


def molecule_case(name: str, atoms, basis="cc-pvdz", aux="cc-pvdz-jkfit"):
    return Case(name, tuple(atoms), shells_for(basis, atoms), shells_for(aux, atoms), "real_spherical", 0, 1)
