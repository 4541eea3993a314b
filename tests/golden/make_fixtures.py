"""Extract the reference's committed golden vectors into small local fixtures.

Reads (as data only; nothing is imported or run from the reference):
  * tests/reference_data/posthf/{h2,lih,water,f_heh}.{json,npz} -- PySCF 2.14 /
    libcint DF tensors, DF Fock matrices and DF-RHF energies;
  * benchmarks/results/issue409-packed-values/references/
    {water,water-tetramer,water-hexadecamer-2s4}-def2-svp-spherical-reference.json
    -- GPU4PySCF 1.8.1 cc-pVDZ / cc-pVDZ-JKFIT DF-RHF energies, forces and
    geometries (the file names say def2-svp; the content is cc-pVDZ).
  * benchmarks/results/issue206-practical-auxiliary/identity/{cc-pvdz,cc-pvdz-jkfit}.json
    -- the reference's checksummed canonical basis records (copied unchanged);
  * tests/reference_data/validation/hf-plus-uhf.json -- PySCF 2.6.2
    exact-integral UHF of HF+ (doublet, STO-3G Cartesian): energy, densities,
    one-electron matrices and orbital energies;
  * tests/data/external_basis/{manifest.json,oracles.npz} -- the water STO-3G
    overlap, kinetic and core-Hamiltonian matrices (Cartesian and spherical,
    two geometries).
The libcint scalar known answers of tests/native/test_density_fitting.cpp:468-499
are copied into libcint_known_answers.json by hand (they are literals in that
test, not a data file).

Usage: python tests/golden/make_fixtures.py /root/reference
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

POSTHF_KEYS = ("raw_three_center", "metric", "df_S", "df_h", "df_F", "df_C", "df_occ", "df_eps")


def main(reference: str) -> None:
    root = Path(reference)
    posthf = root / "tests/reference_data/posthf"
    for name in ("h2", "lih", "water", "f_heh"):
        meta = json.loads((posthf / f"{name}.json").read_text())
        arrays = np.load(posthf / f"{name}.npz", allow_pickle=False)
        np.savez_compressed(HERE / f"posthf_{name}.npz", **{k: arrays[k] for k in POSTHF_KEYS})
        record = {
            "inputs": meta["inputs"],
            "auxiliary_shells": meta["auxiliary_shells"],
            "df_hf_energy": meta["records"]["df"]["hf_energy"],
            "settings": meta["settings"],
            "versions": {"pyscf": meta["versions"]["pyscf"]},
        }
        (HERE / f"posthf_{name}.json").write_text(json.dumps(record, indent=1) + "\n")

    refs = root / "benchmarks/results/issue409-packed-values/references"
    for stem, out in (
        ("water-def2-svp-spherical-reference", "ccpvdz_water"),
        ("water-tetramer-def2-svp-spherical-reference", "ccpvdz_water4"),
        ("water-hexadecamer-2s4-def2-svp-spherical-reference", "ccpvdz_water16"),
    ):
        meta = json.loads((refs / f"{stem}.json").read_text())
        record = {
            key: meta[key]
            for key in (
                "orbital_basis",
                "auxiliary_basis",
                "basis_representation",
                "atoms_bohr",
                "aos",
                "auxiliary_count",
                "energy_hartree",
                "forces_hartree_per_bohr",
                "metric_condition",
                "stock_effective_rank",
                "versions",
            )
        }
        (HERE / f"{out}.json").write_text(json.dumps(record, indent=1) + "\n")

    # PySCF 2.6.2 exact-integral UHF of HF+ (doublet, STO-3G Cartesian).
    meta = json.loads((root / "tests/reference_data/validation/hf-plus-uhf.json").read_text())
    inp, data = meta["inputs"], meta["data"]
    record = {
        "inputs": {k: inp[k] for k in ("atomic_numbers", "coordinates", "shells", "basis_representation", "charge", "multiplicity")},
        "data": {k: data[k] for k in ("energy", "density", "hcore", "overlap", "nuclear_repulsion", "mo_energy", "mo_occ")},
        "versions": {"pyscf": meta["provenance"]["pyscf"]},
    }
    (HERE / "uhf_hf_plus.json").write_text(json.dumps(record, indent=1) + "\n")

    # External-basis oracle: water STO-3G one-electron matrices (BSE basis),
    # Cartesian and spherical, at the case geometry and a changed geometry.
    ext = root / "tests/data/external_basis"
    manifest = json.loads((ext / "manifest.json").read_text())
    arrays = np.load(ext / "oracles.npz", allow_pickle=False)
    record = {}
    for case in manifest["cases"]:
        name = case["name"]
        if not name.startswith("water_"):
            continue
        entry = {"representation": case["representation"], "atoms": case["atoms"], "changed_atoms": case["changed_atoms"]}
        for tag in ("", "changed_"):
            for key in ("overlap", "kinetic", "hcore"):
                entry[tag + key] = arrays[f"{name}_{tag}{key}"].tolist()
        record[name] = entry
    (HERE / "external_water_sto3g.json").write_text(json.dumps(record, indent=1) + "\n")

    identity = root / "benchmarks/results/issue206-practical-auxiliary/identity"
    for name in ("cc-pvdz", "cc-pvdz-jkfit"):
        (HERE / f"canonical_{name}.json").write_bytes((identity / f"{name}.json").read_bytes())


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
