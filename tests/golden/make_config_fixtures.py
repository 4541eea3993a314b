"""Oracle results at the seeded benchmark geometries, for the GPU tests.

The reference ships no fixtures at these geometries, so the FP64 CPU oracle
(oracle/, pinned to the reference's own golden vectors by tests/test_oracle.py)
provides them:
  * (H2O)4, 6 A cube, seed 0 (the plain CPU workload): DF-RHF energy,
    iterations, and J/K checksums at the converged density;
  * members of the 512-molecule ragged batch (ragged_water_batch(), first
    member of each size 1, 2, 4, 8, 12, 16): DF-RHF energies;
  * (H2O)8, 8 A cube, seed 3, N(0, 0.05 A) displacements (the
    differentiation workload): DF-RHF energy and forces.
All cc-pVDZ / cc-pVDZ-JKFIT (spherical), converged to 1e-10 Eh / 1e-8 RMS.
Takes about ten minutes on 8 cores.

Usage: python tests/golden/make_config_fixtures.py
"""

from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import oracle as O  # noqa: E402
from vibeqc_amd import shells_for  # noqa: E402
from vibeqc_amd.geometry import ragged_water_batch, water_cluster  # noqa: E402

SIZES = (1, 2, 4, 8, 12, 16)


def _shells(shells):
    return [O.Shell(s.atom_index, s.l, tuple(p.exponent for p in s.primitives), tuple(p.coefficient for p in s.primitives))
            for s in shells]


def _solve(atoms, forces=False):
    z = [a.atomic_number for a in atoms]
    xyz = np.array([a.position for a in atoms])
    orb, aux = _shells(shells_for("cc-pvdz", atoms)), _shells(shells_for("cc-pvdz-jkfit", atoms))
    problem = O.build_problem(z, xyz, orb, True, aux, True)
    scf = O.df_rhf(problem, energy_tolerance=1e-10, density_tolerance=1e-8)
    out = {"energy": float(scf.energy), "iterations": int(scf.iterations), "nbf": int(problem.S.shape[0])}
    if forces:
        grad = O.df_rhf_gradient(z, xyz, orb, True, aux, True, problem, scf)
        out["forces"] = (-np.asarray(grad)).tolist()
    return out, problem, scf


def main() -> None:
    t0 = time.perf_counter()
    doc = {"generator": "tests/golden/make_config_fixtures.py", "convergence": [1e-10, 1e-8]}
    atoms = water_cluster(4, 6.0, seed=0)
    res, problem, scf = _solve(atoms)
    J, K = O.jk(problem.B, scf.density)
    res.update(trace_J=float(np.trace(J)), trace_K=float(np.trace(K)), norm_J=float(np.linalg.norm(J)),
               norm_K=float(np.linalg.norm(K)), sum_DJ=float(np.sum(scf.density * J)), sum_DK=float(np.sum(scf.density * K)))
    doc["water4_seed0"] = res
    print("water4", res["energy"], res["iterations"], f"{time.perf_counter() - t0:.0f}s", flush=True)
    systems, sizes = ragged_water_batch()
    doc["ragged_batch"] = []
    for n in SIZES:
        k = int(np.argmax(sizes == n))
        res, _, _ = _solve(systems[k])
        res.update(index=k, waters=n)
        doc["ragged_batch"].append(res)
        print("ragged", k, n, res["energy"], f"{time.perf_counter() - t0:.0f}s", flush=True)
    atoms = water_cluster(8, 8.0, seed=3, displacement_angstrom=0.05)
    res, _, _ = _solve(atoms, forces=True)
    doc["water8_displaced"] = res
    print("water8", res["energy"], f"{time.perf_counter() - t0:.0f}s", flush=True)
    (HERE / "config_oracle.json").write_text(json.dumps(doc, indent=1) + "\n")


if __name__ == "__main__":
    main()
