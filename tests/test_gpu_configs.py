"""The benchmark workloads at full size on the HIP device.

  * (H2O)4, 6 A cube, seed 0 (plain FP64 DF-RHF): energy and iteration count
    against the oracle fixture; J/K at the converged density against the
    oracle on the same density (1e-10).
  * (H2O)16 (384 AO / 1856 aux) against the reference's GPU4PySCF fixture:
    FP64 energy within 1e-8 Eh and forces within 1e-7 Eh/bohr; mixed energy
    within 1 uEh and mixed forces within 1e-6 Eh/bohr of FP64.
  * C20H42 single build (490 AO / 2366 aux, 81 Lowdin-orthonormal occupied
    columns): FP32 and BF16X3 J/K against FP64 at full size (max-abs error
    relative to max |X|: J 1e-6 for both, K 1e-5 FP32 and 5e-5 BF16X3); FP64
    J/K against the oracle on the same construction at C4H10.
  * The 512-molecule ragged batch: no failures, mixed within 1 uEh of FP64 per
    molecule, and the first member of each size 1, 2, 4, 8, 12, 16 within
    1e-8 Eh of the oracle fixture in both modes.
  * (H2O)8 displaced (differentiation workload): energy within 1e-8 Eh and
    forces within 1e-7 Eh/bohr of the oracle fixture.
  * (H2O)64 (1536 AO / 7424 aux / 320 occupied): FP32 and BF16X3 J/K against
    FP64 at full size (J 1e-6, K 2e-4 FP32 and 5e-5 BF16X3), mixed SCF within 1 uEh of the FP64 SCF, mixed forces
    within 1e-6 Eh/bohr of FP64 forces, and both sum to zero (1e-8).
"""

from __future__ import annotations

import numpy as np
import pytest

from _cases import ccpvdz_case, molecule_case
from conftest import golden_json

pytestmark = pytest.mark.gpu

TIGHT = dict(energy_tolerance=1e-10, density_tolerance=1e-8)


def _plan(case, **kw):
    from vibeqc_amd import FockBuildSpec, FockPlan

    return FockPlan(case.orbital(), FockBuildSpec.hf(), auxiliary=case.auxiliary(), **kw)


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def test_water4_seed0_against_oracle(oracle):
    from vibeqc_amd.geometry import water_cluster

    ref = golden_json("config_oracle.json")["water4_seed0"]
    case = molecule_case("water4_seed0", water_cluster(4, 6.0, seed=0))
    with _plan(case) as plan:
        r = plan.solve(**TIGHT)
        D = np.asarray(r.density)
        ev = plan.evaluate(D)
    print(f"(H2O)4 seed 0: E {r.energy:.10f} (oracle {ref['energy']:.10f}), iterations {r.iterations} ({ref['iterations']})")
    assert abs(r.energy - ref["energy"]) < 1e-8
    assert abs(r.iterations - ref["iterations"]) <= 1
    J, K = oracle.jk(case.oracle_problem(oracle).B, D)
    assert np.abs(ev.coulomb - J).max() < 1e-10 * np.abs(J).max()
    assert np.abs(ev.exchange - K).max() < 1e-10 * np.abs(K).max()
    assert abs(np.trace(ev.coulomb) - ref["trace_J"]) < 1e-6 and abs(np.trace(ev.exchange) - ref["trace_K"]) < 1e-6


@pytest.mark.timeout(600)
def test_water16_against_reference_fixture():
    case, meta = ccpvdz_case("ccpvdz_water16")
    ref_f = np.array(meta["forces_hartree_per_bohr"])
    with _plan(case) as plan:
        shape = plan.shape()
        assert (shape["nbf"], shape["naux"]) == (meta["aos"], meta["auxiliary_count"])
        r64 = plan.solve(compute_forces=True, **TIGHT)
        plan.set_precision("bf16x3")
        rmx = plan.solve(compute_forces=True, **TIGHT)
    f64, fmx = np.asarray(r64.forces), np.asarray(rmx.forces)
    print(f"(H2O)16: E64 - Eref = {r64.energy - meta['energy_hartree']:.2e}, max |F64 - Fref| = {np.abs(f64 - ref_f).max():.2e}, "
          f"Emixed - E64 = {rmx.energy - r64.energy:.2e}, max |Fmixed - F64| = {np.abs(fmx - f64).max():.2e}")
    assert abs(r64.energy - meta["energy_hartree"]) < 1e-8
    assert np.abs(f64 - ref_f).max() < 1e-7
    assert abs(rmx.energy - r64.energy) < 1e-6
    assert np.abs(fmx - f64).max() < 1e-6


def _alkane_case(n_carbon):
    from vibeqc_amd.basis import cc_pvdz_aux_for_carbon
    from vibeqc_amd.geometry import alkane

    return molecule_case(f"c{n_carbon}", alkane(n_carbon), basis="def2-svp", aux=cc_pvdz_aux_for_carbon())


def test_alkane_fp64_jk_against_oracle(oracle):
    """The C20H42 construction (def2-SVP, scaled oxygen JKFIT for carbon) at
    C4H10, FP64 J/K of a Lowdin-orthonormal occupied factor against the oracle."""
    from vibeqc_amd.geometry import lowdin_occupied

    case = _alkane_case(4)
    P = case.oracle_problem(oracle)
    C = lowdin_occupied(P.S, 17, seed=1)
    J, K = oracle.jk(P.B, 2.0 * C @ C.T)
    with _plan(case) as plan:
        j, k = plan.evaluate_occupied(C)
    assert _rel(j, J) < 1e-10 and _rel(k, K) < 1e-10


@pytest.mark.timeout(300)
def test_c20h42_reduced_precision_jk_full_size():
    from vibeqc_amd.geometry import lowdin_occupied

    case = _alkane_case(20)
    with _plan(case) as plan:
        shape = plan.shape()
        assert (shape["nbf"], shape["naux"]) == (490, 2366)
        S, _ = plan.one_electron()
        C = lowdin_occupied(S, 81, seed=1)
        j64, k64 = plan.evaluate_occupied(C)
        errs = {}
        for precision, tol_k in (("fp32", 1e-5), ("bf16x3", 5e-5)):
            plan.set_precision(precision)
            j, k = plan.evaluate_occupied(C)
            errs[precision] = (_rel(j, j64), _rel(k, k64))
            print(f"C20H42 {precision}: max|dJ| {np.abs(j - j64).max():.2e} (rel {errs[precision][0]:.2e}), "
                  f"max|dK| {np.abs(k - k64).max():.2e} (rel {errs[precision][1]:.2e})")
            assert errs[precision][0] < 1e-6 and errs[precision][1] < tol_k


@pytest.mark.timeout(900)
def test_ragged_batch_512():
    from vibeqc_amd import Calculator
    from vibeqc_amd.geometry import ragged_water_batch

    systems, sizes = ragged_water_batch()
    assert len(systems) == 512
    energies = {}
    for precision in ("mixed", "fp64"):
        with Calculator(precision=precision).prepare_batch(systems, warm_start=False) as batch:
            res = batch.execute(properties=("energy",))
        assert not res.failure_indices, f"{precision}: failed items {res.failure_indices}"
        energies[precision] = np.asarray(res.energies)
    gap = np.abs(energies["mixed"] - energies["fp64"])
    print(f"ragged 512: max |Emixed - E64| = {gap.max():.2e} Eh")
    assert gap.max() < 1e-6
    for ref in golden_json("config_oracle.json")["ragged_batch"]:
        k = ref["index"]
        assert sizes[k] == ref["waters"]
        for precision in ("fp64", "mixed"):
            err = energies[precision][k] - ref["energy"]
            print(f"member {k} (n={ref['waters']}) {precision}: E - E_oracle = {err:.2e}")
            assert abs(err) < 1e-8


def test_water8_displaced_against_oracle():
    from vibeqc_amd.geometry import water_cluster

    ref = golden_json("config_oracle.json")["water8_displaced"]
    case = molecule_case("water8", water_cluster(8, 8.0, seed=3, displacement_angstrom=0.05))
    with _plan(case) as plan:
        r = plan.solve(compute_forces=True, **TIGHT)
    err = np.abs(np.asarray(r.forces) - np.array(ref["forces"])).max()
    print(f"(H2O)8 displaced: E - E_oracle = {r.energy - ref['energy']:.2e}, max |F - F_oracle| = {err:.2e}")
    assert abs(r.energy - ref["energy"]) < 1e-8
    assert err < 1e-7


@pytest.mark.timeout(900)
def test_water64_full_size():
    from vibeqc_amd.geometry import lowdin_occupied, water_cluster

    case = molecule_case("water64", water_cluster(64, 12.4, seed=4))
    with _plan(case) as plan:
        shape = plan.shape()
        assert (shape["nbf"], shape["naux"]) == (1536, 7424)
        S, _ = plan.one_electron()
        C = lowdin_occupied(S, 320, seed=1)
        j64, k64 = plan.evaluate_occupied(C)
        # K sums r ldq = 2.4e6 products per element in FP32 split-K slices of
        # about 2e5, so its error grows with the slice length (measured: FP32
        # 6e-5, BF16X3 1.5e-5 relative); J accumulates in FP64.
        for precision, tol_k in (("fp32", 2e-4), ("bf16x3", 5e-5)):
            plan.set_precision(precision)
            j, k = plan.evaluate_occupied(C)
            print(f"(H2O)64 {precision}: rel J {_rel(j, j64):.2e}, rel K {_rel(k, k64):.2e}")
            assert _rel(j, j64) < 1e-6 and _rel(k, k64) < tol_k
        plan.set_precision("fp64")
        r64 = plan.solve(compute_forces=True, **TIGHT)
        plan.set_precision("bf16x3")
        rmx = plan.solve(compute_forces=True, **TIGHT)
    f64, fmx = np.asarray(r64.forces), np.asarray(rmx.forces)
    print(f"(H2O)64: E64 {r64.energy:.10f} ({r64.iterations} it), Emixed - E64 = {rmx.energy - r64.energy:.2e} "
          f"({rmx.iterations} it), max |Fmixed - F64| = {np.abs(fmx - f64).max():.2e}, "
          f"|sum F64| {np.abs(f64.sum(0)).max():.1e}")
    assert abs(rmx.energy - r64.energy) < 1e-6
    assert np.abs(fmx - f64).max() < 1e-6
    assert np.abs(f64.sum(axis=0)).max() < 1e-8 and np.abs(fmx.sum(axis=0)).max() < 1e-8


@pytest.mark.timeout(300)
@pytest.mark.parametrize("waters,nocc", [(24, 120), (36, 180)])
def test_mid_size_occupied_tiles(waters, nocc):
    """FP32 and BF16X3 J/K against FP64 for occupied widths between the
    C20H42 and (H2O)64 shapes (i tiles of 4 to 6 MFMA column blocks)."""
    from vibeqc_amd.geometry import lowdin_occupied, water_cluster

    case = molecule_case(f"water{waters}", water_cluster(waters, 12.4 * (waters / 64) ** (1 / 3), seed=9))
    with _plan(case) as plan:
        S, _ = plan.one_electron()
        C = lowdin_occupied(S, nocc, seed=3)
        j64, k64 = plan.evaluate_occupied(C)
        for precision, tol_k in (("fp32", 2e-4), ("bf16x3", 5e-5)):
            plan.set_precision(precision)
            j, k = plan.evaluate_occupied(C)
            print(f"(H2O){waters} nocc {nocc} {precision}: rel J {_rel(j, j64):.2e}, rel K {_rel(k, k64):.2e}")
            assert _rel(j, j64) < 1e-6 and _rel(k, k64) < tol_k

