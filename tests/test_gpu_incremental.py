"""The cheaper stages of a mixed (BF16X3) SCF on the device.

A mixed solve starts with single-term BF16 builds while the density change is
above 1e-2, and its FP64 refinement takes K from the last full FP64 build plus
a BF16X3 update of the occupied factor's change (VIBEQC_INCREMENTAL_K=0
disables both).  Forced on the (H2O)16
GPU4PySCF fixture: the energy stays within 1e-8 Eh and the forces within 1e-7
Eh/bohr of the fixture, within 1e-10 Eh / 1e-7 Eh/bohr of the same solve
without it, and both kinds of builds occur.  FP64 solves never use them."""

from __future__ import annotations

import numpy as np
import pytest

from _cases import ccpvdz_case

pytestmark = pytest.mark.gpu

TIGHT = dict(energy_tolerance=1e-10, density_tolerance=1e-8)


def _plan(case):
    from vibeqc_amd import FockBuildSpec, FockPlan

    return FockPlan(case.orbital(), FockBuildSpec.hf(), auxiliary=case.auxiliary())


@pytest.mark.timeout(600)
def test_incremental_refinement_matches_fixture_and_full_builds(monkeypatch):
    case, meta = ccpvdz_case("ccpvdz_water16")
    ref_f = np.array(meta["forces_hartree_per_bohr"])
    runs = {}
    with _plan(case) as plan:
        for mode in ("0", "1"):
            monkeypatch.setenv("VIBEQC_INCREMENTAL_K", mode)
            plan.set_precision("bf16x3")
            r = plan.solve(compute_forces=True, **TIGHT)
            runs[mode] = (r, plan.last_solve_stats())
        monkeypatch.setenv("VIBEQC_INCREMENTAL_K", "1")
        plan.set_precision("fp64")
        r64 = plan.solve(**TIGHT)
        stats64 = plan.last_solve_stats()
    (r0, s0), (r1, s1) = runs["0"], runs["1"]
    f0, f1 = np.asarray(r0.forces), np.asarray(r1.forces)
    print(f"(H2O)16 mixed: full-build refinement E {r0.energy:.12f} ({s0}), incremental E {r1.energy:.12f} ({s1}); "
          f"E1 - Eref {r1.energy - meta['energy_hartree']:.2e}, max |F1 - Fref| {np.abs(f1 - ref_f).max():.2e}, "
          f"max |F1 - F0| {np.abs(f1 - f0).max():.2e}")
    assert s0["incremental_builds"] == 0 and s0["single_term_builds"] == 0
    assert s1["incremental_builds"] >= 1 and s1["single_term_builds"] >= 1
    assert s1["single_term_builds"] < s1["reduced_iterations"]  # three-term builds follow them
    # Incremental builds replace FP64-stage builds after an anchor (a full
    # FP64 build), plus at most the first finalisation rebuild.
    assert s1["incremental_builds"] <= s1["iterations"] - s1["reduced_iterations"]
    assert abs(r1.energy - meta["energy_hartree"]) < 1e-8
    assert np.abs(f1 - ref_f).max() < 1e-7
    # Different schedules converge to densities that agree to the density
    # tolerance (1e-8 RMS), which moves forces (first order in the density)
    # by a few 1e-8 Eh/bohr; energies (second order) agree far closer.
    assert abs(r1.energy - r0.energy) < 1e-10
    assert np.abs(f1 - f0).max() < 1e-7
    assert stats64["incremental_builds"] == 0 and stats64["reduced_iterations"] == 0 and stats64["single_term_builds"] == 0
    assert abs(r64.energy - meta["energy_hartree"]) < 1e-8


@pytest.mark.timeout(300)
def test_sketched_factor_at_small_size(monkeypatch):
    """The sketched projector factor (used for n >= 768 when the projection of
    the previous factor is rejected) forced at 384 AO: same energy as the
    default factor path and as the fixture."""
    case, meta = ccpvdz_case("ccpvdz_water16")
    out = {}
    with _plan(case) as plan:
        for mode in ("0", "1"):
            monkeypatch.setenv("VIBEQC_SKETCH_FACTOR", mode)
            r = plan.solve(**TIGHT)
            out[mode] = (r.energy, plan.last_solve_stats())
    print(f"(H2O)16 FP64: default factors {out['0']}, sketched {out['1']}")
    assert abs(out["1"][0] - out["0"][0]) < 1e-10
    assert abs(out["1"][0] - meta["energy_hartree"]) < 1e-8
    assert abs(out["1"][1]["iterations"] - out["0"][1]["iterations"]) <= 1
    assert out["1"][1]["purified_densities"] >= out["1"][1]["iterations"] - 1
