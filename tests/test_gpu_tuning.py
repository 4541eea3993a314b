"""FockPlan.autotune on the device: it records a launch shape that later
builds use, the shape only regroups sums (J bitwise unchanged, K at rounding
level), and results stay within the precision's tolerance of FP64.  The
registry is cleared around each test so the rest of the suite runs on the
default rules."""

from __future__ import annotations

import numpy as np
import pytest

from _cases import ccpvdz_case

pytestmark = pytest.mark.gpu


@pytest.fixture
def registry():
    from vibeqc_amd import tuning

    tuning.clear()
    yield tuning
    tuning.clear()


def _plan(case):
    from vibeqc_amd import FockBuildSpec, FockPlan

    return FockPlan(case.orbital(), FockBuildSpec.hf(), auxiliary=case.auxiliary())


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.timeout(300)
def test_autotune_records_a_shape_that_builds_use(registry):
    from vibeqc_amd.geometry import lowdin_occupied

    case, _ = ccpvdz_case("ccpvdz_water16")
    nocc = 80
    with _plan(case) as plan:
        S, _ = plan.one_electron()
        C = lowdin_occupied(S, nocc, seed=2)
        j64, k64 = plan.evaluate_occupied(C)
        for precision, tol_k in (("bf16x3", 5e-5), ("fp64", 1e-12)):
            plan.set_precision(precision)
            before = plan.launch_shape(nocc)
            jd, kd = plan.evaluate_occupied(C)
            res = plan.autotune(C, repeat=2)
            print(f"(H2O)16 {precision}: default {before} {res['default_ms']:.3f} ms, "
                  f"chosen u_tile {res['u_tile']} splits {res['syrk_splits']} {res['best_ms']:.3f} ms")
            assert 1 <= res["u_tile"] <= 10 and res["syrk_splits"] >= 1
            assert 0 < res["best_ms"] <= res["default_ms"]
            assert plan.launch_shape(nocc) == {"u_tile": res["u_tile"], "syrk_splits": res["syrk_splits"]}
            j, k = plan.evaluate_occupied(C)
            # The U tile only splits the occupied columns over blocks and the
            # split count only regroups K's k sum: J (from the same rho
            # partials) is bitwise unchanged, K moves at rounding level (FP32
            # slices for BF16X3, FP64 slices for FP64).
            assert np.array_equal(j, jd)
            assert _rel(k, kd) < (tol_k if precision == "bf16x3" else 1e-13)
            assert _rel(j, j64) < (1e-6 if precision == "bf16x3" else 1e-14) and _rel(k, k64) < tol_k
    assert registry.count() == 2


@pytest.mark.timeout(300)
def test_recorded_shape_reaches_new_plans_and_scf(registry):
    """A shape recorded for this device applies to every plan of that shape,
    and an SCF on it converges to the same energy as on the default shape."""
    from vibeqc_amd.geometry import lowdin_occupied

    case, meta = ccpvdz_case("ccpvdz_water4")
    with _plan(case) as plan:
        plan.set_precision("bf16x3")
        S, _ = plan.one_electron()
        nocc = 20
        default = plan.launch_shape(nocc)
        ref = plan.solve(energy_tolerance=1e-10, density_tolerance=1e-8)
        res = plan.autotune(lowdin_occupied(S, nocc, seed=4), repeat=2)
    with _plan(case) as fresh:
        fresh.set_precision("bf16x3")
        assert fresh.launch_shape(nocc) == {"u_tile": res["u_tile"], "syrk_splits": res["syrk_splits"]}
        tuned = fresh.solve(energy_tolerance=1e-10, density_tolerance=1e-8)
    registry.clear()
    with _plan(case) as cleared:
        cleared.set_precision("bf16x3")
        assert cleared.launch_shape(nocc) == default
    print(f"(H2O)4: default {default}, tuned {res}, E {ref.energy:.12f} / {tuned.energy:.12f}")
    assert abs(tuned.energy - ref.energy) < 1e-9
    assert abs(tuned.energy - meta["energy_hartree"]) < 1e-8
