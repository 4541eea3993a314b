"""DF-JK Fock-build benchmark on one MI355X.

One "step" is one complete J + K build at fixed density with every input
resident in HBM (B, the occupied factor, the output buffers).

Headline workload (`value`): (H2O)64 at liquid density (12.4 A cube), cc-pVDZ
(1536 AO) / cc-pVDZ-JKFIT (7424 auxiliary functions), 320 occupied columns
drawn N(0,1) from a seed and Lowdin-orthonormalised in the AO overlap -- the
large single system the mixed/FP64 target is stated on.  FP64, FP32 and BF16X3
(split bfloat16, FP32 accumulation) are timed on the same plan; `value` is the
BF16X3 ("mixed") rate, with FP64 beside it, the J/K max-abs deviation of every
precision from FP64, the roofline of the dominant kernel, and the (H2O)64 SCF
to 1e-10 Eh / 1e-8 RMS in FP64 and in mixed precision (BF16X3 iterations,
then FP64 refinement).

Also reported:
  * c20h42: the single-build C20H42 shape (490 AO, 81 occupied).  The offline
    basis data has no carbon cc-pVDZ / cc-pVDZ-JKFIT, so the orbital basis is
    def2-SVP (the same 490-AO shape) and carbon's fitting set is the oxygen
    JKFIT shells scaled by (6/8)^2 (2366 auxiliary functions);
  * ragged_batch: 512 (H2O)n clusters, n ~ U{1..16}, molecules/s mixed and FP64;
  * scf_water4: (H2O)4 cc-pVDZ SCF seconds per molecule in FP64 and in mixed
    precision with the energy gap;
  * uhf: 4 OH radicals (doublets, cc-pVDZ / cc-pVDZ-JKFIT, 19 AO), UHF
    energy + forces one plan at a time, systems/s, with the UHF oracle's
    systems/s on the host beside it; and the UHF SCF seconds of (H2O)8 ... OH
    (211 AO) on a prepared plan.

Usage: python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "DF-JK Fock builds/s & SCF s/molecule, mixed vs FP64, 1 GPU, ≤1 µEh error"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
PRECISIONS = ("fp64", "fp32", "bf16x3")


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _water64_plan():
    from vibeqc_amd import AOBasis, FockPlan, shells_for
    from vibeqc_amd.geometry import water_cluster

    atoms = water_cluster(64, 12.4, seed=4)
    orb = AOBasis(atoms, shells_for("cc-pvdz", atoms))
    aux = AOBasis(atoms, shells_for("cc-pvdz-jkfit", atoms))
    t0 = time.perf_counter()
    plan = FockPlan(orb, auxiliary=aux)
    return plan, time.perf_counter() - t0, 320


def _alkane_plan():
    from vibeqc_amd import AOBasis, FockPlan, shells_for
    from vibeqc_amd.basis import cc_pvdz_aux_for_carbon
    from vibeqc_amd.geometry import alkane

    atoms = alkane(20)
    orb = AOBasis(atoms, shells_for("def2-svp", atoms))
    aux = AOBasis(atoms, shells_for(cc_pvdz_aux_for_carbon(), atoms))
    t0 = time.perf_counter()
    plan = FockPlan(orb, auxiliary=aux)
    return plan, time.perf_counter() - t0, 81


def _time_precisions(plan, nocc, steps, warmup, precisions=PRECISIONS):
    from vibeqc_amd.geometry import lowdin_occupied

    S, _ = plan.one_electron()
    C = lowdin_occupied(S, nocc, seed=1)
    plan.set_precision("fp64")
    j64, k64 = plan.evaluate_occupied(C)
    out = {}
    for precision in precisions:
        plan.set_precision(precision)
        plan.benchmark_occupied(C, repeat=max(1, warmup))
        ms = plan.benchmark_occupied(C, repeat=steps)
        j, k = plan.evaluate_occupied(C)
        out[precision] = {
            "builds_per_s": 1e3 / ms["wall_ms"],
            "ms_per_build_wall": ms["wall_ms"],
            "phase_ms": ms,
            "max_abs_err_J": float(np.abs(j - j64).max()),
            "max_abs_err_K": float(np.abs(k - k64).max()),
            "rel_err_J": float(np.abs(j - j64).max() / np.abs(j64).max()),
            "rel_err_K": float(np.abs(k - k64).max() / np.abs(k64).max()),
        }
        _log(f"{precision}: {ms['wall_ms']:.3f} ms/build, phases {ms}")
    plan.set_precision("fp64")
    return out


# Device symbols of the kernels each timed phase launches (the phase times are
# HIP events on the plan's stream around exactly these launches).
PHASE_KERNELS = {
    "exchange_u_ms": ("occupied_u", {"fp64": "occupied_u_kernel<double", "fp32": "occupied_u_kernel<float",
                                     "bf16x3": "occupied_u_bf16x3"}),
    "exchange_syrk_ms": ("exchange_syrk", {"fp64": "exchange_syrk_kernel<double", "fp32": "exchange_syrk_kernel<float",
                                           "bf16x3": "exchange_syrk_bf16x3_kernel"}),
    "coulomb_ms": ("coulomb_matrix", {"fp64": "coulomb_matrix_kernel<double", "fp32": "coulomb_matrix_kernel<float",
                                      "bf16x3": "coulomb_matrix_kernel<float"}),
}


def _phase_work(n, naux, npair, nocc, precision):
    """Algorithmic flops and HBM bytes of one launch of each phase.

    U = B C: 2 n^2 A o flops; bytes = packed B once (P A s_B) + U written
    (n o A s_U).  K = U U^T as a lower-triangle SYRK: n (n + 1) A o flops;
    bytes = U read once + K written (n^2 8).  J = B rho: 2 P A flops; bytes =
    packed B once.  s_B = 8 (FP64) or 4 (FP32 copy); s_U = 8, 4 or 2 + 2.
    """
    s_b = 8 if precision == "fp64" else 4
    s_u = 8 if precision == "fp64" else 4
    u_bytes = float(n) * nocc * naux * s_u
    return {
        "exchange_u_ms": (2.0 * n * n * naux * nocc, float(npair) * naux * s_b + u_bytes),
        "exchange_syrk_ms": (float(n) * (n + 1) * naux * nocc, u_bytes + 8.0 * n * n),
        "coulomb_ms": (2.0 * npair * naux, float(npair) * naux * s_b),
    }


def _roofline(phase_ms, n, naux, npair, nocc, precision, traffic=None):
    """Roofline of the dominant (longest) kernel of one build.

    The bound is whichever resource needs more time for the kernel's
    algorithmic work: MFMA at the precision's dense peak (BF16X3 executes three
    bf16 MFMAs per logical product, so its effective peak is 2.5 PFLOP/s / 3)
    or HBM at 8 TB/s.  `achieved` = algorithmic work / HIP-event kernel time.
    """
    work = _phase_work(n, naux, npair, nocc, precision)
    key = max(PHASE_KERNELS, key=lambda k: phase_ms[k])
    label, symbols = PHASE_KERNELS[key]
    flops, nbytes = work[key]
    t = phase_ms[key] * 1e-3
    peak_f = PEAK_TFLOPS[precision] / (3.0 if precision == "bf16x3" else 1.0)
    t_f, t_b = flops / (peak_f * 1e12), nbytes / (HBM_PEAK_GBS * 1e9)
    if t_f >= t_b:
        bound, achieved, peak, unit = "mfma", flops / t / 1e12, peak_f, "TFLOP/s"
    else:
        bound, achieved, peak, unit = "hbm", nbytes / t / 1e9, HBM_PEAK_GBS, "GB/s"
    sym = symbols[precision]
    measured = (traffic or {}).get(sym)
    return {
        "kernel": measured.get("kernel_name", sym) if isinstance(measured, dict) else sym,
        "phase": label,
        "bound": bound,
        "achieved": achieved,
        "peak": peak,
        "unit": unit,
        "frac": achieved / peak,
        "traffic": None if measured is None else measured["bytes"],
        "traffic_detail": measured if measured is not None else (traffic or {}).get("error"),
        "algorithmic_flops": flops,
        "algorithmic_bytes": nbytes,
        "kernel_ms": phase_ms[key],
    }


def _pmc_traffic(workload, precision, copy_mib=256, timeout=240):
    """HBM bytes per launch of this workload's kernels from two rocprofv3 PMC
    runs (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) of a short child
    process, each scaled by what the same counter reads for a device-to-device
    copy of `copy_mib` MiB run first in the same child.  Returns
    {kernel symbol fragment: bytes} for the symbols in PHASE_KERNELS, or an
    {"error": ...} dict."""
    import csv
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if exe is None:
        return {"error": "rocprofv3 not found"}
    per_counter = {}
    scale = {}
    names = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="vqc_pmc_", dir="/tmp")
        cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "run", "--",
               sys.executable, str(ROOT / "bench.py"), "--only", workload, "--precision", precision,
               "--steps", "2", "--warmup", "1", "--copy-mib", str(copy_mib)]
        try:
            proc = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=timeout)
        except subprocess.TimeoutExpired:
            return {"error": f"rocprofv3 --pmc {counter} timed out"}
        files = list(Path(out).rglob("run_counter_collection.csv"))
        if proc.returncode != 0 or not files:
            return {"error": f"rocprofv3 --pmc {counter} exit {proc.returncode}: {proc.stderr[-300:]}"}
        rows = list(csv.DictReader(open(files[0])))
        shutil.rmtree(out, ignore_errors=True)
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        # The calibration copy is the first blit after the fill of its source.
        fill = next(i for i, r in enumerate(rows) if "FillFunctor" in r["Kernel_Name"])
        copy = next(r for r in rows[fill + 1:] if "copyBuffer" in r["Kernel_Name"])
        scale[counter] = copy_mib * 1024.0 / float(copy["Counter_Value"])
        vals = {}
        for _, symbols in PHASE_KERNELS.values():
            sym = symbols[precision]
            v = [float(r["Counter_Value"]) for r in rows if sym in r["Kernel_Name"]]
            if v:
                vals[sym] = sum(v) / len(v) * 1024.0 * scale[counter]
                names[sym] = next(r["Kernel_Name"] for r in rows if sym in r["Kernel_Name"])
        per_counter[counter] = vals
    traffic = {}
    for sym in per_counter["FETCH_SIZE"]:
        traffic[sym] = {
            "bytes": per_counter["FETCH_SIZE"][sym] + per_counter["WRITE_SIZE"].get(sym, 0.0),
            "fetch_bytes": per_counter["FETCH_SIZE"][sym],
            "write_bytes": per_counter["WRITE_SIZE"].get(sym, 0.0),
            "fetch_scale": scale["FETCH_SIZE"],
            "write_scale": scale["WRITE_SIZE"],
            "kernel_name": names.get(sym, sym).replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0],
        }
    return traffic


def _cpu_baseline(n, naux, nocc, fraction, seed=7):
    """The FP64 oracle's occupied-factor J/K (numpy/BLAS) on a bounded sample:
    synthetic B of shape (n, n, naux*fraction), time scaled by 1/fraction (the
    build is linear in naux)."""
    from oracle import oracle as O

    try:
        from threadpoolctl import threadpool_info

        threads = max((p.get("num_threads", 1) for p in threadpool_info() if p.get("user_api") == "blas"), default=1)
    except Exception:
        threads = os.cpu_count() or 1
    sub = max(1, min(naux, int(round(naux * fraction))))
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n, sub))
    B += B.transpose(1, 0, 2)
    C = rng.standard_normal((n, nocc))
    O.jk_occupied(B, C)  # prime
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.jk_occupied(B, C)
        runs.append(time.perf_counter() - t0)
    per_build = float(np.median(runs)) * (naux / sub)
    if sub == naux:
        how = "full size, nothing extrapolated"
    else:
        how = (f"extrapolated (x{naux / sub:.2f}, the build is linear in the auxiliary count) to naux={naux}; "
               f"a full-size FP64 B ({8 * n * n * naux / 1e9:.0f} GB) is beyond a bounded host sample")
    return {
        "value": 1.0 / per_build,
        "unit": "Fock builds/s",
        "cores": int(threads),
        "kind": "port",
        "sample": (
            f"oracle jk_occupied (numpy FP64, OpenBLAS) on synthetic B {n}x{n}x{sub} and C {n}x{nocc}: "
            f"median of 5 after a prime, {how}"
        ),
    }


def _ragged_systems(count=512, seed=2):
    """Ragged batch: (H2O)n with n ~ U{1..16} from a seed, each in a cube of
    side 3 n^(1/3) A (rigid monomers, O-O >= 2.6 A)."""
    from vibeqc_amd.geometry import ragged_water_batch

    return ragged_water_batch(count, seed)


def _ragged_batch(count=512, precisions=("mixed", "fp64")):
    from vibeqc_amd import Calculator

    systems, sizes = _ragged_systems(count)
    out = {"molecules": count, "sizes_histogram": np.bincount(sizes, minlength=17)[1:].tolist()}
    energies = {}
    for precision in precisions:
        calc = Calculator(method="rhf", basis="cc-pvdz", auxiliary_basis="cc-pvdz-jkfit", precision=precision)
        t0 = time.perf_counter()
        with calc.prepare_batch(systems, warm_start=False) as batch:
            t1 = time.perf_counter()
            res = batch.execute(properties=("energy",))
            t2 = time.perf_counter()
        energies[precision] = res.energies
        out[precision] = {
            "molecules_per_s": count / (t2 - t0),
            "prepare_s": t1 - t0,
            "execute_s": t2 - t1,
            "failures": len(res.failure_indices),
            "buckets": len({it.bucket_id for it in res.items}),
            "mean_iterations": float(np.mean([it.iterations for it in res.items])),
        }
        _log(f"ragged batch {precision}: {out[precision]}")
    if len(energies) < 2:
        return out
    err = np.abs(energies["mixed"] - energies["fp64"])
    out["max_abs_energy_error_mixed_vs_fp64"] = float(np.nanmax(err))
    out["mean_abs_energy_error_mixed_vs_fp64"] = float(np.nanmean(err))
    return out


def _ragged_cpu_baseline(sample_sizes=(1, 3, 6, 10)):
    """The oracle's FP64 DF-RHF (C integrals + numpy SCF) on the first member
    of each sampled size of the same ragged batch, then the whole batch's
    rate from a power-law fit t(n) = a n^b of those per-molecule times,
    weighted by the batch's size histogram (the sample spans 1 to 10 waters in
    about 15 s on 16 host cores; the 11- to 16-water members, up to a minute
    each there, are extrapolated)."""
    from oracle import oracle as O
    from vibeqc_amd import shells_for

    all_systems, all_sizes = _ragged_systems()

    def oshells(shells):
        return [O.Shell(s.atom_index, s.l, tuple(p.exponent for p in s.primitives), tuple(p.coefficient for p in s.primitives)) for s in shells]

    times, picks = [], []
    for n in sample_sizes:
        k = int(np.argmax(all_sizes == n))
        atoms = all_systems[k]
        z = [a.atomic_number for a in atoms]
        xyz = np.array([a.position for a in atoms])
        t0 = time.perf_counter()
        problem = O.build_problem(z, xyz, oshells(shells_for("cc-pvdz", atoms)), True,
                                  oshells(shells_for("cc-pvdz-jkfit", atoms)), True)
        O.df_rhf(problem)
        times.append(time.perf_counter() - t0)
        picks.append(k)
    b, log_a = np.polyfit(np.log(sample_sizes), np.log(times), 1)
    per_size = np.exp(log_a) * np.arange(1, 17, dtype=float) ** b
    counts = np.bincount(all_sizes, minlength=17)[1:]
    total = float(np.dot(counts, per_size))
    return {
        "value": len(all_sizes) / total,
        "unit": "molecules/s",
        "cores": os.cpu_count() or 1,
        "kind": "port",
        "sample": f"oracle DF-RHF timed on batch members {picks} (sizes {list(sample_sizes)}: "
        f"{', '.join(f'{t:.2f}' for t in times)} s); the 512-member rate is extrapolated with the fit "
        f"t(n) = {np.exp(log_a):.3g} n^{b:.2f} s over the batch's size histogram",
        "measured_seconds": times,
    }


def _oh_radicals(count=8):
    """OH doublets with O-H from 1.76 to 1.90 bohr and a small off-axis tilt."""
    from vibeqc_amd.basis import Atom

    return [(Atom(8, (0.0, 0.0, 0.0)), Atom(1, (0.0, 0.03 * k, 1.76 + 0.02 * k))) for k in range(count)]


def _uhf_single(repeats=3):
    """UHF energy + forces of the OH radicals one plan at a time (cc-pVDZ /
    cc-pVDZ-JKFIT, 19 AO): the first pass includes plan setup on a cold
    process, the median of the later passes is reported as systems/s."""
    from vibeqc_amd import Calculator

    systems = _oh_radicals(4)
    calc = Calculator(method="uhf", energy_tolerance=1e-12, density_tolerance=1e-10)
    passes, energies, iterations = [], [], []
    for rep in range(repeats + 1):
        t0 = time.perf_counter()
        for atoms in systems:
            r = calc.singlepoint(atoms, multiplicity=2)
            if rep == 0:
                energies.append(r.energy)
                iterations.append(r.iterations)
        passes.append(time.perf_counter() - t0)
    med = float(np.median(passes[1:]))
    # A larger open-shell system: (H2O)8 ... OH (211 AO), SCF seconds on a prepared plan.
    from vibeqc_amd.basis import Atom
    from vibeqc_amd.geometry import water_cluster

    cluster = tuple(water_cluster(8, 6.0, seed=21)) + (Atom(8, (12.0, 0.0, 0.0)), Atom(1, (12.0, 0.0, 1.83)))
    with Calculator(method="uhf").prepare(cluster, multiplicity=2) as plan:
        plan.solve()
        t0 = time.perf_counter()
        big = plan.solve()
        big_s = time.perf_counter() - t0
        big_stats = plan.last_solve_stats()
    out = {
        "workload": "4 OH radicals, UHF doublets, cc-pVDZ (19 AO) / cc-pVDZ-JKFIT, energy + forces, one plan each",
        "systems": len(systems),
        "first_pass_s": passes[0],
        "median_pass_s": med,
        "systems_per_s": len(systems) / med,
        "energies": energies,
        "mean_iterations": float(np.mean(iterations)),
        "water8_oh": {
            "workload": "(H2O)8 ... OH doublet, cc-pVDZ (211 AO) / cc-pVDZ-JKFIT, UHF SCF on a prepared plan",
            "scf_s": big_s,
            "energy": big.energy,
            "iterations": big.iterations,
            "purified_densities": big_stats["purified_densities"],
        },
    }
    _log(f"UHF single plans: {out}")
    return out


def _uhf_cpu_baseline(sample=4):
    """The UHF oracle (C integrals, numpy SCF and gradient) on the same
    radicals, energy + forces."""
    from oracle import oracle as O
    from vibeqc_amd import shells_for

    def oshells(shells):
        return [O.Shell(s.atom_index, s.l, tuple(p.exponent for p in s.primitives), tuple(p.coefficient for p in s.primitives)) for s in shells]

    systems = _oh_radicals()[:sample]
    t0 = time.perf_counter()
    for atoms in systems:
        z = [a.atomic_number for a in atoms]
        xyz = np.array([a.position for a in atoms])
        orb, aux = oshells(shells_for("cc-pvdz", atoms)), oshells(shells_for("cc-pvdz-jkfit", atoms))
        problem = O.build_problem(z, xyz, orb, True, aux, True)
        r = O.df_uhf(problem, 5, 4, energy_tolerance=1e-12, density_tolerance=1e-10)
        O.df_uhf_gradient(z, xyz, orb, True, aux, True, problem, r, (5, 4))
    dt = time.perf_counter() - t0
    return {
        "value": sample / dt,
        "unit": "systems/s",
        "cores": os.cpu_count() or 1,
        "kind": "port",
        "sample": f"UHF oracle energy + forces on the {sample} radicals",
    }


def _scf_stats(plan, result, seconds):
    st = plan.last_solve_stats()
    return {"energy": result.energy, "iterations": result.iterations, "fock_builds": result.fock_builds,
            "reduced_iterations": st["reduced_iterations"], "purified_densities": st["purified_densities"],
            "incremental_builds": st["incremental_builds"], "single_term_builds": st["single_term_builds"],
            "scf_s": seconds}


def _scf_timing(precision):
    """(H2O)4 SCF on a fresh plan, timed after one untimed solve (first
    launches of the SCF kernels)."""
    from vibeqc_amd import AOBasis, FockPlan, shells_for
    from vibeqc_amd.geometry import water_cluster

    atoms = water_cluster(4, 6.0, seed=0)
    orb = AOBasis(atoms, shells_for("cc-pvdz", atoms))
    aux = AOBasis(atoms, shells_for("cc-pvdz-jkfit", atoms))
    t0 = time.perf_counter()
    with FockPlan(orb, auxiliary=aux, precision=precision) as plan:
        t1 = time.perf_counter()
        plan.solve(energy_tolerance=1e-10, density_tolerance=1e-8)
        t2 = time.perf_counter()
        r = plan.solve(energy_tolerance=1e-10, density_tolerance=1e-8)
        out = _scf_stats(plan, r, time.perf_counter() - t2)
    out["setup_s"] = t1 - t0
    return out


def _scf_on_plan(plan, precision):
    """Full SCF (core guess, 1e-10 Eh / 1e-8 RMS) on an existing plan."""
    plan.set_precision(precision)
    t0 = time.perf_counter()
    r = plan.solve(energy_tolerance=1e-10, density_tolerance=1e-8)
    return _scf_stats(plan, r, time.perf_counter() - t0)


def _profile_run(args):
    if args.copy_mib > 0:
        import torch

        count = args.copy_mib * (1 << 20) // 4
        src = torch.ones(count, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        _log(f"copied {args.copy_mib} MiB device-to-device 3 times")
    if args.only == "ragged":
        # The first members of the ragged batch, one precision ("fp64", else mixed).
        precision = "fp64" if args.precision == "fp64" else "mixed"
        print(json.dumps({"workload": "ragged", "results": _ragged_batch(args.molecules, (precision,))}))
        return
    plan, setup, nocc = _alkane_plan() if args.only == "c20h42" else _water64_plan()
    if args.only == "water64_scf":
        # The (H2O)64 SCF legs alone (mixed = the BF16X3 schedule, then FP64).
        precisions = (args.precision,) if args.precision else ("bf16x3", "fp64")
        res = {p: _scf_on_plan(plan, p) for p in precisions}
        for p, r in res.items():
            _log(f"(H2O)64 SCF {p}: {r}")
    else:
        precisions = (args.precision,) if args.precision else PRECISIONS
        res = _time_precisions(plan, nocc, args.steps, args.warmup, precisions)
    plan.close()
    print(json.dumps({"workload": args.only, "setup_s": setup, "results": res}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 PMC byte-count runs")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-molecule ragged batch and the UHF runs")
    ap.add_argument("--no-scf64", action="store_true", help="skip the (H2O)64 FP64 and mixed SCF runs")
    ap.add_argument(
        "--only",
        choices=("c20h42", "water64", "water64_scf", "ragged"),
        help="profiling run: time one workload only (no CPU baseline) and print its phase times; "
        "water64_scf runs the (H2O)64 SCF legs only, ragged the first --molecules members of the ragged batch",
    )
    ap.add_argument("--molecules", type=int, default=64, help="with --only ragged: batch size")
    ap.add_argument("--precision", choices=PRECISIONS, help="with --only: time this precision only")
    ap.add_argument(
        "--copy-mib",
        type=int,
        default=0,
        help="with --only: first run three device-to-device copies of this many MiB "
        "(a known byte count to calibrate hardware byte counters against)",
    )
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one MI355X: the engine runs on one GPU (a multi-GPU run would "
            "shard the auxiliary index and all-reduce J and K). Run with --gpus 1.",
            file=sys.stderr,
        )
        sys.exit(2)
    if args.only:
        _profile_run(args)
        return

    # Headline workload: (H2O)64, the large single system of the mixed/FP64 target.
    wplan, wsetup, wnocc = _water64_plan()
    wshape = wplan.shape()
    wn, wnaux, wnpair = wshape["nbf"], wshape["naux"], wshape["npair"]
    _log(f"(H2O)64 plan: nbf {wn}, naux {wnaux}, setup {wsetup:.2f} s")
    w64 = _time_precisions(wplan, wnocc, args.steps, args.warmup)
    w64_scf = None
    if not args.no_scf64:
        w64_scf = {"fp64": _scf_on_plan(wplan, "fp64"), "mixed": _scf_on_plan(wplan, "bf16x3")}
        w64_scf["energy_diff_mixed_minus_fp64"] = w64_scf["mixed"]["energy"] - w64_scf["fp64"]["energy"]
        w64_scf["mixed_speedup_vs_fp64"] = w64_scf["fp64"]["scf_s"] / w64_scf["mixed"]["scf_s"]
        _log(f"(H2O)64 SCF: {w64_scf}")
    wplan.close()

    aplan, asetup, anocc = _alkane_plan()
    ashape = aplan.shape()
    an, anaux, anpair = ashape["nbf"], ashape["naux"], ashape["npair"]
    _log(f"C20H42 plan: nbf {an}, naux {anaux}, setup {asetup:.2f} s")
    alk = _time_precisions(aplan, anocc, args.steps, args.warmup)
    aplan.close()

    cpu = cpu64 = None
    if not args.no_cpu_baseline:
        cpu64 = _cpu_baseline(wn, wnaux, wnocc, fraction=1.0 / 32.0)
        _log(f"cpu baseline (H2O)64: {cpu64}")
        cpu = _cpu_baseline(an, anaux, anocc, fraction=1.0)
        _log(f"cpu baseline C20H42: {cpu}")
    scf64 = _scf_timing("fp64")
    scfmx = _scf_timing("bf16x3")
    ragged = None
    if not args.no_batch:
        ragged = _ragged_batch()
        if not args.no_cpu_baseline:
            ragged["cpu_baseline"] = _ragged_cpu_baseline()
            _log(f"ragged cpu baseline: {ragged['cpu_baseline']}")
    uhf = None
    if not args.no_batch:
        uhf = _uhf_single()
        if not args.no_cpu_baseline:
            uhf["cpu_baseline"] = _uhf_cpu_baseline()
            _log(f"UHF cpu baseline: {uhf['cpu_baseline']}")
    traffic = traffic64 = None
    if not args.no_traffic:
        traffic64 = _pmc_traffic("water64", "bf16x3")
        _log(f"PMC traffic (H2O)64: {traffic64}")
        traffic = _pmc_traffic("c20h42", "bf16x3")
        _log(f"PMC traffic C20H42: {traffic}")
    line = {
        "metric": METRIC,
        "value": w64["bf16x3"]["builds_per_s"],
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": w64["bf16x3"]["ms_per_build_wall"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3",
        "data": "synthetic",
        "config": {
            "workload": "(H2O)64 at liquid density (12.4 A cube, seed 4), cc-pVDZ (1536 AO) / cc-pVDZ-JKFIT "
            "(7424 aux), DF-JK Fock build at fixed density from 320 Lowdin-orthonormal occupied columns; "
            "mixed = BF16X3 J/K (split bfloat16, FP32 accumulation, J summed in FP64)",
            "nbf": wn,
            "naux": wnaux,
            "nocc": wnocc,
            "parallelism": "single-gpu",
            "setup_s": wsetup,
        },
        "fp64_builds_per_s": w64["fp64"]["builds_per_s"],
        "fp32_builds_per_s": w64["fp32"]["builds_per_s"],
        "mixed_speedup_vs_fp64": w64["bf16x3"]["builds_per_s"] / w64["fp64"]["builds_per_s"],
        "target_mixed_over_fp64": 4.0,
        "precision_results": w64,
        "roofline": _roofline(w64["bf16x3"]["phase_ms"], wn, wnaux, wnpair, wnocc, "bf16x3", traffic64),
        "roofline_fp64": _roofline(w64["fp64"]["phase_ms"], wn, wnaux, wnpair, wnocc, "fp64"),
        "scf": w64_scf,
        "cpu_baseline": cpu64,
        "c20h42": {
            "workload": "C20H42 single DF-JK Fock build at fixed density, def2-SVP (490 AO) / "
            "JKFIT-shaped auxiliary (2366), 81 occupied",
            "nbf": an,
            "naux": anaux,
            "nocc": anocc,
            "setup_s": asetup,
            "mixed_builds_per_s": alk["bf16x3"]["builds_per_s"],
            "fp64_builds_per_s": alk["fp64"]["builds_per_s"],
            "fp32_builds_per_s": alk["fp32"]["builds_per_s"],
            "mixed_speedup_vs_fp64": alk["bf16x3"]["builds_per_s"] / alk["fp64"]["builds_per_s"],
            "precision_results": alk,
            "roofline": _roofline(alk["bf16x3"]["phase_ms"], an, anaux, anpair, anocc, "bf16x3", traffic),
            "cpu_baseline": cpu,
        },
        "ragged_batch": ragged,
        "uhf": uhf,
        "scf_water4": {
            "fp64": scf64,
            "mixed": scfmx,
            "energy_diff_mixed_minus_fp64": scfmx["energy"] - scf64["energy"],
        },
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
