/*
 * fock.h -- prepared DF-JK Fock operator C ABI of the MI355X engine.
 *
 * Same names, struct layouts, status codes and ownership rules as the
 * reference's include/generativeqc/fock.h:1-171 (the resident Krylov response
 * owners at fock.h:172-296 are out of scope).  Plans execute on one HIP device
 * (gfx950): density-fitted J and K with the (P|Q) metric factored by Cholesky
 * (eigen-cutoff fallback when the metric is near-singular), a device-resident
 * RHF SCF loop, and optional reduced-precision Fock builds selected through the
 * vibeqc_ extensions at the end of this file.
 *
 * Conventions kept from the reference:
 *   - D is row-major nbf x nbf and includes the factor 2 for RHF;
 *   - raw J/K are unscaled, F = Hcore + cJ J + cK K;
 *   - output buffers are caller-owned, must be disjoint, and are written only
 *     on success; inputs are copied first and may alias outputs;
 *   - a plan is not re-entrant: serialise calls on one handle and its destroy.
 */
#ifndef GENERATIVEQC_FOCK_H
#define GENERATIVEQC_FOCK_H

#include "generativeqc/generativeqc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct generativeqc_fock_plan generativeqc_fock_plan;
typedef int32_t generativeqc_fock_spin;
enum { GENERATIVEQC_FOCK_RESTRICTED = 0, GENERATIVEQC_FOCK_UNRESTRICTED = 1 };
typedef int32_t generativeqc_fock_operator;
enum {
  GENERATIVEQC_FOCK_FULL_RANGE = 0,
  GENERATIVEQC_FOCK_SHORT_RANGE = 1,
  GENERATIVEQC_FOCK_LONG_RANGE = 2
};
typedef int32_t generativeqc_fock_approximation;
enum { GENERATIVEQC_FOCK_EXACT = 0, GENERATIVEQC_FOCK_DENSITY_FITTED = 1 };

/* Reference: fock.h:25-31.  Only FULL_RANGE, DENSITY_FITTED terms are
 * implemented; other requests return NOT_IMPLEMENTED. */
typedef struct generativeqc_fock_term {
  int32_t present;
  double coefficient;
  generativeqc_fock_operator op;
  double omega;
  generativeqc_fock_approximation approximation;
} generativeqc_fock_term;

/* Reference: fock.h:37-42.  RESTRICTED: D includes double occupation, HF is
 * cJ = 1, cK = -0.5.  UNRESTRICTED: Da and Db have unit occupation, J comes
 * from Da + Db and K from each spin; HF is cJ = 1, cK = -1. */
typedef struct generativeqc_fock_spec {
  uint32_t struct_size, abi_version, spec_version;
  generativeqc_fock_spin spin;
  uint32_t derivative_order;
  generativeqc_fock_term coulomb, exchange;
} generativeqc_fock_spec;

/* Reference: fock.h:51-55.  NULL controls: screening 1e-12, metric cutoff
 * 1e-10, budget 0.  A zero cutoff selects 1e-10. */
typedef struct generativeqc_fock_controls {
  uint32_t struct_size, abi_version;
  double screening_tolerance, metric_relative_threshold;
  uint64_t device_budget_bytes;
} generativeqc_fock_controls;

/* Reference: fock.h:65-68.  A NULL auxiliary uses the orbital basis as the
 * fitting basis.  Failure leaves *output NULL. */
GENERATIVEQC_API generativeqc_status generativeqc_fock_plan_create(
    generativeqc_context* context, const generativeqc_system* system,
    const generativeqc_system* auxiliary, const generativeqc_fock_spec* spec,
    const generativeqc_fock_controls* controls, generativeqc_fock_plan** output);
/* Reference: fock.h:69. */
GENERATIVEQC_API void generativeqc_fock_plan_destroy(generativeqc_fock_plan* plan);
/* Reference: fock.h:71. */
GENERATIVEQC_API const char* generativeqc_fock_plan_last_error(const generativeqc_fock_plan* plan);

/* Reference: fock.h:81-91. */
typedef struct generativeqc_fock_result {
  uint32_t struct_size, abi_version;
  uint64_t matrix_count, gradient_count;
  double* coulomb;
  double* exchange_alpha;
  double* exchange_beta;
  double* fock_alpha;
  double* fock_beta;
  double* gradient;
  double energy_one_electron, energy_two_electron, nuclear_repulsion;
} generativeqc_fock_result;

/* Reference: fock.h:93-95.  The density must be symmetric (K is built from a
 * factor of D); an asymmetric density returns INVALID_ARGUMENT.  Unrestricted
 * plans take the beta density in `beta` and fill exchange_beta / fock_beta. */
GENERATIVEQC_API generativeqc_status generativeqc_fock_plan_evaluate(
    generativeqc_fock_plan* plan, const double* density, uint64_t density_count, const double* beta,
    uint64_t beta_count, generativeqc_fock_result* result);

/* Reference: fock.h:100-104.  NULL: 100 iterations, DIIS 8, 1e-10, 1e-8. */
typedef struct generativeqc_fock_scf_controls {
  uint32_t struct_size, abi_version;
  uint32_t max_iterations, diis_history;
  double energy_tolerance, density_tolerance;
} generativeqc_fock_scf_controls;

/* Reference: fock.h:115-124. */
typedef struct generativeqc_fock_scf_result {
  uint32_t struct_size, abi_version;
  double* density;
  double* forces;
  uint64_t density_count, force_count;
  double energy, energy_change, density_rms;
  uint32_t iterations;
  int32_t initial_density_used;
  uint64_t fock_builds;
} generativeqc_fock_scf_result;

/* Reference: fock.h:134-137.  The whole loop (Fock build, DIIS, eigensolve,
 * density) runs on the device; only per-iteration convergence scalars come
 * back to the host.  With forces non-null (derivative_order 1) the analytic
 * forces -dE/dR of the converged energy are returned.  Unrestricted plans run
 * UHF: initial_density and density are [alpha; beta], 2 nbf^2 values. */
GENERATIVEQC_API generativeqc_status generativeqc_fock_plan_solve(
    generativeqc_fock_plan* plan, const generativeqc_fock_scf_controls* controls,
    const double* initial_density, uint64_t initial_density_count,
    generativeqc_fock_scf_result* result);

/* Reference: fock.h:141-147. */
enum {
  GENERATIVEQC_FOCK_CPU_REFERENCE = 0,
  GENERATIVEQC_FOCK_CUDA_FUSED = 1,
  GENERATIVEQC_FOCK_STANDARD_DF = 2,
  GENERATIVEQC_FOCK_CPU_INDEPENDENT = 3,
  GENERATIVEQC_FOCK_CUDA_INDEPENDENT = 4
};

/* Reference: fock.h:154-166. */
typedef struct generativeqc_fock_diagnostic {
  uint32_t struct_size, abi_version;
  generativeqc_fock_spec requested, resolved;
  generativeqc_backend backend;
  int32_t resolved_schedule, source_schedule;
  uint64_t nbf, coordinate_count, device_bytes, device_budget_bytes;
  uint64_t auxiliary_rank, auxiliary_tile;
  double screening_tolerance, metric_relative_threshold;
  int32_t df_streamed;
  char direct_schedule[96], df_value_backend[48], df_value_mapping[32], df_response_mapping[32];
  char one_electron_value_backend[32], one_electron_value_mapping[32];
  char one_electron_response_mapping[32];
} generativeqc_fock_diagnostic;
/* Reference: fock.h:167-168. */
GENERATIVEQC_API generativeqc_status generativeqc_fock_plan_diagnostic(
    const generativeqc_fock_plan* plan, generativeqc_fock_diagnostic* diagnostic);

/* ------------------------------------------------------------------------
 * Extensions (no counterpart in the reference's fock.h).
 * ------------------------------------------------------------------------ */

/* Arithmetic of the J/K contractions.  FP64 is the reference's only DF
 * precision (src/scf/rhf.cpp:1153-1155).  FP32 stores B in FP32 and runs the K
 * GEMMs on FP32 MFMA with FP32 accumulation; BF16X3 keeps B in FP32 in HBM and
 * splits every operand into hi+lo bfloat16 halves on load, accumulating
 * hi*hi + hi*lo + lo*hi on BF16 MFMA in FP32.  J always accumulates in FP64. */
typedef int32_t vibeqc_precision;
enum { VIBEQC_PRECISION_FP64 = 0, VIBEQC_PRECISION_FP32 = 1, VIBEQC_PRECISION_BF16X3 = 2 };

/* Select the arithmetic of later evaluate/build calls on this plan.  Reduced
 * precision keeps an FP32 copy of B beside the FP64 one.  solve() with a
 * reduced precision iterates in that precision to convergence and then
 * continues with FP64 Fock builds until the FP64 criteria hold (refinement). */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_set_precision(generativeqc_fock_plan* plan,
                                                                    vibeqc_precision precision);

/* Raw J and K from an occupied factor: D = 2 C C^T with C row-major
 * nbf x nocc (the factor used inside the SCF, df_occupied_exchange.cpp:236-300).
 * coulomb/exchange are nbf^2 outputs (either may be NULL). */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_evaluate_occupied(
    generativeqc_fock_plan* plan, const double* occupied, uint64_t nbf, uint64_t nocc,
    double* coulomb, double* exchange);

/* Extension with no counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): vibeqc_fock_plan_evaluate_occupied with a
 * device-resident factor (row-major nbf x nocc, 1 <= nocc <= nbf) and
 * device-resident nbf^2 outputs, at the plan's precision (a reduced precision
 * must have been selected by vibeqc_fock_plan_set_precision, which makes its
 * copy of B).  Every workspace lives in one fixed 128 MiB array per device
 * (held until the process ends); no device or host memory is allocated.
 * Calls whose workspace would exceed it return INVALID_ARGUMENT.  All calls on
 * one device share that array: calls that overlap on one device race (the
 * caller serialises them).  Returns after J and K are written. */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_evaluate_occupied_noalloc(
    generativeqc_fock_plan* plan, const double* occupied_device, uint64_t nbf, uint64_t nocc,
    double* coulomb_device, double* exchange_device);

/* Device-resident timing harness: uploads C once (as above), runs one
 * untimed build, synchronises, then runs `repeat` complete J+K builds on the
 * device with no host transfer in between and synchronises again.
 * kernel_ms (may be NULL, else 5 entries) receives per-build averages:
 * device times from HIP events on the plan's stream, [0] U = B C (with the rho partials),
 * [1] K = U U^T, [2] J = B rho, [3] whole build, and [4] the host wall time
 * between the two synchronisations divided by `repeat`. */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_benchmark_occupied(
    generativeqc_fock_plan* plan, const double* occupied, uint64_t nbf, uint64_t nocc,
    uint32_t repeat, double* kernel_ms);

/* Copies of the plan's one-electron matrices (nbf^2 each, row-major public AO
 * order): the overlap S and Hcore = T + V.  Either pointer may be NULL. */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_one_electron(const generativeqc_fock_plan* plan,
                                                                   double* overlap, double* hcore,
                                                                   uint64_t matrix_count);

/* Sizes of the prepared plan: [0] nbf, [1] auxiliary functions, [2] packed
 * AO pairs nbf(nbf+1)/2, [3] metric rank, [4] 1 if the Cholesky path was used,
 * 0 for the eigen-cutoff path, [5] device bytes held. */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_shape(const generativeqc_fock_plan* plan,
                                                            uint64_t* shape6);

/* Extension (no counterpart in the reference): how the plan honours its
 * device_budget_bytes control.  [0] 1 when the fitted tensor does not fit the
 * budget and is regenerated in auxiliary blocks for every Fock build (the
 * role of the reference's budgeted tiles, density_fitting.cpp:1041-1218),
 * 0 when it is resident; [1] auxiliary functions per block; [2] three-centre
 * row tiles per block.  A streamed plan builds J/K and runs its SCF in FP64;
 * nuclear derivatives and the allocation-free entry need a resident plan
 * (NOT_IMPLEMENTED).  A budget below the metric, its factor and one block
 * fails plan creation with OUT_OF_MEMORY. */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_streaming(const generativeqc_fock_plan* plan,
                                                                uint64_t* streaming3);

/* Extension (no counterpart in the reference): statistics of the plan's last
 * generativeqc_fock_plan_solve, the first min(count, 5) of: [0] SCF
 * iterations, [1] of them with reduced-precision Fock builds, [2] new
 * densities obtained by purification (the rest by diagonalisation), [3]
 * FP64-stage builds of a BF16X3 solve done incrementally (exchange from the
 * last full FP64 build plus a BF16X3 update of the occupied factor's change),
 * [4] BF16X3-stage builds with one bf16 product per logical one (the first
 * iterations of a BF16X3 solve). */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_last_solve_stats(const generativeqc_fock_plan* plan,
                                                                       uint64_t* stats, uint64_t count);

/* Extension (no counterpart in the reference's API; its schedules are tuned
 * at kernel generation time and reused per device, generativeqc.h:1191-1212):
 * measured launch shapes for the occupied-factor J/K build.  Times the build
 * of C (row-major nbf x nocc, as vibeqc_fock_plan_benchmark_occupied) at the
 * plan's precision for the default shape and for alternative U i-tile widths
 * (32-column units; the occupied columns in one, two or three i blocks) and
 * SYRK split-K counts (half and twice the default), `repeat` builds each, and
 * records the fastest for this device architecture, precision, nbf, nocc and
 * padded auxiliary count in the process-wide registry that later builds of
 * that shape (on any plan) use.  shape2 (may be NULL) receives the chosen
 * [U tile, SYRK splits]; ms2 (may be NULL) [best, default] average build ms.
 * A different shape only regroups sums: J is unchanged and K changes at
 * rounding level.  Resident plans only (NOT_IMPLEMENTED when streamed). */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_autotune(generativeqc_fock_plan* plan, const double* occupied,
                                                               uint64_t nbf, uint64_t nocc, uint32_t repeat,
                                                               int32_t* shape2, double* ms2);

/* The [U tile, SYRK splits] that builds of nocc occupied columns use at the
 * plan's current precision (recorded shape, else the default rules). */
GENERATIVEQC_API generativeqc_status vibeqc_fock_plan_launch_shape(const generativeqc_fock_plan* plan, uint64_t nocc,
                                                                   int32_t* shape2);

/* The launch-shape registry itself (host only, no device needed).  Keys are
 * (arch, precision 0 FP64 / 1 FP32 / 2 BF16X3, nbf, occupied columns, ldq =
 * auxiliary count rounded up to 128).  u_tile 0..10 and syrk_splits 0..128,
 * 0 meaning the default rule.  With VIBEQC_TUNING_FILE set, the registry
 * reads that file on first use and rewrites it after every record.  Files
 * are text: the line "vibeqc-jk-tuning 1", then one
 * "<arch> <precision> <nbf> <occupied> <ldq> <u_tile> <syrk_splits> <ms>"
 * line per entry.  load returns *found = 0 for a missing file and
 * INVALID_ARGUMENT for a malformed one; record and save return
 * INTERNAL_ERROR when a file cannot be written. */
GENERATIVEQC_API generativeqc_status vibeqc_tuning_record(const char* arch, int32_t precision, uint64_t nbf,
                                                          uint64_t occupied, uint64_t ldq, int32_t u_tile,
                                                          int32_t syrk_splits, double ms);
GENERATIVEQC_API generativeqc_status vibeqc_tuning_lookup(const char* arch, int32_t precision, uint64_t nbf,
                                                          uint64_t occupied, uint64_t ldq, int32_t* shape2,
                                                          int32_t* found);
GENERATIVEQC_API generativeqc_status vibeqc_tuning_count(uint64_t* count);
GENERATIVEQC_API generativeqc_status vibeqc_tuning_load(const char* path, int32_t* found);
GENERATIVEQC_API generativeqc_status vibeqc_tuning_save(const char* path);
GENERATIVEQC_API generativeqc_status vibeqc_tuning_clear(void);

#ifdef __cplusplus
}
#endif
#endif
