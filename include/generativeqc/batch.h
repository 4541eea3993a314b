/*
 * batch.h -- prepared ragged batch ("fleet") of DF-RHF systems on one HIP
 * device.
 *
 * Same names, struct layouts and semantics as the reference's batch entry
 * points for the RHF density-fitted path (include/generativeqc/generativeqc.h:
 * method descriptor :723-769, batch flags :376-397, input/result descriptors
 * :1158-1183, prepare/destroy/count :1431-1441, execute :1574-1586, warm-start
 * control :1563-1572, Fock-build count :1292-1297; method ids from
 * generated_method_ids.h:11-12).  GENERATIVEQC_METHOD_RHF and
 * GENERATIVEQC_METHOD_UHF with density fitting (CUDA or AUTO mode, executed
 * on the HIP device) are implemented; other methods, KS options and initial
 * guess options return NOT_IMPLEMENTED.  A UHF batch retains [alpha; beta]
 * densities (2 nbf^2 values) as warm states.  precision_mode AUTO runs BF16X3 Fock
 * builds followed by an FP64 refinement to the requested tolerances.
 */
#ifndef GENERATIVEQC_BATCH_H
#define GENERATIVEQC_BATCH_H

#include "generativeqc/generativeqc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t generativeqc_method;
enum { GENERATIVEQC_METHOD_RHF = 1, GENERATIVEQC_METHOD_UHF = 2 };

typedef int32_t generativeqc_density_fitting_mode;
enum {
  GENERATIVEQC_DENSITY_FITTING_NONE = 0,
  GENERATIVEQC_DENSITY_FITTING_CPU_REFERENCE = 1,
  GENERATIVEQC_DENSITY_FITTING_CUDA = 2,
  GENERATIVEQC_DENSITY_FITTING_AUTO = 3
};

typedef int32_t generativeqc_precision_mode;
enum { GENERATIVEQC_PRECISION_FP64 = 0, GENERATIVEQC_PRECISION_AUTO = 1 };

/* Options of other method families; only NULL is accepted here. */
typedef struct generativeqc_ks_options generativeqc_ks_options;
typedef struct generativeqc_initial_guess_options generativeqc_initial_guess_options;

/* Reference: generativeqc.h:723-769.  Zero counts/tolerances select the
 * defaults (100 iterations, DIIS 8, 1e-10 Eh, 1e-8 RMS, screening 1e-12,
 * metric cutoff 1e-10). */
typedef struct generativeqc_method_descriptor {
  uint32_t struct_size;
  uint32_t abi_version;
  generativeqc_method method;
  uint32_t max_iterations;
  uint32_t diis_history;
  double energy_tolerance;
  double density_tolerance;
  double screening_tolerance;
  generativeqc_density_fitting_mode density_fitting_mode;
  const generativeqc_system* density_fitting_auxiliary_basis;
  double density_fitting_relative_threshold;
  uint64_t density_fitting_memory_budget_bytes;
  generativeqc_precision_mode precision_mode;
  uint64_t correlation_memory_budget_bytes;
  double mp2_denominator_threshold;
  const generativeqc_ks_options* ks_options;
  uint32_t ccsd_max_iterations;
  uint32_t ccsd_diis_history;
  double ccsd_energy_tolerance;
  double ccsd_residual_tolerance;
  double ccsd_denominator_threshold;
  double ccsd_damping;
  double ccsd_level_shift;
  uint32_t ccsd_frozen_core;
  const generativeqc_initial_guess_options* initial_guess;
} generativeqc_method_descriptor;

typedef struct generativeqc_batch generativeqc_batch;

/* Reference: generativeqc.h:376-397.  Only warm starts are honoured; the
 * CUDA profiling flags are accepted and ignored. */
typedef uint32_t generativeqc_batch_flags;
enum {
  GENERATIVEQC_BATCH_ENABLE_WARM_STARTS = 1u << 0,
  GENERATIVEQC_BATCH_ENABLE_SHELL_CLASS_PROFILING = 1u << 1,
  GENERATIVEQC_BATCH_ENABLE_INACTIVE_EIGENSOLVER_PROFILING = 1u << 2
};

/* Reference: generativeqc.h:1158-1165. */
typedef struct generativeqc_batch_input_descriptor {
  uint32_t struct_size;
  uint32_t abi_version;
  const double* coordinates;
  uint32_t coordinate_count;
} generativeqc_batch_input_descriptor;

/* Reference: generativeqc.h:1168-1183. */
typedef struct generativeqc_batch_item_result_descriptor {
  uint32_t struct_size;
  uint32_t abi_version;
  generativeqc_status status;
  double energy;
  double* forces;
  uint32_t force_count;
  uint32_t iterations;
  double energy_change;
  double density_rms;
  int32_t converged;
  generativeqc_backend executed_backend;
  uint32_t bucket_id;
  int32_t warm_start_used;
  int32_t warm_start_fallback;
} generativeqc_batch_item_result_descriptor;

/* Reference: generativeqc.h:1431-1437.  Systems may differ in atoms, shells
 * and AO counts.  The auxiliary basis of the descriptor is a template over
 * atom indices: every system must have the template's atom count and uses
 * the template's shells on its own atoms (the reference's topology rule). */
GENERATIVEQC_API generativeqc_status generativeqc_batch_prepare(generativeqc_context* context,
                                                                const generativeqc_system* const* systems,
                                                                uint32_t system_count,
                                                                const generativeqc_method_descriptor* descriptor,
                                                                generativeqc_batch_flags flags,
                                                                generativeqc_batch** batch);
/* Reference: generativeqc.h:1439. */
GENERATIVEQC_API void generativeqc_batch_destroy(generativeqc_batch* batch);
/* Reference: generativeqc.h:1441. */
GENERATIVEQC_API uint32_t generativeqc_batch_get_system_count(const generativeqc_batch* batch);
/* Reference: generativeqc.h:1574-1586.  Structural errors return a status;
 * otherwise each item carries its own status.  inputs may be NULL with
 * input_count 0 to reuse the prepared geometries.  An item with forces NULL
 * and force_count 0 is energy-only; otherwise force_count is 3 natom. */
GENERATIVEQC_API generativeqc_status generativeqc_batch_execute(generativeqc_batch* batch,
                                                                const generativeqc_batch_input_descriptor* inputs,
                                                                uint32_t input_count,
                                                                generativeqc_batch_item_result_descriptor* results,
                                                                uint32_t result_count);
/* Reference: generativeqc.h:1563-1572. */
GENERATIVEQC_API generativeqc_status generativeqc_batch_set_warm_start_updates(generativeqc_batch* batch,
                                                                               int32_t enabled);
/* Reference: generativeqc.h:1292-1297: physical Fock builds of item `index`
 * in the last execution. */
GENERATIVEQC_API generativeqc_status generativeqc_batch_get_last_fock_builds(const generativeqc_batch* batch,
                                                                             uint32_t index, uint64_t* builds);

/* Reference: generativeqc.h:1497-1509: a retained converged density (unit
 * occupation doubled: the restricted D), the coordinates it converged at and
 * its source diagnostics. */
typedef struct generativeqc_hf_warm_state {
  uint32_t struct_size;
  uint32_t abi_version;
  double* density;
  uint64_t density_count;
  double* coordinates;
  uint64_t coordinate_count;
  double energy;
  double energy_change;
  double density_rms;
  int32_t iterations;
  int32_t present;
} generativeqc_hf_warm_state;

/* Reference: generativeqc.h:1511-1514.  With both buffers NULL the call fills
 * the counts, diagnostics and `present`; otherwise the counts must equal the
 * retained sizes and both arrays are copied.  No seed: present = 0, counts 0. */
GENERATIVEQC_API generativeqc_status generativeqc_batch_get_hf_warm_state(const generativeqc_batch* batch,
                                                                          uint32_t index,
                                                                          generativeqc_hf_warm_state* state);
/* Reference: generativeqc.h:1516-1525.  `count` must equal the system count;
 * present = 0 keeps that item's seed.  Every present state is validated
 * (shape, finiteness, symmetry) before any is applied; warm starts must be
 * enabled.  The caller checks scientific identity before calling. */
GENERATIVEQC_API generativeqc_status generativeqc_batch_restore_hf_warm_states(generativeqc_batch* batch,
                                                                               const generativeqc_hf_warm_state* states,
                                                                               uint32_t count);
/* Reference: generativeqc.h:1527-1529: discard every retained seed. */
GENERATIVEQC_API generativeqc_status generativeqc_batch_clear_warm_starts(generativeqc_batch* batch);

/* Extension (no counterpart in the reference): one auxiliary system per batch
 * member, so members need not share an atom topology.  auxiliary_systems may
 * be NULL to fit with each orbital basis itself. */
GENERATIVEQC_API generativeqc_status vibeqc_batch_prepare(generativeqc_context* context,
                                                          const generativeqc_system* const* systems,
                                                          const generativeqc_system* const* auxiliary_systems,
                                                          uint32_t system_count,
                                                          const generativeqc_method_descriptor* descriptor,
                                                          generativeqc_batch_flags flags,
                                                          generativeqc_batch** batch);
/* Extension: shape-bucket id of item `index` (stable for the batch). */
GENERATIVEQC_API generativeqc_status vibeqc_batch_bucket(const generativeqc_batch* batch, uint32_t index,
                                                         uint32_t* bucket_id);

#ifdef __cplusplus
}
#endif
#endif
