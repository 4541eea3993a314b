/*
 * calculation.h -- prepared single-system calculation (the native entry point
 * behind the reference's Calculator.singlepoint) on one HIP device.
 *
 * Same names, struct layouts and semantics as the reference's calculation
 * entry points (include/generativeqc/generativeqc.h: result descriptor
 * :1067-1081, prepare/destroy :1335-1338, execute :1340-1347; behaviour
 * src/api/c_api_calculation.cpp:24-131, descriptor defaults
 * src/methods/hf_method.cpp:28-115).  The method descriptor is the batch's
 * (batch.h).  GENERATIVEQC_METHOD_RHF and GENERATIVEQC_METHOD_UHF with
 * density fitting (CUDA or AUTO mode, executed on the HIP device) are
 * implemented; exact (non-fitted) J/K, the CPU reference backend, KS and
 * initial-guess options return NOT_IMPLEMENTED.  precision_mode AUTO runs
 * BF16X3 Fock builds followed by an FP64 refinement to the requested
 * tolerances.  A calculation is not re-entrant: the caller serialises calls
 * and destruction.
 */
#ifndef GENERATIVEQC_CALCULATION_H
#define GENERATIVEQC_CALCULATION_H

#include "generativeqc/batch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct generativeqc_calculation generativeqc_calculation;

/* Reference: generativeqc.h:1067-1081. */
typedef struct generativeqc_result_descriptor {
  uint32_t struct_size;
  uint32_t abi_version;
  double energy;
  double* forces;
  uint32_t force_count;
  uint32_t iterations;
  double energy_change;
  double density_rms;
  int32_t converged;
  generativeqc_backend executed_backend;
} generativeqc_result_descriptor;

/* Reference: generativeqc.h:1335-1337 (c_api_calculation.cpp:24-46).  The
 * plan (integrals, metric factor, fitted tensor) is built here; *calculation
 * is NULL on failure and the context's last detail says why.  The auxiliary
 * basis of the descriptor is placed on the system's atoms; NULL fits with
 * the orbital basis itself. */
GENERATIVEQC_API generativeqc_status generativeqc_calculation_prepare(generativeqc_context* context,
                                                                      const generativeqc_system* system,
                                                                      const generativeqc_method_descriptor* descriptor,
                                                                      generativeqc_calculation** calculation);
/* Reference: generativeqc.h:1338. */
GENERATIVEQC_API void generativeqc_calculation_destroy(generativeqc_calculation* calculation);
/* Reference: generativeqc.h:1340-1347 (c_api_calculation.cpp:53-131).  A cold
 * SCF from the core-Hamiltonian guess on every call.  forces NULL with
 * force_count 0 requests the energy only (no force work runs); otherwise
 * force_count must be at least 3 natom.  A run that does not converge fills
 * the diagnostics and returns NOT_CONVERGED; forces are written only on
 * success. */
GENERATIVEQC_API generativeqc_status generativeqc_calculation_execute(generativeqc_calculation* calculation,
                                                                      generativeqc_result_descriptor* result);

/* Extension (no counterpart in the reference): J/K precision of the
 * iterations before the FP64 refinement (0 FP64, 1 FP32, 2 BF16X3). */
GENERATIVEQC_API generativeqc_status vibeqc_calculation_set_precision(generativeqc_calculation* calculation,
                                                                      int32_t precision);
/* Extension: the last successful execution's density (nbf^2 values, or
 * 2 nbf^2 as [alpha; beta] for UHF), Fock builds, and plan shape.  With
 * density NULL only the counts are written. */
GENERATIVEQC_API generativeqc_status vibeqc_calculation_last_state(const generativeqc_calculation* calculation,
                                                                   double* density, uint64_t density_count,
                                                                   uint64_t* counts4);

#ifdef __cplusplus
}
#endif
#endif
