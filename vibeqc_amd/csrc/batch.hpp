// Ragged batch ("fleet") of independent closed-shell DF-RHF systems on one
// HIP device: the engine behind generativeqc_batch_prepare / _execute
// (reference generativeqc.h:1431-1586, src/scf/fleet.cpp).
//
// Systems are grouped into shape buckets (nbf, naux, nocc), as the reference
// groups its fleets (fleet.cpp:26-36), and executed largest first in waves of
// 4 x workers systems.  Each host worker owns one HIP stream, one rocBLAS
// handle and one device block cache (blocks recycled only among the worker's
// own buffers, so no block moves between streams), and the SCFs of a wave
// run concurrently on the GPU instead of one after another.  A plan lives only for its wave,
// which bounds device memory by the largest wave rather than the whole batch.
#pragma once

#include <exception>
#include <memory>
#include <vector>

#include "plan.hpp"

namespace vqc {

struct BatchOptions {
  PlanOptions plan;
  ScfOptions scf;
  Precision precision = Precision::FP64;
  bool warm_starts = false;
  bool unrestricted = false;  // UHF members (GENERATIVEQC_METHOD_UHF), else RHF
  int workers = 0;  // 0: VIBEQC_BATCH_WORKERS if set, else 8 (never more than the systems)
};

struct BatchItemOutcome {
  std::exception_ptr error;  // set when the item failed
  ScfOutcome scf;
  int bucket = 0;
  bool warm_start_used = false;
  bool warm_start_fallback = false;
};

// A retained converged density (the seed of the item's next execution: the
// restricted D, or [alpha; beta] for UHF members) with
// the coordinates it was converged at and its source diagnostics
// (reference generativeqc.h:1497-1529).  An empty density: no seed.
struct WarmState {
  std::vector<double> density, coordinates;
  double energy = 0.0, energy_change = 0.0, density_rms = 0.0;
  int iterations = 0;
};

class BatchPlan {
 public:
  // aux[i] is the auxiliary basis of systems[i] (same atoms).
  BatchPlan(int device, std::vector<System> systems, std::vector<System> aux, const BatchOptions& options);
  std::size_t size() const { return systems_.size(); }
  int bucket_of(std::size_t i) const { return bucket_[i]; }
  int bucket_count() const { return buckets_; }
  void set_warm_updates(bool enabled) { warm_updates_ = enabled; }
  bool warm_starts() const { return opts_.warm_starts; }
  bool unrestricted() const { return opts_.unrestricted; }
  int nbf_of(std::size_t i) const { return systems_[i].nbf(); }
  std::size_t atoms_of(std::size_t i) const { return systems_[i].atoms.size(); }
  const WarmState& warm_state(std::size_t i) const { return warm_[i]; }
  // Import seeds: states[i] non-null replaces item i's seed.  Every state is
  // validated (shape, finiteness, symmetry) before any is applied.
  void restore_warm_states(const std::vector<const WarmState*>& states);
  void clear_warm_states();
  // coordinates[i]: replacement flat xyz (Bohr) for system i, or null.
  // forces[i]: whether item i reports forces.
  std::vector<BatchItemOutcome> execute(const std::vector<const std::vector<double>*>& coordinates,
                                        const std::vector<bool>& forces);

 private:
  int device_;
  std::vector<System> systems_, aux_;
  BatchOptions opts_;
  std::vector<int> bucket_;
  int buckets_ = 0;
  std::vector<std::size_t> order_;  // most expensive first
  std::vector<WarmState> warm_;
  bool warm_updates_ = true;
};

}  // namespace vqc
