// extern "C" boundary: the generativeqc_* context / system / Fock-plan ABI
// (reference semantics: src/api/c_api_fock.cpp:22-310, src/api/error.cpp:12-35)
// and the vibeqc_* extensions declared in include/generativeqc/fock.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "basis.hpp"
#include "batch.hpp"
#include "common.hpp"
#include "generativeqc/batch.h"
#include "generativeqc/calculation.h"
#include "generativeqc/fock.h"
#include "generativeqc/generativeqc.h"
#include "noalloc.hpp"
#include "plan.hpp"
#include "tuning.hpp"

struct generativeqc_context {
  int device = 0;
  std::recursive_mutex mutex;
  std::string detail;
};

struct generativeqc_system {
  vqc::System data;
};

struct generativeqc_batch {
  std::unique_ptr<vqc::BatchPlan> plan;
  std::vector<std::size_t> natoms;
  std::vector<uint64_t> fock_builds;
  std::string detail;
};

struct generativeqc_calculation {
  generativeqc_context* context = nullptr;
  std::unique_ptr<vqc::DfPlan> plan;
  vqc::ScfOptions scf;
  bool unrestricted = false;
  std::size_t natom = 0;
  uint64_t budget = 0;
  std::vector<double> density;  // last successful execution
  uint64_t fock_builds = 0;
};

struct generativeqc_fock_plan {
  std::unique_ptr<vqc::DfPlan> plan;
  generativeqc_fock_spec requested{};
  generativeqc_fock_spec resolved{};
  double screening = 1e-12, threshold = 1e-10;
  uint64_t budget = 0;
  // iterations, reduced-precision iterations, purified densities, incremental FP64-stage builds,
  // single-term BF16 builds
  uint64_t solve_stats[5] = {0, 0, 0, 0, 0};
  std::string detail;
};

namespace {

// Exception -> status mapping (reference src/api/error.cpp:12-35).
generativeqc_status map_exception(std::string* detail) {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    if (detail) *detail = "out of host memory";
    return GENERATIVEQC_STATUS_OUT_OF_MEMORY;
  } catch (const vqc::OutOfMemory& e) {
    if (detail) *detail = e.what();
    return GENERATIVEQC_STATUS_OUT_OF_MEMORY;
  } catch (const vqc::NotImplemented& e) {
    if (detail) *detail = e.what();
    return GENERATIVEQC_STATUS_NOT_IMPLEMENTED;
  } catch (const vqc::DeviceError& e) {
    if (detail) *detail = e.what();
    return GENERATIVEQC_STATUS_CUDA_ERROR;
  } catch (const std::invalid_argument& e) {
    if (detail) *detail = e.what();
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  } catch (const std::exception& e) {
    if (detail) *detail = e.what();
    return GENERATIVEQC_STATUS_NUMERICAL_FAILURE;
  } catch (...) {
    if (detail) *detail = "unknown native failure";
    return GENERATIVEQC_STATUS_INTERNAL_ERROR;
  }
}

template <class T>
bool valid_descriptor(const T* d) {
  return d && d->struct_size == sizeof(T) && d->abi_version == GENERATIVEQC_ABI_VERSION;
}

void require(bool condition, const char* message) {
  if (!condition) throw std::invalid_argument(message);
}

std::size_t buffer_bytes(uint64_t count) {
  require(count <= std::numeric_limits<std::size_t>::max() / sizeof(double), "output size overflows address space");
  return static_cast<std::size_t>(count) * sizeof(double);
}

// Output aliasing would make transactional publication ambiguous
// (c_api_fock.cpp:104-118).
void disjoint_outputs(std::initializer_list<std::pair<const void*, std::size_t>> outputs) {
  std::vector<std::pair<std::uintptr_t, std::uintptr_t>> ranges;
  for (const auto& [ptr, bytes] : outputs) {
    if (!ptr) continue;
    const auto first = reinterpret_cast<std::uintptr_t>(ptr);
    require(bytes <= std::numeric_limits<std::uintptr_t>::max() - first, "output range overflows address space");
    ranges.emplace_back(first, first + bytes);
  }
  std::sort(ranges.begin(), ranges.end());
  for (std::size_t i = 1; i < ranges.size(); ++i)
    require(ranges[i - 1].second <= ranges[i].first, "Fock output buffers overlap");
}

void finite(const std::vector<double>& v) {
  for (double x : v)
    if (!std::isfinite(x)) throw std::runtime_error("nonfinite Fock output");
}

bool term_ok(const generativeqc_fock_term& t) {
  return t.present == 0 ||
         (t.op == GENERATIVEQC_FOCK_FULL_RANGE && t.approximation == GENERATIVEQC_FOCK_DENSITY_FITTED);
}

void copy_text(char* dst, std::size_t cap, const char* src) {
  std::strncpy(dst, src, cap - 1);
  dst[cap - 1] = '\0';
}

// RHF/DF options of a method descriptor (reference generativeqc.h:723-769).
vqc::BatchOptions batch_options(const generativeqc_method_descriptor& d, generativeqc_batch_flags flags) {
  if (d.method != GENERATIVEQC_METHOD_RHF && d.method != GENERATIVEQC_METHOD_UHF)
    throw vqc::NotImplemented("the RHF and UHF methods are implemented");
  if (d.density_fitting_mode != GENERATIVEQC_DENSITY_FITTING_CUDA &&
      d.density_fitting_mode != GENERATIVEQC_DENSITY_FITTING_AUTO)
    throw vqc::NotImplemented("density-fitted J/K on the HIP device only (exact J/K and the CPU backend are not implemented)");
  if (d.ks_options || d.initial_guess) throw vqc::NotImplemented("KS and initial-guess options are not implemented");
  require(d.precision_mode == GENERATIVEQC_PRECISION_FP64 || d.precision_mode == GENERATIVEQC_PRECISION_AUTO,
          "unknown precision mode");
  auto nonneg = [](double v) { return std::isfinite(v) && v >= 0.0; };
  require(nonneg(d.energy_tolerance) && nonneg(d.density_tolerance) && nonneg(d.screening_tolerance) &&
              nonneg(d.density_fitting_relative_threshold) && d.density_fitting_relative_threshold < 1.0,
          "method tolerances must be finite and nonnegative");
  vqc::BatchOptions o;
  o.plan.screening = d.screening_tolerance > 0.0 ? d.screening_tolerance : 1e-12;
  o.plan.metric_threshold = d.density_fitting_relative_threshold > 0.0 ? d.density_fitting_relative_threshold : 1e-10;
  if (d.max_iterations) o.scf.max_iterations = d.max_iterations;
  if (d.diis_history) o.scf.diis_history = std::min<uint32_t>(d.diis_history, 32);
  if (d.energy_tolerance > 0.0) o.scf.energy_tolerance = d.energy_tolerance;
  if (d.density_tolerance > 0.0) o.scf.density_tolerance = d.density_tolerance;
  o.precision = d.precision_mode == GENERATIVEQC_PRECISION_AUTO ? vqc::Precision::BF16X3 : vqc::Precision::FP64;
  o.warm_starts = (flags & GENERATIVEQC_BATCH_ENABLE_WARM_STARTS) != 0;
  o.unrestricted = d.method == GENERATIVEQC_METHOD_UHF;
  o.plan.ck = o.unrestricted ? -1.0 : -0.5;  // HF exchange: -1 per spin density, -1/2 of the restricted D
  return o;
}

// The template's shells placed on another system's atoms (same atom count).
vqc::System auxiliary_from_template(const vqc::System& tmpl, const vqc::System& sys) {
  if (tmpl.atoms.size() != sys.atoms.size())
    throw std::invalid_argument("auxiliary template atom topology differs from a batch system");
  vqc::System aux = tmpl;
  aux.atoms = sys.atoms;
  aux.charge = sys.charge;
  aux.multiplicity = sys.multiplicity;
  aux.electrons = sys.electrons;
  return aux;
}

generativeqc_status prepare_batch(generativeqc_context* context, const generativeqc_system* const* systems,
                                  const generativeqc_system* const* auxiliary, uint32_t count,
                                  const generativeqc_method_descriptor* descriptor, generativeqc_batch_flags flags,
                                  generativeqc_batch** batch) {
  if (batch) *batch = nullptr;
  if (!context || !systems || !descriptor || !batch || count == 0) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!valid_descriptor(descriptor)) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  std::lock_guard<std::recursive_mutex> lock(context->mutex);
  try {
    const vqc::BatchOptions opts = batch_options(*descriptor, flags);
    std::vector<vqc::System> orb, aux;
    orb.reserve(count);
    aux.reserve(count);
    for (uint32_t i = 0; i < count; ++i) {
      require(systems[i] != nullptr, "null batch system");
      orb.push_back(systems[i]->data);
      if (auxiliary) {
        require(auxiliary[i] != nullptr, "null auxiliary system");
        aux.push_back(auxiliary_from_template(auxiliary[i]->data, orb.back()));
      } else if (descriptor->density_fitting_auxiliary_basis) {
        aux.push_back(auxiliary_from_template(descriptor->density_fitting_auxiliary_basis->data, orb.back()));
      } else {
        aux.push_back(orb.back());
      }
    }
    auto out = std::make_unique<generativeqc_batch>();
    for (const auto& s : orb) out->natoms.push_back(s.atoms.size());
    out->fock_builds.assign(count, 0);
    out->plan = std::make_unique<vqc::BatchPlan>(context->device, std::move(orb), std::move(aux), opts);
    *batch = out.release();
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&context->detail);
  }
}

}  // namespace

extern "C" {

uint32_t generativeqc_get_abi_version(void) { return GENERATIVEQC_ABI_VERSION; }

const char* generativeqc_get_source_identity(void) {
  return "vibeqc-amd: MI355X (gfx950) HIP DF-JK / RHF engine, hand-written MD integrals + MFMA J/K";
}

const char* generativeqc_status_message(generativeqc_status status) {
  switch (status) {
    case GENERATIVEQC_STATUS_SUCCESS: return "success";
    case GENERATIVEQC_STATUS_INVALID_ARGUMENT: return "invalid argument";
    case GENERATIVEQC_STATUS_ABI_MISMATCH: return "ABI mismatch";
    case GENERATIVEQC_STATUS_NOT_IMPLEMENTED: return "not implemented";
    case GENERATIVEQC_STATUS_NOT_CONVERGED: return "not converged";
    case GENERATIVEQC_STATUS_NUMERICAL_FAILURE: return "numerical failure";
    case GENERATIVEQC_STATUS_CUDA_ERROR: return "device (HIP) error";
    case GENERATIVEQC_STATUS_OUT_OF_MEMORY: return "out of memory";
    case GENERATIVEQC_STATUS_INTERNAL_ERROR: return "internal error";
    case GENERATIVEQC_STATUS_PRECISION_UNAVAILABLE: return "precision provenance unavailable";
    default: return "unknown status";
  }
}

generativeqc_status generativeqc_context_create(const generativeqc_context_descriptor* descriptor,
                                                generativeqc_context** context) {
  if (context) *context = nullptr;
  if (!descriptor || !context) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!valid_descriptor(descriptor)) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  if (descriptor->backend != GENERATIVEQC_BACKEND_CUDA) return GENERATIVEQC_STATUS_NOT_IMPLEMENTED;
  try {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return GENERATIVEQC_STATUS_CUDA_ERROR;
    if (descriptor->device_id < 0 || descriptor->device_id >= count) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    VQC_HIP(hipSetDevice(descriptor->device_id));
    auto c = std::make_unique<generativeqc_context>();
    c->device = descriptor->device_id;
    *context = c.release();
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(nullptr);
  }
}

void generativeqc_context_destroy(generativeqc_context* context) { delete context; }

const char* generativeqc_context_get_last_detail(const generativeqc_context* context) {
  return context ? context->detail.c_str() : "null context";
}

const char* generativeqc_context_last_error(const generativeqc_context* context) {
  return generativeqc_context_get_last_detail(context);
}

generativeqc_status generativeqc_system_create(generativeqc_context* context,
                                               const generativeqc_system_descriptor* descriptor,
                                               generativeqc_system** system) {
  if (system) *system = nullptr;
  if (!context || !descriptor || !system) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!valid_descriptor(descriptor)) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  std::lock_guard<std::recursive_mutex> lock(context->mutex);
  try {
    auto s = std::make_unique<generativeqc_system>();
    const generativeqc_status st = vqc::make_system(*descriptor, s->data, context->detail);
    if (st != GENERATIVEQC_STATUS_SUCCESS) return st;
    *system = s.release();
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&context->detail);
  }
}

void generativeqc_system_destroy(generativeqc_system* system) { delete system; }

generativeqc_status generativeqc_fock_plan_create(generativeqc_context* context, const generativeqc_system* system,
                                                  const generativeqc_system* auxiliary,
                                                  const generativeqc_fock_spec* spec,
                                                  const generativeqc_fock_controls* controls,
                                                  generativeqc_fock_plan** output) {
  if (output) *output = nullptr;
  if (!context || !system || !spec || !output) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!valid_descriptor(spec) || (controls && !valid_descriptor(controls))) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  std::lock_guard<std::recursive_mutex> lock(context->mutex);
  try {
    require(spec->coulomb.present == 0 || spec->coulomb.present == 1, "Fock presence must be zero or one");
    require(spec->exchange.present == 0 || spec->exchange.present == 1, "Fock presence must be zero or one");
    require(std::isfinite(spec->coulomb.coefficient) && std::isfinite(spec->exchange.coefficient),
            "Fock coefficients must be finite");
    const double screening = controls ? controls->screening_tolerance : 1e-12;
    const double requested = controls ? controls->metric_relative_threshold : 0.0;
    require(std::isfinite(screening) && screening >= 0.0, "Fock screening tolerance must be finite and nonnegative");
    require(std::isfinite(requested) && requested >= 0.0 && requested < 1.0,
            "Fock metric cutoff must be finite and in [0, 1)");
    require(spec->spin == GENERATIVEQC_FOCK_RESTRICTED || spec->spin == GENERATIVEQC_FOCK_UNRESTRICTED,
            "unknown Fock spin");
    if (spec->derivative_order > 1)
      throw vqc::NotImplemented("only first nuclear derivatives are implemented");
    if (!term_ok(spec->coulomb) || !term_ok(spec->exchange))
      throw vqc::NotImplemented("only full-range density-fitted J/K terms are implemented");
    if (!spec->coulomb.present && !spec->exchange.present)
      throw std::invalid_argument("a Fock plan requires at least one J or K term");
    vqc::PlanOptions opts;
    opts.screening = screening;
    opts.device_budget = controls ? controls->device_budget_bytes : 0;
    opts.metric_threshold = requested == 0.0 ? 1e-10 : requested;
    opts.coulomb = spec->coulomb.present != 0;
    opts.exchange = spec->exchange.present != 0;
    opts.cj = opts.coulomb ? spec->coulomb.coefficient : 0.0;
    opts.ck = opts.exchange ? spec->exchange.coefficient : 0.0;
    auto plan = std::make_unique<generativeqc_fock_plan>();
    plan->requested = *spec;
    plan->resolved = *spec;
    plan->screening = opts.screening;
    plan->threshold = opts.metric_threshold;
    plan->budget = controls ? controls->device_budget_bytes : 0;
    const vqc::System& aux = auxiliary ? auxiliary->data : system->data;
    plan->plan = std::make_unique<vqc::DfPlan>(context->device, system->data, aux, opts);
    *output = plan.release();
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&context->detail);
  }
}

void generativeqc_fock_plan_destroy(generativeqc_fock_plan* plan) { delete plan; }

const char* generativeqc_fock_plan_last_error(const generativeqc_fock_plan* plan) {
  return plan ? plan->detail.c_str() : "null Fock plan";
}

generativeqc_status generativeqc_fock_plan_evaluate(generativeqc_fock_plan* plan, const double* density,
                                                    uint64_t density_count, const double* beta, uint64_t beta_count,
                                                    generativeqc_fock_result* result) {
  if (!plan || !density || !result) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!valid_descriptor(result)) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  plan->detail.clear();
  try {
    auto& p = *plan->plan;
    const bool unrestricted = plan->resolved.spin == GENERATIVEQC_FOCK_UNRESTRICTED;
    const uint64_t matrix = static_cast<uint64_t>(p.nbf()) * p.nbf();
    require(density_count == matrix && result->matrix_count == matrix &&
                (unrestricted ? beta != nullptr && beta_count == matrix : beta == nullptr && beta_count == 0),
            "Fock density/output shape or spin mismatch");
    const uint64_t coordinates = static_cast<uint64_t>(p.system().atoms.size()) * 3;
    require(result->gradient ? plan->resolved.derivative_order == 1 && result->gradient_count == coordinates
                             : result->gradient_count == 0,
            "Fock gradient capability/count mismatch");
    require(unrestricted || (result->exchange_beta == nullptr && result->fock_beta == nullptr),
            "restricted plans have no beta outputs");
    const auto bytes = buffer_bytes(matrix);
    disjoint_outputs({{result->coulomb, bytes},
                      {result->exchange_alpha, bytes},
                      {result->exchange_beta, bytes},
                      {result->fock_alpha, bytes},
                      {result->fock_beta, bytes},
                      {result->gradient, buffer_bytes(result->gradient_count)},
                      {result, sizeof(*result)}});
    const std::vector<double> d(density, density + matrix);
    const std::vector<double> db = unrestricted ? std::vector<double>(beta, beta + matrix) : std::vector<double>{};
    std::vector<double> gradient;
    vqc::UhfJk jk;
    if (unrestricted) {
      jk = p.evaluate_unrestricted(d, db, result->gradient ? &gradient : nullptr);
    } else {
      vqc::JkResult r = p.evaluate(d, result->gradient ? &gradient : nullptr);
      jk.J = std::move(r.J);
      jk.Ka = std::move(r.K);
    }
    const auto& o = p.options();
    const std::vector<double> h = p.hcore();
    std::vector<double> fock(matrix), fock_b(unrestricted ? matrix : 0);
    double e1 = 0.0, ej = 0.0, ek = 0.0, cj = 0.0, ck = 0.0;
    // Compensated sums as in fock_build.cpp:460-491.
    auto kahan = [](double& sum, double& comp, double v) {
      const double y = v - comp;
      const double t = sum + y;
      comp = (t - sum) - y;
      sum = t;
    };
    double c1 = 0.0;
    for (uint64_t i = 0; i < matrix; ++i) {
      const double total = d[i] + (unrestricted ? db[i] : 0.0);
      const double jpart = o.coulomb ? o.cj * jk.J[i] : 0.0;
      fock[i] = h[i] + jpart + (o.exchange ? o.ck * jk.Ka[i] : 0.0);
      if (unrestricted) fock_b[i] = h[i] + jpart + (o.exchange ? o.ck * jk.Kb[i] : 0.0);
      kahan(e1, c1, total * h[i]);
      if (o.coulomb) kahan(ej, cj, 0.5 * total * o.cj * jk.J[i]);
      if (o.exchange) {
        kahan(ek, ck, 0.5 * d[i] * o.ck * jk.Ka[i]);
        if (unrestricted) kahan(ek, ck, 0.5 * db[i] * o.ck * jk.Kb[i]);
      }
    }
    const double e2 = ej + ek;
    if (!std::isfinite(e1) || !std::isfinite(e2)) throw std::runtime_error("nonfinite Fock energy");
    finite(jk.J);
    finite(jk.Ka);
    finite(jk.Kb);
    finite(fock);
    finite(fock_b);
    if (result->coulomb && o.coulomb) std::copy(jk.J.begin(), jk.J.end(), result->coulomb);
    if (result->exchange_alpha && o.exchange) std::copy(jk.Ka.begin(), jk.Ka.end(), result->exchange_alpha);
    if (unrestricted && result->exchange_beta && o.exchange)
      std::copy(jk.Kb.begin(), jk.Kb.end(), result->exchange_beta);
    if (result->gradient) {
      finite(gradient);
      std::copy(gradient.begin(), gradient.end(), result->gradient);
    }
    if (result->fock_alpha) std::copy(fock.begin(), fock.end(), result->fock_alpha);
    if (unrestricted && result->fock_beta) std::copy(fock_b.begin(), fock_b.end(), result->fock_beta);
    result->energy_one_electron = e1;
    result->energy_two_electron = e2;
    result->nuclear_repulsion = p.nuclear_repulsion();
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&plan->detail);
  }
}

generativeqc_status generativeqc_fock_plan_solve(generativeqc_fock_plan* plan,
                                                 const generativeqc_fock_scf_controls* controls,
                                                 const double* initial_density, uint64_t initial_density_count,
                                                 generativeqc_fock_scf_result* output) {
  if (!plan || !output) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!valid_descriptor(output) || (controls && !valid_descriptor(controls))) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  plan->detail.clear();
  try {
    auto& p = *plan->plan;
    const bool unrestricted = plan->resolved.spin == GENERATIVEQC_FOCK_UNRESTRICTED;
    const uint64_t count = static_cast<uint64_t>(p.nbf()) * p.nbf() * (unrestricted ? 2 : 1);
    require(output->density ? output->density_count == count : output->density_count == 0,
            "SCF density output shape mismatch");
    const uint64_t coordinates = static_cast<uint64_t>(p.system().atoms.size()) * 3;
    require(output->forces ? output->force_count == coordinates && plan->resolved.derivative_order == 1
                           : output->force_count == 0,
            "SCF force output shape or derivative capability mismatch");
    require(initial_density ? initial_density_count == count : initial_density_count == 0,
            "SCF initial density shape mismatch");
    disjoint_outputs({{output->density, buffer_bytes(output->density_count)},
                      {output->forces, buffer_bytes(output->force_count)},
                      {output, sizeof(*output)}});
    vqc::ScfOptions opts;
    if (controls) {
      require(controls->max_iterations > 0 && controls->diis_history > 0 &&
                  std::isfinite(controls->energy_tolerance) && controls->energy_tolerance > 0 &&
                  std::isfinite(controls->density_tolerance) && controls->density_tolerance > 0,
              "SCF counts and finite tolerances must be positive");
      opts.max_iterations = controls->max_iterations;
      opts.diis_history = std::min<uint32_t>(controls->diis_history, 32);
      opts.energy_tolerance = controls->energy_tolerance;
      opts.density_tolerance = controls->density_tolerance;
    }
    std::vector<double> seed;
    if (initial_density) {
      seed.assign(initial_density, initial_density + count);
      finite(seed);
    }
    const vqc::ScfOutcome r =
        unrestricted ? p.solve_unrestricted(opts, initial_density ? &seed : nullptr, output->forces != nullptr)
                     : p.solve(opts, initial_density ? &seed : nullptr, output->forces != nullptr);
    plan->solve_stats[0] = r.iterations;
    plan->solve_stats[1] = r.reduced_iterations;
    plan->solve_stats[2] = r.purified_iterations;
    plan->solve_stats[3] = r.incremental_builds;
    plan->solve_stats[4] = r.single_term_builds;
    if (!r.converged) {
      plan->detail = "Fock SCF did not converge in " + std::to_string(r.iterations) + " iterations";
      return GENERATIVEQC_STATUS_NOT_CONVERGED;
    }
    finite(r.density);
    finite(r.forces);
    if (output->density) std::copy(r.density.begin(), r.density.end(), output->density);
    if (output->forces) std::copy(r.forces.begin(), r.forces.end(), output->forces);
    output->energy = r.energy;
    output->energy_change = r.energy_change;
    output->density_rms = r.density_rms;
    output->iterations = r.iterations;
    output->initial_density_used = initial_density != nullptr;
    output->fock_builds = r.fock_builds;
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&plan->detail);
  }
}

generativeqc_status generativeqc_fock_plan_diagnostic(const generativeqc_fock_plan* plan,
                                                      generativeqc_fock_diagnostic* diagnostic) {
  if (!plan || !diagnostic) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!valid_descriptor(diagnostic)) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  const auto& p = *plan->plan;
  generativeqc_fock_diagnostic d{};
  d.struct_size = sizeof(d);
  d.abi_version = GENERATIVEQC_ABI_VERSION;
  d.requested = plan->requested;
  d.resolved = plan->resolved;
  d.backend = GENERATIVEQC_BACKEND_CUDA;
  d.resolved_schedule = GENERATIVEQC_FOCK_CUDA_INDEPENDENT;
  d.source_schedule = GENERATIVEQC_FOCK_CUDA_INDEPENDENT;
  d.nbf = static_cast<uint64_t>(p.nbf());
  d.coordinate_count = 3 * p.system().atoms.size();
  d.device_bytes = p.device_bytes();
  d.device_budget_bytes = plan->budget;
  d.auxiliary_rank = static_cast<uint64_t>(p.rank());
  d.auxiliary_tile = static_cast<uint64_t>(p.naux());
  d.screening_tolerance = plan->screening;
  d.metric_relative_threshold = plan->threshold;
  d.df_streamed = 0;
  copy_text(d.direct_schedule, sizeof(d.direct_schedule), "none");
  copy_text(d.df_value_backend, sizeof(d.df_value_backend), "hip-gfx950");
  copy_text(d.df_value_mapping, sizeof(d.df_value_mapping),
            p.cholesky() ? "cholesky-packed" : "eigen-cutoff-packed");
  copy_text(d.df_response_mapping, sizeof(d.df_response_mapping),
            plan->resolved.derivative_order ? "shell-pair-deterministic" : "absent");
  copy_text(d.one_electron_value_backend, sizeof(d.one_electron_value_backend), "hip-gfx950");
  copy_text(d.one_electron_value_mapping, sizeof(d.one_electron_value_mapping), "shell-pair");
  copy_text(d.one_electron_response_mapping, sizeof(d.one_electron_response_mapping),
            plan->resolved.derivative_order ? "pair-thread" : "absent");
  *diagnostic = d;
  return GENERATIVEQC_STATUS_SUCCESS;
}

generativeqc_status vibeqc_fock_plan_set_precision(generativeqc_fock_plan* plan, vibeqc_precision precision) {
  if (!plan) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  plan->detail.clear();
  try {
    require(precision == VIBEQC_PRECISION_FP64 || precision == VIBEQC_PRECISION_FP32 ||
                precision == VIBEQC_PRECISION_BF16X3,
            "unknown precision mode");
    plan->plan->set_precision(static_cast<vqc::Precision>(precision));
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&plan->detail);
  }
}

generativeqc_status vibeqc_fock_plan_evaluate_occupied(generativeqc_fock_plan* plan, const double* occupied,
                                                       uint64_t nbf, uint64_t nocc, double* coulomb,
                                                       double* exchange) {
  if (!plan || !occupied) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  plan->detail.clear();
  try {
    auto& p = *plan->plan;
    require(nbf == static_cast<uint64_t>(p.nbf()) && nocc <= nbf, "occupied factor shape mismatch");
    const uint64_t matrix = nbf * nbf;
    disjoint_outputs({{coulomb, buffer_bytes(matrix)}, {exchange, buffer_bytes(matrix)}});
    const std::vector<double> c(occupied, occupied + nbf * nocc);
    const vqc::JkResult jk = p.evaluate_occupied(c, static_cast<int>(nocc));
    finite(jk.J);
    finite(jk.K);
    if (coulomb) std::copy(jk.J.begin(), jk.J.end(), coulomb);
    if (exchange) std::copy(jk.K.begin(), jk.K.end(), exchange);
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&plan->detail);
  }
}

generativeqc_status vibeqc_fock_plan_evaluate_occupied_noalloc(generativeqc_fock_plan* plan,
                                                               const double* occupied_device, uint64_t nbf,
                                                               uint64_t nocc, double* coulomb_device,
                                                               double* exchange_device) {
  if (!plan || !occupied_device || !coulomb_device || !exchange_device) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  plan->detail.clear();
  try {
    auto& p = *plan->plan;
    if (nbf != static_cast<uint64_t>(p.nbf()) || nocc == 0 || nocc > nbf) {
      plan->detail = "occupied factor shape mismatch";
      return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    }
    if (coulomb_device == exchange_device) {
      plan->detail = "Fock output buffers overlap";
      return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    }
    p.evaluate_occupied_into(occupied_device, static_cast<int>(nocc), coulomb_device, exchange_device,
                             vqc::noalloc_scratch(p.device()), vqc::kNoAllocScratchBytes);
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (const std::length_error& e) {
    plan->detail = e.what();
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  } catch (...) {
    return map_exception(&plan->detail);
  }
}

generativeqc_status vibeqc_fock_plan_benchmark_occupied(generativeqc_fock_plan* plan, const double* occupied,
                                                        uint64_t nbf, uint64_t nocc, uint32_t repeat,
                                                        double* kernel_ms) {
  if (!plan || !occupied) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  plan->detail.clear();
  try {
    auto& p = *plan->plan;
    require(nbf == static_cast<uint64_t>(p.nbf()) && nocc > 0 && nocc <= nbf, "occupied factor shape mismatch");
    const std::vector<double> c(occupied, occupied + nbf * nocc);
    const auto ms = p.benchmark_occupied(c, static_cast<int>(nocc), repeat);
    if (kernel_ms) std::copy(ms.begin(), ms.end(), kernel_ms);
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&plan->detail);
  }
}

generativeqc_status vibeqc_fock_plan_one_electron(const generativeqc_fock_plan* plan, double* overlap, double* hcore,
                                                  uint64_t matrix_count) {
  if (!plan) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  try {
    const auto& p = *plan->plan;
    const uint64_t matrix = static_cast<uint64_t>(p.nbf()) * p.nbf();
    require(matrix_count == matrix, "one-electron output shape mismatch");
    disjoint_outputs({{overlap, buffer_bytes(matrix)}, {hcore, buffer_bytes(matrix)}});
    if (overlap) {
      const auto s = p.overlap();
      std::copy(s.begin(), s.end(), overlap);
    }
    if (hcore) {
      const auto h = p.hcore();
      std::copy(h.begin(), h.end(), hcore);
    }
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(nullptr);
  }
}

generativeqc_status vibeqc_fock_plan_shape(const generativeqc_fock_plan* plan, uint64_t* shape6) {
  if (!plan || !shape6) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  const auto& p = *plan->plan;
  shape6[0] = static_cast<uint64_t>(p.nbf());
  shape6[1] = static_cast<uint64_t>(p.naux());
  shape6[2] = static_cast<uint64_t>(p.npair());
  shape6[3] = static_cast<uint64_t>(p.rank());
  shape6[4] = p.cholesky() ? 1 : 0;
  shape6[5] = static_cast<uint64_t>(p.device_bytes());
  return GENERATIVEQC_STATUS_SUCCESS;
}

generativeqc_status vibeqc_fock_plan_streaming(const generativeqc_fock_plan* plan, uint64_t* out3) {
  if (!plan || !out3) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  const auto& p = *plan->plan;
  out3[0] = p.streamed() ? 1 : 0;
  out3[1] = static_cast<uint64_t>(p.auxiliary_block());
  out3[2] = static_cast<uint64_t>(p.row_tiles());
  return GENERATIVEQC_STATUS_SUCCESS;
}

generativeqc_status vibeqc_fock_plan_last_solve_stats(const generativeqc_fock_plan* plan, uint64_t* stats,
                                                      uint64_t count) {
  if (!plan || !stats || count == 0) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  std::copy(plan->solve_stats, plan->solve_stats + std::min<uint64_t>(count, 5), stats);
  return GENERATIVEQC_STATUS_SUCCESS;
}

generativeqc_status vibeqc_fock_plan_autotune(generativeqc_fock_plan* plan, const double* occupied, uint64_t nbf,
                                              uint64_t nocc, uint32_t repeat, int32_t* shape2, double* ms2) {
  if (!plan || !occupied) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  plan->detail.clear();
  try {
    auto& p = *plan->plan;
    require(nbf == static_cast<uint64_t>(p.nbf()) && nocc > 0 && nocc <= nbf, "occupied factor shape mismatch");
    const std::vector<double> c(occupied, occupied + nbf * nocc);
    const vqc::AutotuneResult r = p.autotune(c, static_cast<int>(nocc), repeat);
    if (shape2) {
      shape2[0] = r.u_tile;
      shape2[1] = r.syrk_splits;
    }
    if (ms2) {
      ms2[0] = r.best_ms;
      ms2[1] = r.default_ms;
    }
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&plan->detail);
  }
}

generativeqc_status vibeqc_fock_plan_launch_shape(const generativeqc_fock_plan* plan, uint64_t nocc, int32_t* shape2) {
  if (!plan || !shape2 || nocc == 0) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  try {
    const auto& p = *plan->plan;
    if (nocc > static_cast<uint64_t>(p.nbf())) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    const int r = static_cast<int>(nocc);
    const vqc::JkShape s = p.shape_for(p.precision(), r);
    shape2[0] = s.u_tile > 0 ? s.u_tile : vqc::occupied_tile(r);
    shape2[1] = s.syrk_splits > 0 ? s.syrk_splits : vqc::exchange_default_splits(p.precision(), p.nbf());
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(nullptr);
  }
}

generativeqc_status vibeqc_tuning_record(const char* arch, int32_t precision, uint64_t nbf, uint64_t occupied,
                                         uint64_t ldq, int32_t u_tile, int32_t syrk_splits, double ms) {
  if (!arch || nbf > static_cast<uint64_t>(std::numeric_limits<int>::max()) || occupied > nbf ||
      ldq > static_cast<uint64_t>(std::numeric_limits<int>::max()) || !std::isfinite(ms))
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  try {
    vqc::TuningEntry e;
    e.arch = arch;
    e.precision = precision;
    e.n = static_cast<int>(nbf);
    e.r = static_cast<int>(occupied);
    e.ldq = static_cast<int>(ldq);
    e.shape = vqc::JkShape{u_tile, syrk_splits};
    e.ms = ms;
    vqc::record_shape(e);
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (const std::invalid_argument&) {
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  } catch (const std::runtime_error&) {
    return GENERATIVEQC_STATUS_INTERNAL_ERROR;  // VIBEQC_TUNING_FILE could not be written
  } catch (...) {
    return map_exception(nullptr);
  }
}

generativeqc_status vibeqc_tuning_lookup(const char* arch, int32_t precision, uint64_t nbf, uint64_t occupied,
                                         uint64_t ldq, int32_t* shape2, int32_t* found) {
  if (!arch || !shape2 || !found || nbf > static_cast<uint64_t>(std::numeric_limits<int>::max()) ||
      occupied > nbf || ldq > static_cast<uint64_t>(std::numeric_limits<int>::max()))
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  try {
    vqc::JkShape s;
    *found = vqc::tuned_shape(arch, precision, static_cast<int>(nbf), static_cast<int>(occupied), static_cast<int>(ldq), s)
                 ? 1
                 : 0;
    shape2[0] = *found ? s.u_tile : 0;
    shape2[1] = *found ? s.syrk_splits : 0;
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(nullptr);
  }
}

generativeqc_status vibeqc_tuning_count(uint64_t* count) {
  if (!count) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  try {
    *count = vqc::tuned_shapes().size();
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(nullptr);
  }
}

generativeqc_status vibeqc_tuning_load(const char* path, int32_t* found) {
  if (!path || !found) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  try {
    *found = vqc::load_tuning(path) ? 1 : 0;
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(nullptr);
  }
}

generativeqc_status vibeqc_tuning_save(const char* path) {
  if (!path) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  try {
    vqc::save_tuning(path);
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (const std::invalid_argument&) {
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  } catch (const std::runtime_error&) {
    return GENERATIVEQC_STATUS_INTERNAL_ERROR;  // the file could not be written
  } catch (...) {
    return map_exception(nullptr);
  }
}

generativeqc_status vibeqc_tuning_clear(void) {
  vqc::clear_tuning();
  return GENERATIVEQC_STATUS_SUCCESS;
}

}  // extern "C"

extern "C" {

generativeqc_status generativeqc_batch_prepare(generativeqc_context* context, const generativeqc_system* const* systems,
                                               uint32_t system_count, const generativeqc_method_descriptor* descriptor,
                                               generativeqc_batch_flags flags, generativeqc_batch** batch) {
  return prepare_batch(context, systems, nullptr, system_count, descriptor, flags, batch);
}

generativeqc_status vibeqc_batch_prepare(generativeqc_context* context, const generativeqc_system* const* systems,
                                         const generativeqc_system* const* auxiliary_systems, uint32_t system_count,
                                         const generativeqc_method_descriptor* descriptor,
                                         generativeqc_batch_flags flags, generativeqc_batch** batch) {
  return prepare_batch(context, systems, auxiliary_systems, system_count, descriptor, flags, batch);
}

void generativeqc_batch_destroy(generativeqc_batch* batch) { delete batch; }

uint32_t generativeqc_batch_get_system_count(const generativeqc_batch* batch) {
  return batch ? static_cast<uint32_t>(batch->plan->size()) : 0u;
}

generativeqc_status generativeqc_batch_set_warm_start_updates(generativeqc_batch* batch, int32_t enabled) {
  if (!batch || (enabled != 0 && enabled != 1)) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  batch->plan->set_warm_updates(enabled != 0);
  return GENERATIVEQC_STATUS_SUCCESS;
}

generativeqc_status generativeqc_batch_get_last_fock_builds(const generativeqc_batch* batch, uint32_t index,
                                                            uint64_t* builds) {
  if (!batch || !builds || index >= batch->fock_builds.size()) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  *builds = batch->fock_builds[index];
  return GENERATIVEQC_STATUS_SUCCESS;
}

generativeqc_status generativeqc_batch_get_hf_warm_state(const generativeqc_batch* batch, uint32_t index,
                                                         generativeqc_hf_warm_state* state) {
  if (!batch || !state || index >= batch->plan->size()) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!valid_descriptor(state)) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  const vqc::WarmState& w = batch->plan->warm_state(index);
  const bool present = !w.density.empty();
  if (state->density || state->coordinates) {
    if (!present || !state->density || !state->coordinates || state->density_count != w.density.size() ||
        state->coordinate_count != w.coordinates.size())
      return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    std::copy(w.density.begin(), w.density.end(), state->density);
    std::copy(w.coordinates.begin(), w.coordinates.end(), state->coordinates);
  } else {
    state->density_count = present ? w.density.size() : 0;
    state->coordinate_count = present ? w.coordinates.size() : 0;
  }
  state->energy = present ? w.energy : 0.0;
  state->energy_change = present ? w.energy_change : 0.0;
  state->density_rms = present ? w.density_rms : 0.0;
  state->iterations = present ? w.iterations : 0;
  state->present = present ? 1 : 0;
  return GENERATIVEQC_STATUS_SUCCESS;
}

generativeqc_status generativeqc_batch_restore_hf_warm_states(generativeqc_batch* batch,
                                                              const generativeqc_hf_warm_state* states,
                                                              uint32_t count) {
  if (!batch || !states) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  batch->detail.clear();
  try {
    require(count == batch->plan->size(), "warm state count mismatch");
    std::vector<vqc::WarmState> owned(count);
    std::vector<const vqc::WarmState*> view(count, nullptr);
    for (uint32_t i = 0; i < count; ++i) {
      if (!valid_descriptor(&states[i])) return GENERATIVEQC_STATUS_ABI_MISMATCH;
      const auto& s = states[i];
      require(s.present == 0 || s.present == 1, "warm state presence must be zero or one");
      if (!s.present) continue;
      require(s.density && s.coordinates, "present warm states need both arrays");
      owned[i].density.assign(s.density, s.density + s.density_count);
      owned[i].coordinates.assign(s.coordinates, s.coordinates + s.coordinate_count);
      owned[i].energy = s.energy;
      owned[i].energy_change = s.energy_change;
      owned[i].density_rms = s.density_rms;
      owned[i].iterations = s.iterations;
      view[i] = &owned[i];
    }
    batch->plan->restore_warm_states(view);
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&batch->detail);
  }
}

generativeqc_status generativeqc_batch_clear_warm_starts(generativeqc_batch* batch) {
  if (!batch) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  batch->plan->clear_warm_states();
  return GENERATIVEQC_STATUS_SUCCESS;
}

generativeqc_status vibeqc_batch_bucket(const generativeqc_batch* batch, uint32_t index, uint32_t* bucket_id) {
  if (!batch || !bucket_id || index >= batch->plan->size()) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  *bucket_id = static_cast<uint32_t>(batch->plan->bucket_of(index));
  return GENERATIVEQC_STATUS_SUCCESS;
}

generativeqc_status generativeqc_batch_execute(generativeqc_batch* batch, const generativeqc_batch_input_descriptor* inputs,
                                               uint32_t input_count, generativeqc_batch_item_result_descriptor* results,
                                               uint32_t result_count) {
  if (!batch || !results) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  const std::size_t count = batch->plan->size();
  batch->detail.clear();
  try {
    require(result_count == count, "batch result count mismatch");
    require(inputs ? input_count == count : input_count == 0, "batch input count mismatch");
    for (std::size_t i = 0; i < count; ++i) {
      if (!valid_descriptor(&results[i]) || (inputs && !valid_descriptor(&inputs[i])))
        return GENERATIVEQC_STATUS_ABI_MISMATCH;
      const uint64_t coords = batch->natoms[i] * 3;
      require(results[i].forces ? results[i].force_count == coords : results[i].force_count == 0,
              "batch force output shape mismatch");
      if (inputs) require(inputs[i].coordinates ? inputs[i].coordinate_count == coords : inputs[i].coordinate_count == 0,
                          "batch coordinate shape mismatch");
    }
    std::vector<std::vector<double>> xyz(inputs ? count : 0);
    std::vector<const std::vector<double>*> coordinates(inputs ? count : 0, nullptr);
    for (std::size_t i = 0; inputs && i < count; ++i)
      if (inputs[i].coordinates) {
        xyz[i].assign(inputs[i].coordinates, inputs[i].coordinates + inputs[i].coordinate_count);
        coordinates[i] = &xyz[i];
      }
    std::vector<bool> forces(count);
    for (std::size_t i = 0; i < count; ++i) forces[i] = results[i].forces != nullptr;
    const auto outcomes = batch->plan->execute(coordinates, forces);
    for (std::size_t i = 0; i < count; ++i) {
      auto& r = results[i];
      const auto& o = outcomes[i];
      r.bucket_id = static_cast<uint32_t>(o.bucket);
      r.executed_backend = GENERATIVEQC_BACKEND_CUDA;
      r.warm_start_used = o.warm_start_used;
      r.warm_start_fallback = o.warm_start_fallback;
      batch->fock_builds[i] = o.scf.fock_builds;
      if (o.error) {
        std::string what;
        try {
          std::rethrow_exception(o.error);
        } catch (...) {
          r.status = map_exception(&what);
        }
        // VIBEQC_TRACE set: the failed item's detail on stderr.
        if (std::getenv("VIBEQC_TRACE"))
          std::fprintf(stderr, "{\"trace\": \"batch_item_error\", \"index\": %zu, \"status\": %d, \"detail\": \"%s\"}\n", i,
                       static_cast<int>(r.status), what.c_str());
        r.converged = 0;
        continue;
      }
      r.iterations = o.scf.iterations;
      r.energy_change = o.scf.energy_change;
      r.density_rms = o.scf.density_rms;
      r.converged = o.scf.converged ? 1 : 0;
      if (!o.scf.converged) {
        r.status = GENERATIVEQC_STATUS_NOT_CONVERGED;
        continue;
      }
      r.status = GENERATIVEQC_STATUS_SUCCESS;
      r.energy = o.scf.energy;
      if (r.forces) std::copy(o.scf.forces.begin(), o.scf.forces.end(), r.forces);
    }
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&batch->detail);
  }
}

}  // extern "C"

extern "C" {

generativeqc_status generativeqc_calculation_prepare(generativeqc_context* context, const generativeqc_system* system,
                                                     const generativeqc_method_descriptor* descriptor,
                                                     generativeqc_calculation** calculation) {
  if (!context || !system || !descriptor || !calculation) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  *calculation = nullptr;
  if (!valid_descriptor(descriptor)) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  std::lock_guard<std::recursive_mutex> lock(context->mutex);
  try {
    // The batch's descriptor rules (methods, fitting modes, defaults).
    const vqc::BatchOptions opts = batch_options(*descriptor, 0);
    auto out = std::make_unique<generativeqc_calculation>();
    out->context = context;
    out->scf = opts.scf;
    out->unrestricted = opts.unrestricted;
    out->natom = system->data.atoms.size();
    out->budget = descriptor->density_fitting_memory_budget_bytes;
    const vqc::System aux = descriptor->density_fitting_auxiliary_basis
                                ? auxiliary_from_template(descriptor->density_fitting_auxiliary_basis->data, system->data)
                                : system->data;
    vqc::PlanOptions plan_options = opts.plan;
    plan_options.device_budget = out->budget;
    out->plan = std::make_unique<vqc::DfPlan>(context->device, system->data, aux, plan_options);
    out->plan->set_precision(opts.precision);
    *calculation = out.release();
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&context->detail);
  }
}

void generativeqc_calculation_destroy(generativeqc_calculation* calculation) { delete calculation; }

generativeqc_status generativeqc_calculation_execute(generativeqc_calculation* calculation,
                                                     generativeqc_result_descriptor* result) {
  if (!calculation) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  generativeqc_context* context = calculation->context;
  std::lock_guard<std::recursive_mutex> lock(context->mutex);
  calculation->density.clear();
  calculation->fock_builds = 0;
  if (!result) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!valid_descriptor(result)) return GENERATIVEQC_STATUS_ABI_MISMATCH;
  const bool omit_forces = result->forces == nullptr && result->force_count == 0;
  if (!result->forces && !omit_forces) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  if (!omit_forces && result->force_count < calculation->natom * 3) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  try {
    vqc::DfPlan& plan = *calculation->plan;
    const vqc::ScfOutcome o = calculation->unrestricted ? plan.solve_unrestricted(calculation->scf, nullptr, !omit_forces)
                                                        : plan.solve(calculation->scf, nullptr, !omit_forces);
    result->energy = o.energy;
    result->iterations = o.iterations;
    result->energy_change = o.energy_change;
    result->density_rms = o.density_rms;
    result->converged = o.converged ? 1 : 0;
    result->executed_backend = GENERATIVEQC_BACKEND_CUDA;
    if (!o.converged) {
      context->detail = "SCF did not converge";
      return GENERATIVEQC_STATUS_NOT_CONVERGED;
    }
    if (!omit_forces) {
      if (o.forces.size() > result->force_count) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
      std::copy(o.forces.begin(), o.forces.end(), result->forces);
    }
    calculation->density = o.density;
    calculation->fock_builds = o.fock_builds;
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&context->detail);
  }
}

generativeqc_status vibeqc_calculation_set_precision(generativeqc_calculation* calculation, int32_t precision) {
  if (!calculation || precision < 0 || precision > 2) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  std::lock_guard<std::recursive_mutex> lock(calculation->context->mutex);
  try {
    calculation->plan->set_precision(static_cast<vqc::Precision>(precision));
    return GENERATIVEQC_STATUS_SUCCESS;
  } catch (...) {
    return map_exception(&calculation->context->detail);
  }
}

generativeqc_status vibeqc_calculation_last_state(const generativeqc_calculation* calculation, double* density,
                                                  uint64_t density_count, uint64_t* counts4) {
  if (!calculation || !counts4) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  const auto& d = calculation->density;
  if (density) {
    if (d.empty() || density_count != d.size()) return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    std::copy(d.begin(), d.end(), density);
  }
  counts4[0] = d.size();
  counts4[1] = calculation->fock_builds;
  counts4[2] = static_cast<uint64_t>(calculation->plan->nbf());
  counts4[3] = static_cast<uint64_t>(calculation->plan->naux());
  return GENERATIVEQC_STATUS_SUCCESS;
}

}  // extern "C"
