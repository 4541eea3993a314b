// Ragged batch execution: shape buckets, largest-first scheduling over a pool
// of stream-owning host workers, per-item failure isolation and retained
// warm-start densities (reference semantics: generativeqc.h:1158-1183,
// 1431-1586; fleet.cpp:26-36, 180-196).
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <tuple>

namespace vqc {
namespace {

// Routes this thread's device allocations to a worker's stream and block
// cache for the lifetime of the object; nested scopes restore the enclosing one.
class AllocationScope {
 public:
  AllocationScope(hipStream_t s, BlockCache* cache)
      : previous_(allocation_stream()), previous_cache_(allocation_cache()) {
    allocation_stream() = s;
    allocation_cache() = cache;
  }
  ~AllocationScope() {
    allocation_stream() = previous_;
    allocation_cache() = previous_cache_;
  }
  AllocationScope(const AllocationScope&) = delete;
  AllocationScope& operator=(const AllocationScope&) = delete;

 private:
  hipStream_t previous_;
  BlockCache* previous_cache_;
};

void move_atoms(System& sys, const std::vector<double>& xyz) {
  if (xyz.size() != sys.atoms.size() * 3) throw std::invalid_argument("batch coordinates do not match the atom count");
  for (std::size_t a = 0; a < sys.atoms.size(); ++a)
    for (int x = 0; x < 3; ++x) {
      const double v = xyz[a * 3 + x];
      if (!std::isfinite(v)) throw std::invalid_argument("batch coordinates must be finite");
      sys.atoms[a].xyz[x] = v;
    }
}

}  // namespace

BatchPlan::BatchPlan(int device, std::vector<System> systems, std::vector<System> aux, const BatchOptions& options)
    : device_(device), systems_(std::move(systems)), aux_(std::move(aux)), opts_(options) {
  if (systems_.empty()) throw std::invalid_argument("a batch needs at least one system");
  if (aux_.size() != systems_.size()) throw std::invalid_argument("one auxiliary basis per batch system is required");
  for (std::size_t i = 0; i < systems_.size(); ++i)
    if (aux_[i].atoms.size() != systems_[i].atoms.size())
      throw std::invalid_argument("auxiliary basis atoms do not match the batch system");
  // Shape buckets in order of first appearance.
  std::map<std::tuple<std::size_t, std::size_t, int, int>, int> ids;
  bucket_.resize(systems_.size());
  for (std::size_t i = 0; i < systems_.size(); ++i) {
    // (nbf, naux, n_alpha, n_beta); a restricted member has n_alpha = n_beta.
    const int excess = static_cast<int>(systems_[i].multiplicity) - 1;
    const int na = (systems_[i].electrons + excess) / 2, nb = (systems_[i].electrons - excess) / 2;
    const auto key = std::make_tuple(systems_[i].nbf(), aux_[i].nbf(), na, nb);
    auto it = ids.find(key);
    if (it == ids.end()) it = ids.emplace(key, static_cast<int>(ids.size())).first;
    bucket_[i] = it->second;
  }
  buckets_ = static_cast<int>(ids.size());
  // Largest first: the work of a system grows like nbf^2 naux.
  order_.resize(systems_.size());
  for (std::size_t i = 0; i < order_.size(); ++i) order_[i] = i;
  auto cost = [&](std::size_t i) {
    const double n = static_cast<double>(systems_[i].nbf());
    return n * n * static_cast<double>(aux_[i].nbf());
  };
  std::stable_sort(order_.begin(), order_.end(), [&](std::size_t a, std::size_t b) { return cost(a) > cost(b); });
  warm_.resize(systems_.size());
}

void BatchPlan::restore_warm_states(const std::vector<const WarmState*>& states) {
  if (!opts_.warm_starts) throw std::invalid_argument("restoring seeds requires warm starts enabled");
  if (states.size() != systems_.size()) throw std::invalid_argument("one warm state per batch system is required");
  for (std::size_t i = 0; i < states.size(); ++i) {
    const WarmState* w = states[i];
    if (!w) continue;
    const std::size_t n = static_cast<std::size_t>(systems_[i].nbf());
    const std::size_t spins = opts_.unrestricted ? 2 : 1;
    if (w->density.size() != spins * n * n || w->coordinates.size() != systems_[i].atoms.size() * 3)
      throw std::invalid_argument("warm state shape does not match the batch system");
    double scale = 0.0;
    for (double v : w->density) {
      if (!std::isfinite(v)) throw std::invalid_argument("warm state density must be finite");
      scale = std::max(scale, std::abs(v));
    }
    for (double v : w->coordinates)
      if (!std::isfinite(v)) throw std::invalid_argument("warm state coordinates must be finite");
    for (std::size_t b = 0; b < spins; ++b)
      for (std::size_t r = 0; r < n; ++r)
        for (std::size_t c = 0; c < r; ++c)
          if (std::abs(w->density[b * n * n + r * n + c] - w->density[b * n * n + c * n + r]) >
              1e-10 * std::max(1.0, scale))
            throw std::invalid_argument("warm state density must be symmetric");
    if (!std::isfinite(w->energy) || !std::isfinite(w->energy_change) || !std::isfinite(w->density_rms) ||
        w->density_rms < 0.0 || w->iterations < 0)
      throw std::invalid_argument("warm state diagnostics must be finite");
  }
  for (std::size_t i = 0; i < states.size(); ++i)
    if (states[i]) warm_[i] = *states[i];
}

void BatchPlan::clear_warm_states() {
  for (auto& w : warm_) w = WarmState{};
}

std::vector<BatchItemOutcome> BatchPlan::execute(const std::vector<const std::vector<double>*>& coordinates,
                                                 const std::vector<bool>& forces) {
  const std::size_t count = systems_.size();
  if (!coordinates.empty() && coordinates.size() != count)
    throw std::invalid_argument("batch coordinate count mismatch");
  if (forces.size() != count) throw std::invalid_argument("batch force request count mismatch");
  VQC_HIP(hipSetDevice(device_));
  std::vector<BatchItemOutcome> out(count);
  // Worker count: the option, else VIBEQC_BATCH_WORKERS, else 8.
  int requested = opts_.workers;
  if (requested <= 0)
    if (const char* env = std::getenv("VIBEQC_BATCH_WORKERS")) requested = std::atoi(env);
  const int workers = std::max(1, std::min<int>(requested > 0 ? requested : 8, static_cast<int>(count)));
  // One stream and rocBLAS handle per worker, created before and destroyed
  // after all plans that borrow them.
  std::vector<hipStream_t> streams(workers, nullptr);
  std::vector<rocblas_handle> handles(workers, nullptr);
  // Device memory of each worker's plans (recycled within the worker only).
  std::vector<BlockCache> caches(workers);
  auto release_resources = [&]() {
    for (int w = 0; w < workers; ++w) {
      if (streams[w]) (void)hipStreamSynchronize(streams[w]);
      if (handles[w]) rocblas_destroy_handle(handles[w]);
      if (streams[w]) (void)hipStreamDestroy(streams[w]);
    }
  };
  try {
    for (int w = 0; w < workers; ++w) {
      VQC_HIP(hipStreamCreateWithFlags(&streams[w], hipStreamNonBlocking));
      VQC_BLAS(rocblas_create_handle(&handles[w]));
      VQC_BLAS(rocblas_set_stream(handles[w], streams[w]));
      // Deterministic library kernels (see DfPlan): no atomics.
      VQC_BLAS(rocblas_set_atomics_mode(handles[w], rocblas_atomics_not_allowed));
    }
  } catch (...) {
    release_resources();
    throw;
  }
  // Execution in waves of the largest-first order.  Within a wave the plans
  // are set up one after another (integral generation, metric factorisation,
  // B), then their SCFs run concurrently, one worker per stream, and forces
  // are evaluated one system at a time.  Setup and derivative-integral
  // kernels therefore never share the device with another system's work.
  // Concurrent SCFs are bitwise reproducible.  Setups that overlap other
  // members' work are not: with setups on several hardware queues at once,
  // identical device inputs (checksummed) gave different metric, three-centre
  // and Cholesky results, and with every setup kernel free of per-lane
  // scratch memory, setups overlapping other members' SCFs still turned one
  // energy of 48 into NaN and moved others by up to 1e-6 Eh.  The cause is
  // not established, so setups stay serial.
  const std::size_t wave = static_cast<std::size_t>(workers) * 4;
  std::vector<std::unique_ptr<DfPlan>> plans(count);
  auto setup_item = [&](std::size_t k) {
    const std::size_t i = order_[k];
    const int w = static_cast<int>(k % workers);
    BatchItemOutcome& r = out[i];
    r.bucket = bucket_[i];
    AllocationScope scope(streams[w], &caches[w]);
    try {
      System sys = systems_[i];
      System aux = aux_[i];
      if (!coordinates.empty() && coordinates[i]) {
        move_atoms(sys, *coordinates[i]);
        move_atoms(aux, *coordinates[i]);
      }
      plans[k] = std::make_unique<DfPlan>(device_, sys, aux, opts_.plan, streams[w], handles[w]);
      plans[k]->set_precision(opts_.precision);
      VQC_HIP(hipStreamSynchronize(streams[w]));
    } catch (...) {
      r.error = std::current_exception();
      plans[k].reset();
    }
  };
  auto solve_item = [&](std::size_t k) {
    const std::size_t i = order_[k];
    BatchItemOutcome& r = out[i];
    if (!plans[k]) return;
    try {
      DfPlan& plan = *plans[k];
      const std::vector<double>* seed = nullptr;
      const std::size_t spins = opts_.unrestricted ? 2 : 1;
      if (opts_.warm_starts && warm_[i].density.size() == spins * plan.nbf() * plan.nbf()) {
        seed = &warm_[i].density;
        r.warm_start_used = true;
      }
      auto run = [&](const std::vector<double>* s) {
        return opts_.unrestricted ? plan.solve_unrestricted(opts_.scf, s, false) : plan.solve(opts_.scf, s, false);
      };
      try {
        r.scf = run(seed);
      } catch (const std::invalid_argument&) {
        if (!seed) throw;
        // A retained density the guards reject falls back to the core guess.
        r.warm_start_used = false;
        r.warm_start_fallback = true;
        r.scf = run(nullptr);
      }
    } catch (...) {
      r.error = std::current_exception();
    }
  };
  // Forces and the retained seed of a converged item (on worker k % workers).
  auto finish_item = [&](std::size_t k) {
    const std::size_t i = order_[k];
    BatchItemOutcome& r = out[i];
    if (!plans[k] || r.error || !r.scf.converged) return;
    try {
      if (forces[i]) {
        AllocationScope scope(streams[k % workers], &caches[k % workers]);
        r.scf.forces = plans[k]->converged_forces();
      }
      if (opts_.warm_starts && warm_updates_) {
        WarmState& ws = warm_[i];
        ws.density = r.scf.density;
        if (!coordinates.empty() && coordinates[i]) {
          ws.coordinates = *coordinates[i];
        } else {
          ws.coordinates.clear();
          for (const auto& a : systems_[i].atoms) ws.coordinates.insert(ws.coordinates.end(), a.xyz, a.xyz + 3);
        }
        ws.energy = r.scf.energy;
        ws.energy_change = r.scf.energy_change;
        ws.density_rms = r.scf.density_rms;
        ws.iterations = static_cast<int>(r.scf.iterations);
      }
    } catch (...) {
      r.error = std::current_exception();
    }
  };
  // VIBEQC_TRACE set: one JSON line per wave with its phase times.
  const bool trace = std::getenv("VIBEQC_TRACE") != nullptr;
  using clock = std::chrono::steady_clock;
  auto ms_since = [](clock::time_point t) { return std::chrono::duration<double, std::milli>(clock::now() - t).count(); };
  try {
    for (std::size_t w0 = 0; w0 < count; w0 += wave) {
      const std::size_t w1 = std::min(count, w0 + wave);
      const auto t_setup = clock::now();
      for (std::size_t k = w0; k < w1; ++k) setup_item(k);
      const double setup_ms = ms_since(t_setup);
      const auto t_scf = clock::now();
      auto run = [&](int w) {
        if (hipSetDevice(device_) != hipSuccess) return;
        AllocationScope scope(streams[w], &caches[w]);
        for (std::size_t k = w0 + static_cast<std::size_t>(w); k < w1; k += workers) solve_item(k);
        (void)hipStreamSynchronize(streams[w]);
      };
      std::vector<std::thread> pool;
      for (int w = 1; w < workers; ++w) pool.emplace_back(run, w);
      run(0);
      for (auto& t : pool) t.join();
      const double scf_ms = ms_since(t_scf);
      const auto t_forces = clock::now();
      for (std::size_t k = w0; k < w1; ++k) finish_item(k);
      const double forces_ms = ms_since(t_forces);
      for (std::size_t k = w0; k < w1; ++k) plans[k].reset();
      if (trace)
        std::fprintf(stderr,
                     "{\"trace\": \"batch_wave\", \"first\": %zu, \"systems\": %zu, \"setup_ms\": %.3f, "
                     "\"scf_ms\": %.3f, \"forces_ms\": %.3f}\n",
                     w0, w1 - w0, setup_ms, scf_ms, forces_ms);
    }
  } catch (...) {
    plans.clear();
    release_resources();
    throw;
  }
  release_resources();
  return out;
}

}  // namespace vqc
