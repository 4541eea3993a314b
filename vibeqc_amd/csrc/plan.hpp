// Prepared DF-JK plan: owns the fitted tensor and the one-electron matrices on
// one HIP device and runs Fock builds and the RHF SCF there.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "basis.hpp"
#include "common.hpp"
#include "integrals.hpp"
#include "jk.hpp"
#include "tuning.hpp"

namespace vqc {

struct PlanOptions {
  double screening = 1e-12;
  double metric_threshold = 1e-10;
  double cj = 1.0, ck = -0.5;
  bool coulomb = true, exchange = true;
  // Device bytes the plan may hold (0: no limit).  When the resident fitted
  // tensor does not fit, the plan keeps only the metric factor and
  // regenerates B in auxiliary blocks for every Fock build (streamed plan).
  std::size_t device_budget = 0;
};

struct ScfOptions {
  unsigned max_iterations = 100;
  unsigned diis_history = 8;
  double energy_tolerance = 1e-10;
  double density_tolerance = 1e-8;
};

struct ScfOutcome {
  bool converged = false;
  double energy = 0.0, energy_change = 0.0, density_rms = 0.0;
  unsigned iterations = 0;
  unsigned reduced_iterations = 0;  // iterations run at reduced precision
  unsigned purified_iterations = 0;  // new densities from SP2 purification (the rest diagonalise)
  unsigned incremental_builds = 0;   // FP64-stage builds with K from an anchor plus a BF16X3 update
  unsigned single_term_builds = 0;   // BF16X3-stage builds with one bf16 product per logical one
  uint64_t fock_builds = 0;
  std::vector<double> density;
  std::vector<double> forces;  // 3 natom, -dE/dR, when requested
};

struct JkResult {
  std::vector<double> J, K;
};

// Outcome of DfPlan::autotune: the fastest launch shape (0 fields: the rule's
// choice), its build time and the rule's, both averages of HIP-event times.
struct AutotuneResult {
  JkShape best;
  int u_tile = 0, syrk_splits = 0;  // the tile and split count the best shape ran with
  double best_ms = 0.0, default_ms = 0.0;
  int candidates = 0;
};

// (D + D^T) / 2 of a finite n x n density that is symmetric to 1e-10 of its
// largest element (density-fitted K is built from a factor of D); throws
// std::invalid_argument otherwise.
std::vector<double> symmetrised_density(const std::vector<double>& d, int n);
// Symmetric and antisymmetric parts of a finite n x n density (the reference
// contracts raw, possibly nonsymmetric densities, fock_build.hpp:173); returns
// false, with `anti` left empty, when D is symmetric to 1e-14 of its largest
// element.  Throws std::invalid_argument for non-finite entries.
bool split_density(const std::vector<double>& d, int n, std::vector<double>& sym, std::vector<double>& anti);

// Unrestricted raw J (of the total density) and K of each spin density.
struct UhfJk {
  std::vector<double> J, Ka, Kb;
};

class DfPlan {
 public:
  // With stream/blas non-null the plan runs on that stream and rocBLAS handle
  // (owned by the caller, who keeps them alive longer than the plan);
  // otherwise it creates its own.
  DfPlan(int device, const System& orbital, const System& auxiliary, const PlanOptions& options,
         hipStream_t stream = nullptr, rocblas_handle blas = nullptr);
  ~DfPlan();
  DfPlan(const DfPlan&) = delete;
  DfPlan& operator=(const DfPlan&) = delete;

  int device() const { return device_; }
  int nbf() const { return n_; }
  int naux() const { return naux_; }
  std::size_t npair() const { return npair_; }
  int rank() const { return rank_; }
  // True when B is regenerated per Fock build instead of held on the device.
  bool streamed() const { return streamed_; }
  int auxiliary_block() const { return streamed_ ? qblk_ : ldq_; }
  int row_tiles() const { return streamed_ ? static_cast<int>(tile_shells_.size()) - 1 : 1; }
  bool cholesky() const { return cholesky_; }
  double condition() const { return condition_; }
  double nuclear_repulsion() const { return enuc_; }
  std::size_t device_bytes() const;
  const System& system() const { return orb_; }
  const PlanOptions& options() const { return opts_; }
  void set_precision(Precision p);
  Precision precision() const { return prec_; }

  // One-electron matrices (host copies).
  std::vector<double> hcore() const;
  std::vector<double> overlap() const;

  // Raw J/K for a symmetric density (host in/out).  With `gradient` non-null
  // the build runs in FP64 and *gradient receives the fixed-density
  // two-electron energy derivative dE2/dR (3 natom, Hartree/bohr).
  JkResult evaluate(const std::vector<double>& density, std::vector<double>* gradient = nullptr);
  // Raw J/K for D = 2 C C^T (C row-major n x nocc, host).
  JkResult evaluate_occupied(const std::vector<double>& occupied, int nocc);
  // Raw J/K for D = 2 C C^T from a device-resident occupied factor C
  // (row-major n x nocc) into device outputs J and K (n x n), at the plan's
  // precision, with every workspace carved from `scratch` (`bytes` long) and
  // no device or host allocation; returns after J and K are written.  A
  // reduced precision needs its B copy, made by set_precision.  Throws
  // std::length_error when the workspace does not fit.
  void evaluate_occupied_into(const double* C, int nocc, double* J, double* K, void* scratch, std::size_t bytes);
  // Timing harness (see vibeqc_fock_plan_benchmark_occupied).
  std::array<double, 5> benchmark_occupied(const std::vector<double>& occupied, int nocc, unsigned repeat);
  // Times the occupied J/K build of C (row-major n x nocc) at the plan's
  // precision for the rule's launch shape and for alternative U tile widths
  // (one to three i blocks) and SYRK split counts (half and twice the
  // rule's), `repeat` builds each after one untimed build, and records the
  // fastest in the tuning registry for this device, precision and shape.
  AutotuneResult autotune(const std::vector<double>& occupied, int nocc, unsigned repeat);
  // The launch shape used for builds of r occupied columns at precision p.
  JkShape shape_for(Precision p, int r) const;

  // Unrestricted J/K for spin densities Da, Db (unit occupation, host in/out).
  // With `gradient` non-null the builds run in FP64 and *gradient receives
  // dE2/dR of E2 = 1/2 sum (Da + Db) cJ J + 1/2 sum_s D_s cK K_s.
  UhfJk evaluate_unrestricted(const std::vector<double>& alpha, const std::vector<double>& beta,
                              std::vector<double>* gradient = nullptr);
  // UHF SCF (spin occupations from the system's electrons and multiplicity);
  // initial_density and the outcome's density are [alpha; beta] (2 n^2).
  ScfOutcome solve_unrestricted(const ScfOptions& options, const std::vector<double>* initial_density,
                                bool forces = false);
  // With forces, the outcome carries -dE/dR of the converged energy.
  ScfOutcome solve(const ScfOptions& options, const std::vector<double>* initial_density, bool forces = false);
  // -dE/dR (3 natom) at the density the last successful solve() or
  // solve_unrestricted() converged to; valid until the next operation on this plan.
  std::vector<double> converged_forces();

 private:
  void setup(const System& aux);
  // Factor the metric and apply it to A (npair x ldq, in place), or, with A
  // null, keep the factor as G_ for streamed builds.
  void factor_metric(DeviceArray<double>& M, double* A);
  // Streamed plans (device_budget below the resident footprint).
  std::size_t resident_bytes() const;
  void plan_streaming();
  // B[:, q0:q1] (npair rows, stride ldqb, padding columns zero) into Bblk_.
  void generate_block(int q0, int q1, int ldqb);
  void build_jk_streamed(const double* D, int r, int r_pos, bool exchange);
  void require_resident(const char* what) const;
  void ensure_reduced();
  void ensure_factor(int r);
  // J/K on the device: D (n x n) and the factor already in c64_ with r
  // columns, r_pos of them positive.  Results in J_, K_.
  void build_jk(Precision p, const double* D, int r, int r_pos);
  // FP64-stage build of a mixed BF16X3 solve from the anchor (Kref_, Cref_):
  // J in FP64 from D_, K = Kref + K(C C^T - Cref Cref^T) with the update in
  // BF16X3 from the U of (C + Cref)/2 and C - Cref (C = the factor in c64_,
  // r columns, all positive).  Results in J_, K_; overwrites U_.
  void build_jk_incremental(int r);
  // Whether a BF16X3 solve runs the mixed schedule's two cheaper stages:
  // single-term BF16 builds while the density change is above 1e-2 and
  // incremental FP64-stage builds (VIBEQC_INCREMENTAL_K=0 turns them off).
  bool incremental_enabled() const;
  void upload_factor(const std::vector<double>& occupied, int nocc);
  int factorise_density(const double* D);
  void sync();
  // Gradient pieces (gradient.hip).  df_two_electron_gradient needs the state
  // of an FP64 build of D_ with the factor in c64_ (rho_, FP64 U_); it
  // overwrites U_.  All add into grad (device, 3 natom).
  void df_two_electron_gradient(int r, int r_pos, double* grad);
  // Response weights of 1/2 sum D (cj J + ck K) for the state of the last
  // FP64 build (D_, rho_, factor in c64_, U_), written or added into
  // WA (npair x ldq) and WM (naux x naux).
  void df_gradient_weights(int r, int r_pos, double cj, double ck, bool accumulate, double* WA, double* WM);
  // grad += sum WA d(mu nu|P)/dR + sum WM d(P|Q)/dR; releases WA.
  void contract_df_gradient(DeviceArray<double>& WA, DeviceArray<double>& WM, double* grad);
  // J of D alone (rho_ and J_).
  void coulomb_only(Precision p, const double* D);
  // K += sum_Q B_Q A B_Q for a device n x n matrix A (row-major, in FP64):
  // A = P Q^T from the eigenpairs of A^T A (P = A V diag(s^-1/2),
  // Q = V diag(s^1/2), s the singular values), U_P and U_Q by the occupied
  // U kernel and K += U_P U_Q^T as one GEMM.  Used for the antisymmetric part
  // of a nonsymmetric density; overwrites the occupied-factor workspaces.
  void add_general_exchange(const double* A, double* K);
  // UHF two-electron gradient of device spin densities; with orbitals
  // non-null, orbitals[s] (column-major n x n) holds D_s's factor in its
  // first occupied[s] columns, else D_s is eigen-decomposed.
  void unrestricted_two_electron_gradient(const double* Da, const double* Db, const double* const* orbitals,
                                          const int* occupied, double* grad);
  void one_electron_gradient(const double* D, const double* eps, int nocc, double* grad);
  std::vector<double> nuclear_repulsion_gradient() const;

  int device_;
  std::string arch_;  // device architecture name (gfx...), the tuning registry's key
  bool forcing_ = false;  // autotune in progress: forced_ overrides the registry
  JkShape forced_;
  hipStream_t stream_ = nullptr;
  rocblas_handle blas_ = nullptr;
  bool owns_stream_ = true;
  System orb_, aux_;
  PlanOptions opts_;
  Precision prec_ = Precision::FP64;
  int bf16_terms_ = 3;  // BF16X3 products per logical one in build_jk (1 only inside a large mixed solve)
  int sp2_hint_ = 0;  // SP2 steps of the last purification: the next one's first batch
  int n_ = 0, naux_ = 0, ldq_ = 0, nocc_ = 0;
  std::size_t npair_ = 0;
  int rank_ = 0;
  bool cholesky_ = true;
  double condition_ = 0.0;
  double enuc_ = 0.0;

  DeviceArray<double> boys_;
  DeviceBasis ob_;        // orbital basis as the integral kernels see it
  AuxiliaryLayout al_;    // auxiliary functions in internal order
  DeviceArray<double> L_;  // Cholesky factor of (P|Q), column-major lower (Cholesky path only)
  DeviceArray<double> eps_;  // orbital energies of the last converged solve
  bool converged_state_ = false;  // D_, c64_, rho_, FP64 U_ and eps_ describe that solve
  // The last converged unrestricted solve: [alpha; beta] orbitals (column-major
  // n x n each), orbital energies and densities, and the spin occupations.
  bool converged_unrestricted_ = false;
  DeviceArray<double> uhf_orbitals_, uhf_energies_, uhf_density_;
  int uhf_occ_[2] = {0, 0};
  std::vector<double> unrestricted_forces();
  DeviceArray<double> S_, H_;
  DeviceArray<double> B64_;
  DeviceArray<float> B32_;
  // Fock-build workspaces.
  DeviceArray<double> dpacked_, rho_, rho_partial_, J_, K_, syrk_partial_, D_;
  DeviceArray<double> c64_;
  DeviceArray<float> c32_;
  DeviceArray<uint16_t> ct_hi_, ct_lo_;
  DeviceArray<char> U_;
  int factor_cap_ = 0, ldc_ = 0, ni_ = 1;
  OccupiedOperands ops_;
  DeviceArray<int> flag_;
  // Streamed plans: G_ (naux x naux, column-major) maps A to B (B^T = G A^T;
  // the inverse Cholesky factor, lower triangular, or the symmetric X), the
  // row tiles of the three-centre regeneration (shell ranges) and the
  // per-build block buffers.
  bool streamed_ = false;
  bool g_lower_ = false;
  DeviceArray<double> G_, Ablk_, Bblk_, Jtmp_, Ktmp_;
  // Anchor of incremental FP64-stage builds: K and the factor of the last
  // full FP64 build, and the two update factors.
  DeviceArray<double> Kref_, Cref_, facA_, facB_;
  DeviceArray<char> Ublk_;
  std::vector<int> tile_shells_;  // shell boundaries of the row tiles
  int qblk_ = 0;                  // auxiliary columns per block (multiple of 128)
};

}  // namespace vqc
