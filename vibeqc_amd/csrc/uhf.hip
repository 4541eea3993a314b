// Unrestricted (UHF) density-fitted J/K and SCF on the device.
//
// Raw J comes from the total density, raw K from each spin density
// (density_fitting.cpp:780-801), and F_s = H + cJ J + cK K_s with cK = -1
// for Hartree-Fock (rhf.cpp:451-463).  The SCF follows the reference's host
// UHF driver (mean_field_driver.cpp:346-488): a core-Hamiltonian guess whose
// beta HOMO/LUMO pair is rotated by 45 degrees when the spins are unequal
// (initial_guess/density.cpp:602-623, initial_guess/*.cpp:21-43), one DIIS
// over the stacked alpha/beta Fock matrices and commutator residuals, the
// density change as the RMS over both spin blocks, and a finalisation that
// rebuilds, diagonalises and rebuilds once more
// (mean_field_driver.cpp:201-227).  The energy is
// 1/2 sum_s tr D_s (H + F_s) + E_nuc (reference/mean_field.cpp:26-35).
//
// Every Fock build is the RHF engine's: K_s runs the occupied-factor path on
// the spin's n x n_s orbital block (unit occupation) and J comes from the
// sum of the two fused rho passes, so an unrestricted iteration costs two
// exchange builds and one Coulomb contraction per spin.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

#include "integrals_gradient.hpp"
#include "plan.hpp"
#include "scf.hpp"

namespace vqc {
namespace {

unsigned blocks_for(std::size_t count) { return static_cast<unsigned>((count + 255) / 256); }

// Columns i and j of the column-major n x n matrix W become
// (c w_i + s w_j, -s w_i + c w_j).
__global__ void rotate_columns_kernel(double* __restrict__ W, int n, int i, int j, double c, double s) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  double* wi = W + static_cast<std::size_t>(i) * n;
  double* wj = W + static_cast<std::size_t>(j) * n;
  const double a = wi[row], b = wj[row];
  wi[row] = c * a + s * b;
  wj[row] = -s * a + c * b;
}

// out (column-major n x n) = W[:, k] * e[k] for k < r, zero elsewhere.
__global__ void weighted_columns_kernel(const double* __restrict__ W, const double* __restrict__ e, int n, int r,
                                        double* __restrict__ out) {
  const std::size_t k = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= static_cast<std::size_t>(n) * n) return;
  const int col = static_cast<int>(k / n);
  out[k] = col < r ? W[k] * e[col] : 0.0;
}

// Spin occupations from the electron count and multiplicity
// (initial_guess/*.cpp:45-60).
std::pair<int, int> spin_occupations(const System& sys) {
  const long electrons = sys.electrons;
  const long excess = static_cast<long>(sys.multiplicity) - 1;
  if (electrons < 0 || excess < 0 || excess > electrons || ((electrons + excess) & 1) != 0)
    throw std::invalid_argument("electron count and multiplicity do not define integral UHF occupations");
  return {static_cast<int>((electrons + excess) / 2), static_cast<int>((electrons - excess) / 2)};
}

}  // namespace

void DfPlan::coulomb_only(Precision p, const double* D) {
  if (streamed_) {
    build_jk_streamed(D, 0, 0, false);
    return;
  }
  if (p != Precision::FP64) ensure_reduced();
  const void* B = p == Precision::FP64 ? static_cast<const void*>(B64_.data()) : static_cast<const void*>(B32_.data());
  pack_density(D, n_, dpacked_.data(), stream_);
  coulomb_rho(p, B, dpacked_.data(), npair_, ldq_, rho_.data(), rho_partial_.data(), stream_);
  coulomb_matrix(p, B, rho_.data(), npair_, ldq_, n_, J_.data(), stream_);
}

void DfPlan::unrestricted_two_electron_gradient(const double* Da, const double* Db, const double* const* orbitals,
                                                const int* occupied, double* grad) {
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  const double cj = opts_.coulomb ? opts_.cj : 0.0;
  const double ck = opts_.exchange ? opts_.ck : 0.0;
  DeviceArray<double> WA(npair_ * ldq_), WM(static_cast<std::size_t>(naux_) * naux_);
  // Coulomb weights from the total density (rho_ of Da + Db); they are
  // quadratic in the total, so they cannot be split by spin.
  assemble_fock(Da, Db, Db, 1.0, 0.0, nn, D_.data(), stream_);
  coulomb_only(Precision::FP64, D_.data());
  df_gradient_weights(0, 0, cj, 0.0, false, WA.data(), WM.data());
  if (ck != 0.0) {
    const double* spin[2] = {Da, Db};
    for (int s = 0; s < 2; ++s) {
      copy_device(D_.data(), spin[s], nn, stream_);
      int r = 0, rp = 0;
      if (orbitals && orbitals[s]) {
        r = rp = occupied[s];
        occupied_from_orbitals(orbitals[s], n_, r, ldc_, 1.0, c64_.data(), stream_);
      } else {
        const int packed = factorise_density(D_.data());
        r = packed >> 16;
        rp = packed & 0xffff;
      }
      if (r == 0) continue;
      build_jk(Precision::FP64, D_.data(), r, rp);
      df_gradient_weights(r, rp, 0.0, ck, true, WA.data(), WM.data());
    }
  }
  contract_df_gradient(WA, WM, grad);
}

UhfJk DfPlan::evaluate_unrestricted(const std::vector<double>& alpha, const std::vector<double>& beta,
                                    std::vector<double>* gradient) {
  converged_state_ = false;
  VQC_HIP(hipSetDevice(device_));
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  if (alpha.size() != nn || beta.size() != nn) throw std::invalid_argument("density shape mismatch");
  std::vector<double> a, b, anti[2];
  const bool nonsymmetric[2] = {split_density(alpha, n_, a, anti[0]), split_density(beta, n_, b, anti[1])};
  if ((nonsymmetric[0] || nonsymmetric[1]) && gradient)
    throw std::invalid_argument("the fixed-density derivative requires symmetric densities");
  ensure_factor(1);
  DeviceArray<double> Da(nn), Db(nn);
  Da.upload(a.data(), nn, stream_);
  Db.upload(b.data(), nn, stream_);
  const Precision p = gradient ? Precision::FP64 : prec_;
  UhfJk out;
  out.J.assign(nn, 0.0);
  out.Ka.assign(nn, 0.0);
  out.Kb.assign(nn, 0.0);
  if (opts_.coulomb) {
    assemble_fock(Da.data(), Db.data(), Db.data(), 1.0, 0.0, nn, D_.data(), stream_);
    coulomb_only(p, D_.data());
    J_.download(out.J.data(), nn, stream_);
  }
  if (opts_.exchange) {
    const double* spin[2] = {Da.data(), Db.data()};
    std::vector<double>* dest[2] = {&out.Ka, &out.Kb};
    for (int s = 0; s < 2; ++s) {
      copy_device(D_.data(), spin[s], nn, stream_);
      const int packed = factorise_density(D_.data());
      build_jk(p, D_.data(), packed >> 16, packed & 0xffff);
      if (nonsymmetric[s]) {
        // K of the spin density's antisymmetric part (J sees the symmetric parts only).
        DeviceArray<double> A(nn);
        A.upload(anti[s].data(), nn, stream_);
        add_general_exchange(A.data(), K_.data());
      }
      K_.download(dest[s]->data(), nn, stream_);
    }
  }
  sync();
  if (gradient) {
    require_resident("the fixed-density derivative");
    const std::size_t ng = orb_.atoms.size() * 3;
    DeviceArray<double> g(ng);
    VQC_HIP(hipMemsetAsync(g.data(), 0, g.bytes(), stream_));
    unrestricted_two_electron_gradient(Da.data(), Db.data(), nullptr, nullptr, g.data());
    gradient->assign(ng, 0.0);
    g.download(gradient->data(), ng, stream_);
    sync();
  }
  return out;
}

ScfOutcome DfPlan::solve_unrestricted(const ScfOptions& o, const std::vector<double>* initial_density, bool forces) {
  VQC_HIP(hipSetDevice(device_));
  if (forces) require_resident("nuclear forces");
  converged_state_ = false;
  converged_unrestricted_ = false;
  if (!opts_.coulomb || !opts_.exchange || opts_.cj != 1.0 || opts_.ck != -1.0)
    throw NotImplemented("unrestricted solve requires the standard Hartree-Fock J/K coefficients (1, -1)");
  const auto [na, nb] = spin_occupations(orb_);
  const int n = n_;
  if (na > n || nb > n) throw std::runtime_error("basis has fewer orbitals than required UHF spin occupations");
  const int occ[2] = {na, nb};
  const std::size_t nn = static_cast<std::size_t>(n) * n;
  const double one = 1.0, zero = 0.0;
  ensure_factor(std::max(na, 1));

  // Stacked [alpha; beta] arrays: densities, Fock matrices and orbitals.
  DeviceArray<double> X(nn), V(nn), w2(2 * static_cast<std::size_t>(n)), e(n), T1(nn), T2(nn);
  DeviceArray<double> D(2 * nn), Dn(2 * nn), F(2 * nn), Fe(2 * nn), Jt(nn), Ka(nn), W(2 * nn);
  DeviceArray<rocblas_int> info(1);
  // scalars: [0..3] energy terms, [4] density RMS, [5] count of eigensolver
  // failures (downloaded with the others).
  DeviceArray<double> scalars(8);
  VQC_HIP(hipMemsetAsync(scalars.data(), 0, scalars.bytes(), stream_));
  auto eigh = [&](const double* A, double* vectors, double* values) {
    copy_device(vectors, A, nn, stream_);
    VQC_BLAS(rocsolver_dsyevd(blas_, rocblas_evect_original, rocblas_fill_lower, n, vectors, n, values, e.data(),
                              info.data()));
    note_solver_failure(info.data(), scalars.data() + 5, stream_);
  };
  auto gemm = [&](rocblas_operation ta, rocblas_operation tb, int k, const double* A, int lda, const double* B,
                  int ldb, double* C) {
    VQC_BLAS(rocblas_dgemm(blas_, ta, tb, n, n, k, &one, A, lda, B, ldb, &zero, C, n));
  };
  auto spin_block = [nn](const DeviceArray<double>& a, int s) { return a.data() + static_cast<std::size_t>(s) * nn; };
  auto density_of = [&](const double* orbitals, int nocc, double* Dout) {
    // D_s = C_s[:, :n_s] C_s[:, :n_s]^T (unit occupation); the factor stays in c64_.
    occupied_from_orbitals(orbitals, n, nocc, ldc_, 1.0, c64_.data(), stream_);
    VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, n, n, std::max(nocc, 1), &one,
                           c64_.data(), ldc_, c64_.data(), ldc_, &zero, Dout, n));
  };
  // Orbitals of F_s in the S-orthonormal frame: W_s = X V_s, eigenvalues in w2.
  auto orbitals_of = [&](const double* Fs, int s) {
    gemm(rocblas_operation_none, rocblas_operation_none, n, Fs, n, X.data(), n, T1.data());
    gemm(rocblas_operation_none, rocblas_operation_none, n, X.data(), n, T1.data(), n, T2.data());
    eigh(T2.data(), V.data(), w2.data() + static_cast<std::size_t>(s) * n);
    gemm(rocblas_operation_none, rocblas_operation_none, n, X.data(), n, V.data(), n, W.data() + s * nn);
  };
  // Occupied block of F_s during the iterations: SP2 purification of
  // F~ = X F_s X and a factor G of the projector (projected from the previous
  // iteration's, else a pivoted Cholesky factor; the first
  // n_s columns of W_s are X G), as in the restricted solve; a spin whose
  // purification does not reach an integer trace, or whose factor is rank
  // deficient, is diagonalised instead.  The core guess (its frontier
  // rotation needs the virtual orbital) and the finalisation diagonalise.
  ScfOutcome out;
  const int omax = std::max(na, nb);
  const bool sp2_spin[2] = {na > 0 && na < n && pivoted_cholesky_fits(n, na),
                            nb > 0 && nb < n && pivoted_cholesky_fits(n, nb)};
  // One projector factor per spin (n x n_s at offset s n omax), kept between
  // iterations: each new one is first tried as a projection of the previous.
  const std::size_t gstride = static_cast<std::size_t>(n) * std::max(omax, 1);
  DeviceArray<double> G(2 * gstride), sp2_state(kSp2StateSize), sp2_partial(sp2_partial_size(n));
  DeviceArray<double> projA(gstride), projM(static_cast<std::size_t>(std::max(omax, 1)) * std::max(omax, 1));
  DeviceArray<int> sp2_status(1), proj_status(1);
  DeviceArray<rocblas_int> proj_info(1);
  bool have_factor[2] = {false, false};
  int sp2_hint[2] = {0, 0};
  auto purify = [&](int s) {  // T2 -> projector in V; T1 is scratch
    const double target = static_cast<double>(occ[s]);
    sp2_begin(T2.data(), n, sp2_partial.data(), sp2_state.data(), V.data(), stream_);
    int steps = 0, batch = sp2_hint[s] > 0 ? sp2_hint[s] : 32;
    double hstate[8];
    for (;;) {
      for (int j = 0; j < batch && steps < kSp2MaxSteps; ++j, ++steps) {
        gemm(rocblas_operation_none, rocblas_operation_none, n, V.data(), n, V.data(), n, T1.data());
        sp2_step(V.data(), T1.data(), n, target, sp2_state.data(), steps, stream_);
      }
      sp2_state.download(hstate, 8, stream_);
      const Sp2Summary sum = sp2_summary(hstate, steps);
      if (sum.converged && std::abs(sum.trace - target) < 1e-6) {
        sp2_hint[s] = sum.applied;
        return true;
      }
      if (sum.converged || sum.degenerate || steps >= kSp2MaxSteps) return false;
      batch = 4;
    }
  };
  auto occupied_of = [&](const double* Fs, int s) {
    if (sp2_spin[s]) {
      gemm(rocblas_operation_none, rocblas_operation_none, n, Fs, n, X.data(), n, T1.data());
      gemm(rocblas_operation_none, rocblas_operation_none, n, X.data(), n, T1.data(), n, T2.data());
      if (purify(s)) {
        double* Gs = G.data() + s * gstride;
        if (have_factor[s])
          projected_factor(blas_, V.data(), n, occ[s], Gs, projA.data(), projM.data(), proj_info.data(),
                           proj_status.data(), stream_);
        pivoted_cholesky(V.data(), n, occ[s], Gs, sp2_status.data(), stream_,
                         have_factor[s] ? proj_status.data() : nullptr);
        int status = 1;
        sp2_status.download(&status, 1, stream_);
        have_factor[s] = status == 0;
        if (status == 0) {
          VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, n, occ[s], n, &one, X.data(),
                                 n, Gs, n, &zero, W.data() + s * nn, n));
          ++out.purified_iterations;
          return;
        }
      }
    }
    orbitals_of(Fs, s);
  };
  // X = S^-1/2 (src/scf/reference/linalg.cpp:123-142).
  eigh(S_.data(), V.data(), w2.data());
  {
    std::vector<double> hw(n);
    double hs[6];
    w2.download(hw.data(), n, stream_);
    scalars.download(hs, 6, stream_);
    sync();
    if (hs[5] != 0.0) throw NumericalFailure("overlap eigensolver did not converge");
    if (hw.front() < 1e-10) throw NumericalFailure("overlap matrix is singular or severely linearly dependent");
  }
  scale_columns_inv_sqrt(V.data(), w2.data(), n, T1.data(), stream_);
  gemm(rocblas_operation_none, rocblas_operation_transpose, n, T1.data(), n, V.data(), n, X.data());

  // rank[s] < 0: the factor of D_s comes from an eigendecomposition (seeds).
  int rank[2] = {na, nb};
  if (initial_density) {
    if (initial_density->size() != 2 * nn) throw std::invalid_argument("unrestricted initial density must be 2 n^2");
    for (int s = 0; s < 2; ++s) {
      const std::vector<double> part(initial_density->begin() + s * nn, initial_density->begin() + (s + 1) * nn);
      const std::vector<double> sym = symmetrised_density(part, n);
      copy_to_device(spin_block(D, s), sym.data(), nn * sizeof(double), stream_);
      copy_device(D_.data(), spin_block(D, s), nn, stream_);
      const int packed = factorise_density(D_.data());
      if ((packed >> 16) != (packed & 0xffff))
        throw std::invalid_argument("initial spin densities must be positive semidefinite");
      rank[s] = -1;
    }
  } else {
    // Core guess; beta's frontier pair rotated when the spins differ.
    orbitals_of(H_.data(), 0);
    copy_device(W.data() + nn, W.data(), nn, stream_);
    if (na != nb && nb > 0 && nb < n) {
      const double c = 0.7071067811865476, sn = 0.7071067811865476;
      hipLaunchKernelGGL(rotate_columns_kernel, dim3(blocks_for(n)), dim3(256), 0, stream_, W.data() + nn, n, nb - 1,
                         nb, c, sn);
      VQC_LAUNCH("rotate_columns_kernel");
    }
    for (int s = 0; s < 2; ++s) density_of(W.data() + s * nn, occ[s], spin_block(D, s));
  }

  // One UHF Fock build: F_s = H + J(Da + Db) - K(D_s) from the spin factors.
  auto build_focks = [&](Precision p, const DeviceArray<double>& Dm, const DeviceArray<double>* orbitals,
                         DeviceArray<double>& Fout) {
    for (int s = 0; s < 2; ++s) {
      const double* Ds = spin_block(Dm, s);
      int r = occ[s], rp = occ[s];
      if (!orbitals) {
        copy_device(D_.data(), Ds, nn, stream_);
        const int packed = factorise_density(D_.data());
        r = packed >> 16;
        rp = packed & 0xffff;
      } else {
        occupied_from_orbitals(orbitals->data() + s * nn, n, r, ldc_, 1.0, c64_.data(), stream_);
      }
      build_jk(p, Ds, r, rp);
      if (s == 0) {
        copy_device(Jt.data(), J_.data(), nn, stream_);
        copy_device(Ka.data(), K_.data(), nn, stream_);
      } else {
        assemble_fock(Jt.data(), J_.data(), J_.data(), 1.0, 0.0, nn, Jt.data(), stream_);
      }
    }
    assemble_fock(H_.data(), Jt.data(), Ka.data(), 1.0, -1.0, nn, Fout.data(), stream_);
    assemble_fock(H_.data(), Jt.data(), K_.data(), 1.0, -1.0, nn, Fout.data() + nn, stream_);
  };
  // 1/2 sum_s tr D_s (H + F_s) into scalars[0..3].
  auto energy_terms = [&](const DeviceArray<double>& Dm, const DeviceArray<double>& Fm) {
    dot_pair(Dm.data(), H_.data(), Fm.data(), nn, scalars.data(), stream_);
    dot_pair(Dm.data() + nn, H_.data(), Fm.data() + nn, nn, scalars.data() + 2, stream_);
  };

  const unsigned cap = std::max(1u, o.diis_history);
  const std::size_t n2 = 2 * nn;
  DeviceArray<double> focks(n2 * cap), errors(n2 * cap), gram(static_cast<std::size_t>(cap) * cap), coef(cap);
  DeviceArray<int> slots(cap);
  std::vector<int> order;
  int next_slot = 0;

  const bool reduced = prec_ != Precision::FP64;
  const char* trace_env = std::getenv("VIBEQC_TRACE");
  const bool trace_scf = trace_env && std::strcmp(trace_env, "scf") == 0;
  // Reduced-precision stage and hand-over as in the restricted solve.
  const double e_switch = prec_ == Precision::BF16X3 ? 1e-4 : 1e-6;
  const double d_switch = prec_ == Precision::BF16X3 ? 1e-3 : 1e-5;
  double best_rms = std::numeric_limits<double>::infinity();
  int stalled = 0;
  bool in_reduced = reduced;
  // Single-term first stage, as in the restricted solve: while the density
  // change is above kSingleTermRms the BF16X3 builds use one bf16 product per
  // logical one, then three-term builds continue with a fresh DIIS history.
  constexpr double kSingleTermRms = 1e-2;
  struct TermsReset {
    DfPlan* plan;
    ~TermsReset() { plan->bf16_terms_ = 3; }
  } terms_reset{this};
  bf16_terms_ = reduced && incremental_enabled() ? 1 : 3;
  double previous = std::numeric_limits<double>::infinity();
  bool converged = false;
  unsigned it = 0;
  double energy = 0.0, de = 0.0, rms = 0.0;
  bool from_orbitals = !initial_density;
  for (it = 1; it <= o.max_iterations; ++it) {
    const Precision p = in_reduced ? prec_ : Precision::FP64;
    build_focks(p, D, from_orbitals ? &W : nullptr, F);
    out.fock_builds += 2;
    if (in_reduced) ++out.reduced_iterations;
    if (in_reduced && bf16_terms_ == 1) out.single_term_builds += 2;
    energy_terms(D, F);
    const double* Fuse = F.data();
    if (cap >= 2) {
      int slot;
      if (order.size() < cap) {
        slot = next_slot++;
      } else {
        slot = order.front();
        order.erase(order.begin());
      }
      order.push_back(slot);
      copy_device(focks.data() + static_cast<std::size_t>(slot) * n2, F.data(), n2, stream_);
      for (int s = 0; s < 2; ++s) {
        // Residual F_s D_s S - S D_s F_s.
        gemm(rocblas_operation_none, rocblas_operation_none, n, F.data() + s * nn, n, spin_block(D, s), n, T1.data());
        gemm(rocblas_operation_none, rocblas_operation_none, n, T1.data(), n, S_.data(), n, T2.data());
        antisymmetrise(T2.data(), n, errors.data() + static_cast<std::size_t>(slot) * n2 + s * nn, stream_);
      }
      const int m = static_cast<int>(order.size());
      slots.upload(order.data(), m, stream_);
      diis_gram_row(errors.data(), n2, slot, m, slots.data(), static_cast<int>(cap), gram.data(), stream_);
      if (m >= 2) {
        diis_solve(gram.data(), m, slots.data(), static_cast<int>(cap), coef.data(), stream_);
        diis_extrapolate(focks.data(), n2, m, slots.data(), coef.data(), Fe.data(), stream_);
        Fuse = Fe.data();
      }
    }
    for (int s = 0; s < 2; ++s) {
      occupied_of(Fuse + s * nn, s);
      density_of(W.data() + s * nn, occ[s], spin_block(Dn, s));
    }
    from_orbitals = true;
    rms_difference(Dn.data(), D.data(), n2, scalars.data() + 4, stream_);
    double hs[6];
    scalars.download(hs, 6, stream_);
    sync();
    if (hs[5] != 0.0) throw NumericalFailure("Fock eigensolver did not converge");
    energy = 0.5 * (hs[0] + hs[1] + hs[2] + hs[3]) + enuc_;
    rms = hs[4];
    if (!std::isfinite(energy) || !std::isfinite(rms)) throw NumericalFailure("non-finite SCF iterate");
    if (trace_scf)
      std::fprintf(stderr, "{\"trace\": \"uhf\", \"it\": %u, \"energy\": %.17g, \"rms\": %.17g, \"reduced\": %d}\n", it,
                   energy, rms, in_reduced ? 1 : 0);
    de = std::isfinite(previous) ? std::abs(energy - previous) : std::numeric_limits<double>::infinity();
    std::swap(D, Dn);
    if (in_reduced && bf16_terms_ == 1 && it > 1 && rms < kSingleTermRms) {
      bf16_terms_ = 3;
      order.clear();
      next_slot = 0;
      previous = std::numeric_limits<double>::infinity();
      continue;
    }
    if (in_reduced) {
      const bool settled = it > 1 && de < std::max(e_switch, o.energy_tolerance) &&
                           rms < std::max(d_switch, o.density_tolerance);
      if (rms < 0.5 * best_rms) {
        best_rms = rms;
        stalled = 0;
      } else if (rms < 1e-2) {
        ++stalled;
      }
      if (settled || stalled >= 3 || it >= o.max_iterations / 2) {
        in_reduced = false;
        order.clear();
        next_slot = 0;
        previous = std::numeric_limits<double>::infinity();
        continue;
      }
    } else if (it > 1 && std::isfinite(previous) && de < o.energy_tolerance && rms < o.density_tolerance) {
      converged = true;
      break;
    }
    previous = energy;
  }
  out.iterations = std::min(it, o.max_iterations);
  out.energy = energy;
  out.energy_change = de;
  out.density_rms = rms;
  if (!converged) {
    out.converged = false;
    out.density.resize(n2);
    D.download(out.density.data(), n2, stream_);
    sync();
    return out;
  }
  // Finalisation: rebuild from the converged densities, diagonalise, rebuild.
  build_focks(Precision::FP64, D, &W, F);
  for (int s = 0; s < 2; ++s) {
    orbitals_of(F.data() + s * nn, s);
    density_of(W.data() + s * nn, occ[s], spin_block(D, s));
  }
  build_focks(Precision::FP64, D, &W, F);
  energy_terms(D, F);
  out.fock_builds += 4;
  double hs[6];
  scalars.download(hs, 6, stream_);
  out.density.resize(n2);
  D.download(out.density.data(), n2, stream_);
  sync();
  if (hs[5] != 0.0) throw NumericalFailure("final Fock eigensolver did not converge");
  out.energy = 0.5 * (hs[0] + hs[1] + hs[2] + hs[3]) + enuc_;
  out.converged = std::isfinite(out.energy);
  if (out.converged) {
    uhf_orbitals_ = std::move(W);
    uhf_energies_ = std::move(w2);
    uhf_density_ = std::move(D);
    uhf_occ_[0] = na;
    uhf_occ_[1] = nb;
    converged_unrestricted_ = true;
    converged_state_ = true;
  }
  if (forces && out.converged) out.forces = converged_forces();
  return out;
}

std::vector<double> DfPlan::unrestricted_forces() {
  // -dE/dR (analytic_uhf_forces): one-electron terms of the total density and
  // the total energy-weighted density E = sum_s C_s diag(e_s) C_s^T, the DF
  // two-electron derivative, and the nuclear repulsion.
  const int n = n_;
  const std::size_t nn = static_cast<std::size_t>(n) * n;
  const double one = 1.0;
  const double* W = uhf_orbitals_.data();
  const double* D = uhf_density_.data();
  const std::size_t ng = orb_.atoms.size() * 3;
  DeviceArray<double> g(ng), T(nn), E(nn), Dt(nn);
  VQC_HIP(hipMemsetAsync(g.data(), 0, g.bytes(), stream_));
  for (int s = 0; s < 2; ++s) {
    hipLaunchKernelGGL(weighted_columns_kernel, dim3(blocks_for(nn)), dim3(256), 0, stream_, W + s * nn,
                       uhf_energies_.data() + static_cast<std::size_t>(s) * n, n, uhf_occ_[s], T.data());
    VQC_LAUNCH("weighted_columns_kernel");
    const double b = s == 0 ? 0.0 : 1.0;
    VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_transpose, n, n, std::max(uhf_occ_[s], 1),
                           &one, T.data(), n, W + s * nn, n, &b, E.data(), n));
  }
  assemble_fock(D, D + nn, D + nn, 1.0, 0.0, nn, Dt.data(), stream_);
  compute_one_electron_gradient(orb_, ob_, boys_.data(), Dt.data(), E.data(), g.data(), stream_);
  const double* orbitals[2] = {W, W + nn};
  unrestricted_two_electron_gradient(D, D + nn, orbitals, uhf_occ_, g.data());
  std::vector<double> hg(ng);
  g.download(hg.data(), ng, stream_);
  sync();
  const auto gn = nuclear_repulsion_gradient();
  std::vector<double> forces(ng);
  for (std::size_t k = 0; k < ng; ++k) {
    forces[k] = -(hg[k] + gn[k]);
    if (!std::isfinite(forces[k])) throw NumericalFailure("non-finite analytic force");
  }
  return forces;
}

}  // namespace vqc
